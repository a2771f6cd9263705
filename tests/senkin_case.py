"""The CHEMKIN-II SENKIN ignition cases of gpuFoam's tutorials
(tutorials/{h2,gri}/chemkin/senk.inp: constant pressure, adiabatic) as
test inputs, and the comparison of a computed T(t) with the SENKIN trace
(tests/golden/chemkin/<name>/senkin.json)."""
import json
import os

import numpy as np

from gpufoam_amd import Mechanism

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = {"h2": dict(sample=1e-6, end=1e-3), "gri": dict(sample=1e-5, end=7e-2)}
ATM = 101325.0


def load(name):
    """(mechanism compiled from chem.inp + therm.dat, initial state [nS+2, 1],
    SENKIN record)."""
    d = os.path.join(HERE, "golden", "chemkin", name)
    with open(os.path.join(d, "chem.inp")) as f:
        chem = f.read()
    with open(os.path.join(d, "therm.dat")) as f:
        therm = f.read()
    with open(os.path.join(d, "senkin.json")) as f:
        rec = json.load(f)
    m = Mechanism.from_chemkin(chem, therm, name)
    nS, W = m.n_species, m.arrays["W"]
    X = np.zeros(nS)
    for k, v in rec["mole_fractions"].items():
        X[m.index(k)] = v
    y = np.zeros((nS + 2, 1))
    y[:nS, 0] = X * W / (X * W).sum()
    y[nS, 0] = rec["T0"]
    y[nS + 1, 0] = rec["pressure_atm"] * ATM
    return m, y, rec


def compare(ts, Ts, rec):
    """Ignition time (first crossing of the SENKIN criterion, linear
    interpolation), final temperature, and the largest |T - T_senkin| at the
    SENKIN output times outside the ignition front (before 0.95 t_ign or
    after 1.5 t_ign)."""
    ts, Ts = np.asarray(ts), np.asarray(Ts)
    i = int(np.argmax(Ts > rec["ignition_T"]))
    assert i > 0, "no ignition"
    tig = ts[i - 1] + (rec["ignition_T"] - Ts[i - 1]) * (ts[i] - ts[i - 1]) / (Ts[i] - Ts[i - 1])
    tt, TT = np.array(rec["time"]), np.array(rec["T"])
    away = (tt < 0.95 * rec["ignition_time"]) | (tt > 1.5 * rec["ignition_time"])
    dT = np.abs(np.interp(tt, ts, Ts) - TT)[away].max()
    return tig / rec["ignition_time"] - 1.0, Ts[-1] - TT[-1], dT
