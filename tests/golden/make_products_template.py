"""Writes gpufoam_amd/mechanisms/gri/products_template.json: the state of
stoichiometric CH4-air after 0.1 s of constant-pressure adiabatic reaction
from 1400 K at 1e5 Pa, integrated by the CPU oracle (Rosenbrock34,
relTol 1e-6, absTol 1e-10).  Run once; the JSON is the committed fixture.

    python tests/golden/make_products_template.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from gpufoam_amd import Mechanism, ode_inputs  # noqa: E402
from gpufoam_amd.mechanism import MECH_DIR  # noqa: E402
from gpufoam_amd.workloads import ch4_air  # noqa: E402
from oracle_lib import Oracle  # noqa: E402


def main():
    m = Mechanism.load("gri")
    O = Oracle(m)
    nS = m.n_species
    y = np.zeros(nS + 2)
    y[:nS] = ch4_air(m)
    y[nS], y[nS + 1] = 1400.0, 1e5
    ode = ode_inputs("Rosenbrock34", relTol=1e-6, absTol=1e-10)
    y, dx, st, stats = O.ode_solve(ode, 0.0, 0.1, y, 1e-7)
    assert st == 0, st
    y[:nS] = np.maximum(y[:nS], 0)
    out = {"T_K": y[nS], "p_Pa": 1e5, "t_s": 0.1, "steps": int(stats[1]),
           "Y": {s: float(y[i]) for i, s in enumerate(m.species) if y[i] > 0}}
    with open(os.path.join(MECH_DIR, "gri", "products_template.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"T={y[nS]:.2f} K after {stats[1]} steps")


if __name__ == "__main__":
    main()
