"""Rate types OpenFOAM supports and gpuFoam rejects
(makeMoreGpuChemistryModels.C:59-95): SRI fall-off, chemically activated
Lindemann / Troe / SRI, Landau-Teller, Janev and power series.  There is no
gpuFoam code for them, so no reference vectors: the oracle's restatement is
pinned here by the published closed forms, their limits and finite
differences of the Jacobian's temperature derivative (CPU).

The test mechanism is the H2 mechanism with seven reactions recast to the
new types (`extended_h2`), also used by the GPU tests.
"""
import types

import numpy as np
import pytest

from gpufoam_amd import GfcError, Mechanism, _capi, ode_inputs
from gpufoam_amd.mechanism import chemkin_to_foam, foam_reactions_text
from gpufoam_amd.workloads import make_states
from oracle_lib import Oracle

SRI = [0.6, 250.0, 900.0, 1.2, 0.04]
LT = [7.0, -30.0]
JAN = [0.3, -0.05, 0.004, 1e-4, -2e-5, 1e-6, 0.0, 0.0, 0.0]
PS = [150.0, -4.0e4, 1.0e6, -5.0e8]
# reaction index -> (new type, extra parameters)
RECAST = {
    3: (_capi.RATE_SRI, SRI),            # H + O2 (+M) = HO2, Lindemann -> SRI fall-off
    4: (_capi.RATE_LINDEMANN_CA, None),  # H + O2 (+N2) = HO2
    5: (_capi.RATE_SRI_CA, SRI),         # H + O2 (+H2) = HO2
    22: (_capi.RATE_TROE_CA, None),      # OH + OH (+M) = H2O2
    0: (_capi.RATE_LANDAU_TELLER, LT),   # OH + H2 = H + H2O
    2: (_capi.RATE_JANEV, JAN),          # O + H2 = OH + H
    10: (_capi.RATE_POWER_SERIES, PS),   # H + HO2 = H2 + O2
}


def extended_h2() -> Mechanism:
    base = Mechanism.load("h2")
    a = {k: np.array(v) for k, v in base.arrays.items()}
    for r, (t, x) in RECAST.items():
        a["type"][r] = t
        if x is not None:
            a["extra"][r, :len(x)] = x
    view = types.SimpleNamespace(arrays=a, equations=base.equations, species=base.species,
                                 n_species=base.n_species, n_reactions=base.n_reactions,
                                 Tlow=base.Tlow, Thigh=base.Thigh)
    thermo, _ = base.to_foam()
    return Mechanism(thermo, foam_reactions_text(view), "h2x")


@pytest.fixture(scope="module")
def mech():
    return extended_h2()


def _conc(m, seed=0):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.001, 0.05, m.n_species)


def test_types_compile_and_round_trip(mech):
    a = mech.arrays
    for r, (t, x) in RECAST.items():
        assert a["type"][r] == t
        if x is not None:
            assert np.array_equal(a["extra"][r, :len(x)], x)
    # the written dictionaries compile back to the same tables
    t, rx = mech.to_foam()
    again = Mechanism(t, rx)
    for k in ("type", "k", "kInf", "troe", "extra", "eff"):
        assert np.array_equal(again.arrays[k], a[k]), k


@pytest.mark.parametrize("T", [600.0, 1000.0, 1700.0, 2600.0])
def test_temperature_dependent_closed_forms(mech, T):
    o = Oracle(mech)
    c = _conc(mech)
    k = mech.arrays["k"]
    for r, t in ((0, "LT"), (2, "JAN"), (10, "PS")):
        A, b, Ta = k[r]
        if t == "LT":
            want = A * T ** b * np.exp(-Ta / T + LT[0] * T ** (-1 / 3) + LT[1] * T ** (-2 / 3))
        elif t == "JAN":
            want = A * T ** b * np.exp(-Ta / T) * np.exp(sum(bn * np.log(T) ** n
                                                             for n, bn in enumerate(JAN)))
        else:
            want = A * T ** b * np.exp(-Ta / T + sum(cn / T ** (n + 1)
                                                     for n, cn in enumerate(PS)))
        kf = o.rates(r, 1e5, T, c)[0]
        assert np.isclose(kf, want, rtol=1e-13, atol=0), (t, kf, want)


def test_pressure_dependent_closed_forms_and_limits(mech):
    o = Oracle(mech)
    a = mech.arrays
    T = 1300.0
    ar = lambda p: p[0] * T ** p[1] * np.exp(-p[2] / T)  # noqa: E731
    c = _conc(mech, 3)
    for r in (3, 4, 5, 22):
        k0, kInf = ar(a["k"][r]), ar(a["kInf"][r])
        M = float(a["eff"][r] @ c)
        Pr = k0 / kInf * M
        t = a["type"][r]
        if t in (_capi.RATE_SRI, _capi.RATE_SRI_CA):
            X = 1.0 / (1.0 + np.log10(max(Pr, np.finfo(float).eps)) ** 2)
            F = SRI[3] * (SRI[0] * np.exp(-SRI[1] / T) + np.exp(-T / SRI[2])) ** X * T ** SRI[4]
        elif t == _capi.RATE_TROE_CA:
            al, T3, T1, T2 = a["troe"][r]
            Fc = (1 - al) * np.exp(-T / T3) + al * np.exp(-T / T1) + np.exp(-T2 / T)
            lF = np.log10(Fc)
            cc, nn = -0.4 - 0.67 * lF, 0.75 - 1.27 * lF
            x = (np.log10(Pr) + cc) / (nn - 0.14 * (np.log10(Pr) + cc))
            F = 10 ** (lF / (1 + x * x))
        else:
            F = 1.0
        want = kInf * Pr / (1 + Pr) * F if t == _capi.RATE_SRI else k0 / (1 + Pr) * F
        assert np.isclose(o.rates(r, 1e5, T, c)[0], want, rtol=1e-12), r
    # chemically activated Lindemann: k -> k0 as M -> 0, k -> kInf/M as M -> inf
    r = 4
    k0, kInf = ar(a["k"][r]), ar(a["kInf"][r])
    assert o.rates(r, 1e5, T, np.zeros(mech.n_species))[0] == pytest.approx(k0, rel=1e-14)
    big = np.full(mech.n_species, 1e12)
    M = float(a["eff"][r] @ big)
    assert o.rates(r, 1e5, T, big)[0] == pytest.approx(kInf / M, rel=1e-6)


@pytest.mark.parametrize("r", [0, 2, 10])
def test_temperature_derivative_matches_finite_differences(mech, r):
    """dω/dT of the Landau-Teller, Janev and power-series reactions (their
    exact dk/dT) against central differences of ω."""
    o = Oracle(mech)
    c = _conc(mech, 5)
    for T in (800.0, 1500.0, 2400.0):
        h = T * 1e-6
        wp = o.parameters(r, 1e5, T + h, c)[0][0]
        wm = o.parameters(r, 1e5, T - h, c)[0][0]
        d = o.parameters(r, 1e5, T, c)[0][1]
        assert np.isclose(d, (wp - wm) / (2 * h), rtol=1e-6), (r, T)


def test_rhs_conserves_mass_and_the_solve_runs(mech):
    o = Oracle(mech)
    m0, Y0, dt0, c = make_states("h2_small", 16)
    nS = mech.n_species
    for cell in range(4):
        d = o.derivatives(Y0[:, cell])
        assert abs(d[:nS].sum()) <= 1e-12 * np.abs(d[:nS]).sum()
    ode = ode_inputs("Rosenbrock34", relTol=1e-6, absTol=1e-10)
    Yn, dt, _, st, _ = o.solve_cells(ode, Y0, 1e-5, dt0)
    assert (st[0] == 0).all() and np.all(np.isfinite(Yn))
    assert np.allclose(Yn[:nS].sum(axis=0), 1.0, atol=1e-10)


def test_chemkin_keywords_for_the_new_types():
    body = """2H2+O2(+M)=2H2O(+M)  3.0E+12 0.0 0.0
   LOW / 4.0E+18 -1.0 0.0 /
   SRI / 0.6 250.0 900.0 1.2 0.04 /
2H2+O2(+M)=2H2O(+M)  5.0E+17 0.0 0.0
   HIGH / 2.0E+12 0.0 0.0 /
2H2+O2=2H2O  1.0E+10 0.5 1000.0
   LT / 7.0 -30.0 /
2H2+O2=2H2O  1.0E+10 0.5 1000.0
   JAN / 0.3 -0.05 0.004 1e-4 -2e-5 1e-6 0 0 0 /
2H2+O2=2H2O  1.0E+10 0.5 1000.0
   FIT1 / 150.0 -4.0E4 1.0E6 -5.0E8 /"""
    import test_chemkin as TC
    thermo, reactions = chemkin_to_foam(TC.CHEM.format(body=body), TC._therm())
    m = Mechanism(thermo, reactions)
    a = m.arrays
    assert list(a["type"]) == [_capi.RATE_SRI, _capi.RATE_LINDEMANN_CA, _capi.RATE_LANDAU_TELLER,
                               _capi.RATE_JANEV, _capi.RATE_POWER_SERIES]
    assert list(a["extra"][0, :5]) == SRI
    assert a["k"][1, 0] == 5.0e17 / 1e9 and a["kInf"][1, 0] == 2.0e12 / 1e6  # HIGH = kInf
    assert list(a["extra"][2, :2]) == LT and list(a["extra"][4, :4]) == PS
    with pytest.raises(GfcError) as e:
        chemkin_to_foam(TC.CHEM.format(body="2H2+O2=2H2O  1.0 0.0 0.0\n   LT / 1.0 /"),
                        TC._therm())
    assert e.value.status == 2
