"""The CPU oracle, pinned (CPU only).

1. Golden vectors of the reference's own code (tests/golden/
   reference_vectors.npz, written by tests/golden/make_golden.py from
   gpuFoam's headers built for the host): bit for bit.
2. The reference built for the host (oracle/_ref), where present: bit for
   bit on fresh inputs.
3. Mathematics the reference does not test: closed forms and invariants
   chosen so that a dropped term, a wrong sign or a transposed index fails.
"""
import os

import numpy as np
import pytest

from gpufoam_amd import Mechanism, ode_inputs
from gpufoam_amd.workloads import make_states
from oracle_lib import Oracle, Reference, lu_decompose, lu_solve, ref_available

GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz"))
P_INF = {"h2": 202650.0, "gri": 1.36789e6}
RR = 6.02214e26 * 1.38065e-23


def fill_linear(n):
    dx = 1.0 / n
    return np.array([dx * i + dx for i in range(n)])


@pytest.fixture(scope="module", params=["h2", "gri"])
def mech(request):
    return request.param, Mechanism.load(request.param)


# ------------------------------------------------------------ golden ----

def test_golden_thermo_kc_rates(mech):
    name, m = mech
    O = Oracle(m)
    p = P_INF[name]
    for T in (1000.0, 2500.0):
        got = np.array([O.species_thermo(i, p, T) for i in range(m.n_species)])
        assert np.array_equal(got, GOLDEN[f"{name}_thermo_T{int(T)}"])
        got = np.array([O.Kc(r, p, T) for r in range(m.n_reactions)])
        assert np.array_equal(got, GOLDEN[f"{name}_Kc_T{int(T)}"])
    c = fill_linear(m.n_species)
    got = np.array([O.rates(r, p, 1000.0, c) for r in range(m.n_reactions)])
    assert np.array_equal(got, GOLDEN[f"{name}_rates_T1000"])
    pr = [O.parameters(r, p, 1000.0, c) for r in range(m.n_reactions)]
    assert np.array_equal(np.array([x[0] for x in pr]), GOLDEN[f"{name}_params_omega_dwdT"])
    assert np.array_equal(np.array([x[1] for x in pr]), GOLDEN[f"{name}_params_dCf"])
    assert np.array_equal(np.array([x[2] for x in pr]), GOLDEN[f"{name}_params_dCr"])
    ddc = np.array([x[3] if m.arrays["type"][r] != 0 else np.zeros(m.n_species)
                    for r, x in enumerate(pr)])
    assert np.array_equal(ddc, GOLDEN[f"{name}_params_ddc"])


def test_golden_derivatives_jacobian(mech):
    name, m = mech
    O = Oracle(m)
    y = GOLDEN[f"{name}_y_linear"]
    assert np.array_equal(O.derivatives(y), GOLDEN[f"{name}_derivatives"])
    d, J = O.jacobian(y)
    assert np.array_equal(d, GOLDEN[f"{name}_jacobian_dfdx"])
    assert np.array_equal(J, GOLDEN[f"{name}_jacobian_J"])


@pytest.mark.parametrize("solver", ["Rosenbrock12", "Rosenbrock23", "Rosenbrock34"])
def test_golden_ode_solve(mech, solver):
    """Test-gpuChemistry.C:695-864 case: xEnd 1e-6, dxTry 1e-7, absTol
    1e-12, relTol 0.1, from mechanisms.H initial states."""
    name, m = mech
    O = Oracle(m)
    ode = ode_inputs(solver, absTol=1e-12, relTol=0.1)
    y, dx, st, _ = O.ode_solve(ode, 0.0, 1e-6, GOLDEN[f"{name}_ode_y0"], 1e-7)
    assert st == 0
    assert np.array_equal(y, GOLDEN[f"{name}_ode_{solver}_y"])
    assert dx == GOLDEN[f"{name}_ode_{solver}_dxTry"][0]


@pytest.mark.parametrize("cfg", ["h2_small", "gri_production"])
def test_golden_solve_cells(cfg):
    """singleCellSolver + computeRR on benchmark-law states."""
    m = Mechanism.load("h2" if cfg.startswith("h2") else "gri")
    from gpufoam_amd.workloads import CONFIGS
    c = CONFIGS[cfg]
    O = Oracle(m)
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    Yn, dt, dtmin, stats, rr = O.solve_cells(ode, GOLDEN[f"{cfg}_Y0"], c["deltaT"],
                                             GOLDEN[f"{cfg}_dt0"], rho=GOLDEN[f"{cfg}_rho"],
                                             deltaTChemMax=5e-7, want_rr=True)
    assert np.array_equal(Yn, GOLDEN[f"{cfg}_Ynew"])
    assert np.array_equal(dt, GOLDEN[f"{cfg}_dtChem"])
    assert dtmin == GOLDEN[f"{cfg}_dtMin"][0]
    assert np.array_equal(rr, GOLDEN[f"{cfg}_RR"])
    assert (stats[0] == 0).all() and (stats[2] <= stats[1]).all()


@pytest.mark.parametrize("n", [3, 10, 27, 49])
def test_golden_lu(n):
    lu, piv = lu_decompose(GOLDEN[f"lu{n}_A"])
    assert np.array_equal(lu, GOLDEN[f"lu{n}_LU"])
    assert np.array_equal(piv, GOLDEN[f"lu{n}_piv"])
    assert np.array_equal(lu_solve(lu, piv, np.ones(n)), GOLDEN[f"lu{n}_x"])


# ------------------------------------------------ the reference build ----

needs_ref = pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built")


@needs_ref
def test_ref_composite_thermo(mech):
    """The compiler's reaction thermo equals the one the reference's own
    operators build (gpuThermo.H:354-470)."""
    name, m = mech
    R = Reference(m)
    a = m.arrays
    for r in range(m.n_reactions):
        c = R.composite(r)
        assert c[0] == a["rY"][r] and c[1] == a["rW"][r] and c[2] == a["rTcommon"][r]
        assert np.array_equal(c[5:12], a["rhigh"][r]) and np.array_equal(c[12:19], a["rlow"][r])


@needs_ref
def test_ref_random_states(mech):
    """Bitwise agreement with the reference on random states (T 300-3000 K,
    random composition, including zero and negative mass fractions)."""
    name, m = mech
    O, R = Oracle(m), Reference(m)
    rng = np.random.default_rng(11)
    nS = m.n_species
    for _ in range(5):
        y = np.zeros(nS + 2)
        y[:nS] = rng.random(nS) * (rng.random(nS) < 0.7)
        y[:nS] /= y[:nS].sum()
        y[rng.integers(nS)] = -1e-9
        y[nS], y[nS + 1] = rng.uniform(150, 6000), rng.uniform(1e4, 5e6)
        assert np.array_equal(O.derivatives(y), R.derivatives(y))
        a1, J1 = O.jacobian(y)
        a2, J2 = R.jacobian(y)
        assert np.array_equal(a1, a2) and np.array_equal(J1, J2)


@needs_ref
@pytest.mark.parametrize("cfg,n", [("h2_small", 48), ("gri_production", 24), ("gri_batching", 8)])
def test_ref_solve_cells(cfg, n):
    m, Y0, dt0, c = make_states(cfg, n)
    O, R = Oracle(m), Reference(m)
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    a = O.solve_cells(ode, Y0, c["deltaT"], dt0, threads=4)
    b = R.solve_cells(ode, Y0, c["deltaT"], dt0, threads=4)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]


# ------------------------------------------------------- mathematics ----

def test_thermo_identities(mech):
    """dHa/dT = Cp, dCp/dT = dCpdT, dS/dT = Cp/T (central differences),
    G = H - T S at p = pStd, Hf = Ha(Tstd) of the low polynomial."""
    name, m = mech
    O = Oracle(m)
    p = 1e5
    for i in range(m.n_species):
        for T in (500.0, 1500.0):
            h = 1e-3
            Cp, Ha, Hf, S, G, dCp = O.species_thermo(i, p, T)
            up, dn = O.species_thermo(i, p, T + h), O.species_thermo(i, p, T - h)
            assert np.isclose((up[1] - dn[1]) / (2 * h), Cp, rtol=1e-7)
            assert np.isclose((up[0] - dn[0]) / (2 * h), dCp, rtol=1e-5, atol=1e-9 * abs(Cp))
            assert np.isclose((up[3] - dn[3]) / (2 * h), Cp / T, rtol=1e-6)
            assert np.isclose(G, Ha - T * S, rtol=1e-10, atol=1e-6 * abs(Ha))
        lo = O.species_thermo(i, p, 298.15)
        assert np.isclose(lo[2], lo[1], rtol=1e-12)


def test_kc_detailed_balance(mech):
    """ln Kc = -sum(nu g/RT) + dn ln(pStd/RT) from specie Gibbs energies,
    and dKcdTbyKc = d(ln Kc)/dT."""
    name, m = mech
    O = Oracle(m)
    a = m.arrays
    for T in (900.0, 1800.0):
        g = np.array([O.species_thermo(i, 1e5, T)[4] for i in range(m.n_species)])
        for r in range(m.n_reactions):
            dG, dn = 0.0, 0.0
            for side, sgn in (("lhs", -1), ("rhs", 1)):
                for q in range(a[f"n_{side}"][r]):
                    i = a[f"{side}_index"][r, q]
                    nu = a[f"{side}_stoich"][r, q]
                    dG += sgn * nu * a["W"][i] * g[i]
                    dn += sgn * nu
            lnKc = -dG / (RR * T) + dn * np.log(1e5 / (RR * T))
            Kc, dK = O.Kc(r, 1e5, T)
            if abs(lnKc) < 600:
                assert np.isclose(np.log(Kc), lnKc, rtol=1e-9, atol=1e-8)
            h = 1e-2
            fd = (np.log(O.Kc(r, 1e5, T + h)[0]) - np.log(O.Kc(r, 1e5, T - h)[0])) / (2 * h)
            assert np.isclose(dK, fd, rtol=1e-5, atol=1e-9)


def test_rate_closed_forms():
    """Arrhenius, third-body and Lindemann/Troe rates against textbook
    formulas on the small mechanism of test_mechanism.py."""
    from test_mechanism import R_OK, THERMO
    m = Mechanism(THERMO, R_OK)
    O = Oracle(m)
    c = np.array([0.3, 0.2, 0.1])
    T = 1200.0
    kf = O.rates(0, 1e5, T, c)[0]
    assert np.isclose(kf, 1e10 * T ** 0.5 * np.exp(-1000 / T), rtol=1e-14)
    M = 2.5 * 0.3 + 1 * 0.2 + 12 * 0.1
    assert np.isclose(O.rates(2, 1e5, T, c)[0], M * 2 * np.exp(-10 / T), rtol=1e-14)
    # Troe with k0=1, kInf=2, default efficiency 1.5
    Pr = 1.0 / 2.0 * 1.5 * c.sum()
    Fc = 0.5 * np.exp(-T / 100) + 0.5 * np.exp(-T / 1000) + np.exp(-5000 / T)
    lf = np.log10(Fc)
    cc, nn = -0.4 - 0.67 * lf, 0.75 - 1.27 * lf
    x = (np.log10(Pr) + cc) / (nn - 0.14 * (np.log10(Pr) + cc))
    F = 10 ** (lf / (1 + x * x))
    assert np.isclose(O.rates(3, 1e5, T, c)[0], 2 * Pr / (1 + Pr) * F, rtol=1e-13)
    # Lindemann: F = 1, M with default efficiency 1
    Pr = 0.5 * c.sum()
    assert np.isclose(O.rates(4, 1e5, T, c)[0], 2 * Pr / (1 + Pr), rtol=1e-14)
    # irreversible: kr = 0 ; reversible kr = kf/Kc
    assert O.rates(1, 1e5, T, c)[1] == 0.0
    r0 = O.rates(0, 1e5, T, c)
    assert np.isclose(r0[1], r0[0] / O.Kc(0, 1e5, T)[0], rtol=1e-14)
    # omega = kf*Cf - kr*Cr with the 1.5 real exponent
    om = O.rates(4, 1e5, T, c)
    k4 = O.rates(4, 1e5, T, c)
    Cf = c[1] * c[0] ** 1.5
    Cr = c[2] ** 2
    assert np.isclose(om[2], k4[0] * Cf - k4[1] * Cr, rtol=1e-13)


def test_rhs_conservation(mech):
    """Mass: sum dY/dt = 0.  Energy (constant p): sum Ha dY/dt + Cp dT/dt = 0."""
    name, m = mech
    O = Oracle(m)
    _, Y0, _, _ = make_states("h2_small" if name == "h2" else "gri_production", 8)
    nS = m.n_species
    for k in range(Y0.shape[1]):
        y = Y0[:, k].copy()
        d = O.derivatives(y)
        scale = np.abs(d[:nS]).max()
        assert abs(d[:nS].sum()) <= 1e-12 * scale + 1e-300
        th = np.array([O.species_thermo(i, y[nS + 1], y[nS]) for i in range(nS)])
        CpM = (np.maximum(y[:nS], 0) * th[:, 0]).sum()
        assert abs((th[:, 1] * d[:nS]).sum() + CpM * d[nS]) <= 1e-10 * abs(th[:, 1] * d[:nS]).sum()


def test_reaction_derivatives_match_finite_differences(mech):
    """d(omega)/dc_s = sum of dCf over lhs positions of s - sum of dCr over
    rhs positions of s + ddc_s, and d(omega)/dT = dwdT, against central
    differences of omega (computeReactionParameters, gpuReaction.H:650-702).
    Fall-off reactions are excluded: the reference drops their dPr/dT,
    dF/dT and dM/dc terms on purpose (gpuFallOffReactionRate.H:62-158)."""
    name, m = mech
    O = Oracle(m)
    a = m.arrays
    nS = m.n_species
    rng = np.random.default_rng(3)
    c = rng.uniform(0.05, 1.0, nS)
    p, T = 1e5, 1400.0
    for r in range(m.n_reactions):
        if a["type"][r] > 1:
            continue
        (w, dwdT), dCf, dCr, ddc = O.parameters(r, p, T, c)
        grad = ddc.copy() if a["type"][r] == 1 else np.zeros(nS)
        for q in range(a["n_lhs"][r]):
            grad[a["lhs_index"][r, q]] += dCf[q]
        for q in range(a["n_rhs"][r]):
            grad[a["rhs_index"][r, q]] -= dCr[q]
        fd = np.zeros(nS)
        for s_ in range(nS):
            h = 1e-6 * c[s_]
            cp, cm = c.copy(), c.copy()
            cp[s_] += h
            cm[s_] -= h
            fd[s_] = (O.parameters(r, p, T, cp)[0][0] - O.parameters(r, p, T, cm)[0][0]) / (2 * h)
        scale = np.abs(fd).max() + np.abs(w) * 1e-12
        assert np.abs(grad - fd).max() <= 1e-6 * scale, (r, grad, fd)
        hT = 1e-3
        fdT = (O.parameters(r, p, T + hT, c)[0][0] - O.parameters(r, p, T - hT, c)[0][0]) / (2 * hT)
        assert np.isclose(dwdT, fdT, rtol=1e-6, atol=1e-9 * abs(w) + 1e-300), (r, dwdT, fdT)


def test_jacobian_transforms_hold(mech):
    """The reference's 'fast' Jacobian in mass-fraction form: the T row is
    -(sum_j Ha_j J_ji + Cp_i dT/dt)/CpM (energy conservation of the
    linearisation), and the species block columns conserve mass up to the
    density term: sum_i J_ij = rhoM v_j sum_i dY_i/dt = 0."""
    name, m = mech
    O = Oracle(m)
    nS = m.n_species
    _, Y0, _, _ = make_states("h2_small" if name == "h2" else "gri_production", 3, mech=m)
    for k in range(3):
        y = Y0[:, k].copy()
        d, J = O.jacobian(y)
        th = np.array([O.species_thermo(i, y[nS + 1], y[nS]) for i in range(nS)])
        CpM = (np.maximum(y[:nS], 0) * th[:, 0]).sum()
        Trow = -(th[:, 1] @ J[:nS, :nS] + th[:, 0] * d[nS]) / CpM
        assert np.allclose(J[nS, :nS], Trow, rtol=1e-9, atol=1e-9 * np.abs(Trow).max())
        colsum = J[:nS, :nS].sum(axis=0)
        assert np.abs(colsum).max() <= 1e-9 * np.abs(J[:nS, :nS]).max()
        assert (J[nS + 1] == 0).all() and (J[:, nS + 1] == 0).all()


@pytest.mark.parametrize("n", [1, 2, 5, 17, 54])
def test_lu_solves_and_pivots(n):
    """The solution agrees with numpy; the pivot order is the implicit-
    scaling rule (largest |a|/max|row|, last row wins ties)."""
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) * np.exp(rng.uniform(-6, 6, (n, 1)))
    b = rng.standard_normal(n)
    lu, piv = lu_decompose(A)
    x = lu_solve(lu, piv, b)
    assert np.allclose(A @ x, b, rtol=1e-9, atol=1e-9 * np.abs(b).max())
    # first pivot: argmax of |a_i0|/max_j|a_ij|, ties to the last index
    s = np.abs(A[:, 0]) / np.abs(A).max(axis=1)
    assert piv[0] == np.flatnonzero(s == s.max())[-1]
    # tie: two identical candidate rows -> the later row is chosen
    B = np.eye(3)
    B[0, 0] = B[2, 0] = 1.0
    B[2, 2] = 1.0
    lu2, piv2 = lu_decompose(B)
    assert piv2[0] == 2


@pytest.mark.parametrize("solver,order", [("Rosenbrock12", 2), ("Rosenbrock23", 3),
                                          ("Rosenbrock34", 4)])
def test_rosenbrock_order(solver, order):
    """One step of size h has a local error O(h^(p+1)): halving h divides
    the deviation from a tight reference solution by about 2^(p+1)
    (a mistyped tableau coefficient breaks this).  Species components of an
    H2-air state at 1.5e-8..4e-9 s, where roundoff is well below the error."""
    m = Mechanism.load("h2")
    O = Oracle(m)
    nS = m.n_species
    _, Y0, _, _ = make_states("h2_small", 3, mech=m)
    y0 = Y0[:, 2].copy()
    ode = ode_inputs(solver, absTol=1e-30, relTol=1e-30)
    tight = ode_inputs("Rosenbrock34", absTol=1e-20, relTol=1e-14, maxSteps=100000)
    errs = []
    for h in (1.6e-8, 8e-9, 4e-9):
        y, _ = O.step(ode, y0, h)
        yr, _, st, _ = O.ode_solve(tight, 0.0, h, y0, h / 256)
        assert st == 0
        errs.append(np.abs(y[:nS] - yr[:nS]).max())
    for a, b in zip(errs[:-1], errs[1:]):
        assert 0.6 * 2 ** (order + 1) < a / b < 1.5 * 2 ** (order + 1), errs


def test_counted_work_constants_are_reproducible():
    """gpufoam_amd/flops/<mech>.json (operation counts of the oracle) match
    a fresh count, and the LU / solve counts match their closed forms."""
    import json
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    if not os.path.exists(os.path.join(root, "oracle", "libopcount.so")):
        pytest.skip("oracle/libopcount.so not built")
    committed = {}
    for mech in ("h2", "gri", "syn200"):
        with open(os.path.join(root, "gpufoam_amd", "flops", f"{mech}.json")) as f:
            committed[mech] = json.load(f)
    for mech, d in committed.items():
        n = d["n"]
        # Crout LU: n(n-1)(2n-1)/6 * 2 products and sums, n divisions of the
        # row scale, n(n-1)/2 column scalings
        assert abs(d["lu"] - 2 * n ** 3 / 3) <= 1.5 * n * n + 10 * n
        assert abs(d["solve"] - 2 * n * n) <= 2 * n
        assert d["attempt"] == d["F_jac"] + d["lu"] + 4 * d["solve"] + 2 * d["F_rhs"] + 40 * n
    out = subprocess.run([sys.executable, os.path.join(root, "oracle", "count_flops.py")],
                         capture_output=True, text=True, check=True).stdout
    for mech, d in committed.items():
        with open(os.path.join(root, "gpufoam_amd", "flops", f"{mech}.json")) as f:
            assert json.load(f) == d, mech
    assert "syn200" in out


def test_step_underflow_stops_the_cell():
    """absTol = relTol = 0 makes every normalised error infinite, so
    adaptiveSolve keeps shrinking the step by minScale
    (gpuAdaptiveSolver.H:99-114) until dx < vSmall, where the reference
    prints and asserts.  The oracle reports GFC_CELL_STEP_UNDERFLOW instead,
    after exactly the attempts that take the first step below DBL_MIN, with
    no step accepted and deltaTChem left at its initial value."""
    m, Y0, dt0, c = make_states("h2_small", 16)
    ode = ode_inputs(c["solver"], relTol=0.0, absTol=0.0)
    Yn, dt, _, st, _ = Oracle(m).solve_cells(ode, Y0, c["deltaT"], dt0)
    # attempts: dx0 * minScale^k first below DBL_MIN (binary64 products)
    dx, k = float(dt0[0]), 0
    while True:
        k += 1
        dx *= ode.minScale
        if dx < np.finfo(np.float64).tiny:
            break
    assert (st[0] == 2).all()
    assert (st[1] == k).all() and (st[2] == 0).all()
    nS = m.n_species
    assert np.array_equal(Yn[:nS], np.maximum(Y0[:nS], 0.0))
    assert np.array_equal(Yn[nS:], Y0[nS:])
    assert np.array_equal(dt, dt0)
