"""CPU tests of the discrete adjoint (the library's adjoint code built for
the host, tests/native): every piece against central finite differences of
the oracle, the whole-cell gradient against finite differences of the
replayed accepted-step sequence, and checkpointing invariance.

Finite differences are directional (random directions scaled to each
component's magnitude, rate parameters to their natural scales), with
relative step 1e-6: per-component differences of species at 1e-8 mass
fraction are dominated by round-off and pin nothing.  Species at exactly
zero are left out of central differences (max(Y, 0) has a kink there) and
checked one-sidedly instead."""
import ctypes as C

import numpy as np
import pytest

import adjoint_host as AH
from gpufoam_amd import ode_inputs
from gpufoam_amd.workloads import make_states
from oracle_lib import Oracle

pytestmark = pytest.mark.skipif(not AH.available(), reason="tests/native not built")

EPS = 1e-6


def _setup(config, cell, solver=None):
    m, Y0, dt0, c = make_states(config, 16)
    ode = ode_inputs(solver or c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    return m, Y0[:, cell].copy(), dt0[cell], c, ode


def _dir_check(L, gy, gth, y, tscale, rng, trials=3, tol=1e-7, with_theta=True):
    n = len(y)
    for t in range(trials):
        d = rng.standard_normal(n) * np.abs(y)  # zero species stay fixed
        dth = rng.standard_normal(len(gth)) * tscale * (1.0 if with_theta else 0.0)
        fd = (L(y + EPS * d, EPS * dth) - L(y - EPS * d, -EPS * dth)) / (2 * EPS)
        ad = gy @ d + gth @ dth
        assert abs(ad - fd) <= tol * abs(fd), (t, ad, fd, abs(ad - fd) / abs(fd))


def test_transposed_lu_solve():
    rng = np.random.default_rng(0)
    for n in (1, 5, 11, 54):
        A = rng.standard_normal((n, n)) + 2 * np.eye(n)
        A[0, :] *= 1e3  # force pivoting
        b = rng.standard_normal(n)
        x = b.copy()
        AH.lib().adjh_solve_T(n, AH.P(np.ascontiguousarray(A)), AH.P(x))
        assert np.allclose(A.T @ x, b, rtol=1e-10, atol=1e-10 * np.abs(b).max())


@pytest.mark.parametrize("config,cell", [("h2_small", 3), ("gri_production", 1),
                                         ("gri_production", 5)])
def test_rhs_vjp_matches_oracle_fd(config, cell):
    m, y, _, _, _ = _setup(config, cell)
    nS, n = m.n_species, m.n_species + 1
    p, yn = y[nS + 1], y[:n].copy()
    rng = np.random.default_rng(cell)
    fbar = rng.standard_normal(n)
    gy, gth = np.zeros(n), np.zeros(6 * m.n_reactions)
    AH.lib().adjh_vjp_f(C.pointer(m.view), p, AH.P(yn), AH.P(fbar), AH.P(gy), AH.P(gth))

    def L(yv, dth):
        o = Oracle(AH.PerturbedMech(m, dth))
        return fbar @ o.derivatives(np.concatenate([yv, [p]]))[:n]
    _dir_check(L, gy, gth, yn, AH.theta_scale(m), rng)


@pytest.mark.parametrize("config,cell", [("h2_small", 3), ("gri_production", 1),
                                         ("gri_production", 9)])
def test_jacobian_adjoint_matches_oracle_fd(config, cell):
    """Gradient of <Jbar, J> + <Fbar, F> for the reference's fast Jacobian
    (including its simplified fall-off terms), against the oracle's."""
    m, y, _, _, _ = _setup(config, cell)
    nS, n = m.n_species, m.n_species + 1
    p, yn = y[nS + 1], y[:n].copy()
    rng = np.random.default_rng(10 + cell)
    Jbar, Fbar = rng.standard_normal((n, n)), rng.standard_normal(n)
    gy, gth = np.zeros(n), np.zeros(6 * m.n_reactions)
    AH.lib().adjh_jact(C.pointer(m.view), p, AH.P(yn), AH.P(np.ascontiguousarray(Jbar)),
                       AH.P(Fbar), AH.P(gy), AH.P(gth))

    def L(yv, dth):
        F, J = Oracle(AH.PerturbedMech(m, dth)).jacobian(np.concatenate([yv, [p]]))
        return np.sum(Jbar * J[:n, :n]) + Fbar @ F[:n]
    _dir_check(L, gy, gth, yn, AH.theta_scale(m), rng)


@pytest.mark.parametrize("solver", ["Rosenbrock12", "Rosenbrock23", "Rosenbrock34"])
@pytest.mark.parametrize("config,cell,h", [("h2_small", 2, 1e-6), ("gri_production", 6, 1e-8)])
def test_step_reverse_matches_oracle_fd(solver, config, cell, h):
    m, y, _, _, ode = _setup(config, cell, solver)
    nS, n = m.n_species, m.n_species + 1
    p, yn = y[nS + 1], y[:n].copy()
    rng = np.random.default_rng(20 + cell)
    w = rng.standard_normal(n)
    gy, gth = w.copy(), np.zeros(6 * m.n_reactions)
    AH.lib().adjh_step_reverse(C.pointer(m.view), C.byref(ode), p, AH.P(yn), h, AH.P(gy),
                               AH.P(gth))

    def L(yv, dth):
        out, _ = Oracle(AH.PerturbedMech(m, dth)).step(ode, np.concatenate([yv, [p]]), h)
        return w @ out[:n]
    _dir_check(L, gy, gth, yn, AH.theta_scale(m), rng)


def test_zero_species_take_the_right_derivative():
    """d/dY at Y = 0 is the one-sided (Y >= 0) derivative of the step."""
    m, y, _, _, ode = _setup("h2_small", 3)
    nS, n = m.n_species, m.n_species + 1
    zero = [i for i in range(nS) if y[i] == 0.0]
    assert zero, "the H2 states have zero water and argon"
    p, yn = y[nS + 1], y[:n].copy()
    w = np.random.default_rng(5).standard_normal(n)
    gy, gth = w.copy(), np.zeros(6 * m.n_reactions)
    h = 1e-6
    AH.lib().adjh_step_reverse(C.pointer(m.view), C.byref(ode), p, AH.P(yn), h, AH.P(gy),
                               AH.P(gth))
    o = Oracle(m)

    def L(yv):
        return w @ o.step(ode, np.concatenate([yv, [p]]), h)[0][:n]
    for j in zero:
        e = np.zeros(n)
        e[j] = 1e-6
        fd = (-3 * L(yn) + 4 * L(yn + e) - L(yn + 2 * e)) / 2e-6
        assert abs(gy[j] - fd) <= 1e-6 * abs(fd), (j, gy[j], fd)


@pytest.mark.parametrize("config,cell", [("h2_small", 2), ("gri_production", 6),
                                         ("h2_adjoint", 1)])
def test_cell_gradient_matches_replayed_fd(config, cell):
    """dL/d(Y0, T0, theta) of one CFD step, L = w . final state, against
    central differences of the same accepted steps replayed."""
    m, y, dt0, c, ode = _setup(config, cell)
    if config == "h2_adjoint":
        ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"], maxSteps=400)
    nS, n = m.n_species, m.n_species + 1
    rng = np.random.default_rng(30 + cell)
    w = rng.standard_normal(n)
    yf, gy, gth, hs, N = AH.cell_adjoint(m, ode, y, dt0, c["deltaT"], w, K=1)
    assert N > 0

    def L(yv, dth):
        z = AH.replay(AH.PerturbedMech(m, dth), ode, np.concatenate([yv, [y[nS + 1]]]), hs)
        return w @ z[:n]
    assert np.array_equal(AH.replay(m, ode, y, hs)[:n], yf[:n])
    _dir_check(L, gy, gth, y[:n].copy(), AH.theta_scale(m), rng, tol=1e-6)


def test_replayed_map_is_the_oracle_integration():
    """The map the adjoint differentiates (recorded steps replayed) is the
    oracle's Rosenbrock integration with the same step sizes."""
    m, y, dt0, c, ode = _setup("gri_production", 6)
    nS, n = m.n_species, m.n_species + 1
    _, _, _, hs, N = AH.cell_adjoint(m, ode, y, dt0, c["deltaT"], np.zeros(n))
    o = Oracle(m)
    z = y.copy()
    for h in hs:
        z, _ = o.step(ode, z, h)
    z[:nS] = np.maximum(z[:nS], 0.0)
    r = AH.replay(m, ode, y, hs)
    assert np.allclose(r[:n], z[:n], rtol=1e-12, atol=1e-18)
    # and it is the solve itself: the oracle's full solve ends in the same state
    ref = o.solve_cells(ode, y[:, None], c["deltaT"], np.array([dt0]))[0][:, 0]
    assert np.allclose(r[:n], ref[:n], rtol=1e-12, atol=1e-18)


def test_checkpoint_interval_does_not_change_gradients():
    m, y, dt0, c, ode = _setup("h2_small", 2)
    n = m.n_species + 1
    w = np.random.default_rng(7).standard_normal(n)
    ref = AH.cell_adjoint(m, ode, y, dt0, c["deltaT"], w, K=1)
    for K in (2, 7, 64):
        got = AH.cell_adjoint(m, ode, y, dt0, c["deltaT"], w, K=K)
        assert np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2])


def test_skipped_cell_gradient_is_identity():
    m, y, dt0, c, ode = _setup("h2_small", 2)
    n = m.n_species + 1
    w = np.random.default_rng(8).standard_normal(n)
    _, gy, gth, _, N = AH.cell_adjoint(m, ode, y, dt0, c["deltaT"], w, Treact=1e4)
    assert N == 0 and np.array_equal(gy, w) and not gth.any()


@pytest.mark.parametrize("cells,nS,max_steps", [(65536, 10, 10000), (262144, 53, 10000),
                                                (32768, 200, 10000), (1000, 10, 1)])
def test_checkpoint_interval_sizes_the_whole_workspace(cells, nS, max_steps):
    """The automatic interval is the smallest K whose whole workspace
    (checkpoints, the K-state segment buffer, step sizes, per-cell scratch)
    fits the budget; the workspace is smallest near sqrt(maxSteps), which
    caps the search; nothing fitting returns that cap."""
    from gpufoam_amd import _capi
    L = _capi.lib()
    ws = lambda K: L.gfc_adjoint_workspace_bytes(cells, nS, max_steps, K)
    kcap = min(max_steps, int(np.ceil(np.sqrt(max_steps))))
    # the total falls while the checkpoints dominate, up to the cap
    sizes = [ws(K) for K in range(1, kcap + 1)]
    assert all(b <= a * (1 + 1e-12) for a, b in zip(sizes, sizes[1:]))
    assert ws(1) >= 8.0 * max_steps * (nS + 1) * cells  # checkpoints alone
    for frac in (1.5, 1.0, 0.6, 0.3):
        budget = frac * ws(1)
        K = L.gfc_adjoint_plan_interval(budget, cells, nS, max_steps)
        assert 1 <= K <= kcap
        if ws(K) <= budget:
            assert K == 1 or ws(K - 1) > budget
        else:  # nothing fits: the smallest workspace
            assert K == kcap and all(ws(k) > budget for k in range(1, kcap + 1))
    assert L.gfc_adjoint_plan_interval(1e300, cells, nS, max_steps) == 1
    assert L.gfc_adjoint_plan_interval(0.0, cells, nS, max_steps) == kcap
