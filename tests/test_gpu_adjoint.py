"""GPU tests of the discrete adjoint through the C ABI, one lane per cell
(the default for these H2 batch sizes and for n > 64; set_batched_shape(64)
for GRI, whose default is one wavefront per cell, tests/test_gpu_adjoint_wave.py):
the recording forward is the one-lane solve, the device gradients equal the
host build of the same adjoint code (itself checked against the oracle on
the CPU), and they match central finite differences of the replayed
accepted steps (relative step 1e-6), also through the test-only
torch.autograd.Function."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import adjoint_host as AH  # noqa: E402
from adjoint_autograd import ChemistryStep, replay  # noqa: E402
from gpufoam_amd import ode_inputs  # noqa: E402
from gpufoam_amd.evaluator import Evaluator  # noqa: E402
from gpufoam_amd.workloads import make_states  # noqa: E402

DEV = "cuda"


def _smooth_cells(Y0, nS):
    """Cells whose initial state has at most two species at exactly zero.
    max(Y, 0) inside the mixture has a kink at Y = 0: for unburnt GRI cells
    (fifty species at zero) roundoff-level stage values of those species
    decide which branch the derivative takes, and a 1e-14 change of T0 moves
    the gradient by O(1), so no gradient comparison is meaningful there."""
    return np.sum(Y0[:nS] == 0.0, axis=0) <= 2


def _case(config, ncells, maxSteps=None):
    m, Y0, dt0, c = make_states(config, ncells)
    kw = {} if maxSteps is None else {"maxSteps": maxSteps}
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"], **kw)
    return m, Y0, dt0, c, ode


def _lane_evaluator(m, ode, ncells):
    ev = Evaluator(m, ode, 0.0, max_cells=ncells)
    ev.set_batched_shape(64)  # one lane per cell for every mechanism
    return ev


def _forward(ev, c, Y0, dt0):
    dY0 = torch.tensor(Y0, device=DEV)
    dYn = torch.empty_like(dY0)
    ddt = torch.tensor(dt0, device=DEV)
    st = torch.zeros((3, Y0.shape[1]), dtype=torch.int32, device=DEV)
    ev.adjoint_forward(c["deltaT"], dY0, ddt, Ynew=dYn, stats=st)
    return dY0, dYn, ddt, st


@pytest.mark.parametrize("config", ["h2_small", "gri_production"])
def test_recording_forward_is_the_plain_solve(config):
    m, Y0, dt0, c, ode = _case(config, 96)
    ev = _lane_evaluator(m, ode, 96)
    _, dYn, ddt, st = _forward(ev, c, Y0, dt0)
    Yn, dt, _, stp, _ = ev.solve_host(c["deltaT"], Y0, dt0, lane=True)
    torch.cuda.synchronize()
    assert np.array_equal(dYn.cpu().numpy(), Yn)
    assert np.array_equal(ddt.cpu().numpy(), dt)
    assert np.array_equal(st.cpu().numpy(), stp)


@pytest.mark.parametrize("config,ncells", [("h2_small", 64), ("gri_production", 64)])
def test_device_gradients_equal_host_build(config, ncells):
    m, Y0, dt0, c, ode = _case(config, ncells)
    nS, n = m.n_species, m.n_species + 1
    ev = _lane_evaluator(m, ode, ncells)
    _, _, _, st = _forward(ev, c, Y0, dt0)
    rng = np.random.default_rng(0)
    W = rng.standard_normal((n, ncells))
    Ybar = torch.tensor(W, device=DEV)
    Y0bar = torch.empty_like(Ybar)
    pbar = torch.empty((6 * m.n_reactions, ncells), dtype=torch.float64, device=DEV)
    ev.adjoint_backward(Ybar, Y0bar, pbar)
    torch.cuda.synchronize()
    gY, gP, acc = Y0bar.cpu().numpy(), pbar.cpu().numpy(), st.cpu().numpy()[2]
    compared = 0
    smooth = _smooth_cells(Y0, nS)
    for cell in np.nonzero(smooth)[0]:
        _, hy, hth, _, N = AH.cell_adjoint(m, ode, Y0[:, cell], dt0[cell], c["deltaT"],
                                           W[:, cell])
        if N != acc[cell]:
            continue  # an ulp of exp/log moved a step decision
        compared += 1
        sy = np.abs(hy).max()
        assert np.allclose(gY[:, cell], hy, rtol=1e-8, atol=1e-10 * sy), cell
        sp = np.abs(hth).max() + 1e-300
        assert np.allclose(gP[:, cell], hth, rtol=1e-8, atol=1e-10 * sp), cell
    assert compared >= 0.9 * smooth.sum() and smooth.sum() >= ncells // 3


@pytest.mark.parametrize("config,ncells,maxSteps", [("h2_small", 48, None),
                                                     ("gri_production", 48, None),
                                                     ("h2_adjoint", 16, 2000)])
def test_gradients_match_replayed_central_differences(config, ncells, maxSteps):
    m, Y0, dt0, c, ode = _case(config, ncells, maxSteps)
    nS, n, nR = m.n_species, m.n_species + 1, m.n_reactions
    ev = _lane_evaluator(m, ode, ncells)
    theta0 = ev.get_rate_parameters()
    ev.set_rate_parameters(theta0)
    dY0, _, _, _ = _forward(ev, c, Y0, dt0)
    rng = np.random.default_rng(1)
    W = rng.standard_normal((n, ncells))
    Y0bar = torch.empty((n, ncells), dtype=torch.float64, device=DEV)
    pbar = torch.empty((6 * nR, ncells), dtype=torch.float64, device=DEV)
    ev.adjoint_backward(torch.tensor(W, device=DEV), Y0bar, pbar)
    gY, gP = Y0bar.cpu().numpy(), pbar.cpu().numpy()
    tscale = AH.theta_scale(m).reshape(nR, 6)
    eps = 1e-6
    for trial in range(2):
        D = rng.standard_normal((n, ncells)) * np.abs(Y0[:n])
        Dth = rng.standard_normal((nR, 6)) * tscale * (trial == 1)

        def L(sign):
            Yp = Y0.copy()
            Yp[:n] += sign * eps * D
            ev.set_rate_parameters(theta0 + sign * eps * Dth)
            dYp = torch.tensor(Yp, device=DEV)
            out = torch.empty_like(dYp)
            ev.adjoint_replay(dYp, out)
            torch.cuda.synchronize()
            return np.sum(W * out.cpu().numpy()[:n], axis=0)
        fd = (L(1) - L(-1)) / (2 * eps)
        ad = np.sum(gY * D, axis=0) + gP.T @ Dth.reshape(-1)
        rel = (np.abs(ad - fd) / np.abs(fd))[_smooth_cells(Y0, nS)]
        assert rel.max() <= 1e-5, (trial, rel.max(), int(rel.argmax()))
    ev.set_rate_parameters(theta0)


def test_syn200_gradients_match_replayed_central_differences():
    """The 200-species synthetic mechanism (n = 201, one lane per cell in
    the adjoint), four cells over a short CFD step (2e-9 s), gradients with
    respect to the initial state and every rate parameter against central
    differences of the replayed accepted steps (relative step 1e-6)."""
    m, Y0, dt0, c, ode = _case("syn200", 4)
    c = dict(c, deltaT=2e-9)
    nS, n, nR = m.n_species, m.n_species + 1, m.n_reactions
    print("case syn200 n=4 deltaT=2e-9 adjoint", flush=True)
    ev = Evaluator(m, ode, 0.0, max_cells=4)
    theta0 = ev.get_rate_parameters()
    _, _, _, st = _forward(ev, c, Y0, dt0)
    print(f"accepted steps {st.cpu().numpy()[2].tolist()}", flush=True)
    rng = np.random.default_rng(5)
    W = rng.standard_normal((n, 4))
    Y0bar = torch.empty((n, 4), dtype=torch.float64, device=DEV)
    pbar = torch.empty((6 * nR, 4), dtype=torch.float64, device=DEV)
    ev.adjoint_backward(torch.tensor(W, device=DEV), Y0bar, pbar)
    gY, gP = Y0bar.cpu().numpy(), pbar.cpu().numpy()
    tscale = AH.theta_scale(m).reshape(nR, 6)
    eps = 1e-6
    for trial in range(2):
        D = rng.standard_normal((n, 4)) * np.abs(Y0[:n])
        Dth = rng.standard_normal((nR, 6)) * tscale * (trial == 1)

        def L(sign):
            Yp = Y0.copy()
            Yp[:n] += sign * eps * D
            ev.set_rate_parameters(theta0 + sign * eps * Dth)
            dYp = torch.tensor(Yp, device=DEV)
            out = torch.empty_like(dYp)
            ev.adjoint_replay(dYp, out)
            torch.cuda.synchronize()
            return np.sum(W * out.cpu().numpy()[:n], axis=0)
        fd = (L(1) - L(-1)) / (2 * eps)
        ad = np.sum(gY * D, axis=0) + gP.T @ Dth.reshape(-1)
        rel = np.abs(ad - fd) / np.abs(fd)
        print(f"trial {trial}: max relative gradient error {rel.max():.3g}", flush=True)
        assert rel.max() <= 1e-5, (trial, rel.max(), int(rel.argmax()))
    ev.set_rate_parameters(theta0)


def test_checkpoint_interval_is_invisible_on_device():
    m, Y0, dt0, c, ode = _case("h2_small", 64)
    n = m.n_species + 1
    W = torch.tensor(np.random.default_rng(2).standard_normal((n, 64)), device=DEV)
    out = []
    for K in (1, 9):
        ev = Evaluator(m, ode, 0.0, max_cells=64)
        ev.set_checkpoint_interval(K)
        _forward(ev, c, Y0, dt0)
        assert ev.checkpoint_interval == K
        g = torch.empty_like(W)
        pb = torch.empty((6 * m.n_reactions, 64), dtype=torch.float64, device=DEV)
        ev.adjoint_backward(W, g, pb)
        out.append((g.cpu().numpy(), pb.cpu().numpy()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_skipped_cells_and_empty_launch():
    m, Y0, dt0, c, ode = _case("h2_small", 32)
    n = m.n_species + 1
    ev = Evaluator(m, ode, 1e4, max_cells=32)  # Treact above every cell
    _forward(ev, c, Y0, dt0)
    W = torch.tensor(np.random.default_rng(3).standard_normal((n, 32)), device=DEV)
    g = torch.empty_like(W)
    pb = torch.full((6 * m.n_reactions, 32), 7.0, dtype=torch.float64, device=DEV)
    ev.adjoint_backward(W, g, pb)
    torch.cuda.synchronize()
    assert torch.equal(g, W) and not pb.any()
    e = torch.empty((m.n_species + 2, 0), dtype=torch.float64, device=DEV)
    ev.adjoint_forward(c["deltaT"], e, torch.empty(0, dtype=torch.float64, device=DEV))
    ev.adjoint_backward(e[:n], e[:n])


def test_autograd_function_against_central_differences():
    """Loss L = sum over cells of (T + 1e3 Y_OH) after one step of the
    differentiation workload (H2, deltaT 1e-4), gradients with respect to the
    initial state and every rate parameter."""
    m, Y0, dt0, c, ode = _case("h2_adjoint", 16, 3000)
    nS, n = m.n_species, m.n_species + 1
    iOH = m.species.index("OH")
    ev = Evaluator(m, ode, 0.0, max_cells=16)
    theta0 = torch.tensor(ev.get_rate_parameters())
    state0 = torch.tensor(Y0[:n], device=DEV, requires_grad=True)
    theta = theta0.clone().requires_grad_(True)
    p = torch.tensor(Y0[nS + 1], device=DEV)
    dt = torch.tensor(dt0, device=DEV)
    wvec = torch.zeros(n, dtype=torch.float64, device=DEV)
    wvec[nS] = 1.0
    wvec[iOH] = 1e3
    out = ChemistryStep.apply(state0, theta, ev, c["deltaT"], dt, p)
    loss = (wvec[:, None] * out).sum()
    gs, gt = torch.autograd.grad(loss, [state0, theta])
    rng = np.random.default_rng(4)
    tscale = torch.tensor(AH.theta_scale(m).reshape(-1, 6))
    eps = 1e-6
    for trial in range(3):
        D = torch.tensor(rng.standard_normal((n, 16)) * np.abs(Y0[:n]), device=DEV)
        Dth = torch.tensor(rng.standard_normal(theta0.shape)) * tscale * (trial > 0)
        Lp = (wvec[:, None] * replay(ev, state0 + eps * D, theta0 + eps * Dth, p)).sum()
        Lm = (wvec[:, None] * replay(ev, state0 - eps * D, theta0 - eps * Dth, p)).sum()
        fd = (Lp - Lm).item() / (2 * eps)
        ad = (gs * D).sum().item() + (gt * Dth).sum().item()
        assert abs(ad - fd) <= 1e-5 * abs(fd), (trial, ad, fd)
    ev.set_rate_parameters(theta0.numpy())
