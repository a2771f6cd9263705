"""GPU tests of the discrete adjoint with four cells per wavefront (16-lane
teams, nSpecie+1 <= 16), selected by set_batched_shape(4) and by default
from 16384 cells on.

- The recording solve is the four-cells-per-wavefront batched solve, bit for
  bit (the same kernel code with the recorder compiled in).
- Its gradients agree with the one-lane-per-cell adjoint, itself checked
  against the host build and the oracle.  The two differ in summation order
  (team sums are trees, the third-body part of the Jacobian is summed by
  efficiency set), so the comparison is relative to each cell's largest
  gradient entry: 1e-7 for the initial-state gradients and the
  rate-parameter gradients, on cells whose accepted step counts agree.
- They match central finite differences of the replayed accepted steps
  (relative step 1e-6) to 1e-5, as the one-lane form does.
- The checkpoint interval is all but invisible: with K = 9 the reverse
  re-integrates each segment from its checkpoint with the solve's step code
  compiled into the reverse kernel, and the states it gets differ from the
  recorded ones in the last bits (measured: gradients 3e-15 of the cell's
  largest entry apart from K = 1, every run the same), so K = 1 and K = 9
  are held to 1e-12 of each cell's largest gradient entry.  (The
  one-lane-per-cell form re-integrates with the very function its solve
  runs and is bit-identical.)
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import adjoint_host as AH  # noqa: E402
from gpufoam_amd import ode_inputs  # noqa: E402
from gpufoam_amd.evaluator import Evaluator  # noqa: E402
from gpufoam_amd.workloads import make_states  # noqa: E402

from test_gpu_adjoint import _smooth_cells  # noqa: E402

DEV = "cuda"


def _case(config, ncells, maxSteps=None):
    m, Y0, dt0, c = make_states(config, ncells)
    kw = {} if maxSteps is None else {"maxSteps": maxSteps}
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"], **kw)
    return m, Y0, dt0, c, ode


def _evaluator(m, ode, ncells, cells_per_wave, Treact=0.0):
    ev = Evaluator(m, ode, Treact, max_cells=ncells)
    ev.set_batched_shape(cells_per_wave)
    return ev


def _forward(ev, c, Y0, dt0):
    dY0 = torch.tensor(Y0, device=DEV)
    dYn = torch.empty_like(dY0)
    ddt = torch.tensor(dt0, device=DEV)
    st = torch.zeros((3, Y0.shape[1]), dtype=torch.int32, device=DEV)
    ev.adjoint_forward(c["deltaT"], dY0, ddt, Ynew=dYn, stats=st)
    return dY0, dYn, ddt, st


def _backward(ev, m, W):
    Ybar = torch.tensor(W, device=DEV)
    Y0bar = torch.empty_like(Ybar)
    pbar = torch.empty((6 * m.n_reactions, W.shape[1]), dtype=torch.float64, device=DEV)
    ev.adjoint_backward(Ybar, Y0bar, pbar)
    torch.cuda.synchronize()
    return Y0bar.cpu().numpy(), pbar.cpu().numpy()


@pytest.mark.parametrize("config,ncells", [("h2_small", 256), ("h2_adjoint", 64)])
def test_recording_solve_is_the_four_cell_solve(config, ncells):
    m, Y0, dt0, c, ode = _case(config, ncells)
    print(f"case {config} n={ncells} record vs batched, 4 cells per wavefront", flush=True)
    ev = _evaluator(m, ode, ncells, 4)
    _, dYn, ddt, st = _forward(ev, c, Y0, dt0)
    Yn, dt, _, stb, _ = ev.solve_host(c["deltaT"], Y0, dt0, batched=True)
    torch.cuda.synchronize()
    assert np.array_equal(dYn.cpu().numpy(), Yn)
    assert np.array_equal(ddt.cpu().numpy(), dt)
    assert np.array_equal(st.cpu().numpy(), stb)


def _team_against_lane(m, Y0, dt0, c, ode, ncells, label):
    n = m.n_species + 1
    W = np.random.default_rng(0).standard_normal((n, ncells))
    res = {}
    for cpw in (1, 4):
        print(f"case {label} n={ncells} adjoint, {cpw} cells per wavefront", flush=True)
        ev = _evaluator(m, ode, ncells, cpw)
        _, _, _, st = _forward(ev, c, Y0, dt0)
        res[cpw] = _backward(ev, m, W) + (st.cpu().numpy()[2],)
    (gl, pl, al), (gt, pt, at) = res[1], res[4]
    same = (al == at) & _smooth_cells(Y0, m.n_species)
    assert same.mean() >= 0.8, same.mean()
    sy = np.abs(gl).max(axis=0)
    sp = np.abs(pl).max(axis=0) + 1e-300
    ey = (np.abs(gt - gl).max(axis=0) / sy)[same]
    ep = (np.abs(pt - pl).max(axis=0) / sp)[same]
    print(f"compared {same.sum()} cells: state gradients {ey.max():.2e}, "
          f"rate-parameter gradients {ep.max():.2e} of the cell's largest entry")
    assert ey.max() <= 1e-7 and ep.max() <= 1e-7


@pytest.mark.parametrize("config,ncells", [("h2_small", 256), ("h2_adjoint", 64)])
def test_team_gradients_equal_lane_gradients(config, ncells):
    m, Y0, dt0, c, ode = _case(config, ncells)
    _team_against_lane(m, Y0, dt0, c, ode, ncells, config)


@pytest.mark.parametrize("solver", ["Rosenbrock12", "Rosenbrock23"])
def test_team_gradients_other_solvers(solver):
    """The two- and three-stage step reverses (h2_small states, deltaT 1e-5,
    relTol 1e-5)."""
    m, Y0, dt0, c = make_states("h2_small", 64)
    ode = ode_inputs(solver, relTol=1e-5, absTol=1e-10)
    _team_against_lane(m, Y0, dt0, dict(c, deltaT=1e-5), ode, 64, f"h2_small {solver}")


@pytest.mark.parametrize("config,ncells,maxSteps", [("h2_small", 48, None),
                                                     ("h2_adjoint", 16, 2000)])
def test_team_gradients_match_replayed_central_differences(config, ncells, maxSteps):
    m, Y0, dt0, c, ode = _case(config, ncells, maxSteps)
    nS, n, nR = m.n_species, m.n_species + 1, m.n_reactions
    print(f"case {config} n={ncells} finite differences, 4 cells per wavefront", flush=True)
    ev = _evaluator(m, ode, ncells, 4)
    theta0 = ev.get_rate_parameters()
    ev.set_rate_parameters(theta0)
    _forward(ev, c, Y0, dt0)
    rng = np.random.default_rng(1)
    W = rng.standard_normal((n, ncells))
    gY, gP = _backward(ev, m, W)
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
        print(f"trial {trial}: max relative error {rel.max():.2e}")
        assert rel.max() <= 1e-5, (trial, rel.max(), int(rel.argmax()))
    ev.set_rate_parameters(theta0)


def test_team_checkpoint_interval_changes_gradients_only_by_rounding():
    m, Y0, dt0, c, ode = _case("h2_small", 64)
    n = m.n_species + 1
    W = np.random.default_rng(2).standard_normal((n, 64))
    out = []
    for K in (1, 9):
        print(f"case h2_small n=64 checkpoint interval {K}, 4 cells per wavefront", flush=True)
        ev = _evaluator(m, ode, 64, 4)
        ev.set_checkpoint_interval(K)
        _forward(ev, c, Y0, dt0)
        assert ev.checkpoint_interval == K
        out.append(_backward(ev, m, W))
    for a, b in zip(out[0], out[1]):
        scale = np.abs(a).max(axis=0) + 1e-300
        err = (np.abs(a - b).max(axis=0) / scale).max()
        print(f"K = 1 against K = 9: {err:.2e} of the cell's largest entry")
        assert err <= 1e-12


def test_team_skipped_cells():
    m, Y0, dt0, c, ode = _case("h2_small", 32)
    n = m.n_species + 1
    print("case h2_small n=32 all cells below Treact, 4 cells per wavefront", flush=True)
    ev = _evaluator(m, ode, 32, 4, Treact=1e4)
    _forward(ev, c, Y0, dt0)
    W = np.random.default_rng(3).standard_normal((n, 32))
    g, pb = _backward(ev, m, W)
    assert np.array_equal(g, W) and not pb.any()
