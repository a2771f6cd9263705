"""GPU tests of the discrete adjoint with one wavefront per cell (16 <
nSpecie+1 <= 64; the default for GRI-Mech 3.0, wave_adjoint.hip).

- The recording solve is the batched wavefront solve, bit for bit (the same
  kernel code with the recorder compiled in).
- Its gradients agree with the one-lane-per-cell adjoint (set_batched_shape(64)),
  itself checked against the host build and the oracle.  The two differ in
  summation order (wavefront sums are trees, reaction contributions are
  accumulated in LDS), so the comparison is relative to each cell's largest
  gradient entry: 1e-7 for the initial-state and the rate-parameter
  gradients, on smooth cells whose accepted step counts agree.
- They match central finite differences of the replayed accepted steps
  (relative step 1e-6) to 1e-5.
- The checkpoint interval changes them only by rounding (a re-integrated
  segment reaches the recorded states to the last bits).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import adjoint_host as AH  # noqa: E402
from gpufoam_amd import ode_inputs  # noqa: E402
from gpufoam_amd.evaluator import Evaluator  # noqa: E402
from gpufoam_amd.mechanism import Mechanism  # noqa: E402
from gpufoam_amd.synthetic import syn_states, syn_texts  # noqa: E402
from gpufoam_amd.workloads import make_states  # noqa: E402

from test_gpu_adjoint import _smooth_cells  # noqa: E402

DEV = "cuda"


def _case(config, ncells, maxSteps=None):
    if config in ("syn60s", "syn60m"):
        # syn60m: 90 reactions with 27 distinct third-body efficiency sets,
        # more than the 16 whose products with the Jb factors the reverse
        # precomputes (the per-specie path of D is taken)
        t, r = syn_texts(seed=13, nS=60, nR=45 if config == "syn60s" else 90)
        m = Mechanism(t, r, config)
        Y0, dt0 = syn_states(m, ncells, seed=14)
        # every specie present (the generator leaves most at zero, where
        # max(Y, 0) makes the map non-smooth), renormalised
        nS = m.n_species
        Y0[:nS] += 1e-5
        Y0[:nS] /= Y0[:nS].sum(axis=0)
        c = dict(solver="Rosenbrock34", relTol=1e-6, absTol=1e-10, deltaT=1e-6)
    else:
        m, Y0, dt0, c = make_states(config, ncells)
    kw = {} if maxSteps is None else {"maxSteps": maxSteps}
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"], **kw)
    return m, Y0, dt0, c, ode


def _evaluator(m, ode, ncells, cells_per_wave=None, Treact=0.0):
    ev = Evaluator(m, ode, Treact, max_cells=ncells)
    if cells_per_wave is not None:
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


@pytest.mark.parametrize("config,ncells",
                         [("gri_production", 256), ("syn60s", 32), ("syn60m", 32)])
def test_recording_solve_is_the_wavefront_solve(config, ncells):
    m, Y0, dt0, c, ode = _case(config, ncells)
    print(f"case {config} n={ncells} record vs batched, wavefront per cell", flush=True)
    ev = _evaluator(m, ode, ncells)
    _, dYn, ddt, st = _forward(ev, c, Y0, dt0)
    ev.set_batched_shape(1)
    Yn, dt, _, stb, _ = ev.solve_host(c["deltaT"], Y0, dt0, batched=True)
    torch.cuda.synchronize()
    assert np.array_equal(dYn.cpu().numpy(), Yn)
    assert np.array_equal(ddt.cpu().numpy(), dt)
    assert np.array_equal(st.cpu().numpy(), stb)


@pytest.mark.parametrize("config,ncells",
                         [("gri_production", 64), ("syn60s", 32), ("syn60m", 32)])
def test_wave_gradients_equal_lane_gradients(config, ncells):
    m, Y0, dt0, c, ode = _case(config, ncells)
    n = m.n_species + 1
    W = np.random.default_rng(0).standard_normal((n, ncells))
    res = {}
    for cpw in (None, 64):
        print(f"case {config} n={ncells} adjoint, shape {cpw or 'default'}", flush=True)
        ev = _evaluator(m, ode, ncells, cpw)
        _, _, _, st = _forward(ev, c, Y0, dt0)
        res[cpw] = _backward(ev, m, W) + (st.cpu().numpy()[2],)
    (gw, pw, aw), (gl, pl, al) = res[None], res[64]
    smooth = _smooth_cells(Y0, m.n_species)
    same = (al == aw) & smooth
    assert same.sum() >= 0.8 * smooth.sum() and smooth.sum() >= ncells // 3, (same.sum(),
                                                                             smooth.sum())
    sy = np.abs(gl).max(axis=0)
    sp = np.abs(pl).max(axis=0) + 1e-300
    ey = (np.abs(gw - gl).max(axis=0) / sy)[same]
    ep = (np.abs(pw - pl).max(axis=0) / sp)[same]
    print(f"compared {same.sum()} cells: state gradients {ey.max():.2e}, "
          f"rate-parameter gradients {ep.max():.2e} of the cell's largest entry")
    assert ey.max() <= 1e-7 and ep.max() <= 1e-7


def test_wave_gradients_match_replayed_central_differences():
    m, Y0, dt0, c, ode = _case("gri_production", 48)
    nS, n, nR, ncells = m.n_species, m.n_species + 1, m.n_reactions, 48
    print("case gri_production n=48 finite differences, wavefront per cell", flush=True)
    ev = _evaluator(m, ode, ncells)
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


def test_wave_checkpoint_interval_changes_gradients_only_by_rounding():
    m, Y0, dt0, c, ode = _case("gri_production", 32)
    n = m.n_species + 1
    W = np.random.default_rng(2).standard_normal((n, 32))
    out = []
    for K in (1, 5):
        print(f"case gri_production n=32 checkpoint interval {K}, wavefront per cell",
              flush=True)
        ev = _evaluator(m, ode, 32)
        ev.set_checkpoint_interval(K)
        _forward(ev, c, Y0, dt0)
        assert ev.checkpoint_interval == K
        out.append(_backward(ev, m, W))
    smooth = _smooth_cells(Y0, m.n_species)
    for a, b in zip(out[0], out[1]):
        scale = np.abs(a).max(axis=0) + 1e-300
        err = (np.abs(a - b).max(axis=0) / scale)[smooth].max()
        print(f"K = 1 against K = 5: {err:.2e} of the cell's largest entry")
        assert err <= 1e-10


def test_wave_skipped_cells():
    m, Y0, dt0, c, ode = _case("gri_production", 32)
    n = m.n_species + 1
    print("case gri_production n=32 all cells below Treact, wavefront per cell", flush=True)
    ev = _evaluator(m, ode, 32, Treact=1e4)
    _forward(ev, c, Y0, dt0)
    W = np.random.default_rng(3).standard_normal((n, 32))
    g, pb = _backward(ev, m, W)
    assert np.array_equal(g, W) and not pb.any()
