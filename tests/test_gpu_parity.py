"""Parity of the HIP kernels with the CPU oracle (MI355X).

Three device forms are compared with the oracle on the same seeded states:
  * lane     one lane per cell (GFC_LANE_PER_CELL), the reference's mapping
             and operation order: identical step counts to the oracle;
  * plain    the drop-in launch (no GFC_BATCHED): the wavefront, 16-lane-team
             or workgroup kernel over one grid of cells;
  * batched  the same kernels over a persistent work-stealing grid; its
             results equal the plain form's bit for bit.
Bounds:
  species  |Y_gpu - Y_cpu| <= 1e-6 |Y_cpu| + 1e-10      (rtol 1e-6, absTol)
  T        |T_gpu - T_cpu| <= 1e-8 T_cpu
  RR       from the same Ynew, with rho
  deltaTChem and the step counts: equal wherever the accept/reject sequence
  is the same (ulp-level differences of exp/log and of summation order can
  move a step by one attempt; such cells are reported and bounded)

The oracle is bit-identical to gpuFoam's own code built for the host
(tests/test_oracle.py)."""
import numpy as np
import pytest

from gpufoam_amd import GfcError, ode_inputs
from gpufoam_amd.workloads import make_states

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

RTOL_Y, ATOL_Y, RTOL_T = 1e-6, 1e-10, 1e-8


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gpufoam_amd import _capi
    _capi.lib()  # the HIP library must be present: no fallback


def _oracle_run(cfg, n, **kw):
    from oracle_lib import Oracle
    m, Y0, dt0, c = make_states(cfg, n)
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    rho = np.linspace(0.3, 1.3, n)
    ref = Oracle(m).solve_cells(ode, Y0, c["deltaT"], dt0, rho=rho, want_rr=True,
                                deltaTChemMax=kw.get("deltaTChemMax", 1e300))
    return m, Y0, dt0, c, ode, rho, ref


def _device_run(m, Y0, dt0, c, ode, rho, batched, graph=False, deltaTChemMax=1e300,
                Treact=0.0, cells_per_wave=None, lane=False, factorization=None):
    from gpufoam_amd.evaluator import Evaluator
    n = Y0.shape[1]
    ev = Evaluator(m, ode, Treact, max_cells=n)
    if cells_per_wave is not None:
        ev.set_batched_shape(cells_per_wave)
    if factorization is not None:
        ev.set_factorization(factorization)
    dY0 = torch.tensor(Y0, device="cuda")
    dYn = torch.empty_like(dY0)
    dRR = torch.zeros((m.n_species, n), dtype=torch.float64, device="cuda")
    ddt = torch.tensor(dt0, device="cuda")
    dmin = torch.zeros(1, dtype=torch.float64, device="cuda")
    st = torch.zeros((3, n), dtype=torch.int32, device="cuda")
    drho = torch.tensor(rho, device="cuda")
    ev.solve(c["deltaT"], dY0, ddt, Ynew=dYn, RR=dRR, rho=drho, deltaTMin=dmin, stats=st,
             batched=batched, graph=graph, deltaTChemMax=deltaTChemMax, lane=lane)
    torch.cuda.synchronize()
    return (dYn.cpu().numpy(), ddt.cpu().numpy(), dmin.item(), st.cpu().numpy(),
            dRR.cpu().numpy())


def _check(m, got, ref, tight_dt, Y0=None, rho=None, deltaT=None, dt_rtol=1e-6,
           min_same=None):
    Yg, dtg, dming, stg, rrg = got
    Yo, dto, dmino, sto, rro = ref
    nS = m.n_species
    errY = np.abs(Yg[:nS] - Yo[:nS]) / (RTOL_Y * np.abs(Yo[:nS]) + ATOL_Y)
    errT = np.abs(Yg[nS] - Yo[nS]) / Yo[nS]
    print(f"species error / tolerance max {errY.max():.3g}; T rel max {errT.max():.3g}")
    assert errY.max() <= 1.0
    assert errT.max() <= RTOL_T
    assert np.array_equal(Yg[nS + 1], Yo[nS + 1])  # pressure untouched
    assert (stg[0] == sto[0]).all()                 # status bits
    same = stg[1] == sto[1]
    frac = same.mean()
    print(f"identical step sequence in {frac:.3f} of cells")
    assert frac >= (min_same if min_same is not None else (1.0 if tight_dt else 0.8))
    assert np.allclose(dtg[same], dto[same], rtol=dt_rtol, atol=0)
    if tight_dt:
        assert np.isclose(dming, dmino, rtol=dt_rtol)
    if Y0 is not None:
        # RR = rho (Ynew - Y0) / deltaT of the kernel's own Ynew
        # (gpuKernelEvaluator.cu:84-89), and within the state tolerance of
        # the oracle's RR
        want = rho * (Yg[:nS] - Y0[:nS]) / deltaT
        assert np.allclose(rrg, want, rtol=1e-12, atol=1e-300)
        bound = rho * (RTOL_Y * np.abs(Yo[:nS]) + ATOL_Y) / deltaT
        assert np.all(np.abs(rrg - rro) <= 1.01 * bound)


# lane kernel: one wavefront holds 64 cells, so its samples stay small
LANE_CASES = [("h2_small", 384), ("gri_production", 192), ("gri_batching", 32),
              ("h2_adjoint", 48)]
CASES = [("h2_small", 384), ("gri_production", 2048), ("gri_batching", 512),
         ("h2_adjoint", 48)]


@pytest.mark.parametrize("cfg,n", LANE_CASES)
def test_lane_kernel_matches_oracle(cfg, n):
    print(f"case {cfg} n={n} lane", flush=True)
    m, Y0, dt0, c, ode, rho, ref = _oracle_run(cfg, n)
    got = _device_run(m, Y0, dt0, c, ode, rho, batched=False, lane=True)
    _check(m, got, ref, tight_dt=True, Y0=Y0, rho=rho, deltaT=c["deltaT"])


@pytest.mark.parametrize("cfg,n", CASES)
def test_plain_and_batched_match_oracle(cfg, n):
    """The drop-in launch against the oracle, and the persistent launch of
    the same kernel equal to it bit for bit."""
    print(f"case {cfg} n={n} plain + batched", flush=True)
    m, Y0, dt0, c, ode, rho, ref = _oracle_run(cfg, n)
    got = _device_run(m, Y0, dt0, c, ode, rho, batched=False)
    _check(m, got, ref, tight_dt=False, Y0=Y0, rho=rho, deltaT=c["deltaT"])
    bat = _device_run(m, Y0, dt0, c, ode, rho, batched=True)
    for a, b in zip(got, bat):
        assert np.array_equal(np.asarray(a), np.asarray(b))


@pytest.mark.parametrize("solver", ["Rosenbrock12", "Rosenbrock23"])
@pytest.mark.parametrize("form", ["lane", "plain", "batched"])
def test_other_solvers(solver, form):
    """Ros12 and Ros23 on H2 states with the reference tutorial tolerances
    (tutorials/h2: absTol 1e-12, relTol 0.1 -- here relTol 1e-5)."""
    from oracle_lib import Oracle
    m, Y0, dt0, c = make_states("h2_small", 128)
    ode = ode_inputs(solver, relTol=1e-5, absTol=1e-10)
    rho = np.ones(128)
    ref = Oracle(m).solve_cells(ode, Y0, 1e-5, dt0, rho=rho, want_rr=True)
    got = _device_run(m, Y0, dt0, c | {"deltaT": 1e-5}, ode, rho,
                      batched=form == "batched", lane=form == "lane")
    _check(m, got, ref, tight_dt=form == "lane")


@pytest.mark.parametrize("solver", ["Rosenbrock12", "Rosenbrock23"])
def test_four_cells_per_wavefront_other_solvers(solver):
    """The two- and three-stage tableaux through the 16-lane-team form
    (set_batched_shape(4)), as test_other_solvers."""
    from oracle_lib import Oracle
    m, Y0, dt0, c = make_states("h2_small", 128)
    ode = ode_inputs(solver, relTol=1e-5, absTol=1e-10)
    rho = np.ones(128)
    ref = Oracle(m).solve_cells(ode, Y0, 1e-5, dt0, rho=rho, want_rr=True)
    print(f"case h2_small n=128 {solver} cells per wavefront 4", flush=True)
    got = _device_run(m, Y0, dt0, c | {"deltaT": 1e-5}, ode, rho, batched=True,
                      cells_per_wave=4)
    _check(m, got, ref, tight_dt=False)


@pytest.mark.parametrize("cfg,n", [("h2_small", 384), ("h2_adjoint", 96)])
def test_four_cells_per_wavefront_matches_oracle(cfg, n):
    """The 16-lane-team form (four cells per wavefront, selected with
    gfc_evaluator_set_batched_shape(4, ...)) on H2 states; cells of a
    wavefront take independent step sequences."""
    from gpufoam_amd.evaluator import Evaluator
    m, Y0, dt0, c, ode, rho, ref = _oracle_run(cfg, n)
    ev = Evaluator(m, ode, 0.0, max_cells=n)
    ev.set_batched_shape(4, 0)
    print(f"case {cfg} n={n} cells per wavefront 4", flush=True)
    dY0 = torch.tensor(Y0, device="cuda")
    dYn = torch.empty_like(dY0)
    dRR = torch.zeros((m.n_species, n), dtype=torch.float64, device="cuda")
    ddt = torch.tensor(dt0, device="cuda")
    dmin = torch.zeros(1, dtype=torch.float64, device="cuda")
    st = torch.zeros((3, n), dtype=torch.int32, device="cuda")
    ev.solve(c["deltaT"], dY0, ddt, Ynew=dYn, RR=dRR, rho=torch.tensor(rho, device="cuda"),
             deltaTMin=dmin, stats=st, batched=True)
    torch.cuda.synchronize()
    got = (dYn.cpu().numpy(), ddt.cpu().numpy(), dmin.item(), st.cpu().numpy(),
           dRR.cpu().numpy())
    _check(m, got, ref, tight_dt=False, Y0=Y0, rho=rho, deltaT=c["deltaT"])
    # repeated launches and graph replays give the same bits
    outs = []
    for graph in (False, True, True):
        Yn = torch.empty_like(dY0)
        d = torch.tensor(dt0, device="cuda")
        ev.solve(c["deltaT"], dY0, d, Ynew=Yn, batched=True, graph=graph)
        torch.cuda.synchronize()
        outs.append((Yn.cpu().numpy(), d.cpu().numpy()))
    for o in outs:
        assert np.array_equal(o[0], got[0]) and np.array_equal(o[1], got[1])


def test_batched_is_deterministic_and_graph_replays():
    m, Y0, dt0, c = make_states("gri_production", 96)
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    rho = np.ones(96)
    a = _device_run(m, Y0, dt0, c, ode, rho, batched=True)
    b = _device_run(m, Y0, dt0, c, ode, rho, batched=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    from gpufoam_amd.evaluator import Evaluator
    ev = Evaluator(m, ode, 0.0, max_cells=96)
    dY0 = torch.tensor(Y0, device="cuda")
    outs = []
    for rep in range(3):  # capture, then two replays, with host work between
        dYn = torch.empty_like(dY0)
        ddt = torch.tensor(dt0, device="cuda")
        ev.solve(c["deltaT"], dY0, ddt, Ynew=dYn, batched=True, graph=True)
        torch.cuda.synchronize()
        outs.append((dYn.cpu().numpy(), ddt.cpu().numpy()))
        _ = np.linalg.norm(outs[-1][0])
    for o in outs:
        assert np.array_equal(o[0], a[0]) and np.array_equal(o[1], a[1])


def test_plain_lane_shape_equals_plain():
    """The batched launcher with one lane per cell runs the lane kernel."""
    from gpufoam_amd.evaluator import Evaluator
    m, Y0, dt0, c = make_states("h2_small", 100)
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    a = _device_run(m, Y0, dt0, c, ode, np.ones(100), batched=False, lane=True)
    ev = Evaluator(m, ode, 0.0, max_cells=100)
    ev.set_batched_shape(64)
    Yn, dt, dmin, st, _ = ev.solve_host(c["deltaT"], Y0, dt0, batched=True)
    assert np.array_equal(Yn, a[0]) and np.array_equal(dt, a[1])


def test_treact_gate_and_clamp():
    """Cells at or below Treact are not integrated and get deltaTChem =
    deltaT (singleCellSolver.H:31,48-50); deltaTChem is clamped to
    deltaTChemMax after deltaTMin is taken (gpuKernelEvaluator.cu:91-94)."""
    from gpufoam_amd.evaluator import Evaluator
    m, Y0, dt0, c = make_states("h2_small", 64)
    nS = m.n_species
    Treact = 1200.0
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    for batched, lane in ((False, True), (False, False), (True, False)):
        ev = Evaluator(m, ode, Treact, max_cells=64)
        Yn, dt, dmin, st, rr = ev.solve_host(c["deltaT"], Y0, dt0, deltaTChemMax=2e-6,
                                             want_rr=True, batched=batched, lane=lane)
        cold = Y0[nS] <= Treact
        assert cold.any() and (~cold).any()
        assert np.array_equal(Yn[:, cold], Y0[:, cold])
        assert (st[0][cold] == 4).all() and (st[1][cold] == 0).all()
        assert (rr[:, cold] == 0).all()
        assert np.all(dt <= 2e-6)
        assert np.isclose(dmin, min(c["deltaT"], dmin))


def test_host_pointer_api_matches_reference_interface():
    """computeRR / computeYNew keep the reference's cell-major layout."""
    from gpufoam_amd.evaluator import Evaluator
    from oracle_lib import Oracle
    m, Y0, dt0, c = make_states("h2_small", 40)
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    ev = Evaluator(m, ode, 0.0, max_cells=64)
    rho = np.linspace(0.5, 1.5, 40)
    RR, dtn, dtmin = ev.computeRR(c["deltaT"], 5e-7, rho, dt0, Y0.T.reshape(-1))
    _, dto, dtmino, _, rro = Oracle(m).solve_cells(ode, Y0, c["deltaT"], dt0, rho=rho,
                                                   want_rr=True, deltaTChemMax=5e-7)
    RR = RR.reshape(40, m.n_species).T
    scale = np.abs(rro).max()
    assert np.allclose(RR, rro, rtol=1e-5, atol=1e-9 * scale)
    assert np.allclose(dtn, dto, rtol=1e-6) and np.isclose(dtmin, dtmino, rtol=1e-6)
    Ynew, dty = ev.computeYNew(c["deltaT"], 1e300, dt0, Y0.T.reshape(-1))
    assert Ynew.shape == (40 * (m.n_species + 2),)


def test_in_place_ld_and_empty():
    from gpufoam_amd.evaluator import Evaluator
    m, Y0, dt0, c = make_states("h2_small", 70)
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    ev = Evaluator(m, ode, 0.0, max_cells=128)
    a = ev.solve_host(c["deltaT"], Y0, dt0)
    # leading dimension 100 > nCells, Ynew aliasing Y0
    big = np.zeros((m.n_species + 2, 100))
    big[:, :70] = Y0
    dY = torch.tensor(big, device="cuda")
    ddt = torch.tensor(dt0, device="cuda")
    ev.solve(c["deltaT"], dY, ddt, Ynew=dY, ldY=100)
    torch.cuda.synchronize()
    assert np.array_equal(dY.cpu().numpy()[:, :70], a[0])
    assert np.array_equal(ddt.cpu().numpy(), a[1])
    # zero cells: deltaTMin = +inf, nothing else touched
    dmin = torch.zeros(1, dtype=torch.float64, device="cuda")
    ev.solve(c["deltaT"], dY[:, :0], ddt[:0], deltaTMin=dmin, ldY=100)
    torch.cuda.synchronize()
    assert dmin.item() == float("inf")
    with pytest.raises(GfcError) as e:
        ev.solve_host(c["deltaT"], np.zeros((m.n_species + 2, 200)), np.ones(200))
    assert e.value.status == 9


def test_max_steps_is_reported():
    """Running out of maxSteps is silent in the reference
    (gpuODESolverBase.H:80-111); here it sets a status bit."""
    from gpufoam_amd.evaluator import Evaluator
    from oracle_lib import Oracle
    m, Y0, dt0, c = make_states("h2_small", 32)
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"], maxSteps=3)
    ref = Oracle(m).solve_cells(ode, Y0, c["deltaT"], dt0)
    for batched, lane in ((False, True), (False, False), (True, False)):
        ev = Evaluator(m, ode, 0.0, max_cells=32)
        Yn, dt, _, st, _ = ev.solve_host(c["deltaT"], Y0, dt0, batched=batched, lane=lane)
        assert np.array_equal(st[0], ref[3][0])
        assert (st[0] & 1).any()


def test_step_underflow_is_reported():
    """With absTol = relTol = 0 every error estimate is infinite, so
    adaptiveSolve shrinks the step by minScale until dx < vSmall
    (gpuAdaptiveSolver.H:111-114, where the reference prints and asserts);
    here the cell stops with GFC_CELL_STEP_UNDERFLOW, as in the oracle."""
    from gpufoam_amd.evaluator import Evaluator
    from oracle_lib import Oracle
    m, Y0, dt0, c = make_states("h2_small", 64)
    ode = ode_inputs(c["solver"], relTol=0.0, absTol=0.0)
    ref = Oracle(m).solve_cells(ode, Y0, c["deltaT"], dt0)
    assert (ref[3][0] & 2).all()
    for batched, lane in ((False, True), (False, False), (True, False)):
        print(f"case h2_small n=64 underflow batched={batched} lane={lane}", flush=True)
        ev = Evaluator(m, ode, 0.0, max_cells=64)
        Yn, dt, _, st, _ = ev.solve_host(c["deltaT"], Y0, dt0, batched=batched, lane=lane)
        assert np.array_equal(st[0], ref[3][0])
        assert np.array_equal(st[1], ref[3][1]) and np.array_equal(st[2], ref[3][2])


def test_wavefront_kernel_with_many_efficiency_sets():
    """A 60-species mechanism with 94 distinct third-body efficiency sets
    (n = 61: the 64-row wavefront kernel).  The kernel's LDS table of
    third-body concentrations once held 64 sets, and the sets beyond it
    overwrote the Gibbs tables; this solve pins the 128-set table."""
    from gpufoam_amd.mechanism import Mechanism
    from gpufoam_amd.synthetic import syn_states, syn_texts
    from oracle_lib import Oracle
    t, r = syn_texts(seed=23, nS=60, nR=300)
    m = Mechanism(t, r, "syn60")
    Y0, dt0 = syn_states(m, 16, seed=24)
    c = dict(solver="Rosenbrock34", relTol=1e-6, absTol=1e-10, deltaT=1e-6)
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    ref = Oracle(m).solve_cells(ode, Y0, c["deltaT"], dt0)
    rho = np.linspace(0.5, 1.1, 16)
    print("case syn60 n=16 batched, wavefront per cell", flush=True)
    got = _device_run(m, Y0, dt0, c, ode, rho, batched=True, cells_per_wave=1)
    from test_gpu_block import _check_conditioned
    _check_conditioned(m, got, ref, Y0, dt0, c, ode)
