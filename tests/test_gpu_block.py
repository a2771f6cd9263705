"""The workgroup-per-cell kernel (mechanisms with 64 < nSpecie+1 <= 208) on
MI355X: the batched form of the 200-species synthetic mechanism and of a
100-species one (the two register-block sizes of the kernel), against the
oracle, plus its determinism and its graph replay.

Tolerances are those of test_gpu_parity.py: species within rtol 1e-6 +
absTol 1e-10, temperature within 1e-8 relative, identical status bits and
step sequences, RR from the kernel's own final state.  The chemical time
step is checked to 1e-4 relative instead of 1e-6: it scales with the
error estimate to the power -1/4, and the estimate is a difference of stage
vectors whose rounding follows the summation order, which here is blocked
(32-column Jacobian assembly, 32-position triangular solves) instead of the
reference's sequential order; for n = 201 that moves it by up to ~1e-5.
For the same reason a cell near an accept/reject boundary can take a
different number of attempts than the oracle.  Some cells of the synthetic
mechanisms are that sensitive even to the last-bit differences between the
device and host exp/log: the plain kernel, which keeps the reference's
operation order, departs from the oracle on the same cells of the
100-species set (cells 5, 9 and 10 of seed 11, one of them 3.6x the species
tolerance).  Such a cell ends at a different, equally accurate state; the
synthetic-mechanism checks therefore hold the tolerances above on every
cell whose step sequence equals the oracle's, and bound the others at 10x
the species tolerance and 1e-7 in temperature."""
import numpy as np
import pytest

from gpufoam_amd import ode_inputs
from gpufoam_amd.mechanism import Mechanism
from gpufoam_amd.synthetic import syn_states, syn_texts
from gpufoam_amd.workloads import make_states

from test_gpu_parity import _check, _device_run

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gpufoam_amd import _capi
    _capi.lib()


def _syn100():
    t, r = syn_texts(seed=7, nS=100, nR=600)
    return Mechanism(t, r, "syn100")


def _check_sensitive(m, got, ref, Y0, rho, deltaT, min_same=0.5):
    Yg, dtg, _, stg, rrg = got
    Yo, dto, _, sto, _ = ref
    nS = m.n_species
    errY = (np.abs(Yg[:nS] - Yo[:nS]) / (1e-6 * np.abs(Yo[:nS]) + 1e-10)).max(axis=0)
    errT = np.abs(Yg[nS] - Yo[nS]) / Yo[nS]
    same = stg[1] == sto[1]
    print(f"identical step sequence in {same.mean():.3f} of cells; species error / "
          f"tolerance max {errY[same].max():.3g} there, {errY.max():.3g} overall")
    assert (stg[0] == sto[0]).all()
    assert np.array_equal(Yg[nS + 1], Yo[nS + 1])
    assert same.mean() >= min_same
    assert errY[same].max() <= 1.0 and errT[same].max() <= 1e-8
    assert np.allclose(dtg[same], dto[same], rtol=1e-4, atol=0)
    assert errY.max() <= 10.0 and errT.max() <= 1e-7
    assert np.allclose(rrg, rho * (Yg[:nS] - Y0[:nS]) / deltaT, rtol=1e-12, atol=1e-300)


def _oracle(m, Y0, dt0, c, rho, Treact=0.0):
    from oracle_lib import Oracle
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    ref = Oracle(m).solve_cells(ode, Y0, c["deltaT"], dt0, rho=rho, want_rr=True,
                                Treact=Treact)
    return ode, ref


def test_block_kernel_syn200_matches_oracle():
    """Twelve cells of the 200-species state law (n = 201: 13 x 7 register
    blocks); two of them below Treact are skipped."""
    m, Y0, dt0, c = make_states("syn200", 16)
    keep = [0, 1, 3, 5, 6, 8, 9, 11, 12, 13, 14, 15]
    Y0, dt0 = np.ascontiguousarray(Y0[:, keep]), dt0[keep]
    nS = m.n_species
    Treact = float(np.sort(Y0[nS])[1]) + 1e-9  # the two coolest cells stay put
    rho = np.linspace(0.4, 1.2, len(keep))
    ode, ref = _oracle(m, Y0, dt0, c, rho, Treact)
    print(f"case syn200 n={len(keep)} batched Treact={Treact}", flush=True)
    got = _device_run(m, Y0, dt0, c, ode, rho, batched=True, Treact=Treact)
    assert ((got[3][0] & 4) != 0).sum() == 2  # GFC_CELL_SKIPPED on two cells
    _check(m, got, ref, tight_dt=True, Y0=Y0, rho=rho, deltaT=c["deltaT"], dt_rtol=1e-4)


def test_block_kernel_syn100_matches_oracle():
    """A 100-species / 600-reaction synthetic mechanism (n = 101: the
    8 x 4 register-block instantiation)."""
    m = _syn100()
    Y0, dt0 = syn_states(m, 12, seed=11)
    c = dict(solver="Rosenbrock34", relTol=1e-6, absTol=1e-10, deltaT=1e-6)
    rho = np.linspace(0.4, 1.2, Y0.shape[1])
    ode, ref = _oracle(m, Y0, dt0, c, rho)
    print("case syn100 n=12 batched", flush=True)
    got = _device_run(m, Y0, dt0, c, ode, rho, batched=True)
    _check_sensitive(m, got, ref, Y0, rho, c["deltaT"])


def test_block_kernel_deterministic_and_graph_replay():
    """No atomics in the arithmetic: repeated eager launches and replays of
    the captured graph give bit-identical states, chemical time steps and
    step counts."""
    from gpufoam_amd.evaluator import Evaluator
    m, Y0, dt0, c = make_states("syn200", 8)
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    ev = Evaluator(m, ode, 0.0, max_cells=8)
    dY0 = torch.tensor(Y0, device="cuda")
    runs = []
    for graph in (False, False, True, True, True):
        print(f"case syn200 n=8 graph={graph}", flush=True)
        Yn = torch.empty_like(dY0)
        dt = torch.tensor(dt0, device="cuda")
        st = torch.zeros((3, 8), dtype=torch.int32, device="cuda")
        ev.solve(c["deltaT"], dY0, dt, Ynew=Yn, stats=st, batched=True, graph=graph)
        torch.cuda.synchronize()
        runs.append((Yn.cpu().numpy(), dt.cpu().numpy(), st.cpu().numpy()))
    for r in runs[1:]:
        for a, b in zip(runs[0], r):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("solver", ["Rosenbrock12", "Rosenbrock23"])
def test_block_kernel_other_solvers(solver):
    """The two- and three-stage Rosenbrock tableaux through the
    workgroup-per-cell kernel (100-species mechanism, relTol 1e-5)."""
    m = _syn100()
    Y0, dt0 = syn_states(m, 8, seed=12)
    c = dict(solver=solver, relTol=1e-5, absTol=1e-10, deltaT=1e-6)
    rho = np.linspace(0.5, 1.1, Y0.shape[1])
    ode, ref = _oracle(m, Y0, dt0, c, rho)
    print(f"case syn100 n=8 {solver}", flush=True)
    got = _device_run(m, Y0, dt0, c, ode, rho, batched=True)
    _check_sensitive(m, got, ref, Y0, rho, c["deltaT"])


def test_block_kernel_reports_max_steps():
    """Running out of maxSteps sets the status bit, as in the oracle."""
    from gpufoam_amd.evaluator import Evaluator
    from oracle_lib import Oracle
    m = _syn100()
    Y0, dt0 = syn_states(m, 8, seed=13)
    ode = ode_inputs("Rosenbrock34", relTol=1e-6, absTol=1e-10, maxSteps=3)
    ref = Oracle(m).solve_cells(ode, Y0, 1e-6, dt0)
    print("case syn100 n=8 maxSteps=3", flush=True)
    ev = Evaluator(m, ode, 0.0, max_cells=8)
    _, _, _, st, _ = ev.solve_host(1e-6, Y0, dt0, batched=True)
    assert np.array_equal(st[0], ref[3][0])
    assert (st[0] & 1).any()
