"""Two wavefronts per cell (pair_kernel.hip; set_batched_shape(2)) against
the oracle, with the bounds of test_gpu_parity.py: species within
1e-6 |Y| + 1e-10, T within 1e-8 relative, status bits equal, at least 80 %
of the cells on the oracle's step sequence (the kernels sum reaction
contributions in another order than the reference's sequential loop, so a
last-bit difference can move a step by one attempt).  The persistent and
the drop-in launches of the kernel give the same bits, and repeated
launches and graph replays are bitwise identical (every LDS accumulation of
the kernel has a fixed order)."""
import numpy as np
import pytest

from gpufoam_amd import _capi, ode_inputs
from gpufoam_amd.mechanism import Mechanism
from gpufoam_amd.synthetic import syn_states, syn_texts
from gpufoam_amd.workloads import make_states

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from test_gpu_parity import _check, _device_run  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _capi.lib()


def _syn60_small():
    """60 species, 45 reactions, 12 efficiency sets: the padded 64-row
    instantiation, with one reaction pass (the second wavefront has none)."""
    t, r = syn_texts(seed=13, nS=60, nR=45)
    m = Mechanism(t, r, "syn60s")
    Y0, dt0 = syn_states(m, 32, seed=14)
    c = dict(solver="Rosenbrock34", relTol=1e-6, absTol=1e-10, deltaT=1e-6)
    return m, Y0, dt0, c


def _case(cfg, n):
    if cfg == "syn60s":
        m, Y0, dt0, c = _syn60_small()
    else:
        m, Y0, dt0, c = make_states(cfg, n)
    return m, Y0, dt0, c


@pytest.mark.parametrize("cfg,n", [("gri_production", 2048), ("gri_batching", 512),
                                   ("syn60s", 32)])
def test_pair_kernel_matches_oracle(cfg, n):
    from oracle_lib import Oracle
    m, Y0, dt0, c = _case(cfg, n)
    n = Y0.shape[1]
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    rho = np.linspace(0.3, 1.3, n)
    ref = Oracle(m).solve_cells(ode, Y0, c["deltaT"], dt0, rho=rho, want_rr=True)
    print(f"case {cfg} n={n} two wavefronts per cell, persistent", flush=True)
    got = _device_run(m, Y0, dt0, c, ode, rho, batched=True, cells_per_wave=2)
    _check(m, got, ref, tight_dt=False, Y0=Y0, rho=rho, deltaT=c["deltaT"])
    print(f"case {cfg} n={n} two wavefronts per cell, drop-in", flush=True)
    plain = _device_run(m, Y0, dt0, c, ode, rho, batched=False, cells_per_wave=2)
    for a, b in zip(got, plain):
        assert np.array_equal(np.asarray(a), np.asarray(b))


@pytest.mark.parametrize("solver", ["Rosenbrock12", "Rosenbrock23"])
def test_pair_kernel_other_solvers(solver):
    from oracle_lib import Oracle
    m, Y0, dt0, c = make_states("gri_production", 256)
    c = dict(c, solver=solver)
    ode = ode_inputs(solver, relTol=c["relTol"], absTol=c["absTol"])
    rho = np.linspace(0.3, 1.3, 256)
    ref = Oracle(m).solve_cells(ode, Y0, c["deltaT"], dt0, rho=rho, want_rr=True)
    print(f"case gri_production n=256 {solver} two wavefronts per cell", flush=True)
    got = _device_run(m, Y0, dt0, c, ode, rho, batched=True, cells_per_wave=2)
    _check(m, got, ref, tight_dt=False, Y0=Y0, rho=rho, deltaT=c["deltaT"])


def test_pair_kernel_is_deterministic_and_graph_replays():
    from gpufoam_amd.evaluator import Evaluator
    m, Y0, dt0, c = make_states("gri_production", 1024)
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    ev = Evaluator(m, ode, 0.0, max_cells=1024)
    ev.set_batched_shape(2)
    assert ev.batched_shape()[0] == 2
    dY0 = torch.tensor(Y0, device="cuda")
    ddt0 = torch.tensor(dt0, device="cuda")
    outs = []
    for graph in (False, False, True, True, True):
        print(f"case gri_production n=1024 pair graph={graph}", flush=True)
        dYn = torch.empty_like(dY0)
        ddt = ddt0.clone()
        st = torch.zeros((3, 1024), dtype=torch.int32, device="cuda")
        ev.solve(c["deltaT"], dY0, ddt, Ynew=dYn, stats=st, batched=True, graph=graph)
        torch.cuda.synchronize()
        outs.append((dYn.cpu().numpy(), ddt.cpu().numpy(), st.cpu().numpy()))
    for o in outs[1:]:
        for a, b in zip(o, outs[0]):
            assert np.array_equal(a, b)


def test_pair_shape_is_refused_where_it_cannot_run():
    from gpufoam_amd import GfcError
    from gpufoam_amd.evaluator import Evaluator
    m, _, _, c = make_states("h2_small", 4)
    ev = Evaluator(m, ode_inputs(c["solver"]), 0.0, max_cells=4)
    with pytest.raises(GfcError):
        ev.set_batched_shape(2)
