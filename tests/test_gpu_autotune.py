"""Launch-shape autotuner of the batched solve (MI355X)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from gpufoam_amd import GfcError, ode_inputs  # noqa: E402
from gpufoam_amd.autotune import autotune  # noqa: E402
from gpufoam_amd.evaluator import Evaluator  # noqa: E402
from gpufoam_amd.workloads import make_states  # noqa: E402


@pytest.mark.parametrize("config", ["h2_small", "gri_production"])
def test_autotune_sets_the_fastest_shape_and_keeps_results(config):
    m, Y0, dt0, c = make_states(config, 512)
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    ev = Evaluator(m, ode, 0.0, max_cells=512)
    dY0 = torch.tensor(Y0, device="cuda")
    ddt = torch.tensor(dt0, device="cuda")
    times = autotune(ev, c["deltaT"], dY0, ddt, sample=256, repeats=1)
    best = times.pop("best")
    print(config, {k: round(v, 3) for k, v in times.items()}, "best", best)
    assert best == min(times, key=times.get)
    assert ev.batched_shape()[:2] == best
    # every shape computes the same chemistry (within the parity bound)
    ref = ev.solve_host(c["deltaT"], Y0, dt0)[0]
    for shape in times:
        ev.set_batched_shape(*shape)
        got = ev.solve_host(c["deltaT"], Y0, dt0, batched=True)[0]
        nS = m.n_species
        assert np.all(np.abs(got[:nS] - ref[:nS]) <= 1e-6 * np.abs(ref[:nS]) + 1e-10)
        assert np.all(np.abs(got[nS] - ref[nS]) <= 1e-8 * ref[nS])


def test_invalid_shapes_are_rejected():
    m, Y0, dt0, c = make_states("h2_small", 8)
    ev = Evaluator(m, ode_inputs(c["solver"]), 0.0, max_cells=8)
    for bad in ((2, 0, 0), (16, 0, 0), (1, -1, 0), (1, 0, 8)):
        with pytest.raises(GfcError) as e:
            ev.set_batched_shape(*bad)
        assert e.value.status == 1
    ev.set_batched_shape(4, 0)  # 16-lane teams: H2 has n = 11
    # four cells per wavefront needs n <= 16: not GRI
    m, Y0, dt0, c = make_states("gri_production", 8)
    ev = Evaluator(m, ode_inputs(c["solver"]), 0.0, max_cells=8)
    with pytest.raises(GfcError) as e:
        ev.set_batched_shape(4, 0)
    assert e.value.status == 1


def test_autotune_factorization_axis_on_the_workgroup_kernel():
    """For a mechanism on the workgroup kernel (100-species synthetic,
    n = 101) the tuner also times the three LU variants and keeps the
    fastest; all three give bit-identical states."""
    from gpufoam_amd import _capi
    from gpufoam_amd.mechanism import Mechanism
    from gpufoam_amd.synthetic import syn_states, syn_texts
    t, r = syn_texts(seed=7, nS=100, nR=600)
    m = Mechanism(t, r, "syn100")
    Y0, dt0 = syn_states(m, 64, seed=11)
    ode = ode_inputs("Rosenbrock34", relTol=1e-6, absTol=1e-10)
    ev = Evaluator(m, ode, 0.0, max_cells=64)
    dY0 = torch.tensor(Y0, device="cuda")
    ddt = torch.tensor(dt0, device="cuda")
    print("case syn100 n=64 autotune with LU variants", flush=True)
    times = autotune(ev, 1e-6, dY0, ddt, sample=0, repeats=1, shapes=[(1, 0)])
    lu = {k[1]: v for k, v in times.items() if isinstance(k, tuple) and k[0] == "lu"}
    print({k: round(v, 3) for k, v in lu.items()}, "best", times["best_lu"])
    assert set(lu) == {_capi.LU_COLUMN, _capi.LU_PANEL, _capi.LU_PANEL_MFMA}
    assert times["best_lu"] == min(lu, key=lu.get)
    assert ev.factorization() == times["best_lu"]
    outs = []
    for v in lu:
        ev.set_factorization(v)
        outs.append(ev.solve_host(1e-6, Y0, dt0, batched=True)[0])
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
