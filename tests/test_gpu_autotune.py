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
