"""The H2 mechanism with seven reactions recast to the rate types gpuFoam
rejects (SRI fall-off, chemically activated Lindemann/Troe/SRI,
Landau-Teller, Janev, power series; test_extended_rates.extended_h2) on
the device: every launch form runs the one-lane kernel for such a
mechanism and matches the oracle with identical step counts (the bound of
test_gpu_parity.py); the chemistry queries match the oracle; the discrete
adjoint reports that it does not cover these types."""
import numpy as np
import pytest

from gpufoam_amd import GfcError, ode_inputs
from gpufoam_amd.workloads import make_states

from test_extended_rates import extended_h2
from test_gpu_parity import _check, _device_run

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gpufoam_amd import _capi
    _capi.lib()


def test_extended_mechanism_solve_matches_oracle():
    from oracle_lib import Oracle
    m = extended_h2()
    _, Y0, dt0, c = make_states("h2_small", 256)
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    rho = np.linspace(0.3, 1.3, 256)
    ref = Oracle(m).solve_cells(ode, Y0, c["deltaT"], dt0, rho=rho, want_rr=True)
    outs = []
    for batched, lane in ((False, True), (False, False), (True, False)):
        print(f"case h2x n=256 batched={batched} lane={lane}", flush=True)
        got = _device_run(m, Y0, dt0, c, ode, rho, batched=batched, lane=lane)
        _check(m, got, ref, tight_dt=True, Y0=Y0, rho=rho, deltaT=c["deltaT"])
        outs.append(got)
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(np.asarray(a), np.asarray(b))


def test_extended_mechanism_queries_match_oracle():
    from gpufoam_amd.evaluator import Evaluator
    from oracle_lib import Oracle
    m = extended_h2()
    _, Y0, _, c = make_states("h2_small", 128)
    rho = np.linspace(0.4, 1.4, 128)
    o = Oracle(m)
    ev = Evaluator(m, ode_inputs(c["solver"]), 0.0, max_cells=128)
    dY, drho = torch.tensor(Y0, device="cuda"), torch.tensor(rho, device="cuda")
    RR = torch.empty((m.n_species, 128), dtype=torch.float64, device="cuda")
    tc = torch.empty(128, dtype=torch.float64, device="cuda")
    rr = torch.empty(128, dtype=torch.float64, device="cuda")
    ev.calculate(drho, dY, RR)
    ev.tc(drho, dY, tc)
    torch.cuda.synchronize()
    scale = sum(np.abs(o.specie_reaction_rr(r, rho, Y0)) for r in range(m.n_reactions))
    assert np.all(np.abs(RR.cpu().numpy() - o.calculate(rho, Y0)) <= 1e-12 * scale)
    assert np.allclose(tc.cpu().numpy(), o.tc(rho, Y0), rtol=1e-11, atol=0)
    for r in (0, 2, 3, 4, 5, 10, 22):
        ev.reactionRR(r, drho, dY, rr)
        torch.cuda.synchronize()
        want = o.reaction_rr(r, rho, Y0)
        s = np.array([np.abs(o.omega_fr(r, Y0[-1, i], Y0[-2, i], rho[i] * Y0[:-2, i] /
                                        m.arrays["W"])).sum() for i in range(128)])
        assert np.all(np.abs(rr.cpu().numpy() - want) <= 1e-12 * s), r


def test_extended_mechanism_adjoint_and_shapes_are_refused():
    from gpufoam_amd.evaluator import Evaluator
    m = extended_h2()
    _, Y0, dt0, c = make_states("h2_small", 16)
    ev = Evaluator(m, ode_inputs(c["solver"]), 0.0, max_cells=16)
    with pytest.raises(GfcError) as e:
        ev.adjoint_forward(c["deltaT"], torch.tensor(Y0, device="cuda"),
                           torch.tensor(dt0, device="cuda"))
    assert e.value.status == 3
    with pytest.raises(GfcError) as e:
        ev.set_batched_shape(4)
    assert e.value.status == 1
