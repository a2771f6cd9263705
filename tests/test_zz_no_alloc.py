"""gfc_solve_noalloc, the allocation-free persistent solve, against
gfc_solve's persistent form through the C ABI: two calls in a row on one
evaluator, the second with fewer cells (the shared device array keeps the
first call's queue counter and minimum, which each call must clear), for
the wavefront kernel (GRI, H2) and the four-cells-per-wavefront kernel
(H2); results are bit-identical.  Mechanisms beyond the array's reach are
rejected with GFC_ERR_CAPACITY."""
import numpy as np
import pytest

from gpufoam_amd import GfcError, _capi, ode_inputs
from gpufoam_amd.workloads import make_states

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

DEV = "cuda"


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _capi.lib()


def _call(ev, fn, deltaT, Y0, dt0, n):
    """One solve of the first n cells (ldY = the full width)."""
    L = _capi.lib()
    ld = Y0.shape[1]
    Yn = torch.full_like(Y0, -1.0)
    dt = dt0.clone()
    dmin = torch.zeros(1, dtype=torch.float64, device=DEV)
    st = torch.full((3, ld), -7, dtype=torch.int32, device=DEV)
    rho = torch.linspace(0.3, 1.3, ld, dtype=torch.float64, device=DEV)
    RR = torch.full((Y0.shape[0] - 2, ld), -1.0, dtype=torch.float64, device=DEV)
    stream = torch.cuda.current_stream().cuda_stream
    args = (ev.handle, deltaT, 1e300, n, ld, rho.data_ptr(), Y0.data_ptr(), Yn.data_ptr(),
            RR.data_ptr(), dt.data_ptr(), dmin.data_ptr(), st.data_ptr(), stream)
    if fn == "noalloc":
        _capi.check(L.gfc_solve_noalloc(*args))
    else:
        _capi.check(L.gfc_solve(*args, _capi.BATCHED))
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (Yn, dt, dmin, st, RR)]


@pytest.mark.parametrize("cfg,n1,n2,cpw", [("gri_production", 640, 200, 0),
                                           ("h2_small", 512, 130, 0),
                                           ("h2_small", 512, 130, 4)])
def test_noalloc_equals_persistent_solve(cfg, n1, n2, cpw):
    from gpufoam_amd.evaluator import Evaluator
    print(f"case {cfg} cells {n1} then {n2} cells per wavefront {cpw}", flush=True)
    m, Y0, dt0, c = make_states(cfg, n1)
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    ev = Evaluator(m, ode, 0.0, max_cells=n1)
    if cpw:
        ev.set_batched_shape(cpw)
    dY0 = torch.tensor(Y0, device=DEV)
    ddt0 = torch.tensor(dt0, device=DEV)
    for n in (n1, n2):
        a = _call(ev, "noalloc", c["deltaT"], dY0, ddt0, n)
        b = _call(ev, "solve", c["deltaT"], dY0, ddt0, n)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        # cells beyond n are untouched, and the minimum is over the n cells
        assert (a[0][:, n:] == -1.0).all() and (a[3][:, n:] == -7).all()
        assert a[2][0] == b[2][0] > 0
        assert (a[3][1, :n] > 0).all()


def test_noalloc_rejects_what_it_cannot_hold():
    from gpufoam_amd.evaluator import Evaluator
    m, Y0, dt0, c = make_states("h2_small", 64)
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    ev = Evaluator(m, ode, 0.0, max_cells=64)
    dY0 = torch.tensor(Y0, device=DEV)
    ddt0 = torch.tensor(dt0, device=DEV)
    L = _capi.lib()
    # more cells than the evaluator holds
    with pytest.raises(GfcError) as e:
        _capi.check(L.gfc_solve_noalloc(ev.handle, c["deltaT"], 1e300, 65, 65, None,
                                        dY0.data_ptr(), None, None, ddt0.data_ptr(), None,
                                        None, None))
    assert e.value.status == 9
    # one lane per cell keeps per-cell scratch: not served
    ev.set_batched_shape(64)
    with pytest.raises(GfcError) as e:
        _capi.check(L.gfc_solve_noalloc(ev.handle, c["deltaT"], 1e300, 64, 64, None,
                                        dY0.data_ptr(), None, None, ddt0.data_ptr(), None,
                                        None, None))
    assert e.value.status == 9
    # the workgroup-per-cell mechanism (200 species)
    m, Y0, dt0, c = make_states("syn200", 4)
    ev = Evaluator(m, ode, 0.0, max_cells=4)
    dY0 = torch.tensor(Y0, device=DEV)
    ddt0 = torch.tensor(dt0, device=DEV)
    with pytest.raises(GfcError) as e:
        _capi.check(L.gfc_solve_noalloc(ev.handle, c["deltaT"], 1e300, 4, 4, None,
                                        dY0.data_ptr(), None, None, ddt0.data_ptr(), None,
                                        None, None))
    assert e.value.status == 9
    torch.cuda.synchronize()
