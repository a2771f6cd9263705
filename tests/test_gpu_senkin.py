"""The CHEMKIN-II SENKIN ignition cases of the tutorials (H2: 2 atm,
1000 K, 1 ms; GRI: 13.5 atm, 1000 K, 70 ms; constant pressure) integrated
on the device through gfc_solve with the mechanisms compiled from the
tutorials' CHEMKIN inputs, in CFD steps of 1e-6 s (H2) and 1e-5 s (GRI),
Rosenbrock34 at relTol 1e-6 / absTol 1e-10: ignition time within 0.5 %,
final temperature within 1 K and T(t) within 2 K outside the ignition front
of CHEMKIN-II (senk.out), and no cell status raised."""
import numpy as np
import pytest

from gpufoam_amd import ode_inputs

import senkin_case as SC

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gpufoam_amd import _capi
    _capi.lib()


@pytest.mark.parametrize("name", ["h2", "gri"])
@pytest.mark.parametrize("batched", [False, True])
def test_senkin_ignition_on_the_device(name, batched):
    from gpufoam_amd.evaluator import Evaluator
    print(f"case senkin {name} batched={batched}", flush=True)
    m, y, rec = SC.load(name)
    nS = m.n_species
    ode = ode_inputs("Rosenbrock34", relTol=1e-6, absTol=1e-10)
    ev = Evaluator(m, ode, 0.0, max_cells=1)
    cs = SC.CASES[name]
    n = int(round(cs["end"] / cs["sample"]))
    dY = torch.tensor(y, device="cuda")
    ddt = torch.full((1,), 1e-8, dtype=torch.float64, device="cuda")
    st = torch.zeros((3, 1), dtype=torch.int32, device="cuda")
    T = torch.empty(n + 1, dtype=torch.float64, device="cuda")
    bad = torch.zeros(1, dtype=torch.int32, device="cuda")
    T[0] = dY[nS, 0]
    for k in range(n):
        ev.solve(cs["sample"], dY, ddt, Ynew=dY, stats=st, batched=batched)
        T[k + 1] = dY[nS, 0]
        bad |= st[0]
    torch.cuda.synchronize()
    assert bad.item() == 0
    ts = np.arange(n + 1) * cs["sample"]
    dtig, dTend, dT = SC.compare(ts, T.cpu().numpy(), rec)
    print(f"{name}: ignition {dtig:+.3%}, final T {dTend:+.2f} K, |dT| away from the front "
          f"{dT:.2f} K")
    assert abs(dtig) < 5e-3 and abs(dTend) < 1.0 and dT < 2.0
