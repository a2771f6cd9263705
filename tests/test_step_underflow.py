"""Step-size underflow (gpuAdaptiveSolver.H:111-114: the reference prints
and asserts when dx < vSmall).  Here the cell stops with
GFC_CELL_STEP_UNDERFLOW and keeps its last accepted state.  With relTol =
absTol = 0 every attempt's error is infinite, so each retry multiplies the
step by minScale until it falls below vSmall: the attempt count follows
from the initial step alone."""
import numpy as np
import pytest

import adjoint_host as AH
from gpufoam_amd import _capi, ode_inputs
from gpufoam_amd.workloads import make_states
from oracle_lib import Oracle


def underflow_attempts(dx0: float, min_scale: float = 0.2,
                       vsmall: float = float(np.finfo(np.float64).tiny)) -> int:
    """Attempts until dx (multiplied by min_scale after each rejection, in
    this order of rounding) drops below vSmall (DBL_MIN, gpu_constants.H)."""
    dx, k = dx0, 0
    while True:
        k += 1
        dx *= min_scale
        if dx < vsmall:
            return k


def zero_tolerance_case(config="h2_small", cells=4):
    m, Y0, dt0, c = make_states(config, cells)
    ode = ode_inputs(c["solver"], relTol=0.0, absTol=0.0)
    return m, Y0, dt0, c, ode


def test_oracle_reports_step_underflow():
    m, Y0, dt0, c, ode = zero_tolerance_case()
    assert ode.minScale == 0.2
    Yn, dt, _, stats, _ = Oracle(m).solve_cells(ode, Y0, c["deltaT"], dt0)
    for i in range(Y0.shape[1]):
        assert stats[0, i] & _capi.CELL_STEP_UNDERFLOW
        assert stats[2, i] == 0
        # the first step is dtChem, or deltaT when dtChem overshoots
        assert stats[1, i] == underflow_attempts(min(dt0[i], c["deltaT"]))
    # nothing accepted: the state is the initial one, clamped at zero
    nS = m.n_species
    np.testing.assert_array_equal(Yn[:nS], np.maximum(Y0[:nS], 0.0))
    np.testing.assert_array_equal(Yn[nS:], Y0[nS:])


@pytest.mark.skipif(not AH.available(), reason="tests/native not built")
def test_library_solver_reports_step_underflow_on_the_host():
    """The library's one-lane solver (the device source built for the host)
    stops the cell with the status bit, as the oracle does."""
    m, Y0, dt0, c, ode = zero_tolerance_case()
    ref = Oracle(m).solve_cells(ode, Y0, c["deltaT"], dt0)
    for i in range(Y0.shape[1]):
        y, dt, st = AH.cell_solve(m, ode, Y0[:, i], dt0[i], c["deltaT"])
        assert st[0] == _capi.CELL_STEP_UNDERFLOW
        assert st[1] == ref[3][1, i] and st[2] == 0
        np.testing.assert_array_equal(y[:m.n_species + 1], ref[0][:m.n_species + 1, i])


@pytest.mark.gpu
@pytest.mark.parametrize("config,cells,form", [
    ("h2_small", 64, "batched4"), ("gri_production", 64, "dropin"),
    ("gri_production", 64, "batched"), ("syn200", 4, "batched")])
def test_every_kernel_family_reports_step_underflow(config, cells, form):
    """The 16-lane-team, wavefront (n = 54) and workgroup kernels stop the
    cell with GFC_CELL_STEP_UNDERFLOW after the same attempts as the oracle
    and return the clamped initial state (the one-lane and H2 wavefront
    forms: test_gpu_parity.py::test_step_underflow_is_reported).  GRI's
    cold unburnt cells (T < 420 K) have frozen chemistry: every stage is
    exactly zero, the error estimate 0/0 is skipped by the reference's
    max (err = 0), the steps are accepted and the cell ends at maxSteps,
    in the oracle as on the device."""
    import torch
    from gpufoam_amd.evaluator import Evaluator
    print(f"[underflow] {config} {cells} cells, {form}", flush=True)
    m, Y0, dt0, c, ode = zero_tolerance_case(config, cells)
    ref = Oracle(m).solve_cells(ode, Y0, c["deltaT"], dt0)
    ev = Evaluator(m, ode, 0.0, max_cells=cells)
    if form == "batched4":
        ev.set_batched_shape(4)
    dY0 = torch.tensor(Y0, device="cuda:0")
    dYn = torch.empty_like(dY0)
    ddt = torch.tensor(dt0, device="cuda:0")
    st = torch.zeros((3, cells), dtype=torch.int32, device="cuda:0")
    ev.solve(c["deltaT"], dY0, ddt, Ynew=dYn, stats=st, batched=form.startswith("batched"),
             lane=form == "lane")
    torch.cuda.synchronize()
    stats = st.cpu().numpy()
    under = (ref[3][0] & _capi.CELL_STEP_UNDERFLOW) != 0
    assert under.sum() >= cells // 2, ref[3][0]
    np.testing.assert_array_equal(stats[0], ref[3][0])
    np.testing.assert_array_equal(stats[1][under], ref[3][1][under])
    # a frozen component can let a few steps through before the step collapses
    np.testing.assert_array_equal(stats[2][under], ref[3][2][under])
    # nothing accepted: the clamped initial state, bit for bit
    none = under & (ref[3][2] == 0)
    assert none.any()
    np.testing.assert_array_equal(dYn.cpu().numpy()[:, none], ref[0][:, none])
    np.testing.assert_array_equal(ddt.cpu().numpy()[none], ref[1][none])
