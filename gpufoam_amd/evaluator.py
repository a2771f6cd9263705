"""Host-side face of the chemistry evaluator.

``Evaluator`` mirrors FoamGpu::GpuKernelEvaluator
(gpu_chemistry/src/gpuKernelEvaluator/gpuKernelEvaluator/gpuKernelEvaluator.H:18-85):
construct it once per mechanism with the odeCoeffs and Treact of
chemistryProperties, then call ``computeRR`` (or ``computeYNew``) once per
CFD time step.  The reference methods take and return host vectors with the
state packed cell-major, [cell][Y..., T, p]; they are kept with the same
meaning here.  ``solve`` is the direct, allocation-free entry for device
tensors in the library's structure-of-arrays layout.

torch is used only for device memory and streams; all chemistry runs in
libgpufoam_amd's HIP kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from .mechanism import Mechanism

_vp = C.c_void_p


def _ptr(a):
    """Device (torch) or host (numpy) data pointer, None for None."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


def _check_device(name, t, rows, ld, dtype="float64", device=None, min_cols=None):
    """Rejects a device buffer the kernels would index out of bounds: it must
    be a contiguous torch tensor of `dtype` with `rows` rows of exactly `ld`
    columns (at least `min_cols` when given instead), on `device`."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name}: expected a torch tensor, got {type(t).__name__}")
    if str(t.dtype) != f"torch.{dtype}":
        raise ValueError(f"{name}: dtype {t.dtype}, expected torch.{dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: must be contiguous")
    if t.device.type != "cuda" or (device is not None and t.device != device):
        raise ValueError(f"{name}: on {t.device}, expected {device or 'a cuda device'}")
    if t.dim() != 2 or t.shape[0] < rows:
        raise ValueError(f"{name}: shape {tuple(t.shape)}, expected [{rows}, {ld}]")
    if min_cols is not None:
        if t.shape[1] < min_cols:
            raise ValueError(f"{name}: {t.shape[1]} columns, at least {min_cols} needed")
    elif t.shape[1] != ld:
        raise ValueError(f"{name}: {t.shape[1]} columns, expected exactly {ld}")


class Evaluator:
    def __init__(self, mech: Mechanism, ode: _capi.OdeInputs, Treact: float = 0.0,
                 max_cells: int = 1 << 16):
        self.mech = mech
        self.ode = ode
        self.Treact = Treact
        self.max_cells = max_cells
        self.nS = mech.n_species
        self.nE = self.nS + 2
        h = _vp()
        _capi.check(_capi.lib().gfc_evaluator_create(mech.handle, C.byref(ode), Treact,
                                                    max_cells, C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _capi.lib().gfc_evaluator_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    # ----------------------------------------------------------- batched
    def set_batched_shape(self, cells_per_wave: int, waves_per_block: int = 0,
                          chunk: int = 0) -> None:
        _capi.check(_capi.lib().gfc_evaluator_set_batched_shape(
            self._h, cells_per_wave, waves_per_block, chunk))

    def batched_shape(self) -> tuple[int, int, int]:
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _capi.check(_capi.lib().gfc_evaluator_get_batched_shape(
            self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def set_factorization(self, variant: int) -> None:
        """LU variant of the workgroup-per-cell kernel: _capi.LU_COLUMN,
        LU_PANEL (bit-identical factors, one barrier per four columns) or
        LU_PANEL_MFMA (rank-4 trailing updates on the matrix cores)."""
        _capi.check(_capi.lib().gfc_evaluator_set_factorization(self._h, variant))

    def factorization(self) -> int:
        v = C.c_int()
        _capi.check(_capi.lib().gfc_evaluator_get_factorization(self._h, C.byref(v)))
        return v.value

    # -------------------------------------------------------------- solve
    def solve(self, deltaT: float, Y0, dtChem, Ynew=None, RR=None, rho=None,
              deltaTMin=None, stats=None, deltaTChemMax: float = 1e300,
              batched: bool = False, graph: bool = False, stream=None,
              ldY: int | None = None, lane: bool = False) -> None:
        """One CFD step on device tensors (SoA layout, see gpufoam_amd.h).

        Y0 [nS+2, ld], Ynew like Y0 (may be Y0), RR [nS, ld], rho [n],
        dtChem [n] (in/out), deltaTMin [1], stats int32 [3, ld].
        batched: the persistent work-stealing launch; otherwise one launch
        over the cells (the drop-in form), with lane=True one lane per cell.
        Asynchronous on ``stream`` (default: torch's current stream)."""
        n = dtChem.shape[0]
        ld = ldY if ldY is not None else Y0.shape[1]
        flags = 0
        if batched:
            flags |= _capi.BATCHED
        if graph:
            flags |= _capi.GRAPH
        if lane:
            flags |= _capi.LANE_PER_CELL
        if stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        _capi.check(_capi.lib().gfc_solve(
            self._h, deltaT, deltaTChemMax, n, ld, _ptr(rho), _ptr(Y0), _ptr(Ynew),
            _ptr(RR), _ptr(dtChem), _ptr(deltaTMin), _ptr(stats), stream, flags))

    def solve_host(self, deltaT: float, Y0: np.ndarray, dtChem: np.ndarray,
                   rho: np.ndarray | None = None, deltaTChemMax: float = 1e300,
                   want_rr: bool = False, batched: bool = False, lane: bool = False):
        """Host-array solve (staged through the evaluator's device buffers).
        Returns (Ynew, dtChem_new, deltaTMin, stats, RR)."""
        Y0 = np.ascontiguousarray(Y0, dtype=np.float64)
        n = Y0.shape[1]
        Yn = np.zeros_like(Y0)
        dt = np.array(dtChem, dtype=np.float64).copy()
        dtmin = np.zeros(1)
        stats = np.zeros((3, n), dtype=np.int32)
        rr = np.zeros((self.nS, n)) if want_rr else None
        rho_a = np.ascontiguousarray(rho if rho is not None else np.ones(n), dtype=np.float64)
        flags = (_capi.HOST_POINTERS | (_capi.BATCHED if batched else 0) |
                 (_capi.LANE_PER_CELL if lane else 0))
        _capi.check(_capi.lib().gfc_solve(
            self._h, deltaT, deltaTChemMax, n, n, _ptr(rho_a) if want_rr else None,
            _ptr(Y0), _ptr(Yn), _ptr(rr), _ptr(dt), _ptr(dtmin), _ptr(stats), None, flags))
        return Yn, dt, dtmin[0], stats, rr

    # ---------------------------------------------------- rate parameters
    @property
    def n_params(self) -> int:
        return _capi.lib().gfc_evaluator_param_count(self._h)

    def get_rate_parameters(self) -> np.ndarray:
        """[nR, 6]: ln|A|, beta, Ta of k (k0 for fall-off) and of kInf."""
        out = np.zeros(self.n_params)
        _capi.check(_capi.lib().gfc_evaluator_get_rate_parameters(
            self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out.reshape(-1, 6)

    def set_rate_parameters(self, params) -> None:
        a = np.ascontiguousarray(params, dtype=np.float64).reshape(-1)
        if a.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {a.size}")
        _capi.check(_capi.lib().gfc_evaluator_set_rate_parameters(
            self._h, a.ctypes.data_as(C.POINTER(C.c_double))))

    # ------------------------------------------------------------ adjoint
    def set_checkpoint_interval(self, K: int) -> None:
        _capi.check(_capi.lib().gfc_adjoint_set_checkpoint_interval(self._h, K))

    @property
    def checkpoint_interval(self) -> int:
        return _capi.lib().gfc_adjoint_checkpoint_interval(self._h)

    def adjoint_forward(self, deltaT: float, Y0, dtChem, Ynew=None, stats=None,
                        stream=None) -> None:
        """Recording forward solve on device tensors (SoA [nS+2, ld]);
        Ynew and stats, when given, share Y0's leading dimension."""
        n = dtChem.shape[0]
        ld = Y0.shape[1]
        _check_device("Y0", Y0, self.nE, ld, min_cols=n)
        _check_device("dtChem", dtChem.reshape(1, -1), 1, n, device=Y0.device)
        if Ynew is not None:
            _check_device("Ynew", Ynew, self.nE, ld, device=Y0.device)
        if stats is not None:
            _check_device("stats", stats, 3, ld, dtype="int32", device=Y0.device)
        if stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        _capi.check(_capi.lib().gfc_adjoint_forward(
            self._h, deltaT, n, Y0.shape[1], _ptr(Y0), _ptr(dtChem), _ptr(Ynew),
            _ptr(stats), stream))

    def adjoint_backward(self, Ybar, Y0bar, paramBar=None, stream=None) -> None:
        """Ybar, Y0bar [nS+1, ld]; paramBar [6 nR, ld] or None, all with
        the same leading dimension ld (the kernels index Y0bar and paramBar
        with Ybar's)."""
        ld = Ybar.shape[1] if Ybar.dim() == 2 else -1
        _check_device("Ybar", Ybar, self.nS + 1, ld)
        _check_device("Y0bar", Y0bar, self.nS + 1, ld, device=Ybar.device)
        if paramBar is not None:
            _check_device("paramBar", paramBar, 6 * self.mech.n_reactions, ld,
                          device=Ybar.device)
        if stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        _capi.check(_capi.lib().gfc_adjoint_backward(
            self._h, _ptr(Ybar), Ybar.shape[1], _ptr(Y0bar), _ptr(paramBar), stream))

    def adjoint_replay(self, Y0, Ynew, stream=None) -> None:
        """Re-integrates the recorded accepted steps from Y0 into Ynew
        (both [nS+2, ld] with the same ld)."""
        ld = Y0.shape[1] if Y0.dim() == 2 else -1
        _check_device("Y0", Y0, self.nE, ld)
        _check_device("Ynew", Ynew, self.nE, ld, device=Y0.device)
        if stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        _capi.check(_capi.lib().gfc_adjoint_replay(
            self._h, _ptr(Y0), Y0.shape[1], _ptr(Ynew), stream))

    # ------------------------------------------------- device evaluations
    # One kernel family's own derivative, Jacobian and LU solve
    # (gpuODESystem.H:164-281, ludecompose.H:9-140) on device tensors; shape
    # as _capi.SHAPE_* (0: the family the batched solve uses).
    def eval_derivatives(self, Y, shape: int = 0, stream=None):
        """Y [nS+2, n] -> dYdt [nS+1, n] (species, T)."""
        import torch
        n = Y.shape[1]
        _check_device("Y", Y, self.nE, n)
        d = torch.empty((self.nS + 1, n), dtype=torch.float64, device=Y.device)
        _capi.check(_capi.lib().gfc_eval_derivatives(
            self._h, shape, n, n, _ptr(Y), _ptr(d), self._stream(stream)))
        return d

    def eval_jacobian(self, Y, shape: int = 0, stream=None):
        """Y [nS+2, n] -> (dYdt [nS+1, n], J [n, nS+1, nS+1])."""
        import torch
        n = Y.shape[1]
        _check_device("Y", Y, self.nE, n)
        m = self.nS + 1
        d = torch.empty((m, n), dtype=torch.float64, device=Y.device)
        J = torch.empty((n, m, m), dtype=torch.float64, device=Y.device)
        _capi.check(_capi.lib().gfc_eval_jacobian(
            self._h, shape, n, n, _ptr(Y), _ptr(d), _ptr(J), self._stream(stream)))
        return d, J

    def eval_lu_solve(self, M, b, shape: int = 0, stream=None):
        """M [n, nS+1, nS+1] (row-major per cell), b [nS+1, n] -> x [nS+1, n]."""
        import torch
        n = b.shape[1]
        m = self.nS + 1
        _check_device("b", b, m, n)
        _check_device("M", M.reshape(n * m, m), n * m, m, device=b.device)
        x = torch.empty_like(b)
        _capi.check(_capi.lib().gfc_eval_lu_solve(
            self._h, shape, n, n, _ptr(M), _ptr(b), _ptr(x), self._stream(stream)))
        return x

    # ------------------------------------------------------------ queries
    # gpuChemistryModel's queries at the current state
    # (gpuChemistryModel.C:146-466) on device tensors: Y [nS+2, ld] SoA,
    # rho [n]; results are written into the given output tensors.
    def _stream(self, stream):
        if stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        return stream

    def _state_checks(self, rho, Y):
        n = rho.shape[0]
        _check_device("Y", Y, self.nE, Y.shape[1] if Y.dim() == 2 else -1, min_cols=n)
        _check_device("rho", rho.reshape(1, -1), 1, n, device=Y.device)
        return n, Y.shape[1]

    def calculate(self, rho, Y, RR, stream=None) -> None:
        """RR [nS, ld] = W_i sum_r nu_ir omega_r (calculate, :146-183)."""
        n, ld = self._state_checks(rho, Y)
        _check_device("RR", RR, self.nS, ld, device=Y.device)
        _capi.check(_capi.lib().gfc_calculate(self._h, n, ld, _ptr(rho), _ptr(Y), _ptr(RR),
                                              self._stream(stream), 0))

    def tc(self, rho, Y, out, stream=None) -> None:
        """out [n] = chemical time scale (tc, :366-443)."""
        n, ld = self._state_checks(rho, Y)
        _check_device("tc", out.reshape(1, -1), 1, n, device=Y.device)
        _capi.check(_capi.lib().gfc_tc(self._h, n, ld, _ptr(rho), _ptr(Y), _ptr(out),
                                       self._stream(stream), 0))

    def reactionRR(self, reaction: int, rho, Y, out, stream=None) -> None:
        """out [n] = omega of one reaction (reactionRR, :186-240)."""
        n, ld = self._state_checks(rho, Y)
        _check_device("rr", out.reshape(1, -1), 1, n, device=Y.device)
        _capi.check(_capi.lib().gfc_reaction_rr(self._h, reaction, n, ld, _ptr(rho),
                                                _ptr(Y), _ptr(out), self._stream(stream), 0))

    def specieReactionRR(self, reaction: int, rho, Y, RR, stream=None) -> None:
        """RR [nS, ld] = W_i nu_ir omega_r of one reaction (specieReactionRR,
        :242-317)."""
        n, ld = self._state_checks(rho, Y)
        _check_device("RR", RR, self.nS, ld, device=Y.device)
        _capi.check(_capi.lib().gfc_specie_reaction_rr(self._h, reaction, n, ld, _ptr(rho),
                                                       _ptr(Y), _ptr(RR),
                                                       self._stream(stream), 0))

    def Qdot(self, RR, out, stream=None) -> None:
        """out [n] = -sum_i hf_i RR_i (Qdot, :445-466)."""
        n = out.shape[0]
        _check_device("RR", RR, self.nS, RR.shape[1] if RR.dim() == 2 else -1, min_cols=n)
        _check_device("Qdot", out.reshape(1, -1), 1, n, device=RR.device)
        _capi.check(_capi.lib().gfc_qdot(self._h, n, RR.shape[1], _ptr(RR), _ptr(out),
                                         self._stream(stream), 0))

    # ------------------------------------------------ reference interface
    def computeYNew(self, deltaT: float, deltaTChemMax: float, deltaTChem, Y):
        """GpuKernelEvaluator::computeYNew (gpuKernelEvaluator.cu:35-63):
        Y is the cell-major [nCells*(nS+2)] host vector; returns (YNew in the
        same layout, deltaTChem new, unclamped as in the reference)."""
        Yc = np.asarray(Y, dtype=np.float64).reshape(-1, self.nE)
        Yn, dt, _, _, _ = self.solve_host(deltaT, Yc.T, deltaTChem)
        return Yn.T.reshape(-1).copy(), dt

    def computeRR(self, deltaT: float, deltaTChemMax: float, rho, deltaTChem, Y):
        """GpuKernelEvaluator::computeRR (gpuKernelEvaluator.cu:65-97):
        returns (RR cell-major [nCells*nS], deltaTChem new clamped to
        deltaTChemMax, deltaTMin before clamping)."""
        Yc = np.asarray(Y, dtype=np.float64).reshape(-1, self.nE)
        _, dt, dtmin, _, rr = self.solve_host(deltaT, Yc.T, deltaTChem,
                                              rho=np.asarray(rho, dtype=np.float64),
                                              deltaTChemMax=deltaTChemMax, want_rr=True)
        return rr.T.reshape(-1).copy(), dt, dtmin
