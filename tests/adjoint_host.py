"""ctypes face of tests/native/libadjhost.so: the library's per-lane physics
and discrete-adjoint code built for the host (test-only), plus helpers that
perturb a mechanism's rate parameters for finite differences against the
oracle."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from gpufoam_amd import _capi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "native", "libadjhost.so")
_dp = C.POINTER(C.c_double)
_VP = C.POINTER(_capi.MechView)
_OP = C.POINTER(_capi.OdeInputs)
_lib = None


def available() -> bool:
    return os.path.exists(LIB)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        L = C.CDLL(LIB)
        sigs = {
            "adjh_vjp_f": [_VP, C.c_double, _dp, _dp, _dp, _dp],
            "adjh_jact": [_VP, C.c_double, _dp, _dp, _dp, _dp, _dp],
            "adjh_jacobian": [_VP, C.c_double, _dp, _dp, _dp],
            "adjh_step": [_VP, _OP, C.c_double, _dp, C.c_double, _dp],
            "adjh_step_reverse": [_VP, _OP, C.c_double, _dp, C.c_double, _dp, _dp],
            "adjh_solve_T": [C.c_int, _dp, _dp],
            "adjh_cell": [_VP, _OP, C.c_double, C.c_double, _dp, _dp, C.c_int, _dp, _dp, _dp],
            "adjh_replay": [_VP, _OP, _dp, C.c_int, _dp],
            "adjh_cell_solve": [_VP, _OP, C.c_double, _dp, _dp, C.POINTER(C.c_int)],
        }
        for name, args in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int if name == "adjh_cell" else None
        _lib = L
    return _lib


def P(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


class PerturbedMech:
    """A mechanism view whose rate parameters are shifted by dtheta
    ([6 nR]: d ln|A|, d beta, d Ta of k, then of kInf for fall-off)."""

    def __init__(self, mech, dtheta=None):
        nR = mech.n_reactions
        self.n_species = mech.n_species
        k = np.ctypeslib.as_array(mech.view.k, shape=(3 * nR,)).copy()
        kI = np.ctypeslib.as_array(mech.view.kInf, shape=(3 * nR,)).copy()
        if dtheta is not None:
            t = np.asarray(dtheta, dtype=np.float64).reshape(nR, 6)
            fo = mech.arrays["type"] >= 2
            k[0::3] *= np.exp(t[:, 0])
            k[1::3] += t[:, 1]
            k[2::3] += t[:, 2]
            kI[0::3] = np.where(fo, kI[0::3] * np.exp(t[:, 3]), kI[0::3])
            kI[1::3] = np.where(fo, kI[1::3] + t[:, 4], kI[1::3])
            kI[2::3] = np.where(fo, kI[2::3] + t[:, 5], kI[2::3])
        self._k, self._kI = k, kI
        v = _capi.MechView.from_buffer_copy(mech.view)
        v.k = k.ctypes.data_as(_dp)
        v.kInf = kI.ctypes.data_as(_dp)
        self.view = v


def theta_scale(mech) -> np.ndarray:
    """Natural perturbation scale of each rate parameter (0 where unused)."""
    nR = mech.n_reactions
    a = mech.arrays
    s = np.zeros((nR, 6))
    s[:, 0] = 1.0
    s[:, 1] = 0.1
    s[:, 2] = np.maximum(np.abs(a["k"][:, 2]), 100.0)
    fo = a["type"] >= 2
    s[fo, 3] = 1.0
    s[fo, 4] = 0.1
    s[fo, 5] = np.maximum(np.abs(a["kInf"][fo, 2]), 100.0)
    return s.reshape(-1)


def cell_adjoint(mech, ode, y0, dtChem, deltaT, ybar, K=1, Treact=0.0):
    """Recording forward + reverse of one cell on the host.  Returns
    (final state, dL/dy0 [nS+1], dL/dtheta [6 nR], step sizes, nAccepted)."""
    y = np.array(y0, dtype=np.float64)
    dt = np.array([dtChem], dtype=np.float64)
    yb = np.array(ybar, dtype=np.float64)
    th = np.zeros(6 * mech.n_reactions)
    hs = np.zeros(ode.maxSteps)
    N = lib().adjh_cell(C.pointer(mech.view), C.byref(ode), Treact, deltaT, P(y), P(dt), K,
                        P(yb), P(th), P(hs))
    return y, yb, th, hs[:max(N, 0)], N


def replay(mech, ode, y0, hs) -> np.ndarray:
    """Re-integrates the step sizes hs from y0 (host build of the lane code);
    `mech` is a Mechanism or PerturbedMech (kept alive during the call)."""
    z = np.array(y0, dtype=np.float64)
    h = np.ascontiguousarray(hs, dtype=np.float64)
    lib().adjh_replay(C.pointer(mech.view), C.byref(ode), P(z), len(h), P(h))
    return z


def cell_solve(mech, ode, y0, dtChem, deltaT):
    """One cell through the library's one-lane solver built for the host.
    Returns (final state, dtChem, [status, attempts, accepted])."""
    y = np.array(y0, dtype=np.float64)
    dt = np.array([dtChem], dtype=np.float64)
    st = np.zeros(3, dtype=np.int32)
    lib().adjh_cell_solve(C.pointer(mech.view), C.byref(ode), deltaT, P(y), P(dt),
                          st.ctypes.data_as(C.POINTER(C.c_int)))
    return y, dt[0], st
