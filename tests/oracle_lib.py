"""ctypes access to the CPU oracle (oracle/liboracle.so) and to the
reference's own chemistry headers built for the host (oracle/_ref/libgfref.so).

Test infrastructure: only tests/, bench.py's CPU baseline and
__graft_entry__.smoke() use this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from gpufoam_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libgfref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p
_MV = C.POINTER(_capi.MechView)
_OI = C.POINTER(_capi.OdeInputs)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


_oracle = None
_ref = None


def oracle() -> C.CDLL:
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise ImportError(f"{ORACLE_SO} missing: run `make -C oracle`")
        L = C.CDLL(ORACLE_SO)
        L.orc_workspace_doubles.restype = C.c_long
        L.orc_workspace_doubles.argtypes = [C.c_int]
        L.orc_species_thermo.argtypes = [_MV, C.c_int, C.c_double, C.c_double, _dp]
        L.orc_reaction_Kc.argtypes = [_MV, C.c_int, C.c_double, C.c_double, _dp]
        L.orc_reaction_rates.argtypes = [_MV, C.c_int, C.c_double, C.c_double, _dp, _dp]
        L.orc_reaction_parameters.argtypes = [_MV, C.c_int, C.c_double, C.c_double, _dp,
                                              _dp, _dp, _dp, _dp]
        L.orc_derivatives.argtypes = [_MV, _dp, _dp, _dp]
        L.orc_jacobian.argtypes = [_MV, _dp, _dp, _dp, _dp]
        L.orc_lu_decompose.argtypes = [C.c_int, _dp, _ip, _dp]
        L.orc_lu_backsubstitute.argtypes = [C.c_int, _dp, _ip, _dp]
        L.orc_rosenbrock_step.restype = C.c_double
        L.orc_rosenbrock_step.argtypes = [_MV, _OI, _dp, _dp, C.c_double, _dp, _dp]
        L.orc_ode_solve.restype = C.c_int
        L.orc_ode_solve.argtypes = [_MV, _OI, C.c_double, C.c_double, _dp, _dp, _dp, _ip]
        L.orc_solve_cells.restype = C.c_int
        L.orc_solve_cells.argtypes = [_MV, _OI, C.c_double, C.c_double, C.c_double,
                                      C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp,
                                      _ip, C.c_int]
        L.orc_reaction_omega_fr.argtypes = [_MV, C.c_int, C.c_double, C.c_double, _dp, _dp]
        for name in ("orc_calculate", "orc_tc"):
            getattr(L, name).argtypes = [_MV, C.c_int, C.c_int, _dp, _dp, _dp]
        for name in ("orc_reaction_rr", "orc_specie_reaction_rr"):
            getattr(L, name).argtypes = [_MV, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_qdot.argtypes = [_MV, C.c_int, C.c_int, _dp, _dp]
        _oracle = L
    return _oracle


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref() -> C.CDLL:
    global _ref
    if _ref is None:
        L = C.CDLL(REF_SO)
        L.gfr_create.restype = _vp
        L.gfr_create.argtypes = [_MV, _dp, _dp]
        L.gfr_destroy.argtypes = [_vp]
        L.gfr_composite_thermo.argtypes = [_vp, C.c_int, _dp]
        L.gfr_species_thermo.argtypes = [_vp, C.c_int, C.c_double, C.c_double, _dp]
        L.gfr_reaction_Kc.argtypes = [_vp, C.c_int, C.c_double, C.c_double, _dp]
        L.gfr_reaction_rates.argtypes = [_vp, C.c_int, C.c_double, C.c_double, _dp, _dp]
        L.gfr_reaction_parameters.argtypes = [_vp, C.c_int, C.c_double, C.c_double, _dp,
                                              _dp, _dp, _dp, _dp]
        L.gfr_derivatives.argtypes = [_vp, _dp, _dp]
        L.gfr_jacobian.argtypes = [_vp, _dp, _dp, _dp]
        L.gfr_lu_decompose.argtypes = [C.c_int, _dp, _ip]
        L.gfr_lu_backsubstitute.argtypes = [C.c_int, _dp, _ip, _dp]
        L.gfr_ode_solve.restype = C.c_int
        L.gfr_ode_solve.argtypes = [_vp, _OI, C.c_double, C.c_double, _dp, _dp]
        L.gfr_solve_cells.restype = C.c_int
        L.gfr_solve_cells.argtypes = [_vp, _OI, C.c_double, C.c_double, C.c_double,
                                      C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp,
                                      C.c_int]
        _ref = L
    return _ref


class Oracle:
    """Python face of the C oracle for one mechanism."""

    def __init__(self, mech):
        self.m = mech
        self.L = oracle()
        self.v = C.pointer(mech.view)
        self.nS = mech.n_species
        self.nE = self.nS + 2
        self.work = np.zeros(self.L.orc_workspace_doubles(self.nS) + 16)

    def species_thermo(self, i, p, T):
        o = np.zeros(6)
        self.L.orc_species_thermo(self.v, i, p, T, dptr(o))
        return o

    def Kc(self, r, p, T):
        o = np.zeros(2)
        self.L.orc_reaction_Kc(self.v, r, p, T, dptr(o))
        return o

    def rates(self, r, p, T, c):
        o = np.zeros(3)
        self.L.orc_reaction_rates(self.v, r, p, T, dptr(np.ascontiguousarray(c)), dptr(o))
        return o

    def parameters(self, r, p, T, c):
        o, dCf, dCr, ddc = np.zeros(2), np.zeros(10), np.zeros(10), np.zeros(self.nE)
        self.L.orc_reaction_parameters(self.v, r, p, T, dptr(np.ascontiguousarray(c)),
                                       dptr(o), dptr(dCf), dptr(dCr), dptr(ddc))
        return o, dCf, dCr, ddc[: self.nS]

    def derivatives(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        d = np.zeros(self.nE)
        self.L.orc_derivatives(self.v, dptr(y), dptr(d), dptr(self.work))
        return d

    def jacobian(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        d = np.zeros(self.nE)
        J = np.zeros((self.nE, self.nE))
        self.L.orc_jacobian(self.v, dptr(y), dptr(d), dptr(J), dptr(self.work))
        return d, J

    def step(self, ode, y0, dx):
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        dydx0 = self.derivatives(y0)
        y = np.zeros(self.nE)
        err = self.L.orc_rosenbrock_step(self.v, C.byref(ode), dptr(y0), dptr(dydx0),
                                         dx, dptr(y), dptr(self.work))
        return y, err

    def ode_solve(self, ode, x0, x1, y, dxTry):
        y = np.array(y, dtype=np.float64)
        dx = np.array([dxTry], dtype=np.float64)
        stats = np.zeros(2, dtype=np.int32)
        st = self.L.orc_ode_solve(self.v, C.byref(ode), x0, x1, dptr(y), dptr(dx),
                                  dptr(self.work), iptr(stats))
        return y, dx[0], st, stats

    # chemistry queries (gpuChemistryModel.C:146-466); Y [nS+2, n] SoA
    def omega_fr(self, r, p, T, c):
        out = np.zeros(2)
        self.L.orc_reaction_omega_fr(self.v, r, p, T, dptr(np.ascontiguousarray(c, float)),
                                     dptr(out))
        return out

    def calculate(self, rho, Y):
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        n = Y.shape[1]
        RR = np.zeros((self.nS, n))
        self.L.orc_calculate(self.v, n, n, dptr(np.ascontiguousarray(rho, float)), dptr(Y),
                             dptr(RR))
        return RR

    def tc(self, rho, Y):
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        n = Y.shape[1]
        out = np.zeros(n)
        self.L.orc_tc(self.v, n, n, dptr(np.ascontiguousarray(rho, float)), dptr(Y),
                      dptr(out))
        return out

    def reaction_rr(self, r, rho, Y):
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        n = Y.shape[1]
        out = np.zeros(n)
        self.L.orc_reaction_rr(self.v, r, n, n, dptr(np.ascontiguousarray(rho, float)),
                               dptr(Y), dptr(out))
        return out

    def specie_reaction_rr(self, r, rho, Y):
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        n = Y.shape[1]
        RR = np.zeros((self.nS, n))
        self.L.orc_specie_reaction_rr(self.v, r, n, n, dptr(np.ascontiguousarray(rho, float)),
                                      dptr(Y), dptr(RR))
        return RR

    def qdot(self, RR):
        RR = np.ascontiguousarray(RR, dtype=np.float64)
        n = RR.shape[1]
        out = np.zeros(n)
        self.L.orc_qdot(self.v, n, n, dptr(RR), dptr(out))
        return out

    def solve_cells(self, ode, Y0, deltaT, dtChem, rho=None, Treact=0.0,
                    deltaTChemMax=1e300, threads=0, want_rr=False):
        """Y0: [nE, nCells] SoA.  Returns (Ynew, dtChem, dtMin, stats, RR)."""
        Y0 = np.ascontiguousarray(Y0, dtype=np.float64)
        n = Y0.shape[1]
        Yn = np.zeros_like(Y0)
        dt = np.array(dtChem, dtype=np.float64).copy()
        dtmin = np.zeros(1)
        stats = np.zeros((3, n), dtype=np.int32)
        rr = np.zeros((self.nS, n)) if want_rr else None
        rho_a = np.ascontiguousarray(rho if rho is not None else np.ones(n), dtype=np.float64)
        self.L.orc_solve_cells(self.v, C.byref(ode), Treact, deltaT, deltaTChemMax, n, n,
                               dptr(rho_a), dptr(Y0), dptr(Yn),
                               dptr(rr) if rr is not None else None,
                               dptr(dt), dptr(dtmin), iptr(stats), threads)
        return Yn, dt, dtmin[0], stats, rr


class Reference:
    """The reference's header-only code built for the host (oracle/_ref)."""

    def __init__(self, mech):
        self.m = mech
        self.L = ref()
        self.nS = mech.n_species
        self.nE = self.nS + 2
        self._rh = np.ascontiguousarray(mech.raw_high)
        self._rl = np.ascontiguousarray(mech.raw_low)
        self.h = self.L.gfr_create(C.pointer(mech.view), dptr(self._rh), dptr(self._rl))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.gfr_destroy(self.h)
            self.h = None

    def composite(self, r):
        o = np.zeros(19)
        self.L.gfr_composite_thermo(self.h, r, dptr(o))
        return o

    def species_thermo(self, i, p, T):
        o = np.zeros(6)
        self.L.gfr_species_thermo(self.h, i, p, T, dptr(o))
        return o

    def Kc(self, r, p, T):
        o = np.zeros(2)
        self.L.gfr_reaction_Kc(self.h, r, p, T, dptr(o))
        return o

    def rates(self, r, p, T, c):
        o = np.zeros(3)
        self.L.gfr_reaction_rates(self.h, r, p, T, dptr(np.ascontiguousarray(c)), dptr(o))
        return o

    def parameters(self, r, p, T, c):
        o, dCf, dCr, ddc = np.zeros(2), np.zeros(10), np.zeros(10), np.zeros(self.nS)
        self.L.gfr_reaction_parameters(self.h, r, p, T, dptr(np.ascontiguousarray(c)),
                                       dptr(o), dptr(dCf), dptr(dCr), dptr(ddc))
        return o, dCf, dCr, ddc

    def derivatives(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        d = np.zeros(self.nE)
        self.L.gfr_derivatives(self.h, dptr(y), dptr(d))
        return d

    def jacobian(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        d = np.zeros(self.nE)
        J = np.zeros((self.nE, self.nE))
        self.L.gfr_jacobian(self.h, dptr(y), dptr(d), dptr(J))
        return d, J

    def ode_solve(self, ode, x0, x1, y, dxTry):
        y = np.array(y, dtype=np.float64)
        dx = np.array([dxTry], dtype=np.float64)
        st = self.L.gfr_ode_solve(self.h, C.byref(ode), x0, x1, dptr(y), dptr(dx))
        return y, dx[0], st

    def solve_cells(self, ode, Y0, deltaT, dtChem, rho=None, Treact=0.0,
                    deltaTChemMax=1e300, threads=0, want_rr=False):
        Y0 = np.ascontiguousarray(Y0, dtype=np.float64)
        n = Y0.shape[1]
        Yn = np.zeros_like(Y0)
        dt = np.array(dtChem, dtype=np.float64).copy()
        dtmin = np.zeros(1)
        rr = np.zeros((self.nS, n)) if want_rr else None
        rho_a = np.ascontiguousarray(rho if rho is not None else np.ones(n), dtype=np.float64)
        self.L.gfr_solve_cells(self.h, C.byref(ode), Treact, deltaT, deltaTChemMax, n, n,
                               dptr(rho_a), dptr(Y0), dptr(Yn),
                               dptr(rr) if rr is not None else None,
                               dptr(dt), dptr(dtmin), threads)
        return Yn, dt, dtmin[0], rr


def lu_decompose(a, lib="oracle"):
    a = np.array(a, dtype=np.float64, order="C")
    n = a.shape[0]
    piv = np.zeros(n, dtype=np.int32)
    if lib == "oracle":
        vv = np.zeros(n)
        oracle().orc_lu_decompose(n, dptr(a), iptr(piv), dptr(vv))
    else:
        ref().gfr_lu_decompose(n, dptr(a), iptr(piv))
    return a, piv


def lu_solve(lu, piv, b, lib="oracle"):
    b = np.array(b, dtype=np.float64)
    n = lu.shape[0]
    lu = np.ascontiguousarray(lu)
    piv = np.ascontiguousarray(piv, dtype=np.int32)
    if lib == "oracle":
        oracle().orc_lu_backsubstitute(n, dptr(lu), iptr(piv), dptr(b))
    else:
        ref().gfr_lu_backsubstitute(n, dptr(lu), iptr(piv), dptr(b))
    return b
