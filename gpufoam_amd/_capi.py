"""ctypes bindings of the gpufoam_amd C ABI (include/gpufoam_amd.h).

The shared library is built in-tree by ``gpufoam_amd.build.build()`` into
``gpufoam_amd/libgpufoam_amd.so``.  Loading fails loudly when it is missing:
there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# The debug bounds-check build of the same sources (include/gpufoam_amd.h,
# gfc_bounds_checked); the environment variable GFC_LIBRARY selects a
# library file in place of the default one.
BOUNDS_LIB_PATH = os.path.join(_HERE, "libgpufoam_amd_bounds.so")
LIB_PATH = os.environ.get("GFC_LIBRARY") or os.path.join(_HERE, "libgpufoam_amd.so")

MAX_REACTION_WIDTH = 10
NO_INT_EXPONENT = 214748364

GFC_OK = 0
STATUS_NAMES = {
    1: "INVALID_ARGUMENT", 2: "PARSE", 3: "UNSUPPORTED", 4: "MASS_IMBALANCE",
    5: "TOO_LONG", 6: "TOO_MANY_SPECIES", 7: "EFFICIENCIES", 8: "SOLVER_NAME",
    9: "CAPACITY", 10: "HIP", 11: "NOT_BUILT", 12: "LTS", 13: "BOUNDS",
}

RATE_ARRHENIUS, RATE_THIRD_BODY, RATE_LINDEMANN, RATE_TROE = 0, 1, 2, 3
RATE_SRI, RATE_LINDEMANN_CA, RATE_TROE_CA, RATE_SRI_CA = 4, 5, 6, 7
RATE_LANDAU_TELLER, RATE_JANEV, RATE_POWER_SERIES = 8, 9, 10
EXTRA_PARAMS = 9
ROSENBROCK12, ROSENBROCK23, ROSENBROCK34 = 12, 23, 34
SOLVER_NAMES = {"Rosenbrock12": 12, "Rosenbrock23": 23, "Rosenbrock34": 34}

CELL_MAX_STEPS = 0x1
CELL_STEP_UNDERFLOW = 0x2
CELL_SKIPPED = 0x4

HOST_POINTERS = 0x1
BATCHED = 0x2
GRAPH = 0x4
LANE_PER_CELL = 0x8
# factorisation of the workgroup-per-cell kernel
LU_COLUMN, LU_PANEL, LU_PANEL_MFMA = 0, 1, 2
# kernel families of the device-level evaluations
SHAPE_AUTO, SHAPE_WAVE, SHAPE_PAIR, SHAPE_TEAM16, SHAPE_WORKGROUP = 0, 1, 2, 4, 512


class GfcError(RuntimeError):
    """A non-zero gfc_status, with the library's message."""

    def __init__(self, status: int, message: str):
        self.status = status
        super().__init__(f"gfc status {status} ({STATUS_NAMES.get(status, '?')}): {message}")


class OdeInputs(C.Structure):
    _fields_ = [
        ("solver", C.c_int), ("absTol", C.c_double), ("relTol", C.c_double),
        ("maxSteps", C.c_int), ("safeScale", C.c_double), ("alphaInc", C.c_double),
        ("alphaDec", C.c_double), ("minScale", C.c_double), ("maxScale", C.c_double),
    ]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class MechView(C.Structure):
    _fields_ = [
        ("n_species", C.c_int), ("n_reactions", C.c_int),
        ("Tlow", C.c_double), ("Thigh", C.c_double),
        ("W", _dp), ("sTlow", _dp), ("sThigh", _dp), ("sTcommon", _dp),
        ("high", _dp), ("low", _dp),
        ("type", _ip), ("irreversible", _ip), ("k", _dp), ("kInf", _dp),
        ("troe", _dp), ("eff", _dp), ("n_lhs", _ip), ("n_rhs", _ip),
        ("lhs_index", _ip), ("lhs_stoich", _dp), ("lhs_exp", _dp), ("lhs_exp_int", _ip),
        ("rhs_index", _ip), ("rhs_stoich", _dp), ("rhs_exp", _dp), ("rhs_exp_int", _ip),
        ("rY", _dp), ("rW", _dp), ("rTcommon", _dp), ("rhigh", _dp), ("rlow", _dp),
        ("extra", _dp),
    ]


_lib = None


def lib() -> C.CDLL:
    """The loaded product library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        sigs = {
            "gfc_last_error": (C.c_char_p, []),
            "gfc_bounds_checked": (C.c_int, []),
            "gfc_mechanism_from_foam": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(vp)]),
            "gfc_mechanism_destroy": (None, [vp]),
            "gfc_chemkin_to_foam": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_size_t,
                                              C.POINTER(C.c_size_t), C.c_char_p, C.c_size_t,
                                              C.POINTER(C.c_size_t)]),
            "gfc_mechanism_from_chemkin": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(vp)]),
            "gfc_mechanism_view": (C.c_int, [vp, C.POINTER(MechView)]),
            "gfc_mechanism_species_name": (C.c_char_p, [vp, C.c_int]),
            "gfc_mechanism_reaction_equation": (C.c_char_p, [vp, C.c_int]),
            "gfc_mechanism_raw_thermo": (C.c_int, [vp, C.c_int, _dp, _dp]),
            "gfc_mechanism_species_index": (C.c_int, [vp, C.c_char_p]),
            "gfc_ode_inputs_from_foam": (C.c_int, [C.c_char_p, C.POINTER(OdeInputs), _dp]),
            "gfc_solver_from_name": (C.c_int, [C.c_char_p, _ip]),
            "gfc_chemistry_switches_from_foam": (C.c_int, [C.c_char_p, _ip, _dp, _dp]),
            "gfc_ode_inputs_default": (None, [C.c_int, C.POINTER(OdeInputs)]),
            "gfc_evaluator_create": (C.c_int, [vp, C.POINTER(OdeInputs), C.c_double,
                                              C.c_int, C.POINTER(vp)]),
            "gfc_evaluator_destroy": (None, [vp]),
            "gfc_solve": (C.c_int, [vp, C.c_double, C.c_double, C.c_int, C.c_int,
                                    vp, vp, vp, vp, vp, vp, vp, vp, C.c_uint]),
            "gfc_evaluator_set_batched_shape": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
            "gfc_evaluator_get_batched_shape": (C.c_int, [vp, _ip, _ip, _ip]),
            "gfc_evaluator_set_factorization": (C.c_int, [vp, C.c_int]),
            "gfc_evaluator_get_factorization": (C.c_int, [vp, _ip]),
            "gfc_eval_derivatives": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp]),
            "gfc_eval_jacobian": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp]),
            "gfc_eval_lu_solve": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp]),
            "gfc_evaluator_param_count": (C.c_int, [vp]),
            "gfc_evaluator_get_rate_parameters": (C.c_int, [vp, _dp]),
            "gfc_evaluator_set_rate_parameters": (C.c_int, [vp, _dp]),
            "gfc_adjoint_set_checkpoint_interval": (C.c_int, [vp, C.c_int]),
            "gfc_adjoint_checkpoint_interval": (C.c_int, [vp]),
            "gfc_adjoint_workspace_bytes": (C.c_double, [C.c_int, C.c_int, C.c_int, C.c_int]),
            "gfc_adjoint_plan_interval": (C.c_int, [C.c_double, C.c_int, C.c_int, C.c_int]),
            "gfc_adjoint_forward": (C.c_int, [vp, C.c_double, C.c_int, C.c_int, vp, vp, vp,
                                              vp, vp]),
            "gfc_adjoint_backward": (C.c_int, [vp, vp, C.c_int, vp, vp, vp]),
            "gfc_adjoint_replay": (C.c_int, [vp, vp, C.c_int, vp, vp]),
            "gfc_calculate": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, vp, vp, C.c_uint]),
            "gfc_tc": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, vp, vp, C.c_uint]),
            "gfc_reaction_rr": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp,
                                          C.c_uint]),
            "gfc_specie_reaction_rr": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp,
                                                 vp, C.c_uint]),
            "gfc_qdot": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, vp, C.c_uint]),
            "gfc_solve_noalloc": (C.c_int, [vp, C.c_double, C.c_double, C.c_int, C.c_int,
                                            vp, vp, vp, vp, vp, vp, vp, vp]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols() -> list[str]:
    """Names declared in include/gpufoam_amd.h (checked by the CPU tests)."""
    return [
        "gfc_last_error", "gfc_bounds_checked", "gfc_mechanism_from_foam",
        "gfc_mechanism_destroy",
        "gfc_mechanism_view", "gfc_mechanism_species_name",
        "gfc_mechanism_reaction_equation", "gfc_mechanism_raw_thermo",
        "gfc_mechanism_species_index", "gfc_ode_inputs_from_foam",
        "gfc_solver_from_name", "gfc_ode_inputs_default", "gfc_evaluator_create",
        "gfc_evaluator_destroy", "gfc_solve", "gfc_evaluator_set_batched_shape",
        "gfc_evaluator_get_batched_shape", "gfc_evaluator_set_factorization",
        "gfc_evaluator_get_factorization", "gfc_eval_derivatives", "gfc_eval_jacobian",
        "gfc_eval_lu_solve", "gfc_evaluator_param_count",
        "gfc_evaluator_get_rate_parameters", "gfc_evaluator_set_rate_parameters",
        "gfc_adjoint_set_checkpoint_interval", "gfc_adjoint_checkpoint_interval",
        "gfc_adjoint_workspace_bytes", "gfc_adjoint_plan_interval",
        "gfc_adjoint_forward", "gfc_adjoint_backward", "gfc_adjoint_replay",
        "gfc_calculate", "gfc_tc", "gfc_reaction_rr", "gfc_specie_reaction_rr", "gfc_qdot",
        "gfc_chemistry_switches_from_foam", "gfc_solve_noalloc",
        "gfc_chemkin_to_foam", "gfc_mechanism_from_chemkin",
    ]


def check(status: int) -> None:
    if status != GFC_OK:
        raise GfcError(status, lib().gfc_last_error().decode())
