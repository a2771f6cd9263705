"""The debug bounds-check build on the CPU: the check helper's semantics
(host build of gpufoam_amd/csrc/bounds.hpp with and without
GFC_BOUNDS_CHECK), and the bounds-check library exporting the same C ABI
as the default one and reporting itself."""
from __future__ import annotations

import ctypes as C
import os

import pytest

from gpufoam_amd import _capi

NATIVE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native")


def _helper(name):
    path = os.path.join(NATIVE, name)
    if not os.path.exists(path):
        pytest.fail(f"{path} is missing: run `make -C tests/native`")
    L = C.CDLL(path)
    L.bh_bound.restype = C.c_int
    L.bh_bound.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_uint), C.c_uint]
    L.bh_checked.restype = C.c_int
    return L


def test_checked_helper_flags_and_clamps():
    L = _helper("libboundshost.so")
    assert L.bh_checked() == 1
    w = C.c_uint(0)
    # in range: the index itself, no flag
    for i, n in ((0, 1), (5, 6), (63, 64)):
        assert L.bh_bound(i, n, C.byref(w), 4) == i
    assert w.value == 0
    # out of range: 0 and the check's bit, bits accumulate
    assert L.bh_bound(6, 6, C.byref(w), 4) == 0
    assert w.value == 4
    assert L.bh_bound(-1, 6, C.byref(w), 16) == 0
    assert w.value == 4 | 16
    # no word: still clamped
    assert L.bh_bound(100, 6, None, 1) == 0


def test_default_helper_is_the_identity():
    L = _helper("libboundsoff.so")
    assert L.bh_checked() == 0
    w = C.c_uint(0)
    assert L.bh_bound(6, 6, C.byref(w), 4) == 6
    assert L.bh_bound(-1, 6, C.byref(w), 4) == -1
    assert w.value == 0


def test_bounds_library_exports_the_abi_and_reports_itself():
    if not os.path.exists(_capi.BOUNDS_LIB_PATH):
        pytest.fail(f"{_capi.BOUNDS_LIB_PATH} is missing: build it with gpufoam_amd.build")
    B = C.CDLL(_capi.BOUNDS_LIB_PATH)
    for name in _capi.exported_symbols():
        assert hasattr(B, name), name
    B.gfc_bounds_checked.restype = C.c_int
    assert B.gfc_bounds_checked() == 1
    assert _capi.lib().gfc_bounds_checked() == 0
    assert _capi.STATUS_NAMES[13] == "BOUNDS"
