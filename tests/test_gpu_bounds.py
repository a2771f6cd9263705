"""The debug bounds-check build on the GPU: every kernel family's small case
(tests/bounds_cases.py) runs under the checks without a failed check, and
its results equal the default build's bit for bit (the checks only read
indices).  Both libraries run in child processes, one after the other."""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np
import pytest

from gpufoam_amd import _capi

HERE = os.path.dirname(os.path.abspath(__file__))
pytestmark = pytest.mark.gpu


def _run(lib_path: str, out: str) -> None:
    env = dict(os.environ, GFC_LIBRARY=lib_path)
    r = subprocess.run([sys.executable, "-u", os.path.join(HERE, "bounds_cases.py"), out],
                       env=env, capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr[-2000:]


def test_bounds_checked_build_matches_default_build(tmp_path):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(_capi.BOUNDS_LIB_PATH):
        pytest.fail(f"{_capi.BOUNDS_LIB_PATH} is missing: build it with gpufoam_amd.build")
    default = os.path.join(_capi._HERE, "libgpufoam_amd.so")
    a, b = str(tmp_path / "default.npz"), str(tmp_path / "checked.npz")
    print("case: default build", flush=True)
    _run(default, a)
    print("case: bounds-check build", flush=True)
    _run(_capi.BOUNDS_LIB_PATH, b)
    ra, rb = np.load(a), np.load(b)
    assert int(ra["checked"][0]) == 0 and int(rb["checked"][0]) == 1
    assert sorted(ra.files) == sorted(rb.files)
    for k in ra.files:
        if k == "checked":
            continue
        assert np.array_equal(ra[k], rb[k], equal_nan=True), k
