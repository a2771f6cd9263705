"""The command-line contract of bench.py that holds without a GPU: more than
one GPU is refused with a message before any device work, and the options
that select the workload and the adjoint-cost sample are listed."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = _bench("--gpus", "2")
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_options_are_listed():
    r = _bench("--help")
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--config", "--cells", "--adjoint-cells",
                "--adjoint-config", "--adjoint-dt", "--autotune"):
        assert opt in r.stdout, opt
