"""The C++ interface (include/gpufoam_amd.hpp, the FoamGpu::GpuKernelEvaluator
mirror) driven from a compiled C++ program the way gpuChemistryModel drives
the reference evaluator: cell-major Y, computeRR and computeYNew, compared
with the oracle's computeRR semantics (gpuKernelEvaluator.cu:65-97)."""
import os
import subprocess

import numpy as np
import pytest

from gpufoam_amd.mechanism import MECH_DIR
from gpufoam_amd.workloads import make_states

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "native", "evaluator_cpp")
PROPS = """odeCoeffs { solver Rosenbrock34; absTol 1e-10; relTol 1e-6; }
Treact 0;
"""


def test_cpp_program_is_built():
    assert os.path.exists(EXE), "build() compiles tests/native/evaluator_cpp"
    r = subprocess.run([EXE], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
def test_cpp_evaluator_matches_oracle(tmp_path):
    from oracle_lib import Oracle
    from gpufoam_amd import ode_inputs
    m, Y0, dt0, c = make_states("gri_production", 64)
    nS, nE, n = m.n_species, m.n_species + 2, 64
    rho = np.linspace(0.2, 1.2, n)
    dtmax = 3e-7
    (tmp_path / "props").write_text(PROPS)
    with open(tmp_path / "in.bin", "wb") as f:
        np.array([n, nE], dtype=np.int32).tofile(f)
        np.array([c["deltaT"], dtmax]).tofile(f)
        np.ascontiguousarray(Y0.T).tofile(f)
        rho.tofile(f)
        dt0.tofile(f)
    mdir = os.path.join(MECH_DIR, "gri")
    r = subprocess.run([EXE, os.path.join(mdir, "thermo"), os.path.join(mdir, "reactions"),
                        str(tmp_path / "props"), str(tmp_path / "in.bin"),
                        str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "size error: status 1" in r.stdout
    out = np.fromfile(tmp_path / "out.bin")
    RR = out[:n * nS].reshape(n, nS)
    dtRR = out[n * nS:n * nS + n]
    dtMin = out[n * nS + n]
    YNew = out[n * nS + n + 1:n * nS + n + 1 + n * nE].reshape(n, nE)
    dtY = out[n * nS + n + 1 + n * nE:]
    ode = ode_inputs("Rosenbrock34", relTol=1e-6, absTol=1e-10)
    ref = Oracle(m).solve_cells(ode, Y0, c["deltaT"], dt0, rho=rho, want_rr=True,
                                deltaTChemMax=dtmax)
    Yo = ref[0]
    assert np.all(np.abs(YNew.T[:nS] - Yo[:nS]) <= 1e-6 * np.abs(Yo[:nS]) + 1e-10)
    assert np.all(np.abs(YNew.T[nS] - Yo[nS]) <= 1e-8 * Yo[nS])
    bound = rho * (1e-6 * np.abs(Yo[:nS]) + 1e-10) / c["deltaT"]
    assert np.all(np.abs(RR.T - ref[4]) <= 1.01 * bound)
    # computeRR clamps the chemistry time step, computeYNew does not
    assert np.all(dtRR <= dtmax) and np.allclose(dtRR, np.minimum(dtY, dtmax))
    assert np.isclose(dtMin, dtY.min())
    assert np.isclose(dtMin, ref[2], rtol=1e-6)
