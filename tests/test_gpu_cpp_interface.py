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


MODEL_EXE = os.path.join(HERE, "native", "chemistry_model_cpp")
MODEL_PROPS = """chemistry on;
initialChemicalTimeStep 1e-7;
maxChemicalTimeStep 4e-7;
odeCoeffs { solver Rosenbrock34; absTol 1e-10; relTol 1e-6; }
"""


def test_cpp_model_program_is_built():
    assert os.path.exists(MODEL_EXE), "build() compiles tests/native/chemistry_model_cpp"


def _run_model(tmp_path, m, Y0, rho, deltaT, reaction, props):
    nS, n = m.n_species, Y0.shape[1]
    (tmp_path / "props").write_text(props)
    with open(tmp_path / "in.bin", "wb") as f:
        np.array([n, nS, reaction], dtype=np.int32).tofile(f)
        np.array([deltaT]).tofile(f)
        np.ascontiguousarray(Y0).tofile(f)
        rho.tofile(f)
    mdir = os.path.join(MECH_DIR, m.name)
    r = subprocess.run([MODEL_EXE, os.path.join(mdir, "thermo"), os.path.join(mdir, "reactions"),
                        str(tmp_path / "props"), str(tmp_path / "in.bin"),
                        str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = np.fromfile(tmp_path / "out.bin")
    sizes = [1, nS * n, n, n, nS * n, n, n, nS * n]
    parts, o = [], 0
    for s in sizes:
        parts.append(out[o:o + s])
        o += s
    assert o == out.size
    ret, RR, dt, qdot, RRc, tc, rr, srr = parts
    return (ret[0], RR.reshape(nS, n), dt, qdot, RRc.reshape(nS, n), tc, rr,
            srr.reshape(nS, n), r.stdout)


@pytest.mark.gpu
@pytest.mark.parametrize("deltaT", [1e-9, 1e-5])
def test_cpp_chemistry_model_matches_oracle(tmp_path, deltaT):
    """ChemistryModel::solve(deltaT) returns min(deltaTMin, 2 deltaT)
    (gpuChemistryModel.C:320-326) with deltaTMin taken before the
    maxChemicalTimeStep clamp (:336-364); solve(scalarField) throws
    std::logic_error (:328-333); the queries match the oracle."""
    from oracle_lib import Oracle
    from gpufoam_amd import ode_inputs
    m, Y0, dt0, c = make_states("h2_small", 64)
    nS = m.n_species
    rho = np.linspace(0.2, 1.2, 64)
    r0 = 5
    ret, RR, dt, qdot, RRc, tc, rr, srr, stdout = _run_model(tmp_path, m, Y0, rho, deltaT, r0,
                                                              MODEL_PROPS)
    assert "lts: status 12" in stdout and "LTS time step not yet supported" in stdout
    o = Oracle(m)
    ode = ode_inputs("Rosenbrock34", relTol=1e-6, absTol=1e-10)
    ref = o.solve_cells(ode, Y0, deltaT, np.full(64, 1e-7), rho=rho, want_rr=True,
                        deltaTChemMax=4e-7)
    print(f"deltaT {deltaT}: solve returned {ret:.6g}, oracle deltaTMin {ref[2]:.6g}")
    assert np.isclose(ret, min(ref[2], 2 * deltaT), rtol=1e-6)
    if deltaT == 1e-9:
        assert ret == 2 * deltaT
    else:
        assert ret < 2 * deltaT
    assert np.all(dt <= 4e-7)
    bound = rho * (1e-6 * np.abs(ref[0][:nS]) + 1e-10) / deltaT
    assert np.all(np.abs(RR - ref[4]) <= 1.01 * bound)
    hf = np.array([o.species_thermo(i, 1e5, 1000.0)[2] for i in range(nS)])
    assert np.allclose(qdot, o.qdot(RR), rtol=1e-12,
                       atol=1e-13 * np.abs(hf[:, None] * RR).sum(axis=0).max())
    scale = sum(np.abs(o.specie_reaction_rr(r, rho, Y0)) for r in range(m.n_reactions))
    assert np.all(np.abs(RRc - o.calculate(rho, Y0)) <= 1e-12 * scale)
    assert np.allclose(tc, o.tc(rho, Y0), rtol=1e-11, atol=0)
    assert np.allclose(rr, o.reaction_rr(r0, rho, Y0), rtol=1e-12, atol=1e-300)
    assert np.allclose(srr, o.specie_reaction_rr(r0, rho, Y0), rtol=1e-12, atol=1e-300)


@pytest.mark.gpu
def test_cpp_chemistry_model_switched_off(tmp_path):
    """chemistry off: solve computes nothing and returns min(great, 2 deltaT);
    RR and Qdot stay zero, tc is `small` everywhere (gpuChemistryModel.C:147,
    338, 375-378, 452)."""
    m, Y0, dt0, c = make_states("h2_small", 16)
    rho = np.ones(16)
    props = MODEL_PROPS.replace("chemistry on;", "chemistry off;")
    ret, RR, dt, qdot, RRc, tc, rr, srr, stdout = _run_model(tmp_path, m, Y0, rho, 1e-5, 0,
                                                              props)
    assert ret == 2e-5
    assert not RR.any() and not qdot.any() and not RRc.any() and not rr.any() and not srr.any()
    assert np.all(tc == np.finfo(np.float64).eps)
    assert np.all(dt == 1e-7)
