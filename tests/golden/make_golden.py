"""Writes tests/golden/reference_vectors.npz: input/output vectors of the
reference's own chemistry code (gpuFoam's header-only device code built for
the host, oracle/_ref/libgfref.so) on the cases of its test suite
(gpu_chemistry/unittest/tests/Test-gpuChemistry.C) and on small samples of
the benchmark states.

Run in a checkout where oracle/_ref is built (make -C oracle):

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from gpufoam_amd import Mechanism, ode_inputs  # noqa: E402
from gpufoam_amd.workloads import make_states  # noqa: E402
from oracle_lib import Reference, lu_decompose, lu_solve  # noqa: E402

# Test-gpuChemistry.C conditions (mechanisms.H:25-104)
P_INF = {"h2": 202650.0, "gri": 1.36789e6}
INITIAL = {"h2": {"H2": 0.2, "O2": 0.2, "N2": 0.6},
           "gri": {"CH4": 0.2, "O2": 0.5, "N2": 0.3}}


def fill_linear(n):
    """test_utilities.H:23-34"""
    dx = 1.0 / n
    return np.array([dx * i + dx for i in range(n)])


def main():
    out = {}
    for name in ("h2", "gri"):
        m = Mechanism.load(name)
        R = Reference(m)
        nS, nE = m.n_species, m.n_species + 2
        p = P_INF[name]
        for T in (1000.0, 2500.0):
            out[f"{name}_thermo_T{int(T)}"] = np.array([R.species_thermo(i, p, T) for i in range(nS)])
            out[f"{name}_Kc_T{int(T)}"] = np.array([R.Kc(r, p, T) for r in range(m.n_reactions)])
        c = fill_linear(nS)
        out[f"{name}_rates_T1000"] = np.array([R.rates(r, p, 1000.0, c) for r in range(m.n_reactions)])
        pr = [R.parameters(r, p, 1000.0, c) for r in range(m.n_reactions)]
        out[f"{name}_params_omega_dwdT"] = np.array([x[0] for x in pr])
        out[f"{name}_params_dCf"] = np.array([x[1] for x in pr])
        out[f"{name}_params_dCr"] = np.array([x[2] for x in pr])
        out[f"{name}_params_ddc"] = np.array(
            [x[3] if m.arrays["type"][r] != 0 else np.zeros(nS) for r, x in enumerate(pr)])
        y = fill_linear(nE)
        y[nS], y[nS + 1] = 1000.0, p
        out[f"{name}_y_linear"] = y
        out[f"{name}_derivatives"] = R.derivatives(y)
        d, J = R.jacobian(y)
        out[f"{name}_jacobian_dfdx"] = d
        out[f"{name}_jacobian_J"] = J
        y0 = np.zeros(nE)
        for s, v in INITIAL[name].items():
            y0[m.index(s)] = v
        y0[nS], y0[nS + 1] = 1000.0, p
        out[f"{name}_ode_y0"] = y0
        for solver in ("Rosenbrock12", "Rosenbrock23", "Rosenbrock34"):
            ode = ode_inputs(solver, absTol=1e-12, relTol=0.1)
            yy, dx, st = R.ode_solve(ode, 0.0, 1e-6, y0, 1e-7)
            assert st == 0
            out[f"{name}_ode_{solver}_y"] = yy
            out[f"{name}_ode_{solver}_dxTry"] = np.array([dx])
    # batch of benchmark-law states through singleCellSolver + computeRR
    for cfg, n in (("h2_small", 64), ("gri_production", 32)):
        m, Y0, dt0, c = make_states(cfg, n)
        R = Reference(m)
        ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
        rho = np.linspace(0.2, 1.2, n)
        Yn, dt, dtmin, rr = R.solve_cells(ode, Y0, c["deltaT"], dt0, rho=rho,
                                          deltaTChemMax=5e-7, want_rr=True)
        out[f"{cfg}_Y0"] = Y0
        out[f"{cfg}_dt0"] = dt0
        out[f"{cfg}_rho"] = rho
        out[f"{cfg}_Ynew"] = Yn
        out[f"{cfg}_dtChem"] = dt
        out[f"{cfg}_dtMin"] = np.array([dtmin])
        out[f"{cfg}_RR"] = rr
    # LU (Test-gpuChemistry.C:481-557: sizes 3..49, source = 1)
    rng = np.random.default_rng(7)
    for n in (3, 10, 27, 49):
        A = rng.random((n, n))
        lu, piv = lu_decompose(A, "ref")
        out[f"lu{n}_A"] = A
        out[f"lu{n}_LU"] = lu
        out[f"lu{n}_piv"] = piv
        out[f"lu{n}_x"] = lu_solve(lu, piv, np.ones(n), "ref")
    np.savez_compressed(os.path.join(HERE, "reference_vectors.npz"), **out)
    print(f"{len(out)} arrays written")


if __name__ == "__main__":
    main()
