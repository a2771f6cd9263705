"""Writes gpufoam_amd/flops/<mechanism>.json: the algorithmic FP64 work of
one Rosenbrock34 step attempt, from operation counts of the oracle (the
restated reference arithmetic, oracle/libopcount.so) on sample states of
each mechanism's workload.  Counted: +, -, *, / and transcendental calls
(exp, log, log10, pow, sqrt), one each; the transcendentals are also listed
separately.  n = nSpecie+1 is the effective dimension (the reference's
pressure row and column are identically zero).

  F_attempt = F_jac + LU(n) + 4 solve(n) + 2 F_rhs + 40 n   (Rosenbrock34)
  F_accept  = F_rhs   (the derivative at the start of each accepted step)

    python oracle/count_flops.py        (after make -C oracle opcount)
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from gpufoam_amd import _capi  # noqa: E402
from gpufoam_amd.workloads import make_states  # noqa: E402
from oracle_lib import Oracle  # noqa: E402

L = C.CDLL(os.path.join(HERE, "libopcount.so"))
u64 = C.POINTER(C.c_uint64)
dp = C.POINTER(C.c_double)
L.opc_count_derivatives.argtypes = [C.POINTER(_capi.MechView), dp, u64]
L.opc_count_jacobian.argtypes = [C.POINTER(_capi.MechView), dp, u64]
L.opc_count_lu.argtypes = [C.c_int, dp, u64, u64]

WORKLOADS = {"h2": "h2_small", "gri": "gri_production", "syn200": "syn200", "syn100": "syn100"}
C_V = 40  # stage vector operations and the error norm, per component


def counts(fn, *args):
    out = np.zeros(4, dtype=np.uint64)
    fn(*args, out.ctypes.data_as(u64))
    return out.astype(np.int64)


def main(samples: int = 16):
    os.makedirs(os.path.join(ROOT, "gpufoam_amd", "flops"), exist_ok=True)
    for mech, cfg in WORKLOADS.items():
        m, Y0, _, c = make_states(cfg, samples)
        v = C.pointer(m.view)
        nS = m.n_species
        n = nS + 1
        orc = Oracle(m)
        f_rhs, f_jac, lu, sol = [], [], [], []
        for cell in range(samples):
            y = np.ascontiguousarray(Y0[:, cell])
            f_rhs.append(counts(L.opc_count_derivatives, v, y.ctypes.data_as(dp)))
            f_jac.append(counts(L.opc_count_jacobian, v, y.ctypes.data_as(dp)))
            _, J = orc.jacobian(y)
            A = np.ascontiguousarray(np.eye(n) / (0.5 * 1e-7) - J[:n, :n])
            a1, a2 = np.zeros(4, dtype=np.uint64), np.zeros(4, dtype=np.uint64)
            L.opc_count_lu(n, A.ctypes.data_as(dp), a1.ctypes.data_as(u64), a2.ctypes.data_as(u64))
            lu.append(a1.astype(np.int64))
            sol.append(a2.astype(np.int64))
        mean = lambda xs: np.mean(np.array(xs), axis=0)
        tot = lambda x: int(round(float(x.sum())))
        F_rhs, F_jac, LU, SOL = mean(f_rhs), mean(f_jac), mean(lu), mean(sol)
        attempt = tot(F_jac) + tot(LU) + 4 * tot(SOL) + 2 * tot(F_rhs) + C_V * n
        out = {
            "mechanism": mech, "species": nS, "reactions": m.n_reactions, "n": n,
            "states": f"{samples} cells of the {cfg} state law",
            "F_rhs": tot(F_rhs), "F_jac": tot(F_jac), "lu": tot(LU), "solve": tot(SOL),
            "c_v": C_V, "attempt": attempt, "accept": tot(F_rhs),
            "transcendentals": {"F_rhs": int(round(F_rhs[3])), "F_jac": int(round(F_jac[3]))},
            "breakdown": {k: {"add": int(x[0]), "mul": int(x[1]), "div": int(x[2]),
                              "trans": int(x[3])}
                          for k, x in (("F_rhs", F_rhs), ("F_jac", F_jac), ("lu", LU),
                                       ("solve", SOL))},
            "spread_F_rhs": [tot(min(f_rhs, key=lambda z: z.sum())),
                             tot(max(f_rhs, key=lambda z: z.sum()))],
        }
        path = os.path.join(ROOT, "gpufoam_amd", "flops", f"{mech}.json")
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
        print(mech, {k: out[k] for k in ("F_rhs", "F_jac", "lu", "solve", "attempt")})


if __name__ == "__main__":
    main()
