"""Algorithmic FP64 work of one Rosenbrock step attempt, per mechanism.

The count follows the operations of gpuFoam's per-cell path as written
(gpu_chemistry/src/gpuOde, gpuReaction, gpuThermo), one flop per add, sub,
mul, div, fma-half and per transcendental call (exp, log, pow, log10 count
1 each; their instruction cost on the GPU is far higher, so this is a
conservative lower bound of the useful work).  n = nSpecie + 1 is the
effective dimension (the pressure row/column of the reference's Jacobian is
identically zero).

  F_attempt = F_jac + n^2 + 2n^3/3 + stages * 2n^2 + r * F_rhs + 40 n
  (r = 2 RHS evaluations per attempt for Rosenbrock34, 1 for 12 and 23)
  F_accept  = F_rhs (the derivative at y0 of adaptiveSolve)

The reference evaluates the Jacobian on every attempt; kernels that reuse it
across rejected attempts still report the reference's work.
"""
from __future__ import annotations

import numpy as np


def _reaction_rhs_flops(a, r, nS) -> int:
    t = int(a["type"][r])
    nl, nr = int(a["n_lhs"][r]), int(a["n_rhs"][r])
    f = 6 + 2  # Arrhenius: pow, exp, 2 mul, div, ...
    if t >= 1:
        f += 2 * nS  # third-body sum
    if t >= 2:
        f += 8 + 6  # second Arrhenius, Pr, blend
    if t == 3:
        f += 30  # Troe: 3 exp, 2 log10, pow and ~24 arithmetic
    if not a["irreversible"][r]:
        f += 25  # Kc: log, Gstd polynomial (~15), exp, pow, div
    f += 2 * (nl + nr) + 3  # concentration powers (integer) and omega
    f += 2 * (nl + nr)      # accumulation into dN/dt
    return f


def rhs_flops(mech) -> int:
    a = mech.arrays
    nS = mech.n_species
    f = nS * 40 + 10  # mixture: Cp/Ha polynomials, densities, concentrations
    f += sum(_reaction_rhs_flops(a, r, nS) for r in range(mech.n_reactions))
    f += nS * 3 + 4   # dY/dt scaling and dT/dt
    return f


def jacobian_flops(mech) -> int:
    a = mech.arrays
    nS = mech.n_species
    f = nS * 55 + 10  # mixture with dCp/dT and alphav
    for r in range(mech.n_reactions):
        nl, nr = int(a["n_lhs"][r]), int(a["n_rhs"][r])
        w = nl + nr
        f += _reaction_rhs_flops(a, r, nS) + 6           # rate, dk/dT
        f += (nl * nl + nr * nr) * 2                       # dCf/dc, dCr/dc
        f += 2 * (nl * w + nr * w) + 2 * w                 # sparse entries, dT column
        if not a["irreversible"][r]:
            f += 12                                        # dKc/dT, dkr/dT
        if a["type"][r] >= 1:
            f += 2 * nS + 2 * w * nS                       # dense third-body rows
    f += nS * nS * 7 + nS * 6       # ddY_dcdY
    f += nS * nS * 2 + nS * 4       # ddT_dtdY
    f += nS * 4 + 6                 # ddT_dtdT
    return f


def attempt_flops(mech, stages: int = 4) -> dict:
    n = mech.n_species + 1
    F_f = rhs_flops(mech)
    F_J = jacobian_flops(mech)
    lu = int(np.ceil(2 * n ** 3 / 3)) + n * n // 2 + n
    solves = stages * 2 * n * n
    n_rhs = {2: 1, 3: 1, 4: 2}[stages]
    vec = 40 * n
    return {"F_rhs": F_f, "F_jac": F_J, "lu": lu, "solves": solves,
            "attempt": F_J + n * n + lu + solves + n_rhs * F_f + vec,
            "accept": F_f}


def work_flops(mech, attempts: np.ndarray, accepted: np.ndarray, stages: int = 4) -> float:
    w = attempt_flops(mech, stages)
    return float(w["attempt"] * attempts.sum() + w["accept"] * accepted.sum())
