"""Algorithmic FP64 work of one Rosenbrock step attempt, per mechanism.

The per-mechanism constants come from operation counts of the CPU
restatement of gpuFoam's per-cell path (the reference arithmetic of
gpu_chemistry/src/gpuOde, gpuReaction, gpuThermo): gpufoam_amd/flops/<mech>.json,
written by oracle/count_flops.py.  One flop per +, -, *, / and per
transcendental call (exp, log, log10, pow, sqrt); their instruction cost on
the GPU is far higher, so this is a conservative count of the useful work.
n = nSpecie + 1 is the effective dimension (the pressure row and column of
the reference's Jacobian are identically zero).

  F_attempt = F_jac + LU(n) + s * solve(n) + r * F_rhs + c_v * n
  F_accept  = F_rhs  (the derivative at y0 of adaptiveSolve)

with s = 4 back-substitutions and r = 2 extra right-hand sides for
Rosenbrock34 (3 and 1 for Rosenbrock23, 2 and 1 for Rosenbrock12).  The
reference evaluates the Jacobian on every attempt; kernels that reuse it
across rejected attempts still report the reference's work.  Mechanisms
without a counted file use a model of the same terms (estimate()).
"""
from __future__ import annotations

import json
import os

import numpy as np

FLOPS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "flops")
_SOLVES = {2: 2, 3: 3, 4: 4}
_RHS = {2: 1, 3: 1, 4: 2}


def counted(mech) -> dict | None:
    path = os.path.join(FLOPS_DIR, f"{mech.name}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if d["species"] != mech.n_species or d["reactions"] != mech.n_reactions:
        return None
    return d


def estimate(mech) -> dict:
    """Term model for mechanisms without counted constants."""
    a = mech.arrays
    nS = mech.n_species
    n = nS + 1
    F_rhs = nS * 40 + 10 + nS * 3 + 4
    F_jac = nS * 55 + 10 + nS * nS * 9 + nS * 14 + 6
    for r in range(mech.n_reactions):
        t = int(a["type"][r])
        nl, nr = int(a["n_lhs"][r]), int(a["n_rhs"][r])
        w = nl + nr
        f = 8 + (2 * nS if t >= 1 else 0) + (14 if t >= 2 else 0) + (30 if t == 3 else 0)
        f += (25 if not a["irreversible"][r] else 0) + 4 * w + 3
        F_rhs += f
        F_jac += f + 6 + (nl * nl + nr * nr) * 2 + 2 * w * w + 2 * w
        F_jac += (12 if not a["irreversible"][r] else 0) + ((2 * nS + 2 * w * nS) if t >= 1 else 0)
    return {"F_rhs": F_rhs, "F_jac": F_jac, "lu": int(np.ceil(2 * n ** 3 / 3)) + n * n // 2 + n,
            "solve": 2 * n * n, "c_v": 40, "n": n}


def attempt_flops(mech, stages: int = 4) -> dict:
    d = counted(mech) or estimate(mech)
    n = mech.n_species + 1
    att = (d["F_jac"] + d["lu"] + _SOLVES[stages] * d["solve"] + _RHS[stages] * d["F_rhs"]
           + d["c_v"] * n)
    return {"F_rhs": d["F_rhs"], "F_jac": d["F_jac"], "lu": d["lu"],
            "solves": _SOLVES[stages] * d["solve"], "attempt": att, "accept": d["F_rhs"],
            "source": "counted" if counted(mech) else "estimate"}


def work_flops(mech, attempts: np.ndarray, accepted: np.ndarray, stages: int = 4) -> float:
    w = attempt_flops(mech, stages)
    return float(w["attempt"] * np.sum(attempts, dtype=np.float64) +
                 w["accept"] * np.sum(accepted, dtype=np.float64))
