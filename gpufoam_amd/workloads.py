"""Synthetic cell states for the benchmark and parity configurations.

Every generator is seeded (numpy.random.default_rng) and returns the state in
the library's structure-of-arrays layout: Y0[k, cell] for k < nSpecie mass
fractions, Y0[nSpecie] temperature, Y0[nSpecie+1] pressure, plus the initial
chemical time step per cell.

Configurations (one CFD time step each):
  h2_small       10 species / 27 reactions, 4096 cells, Rosenbrock34
                 relTol 1e-6 absTol 1e-10, deltaT 1e-5
  gri_production 53 species / 325 reactions, 262144 cells, Rosenbrock34,
                 deltaT 1e-6; half unburnt CH4-air, half near-equilibrium
                 products
  gri_batching   1048576 cells, 90% near-equilibrium, 10% igniting
                 stoichiometric CH4-air at 1100-1300 K, deltaT 1e-3
  h2_adjoint     65536 cells of the h2_small state law, deltaT 1e-4
  syn200         200 species / 1200 reactions, 32768 cells, deltaT 1e-6
  syn100         100 species / 600 reactions (the workgroup kernel's
                 8 x 4 register block), 16384 cells, deltaT 1e-6
"""
from __future__ import annotations

import json
import os

import numpy as np

from .mechanism import MECH_DIR, Mechanism

P_ATM = 1e5
DTCHEM0 = 1e-7  # initialChemicalTimeStep of the reference tutorials

CONFIGS = {
    "h2_small": dict(mech="h2", cells=4096, solver="Rosenbrock34", relTol=1e-6,
                     absTol=1e-10, deltaT=1e-5, seed=1001),
    "gri_production": dict(mech="gri", cells=262144, solver="Rosenbrock34",
                           relTol=1e-6, absTol=1e-10, deltaT=1e-6, seed=1002),
    "gri_batching": dict(mech="gri", cells=1048576, solver="Rosenbrock34",
                         relTol=1e-6, absTol=1e-10, deltaT=1e-3, seed=1003),
    "h2_adjoint": dict(mech="h2", cells=65536, solver="Rosenbrock34", relTol=1e-6,
                       absTol=1e-10, deltaT=1e-4, seed=1004),
    "syn200": dict(mech="syn200", cells=32768, solver="Rosenbrock34", relTol=1e-6,
                   absTol=1e-10, deltaT=1e-6, seed=1005),
    "syn100": dict(mech="syn100", cells=16384, solver="Rosenbrock34", relTol=1e-6,
                   absTol=1e-10, deltaT=1e-6, seed=1006),
}


def molar_to_mass(X: np.ndarray, W: np.ndarray) -> np.ndarray:
    """Mole fractions [nS, n] -> mass fractions."""
    m = X * W[:, None]
    return m / m.sum(axis=0, keepdims=True)


def h2_states(mech: Mechanism, n: int, seed: int, p: float = P_ATM):
    """T~U(1000,1400) K; H2:O2:N2 = 2phi:1:3.76 (molar), phi~U(0.5,2);
    H, O, OH, HO2, H2O2 each log-uniform in [1e-8, 1e-4] mass fraction;
    renormalised."""
    rng = np.random.default_rng(seed)
    nS = mech.n_species
    W = mech.arrays["W"]
    phi = rng.uniform(0.5, 2.0, n)
    X = np.zeros((nS, n))
    X[mech.index("H2")] = 2 * phi
    X[mech.index("O2")] = 1.0
    X[mech.index("N2")] = 3.76
    Y = molar_to_mass(X, W)
    for s in ("H", "O", "OH", "HO2", "H2O2"):
        Y[mech.index(s)] = 10.0 ** rng.uniform(-8, -4, n)
    Y /= Y.sum(axis=0, keepdims=True)
    Y0 = np.zeros((nS + 2, n))
    Y0[:nS] = Y
    Y0[nS] = rng.uniform(1000.0, 1400.0, n)
    Y0[nS + 1] = p
    return Y0, np.full(n, DTCHEM0)


def ch4_air(mech: Mechanism, phi: float = 1.0) -> np.ndarray:
    """Mass fractions of CH4-air at equivalence ratio phi (CH4:O2:N2 =
    phi:2:7.52 molar)."""
    X = np.zeros((mech.n_species, 1))
    X[mech.index("CH4")] = phi
    X[mech.index("O2")] = 2.0
    X[mech.index("N2")] = 7.52
    return molar_to_mass(X, mech.arrays["W"])[:, 0]


def products_template(mech: Mechanism) -> np.ndarray:
    """Near-equilibrium CH4-air products (mass fractions), made once by
    integrating stoichiometric CH4-air from 1400 K at 1 atm for 0.1 s
    (tests/golden/make_products_template.py)."""
    with open(os.path.join(MECH_DIR, mech.name, "products_template.json")) as f:
        d = json.load(f)
    Y = np.zeros(mech.n_species)
    for k, v in d["Y"].items():
        Y[mech.index(k)] = v
    return Y


def _jitter(rng, Y: np.ndarray, frac: float = 0.05) -> np.ndarray:
    Y = Y * (1.0 + rng.uniform(-frac, frac, Y.shape))
    return Y / Y.sum(axis=0, keepdims=True)


def gri_production_states(mech: Mechanism, n: int, seed: int, p: float = P_ATM):
    """50% unburnt stoichiometric CH4-air, T~U(300,900) K; 50% products,
    T~U(1800,2400) K; mass fractions jittered by +-5% and renormalised;
    randomly interleaved."""
    rng = np.random.default_rng(seed)
    nS = mech.n_species
    unburnt = ch4_air(mech)
    prod = products_template(mech)
    burnt = rng.permutation(n) < n // 2
    Y = np.where(burnt[None, :], prod[:, None], unburnt[:, None])
    Y = _jitter(rng, Y)
    T = np.where(burnt, rng.uniform(1800, 2400, n), rng.uniform(300, 900, n))
    Y0 = np.zeros((nS + 2, n))
    Y0[:nS] = Y
    Y0[nS] = T
    Y0[nS + 1] = p
    return Y0, np.full(n, DTCHEM0)


def gri_batching_states(mech: Mechanism, n: int, seed: int, p: float = P_ATM,
                        igniting_fraction: float = 0.1):
    """90% near-equilibrium products (T~U(1800,2400) K) and 10% igniting
    stoichiometric CH4-air (T~U(1100,1300) K), randomly interleaved."""
    rng = np.random.default_rng(seed)
    nS = mech.n_species
    prod = products_template(mech)
    fresh = ch4_air(mech)
    ign = rng.permutation(n) < int(round(igniting_fraction * n))
    Y = np.where(ign[None, :], fresh[:, None], prod[:, None])
    Y = _jitter(rng, Y)
    T = np.where(ign, rng.uniform(1100, 1300, n), rng.uniform(1800, 2400, n))
    Y0 = np.zeros((nS + 2, n))
    Y0[:nS] = Y
    Y0[nS] = T
    Y0[nS + 1] = p
    return Y0, np.full(n, DTCHEM0), ign


def make_states(config: str, n: int | None = None, mech: Mechanism | None = None):
    """(mechanism, Y0, dtChem0, config dict) for a named configuration,
    optionally with fewer cells drawn from the same seeded law."""
    cfg = dict(CONFIGS[config])
    m = mech or Mechanism.load(cfg["mech"])
    nc = n or cfg["cells"]
    if cfg["mech"] == "h2":
        Y0, dt = h2_states(m, nc, cfg["seed"])
    elif config == "gri_production":
        Y0, dt = gri_production_states(m, nc, cfg["seed"])
    elif config == "gri_batching":
        Y0, dt, _ = gri_batching_states(m, nc, cfg["seed"])
    else:
        from .synthetic import syn_states
        Y0, dt = syn_states(m, nc, cfg["seed"])
    return m, Y0, dt, cfg
