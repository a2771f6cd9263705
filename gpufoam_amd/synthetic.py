"""Seeded synthetic mechanism of 200 species / 1200 reactions and its cell
states (the scale workload).

Species: copies of the 53 GRI-Mech 3.0 species (copy k of species s is
named s for k = 0 and s_k otherwise; 200 = 3 full copies + 41 species of a
fourth), with the molecular weight and element content of the original and
every NASA-7 coefficient a1..a7 of both ranges multiplied by 1 + U(-0.01,
0.01).

Reactions: each takes the stoichiometry of a randomly drawn GRI reaction
(its reversibility too) and maps every species to a random copy of the same
base species, which keeps the mass balance exact (the compiler checks it,
as makeGpuReactions.H:48-54 does).  Rate types: 70 % Arrhenius, 20 % third
body with efficiencies ~ U(0.5, 5) for every species, 10 % Troe fall-off.
Rates: log10 A ~ U(8, 16), beta ~ U(-1, 2), Ea ~ U(0, 2.5e5) J/mol with
Ta = Ea / 8.314 K; Troe alpha ~ U(0.1, 0.9), T*** ~ U(50, 500) K,
T* ~ U(500, 5000) K, T** ~ U(2000, 9000) K (k0 and kInf drawn alike).

States: the h2_small state law with methane as the fuel: T ~ U(1000, 1400) K,
CH4:O2:N2 = phi/2 : 1 : 3.76 (molar), phi ~ U(0.5, 2), H, O, OH, HO2 and H2O2
log-uniform in [1e-8, 1e-4] mass fraction, renormalised; p = 1e5 Pa.
"""
from __future__ import annotations

import numpy as np

from .mechanism import Mechanism, _num, foam_thermo_text

N_SPECIES = 200
N_REACTIONS = 1200
SEED = 1005
R_EA = 8.314  # J/mol/K used to turn Ea into Ta


def _names(base_species, nS):
    nb = len(base_species)
    return [base_species[i % nb] if i < nb else f"{base_species[i % nb]}_{i // nb}"
            for i in range(nS)]


def syn_texts(seed: int = SEED, nS: int = N_SPECIES, nR: int = N_REACTIONS):
    """(thermo, reactions) OpenFOAM dictionary texts of the synthetic mechanism."""
    gri = Mechanism.load("gri")
    rng = np.random.default_rng(seed)
    nb = gri.n_species
    a = gri.arrays
    names = _names(gri.species, nS)
    base = np.arange(nS) % nb
    copies = [[i for i in range(nS) if base[i] == b] for b in range(nb)]
    high = gri.raw_high[base] * (1.0 + rng.uniform(-0.01, 0.01, (nS, 7)))
    low = gri.raw_low[base] * (1.0 + rng.uniform(-0.01, 0.01, (nS, 7)))
    thermo = foam_thermo_text(names, a["W"][base], a["sTlow"][base], a["sThigh"][base],
                              a["sTcommon"][base], high, low)

    def side(r, prefix):
        n = int(a[f"n_{prefix}"][r])
        terms = []
        for q in range(n):
            s = int(a[f"{prefix}_index"][r, q])
            st = float(a[f"{prefix}_stoich"][r, q])
            nm = names[rng.choice(copies[s])]
            terms.append(nm if st == 1.0 else f"{_num(st)}{nm}")
        return " + ".join(terms)

    def arrh(prefix):
        A = 10.0 ** rng.uniform(8.0, 16.0)
        beta = rng.uniform(-1.0, 2.0)
        Ta = rng.uniform(0.0, 2.5e5) / R_EA
        return [f"{prefix}A               {_num(A)};",
                f"{prefix}beta            {_num(beta)};",
                f"{prefix}Ta              {_num(Ta)};"]

    def effs(prefix):
        e = rng.uniform(0.5, 5.0, nS)
        return ([f"{prefix}coeffs", f"{nS}", "("] +
                [f"({s} {_num(x)})" for s, x in zip(names, e)] + [")", ";"])

    out = ["reactions", "{"]
    p = "        "
    for j in range(nR):
        t = int(rng.integers(gri.n_reactions))
        rev = not bool(a["irreversible"][t])
        u = rng.random()
        kind = "Arrhenius" if u < 0.7 else ("ThirdBodyArrhenius" if u < 0.9 else
                                             "ArrheniusTroeFallOff")
        eq = f"{side(t, 'lhs')} = {side(t, 'rhs')}"
        tn = ("reversible" if rev or kind != "Arrhenius" else "irreversible") + kind
        out += [f"    syn-{j}", "    {", f"{p}type            {tn};",
                f"{p}reaction        \"{eq}\";"]
        if kind == "Arrhenius":
            out += arrh(p)
        elif kind == "ThirdBodyArrhenius":
            out += arrh(p) + effs(p)
        else:
            out += [p + "k0", p + "{"] + arrh(p + "    ") + [p + "}"]
            out += [p + "kInf", p + "{"] + arrh(p + "    ") + [p + "}"]
            out += [p + "F", p + "{",
                    f"{p}    alpha           {_num(rng.uniform(0.1, 0.9))};",
                    f"{p}    Tsss            {_num(rng.uniform(50.0, 500.0))};",
                    f"{p}    Ts              {_num(rng.uniform(500.0, 5000.0))};",
                    f"{p}    Tss             {_num(rng.uniform(2000.0, 9000.0))};",
                    p + "}", p + "thirdBodyEfficiencies", p + "{"] + effs(p + "    ") + [p + "}"]
        out += ["    }"]
    out += ["}", f"Tlow            {_num(gri.Tlow)};", f"Thigh           {_num(gri.Thigh)};", ""]
    return thermo, "\n".join(out)


def syn_mechanism(seed: int = SEED) -> Mechanism:
    t, r = syn_texts(seed)
    return Mechanism(t, r, "syn200")


def syn_states(mech: Mechanism, n: int, seed: int, p: float = 1e5):
    """Cell states of the synthetic mechanism (see the module docstring)."""
    from .workloads import DTCHEM0, molar_to_mass
    rng = np.random.default_rng(seed)
    nS = mech.n_species
    W = mech.arrays["W"]
    phi = rng.uniform(0.5, 2.0, n)
    X = np.zeros((nS, n))
    X[mech.index("CH4")] = phi / 2.0
    X[mech.index("O2")] = 1.0
    X[mech.index("N2")] = 3.76
    Y = molar_to_mass(X, W)
    for s in ("H", "O", "OH", "HO2", "H2O2"):
        Y[mech.index(s)] = 10.0 ** rng.uniform(-8.0, -4.0, n)
    Y /= Y.sum(axis=0, keepdims=True)
    T = rng.uniform(1000.0, 1400.0, n)
    Y0 = np.concatenate([Y, T[None, :], np.full((1, n), p)], axis=0)
    return Y0, np.full(n, DTCHEM0)
