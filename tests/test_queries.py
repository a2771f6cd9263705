"""The chemistry queries at the current state -- calculate, reactionRR,
specieReactionRR, tc and Qdot (gpuChemistryModel.C:146-466) -- restated in
the oracle, pinned on the CPU:

  * reactionRR against gpuReaction::omega of the reference's own headers
    built for the host (bit for bit), which the reference's tests hold equal
    to OpenFOAM's Reaction::omega (Test-gpuChemistry.C:384-388);
  * the forward/reverse split of omega against kf, kr and the
    concentration products formed independently here;
  * specieReactionRR against the stoichiometry (W_i nu_ir omega_r), and
    calculate against the sum of specieReactionRR over reactions;
  * tc against its closed form on a one-reaction mechanism
    (first-order A -> A': tc = sum(c) / (k c_A)) and its zero-rate case;
  * Qdot against the formation enthalpies of the reference build, with
    hf(H2O) checked against the tabulated -241.8 kJ/mol.
"""
import numpy as np
import pytest

from gpufoam_amd import Mechanism
from gpufoam_amd.workloads import make_states
from oracle_lib import Oracle, Reference, ref_available

import test_mechanism as TM


def _query_states(mech_cfg, n, seed=7):
    """Benchmark states with a few mass fractions pushed slightly negative
    (the queries do not clip c) and a density field."""
    m, Y0, _, _ = make_states(mech_cfg, n)
    rng = np.random.default_rng(seed)
    nS = m.n_species
    Y = Y0.copy()
    neg = rng.random((nS, n)) < 0.02
    Y[:nS][neg] = -1e-12 * rng.random(neg.sum())
    rho = rng.uniform(0.2, 1.5, n)
    return m, Y, rho


def _conc(m, Y, rho, cell):
    nS = m.n_species
    return rho[cell] * Y[:nS, cell] / m.arrays["W"]


def _net_nu(m, r):
    a = m.arrays
    nu = np.zeros(m.n_species)
    for q in range(int(a["n_lhs"][r])):
        nu[a["lhs_index"][r, q]] -= a["lhs_stoich"][r, q]
    for q in range(int(a["n_rhs"][r])):
        nu[a["rhs_index"][r, q]] += a["rhs_stoich"][r, q]
    return nu


@pytest.mark.skipif(not ref_available(), reason="reference build absent")
@pytest.mark.parametrize("cfg", ["h2_small", "gri_production"])
def test_reaction_rr_is_the_reference_omega(cfg):
    m, Y, rho = _query_states(cfg, 12)
    o, ref = Oracle(m), Reference(m)
    nS = m.n_species
    for r in range(m.n_reactions):
        got = o.reaction_rr(r, rho, Y)
        for cell in range(Y.shape[1]):
            want = ref.rates(r, Y[nS + 1, cell], Y[nS, cell], _conc(m, Y, rho, cell))[2]
            assert got[cell] == want or (np.isnan(got[cell]) and np.isnan(want)), (r, cell)


@pytest.mark.parametrize("cfg", ["h2_small", "gri_production"])
def test_omega_split_into_forward_and_reverse(cfg):
    m, Y, rho = _query_states(cfg, 6)
    o = Oracle(m)
    a = m.arrays
    nS = m.n_species
    for cell in range(Y.shape[1]):
        c = _conc(m, Y, rho, cell)
        p, T = Y[nS + 1, cell], Y[nS, cell]
        for r in range(m.n_reactions):
            kf, kr, w = o.rates(r, p, T, c)
            wf, wr = o.omega_fr(r, p, T, c)
            assert wf - wr == w
            Cf = np.prod([c[a["lhs_index"][r, q]] ** a["lhs_exp"][r, q]
                          for q in range(int(a["n_lhs"][r]))])
            Cr = np.prod([c[a["rhs_index"][r, q]] ** a["rhs_exp"][r, q]
                          for q in range(int(a["n_rhs"][r]))])
            assert np.isclose(wf, kf * Cf, rtol=1e-13, atol=0)
            assert np.isclose(wr, kr * Cr, rtol=1e-13, atol=0)
            if a["irreversible"][r]:
                assert wr == 0.0


@pytest.mark.parametrize("cfg", ["h2_small", "gri_production"])
def test_specie_rates_follow_the_stoichiometry(cfg):
    m, Y, rho = _query_states(cfg, 24)
    o = Oracle(m)
    W = m.arrays["W"]
    total = np.zeros((m.n_species, Y.shape[1]))
    scale = np.zeros_like(total)
    for r in range(m.n_reactions):
        w = o.reaction_rr(r, rho, Y)
        srr = o.specie_reaction_rr(r, rho, Y)
        want = (W * _net_nu(m, r))[:, None] * w[None, :]
        assert np.allclose(srr, want, rtol=1e-15, atol=0)
        total += srr
        scale += np.abs(srr)
    RR = o.calculate(rho, Y)
    assert np.all(np.abs(RR - total) <= 1e-13 * scale)
    # mass balance: sum_i RR_i = sum_r omega_r sum_i nu_ir W_i
    imb = sum(o.reaction_rr(r, rho, Y) * float(np.dot(_net_nu(m, r), W))
              for r in range(m.n_reactions))
    assert np.allclose(RR.sum(axis=0), imb, rtol=1e-9, atol=1e-12 * scale.sum(axis=0).max())


def test_tc_closed_form_first_order():
    """One irreversible first-order reaction H2O -> H2O with rate constant
    k: wf = k c_H2O counted once on the right side, wr = 0, so
    tc = sum(c) / (k c_H2O); with no H2O every rate is zero and tc is
    vGreat = DBL_MAX/10."""
    k = 37.5
    rx = TM.reactions(
        f'r0 {{ type irreversibleArrhenius; reaction "H2O = H2O"; A {k}; beta 0; Ta 0; }}')
    m = Mechanism(TM.THERMO, rx)
    o = Oracle(m)
    n = 5
    Y = np.zeros((5, n))
    Y[0] = [0.1, 0.2, 0.0, 0.5, 0.0]
    Y[1] = [0.3, 0.1, 0.5, 0.0, 0.0]
    Y[2] = 1.0 - Y[0] - Y[1]
    Y[2, 4] = 0.0
    Y[3] = 1500.0
    Y[4] = 1e5
    rho = np.array([0.3, 0.5, 0.8, 1.1, 0.9])
    tc = o.tc(rho, Y)
    W = m.arrays["W"]
    c = rho[None, :] * Y[:3] / W[:, None]
    with np.errstate(divide="ignore", invalid="ignore"):
        want = c.sum(axis=0) / (k * c[2])
    assert np.allclose(tc[:4], want[:4], rtol=4e-15, atol=0)
    assert tc[4] == np.finfo(np.float64).max / 10
    # the net rate of H2O -> H2O cancels exactly
    assert not o.calculate(rho, Y).any()


@pytest.mark.skipif(not ref_available(), reason="reference build absent")
@pytest.mark.parametrize("cfg", ["h2_small", "gri_production"])
def test_qdot_from_formation_enthalpies(cfg):
    m, Y, rho = _query_states(cfg, 16)
    o, ref = Oracle(m), Reference(m)
    RR = o.calculate(rho, Y)
    hf = np.array([ref.species_thermo(i, 1e5, 1000.0)[2] for i in range(m.n_species)])
    q = o.qdot(RR)
    want = -(hf[:, None] * RR).sum(axis=0)
    assert np.allclose(q, want, rtol=1e-12, atol=1e-12 * np.abs(hf[:, None] * RR).sum(axis=0).max())
    # hf is the enthalpy of formation per unit mass: H2O at 298.15 K is
    # -241.826 kJ/mol (gas), H2 and O2 are reference elements (0)
    iH2O = m.index("H2O")
    W = m.arrays["W"]
    assert abs(hf[iH2O] * W[iH2O] / 1e6 - (-241.826)) < 0.1  # kJ/mol
    for s in ("H2", "O2"):
        assert abs(hf[m.index(s)] * W[m.index(s)] / 1e6) < 0.01


def test_tc_weights_forward_by_products_and_reverse_by_reactants():
    """One reversible reaction 2H2 + O2 = 2H2O: the forward rate is counted
    with the right side's coefficients (2), the reverse with the left
    side's (3): tc = sum(c) (2 wf + 3 wr) / ((2 wf)^2 + (3 wr)^2)."""
    rx = TM.reactions(
        'r0 { type reversibleArrhenius; reaction "2H2 + O2 = 2H2O"; A 1e10; beta 0.5; '
        'Ta 1000; }')
    m = Mechanism(TM.THERMO, rx)
    o = Oracle(m)
    rng = np.random.default_rng(3)
    n = 8
    Y = np.zeros((5, n))
    Y[:3] = rng.dirichlet([1, 1, 1], n).T
    Y[3] = rng.uniform(1500, 3500, n)
    Y[4] = 1e5
    rho = rng.uniform(0.1, 1.0, n)
    tc = o.tc(rho, Y)
    W = m.arrays["W"]
    for cell in range(n):
        c = rho[cell] * Y[:3, cell] / W
        wf, wr = o.omega_fr(0, 1e5, Y[3, cell], c)
        assert wf > 0 and wr > 0
        want = c.sum() * (2 * wf + 3 * wr) / ((2 * wf) ** 2 + (3 * wr) ** 2)
        assert np.isclose(tc[cell], want, rtol=1e-14, atol=0)
