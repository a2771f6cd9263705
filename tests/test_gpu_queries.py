"""The chemistry queries at the current state on the device (gfc_calculate,
gfc_reaction_rr, gfc_specie_reaction_rr, gfc_tc, gfc_qdot) against the
oracle (gpuChemistryModel.C:146-466), on benchmark states with a few mass
fractions pushed slightly negative (the queries do not clip c).

Bounds (the device evaluates exp/log/pow with its own last-ulp rounding and
sums species and reactions in another order; a net rate is a difference of
its forward and reverse parts, so its scale is s_r = |omegaf_r| + |omegar_r|):
  rates           |d| <= 1e-12 s_r
  specie rates    |d| <= 1e-12 W_i sum_r |nu_ir| s_r
  tc              |d| <= 1e-11 tc
  Qdot            |d| <= 1e-13 sum_i |hf_i RR_i|   (same RR input)
The host-pointer path returns the device-pointer path's bits."""
import numpy as np
import pytest

from gpufoam_amd import ode_inputs
from gpufoam_amd.workloads import make_states

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

DEV = "cuda"


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gpufoam_amd import _capi
    _capi.lib()


def _states(cfg, n, seed=11):
    m, Y0, _, c = make_states(cfg, n)
    rng = np.random.default_rng(seed)
    nS = m.n_species
    Y = Y0.copy()
    neg = rng.random((nS, n)) < 0.02
    Y[:nS][neg] = -1e-12 * rng.random(neg.sum())
    rho = rng.uniform(0.2, 1.5, n)
    return m, Y, rho, c


@pytest.mark.parametrize("cfg,n", [("h2_small", 256), ("gri_production", 512), ("syn200", 64)])
def test_queries_match_oracle(cfg, n):
    from gpufoam_amd.evaluator import Evaluator
    from oracle_lib import Oracle
    print(f"case {cfg} n={n} queries", flush=True)
    m, Y, rho, c = _states(cfg, n)
    o = Oracle(m)
    nS, nR = m.n_species, m.n_reactions
    ev = Evaluator(m, ode_inputs(c["solver"]), 0.0, max_cells=n)
    dY = torch.tensor(Y, device=DEV)
    drho = torch.tensor(rho, device=DEV)

    # the cancellation scale s_r of every reaction and cell, and of every
    # specie rate
    W = m.arrays["W"]
    S = np.zeros((nR, n))
    for cell in range(n):
        cc = rho[cell] * Y[:nS, cell] / W
        for r in range(nR):
            wf, wr = o.omega_fr(r, Y[nS + 1, cell], Y[nS, cell], cc)
            S[r, cell] = abs(wf) + abs(wr)
    a = m.arrays
    absnu = np.zeros((nR, nS))
    for r in range(nR):
        for side in ("lhs", "rhs"):
            for q in range(int(a[f"n_{side}"][r])):
                absnu[r, a[f"{side}_index"][r, q]] += a[f"{side}_stoich"][r, q]
    scale = W[:, None] * (absnu.T @ S)

    # calculate: specie rates of the whole mechanism
    RR = torch.empty((nS, n), dtype=torch.float64, device=DEV)
    ev.calculate(drho, dY, RR)
    torch.cuda.synchronize()
    RRg = RR.cpu().numpy()
    RRo = o.calculate(rho, Y)
    reactions = range(nR) if nR <= 400 else range(0, nR, 7)
    err = np.abs(RRg - RRo) / (scale + 1e-300)
    print(f"calculate: max |d| / scale {err.max():.3g}")
    assert np.all(np.abs(RRg - RRo) <= 1e-12 * scale)

    # reactionRR and specieReactionRR of (a sample of) the reactions
    out = torch.empty(n, dtype=torch.float64, device=DEV)
    SR = torch.empty((nS, n), dtype=torch.float64, device=DEV)
    worst = 0.0
    for r in reactions:
        ev.reactionRR(r, drho, dY, out)
        ev.specieReactionRR(r, drho, dY, SR)
        torch.cuda.synchronize()
        w, wo = out.cpu().numpy(), o.reaction_rr(r, rho, Y)
        assert np.all(np.abs(w - wo) <= 1e-12 * S[r]), r
        s, so = SR.cpu().numpy(), o.specie_reaction_rr(r, rho, Y)
        assert np.all(np.abs(s - so) <= 1e-12 * (W * absnu[r])[:, None] * S[r][None, :]), r
        worst = max(worst, float((np.abs(w - wo) / (S[r] + 1e-300)).max()))
    print(f"reactionRR: max |d| / s_r {worst:.3g}")

    # tc
    tc = torch.empty(n, dtype=torch.float64, device=DEV)
    ev.tc(drho, dY, tc)
    torch.cuda.synchronize()
    tcg, tco = tc.cpu().numpy(), o.tc(rho, Y)
    print(f"tc: max rel {(np.abs(tcg - tco) / np.abs(tco)).max():.3g}")
    assert np.all(np.abs(tcg - tco) <= 1e-11 * np.abs(tco))

    # Qdot of the oracle's RR
    q = torch.empty(n, dtype=torch.float64, device=DEV)
    ev.Qdot(torch.tensor(RRo, device=DEV), q)
    torch.cuda.synchronize()
    qo = o.qdot(RRo)
    hf = np.array([o.species_thermo(i, 1e5, 1000.0)[2] for i in range(nS)])
    qs = np.abs(hf[:, None] * RRo).sum(axis=0)
    assert np.all(np.abs(q.cpu().numpy() - qo) <= 1e-13 * qs + 1e-300)


def test_host_pointer_queries_equal_device_pointer_queries():
    from gpufoam_amd import _capi
    from gpufoam_amd.evaluator import Evaluator
    m, Y, rho, c = _states("gri_production", 96)
    nS, n = m.n_species, 96
    ev = Evaluator(m, ode_inputs(c["solver"]), 0.0, max_cells=128)
    L = _capi.lib()
    dY, drho = torch.tensor(Y, device=DEV), torch.tensor(rho, device=DEV)
    RRd = torch.empty((nS, n), dtype=torch.float64, device=DEV)
    tcd = torch.empty(n, dtype=torch.float64, device=DEV)
    ev.calculate(drho, dY, RRd)
    ev.tc(drho, dY, tcd)
    torch.cuda.synchronize()
    RRh, tch, qh = np.zeros((nS, n)), np.zeros(n), np.zeros(n)
    Yc = np.ascontiguousarray(Y)
    H = _capi.HOST_POINTERS
    _capi.check(L.gfc_calculate(ev.handle, n, n, rho.ctypes.data, Yc.ctypes.data,
                                RRh.ctypes.data, None, H))
    _capi.check(L.gfc_tc(ev.handle, n, n, rho.ctypes.data, Yc.ctypes.data,
                         tch.ctypes.data, None, H))
    _capi.check(L.gfc_qdot(ev.handle, n, n, RRh.ctypes.data, qh.ctypes.data, None, H))
    assert np.array_equal(RRh, RRd.cpu().numpy())
    assert np.array_equal(tch, tcd.cpu().numpy())
    qd = torch.empty(n, dtype=torch.float64, device=DEV)
    ev.Qdot(RRd, qd)
    torch.cuda.synchronize()
    assert np.array_equal(qh, qd.cpu().numpy())
    # argument errors come back as status codes
    from gpufoam_amd import GfcError
    with pytest.raises(GfcError) as e:
        _capi.check(L.gfc_reaction_rr(ev.handle, m.n_reactions, n, n, rho.ctypes.data,
                                      Yc.ctypes.data, tch.ctypes.data, None, H))
    assert e.value.status == 1
    with pytest.raises(GfcError) as e:
        _capi.check(L.gfc_tc(ev.handle, 200, 200, rho.ctypes.data, Yc.ctypes.data,
                             tch.ctypes.data, None, H))
    assert e.value.status == 9
