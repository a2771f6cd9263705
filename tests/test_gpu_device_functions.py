"""Device-level tests of the batched kernels' own physics and LU, function by
function, as the reference's [GPU] ladder tests its device code
(unittest/tests/Test-gpuChemistry.C:521-691: ludecompose, and
gpuODESystem::derivatives / ::jacobian at 1e-9).

Each kernel family (one wavefront per cell, 16-lane teams, one workgroup
per cell) evaluates the derivative and the Jacobian with the same device
functions its solve calls (gfc_eval_derivatives / gfc_eval_jacobian), and
factorises and solves with its own LU (gfc_eval_lu_solve); the results are
held to the oracle at the same states:

* derivative and Jacobian entries: |gpu - oracle| <= 1e-9 |oracle| +
  1e-13 max|row| (element-wise relative 1e-9, the reference's bound, plus an
  allowance of 1e-13 of the row's largest entry for entries that are sums
  with cancellation: the kernels sum reaction contributions in another
  order than the reference's sequential loop);
* LU solves of the Rosenbrock matrices I/(gamma h) - J of those states and
  of random matrices: |x - x_oracle| <= max(1e-9, n eps cond_inf(M))
  max|x_oracle|, with the oracle's LUDecompose / LUBacksubstitute
  (ludecompose.H:9-140) as the reference (two backward-stable solves of a
  matrix of condition number cond differ by up to about n eps cond; the
  Rosenbrock matrix of the cell above the temperature range reaches 1e14),
  and the scaled residual |M x - b| / (max|M| max|x|) <= 1e-12.

Mechanisms: H2 (n = 11: the wavefront and 16-lane kernels), GRI-Mech 3.0
(n = 54: wavefront, and two wavefronts per cell), a 60-species synthetic
mechanism (n = 61: the 64-row wavefront kernel), a 60-species one with 45
reactions (the padded 64-row two-wavefront kernel), and the 100- and 200-species synthetic ones (n = 101,
201: the two workgroup register blocks, every factorisation variant).  No
bundled mechanism has 16 < n <= 32, the 32-row wavefront instantiation's
range (the synthetic generator needs every GRI species).
"""
import numpy as np
import pytest

from gpufoam_amd import _capi
from gpufoam_amd.mechanism import Mechanism
from gpufoam_amd.synthetic import syn_states, syn_texts
from gpufoam_amd.workloads import make_states

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

REL, ROWABS = 1e-9, 1e-13
EPS = np.finfo(np.float64).eps


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _capi.lib()


def _syn(nS, nR, seed):
    t, r = syn_texts(seed=seed, nS=nS, nR=nR)
    return Mechanism(t, r, f"syn{nS}")


def _case(name):
    if name == "h2":
        m, Y0, _, _ = make_states("h2_small", 48)
    elif name == "gri":
        m, Y0, _, _ = make_states("gri_production", 32)
    elif name == "syn60":
        m = _syn(60, 300, 23)
        Y0, _ = syn_states(m, 24, seed=24)
    elif name == "syn60s":
        # 45 reactions, 12 efficiency sets: the padded two-wavefront kernel
        t, r = syn_texts(seed=13, nS=60, nR=45)
        m = Mechanism(t, r, "syn60s")
        Y0, _ = syn_states(m, 16, seed=14)
    elif name == "syn100":
        m = _syn(100, 600, 7)
        Y0, _ = syn_states(m, 16, seed=11)
    else:
        m, Y0, _, _ = make_states("syn200", 12)
    Y0 = np.ascontiguousarray(Y0)
    # one cell above the reaction temperature range: the derivative
    # evaluates the rates at the clipped temperature, the Jacobian does not
    Y0[m.n_species, 1] = m.Thigh + 150.0
    return m, Y0


SHAPES = [("h2", _capi.SHAPE_WAVE), ("h2", _capi.SHAPE_TEAM16), ("gri", _capi.SHAPE_WAVE),
          ("gri", _capi.SHAPE_PAIR), ("syn60", _capi.SHAPE_WAVE), ("syn60s", _capi.SHAPE_PAIR),
          ("syn100", _capi.SHAPE_WORKGROUP), ("syn200", _capi.SHAPE_WORKGROUP)]


def _evaluator(m, n):
    from gpufoam_amd import ode_inputs
    from gpufoam_amd.evaluator import Evaluator
    return Evaluator(m, ode_inputs("Rosenbrock34", relTol=1e-6, absTol=1e-10), 0.0,
                     max_cells=n)


def _close(got, want, what):
    """Element-wise relative bound with a row-scaled floor (module
    docstring); rows are the last axis."""
    scale = np.abs(want).max(axis=-1, keepdims=True)
    err = np.abs(got - want)
    bound = REL * np.abs(want) + ROWABS * scale
    worst = (err / np.where(bound > 0, bound, 1.0)).max()
    rel = (err / np.where(np.abs(want) > 0, np.abs(want), 1.0)).max()
    print(f"  {what}: max err/bound {worst:.3g}, max element-wise rel {rel:.3g}")
    assert np.all(err <= bound), what
    assert np.array_equal(np.isfinite(got), np.isfinite(want)), what


@pytest.mark.parametrize("name,shape", SHAPES)
def test_derivatives_and_jacobian_match_oracle(name, shape):
    from oracle_lib import Oracle
    m, Y0 = _case(name)
    nS, n = m.n_species, m.n_species + 1
    ncell = Y0.shape[1]
    o = Oracle(m)
    ev = _evaluator(m, ncell)
    print(f"case {name} shape {shape}: {ncell} cells, n = {n}", flush=True)
    dY = torch.tensor(Y0, device="cuda")
    d = ev.eval_derivatives(dY, shape).cpu().numpy()
    dj, J = ev.eval_jacobian(dY, shape)
    torch.cuda.synchronize()
    dj, J = dj.cpu().numpy(), J.cpu().numpy()
    dref = np.stack([o.derivatives(Y0[:, c])[:n] for c in range(ncell)], axis=1)
    jref = [o.jacobian(Y0[:, c]) for c in range(ncell)]
    djref = np.stack([r[0][:n] for r in jref], axis=1)
    Jref = np.stack([r[1][:n, :n] for r in jref])
    # species part and temperature part of the derivative separately
    _close(d[:nS].T, dref[:nS].T, "derivative (species)")
    _close(d[nS:].T, dref[nS:].T, "derivative (T)")
    _close(dj[:nS].T, djref[:nS].T, "Jacobian pass derivative (species)")
    _close(dj[nS:].T, djref[nS:].T, "Jacobian pass derivative (T)")
    _close(J, Jref, "Jacobian")


def _matrices(m, Y0, ncell, seed):
    """Rosenbrock matrices I/(gamma h) - J of the states (h log-uniform in
    [1e-9, 1e-5]) for the first half of the cells, random matrices with
    entries in [-1, 1) plus n on the diagonal for the rest."""
    from oracle_lib import Oracle
    o = Oracle(m)
    n = m.n_species + 1
    rng = np.random.default_rng(seed)
    M = np.empty((ncell, n, n))
    for c in range(ncell):
        if c < ncell // 2:
            h = 10.0 ** rng.uniform(-9, -5)
            M[c] = np.eye(n) / (0.5 * h) - o.jacobian(Y0[:, c])[1][:n, :n]
        else:
            M[c] = rng.uniform(-1, 1, (n, n)) + n * np.eye(n)
    b = rng.uniform(-1, 1, (n, ncell))
    return M, b


@pytest.mark.parametrize("name,shape", SHAPES)
def test_lu_solve_matches_oracle(name, shape):
    from oracle_lib import lu_decompose, lu_solve
    m, Y0 = _case(name)
    n = m.n_species + 1
    ncell = Y0.shape[1]
    M, b = _matrices(m, Y0, ncell, seed=5)
    xref = np.stack([lu_solve(*lu_decompose(M[c]), b[:, c]) for c in range(ncell)], axis=1)
    cond = np.array([np.linalg.cond(M[c], np.inf) for c in range(ncell)])
    variants = ([_capi.LU_COLUMN, _capi.LU_PANEL, _capi.LU_PANEL_MFMA]
                if shape == _capi.SHAPE_WORKGROUP else [None])
    ev = _evaluator(m, ncell)
    dM = torch.tensor(M, device="cuda")
    db = torch.tensor(b, device="cuda")
    xs = []
    for v in variants:
        if v is not None:
            ev.set_factorization(v)
        print(f"case {name} shape {shape} LU {v}: {ncell} matrices, n = {n}", flush=True)
        x = ev.eval_lu_solve(dM, db, shape).cpu().numpy()
        err = np.abs(x - xref).max(axis=0) / np.abs(xref).max(axis=0)
        res = max(np.abs(M[c] @ x[:, c] - b[:, c]).max() /
                  (np.abs(M[c]).max() * np.abs(x[:, c]).max()) for c in range(ncell))
        bound = np.maximum(1e-9, n * EPS * cond)
        print(f"  max |x - x_oracle| / max|x_oracle| {err.max():.3g} (cell {err.argmax()}, "
              f"cond {cond[err.argmax()]:.3g}); max of error / bound {(err / bound).max():.3g}; "
              f"scaled residual {res:.3g}")
        assert np.all(err <= bound)
        assert res <= 1e-12
        xs.append(x)
    # the factorisation variants agree bit for bit
    for x in xs[1:]:
        assert np.array_equal(x, xs[0])


def test_shapes_the_mechanism_cannot_take_are_refused():
    from gpufoam_amd import GfcError
    m, Y0 = _case("h2")
    ev = _evaluator(m, Y0.shape[1])
    dY = torch.tensor(Y0, device="cuda")
    with pytest.raises(GfcError) as e:
        ev.eval_derivatives(dY, _capi.SHAPE_WORKGROUP)
    assert e.value.status == 3  # GFC_ERR_UNSUPPORTED
    with pytest.raises(GfcError):
        ev.eval_derivatives(dY, 7)
