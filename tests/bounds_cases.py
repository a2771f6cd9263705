"""Runs one small case of every kernel family (drop-in, persistent
wavefront, two-wavefront, one-lane, four-cells-per-wavefront, workgroup
kernels; the wavefront and four-cell adjoints) through the library that
gpufoam_amd._capi loads, and writes the results to an .npz file.

``python tests/bounds_cases.py OUT.npz`` with GFC_LIBRARY naming the
bounds-check build runs the same cases under its checks; the GPU test
compares both files bit for bit.  Each case prints its name before it
runs."""
from __future__ import annotations

import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402
import torch  # noqa: E402

from gpufoam_amd import _capi, ode_inputs  # noqa: E402
from gpufoam_amd.evaluator import Evaluator  # noqa: E402
from gpufoam_amd.mechanism import Mechanism  # noqa: E402
from gpufoam_amd.synthetic import syn_states, syn_texts  # noqa: E402
from gpufoam_amd.workloads import make_states  # noqa: E402

DEV = "cuda:0"


def _solve(m, Y0, dt0, c, shape, batched, cells):
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    ev = Evaluator(m, ode, 0.0, max_cells=cells)
    if shape is not None:
        ev.set_batched_shape(shape)
    dY0 = torch.tensor(Y0, device=DEV)
    dYn = torch.empty_like(dY0)
    ddt = torch.tensor(dt0, device=DEV)
    st = torch.zeros((3, Y0.shape[1]), dtype=torch.int32, device=DEV)
    ev.solve(c["deltaT"], dY0, ddt, Ynew=dYn, stats=st, batched=batched)
    torch.cuda.synchronize()
    return {"Y": dYn.cpu().numpy(), "dt": ddt.cpu().numpy(), "stats": st.cpu().numpy()}


def _adjoint(m, Y0, dt0, c, shape, cells):
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    ev = Evaluator(m, ode, 0.0, max_cells=cells)
    if shape is not None:
        ev.set_batched_shape(shape)
    dY0 = torch.tensor(Y0, device=DEV)
    dYn = torch.empty_like(dY0)
    ddt = torch.tensor(dt0, device=DEV)
    ev.adjoint_forward(c["deltaT"], dY0, ddt, Ynew=dYn)
    n = m.n_species + 1
    W = torch.zeros((n, cells), dtype=torch.float64, device=DEV)
    W[m.n_species] = 1.0
    g = torch.empty_like(W)
    pb = torch.empty((6 * m.n_reactions, cells), dtype=torch.float64, device=DEV)
    ev.adjoint_backward(W, g, pb)
    torch.cuda.synchronize()
    return {"Y": dYn.cpu().numpy(), "g": g.cpu().numpy(), "pb": pb.cpu().numpy()}


def run_cases() -> dict:
    out = {"checked": np.array([_capi.lib().gfc_bounds_checked()])}

    def case(name, fn):
        print(f"[bounds case] {name}", flush=True)
        for k, v in fn().items():
            out[f"{name}/{k}"] = v

    m, Y0, dt0, c = make_states("gri_production", 128)
    case("gri drop-in", lambda: _solve(m, Y0, dt0, c, None, False, 128))
    case("gri wavefront", lambda: _solve(m, Y0, dt0, c, _capi.SHAPE_WAVE, True, 128))
    case("gri two-wavefront", lambda: _solve(m, Y0, dt0, c, _capi.SHAPE_PAIR, True, 128))
    case("gri lane", lambda: _solve(m, Y0[:, :64], dt0[:64], c, 64, True, 64))
    case("gri adjoint", lambda: _adjoint(m, Y0[:, :32], dt0[:32], c, None, 32))
    mh, Yh, dth, ch = make_states("h2_small", 128)
    case("h2 four-cell", lambda: _solve(mh, Yh, dth, ch, _capi.SHAPE_TEAM16, True, 128))
    case("h2 four-cell adjoint",
         lambda: _adjoint(mh, Yh[:, :64], dth[:64], ch, _capi.SHAPE_TEAM16, 64))
    t, r = syn_texts(seed=7, nS=100, nR=600)
    ms = Mechanism(t, r, "syn100")
    Ys, dts = syn_states(ms, 8, seed=11)
    cs = dict(solver="Rosenbrock34", relTol=1e-6, absTol=1e-10, deltaT=1e-6)
    case("syn100 workgroup", lambda: _solve(ms, Ys, dts, cs, None, True, 8))
    return out


if __name__ == "__main__":
    res = run_cases()
    np.savez(sys.argv[1], **res)
    print("[bounds case] done", flush=True)
