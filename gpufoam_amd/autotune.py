"""Launch-shape autotuner of the batched solve.

The search space is the evaluator's batched shapes (gfc_evaluator_set_batched_shape):
one wavefront per cell with 1, 2, 4, ... resident wavefronts per compute
unit up to what fits, one cell per two wavefronts (32 < nSpecie+1 <= 64),
four cells per wavefront (16-lane teams, for mechanisms with
nSpecie+1 <= 16), and one lane per cell; for mechanisms that
run on the workgroup-per-cell kernel (64 < nSpecie+1 <= 208) also its LU
variant (gfc_evaluator_set_factorization: column by column, four-column
panels with vector-FMA or MFMA trailing updates).  Each candidate is timed
with HIP events on one CFD step of a sample of the caller's own cells (the
chemical time step reset before every run, so every candidate does the same
work), and the fastest is set on the evaluator.
"""
from __future__ import annotations

import numpy as np


def autotune(ev, deltaT: float, Y0, dtChem, sample: int = 4096, repeats: int = 2,
             seed: int = 0, shapes=None, factorizations=None) -> dict:
    """Times the batched solve for each launch shape on up to `sample`
    cells drawn from Y0 [nS+2, ld] / dtChem [n] (torch device tensors) and
    sets the fastest on `ev`.  Returns {(cellsPerWave, wavesPerCU): ms},
    the chosen shape under the key "best", and, where the workgroup kernel
    runs, {("lu", variant): ms} with the chosen variant under "best_lu"."""
    import torch
    n = dtChem.shape[0]
    if sample and sample < n:
        idx = torch.as_tensor(np.sort(np.random.default_rng(seed).choice(n, sample,
                                                                          replace=False)),
                              device=Y0.device)
        Ys = Y0[:, :n].index_select(1, idx).contiguous()
        dts = dtChem.index_select(0, idx).contiguous()
    else:
        Ys, dts = Y0[:, :n].contiguous(), dtChem.clone()
    Yn = torch.empty_like(Ys)
    if shapes is None:
        shapes = ([(1, w) for w in (1, 2, 4, 8, 16)] +
                  [(1, 0), (2, 0), (4, 0), (64, 0)])
    from ._capi import LU_COLUMN, LU_PANEL, LU_PANEL_MFMA, GfcError

    def timed():
        d = dts.clone()
        ev.solve(deltaT, Ys, d, Ynew=Yn, batched=True)  # warm-up (and graph-free)
        ms = []
        for _ in range(repeats):
            d.copy_(dts)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            ev.solve(deltaT, Ys, d, Ynew=Yn, batched=True)
            e1.record()
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        return min(ms)

    times = {}
    for cpw, wpc in shapes:
        try:
            ev.set_batched_shape(cpw, wpc)
        except GfcError:
            continue  # a shape this mechanism cannot run (4 cells need n <= 16,
            # two wavefronts per cell 32 < n <= 64)
        times[(cpw, wpc)] = timed()
    best = min(times, key=times.get)
    ev.set_batched_shape(*best)
    times["best"] = best
    n1 = ev.nS + 1
    if factorizations is None:
        factorizations = ([LU_COLUMN, LU_PANEL, LU_PANEL_MFMA]
                          if 64 < n1 <= 208 and best[0] == 1 else [])
    if factorizations:
        lu = {}
        for v in factorizations:
            ev.set_factorization(v)
            lu[v] = timed()
            times[("lu", v)] = lu[v]
        times["best_lu"] = min(lu, key=lu.get)
        ev.set_factorization(times["best_lu"])
    return times
