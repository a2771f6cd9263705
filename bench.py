"""Benchmark: chemistry cells per second for one CFD time step on one MI355X.

Workload (default): the 53-specie / 325-reaction GRI methane-air mechanism,
262144 cells in FP64, Rosenbrock34 with relTol 1e-6 and absTol 1e-10, one
CFD step deltaT = 1e-6 s at p = 1e5 Pa; half unburnt CH4-air at
300-900 K, half near-equilibrium products at 1800-2400 K, mass fractions
jittered by +-5% (synthetic, seeded).  A timed step is one full chemistry
solve of all cells (the batched persistent kernel replayed as a hipGraph),
with the chemical time step reset to its initial value before each step so
every step does the same work.  Inputs are resident in HBM.

Printed (one JSON line): the headline metric and, beside it,
  * the drop-in launch (the same per-cell kernel over one grid of cells,
    no work stealing) and the one-lane-per-cell kernel (the reference's
    mapping) on the same cells;
  * the batching workload (GRI, 1048576 cells, 10% igniting, deltaT 1e-3):
    persistent + hipGraph against the drop-in launch, with its attempt
    histogram;
  * the FP64 roofline fraction of the batched kernel;
  * the CPU baseline (gpuFoam's own per-cell code built for the host, or
    the oracle restatement where that build is absent) on the host cores
    over a bounded sample of the same cells, and the error of the GPU
    result against that CPU result on the sample;
  * the adjoint/forward cost on the differentiation workload (10-species
    H2, 65536 cells, deltaT 1e-4) and on 16384 cells of the 53-species
    workload.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--config gri_production]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)


def progress(msg: str) -> None:
    """Stage line on stderr (stdout carries only the JSON result)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(m, Y0, dt0, c, config: str, sample: int, threads: int,
                 gpu_Ynew: np.ndarray) -> dict:
    """gpuFoam's per-cell solver built for the host (oracle/_ref, shipped
    with the tree) or, where that build is absent, the oracle, timed on the
    first `sample` cells of the benchmarked workload.  The CPU result on
    the sample is also the yardstick of the GPU result's error."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    from gpufoam_amd import _capi, ode_inputs
    import oracle_lib
    Ys = np.ascontiguousarray(Y0[:, :sample])
    dts = np.ascontiguousarray(dt0[:sample])
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    if oracle_lib.ref_available():
        runner, kind = oracle_lib.Reference(m), "reference"
    else:
        runner, kind = oracle_lib.Oracle(m), "port"
        progress("cpu baseline: oracle/_ref/libgfref.so absent, timing the oracle (kind port)")
    runner.solve_cells(ode, Ys[:, :64], c["deltaT"], dts[:64], threads=threads)  # warm
    t = time.perf_counter()
    ref = runner.solve_cells(ode, Ys, c["deltaT"], dts, threads=threads)
    el = time.perf_counter() - t
    nS = m.n_species
    Yc, Yg = ref[0], gpu_Ynew[:, :sample]
    errY = np.abs(Yg[:nS] - Yc[:nS]) / (1e-6 * np.abs(Yc[:nS]) + 1e-10)
    errT = np.abs(Yg[nS] - Yc[nS]) / Yc[nS]
    return {"value": sample / el, "unit": "cells/s", "cores": threads, "kind": kind,
            "sample": f"first {sample} cells of the benchmarked {config} cells, one CFD "
                      f"step ({el:.1f} s wall on {threads} threads)",
            "max_err_over_tol": float(errY.max()), "max_T_rel": float(errT.max()),
            "error_definition": "GPU batched result vs this CPU result on the sample: "
                                "max |dY|/(1e-6|Y|+1e-10) over species, max |dT|/T"}


def adjoint_cost(n_cells: int, config: str = "h2_adjoint", deltaT: float = 0.0) -> dict:
    """Adjoint/forward cost on a workload (by default the differentiation
    workload, H2 with deltaT 1e-4): recording forward + reverse sweep versus
    the plain and the batched forward solve of the same cells; loss
    sum(T + 1e3 Y_OH), or sum(T) for mechanisms without OH.  deltaT > 0
    replaces the workload's step."""
    import torch
    from gpufoam_amd import _capi, ode_inputs
    from gpufoam_amd.evaluator import Evaluator
    from gpufoam_amd.workloads import make_states
    m, Y0, dt0, c = make_states(config, n_cells)
    if deltaT > 0:
        c = dict(c, deltaT=deltaT)
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    ev = Evaluator(m, ode, 0.0, max_cells=n_cells)
    dY0 = torch.tensor(Y0, device="cuda:0")
    dYn = torch.empty_like(dY0)
    ddt0 = torch.tensor(dt0, device="cuda:0")
    n = m.n_species + 1
    W = torch.zeros((n, n_cells), dtype=torch.float64, device="cuda:0")
    W[m.n_species] = 1.0
    has_oh = "OH" in m.species
    if has_oh:
        W[m.index("OH")] = 1e3
    g = torch.empty_like(W)
    pb = torch.empty((6 * m.n_reactions, n_cells), dtype=torch.float64, device="cuda:0")

    def timed(fn):
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3

    ddt = ddt0.clone()
    ev.solve(c["deltaT"], dY0[:, :64], ddt[:64], Ynew=dYn[:, :64], ldY=n_cells)  # warm
    ddt.copy_(ddt0)
    t_plain = timed(lambda: ev.solve(c["deltaT"], dY0, ddt, Ynew=dYn))
    progress(f"adjoint workload: drop-in forward {t_plain:.2f} s")
    ddt.copy_(ddt0)
    t_batched = timed(lambda: ev.solve(c["deltaT"], dY0, ddt, Ynew=dYn, batched=True))
    progress(f"adjoint workload: batched forward {t_batched:.2f} s")
    ddt.copy_(ddt0)
    t_fwd = timed(lambda: ev.adjoint_forward(c["deltaT"], dY0, ddt, Ynew=dYn))
    progress(f"adjoint workload: recording forward {t_fwd:.2f} s")
    t_bwd = timed(lambda: ev.adjoint_backward(W, g, pb))
    progress(f"adjoint workload: reverse sweep {t_bwd:.2f} s")
    return {"workload": config, "cells": n_cells, "deltaT": c["deltaT"],
            "loss": "sum(T + 1e3*Y_OH)" if has_oh else "sum(T)",
            "params": 6 * m.n_reactions,
            "forward_plain_s": t_plain, "forward_batched_s": t_batched,
            "adjoint_forward_s": t_fwd, "adjoint_reverse_s": t_bwd,
            "checkpoint_interval": ev.checkpoint_interval,
            "adjoint_over_batched_forward": (t_fwd + t_bwd) / t_batched,
            "adjoint_over_plain_forward": (t_fwd + t_bwd) / t_plain,
            "target": "adjoint_over_batched_forward <= 4"}


def batching_block(n_cells: int, graph: bool) -> dict:
    """The batching workload: persistent work-stealing launch (replayed as a
    hipGraph) against the drop-in launch of the same kernel, one CFD step
    each on the same cells, and the per-cell attempt histogram."""
    import torch
    from gpufoam_amd import _capi, ode_inputs
    from gpufoam_amd.evaluator import Evaluator
    from gpufoam_amd.workloads import make_states
    m, Y0, dt0, c = make_states("gri_batching", n_cells)
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    ev = Evaluator(m, ode, 0.0, max_cells=n_cells)
    dY0 = torch.tensor(Y0, device="cuda:0")
    del Y0
    dYn = torch.empty_like(dY0)
    ddt0 = torch.tensor(dt0, device="cuda:0")
    ddt = ddt0.clone()
    st = torch.zeros((3, n_cells), dtype=torch.int32, device="cuda:0")
    ev.solve(c["deltaT"], dY0[:, :256], ddt[:256], Ynew=dYn[:, :256], ldY=n_cells)  # warm
    ev.solve(c["deltaT"], dY0[:, :256], ddt[:256], Ynew=dYn[:, :256], ldY=n_cells,
             batched=True)
    torch.cuda.synchronize()

    def timed(batched):
        ddt.copy_(ddt0)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        ev.solve(c["deltaT"], dY0, ddt, Ynew=dYn, stats=st, batched=batched,
                 graph=graph and batched)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3

    t_b = timed(True)
    progress(f"batching workload: persistent {t_b:.2f} s")
    att = st[1].cpu().numpy()
    t_p = timed(False)
    progress(f"batching workload: drop-in {t_p:.2f} s")
    q = np.percentile(att, [50, 90, 99])
    return {"workload": "gri_batching", "cells": n_cells, "deltaT": c["deltaT"],
            "batched_s": t_b, "plain_s": t_p,
            "batched_cells_per_s": n_cells / t_b, "plain_cells_per_s": n_cells / t_p,
            "batched_over_plain": t_p / t_b,
            "attempts": {"mean": float(att.mean()), "p50": float(q[0]), "p90": float(q[1]),
                         "p99": float(q[2]), "max": int(att.max())}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="gri_production")
    ap.add_argument("--cells", type=int, default=0, help="override the cell count")
    ap.add_argument("--plain-sample", type=int, default=0,
                    help="cells for the drop-in and lane-kernel lines (0: the whole workload)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-lane", action="store_true",
                    help="skip the one-lane-per-cell line")
    ap.add_argument("--batching-cells", type=int, default=1048576,
                    help="cells of the batching workload block (0: skip)")
    ap.add_argument("--cpu-sample", type=int, default=65536)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--cells-per-wave", type=int, choices=[0, 1, 2, 4, 64], default=0,
                    help="launch shape of the batched solve: 0 automatic, 1 a wavefront per "
                         "cell, 2 two wavefronts per cell, 4 four cells per wavefront, 64 one "
                         "lane per cell")
    ap.add_argument("--queue-chunk", type=int, choices=[0, 1], default=0,
                    help="cells per work-queue ticket of the wavefront kernel: 0 automatic "
                         "(16-cell chunks), 1 one cell per ticket")
    ap.add_argument("--factorization", choices=["column", "panel", "mfma"], default=None,
                    help="LU variant of the workgroup-per-cell kernel (64 < nSpecie+1 <= 208)")
    ap.add_argument("--adjoint-cells", type=int, default=65536,
                    help="cells of the differentiation workload for the adjoint cost (0: skip)")
    ap.add_argument("--adjoint-config", default="h2_adjoint",
                    help="workload of the adjoint cost (default: the differentiation workload)")
    ap.add_argument("--adjoint-gri-cells", type=int, default=16384,
                    help="cells of the 53-species workload for a second adjoint cost (0: skip)")
    ap.add_argument("--adjoint-dt", type=float, default=0.0,
                    help="CFD step of the adjoint cost in place of the workload's (0: keep)")
    ap.add_argument("--autotune", type=int, default=0, metavar="CELLS",
                    help="run the launch-shape autotuner on CELLS cells of the workload "
                         "first and time the batched solve with the shape it picks")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB three times "
                         "(the known-size kernel that scales the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the chemistry of each cell is "
              "independent, so N GPUs would each run this same kernel on their "
              "share of cells; no multi-GPU figure is produced.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_calibration:
        nb = args.copy_calibration << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": nb, "copy_bytes_written": nb, "calls": 3}))
        return
    from gpufoam_amd import _capi, ode_inputs
    from gpufoam_amd.evaluator import Evaluator
    from gpufoam_amd.workloads import CONFIGS, make_states
    from gpufoam_amd.workmodel import attempt_flops, work_flops

    cfg = CONFIGS[args.config]
    n = args.cells or cfg["cells"]
    m, Y0, dt0, c = make_states(args.config, n)
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    dev = torch.device("cuda:0")
    dY0 = torch.tensor(Y0, device=dev)
    dYn = torch.empty_like(dY0)
    ddt0 = torch.tensor(dt0, device=dev)
    ddt = ddt0.clone()
    dmin = torch.zeros(1, dtype=torch.float64, device=dev)
    stats = torch.zeros((3, n), dtype=torch.int32, device=dev)
    ev = Evaluator(m, ode, 0.0, max_cells=n)
    if args.queue_chunk or args.cells_per_wave:
        ev.set_batched_shape(args.cells_per_wave, 0, args.queue_chunk)
    if args.factorization:
        ev.set_factorization({"column": _capi.LU_COLUMN, "panel": _capi.LU_PANEL,
                              "mfma": _capi.LU_PANEL_MFMA}[args.factorization])
    graph = not args.no_graph
    tuned = None
    if args.autotune:
        from gpufoam_amd.autotune import autotune
        progress(f"autotuning on {args.autotune} cells")
        times = autotune(ev, c["deltaT"], dY0, ddt0, sample=args.autotune, repeats=2)
        best = times.pop("best")
        best_lu = times.pop("best_lu", None)
        tuned = {"sample_cells": args.autotune, "best": list(best),
                 "ms": {f"{k[0]}x{k[1]}": round(v, 3) for k, v in times.items()}}
        if best_lu is not None:
            tuned["best_factorization"] = ["column", "panel", "mfma"][best_lu]
        progress(f"autotuner: {tuned}")

    def step(with_stats=False):
        ddt.copy_(ddt0)
        ev.solve(c["deltaT"], dY0, ddt, Ynew=dYn, deltaTMin=dmin,
                 stats=stats if with_stats else None, batched=True, graph=graph)

    # first call with statistics (work accounting), then warmup
    progress(f"{args.config}: {n} cells, statistics run")
    step(with_stats=True)
    torch.cuda.synchronize()
    st = stats.cpu().numpy()
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = e0.elapsed_time(e1) / args.steps
    value = n / (ms_per_step / 1e3)
    progress(f"batched: {ms_per_step:.1f} ms per step")

    stages = {12: 2, 23: 3, 34: 4}[ode.solver]
    flops = work_flops(m, st[1], st[2], stages)
    achieved = flops / (ms_per_step / 1e3) / 1e12

    Yg = dYn.cpu().numpy()

    # the drop-in launch and the one-lane-per-cell kernel on the same states
    def plain_line(lane: bool):
        ns = min(args.plain_sample or n, n)
        dY0p = dY0[:, :ns].contiguous()
        dYnp = torch.empty_like(dY0p)
        ddtp = ddt0[:ns].clone()
        ev.solve(c["deltaT"], dY0p[:, :64], ddtp[:64], Ynew=dYnp[:, :64], ldY=ns,
                 lane=lane)  # warm
        torch.cuda.synchronize()
        ddtp.copy_(ddt0[:ns])
        p0, p1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        p0.record(stream)
        ev.solve(c["deltaT"], dY0p, ddtp, Ynew=dYnp, ldY=ns, lane=lane)
        p1.record(stream)
        torch.cuda.synchronize()
        return ns / (p0.elapsed_time(p1) / 1e3), ns

    plain_value = lane_value = None
    ns = nl = 0
    if not args.no_plain:
        plain_value, ns = plain_line(False)
        progress(f"drop-in: {plain_value:.0f} cells/s")
        if not args.no_lane:
            lane_value, nl = plain_line(True)
            progress(f"one lane per cell: {lane_value:.0f} cells/s")

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(m, Y0, dt0, c, args.config, min(args.cpu_sample, n), threads, Yg)
        progress(f"cpu baseline: {cpu['value']:.0f} cells/s on {threads} threads; "
                 f"GPU error {cpu['max_err_over_tol']:.3g} of tol, T {cpu['max_T_rel']:.3g}")
    del Yg

    adjoint = None
    if args.adjoint_cells:
        progress(f"adjoint cost on {args.adjoint_cells} cells of {args.adjoint_config}")
        adjoint = adjoint_cost(args.adjoint_cells, args.adjoint_config, args.adjoint_dt)
    adjoint_gri = None
    if args.adjoint_gri_cells:
        progress(f"adjoint cost on {args.adjoint_gri_cells} cells of gri_production")
        adjoint_gri = adjoint_cost(args.adjoint_gri_cells, "gri_production")

    batching = None
    if args.batching_cells:
        del dY0, dYn
        ev = None
        torch.cuda.empty_cache()
        progress(f"batching workload: {args.batching_cells} cells")
        batching = batching_block(args.batching_cells, graph)

    traffic = None
    tfile = os.path.join(HERE, "profiles", f"hbm_traffic_{args.config}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("bytes_per_launch")

    out = {
        "metric": "chem cells/s per CFD \u0394t, 53sp FP64 rtol1e-6: plain vs batched; adjoint/forward cost",
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.config, "mechanism": cfg["mech"],
                   "species": m.n_species, "reactions": m.n_reactions, "cells": n,
                   "solver": c["solver"], "relTol": c["relTol"], "absTol": c["absTol"],
                   "deltaT": c["deltaT"], "graph": graph,
                   "factorization": args.factorization or "default",
                   "queue_chunk": args.queue_chunk,
                   "cells_per_wave": args.cells_per_wave or "automatic"},
        "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": traffic,
                     "work_model_flops_per_attempt": attempt_flops(m, stages)["attempt"],
                     "work_model": "operation counts of the CPU restatement: +,-,*,/ and "
                                   "each exp/log/pow = 1 flop; one Jacobian, LU and "
                                   f"{stages} solves per attempt, one extra RHS per accepted "
                                   "step, from the kernel's own attempt/accept counters; "
                                   "peak = MI355X FP64 vector (no MFMA on this path)"},
        "cpu_baseline": cpu,
        "max_err_over_tol": cpu["max_err_over_tol"] if cpu else None,
        "max_T_rel": cpu["max_T_rel"] if cpu else None,
        "plain_cells_per_s": plain_value,
        "plain_sample_cells": ns,
        "batched_over_plain": value / plain_value if plain_value else None,
        "lane_cells_per_s": lane_value,
        "lane_sample_cells": nl,
        "batched_over_lane": value / lane_value if lane_value else None,
        "batching": batching,
        "adjoint": adjoint,
        "adjoint_gri": adjoint_gri,
        "autotune": tuned,
        "attempts_per_cell": float(st[1].mean()),
        "accepted_per_cell": float(st[2].mean()),
        "wall_s": wall,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
