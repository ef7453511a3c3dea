#!/usr/bin/env python3
"""Throughput of the explicit MPM step on one MI355X.

Workload: the 3D granular column collapse of 2,048,000 particles (0.4 x 0.4 x
0.2 m Drucker-Prager column, h = 5 mm, 2x2x2 particles per cell, 1.2 x 1.2 x
0.25 m box, dt = 0.2 h / c_p), FP32 bulk with FP64 positions and det F,
particles cell-sorted every 10 steps.  A step is one full MUSL timestep of
every particle (both P2G passes, both grid solves, G2P, Drucker-Prager).

Prints one JSON line: particle-steps/s over the timed steps, the roofline of
the dominant kernel (algorithmic bytes per launch / its mean HIP-event
duration on the context stream, against the 8 TB/s HBM peak) and the CPU
baseline (the FP64 oracle restatement of the reference's CPU path, OpenMP on
the host cores, on a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s (3D DP column, %HBM peak); batched & surrogate vs plain at tol"

# Algorithmic bytes (FP32 bulk, FP64 positions and det F, 3D): per particle and
# per active node, for each kernel of the step.  k_g2p2 includes the +64 B of
# the ϵijs / ϵv / ϵk outputs the download contract needs every step.
ALGO_BYTES = {
    "k_p2g": (68, 28),
    "k_grid1": (0, 77),
    "k_g2p1_p2g2": (64, 48),
    "k_grid2": (0, 29),
    "k_g2p2": (208, 12),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--precision", default="FP32", choices=["FP32", "FP64"])
    ap.add_argument("--sort-interval", type=int, default=10)
    ap.add_argument("--h", type=float, default=0.005, help="grid spacing (0.005 = 2,048,000 particles)")
    ap.add_argument("--cpu-steps", type=int, default=6, help="oracle steps for the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(cores, 16)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    return ap.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {a.gpus}); "
              "multi-GPU sharding is described in DESIGN.md but not built.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import __graft_entry__ as ge

    mpm = ge.load_package()
    t_setup = time.time()
    args, grid, mp, attr, bc = mpm.scenes.column_collapse_3d(h=a.h, eps=a.precision)
    dt = float(args.ΔT)
    model = mpm.HIPModel(grid, mp, attr, bc, sort_interval=a.sort_interval)
    sargs = model.step_args(args)
    print(f"# setup {time.time() - t_setup:.1f}s: np={mp.np} ni={grid.ni} dt={dt:.4e}",
          file=sys.stderr, flush=True)

    Ti = model.run(sargs, dt, 0.0, a.warmup)
    model.sync()
    n_active = model.count_active_nodes()

    model.profile_enable(True)
    model.sync()
    t0 = time.perf_counter()
    Ti = model.run(sargs, dt, Ti, a.steps)
    model.sync()
    t1 = time.perf_counter()
    prof = model.profile_read()
    model.profile_enable(False)
    elapsed = t1 - t0
    n = int(mp.np)
    value = n * a.steps / elapsed
    print(f"# timed {a.steps} steps in {elapsed:.3f}s, active nodes {n_active}", file=sys.stderr,
          flush=True)

    # roofline of the dominant kernel
    kernels = {}
    for name, (ms, cnt) in prof.items():
        if cnt == 0:
            continue
        per_p, per_n = ALGO_BYTES.get(name, (0, 0))
        algo = per_p * n + per_n * n_active
        avg_s = ms * 1e-3 / cnt
        kernels[name] = dict(avg_us=avg_s * 1e6, launches=cnt, total_ms=ms,
                             algo_bytes=algo, gbs=(algo / avg_s / 1e9) if algo else None)
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    dk = kernels[dom]
    for name, k in sorted(kernels.items(), key=lambda kv: -kv[1]["total_ms"]):
        g = f"{k['gbs']:.0f} GB/s" if k["gbs"] else "-"
        print(f"#   {name:14s} {k['avg_us']:9.1f} us x {k['launches']:5d}  {g}",
              file=sys.stderr, flush=True)

    step_bytes = sum(ALGO_BYTES[k][0] * n + ALGO_BYTES[k][1] * n_active for k in ALGO_BYTES)
    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(a, args, grid, mp, attr, bc)
    model.close()

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if a.precision == "FP32" else "f64",
        "data": "synthetic",
        "config": {
            "workload": "3D granular column collapse, Drucker-Prager, uGIMP MUSL, "
                        f"{n} particles, h={a.h} m, 2x2x2 ppc, grid {grid.nnx}x{grid.nny}x{grid.nnz} "
                        f"nodes, dt={dt:.4e} s",
            "precision": a.precision + (" bulk, FP64 positions and det F" if a.precision == "FP32" else ""),
            "particles": n,
            "grid_nodes": int(grid.ni),
            "active_nodes": n_active,
            "sort_interval": a.sort_interval,
            "step_algorithmic_bytes": step_bytes,
            "step_hbm_fraction": step_bytes / (elapsed / a.steps) / (HBM_PEAK_GBS * 1e9),
            "kernels_us": {k: round(v["avg_us"], 2) for k, v in kernels.items()},
        },
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": dk["gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (dk["gbs"] / HBM_PEAK_GBS) if dk["gbs"] else None,
            "traffic": None,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(a, args, grid, mp, attr, bc):
    """The FP64 oracle (CPU restatement of the reference step) on the same scene."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    threads = a.cpu_threads or min(os.cpu_count() or 1, 16)
    o = oracle.Oracle(args, grid, mp, attr, bc, nthreads=threads)
    dt = float(args.ΔT)
    o.run(dt, 0.0, 1)  # first touch of every array
    t0 = time.perf_counter()
    o.run(dt, dt, a.cpu_steps)
    el = time.perf_counter() - t0
    n = int(mp.np)
    return {"value": n * a.cpu_steps / el, "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"same 3D column ({n} particles), {a.cpu_steps} FP64 steps after one "
                      f"untimed step, OpenMP {threads} threads, {el:.1f} s"}


if __name__ == "__main__":
    main()
