#!/usr/bin/env python3
"""Throughput of the explicit MPM step on one MI355X.

Workload: the 3D granular column collapse of 2,048,000 particles (0.4 x 0.4 x
0.2 m Drucker-Prager column, h = 5 mm, 2x2x2 particles per cell, 1.2 x 1.2 x
0.25 m box, dt = 0.2 h / c_p), FP32 bulk with FP64 positions and det F,
particles cell-sorted every 10 steps.  A step is one full MUSL timestep of
every particle (both P2G passes, both grid solves, G2P, Drucker-Prager).

Prints one JSON line: particle-steps/s over the timed steps, the roofline of
P2G #1 (algorithmic bytes per launch / its mean HIP-event duration on the
context stream inside the timed region, against the 8 TB/s HBM peak) with
G2P #2 beside it, and the CPU baseline (the FP64 oracle restatement of the
reference's CPU path, OpenMP on every core this process may use and on one
thread, on a bounded sample of the same workload).
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
    "k_g2p1_p2g2": (64, 36),  # re-map gather + momentum scatter in one pass (24 + 12 per node)
    "k_p2g2": (0, 12),        # re-map momentum scatter when split (its particle reads are overhead)
    "k_grid2": (0, 29),
    "k_g2p2": (208, 12),
}
# Kernels outside the step's compulsory traffic (the sort is an implementation
# choice and counts against the step's fraction): their own minimum bytes, for
# the per-kernel table only.  k_sort: positions in, key and rank out, then key
# and rank in, order and sorted key out (48 B); k_permute: every permuted FP32
# stream read and written once (124 B each way) plus the order (4 B); MVL halves
# and the learned return: the particle fields each reads and writes.
AUX_BYTES = {
    "k_sort": (48, 0),
    "k_permute": (252, 0),
    "k_mvl1": (56, 0),
    "k_mvl2": (48, 0),
    "k_learned_return": (88, 0),
}
# The kernel the roofline figure is quoted for: P2G #1, the per-kernel gate of
# the target (P2G and G2P at >= 50 % of HBM peak); G2P #2 is reported beside it.
ROOFLINE_KERNEL = "k_p2g"
SECOND_KERNEL = "k_g2p2"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--advance", type=int, default=0,
                    help="untimed steps before the warm-up (to measure a later phase of the collapse)")
    ap.add_argument("--precision", default="FP32", choices=["FP32", "FP64"])
    ap.add_argument("--sort-interval", type=int, default=10)
    ap.add_argument("--h", type=float, default=0.005, help="grid spacing (0.005 = 2,048,000 particles)")
    ap.add_argument("--cpu-steps", type=int, default=40,
                    help="oracle steps for the CPU baseline (about 12 s on 16 cores)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0 = every core this process may run on (sched_getaffinity)")
    ap.add_argument("--cpu-steps-1t", type=int, default=2,
                    help="oracle steps of the single-thread CPU sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc child runs (FETCH_SIZE, WRITE_SIZE)")
    ap.add_argument("--no-autotune", action="store_true", help="skip the autotuner report")
    ap.add_argument("--no-full-run", action="store_true",
                    help="skip the whole 20,000-step collapse (captured graphs, no events)")
    ap.add_argument("--full-steps", type=int, default=20_000)
    ap.add_argument("--full-sort-interval", type=int, default=5,
                    help="sort interval of the whole run (5 is its optimum over 3..15 on MI355X)")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--ens-instances", type=int, default=256)
    ap.add_argument("--ens-steps", type=int, default=200)
    ap.add_argument("--ens-seq-steps", type=int, default=40)
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned-return leg")
    ap.add_argument("--sur-steps", type=int, default=200)
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the FP64 / volumetric-locking variants of the timed workload")
    ap.add_argument("--no-schedules", action="store_true",
                    help="skip the USL/USF schedule leg (same workload under the other schedules)")
    ap.add_argument("--no-slope", action="store_true",
                    help="skip the 20M-particle GRF slope leg (exact and learned return)")
    ap.add_argument("--slope-h", type=float, default=0.5, help="slope grid spacing (0.5 = 20M particles)")
    ap.add_argument("--slope-k0", type=int, default=400, help="damped K0 steps before the runout")
    ap.add_argument("--slope-steps", type=int, default=200, help="timed runout steps per form")
    ap.add_argument("--scatter", default=None,
                    choices=["rounds", "rounds_layered", "owned", "owned_layers", "pairs"],
                    help="tile-scatter form (default: the library's)")
    ap.add_argument("--remap-scatter", default=None,
                    choices=["rounds", "rounds_layered", "owned", "owned_layers", "pairs"],
                    help="the re-map momentum scatter's form (default: the library's)")
    ap.add_argument("--fused-remap", type=int, default=None, choices=[0, 1],
                    help="one-pass MUSL re-map (default: the library's)")
    ap.add_argument("--tile-partials", type=int, default=None, choices=[0, 1],
                    help="tile partials summed by the grid solves (default: the library's)")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    return ap.parse_args()


# Counter calibration: FETCH_SIZE / WRITE_SIZE of the streaming triad kernel
# (a = b + s c over three float4 arrays of TRIAD_BYTES each: reads 2x, writes 1x).
TRIAD_BYTES = 256 << 20
KERNEL_RE = {
    "k_p2g": r"mpm::k_p2g<|mpm::k_scatter_(tiled|owned)<\d, \d, \d, \w+, 0[,>]|mpm::k_scatter_pairs<\d, \d, 0>",
    "k_grid1": r"mpm::k_grid1<",
    "k_g2p1_p2g2": (r"mpm::k_g2p1(_p2g2)?(_tiled)?<|mpm::k_scatter_tiled<\d, \d, \d, \w+, [23][,>]"
                    r"|mpm::k_remap_owned<"),
    "k_p2g2": r"mpm::k_scatter_(tiled|owned)<\d, \d, \d, \w+, 1[,>]|mpm::k_scatter_pairs<\d, \d, 1>",
    "k_grid2": r"mpm::k_grid2<",
    "k_g2p2": r"mpm::k_g2p2(_tiled)?<",
}


def apply_options(a, model):
    """The kernel-variant options given on the command line."""
    if a.scatter is not None:
        model.set_option("scatter", a.scatter)
    if a.remap_scatter is not None:
        model.set_option("remap_scatter", a.remap_scatter)
    if a.fused_remap is not None:
        model.set_option("fused_remap", a.fused_remap)
    if a.tile_partials is not None:
        model.set_option("tile_partials", a.tile_partials)


def variant_args(a):
    out = []
    for flag, v in (("--scatter", a.scatter), ("--remap-scatter", a.remap_scatter),
                    ("--fused-remap", a.fused_remap),
                    ("--tile-partials", a.tile_partials)):
        if v is not None:
            out += [flag, str(v)]
    return out


def pmc_child(a):
    """Runs under rocprofv3 --pmc: a few steps of the workload, then the triad."""
    import __graft_entry__ as ge

    mpm = ge.load_package()
    args, grid, mp, attr, bc = mpm.scenes.column_collapse_3d(h=a.h, eps=a.precision)
    model = mpm.HIPModel(grid, mp, attr, bc, sort_interval=a.sort_interval)
    apply_options(a, model)
    sargs = model.step_args(args)
    model.run(sargs, float(args.ΔT), 0.0, 14)
    model.sync()
    model.close()
    print("# pmc child: 14 steps done", flush=True)
    mpm.Tpeak(nbytes=TRIAD_BYTES, reps=2)
    print("# pmc child: triad done", flush=True)


def counter_means(csv_path):
    """Mean counter value (KiB) per kernel-name regex over its dispatches."""
    import csv
    import re

    vals = {}
    with open(csv_path) as f:
        for r in csv.DictReader(f):
            vals.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    out = {}
    pats = dict(KERNEL_RE, triad=r"k_triad")
    for key, pat in pats.items():
        xs = [v for name, vs in vals.items() if re.search(pat, name) for v in vs]
        if xs:
            xs.sort()
            out[key] = xs[len(xs) // 2]  # median dispatch
    return out


def measure_traffic(a):
    """HBM bytes per launch of each step kernel from two rocprofv3 --pmc child
    runs (FETCH_SIZE, then WRITE_SIZE), each scaled by what the same counter
    reads for the triad kernel of known size.  None when unavailable."""
    import glob
    import shutil
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None, "rocprofv3 not found"
    env = dict(os.environ, TMPDIR="/tmp")
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="mpm_pmc_", dir="/tmp")
        cmd = ["timeout", "-s", "KILL", "120", prof, "--pmc", counter, "--output-format", "csv",
               "-d", d, "-o", "run", "--", sys.executable, os.path.abspath(__file__), "--pmc-child",
               "--h", str(a.h), "--precision", a.precision, "--sort-interval", str(a.sort_interval),
               *variant_args(a)]
        print(f"# traffic: rocprofv3 --pmc {counter}", file=sys.stderr, flush=True)
        r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True)
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            shutil.rmtree(d, ignore_errors=True)
            return None, f"{counter}: rc={r.returncode}, no counter file; " + r.stderr[-300:]
        res[counter] = counter_means(files[0])
        shutil.rmtree(d, ignore_errors=True)
    fe, wr = res["FETCH_SIZE"], res["WRITE_SIZE"]
    if "triad" not in fe or "triad" not in wr:
        return None, "triad dispatch missing from the counter files"
    f_fetch = 2 * TRIAD_BYTES / (fe["triad"] * 1024.0)
    f_write = TRIAD_BYTES / (wr["triad"] * 1024.0)
    out = {"fetch_scale": f_fetch, "write_scale": f_write, "kernels": {}}
    for k in KERNEL_RE:
        if k in fe and k in wr:
            out["kernels"][k] = {"fetch_bytes": fe[k] * 1024.0 * f_fetch,
                                 "write_bytes": wr[k] * 1024.0 * f_write}
    return out, ""


def full_run_leg(a, mpm):
    """The whole collapse from rest, 20,000 steps (the horizon the 3D column is
    defined with), with captured step graphs and no per-kernel events: the
    wall-clock cost of the job.  The spreading deposit costs more per step than
    the first steps that the timed region above covers (fewer particles per
    tile, rounds with fewer lanes), so this is the job-level figure."""
    import numpy as np
    args, grid, mp, attr, bc = mpm.scenes.column_collapse_3d(h=a.h, eps=a.precision)
    dt = float(args.ΔT)
    m = mpm.HIPModel(grid, mp, attr, bc, sort_interval=a.full_sort_interval)
    try:
        sa = m.step_args(args)
        seg, Ti, done = [], 0.0, 0
        m.sync()
        t0 = time.perf_counter()
        while done < a.full_steps:
            k = min(2000, a.full_steps - done)
            ts = time.perf_counter()
            Ti = m.run(sa, dt, Ti, k, graph=True)
            m.sync()
            seg.append(round((time.perf_counter() - ts) / k * 1e3, 4))
            done += k
            print(f"# full run: {done} steps", file=sys.stderr, flush=True)
        T = time.perf_counter() - t0
        x = m.download()["xi"]
    finally:
        m.close()
    n = int(mp.np)
    x0 = np.asarray(mp.ξ, dtype=np.float64)
    res = {"steps": a.full_steps, "seconds": T, "ms_per_step": T / a.full_steps * 1e3,
           "particle_steps_per_s": n * a.full_steps / T, "sort_interval": a.full_sort_interval,
           "graph": True, "ms_per_step_per_2000": seg,
           "runout_m": float(np.abs(x[:, :2]).max() - np.abs(x0[:, :2]).max())}
    print(f"# full run: {a.full_steps} steps in {T:.2f} s, {res['ms_per_step']:.4f} ms/step, "
          f"{res['particle_steps_per_s']:.3e} particle-steps/s", file=sys.stderr, flush=True)
    return res


def main():
    a = parse()
    if a.pmc_child:
        pmc_child(a)
        return
    if a.gpus != 1:
        print(f"bench.py: the solver runs on one GPU only (asked for --gpus {a.gpus}); "
              "spatial-slab sharding across GPUs is not built.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import __graft_entry__ as ge

    mpm = ge.load_package()
    t_setup = time.time()
    args, grid, mp, attr, bc = mpm.scenes.column_collapse_3d(h=a.h, eps=a.precision)
    dt = float(args.ΔT)
    model = mpm.HIPModel(grid, mp, attr, bc, sort_interval=a.sort_interval)
    apply_options(a, model)
    sargs = model.step_args(args)
    print(f"# setup {time.time() - t_setup:.1f}s: np={mp.np} ni={grid.ni} dt={dt:.4e}",
          file=sys.stderr, flush=True)

    n = int(mp.np)
    Ti = 0.0
    if a.advance > 0:
        # eager launches: the profiler's kernel tracing does not take captured graphs here
        Ti = model.run(sargs, dt, Ti, a.advance)
        print(f"# advanced {a.advance} steps", file=sys.stderr, flush=True)
    Ti = model.run(sargs, dt, Ti, a.warmup)
    model.sync()
    n_active = model.count_active_nodes()
    dom = ROOFLINE_KERNEL

    def table(prof):
        kern = {}
        split = prof.get("k_p2g2", (0, 0))[1] > 0
        for name, (ms, cnt) in prof.items():
            if cnt == 0:
                continue
            per_p, per_n = ALGO_BYTES.get(name) or AUX_BYTES.get(name, (0, 0))
            if name == "k_g2p1_p2g2" and split:
                per_n -= ALGO_BYTES["k_p2g2"][1]  # the gather alone: vsT, as
            algo = per_p * n + per_n * n_active
            avg_s = ms * 1e-3 / cnt
            kern[name] = dict(avg_us=avg_s * 1e6, launches=cnt, total_ms=ms,
                              algo_bytes=algo, gbs=(algo / avg_s / 1e9) if algo else None)
        return kern

    # timed region: HIP events (on the context stream) around P2G #1's launches only
    model.set_option("profile_kernels", [dom])
    model.profile_enable(True)
    model.sync()
    t0 = time.perf_counter()
    Ti = model.run(sargs, dt, Ti, a.steps)
    model.sync()
    t1 = time.perf_counter()
    prof = model.profile_read()
    model.profile_enable(False)
    model.set_option("profile_kernels", None)
    elapsed = t1 - t0
    value = n * a.steps / elapsed
    print(f"# timed {a.steps} steps in {elapsed:.3f}s, active nodes {n_active}", file=sys.stderr,
          flush=True)
    dk = table(prof)[dom]

    # per-kernel table: the same number of steps right after the timed region,
    # with events on every kernel
    model.profile_enable(True)
    model.sync()
    t0 = time.perf_counter()
    Ti = model.run(sargs, dt, Ti, a.steps)
    model.sync()
    ms_all_events = (time.perf_counter() - t0) / a.steps * 1e3
    kernels = table(model.profile_read())
    model.profile_enable(False)
    print(f"# roofline kernel {dom}: {dk['avg_us']:.1f} us in the timed region; per-kernel table of "
          f"the next {a.steps} steps ({ms_all_events:.4f} ms/step with events on every kernel):",
          file=sys.stderr, flush=True)
    for name, k in sorted(kernels.items(), key=lambda kv: -kv[1]["total_ms"]):
        g = f"{k['gbs']:.0f} GB/s" if k["gbs"] else "-"
        print(f"#   {name:14s} {k['avg_us']:9.1f} us x {k['launches']:5d}  {g}",
              file=sys.stderr, flush=True)

    # the achievable streaming bandwidth on this GPU (triad a = b + c, 1 GiB arrays)
    probe_gbs = mpm.Tpeak(nbytes=1 << 30, reps=5)
    print(f"# triad probe: {probe_gbs:.0f} GB/s", file=sys.stderr, flush=True)
    # k_g2p1_p2g2's entry already holds the re-map scatter's 12 B per node
    step_bytes = sum(ALGO_BYTES[k][0] * n + ALGO_BYTES[k][1] * n_active for k in ALGO_BYTES
                     if k != "k_p2g2")
    tuned = None
    if not a.no_autotune:
        best, tab, Ti = mpm.autotune(model, sargs, dt, Ti, steps=40, warm=10,
                                     scatters=mpm.SCATTER_FORMS + (("rounds", "rounds"),))

        def _name(sc):
            return "+".join(sc) if isinstance(sc, tuple) else sc

        tuned = {"best": dict(best, scatter=_name(best["scatter"]) if best["scatter"] else None),
                 "ms_per_step": {f"sort{k}_{_name(sc) if t else 'plain'}": round(s * 1e3, 4)
                                 for k, t, sc, s in tab},
                 "note": "a form name sets P2G #1's scatter (the re-map's automatic); "
                         "'a+b' sets P2G #1's to a and the re-map's to b"}
        print(f"# autotune: {best}", file=sys.stderr, flush=True)
    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(a, args, grid, mp, attr, bc)
    model.close()
    traffic, traffic_note = None, "--no-traffic"
    if not a.no_traffic:
        traffic, traffic_note = measure_traffic(a)
        if traffic is None:
            print(f"# traffic unavailable: {traffic_note}", file=sys.stderr, flush=True)
    dom_traffic = sec_traffic = None
    if traffic and dom in traffic["kernels"]:
        t = traffic["kernels"][dom]
        dom_traffic = t["fetch_bytes"] + t["write_bytes"]
    if traffic and SECOND_KERNEL in traffic["kernels"]:
        t = traffic["kernels"][SECOND_KERNEL]
        sec_traffic = t["fetch_bytes"] + t["write_bytes"]
    if traffic:
        for k, t in traffic["kernels"].items():
            print(f"#   traffic {k:14s} fetch {t['fetch_bytes'] / 1e6:8.1f} MB  write "
                  f"{t['write_bytes'] / 1e6:8.1f} MB  (algorithmic "
                  f"{kernels[k]['algo_bytes'] / 1e6 if k in kernels else 0:.1f} MB)",
                  file=sys.stderr, flush=True)

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
            "first_timed_step": a.advance + a.warmup,
            "grid_nodes": int(grid.ni),
            "active_nodes": n_active,
            "sort_interval": a.sort_interval,
            "kernel_variant": {"scatter": a.scatter or "pairs (default in FP32 3D)",
                               "remap_scatter": a.remap_scatter or "owned (default in FP32 3D)",
                               "fused_remap": bool(a.fused_remap) if a.fused_remap is not None
                               else "off (default)",
                               "tile_partials": bool(a.tile_partials) if a.tile_partials is not None
                               else "on (default)"},
            "step_algorithmic_bytes": step_bytes,
            "step_hbm_fraction": step_bytes / (elapsed / a.steps) / (HBM_PEAK_GBS * 1e9),
            "hbm_triad_probe_gbs": probe_gbs,
            "kernels_us": {k: round(v["avg_us"], 2) for k, v in kernels.items()},
            "kernels_us_window": f"the {a.steps} steps after the timed region, events on every "
                                 f"kernel ({ms_all_events:.4f} ms/step)",
        },
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": dk["gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (dk["gbs"] / HBM_PEAK_GBS) if dk["gbs"] else None,
            "traffic": dom_traffic,
            "traffic_unit": "bytes per launch (FETCH_SIZE + WRITE_SIZE, triad-calibrated)",
            "algorithmic_bytes": dk["algo_bytes"],
            "avg_launch_us": dk["avg_us"],
            "launches": dk["launches"],
            "timing": "HIP events on the context stream around every launch of this kernel "
                      "inside the timed region (no other kernel carries events there)",
        },
        "roofline_g2p2": _second_roofline(kernels.get(SECOND_KERNEL), sec_traffic),
        "cpu_baseline": cpu,
        "autotune": tuned,
    }
    if not a.no_full_run:
        out["full_run"] = full_run_leg(a, mpm)
    if not a.no_ensemble:
        out["batched"] = ensemble_leg(a, mpm)
    if not a.no_surrogate:
        out["surrogate"] = surrogate_leg(a, mpm)
    if not a.no_variants:
        out["variants"] = variants_leg(a, mpm)
    if not a.no_schedules:
        out["schedules"] = schedules_leg(a, mpm, value)
    if not a.no_slope:
        out["slope"] = slope_leg(a, mpm)
    print(json.dumps(out), flush=True)


def _second_roofline(k, traffic):
    """G2P #2 (the other kernel of the per-kernel gate) from the per-kernel table."""
    if not k or not k["gbs"]:
        return None
    return {"kernel": SECOND_KERNEL, "bound": "hbm", "achieved": k["gbs"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": k["gbs"] / HBM_PEAK_GBS, "traffic": traffic,
            "algorithmic_bytes": k["algo_bytes"], "avg_launch_us": k["avg_us"],
            "launches": k["launches"],
            "timing": "HIP events around every kernel of the steps after the timed region"}


def ensemble_leg(a, mpm):
    """Batched ensemble of independent 2D column collapses packed into one grid
    (one launch per kernel for all instances, graph-captured steps) against the
    same instances solved one after another, each on its own grid."""
    import numpy as np

    t = time.time()
    args, grid, mp, attr, bc, lay = mpm.scenes.ensemble_2d(n_instances=a.ens_instances)
    dt = float(args.ΔT)
    ninst = len(lay["ranges"])
    print(f"# ensemble: {ninst} instances, {mp.np} particles, {grid.ni} nodes "
          f"(setup {time.time() - t:.1f}s)", file=sys.stderr, flush=True)
    x0 = np.asarray(mp.ξ, dtype=np.float64)
    m = mpm.HIPModel(grid, mp, attr, bc, sort_interval=10)
    sa = m.step_args(args)
    res = {}
    Ti = m.run(sa, dt, 0.0, 20, graph=True)
    for graph in (False, True):
        m.sync()
        t0 = time.perf_counter()
        Ti = m.run(sa, dt, Ti, a.ens_steps, graph=graph)
        m.sync()
        el = time.perf_counter() - t0
        res["graph" if graph else "eager"] = ninst * a.ens_steps / el
    st = m.download()
    m.close()
    steps_done = 20 + 2 * a.ens_steps
    run = []
    for k, (lo, hi) in enumerate(lay["ranges"]):
        ox = lay["origin"][k][0]
        run.append(mpm.scenes.runout(st["xi"][lo:hi], x0[lo:hi]) - ox)
    run = np.array(run)
    seq = {}
    for graph in (False, True):
        tot = 0.0
        for k in range(ninst):
            sc = mpm.scenes.ensemble_instance(lay, k)
            s = mpm.HIPModel(sc[1], sc[2], sc[3], sc[4], sort_interval=10)
            ss = s.step_args(sc[0])
            T = s.run(ss, dt, 0.0, 10, graph=graph)
            s.sync()
            t0 = time.perf_counter()
            s.run(ss, dt, T, a.ens_seq_steps, graph=graph)
            s.sync()
            tot += time.perf_counter() - t0
            s.close()
        seq["graph" if graph else "eager"] = ninst * a.ens_seq_steps / tot
    best_seq = max(seq.values())
    out = {
        "metric": "instance-steps/s",
        "workload": f"{ninst} independent 2D DP column collapses (aspect U(0.5,4), 5k-60k particles, "
                    f"phi U(20,40) deg, cohesion lognormal median 1 kPa), {mp.np} particles, FP32",
        "batched_graph": res["graph"],
        "batched_eager": res["eager"],
        "sequential_eager": seq["eager"],
        "sequential_graph": seq["graph"],
        "speedup_vs_best_sequential": res["graph"] / best_seq,
        "particle_steps_per_s": res["graph"] * mp.np / ninst,
        "runout_after_steps": steps_done,
        "runout_m": {"min": float(run.min()), "median": float(np.median(run)),
                     "max": float(run.max())},
    }
    print(f"# ensemble: batched {res['graph']:.0f} (graph) / {res['eager']:.0f} (eager) "
          f"instance-steps/s; sequential {seq['eager']:.0f} / {seq['graph']:.0f}", file=sys.stderr,
          flush=True)
    return out


def _timed_run(a, mpm, **scene):
    """Warm-up + timed steps of the 3D column with the given scene options, no
    events; returns (particle-steps/s, ms/step)."""
    args, grid, mp, attr, bc = mpm.scenes.column_collapse_3d(h=a.h, **scene)
    dt = float(args.ΔT)
    m = mpm.HIPModel(grid, mp, attr, bc, sort_interval=a.sort_interval)
    try:
        sa = m.step_args(args)
        Ti = m.run(sa, dt, 0.0, a.warmup)
        m.sync()
        t0 = time.perf_counter()
        m.run(sa, dt, Ti, a.steps)
        m.sync()
        el = time.perf_counter() - t0
    finally:
        m.close()
    return int(mp.np) * a.steps / el, el / a.steps * 1e3


def variants_leg(a, mpm):
    """The timed workload in the FP64 parity mode and with volumetric locking
    (MVL, utils_OS.jl:1162-1256) on: the settings of the reference's own 3D
    collapse notebook run (FP64, MVL on)."""
    res = {}
    for name, kw in (("fp64", dict(eps="FP64")), ("fp64_mvl", dict(eps="FP64", MVL=True)),
                     ("fp32_mvl", dict(eps="FP32", MVL=True))):
        v, ms = _timed_run(a, mpm, **kw)
        res[name] = {"particle_steps_per_s": v, "ms_per_step": ms}
        print(f"# variant {name}: {ms:.4f} ms/step, {v:.3e} particle-steps/s", file=sys.stderr,
              flush=True)
    return res


def schedules_leg(a, mpm, musl_value):
    """The timed workload under the USL and USF schedules (src/solvers/OS_USL.jl,
    OS_USF.jl): same particles, sort interval, warm-up and step count as the
    MUSL line, no events."""
    import numpy as np

    res = {"MUSL": {"particle_steps_per_s": musl_value}}
    for scheme in ("USL", "USF"):
        args, grid, mp, attr, bc = mpm.scenes.column_collapse_3d(h=a.h, eps=a.precision,
                                                                  scheme=scheme)
        dt = float(args.ΔT)
        m = mpm.HIPModel(grid, mp, attr, bc, sort_interval=a.sort_interval)
        try:
            sa = m.step_args(args)
            Ti = m.run(sa, dt, 0.0, a.warmup)
            m.sync()
            t0 = time.perf_counter()
            m.run(sa, dt, Ti, a.steps)
            m.sync()
            el = time.perf_counter() - t0
            st = m.download()
        finally:
            m.close()
        n = int(mp.np)
        finite = bool(np.isfinite(st["xi"]).all() and np.isfinite(st["sigma"]).all())
        res[scheme] = {"particle_steps_per_s": n * a.steps / el, "ms_per_step": el / a.steps * 1e3,
                       "finite": finite}
        print(f"# schedule {scheme}: {res[scheme]['ms_per_step']:.4f} ms/step, "
              f"{res[scheme]['particle_steps_per_s']:.3e} particle-steps/s", file=sys.stderr,
              flush=True)
    return res


def surrogate_leg(a, mpm):
    """The learned 3D Mohr-Coulomb return against the exact return on the 3D
    column: samples logged from a plain run, network trained on the GPU, then
    both forms timed from the same state; fallback rate and runout difference."""
    import numpy as np

    sc = mpm.scenes.column_collapse_3d(h=a.h, eps="FP32", constitutive="mohrcoulomb",
                                       phi_deg=30.0, cohesion=0.0)
    args, grid, mp, attr, bc = sc
    dt = float(args.ΔT)
    x0 = np.asarray(mp.ξ, dtype=np.float64)
    n = int(mp.np)
    plain = mpm.HIPModel(grid, mp, attr, bc, sort_interval=a.sort_interval)
    sa = plain.step_args(args)
    # samples over the whole span the learned form runs (2 x sur_steps from rest)
    _, X, Y = plain.run_logged(sa, dt, 0.0, 2 * a.sur_steps, chunk=max(1, a.sur_steps // 4),
                               capacity=250_000, stride=13)
    if len(X) < 1000:
        plain.close()
        return {"skipped": f"only {len(X)} return-mapping samples in {a.sur_steps} steps"}
    S = mpm.Surrogate(batch_capacity=4096, seed=0)
    S.fit_normalisation(X, Y)
    t1 = time.perf_counter()
    hist = S.train(X, Y, steps=2000, batch=4096, lr=2e-3, seed=1)
    t_train = time.perf_counter() - t1
    print(f"# surrogate: {len(X)} samples, trained 2000 steps in {t_train:.1f}s, loss "
          f"{hist[:20].mean():.4f} -> {hist[-20:].mean():.5f}", file=sys.stderr, flush=True)
    # timed: each form runs 2 x sur_steps from the initial state; the second half is timed
    plain.close()
    times, results = {}, {}
    for mode in ("exact", "learned"):
        m = mpm.HIPModel(grid, mp, attr, bc, sort_interval=a.sort_interval)
        if mode == "learned":
            m.set_surrogate(S, tol=1e-3)
        T = m.run(sa, dt, 0.0, a.sur_steps)
        m.sync()
        m.profile_enable(True)
        t0 = time.perf_counter()
        m.run(sa, dt, T, a.sur_steps)
        m.sync()
        times[mode] = time.perf_counter() - t0
        prof = m.profile_read()
        m.profile_enable(False)
        st = m.download()
        results[mode] = (st, prof, m.surrogate_stats() if mode == "learned" else (0, 0))
        m.close()
    S.close()
    st_e, st_l = results["exact"][0], results["learned"][0]
    r_e = mpm.scenes.runout(st_e["xi"], x0)
    r_l = mpm.scenes.runout(st_l["xi"], x0)
    plastic, fb = results["learned"][2]
    ms_k, cnt_k = results["learned"][1].get("k_learned_return", (0.0, 0))
    flops = mpm.surrogate.Surrogate.FLOPS_PER_SAMPLE * n
    tflops = flops / (ms_k * 1e-3 / cnt_k) / 1e12 if cnt_k else None
    out = {
        "workload": f"3D Mohr-Coulomb column (phi 30 deg, c 0), {n} particles, FP32, "
                    f"{a.sur_steps} timed steps after {a.sur_steps}",
        "exact_particle_steps_per_s": n * a.sur_steps / times["exact"],
        "learned_particle_steps_per_s": n * a.sur_steps / times["learned"],
        "speed_ratio_learned_over_exact": times["exact"] / times["learned"],
        "yielding_particle_steps": plastic,
        "fallback_rate": fb / plastic if plastic else None,
        "residual_tol": 1e-3,
        "runout_rel_diff": r_l / r_e - 1.0,
        "displacement_rel_l2": float(np.linalg.norm(st_l["xi"] - st_e["xi"]) /
                                     np.linalg.norm(st_e["xi"] - x0)),
        "samples": int(len(X)),
        "train_seconds": t_train,
        "roofline": {"kernel": "k_learned_return", "bound": "mfma",
                     "achieved": tflops, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": tflops / 2500.0 if tflops else None,
                     "flops_per_launch": flops},
    }
    print(f"# surrogate: exact {out['exact_particle_steps_per_s']:.3e}, learned "
          f"{out['learned_particle_steps_per_s']:.3e} particle-steps/s, fallback "
          f"{out['fallback_rate']}, runout diff {out['runout_rel_diff']:+.2e}, MLP kernel "
          f"{tflops} TFLOP/s", file=sys.stderr, flush=True)
    return out


def slope_leg(a, mpm):
    """The seeded GRF slope at its configured size (20M particles at h = 0.5 m,
    one Mohr-Coulomb material per initial cell): a damped K0 phase, then the
    undamped runout timed from that K0 state with the exact 3D Mohr-Coulomb
    return and with the learned return (trained on return samples logged on the
    slope itself), each in a fresh context resumed from the K0 state, with
    captured step graphs."""
    import numpy as np

    t_setup = time.perf_counter()
    args, grid, mp, attr, bc, lay = mpm.scenes.slope_3d(h=a.slope_h, seed=1, k0_steps=a.slope_k0)
    dt = lay["dt"]
    n = int(mp.np)
    K = a.full_sort_interval
    print(f"# slope: {n} particles, {lay['nmat']} materials, {int(grid.ni)} nodes "
          f"(setup {time.perf_counter() - t_setup:.1f}s)", file=sys.stderr, flush=True)
    m = mpm.HIPModel(grid, mp, attr, bc, sort_interval=K)
    try:
        t0 = time.perf_counter()
        Ti = m.run(m.step_args(args), dt, 0.0, lay["k0_steps"], graph=True)
        m.sync()
        t_k0 = time.perf_counter() - t0
        k0 = m.download()
        ra = m.step_args(lay["runout_args"])
        # samples over the whole span the timed forms run (20 + slope_steps)
        _, X, Y = m.run_logged(ra, dt, Ti, 20 + a.slope_steps, chunk=max(1, a.slope_steps // 8),
                               capacity=1_000_000, stride=29, graph=True)
    finally:
        m.close()
    print(f"# slope: K0 {lay['k0_steps']} steps in {t_k0:.2f}s, {len(X)} return samples",
          file=sys.stderr, flush=True)
    snap = dict(coords=k0["xi"], velocity_s=k0["vs"], stress=k0["sigma"], strain_s=k0["eps_s"],
                eqstrain=k0["eps_q"], ekstrain=k0["eps_k"], volume=k0["omega"], mass_s=k0["ms"],
                time=np.asarray(Ti))
    S = None
    if len(X) >= 1000:
        S = mpm.Surrogate(batch_capacity=4096, seed=0)
        S.fit_normalisation(X, Y)
        hist = S.train(X, Y, steps=2000, batch=4096, lr=2e-3, seed=1)
        print(f"# slope: trained on {len(X)} samples, loss {hist[:20].mean():.4f} -> "
              f"{hist[-20:].mean():.5f}", file=sys.stderr, flush=True)
    res = {}
    for mode in ("exact", "learned"):
        if mode == "learned" and S is None:
            continue
        m = mpm.HIPModel(grid, mp, attr, bc, sort_interval=K)
        try:
            T = m.upload_snapshot(snap, mp.Ω0)
            if mode == "learned":
                m.set_surrogate(S, tol=1e-3)
            T = m.run(ra, dt, T, 20, graph=True)
            m.sync()
            t0 = time.perf_counter()
            m.run(ra, dt, T, a.slope_steps, graph=True)
            m.sync()
            t = time.perf_counter() - t0
            res[mode] = (t, m.download()["xi"], m.surrogate_stats() if mode == "learned" else (0, 0))
        finally:
            m.close()
        print(f"# slope: {mode} {a.slope_steps} runout steps in {t:.2f}s, "
              f"{n * a.slope_steps / t:.3e} particle-steps/s", file=sys.stderr, flush=True)
    if S is not None:
        S.close()
    out = {"workload": f"GRF slope, {n} particles (h = {a.slope_h} m, 8 per cell), "
                       f"{lay['nmat']} Mohr-Coulomb materials, {int(grid.ni)} grid nodes, FP32, "
                       f"K0 {lay['k0_steps']} steps, then {a.slope_steps} timed runout steps "
                       f"after 20, sort every {K}, captured graphs",
           "particles": n,
           "k0_seconds": t_k0,
           "exact_particle_steps_per_s": n * a.slope_steps / res["exact"][0]}
    if "learned" in res:
        plastic, fb = res["learned"][2]
        xe, xl = res["exact"][1], res["learned"][1]
        out.update({"learned_particle_steps_per_s": n * a.slope_steps / res["learned"][0],
                    "speed_ratio_learned_over_exact": res["exact"][0] / res["learned"][0],
                    "fallback_rate": fb / plastic if plastic else None,
                    "displacement_rel_l2": float(np.linalg.norm(xl - xe) /
                                                 max(np.linalg.norm(xe - k0["xi"]), 1e-300)),
                    "samples": int(len(X))})
    return out


def cpu_baseline(a, args, grid, mp, attr, bc):
    """The FP64 oracle (CPU restatement of the reference step) on the same scene,
    on every core this process may use and on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    threads = a.cpu_threads or allowed
    n = int(mp.np)
    dt = float(args.ΔT)
    res = {}
    for nt, steps in ((threads, a.cpu_steps), (1, a.cpu_steps_1t)):
        o = oracle.Oracle(args, grid, mp, attr, bc, nthreads=nt)
        o.run(dt, 0.0, 1)  # first touch of every array
        t0 = time.perf_counter()
        o.run(dt, dt, steps)
        res[nt] = (steps, time.perf_counter() - t0)
        del o
        print(f"# cpu baseline: {nt} thread(s), {steps} steps in {res[nt][1]:.1f} s",
              file=sys.stderr, flush=True)
    steps, el = res[threads]
    s1, e1 = res[1]
    return {"value": n * steps / el, "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"same 3D column ({n} particles), {steps} FP64 steps after one untimed "
                      f"step, OpenMP {threads} threads (every core this process may run on), "
                      f"{el:.1f} s",
            "single_thread": {"value": n * s1 / e1, "cores": 1,
                              "sample": f"{s1} FP64 steps after one untimed step, {e1:.1f} s"}}


if __name__ == "__main__":
    main()
