"""Full-run checks on the GPU.

* The plane-strain column collapse at full resolution (12,800 particles,
  h = 2.5 mm, dt = 2e-5 s, 50,000 steps = 1 s) in FP64 against the
  experimental free surface the reference ships, and FP32 (FP64 positions and
  det F) against FP64: runout within 1 %, kinetic-energy history and deposit
  profile close.
* The 3D column at full size (2,048,000 particles): size-independent
  properties (mass conserved exactly, every particle finite and inside the
  grid, FP32 vs FP64 agreement over a short horizon).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def runout(x, q=0.999):
    """Robust runout: the q-quantile of the final particle x positions."""
    return float(np.quantile(x[:, 0], q))


def _run_2d(mpm, eps, chunk=1000, total=50_000):
    args, grid, mp, attr, bc = mpm.scenes.column_collapse_2d(eps=eps)
    print(f"case: collapse2d eps={eps} np={mp.np} steps={total}", flush=True)
    model = mpm.HIPModel(grid, mp, attr, bc, sort_interval=0)
    ke = []
    try:
        sargs = model.step_args(args)
        dt = float(args.ΔT)
        Ti = 0.0
        for _ in range(total // chunk):
            Ti = model.run(sargs, dt, Ti, chunk)
            s = model.download()
            ke.append(mpm.scenes.kinetic_energy(s["ms"], s["vs"]))
        model.sync()
    finally:
        model.close()
    return s, np.array(ke)


def test_collapse2d_full_fp64_vs_experiment_and_fp32(mpm):
    s64, ke64 = _run_2d(mpm, "FP64")
    s32, ke32 = _run_2d(mpm, "FP32")
    x64, x32 = s64["xi"], s32["xi"]
    surf = np.loadtxt(os.path.join(GOLDEN, "collapse2d_surface.csv"), delimiter=",")
    err = []
    for xe, ye in surf:
        if xe > 0.40:
            continue
        sel = np.abs(x64[:, 0] - xe) < 0.0025
        err.append(x64[sel, 1].max() - ye)
    err = np.array(err)
    print("surface error rms", np.sqrt(np.mean(err ** 2)), "max", np.abs(err).max())
    assert np.sqrt(np.mean(err ** 2)) < 0.004 and np.abs(err).max() < 0.008
    r64, r32 = runout(x64), runout(x32)
    print("runout fp64", r64, "fp32", r32, "rel", abs(r32 - r64) / r64)
    # the FP32 run's atomic summation order differs from run to run, and this
    # 50,000-step granular flow amplifies it: two runs of one build measured
    # 0.22 % and 0.68 % for this 99.9 % quantile of x (a few particle spacings)
    assert abs(r32 - r64) / r64 < 0.01
    # kinetic-energy history (sampled every 1000 steps)
    assert np.max(np.abs(ke32 - ke64)) < 0.05 * np.max(ke64)
    # deposit profiles: max height in 5 mm bins
    bins = np.arange(0.0, 0.5, 0.005)
    h64 = np.array([x64[(x64[:, 0] >= b) & (x64[:, 0] < b + 0.005), 1].max(initial=0) for b in bins])
    h32 = np.array([x32[(x32[:, 0] >= b) & (x32[:, 0] < b + 0.005), 1].max(initial=0) for b in bins])
    print("deposit profile max diff", np.abs(h64 - h32).max())
    assert np.abs(h64 - h32).max() < 0.005


def test_column3d_full_size_invariants(mpm):
    out = {}
    for eps in ("FP64", "FP32"):
        args, grid, mp, attr, bc = mpm.scenes.column_collapse_3d(eps=eps)
        print(f"case: column3d eps={eps} np={mp.np}", flush=True)
        model = mpm.HIPModel(grid, mp, attr, bc, sort_interval=10)
        try:
            sargs = model.step_args(args)
            model.run(sargs, float(args.ΔT), 0.0, 50)
            model.sync()
            out[eps] = model.download()
        finally:
            model.close()
        s = out[eps]
        m0 = float(np.sum(np.asarray(mp.Ω, dtype=np.float64) * np.asarray(mp.ρs, dtype=np.float64)))
        assert abs(s["ms"].sum() - m0) < 1e-6 * m0
        assert np.all(np.isfinite(s["xi"])) and np.all(np.isfinite(s["sigma"]))
        assert s["xi"][:, 2].min() > -0.01
    x0 = np.asarray(mp.ξ, dtype=np.float64)
    d64, d32 = out["FP64"]["xi"] - x0, out["FP32"]["xi"] - x0
    rel = np.linalg.norm(d32 - d64) / np.linalg.norm(d64)
    print("3D FP32 vs FP64 displacement rel L2 after 50 steps:", rel)
    assert rel < 1e-3
    rs = np.linalg.norm(out["FP32"]["sigma"] - out["FP64"]["sigma"]) / np.linalg.norm(out["FP64"]["sigma"])
    print("stress rel L2:", rs)
    assert rs < 2e-3
