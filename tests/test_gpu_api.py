"""The reference's API entry on the GPU, against the oracle.

* materialpointsolver(args, grid, mp, attr, bc) (src/solver.jl:37-131 through
  this package's submit_work: upload, the ``while Ti < Ttol`` loop with
  ``Ti += ΔT``, download into the caller's particle struct, free) on a short
  horizon of the 2D column collapse, FP64, against the oracle run of the same
  number of steps: relative L2 <= 1e-9 on the fields the download returns.
* Snapshot resume: a context resumed from a snapshot group written after 40
  steps and run 40 more steps against the oracle's uninterrupted 80-step run.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))


@pytest.mark.parametrize("sort_interval", [0, 10])
def test_materialpointsolver_entry_vs_oracle(mpm, orc, tmp_path, sort_interval):
    dt, nsteps = 2e-5, 300
    sc = mpm.scenes.column_collapse_2d(h=0.005, dt=dt, nsteps=nsteps, eps="FP64",
                                       project_name="entry", project_path=str(tmp_path),
                                       sort_interval=sort_interval)
    args, grid, mp, attr, bc = sc
    x0 = np.array(mp.ξ, dtype=np.float64, copy=True)
    ref = orc.Oracle(args, grid, mp, attr, bc, nthreads=8)
    print(f"case: materialpointsolver 2D np={mp.np} Ttol={args.Ttol} sort={sort_interval}",
          flush=True)
    mpm.materialpointsolver(args, grid, mp, attr, bc)  # fills mp in place (device2host!)
    n = args.iter_num
    Ti, k = 0.0, 0  # the reference's clock: Ti += ΔT in T2 while Ti < Ttol
    while Ti < args.Ttol:
        Ti, k = Ti + dt, k + 1
    assert n == k
    ref.run(dt, 0.0, n)
    want = ref.state()
    got = dict(xi=np.asarray(mp.ξ) - x0, vs=np.asarray(mp.vs), sigma=np.asarray(mp.σij),
               eps_q=np.asarray(mp.ϵq), omega=np.asarray(mp.Ω))
    want = dict(xi=want["xi"] - x0, vs=want["vs"], sigma=want["sigma"], eps_q=want["eps_q"],
                omega=want["omega"])
    for k in got:
        r = _rel(got[k], want[k])
        print(f"  {k}: rel L2 {r:.1e}", flush=True)
        assert r <= 1e-9, k


def test_snapshot_resume_vs_oracle(mpm, orc, tmp_path):
    args, grid, mp, attr, bc = mpm.scenes.column_collapse_3d(
        h=0.02, width=0.16, height=0.08, box_xy=0.4, box_z=0.12, nsteps=100, eps="FP64",
        project_name="resume", project_path=str(tmp_path))
    dt = float(args.ΔT)
    print(f"case: resume 3D np={mp.np}", flush=True)
    a = mpm.HIPModel(grid, mp, attr, bc, sort_interval=5)
    try:
        Ti = a.run(a.step_args(args), dt, 0.0, 40)
        snap = a.download()
    finally:
        a.close()
    w = mpm.solver.SnapshotWriter(args)
    host = type("MP", (), {})()
    host.σij, host.ϵijs, host.ϵq, host.ϵk = snap["sigma"], snap["eps_s"], snap["eps_q"], snap["eps_k"]
    host.ϵv, host.ξ, host.vs, host.Ω, host.ms = snap["eps_v"], snap["xi"], snap["vs"], snap["omega"], snap["ms"]
    w.group(1, host, Ti)
    b = mpm.HIPModel(grid, mp, attr, bc, sort_interval=5)
    try:
        T0 = b.upload_snapshot(str(tmp_path / "resume" / "resume_snapshots" / "group1.npz"), mp.Ω0)
        assert T0 == Ti
        b.run(b.step_args(args), dt, T0, 40)
        got = b.download()
    finally:
        b.close()
    ref = orc.Oracle(args, grid, mp, attr, bc, nthreads=8)
    ref.run(dt, 0.0, 80)
    want = ref.state()
    x0 = np.asarray(mp.ξ, dtype=np.float64)
    for k in ("xi", "vs", "sigma", "eps_q", "omega"):
        r = _rel(got[k] - (x0 if k == "xi" else 0), want[k] - (x0 if k == "xi" else 0))
        print(f"  {k}: rel L2 {r:.1e}", flush=True)
        assert r <= 1e-9, k
