"""Graph-captured stepping and the packed ensemble, through the C ABI.

* mpm_run_graph replays captured step graphs between sorts; its results must
  equal mpm_run's (FP64 mode, up to atomic summation order: relative L2
  <= 1e-9), checked after every chunk of a run cut into uneven chunks, so each
  captured graph is replayed several times with eager sorts and downloads
  (synchronisations) in between.
* ensemble_2d packs independent instances into one grid; each packed instance
  must evolve like the same instance alone on its own grid (FP64, relative L2
  of the displacement <= 1e-8: positions differ by the packing offset, so
  rounding differs at 1e-16 and grows only mildly over 150 steps).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))


def _graph_vs_eager(mpm, scene, chunks, sort_interval, Te=None):
    args, grid, mp, attr, bc = scene
    x0 = np.asarray(mp.ξ, dtype=np.float64)
    dt = float(args.ΔT)
    eager = mpm.HIPModel(grid, mp, attr, bc, sort_interval=sort_interval)
    graph = mpm.HIPModel(grid, mp, attr, bc, sort_interval=sort_interval)
    try:
        sa = eager.step_args(args)
        Te_, Tg = 0.0, 0.0
        for c in chunks:
            Te_ = eager.run(sa, dt, Te_, c)
            Tg = graph.run(sa, dt, Tg, c, graph=True)
            assert Tg == Te_
            a, b = eager.download(), graph.download()
            for k in ("xi", "vs", "sigma", "eps_q"):
                ref = b[k] - x0 if k == "xi" else b[k]
                got = a[k] - x0 if k == "xi" else a[k]
                r = _rel(ref, got)
                assert r <= 1e-9, (k, c, r)
            print(f"  chunk {c}: xi rel L2 {_rel(b['xi'] - x0, a['xi'] - x0):.1e}", flush=True)
    finally:
        eager.close()
        graph.close()


def test_graph_run_matches_eager_2d(mpm):
    sc = mpm.scenes.column_collapse_2d(h=0.01, width=0.2, height=0.1, box_x=0.5, box_y=0.12,
                                       dt=5e-5, nsteps=100, eps="FP64")
    _graph_vs_eager(mpm, sc, [7, 9, 12, 3, 20], sort_interval=5)


def test_graph_run_matches_eager_3d_with_ramp(mpm):
    sc = mpm.scenes.column_collapse_3d(h=0.02, width=0.16, height=0.08, box_xy=0.4, box_z=0.12,
                                       nsteps=100, eps="FP64", Te=0.002)
    # the first ~10 steps are in the gravity ramp (eager), the rest replay graphs
    _graph_vs_eager(mpm, sc, [13, 8, 25], sort_interval=6)


def test_graph_run_unsorted_and_mc3d(mpm):
    sc = mpm.scenes.column_collapse_3d(h=0.02, width=0.16, height=0.08, box_xy=0.4, box_z=0.12,
                                       nsteps=100, eps="FP64", constitutive="mohrcoulomb",
                                       cohesion=200.0)
    _graph_vs_eager(mpm, sc, [10, 10, 5], sort_interval=0)


@pytest.mark.parametrize("scheme", ["USL", "USF"])
def test_graph_run_matches_eager_schemes(mpm, scheme):
    # USF graphs hold body, then head + body; each run of steps starts with an
    # eager head before the sort
    sc = mpm.scenes.column_collapse_3d(h=0.02, width=0.16, height=0.08, box_xy=0.4, box_z=0.12,
                                       nsteps=100, eps="FP64", scheme=scheme, MVL=True)
    _graph_vs_eager(mpm, sc, [1, 9, 12, 3, 20], sort_interval=5)


def test_packed_instances_evolve_independently(mpm):
    kw = dict(n_instances=5, seed=11, np_range=(400, 2500), row_width=1.0, eps="FP64")
    args, grid, mp, attr, bc, lay = mpm.scenes.ensemble_2d(**kw)
    nsteps = 150
    dt = float(args.ΔT)
    x0 = np.asarray(mp.ξ, dtype=np.float64)
    m = mpm.HIPModel(grid, mp, attr, bc, sort_interval=10)
    try:
        m.run(m.step_args(args), dt, 0.0, nsteps, graph=True)
        packed = m.download()
    finally:
        m.close()
    for k, (a, b) in enumerate(lay["ranges"]):
        sa, sg, smp, sattr, sbc = mpm.scenes.ensemble_instance(lay, k, eps="FP64")
        assert smp.np == b - a
        s = mpm.HIPModel(sg, smp, sattr, sbc, sort_interval=10)
        try:
            s.run(s.step_args(sa), dt, 0.0, nsteps)
            alone = s.download()
        finally:
            s.close()
        ox, oy = lay["origin"][k]
        d_packed = packed["xi"][a:b] - x0[a:b]
        d_alone = alone["xi"] - np.asarray(smp.ξ, dtype=np.float64)
        r = _rel(d_packed, d_alone)
        rs = _rel(packed["sigma"][a:b], alone["sigma"])
        print(f"instance {k}: {b - a} particles, displacement rel L2 {r:.1e}, stress {rs:.1e}",
              flush=True)
        assert np.abs(d_alone).max() > 1e-5  # it moved
        assert r <= 1e-8 and rs <= 1e-8


def test_slope_k0_then_runout(mpm):
    # the GRF slope at 4 m cells: damped K0 phase with the gravity ramp, then the
    # undamped runout with graph-captured steps; mass moves downslope
    args, grid, mp, attr, bc, lay = mpm.scenes.slope_3d(h=4.0, seed=3, k0_steps=200)
    x0 = np.asarray(mp.ξ, dtype=np.float64)
    dt = lay["dt"]
    m = mpm.HIPModel(grid, mp, attr, bc, sort_interval=10)
    try:
        Ti = m.run(m.step_args(args), dt, 0.0, lay["k0_steps"], graph=True)
        k0 = m.download()
        Ti = m.run(m.step_args(lay["runout_args"]), dt, Ti, 400, graph=True)
        m.sync()
        out = m.download()
    finally:
        m.close()
    print(f"{mp.np} particles, {lay['nmat']} materials; K0 mean |dz| "
          f"{np.abs(k0['xi'][:, 2] - x0[:, 2]).mean():.3e} m; runout mean dx "
          f"{(out['xi'][:, 0] - x0[:, 0]).mean():.3e} m, max {(out['xi'][:, 0] - x0[:, 0]).max():.2f} m",
          flush=True)
    assert np.all(np.isfinite(out["xi"])) and np.all(np.isfinite(out["sigma"]))
    assert (out["xi"][:, 0] - x0[:, 0]).mean() > 0.01  # slides downslope
    # compressive mean stress at depth after K0
    sm = out["sigma_m"]
    assert np.median(sm) < 0


def test_autotune_on_the_device(mpm):
    args, grid, mp, attr, bc = mpm.scenes.column_collapse_3d(
        h=0.02, width=0.16, height=0.08, box_xy=0.4, box_z=0.12, nsteps=100, eps="FP32")
    m = mpm.HIPModel(grid, mp, attr, bc, sort_interval=10)
    try:
        sa = m.step_args(args)
        best, table, Ti = mpm.autotune(m, sa, float(args.ΔT), 0.0, sort_intervals=(5, 20),
                                       steps=5, warm=2, scatters=("owned", "owned_layers", "rounds"))
        assert len(table) == 8 and best["sort_interval"] in (5, 20)
        print("autotune:", best, [(k, t, sc, f"{s * 1e6:.0f} us") for k, t, sc, s in table],
              flush=True)
        m.run(sa, float(args.ΔT), Ti, 5, graph=True)
        st = m.download()
        assert np.all(np.isfinite(st["xi"]))
    finally:
        m.close()


@pytest.mark.parametrize("eps", ["FP64", "FP32"])
def test_resume_from_snapshot(mpm, tmp_path, eps):
    # run 40 steps, write a snapshot group, run 40 more; a new context resumed
    # from the group and run 40 steps lands on the same state (FP64: to atomic
    # summation order; FP32: at the tolerance of its arithmetic)
    args, grid, mp, attr, bc = mpm.scenes.column_collapse_3d(
        h=0.02, width=0.16, height=0.08, box_xy=0.4, box_z=0.12, nsteps=100, eps=eps,
        project_name="resume", project_path=str(tmp_path))
    dt = float(args.ΔT)
    a = mpm.HIPModel(grid, mp, attr, bc, sort_interval=5)
    try:
        sa = a.step_args(args)
        Ti = a.run(sa, dt, 0.0, 40)
        snap = a.download()
        w = mpm.solver.SnapshotWriter(args)
        host = type("MP", (), {})()
        host.σij, host.ϵijs, host.ϵq, host.ϵk = snap["sigma"], snap["eps_s"], snap["eps_q"], snap["eps_k"]
        host.ϵv, host.ξ, host.vs, host.Ω, host.ms = snap["eps_v"], snap["xi"], snap["vs"], snap["omega"], snap["ms"]
        w.group(1, host, Ti)
        a.run(sa, dt, Ti, 40)
        ref = a.download()
    finally:
        a.close()
    b = mpm.HIPModel(grid, mp, attr, bc, sort_interval=5)
    try:
        T0 = b.upload_snapshot(str(tmp_path / "resume" / "resume_snapshots" / "group1.npz"), mp.Ω0)
        assert T0 == Ti
        b.run(b.step_args(args), dt, T0, 40)
        got = b.download()
    finally:
        b.close()
    tol = 1e-9 if eps == "FP64" else 2e-4
    x0 = np.asarray(mp.ξ, dtype=np.float64)
    for k in ("xi", "vs", "sigma", "eps_q", "omega"):
        r = _rel(got[k] - (x0 if k == "xi" else 0), ref[k] - (x0 if k == "xi" else 0))
        print(f"  {k}: rel L2 {r:.1e}", flush=True)
        assert r <= tol, k


def test_profile_mask_times_only_the_named_kernels(mpm):
    # MPM_OPT_PROFILE_MASK: events around the masked kernels only; 0 = all
    sc = mpm.scenes.column_collapse_3d(h=0.02, width=0.16, height=0.08, box_xy=0.4, box_z=0.12,
                                       nsteps=20)
    args, grid, mp, attr, bc = sc
    dt = float(args.ΔT)
    m = mpm.HIPModel(grid, mp, attr, bc, sort_interval=4)
    try:
        sa = m.step_args(args)
        m.profile_enable(True)
        T = m.run(sa, dt, 0.0, 8)
        every = m.profile_read()
        m.profile_enable(False)
        assert every["k_p2g"][1] == 8 and every["k_grid1"][1] == 8 and every["k_permute"][1] == 2
        assert every["k_sort"][1] > 0
        m.set_option("profile_kernels", ["k_p2g", "k_sort"])
        m.profile_enable(True)
        T = m.run(sa, dt, T, 8)
        some = m.profile_read()
        m.profile_enable(False)
        assert some["k_p2g"][1] == 8 and some["k_sort"][1] == every["k_sort"][1]
        assert some["k_p2g"][0] > 0.0
        assert all(c == 0 for k, (_, c) in some.items() if k not in ("k_p2g", "k_sort"))
        with pytest.raises(ValueError):
            m.set_option("profile_kernels", ["no_such_kernel"])
        m.set_option("profile_kernels", None)
        m.profile_enable(True)
        m.run(sa, dt, T, 4)
        again = m.profile_read()
        m.profile_enable(False)
        assert again["k_grid2"][1] == 4
    finally:
        m.close()
