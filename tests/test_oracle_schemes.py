"""Oracle pins for the USL and USF schedules (src/solvers/OS_USL.jl,
src/solvers/OS_USF.jl) next to MUSL (src/solvers/OS_MUSL.jl).

* USL/USF grid solves are the MUSL solves plus the nodal displacement
  dus = v_new dt (OS_USL.jl:63-100, 225-269).
* USL :a gathers the NEW nodal velocity for the PIC term
  (doublemapping1_a_USL_OS!, OS_USL.jl:271-313), so with PIC = 1 a free
  particle under gravity gains 2 g dt in one step, against g dt under MUSL :a
  and USL :v (closed forms).
* The USL strain increment comes from the grid velocity of the same step: an
  affine velocity field on a symmetric lattice gives dF = A dt exactly.
* USF is USL with the stress update moved to the start of the next step
  (OS_USF.jl:58-97): after n steps its positions and velocities equal USL's
  after n steps and its stresses and strains equal USL's after n - 1 steps,
  bit for bit in the serial oracle.
"""
import numpy as np
import pytest


def _free_block(mpm, dim, scheme, va="a", n=4, h=0.1, pic=0.0, gravity=-9.8, jitter=True,
                constitutive="linearelastic"):
    """A stress-free block far from every wall (no boundary node in reach)."""
    kw = dict(h=h, width=n * h, height=n * h, pad=2, nsteps=1, constitutive=constitutive,
              va=va, scheme=scheme, gravity=gravity)
    if dim == 2:
        sc = mpm.scenes.column_collapse_2d(box_x=12 * h, box_y=12 * h, dt=1e-4, **kw)
    else:
        sc = mpm.scenes.column_collapse_3d(box_xy=12 * h, box_z=12 * h, eps="FP64", dt=1e-4, **kw)
    args, grid, mp, attr, bc = sc
    shift = np.full(dim, 4.0 * h)
    if dim == 3:
        shift[:2] = 0.0
    rng = np.random.default_rng(0)
    mp.ξ[...] = mp.ξ + shift
    if jitter:
        mp.ξ[...] = mp.ξ + rng.uniform(-0.2 * h, 0.2 * h, size=mp.ξ.shape)
    args.FLIP, args.PIC = 1.0 - pic, pic
    return args, grid, mp, attr, bc


@pytest.mark.parametrize("va", [0, 1])
def test_usl_grid_solve_is_musl_solve_plus_displacement(mpm, orc, va):
    sc = _free_block(mpm, 3, "USL")
    o = orc.Oracle(*sc)
    a, m = o.a, o.m
    rng = np.random.default_rng(4)
    ni = a["g_ms"].shape[0]
    ms = rng.uniform(0.5, 2.0, ni)
    ms[rng.random(ni) < 0.2] = 0.0
    ps = rng.normal(size=a["g_ps"].shape)
    fs = rng.normal(size=a["g_fs"].shape)
    m.va, m.zeta_s = va, 0.1
    dt = 1e-3
    out = {}
    for scheme in (0, 1):
        m.scheme = scheme
        a["g_ms"][...] = ms
        a["g_ps"][...] = ps
        a["g_fs"][...] = fs
        a["g_dus"][...] = 7.0
        o.solvegrid(dt)
        out[scheme] = {k: a[k].copy() for k in ("g_vs", "g_vsT", "g_as", "g_dus", "g_ps")}
    mu, us = out[0], out[1]
    np.testing.assert_array_equal(us["g_vsT"], mu["g_vsT"])
    np.testing.assert_array_equal(us["g_vs"], mu["g_vs"])
    if va == 0:
        np.testing.assert_array_equal(us["g_as"], mu["g_as"])
    np.testing.assert_array_equal(us["g_dus"], us["g_vsT"] * dt)
    assert np.all(mu["g_dus"] == 7.0)          # MUSL leaves dus to the re-map
    assert np.all(mu["g_ps"] == 0.0)           # MUSL clears p for the re-map scatter
    # USL's p keeps P2G's momentum (with the :a form's boundary values on fixed axes)
    assert np.abs(us["g_ps"]).max() > 0.1


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("scheme,va,gain", [("MUSL", "a", 1), ("USL", "a", 2), ("USF", "a", 2),
                                            ("MUSL", "v", 1), ("USL", "v", 1), ("USF", "v", 1)])
def test_pic_velocity_closed_form(mpm, orc, dim, scheme, va, gain):
    # PIC = 1 from a uniform velocity under gravity only: MUSL and the :v forms
    # give v0 + g dt; USL/USF :a add the acceleration on top of the already
    # accelerated nodal velocity, v0 + 2 g dt.  Every form moves x by
    # dt (v0 + g dt).
    g = -9.8
    sc = _free_block(mpm, dim, scheme, va=va, pic=1.0, gravity=g)
    args = sc[0]
    o = orc.Oracle(*sc)
    v0 = np.array([0.3, -0.2, 0.1][:dim])
    o.a["vs"][...] = v0
    x0 = o.a["xi"].copy()
    dt = float(args.ΔT)
    o.procedure(dt, 0.0)
    dv = np.zeros(dim)
    dv[-1] = g * dt
    v, x = o.a["vs"], o.a["xi"]
    np.testing.assert_allclose(v, np.broadcast_to(v0 + gain * dv, v.shape), rtol=0, atol=1e-14)
    np.testing.assert_allclose(x - x0, np.broadcast_to(dt * (v0 + dv), x.shape), rtol=0,
                               atol=1e-16)


@pytest.mark.parametrize("dim", [2, 3])
def test_usl_affine_velocity_gives_exact_strain_increment(mpm, orc, dim):
    # v = b + A x on a regular quarter-point lattice (no jitter): a node whose
    # whole support lies in the block receives p = m (b + A x_node) by symmetry,
    # so every particle whose stencil holds only such nodes gets dF = A dt
    h = 0.1
    sc = _free_block(mpm, dim, "USL", n=8, h=h, jitter=False, gravity=-9.8)
    args, grid, mp, attr, bc = sc
    o = orc.Oracle(*sc)
    a = o.a
    rng = np.random.default_rng(2)
    A = rng.normal(scale=0.5, size=(dim, dim))
    b = rng.normal(scale=0.5, size=dim)
    x0 = a["xi"].copy()
    a["vs"][...] = x0 @ A.T + b
    dt = float(args.ΔT)
    o.procedure(dt, 0.0)
    lo, hi = x0.min(axis=0), x0.max(axis=0)
    inner = np.all((x0 > lo + 2.6 * h) & (x0 < hi - 2.6 * h), axis=1)
    assert inner.sum() >= (8 if dim == 3 else 4)
    dF = a["dF"].reshape(-1, dim, dim)[inner]
    np.testing.assert_allclose(dF, np.broadcast_to(A * dt, dF.shape), rtol=0, atol=1e-15)
    # positions move with the accelerated grid velocity: dt (b + A x + g dt e_z)
    g = np.zeros(dim)
    g[-1] = -9.8 * dt
    want = dt * (x0[inner] @ A.T + b + g)
    np.testing.assert_allclose(a["xi"][inner] - x0[inner], want, rtol=0, atol=1e-15)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("va", ["a", "v"])
@pytest.mark.parametrize("constitutive", ["linearelastic", "druckerprager"])
def test_usf_is_usl_with_the_stress_update_deferred(mpm, orc, dim, va, constitutive):
    if dim == 2:
        mk = lambda s: mpm.scenes.column_collapse_2d(  # noqa: E731
            h=0.02, width=0.2, height=0.1, box_x=0.5, box_y=0.14, dt=1e-4, nsteps=30,
            constitutive=constitutive, va=va, scheme=s)
    else:
        mk = lambda s: mpm.scenes.column_collapse_3d(  # noqa: E731
            h=0.04, width=0.16, height=0.08, box_xy=0.4, box_z=0.16, eps="FP64", nsteps=30,
            constitutive=constitutive, va=va, scheme=s)
    usl, usf = orc.Oracle(*mk("USL")), orc.Oracle(*mk("USF"))
    dt = float(mk("USL")[0].ΔT)
    s0 = usf.a["sig"].copy()
    usf.procedure(dt, 0.0)
    np.testing.assert_array_equal(usf.a["sig"], s0)    # first USF stress update: zero increment
    prev = usl.state()
    usl.procedure(dt, 0.0)
    Ti = dt
    for k in range(2, 13):
        prev = usl.state()
        Fprev = usl.a["F"].copy()
        usl.procedure(dt, Ti)
        usf.procedure(dt, Ti)
        Ti += dt
        cur, got = usl.state(), usf.state()
        for key in ("xi", "vs"):
            np.testing.assert_array_equal(got[key], cur[key], err_msg=f"{key} step {k}")
        for key in ("sigma", "eps_s", "eps_q", "eps_v", "omega"):
            np.testing.assert_array_equal(got[key], prev[key], err_msg=f"{key} step {k}")
        np.testing.assert_array_equal(usf.a["F"], Fprev)
    # and the schedules really differ from MUSL
    musl = orc.Oracle(*mk("MUSL"))
    musl.run(dt, 0.0, 12)
    assert np.abs(musl.state()["sigma"] - cur["sigma"]).max() > 0
