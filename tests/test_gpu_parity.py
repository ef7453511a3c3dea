"""HIP path vs the FP64 oracle on the same inputs, through the C ABI.

FP64 mode must agree with the oracle to rounding (relative L2 <= 1e-9 over a
short horizon; the oracle and the device sum the scatters in different orders
and the device tracks det F as a product of det(I + ΔF)).  FP32 mode (FP32
bulk, FP64 positions and det F) is held to the tolerances written per test.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def _compare(mpm, orc, scene, nsteps, tol, sort_interval=0, keys=("xi", "vs", "sigma", "eps_s",
                                                                   "eps_q", "omega"),
             scatter=None, fused_remap=None, tile_partials=None, remap_scatter=None):
    args, grid, mp, attr, bc = scene
    print(f"case: dim={grid.dim} eps={mp.ϵ} np={mp.np} const={args.constitutive} va={args.va} "
          f"MVL={args.MVL} NIC={grid.NIC} steps={nsteps} sort={sort_interval} "
          f"scatter={scatter or 'default'} fused_remap={fused_remap} "
          f"tile_partials={tile_partials} remap_scatter={remap_scatter or 'default'}", flush=True)
    model = mpm.HIPModel(grid, mp, attr, bc, sort_interval=sort_interval)
    try:
        if scatter is not None:
            model.set_option("scatter", scatter)
        if remap_scatter is not None:
            model.set_option("remap_scatter", remap_scatter)
        if fused_remap is not None:
            model.set_option("fused_remap", fused_remap)
        if tile_partials is not None:
            model.set_option("tile_partials", tile_partials)
        sargs = model.step_args(args)
        model.run(sargs, float(args.ΔT), 0.0, nsteps)
        model.sync()
        got = model.download()
    finally:
        model.close()
    ref = orc.Oracle(args, grid, mp, attr, bc)
    ref.run(float(args.ΔT), 0.0, nsteps)
    want = ref.state()
    x0 = np.asarray(mp.ξ, dtype=np.float64)
    out = {}
    for k in keys:
        if k == "xi":
            r = _rel(got[k] - x0, want[k] - x0)
        else:
            r = _rel(got[k], want[k])
        out[k] = r
    print("  rel L2:", {k: f"{v:.2e}" for k, v in out.items()}, flush=True)
    for k, r in out.items():
        t = tol[k] if isinstance(tol, dict) else tol
        assert r <= t, f"{k}: rel L2 {r:.3e} > {t:g}"
    return got, want


def _c2(mpm, **kw):
    base = dict(h=0.01, width=0.2, height=0.1, box_x=0.5, box_y=0.12, dt=5e-5, nsteps=200)
    base.update(kw)
    return mpm.scenes.column_collapse_2d(**base)


def _c3(mpm, **kw):
    base = dict(h=0.02, width=0.16, height=0.08, box_xy=0.4, box_z=0.12, nsteps=100, eps="FP64")
    base.update(kw)
    return mpm.scenes.column_collapse_3d(**base)


# sort_interval 0 runs the plain scatter kernels, > 0 the LDS tile-binned ones
@pytest.mark.parametrize("sort", [0, 5])
@pytest.mark.parametrize("constitutive", ["linearelastic", "druckerprager", "mohrcoulomb"])
@pytest.mark.parametrize("va", ["a", "v"])
def test_fp64_2d(mpm, orc, constitutive, va, sort):
    _compare(mpm, orc, _c2(mpm, constitutive=constitutive, va=va, cohesion=200.0), 200, 1e-9,
             sort_interval=sort)


@pytest.mark.parametrize("sort", [0, 5])
@pytest.mark.parametrize("constitutive", ["linearelastic", "druckerprager"])
@pytest.mark.parametrize("va", ["a", "v"])
def test_fp64_3d(mpm, orc, constitutive, va, sort):
    _compare(mpm, orc, _c3(mpm, constitutive=constitutive, va=va), 100, 1e-9, sort_interval=sort)


@pytest.mark.parametrize("sort", [0, 5])
@pytest.mark.parametrize("va", ["a", "v"])
def test_fp64_3d_mohrcoulomb(mpm, orc, va, sort):
    # 3D Mohr-Coulomb is defined by this library (principal-space return with
    # dilation 5 deg; the reference's 3D mcP! is a no-op); every particle of
    # this column yields within the 100 steps
    _compare(mpm, orc, _c3(mpm, constitutive="mohrcoulomb", va=va, cohesion=200.0, psi_deg=5.0),
             100, 1e-9, sort_interval=sort)


# The three tile-scatter forms (base-owned register sums, the same one z
# layer per pass, rank rounds into an LDS window) against the oracle.  Sort
# interval 5 keeps most particles on their sorted base; 400 sorts once at the
# start, so over 200 collapse steps many particles drift to other bases, other
# tiles' windows and out of every window (the drift lists and their global
# atomic fallback carry them).
@pytest.mark.parametrize("sort", [5, 400])
@pytest.mark.parametrize("scatter", ["owned", "owned_layers", "rounds"])
def test_fp64_scatter_forms(mpm, orc, scatter, sort):
    _compare(mpm, orc, _c2(mpm, constitutive="druckerprager", cohesion=200.0), 200, 1e-9,
             sort_interval=sort, scatter=scatter)
    _compare(mpm, orc, _c3(mpm, constitutive="druckerprager"), 100, 1e-9, sort_interval=sort,
             scatter=scatter)


# Tile partials (windows stored per tile and summed by the grid solves) and the
# atomic window flush, with every scatter form, against the oracle
@pytest.mark.parametrize("partials", [True, False])
@pytest.mark.parametrize("scatter", ["owned", "owned_layers", "rounds"])
def test_tile_partials_and_atomic_flush(mpm, orc, scatter, partials):
    for sort in (5, 400):
        _compare(mpm, orc, _c2(mpm, constitutive="druckerprager", cohesion=200.0), 200, 1e-9,
                 sort_interval=sort, scatter=scatter, tile_partials=partials)
        _compare(mpm, orc, _c3(mpm, constitutive="druckerprager"), 100, 1e-9,
                 sort_interval=sort, scatter=scatter, tile_partials=partials)
    _compare(mpm, orc, _c3(mpm, basis="linear", scheme="USL"), 60, 1e-9, sort_interval=5,
             scatter=scatter, tile_partials=partials)


@pytest.mark.parametrize("scatter", ["rounds", "owned"])
@pytest.mark.parametrize("fused", [True, False])
def test_remap_one_pass_and_split(mpm, orc, fused, scatter):
    # the one-pass MUSL re-map (gather, move and momentum scatter in one tile
    # kernel: in the rank rounds, or base-owned) against the gather +
    # momentum-scatter pair, both against the oracle; FP64 in 2D and 3D, FP32
    # in 3D; sort interval 400 adds drifted particles
    tol = dict(xi=2e-4, vs=5e-4, sigma=2e-3, eps_s=5e-4, eps_q=2e-3, omega=1e-5)
    kw = dict(fused_remap=fused, scatter=scatter, remap_scatter=scatter)
    for sort in (5, 400):
        _compare(mpm, orc, _c2(mpm, constitutive="druckerprager", cohesion=200.0), 200, 1e-9,
                 sort_interval=sort, **kw)
        _compare(mpm, orc, _c2(mpm, constitutive="druckerprager", cohesion=200.0, va="v"), 200,
                 1e-9, sort_interval=sort, **kw)
        _compare(mpm, orc, _c3(mpm, eps="FP32", constitutive="druckerprager"), 100, tol,
                 sort_interval=sort, **kw)
        _compare(mpm, orc, _c3(mpm, constitutive="druckerprager", va="v"), 100, 1e-9,
                 sort_interval=sort, **kw)
    _compare(mpm, orc, _c3(mpm, basis="linear"), 60, 1e-9, sort_interval=5, **kw)
    _compare(mpm, orc, _c3(mpm, basis="linear", scheme="USL"), 60, 1e-9, sort_interval=5, **kw)


@pytest.mark.parametrize("scatter",
                         ["owned", "owned_layers", "rounds", "rounds_layered", "pairs"])
def test_fp32_scatter_forms(mpm, orc, scatter):
    tol = dict(xi=2e-4, vs=5e-4, sigma=2e-3, eps_s=5e-4, eps_q=2e-3, omega=1e-5)
    _compare(mpm, orc, _c3(mpm, eps="FP32", constitutive="druckerprager"), 100, tol,
             sort_interval=10, scatter=scatter)
    _compare(mpm, orc, _c3(mpm, eps="FP32", constitutive="druckerprager"), 100, tol,
             sort_interval=400, scatter=scatter)


# The re-map's momentum scatter in each form, with P2G #1 in the rank rounds
# (FP32 3D uGIMP takes the base-owned form by default), sorted often and rarely
@pytest.mark.parametrize("remap_scatter",
                         ["rounds", "owned", "owned_layers", "rounds_layered", "pairs"])
def test_fp32_remap_scatter_forms(mpm, orc, remap_scatter):
    tol = dict(xi=2e-4, vs=5e-4, sigma=2e-3, eps_s=5e-4, eps_q=2e-3, omega=1e-5)
    for sort in (10, 400):
        _compare(mpm, orc, _c3(mpm, eps="FP32", constitutive="druckerprager"), 100, tol,
                 sort_interval=sort, scatter="rounds", remap_scatter=remap_scatter)
    _compare(mpm, orc, _c3(mpm), 60, 1e-9, sort_interval=5, scatter="rounds",
             remap_scatter=remap_scatter)


# The layer-window rank rounds (FP32 3D uGIMP) with tile partials and with the
# atomic flush, sorted often and rarely (drift out of the windows and onto
# bases other lanes hold), P2G #1 and the re-map scatter, :a and :v
@pytest.mark.parametrize("partials", [True, False])
def test_fp32_rounds_layered(mpm, orc, partials):
    tol = dict(xi=2e-4, vs=5e-4, sigma=2e-3, eps_s=5e-4, eps_q=2e-3, omega=1e-5)
    for sort in (5, 400):
        for va in ("a", "v"):
            _compare(mpm, orc, _c3(mpm, eps="FP32", constitutive="druckerprager", va=va), 100,
                     tol, sort_interval=sort, scatter="rounds_layered", tile_partials=partials)


# The pair rounds (FP32 3D uGIMP) with tile partials and with the atomic
# flush, sorted often and rarely (drift off the sort-time bases and out of the
# windows), P2G #1 and the re-map scatter, :a and :v
@pytest.mark.parametrize("partials", [True, False])
def test_fp32_pairs(mpm, orc, partials):
    tol = dict(xi=2e-4, vs=5e-4, sigma=2e-3, eps_s=5e-4, eps_q=2e-3, omega=1e-5)
    for sort in (5, 400):
        for va in ("a", "v"):
            _compare(mpm, orc, _c3(mpm, eps="FP32", constitutive="druckerprager", va=va), 100,
                     tol, sort_interval=sort, scatter="pairs", remap_scatter="pairs",
                     tile_partials=partials)


# Pair rounds on dual cells of 64 particles (4x4x4 per cell): pair groups up
# to rank 29, then ranks 30..63 one lane per base (rank 31 on in the merged
# last group of the sort), against the plain global-atomic kernels, which sum
# the same terms in another order.
def test_fp32_pairs_dense_cells(mpm):
    args, grid, mp, attr, bc = mpm.scenes.column_collapse_3d(
        h=0.02, width=0.08, height=0.04, box_xy=0.24, box_z=0.08, eps="FP32", nsteps=10)
    lp = 0.005
    w2 = 0.04
    lat = lambda lo, hi: np.arange(lo + lp / 2, hi, lp)
    pts = mpm.types.meshbuilder(lat(-w2, w2), lat(-w2, w2), lat(0.0, 0.04))
    mp = mpm.types.UserParticle3D(ϵ="FP32", NIC=27, dx=lp, dy=lp, dz=lp, ξ=pts,
                                  ρs=np.full(len(pts), mpm.scenes.RHO))
    attr = mpm.types.UserProperty(ϵ="FP32", nid=np.ones(mp.np), ν=[mpm.scenes.NU],
                                  Es=[mpm.scenes.E], Gs=[mpm.scenes.GS], Ks=[mpm.scenes.KS],
                                  σt=[0.0], ϕ=[np.radians(mpm.scenes.PHI_DEG)], ϕr=[0.0],
                                  ψ=[0.0], c=[0.0], cr=[0.0], Hp=[0.0])
    dt = float(args.ΔT)
    x0 = np.asarray(mp.ξ, dtype=np.float64)
    tiled = mpm.HIPModel(grid, mp, attr, bc, sort_interval=10)
    plain = mpm.HIPModel(grid, mp, attr, bc, sort_interval=10)
    try:
        tiled.set_option("scatter", "pairs")
        tiled.set_option("remap_scatter", "pairs")
        plain.set_option("tiles", False)
        sa = tiled.step_args(args)
        T = 0.0
        for nsteps, tol in ((1, 1e-5), (9, 1e-4)):
            tiled.run(sa, dt, T, nsteps)
            T = plain.run(sa, dt, T, nsteps)
            a, b = tiled.download(), plain.download()
            for k in ("xi", "vs", "sigma"):
                ga = a[k] - x0 if k == "xi" else a[k]
                gb = b[k] - x0 if k == "xi" else b[k]
                r = _rel(ga, gb)
                print(f"  pairs, 64 per cell, after {nsteps} more step(s): {k} rel L2 {r:.2e}",
                      flush=True)
                assert r <= tol, (k, nsteps, r)
    finally:
        tiled.close()
        plain.close()


# Mohr-Coulomb softening.  2D: cohesion max(c + Hp ϵq, cr) (mohrcoulomb.jl:36;
# the reference's 2D mcP! ignores ϕr).  3D: cohesion the same, friction
# 19.8 -> 12 deg and dilation 5 -> 0 deg linearly over the plastic strain in
# which cohesion softens (eqr = (cr - c) / Hp = 0.01; mohr_coulomb.tcc:8-25).
# Both columns end with particles before, inside and past the softening range.
@pytest.mark.parametrize("sort", [0, 5])
def test_fp64_mohrcoulomb_softening(mpm, orc, sort):
    soft = dict(constitutive="mohrcoulomb", cohesion=200.0, Hp=-1.8e4, cr=20.0, phir_deg=12.0,
                psi_deg=5.0)
    for sc, n in ((_c2(mpm, **soft), 200), (_c3(mpm, **soft), 100)):
        got, want = _compare(mpm, orc, sc, n, 1e-9, sort_interval=sort)
        eq = want["eps_q"]
        inside = np.mean((eq > 1e-3) & (eq < 0.01))
        print(f"  softening: {inside:.2f} of the particles inside the range, "
              f"{np.mean(eq >= 0.01):.2f} past it", flush=True)
        assert inside > 0.05 and np.mean(eq >= 0.01) > 0.005


# USL and USF schedules (OS_USL.jl, OS_USF.jl); PIC = 0.3 so that the USL :a
# re-map's PIC term (gathered from the new nodal velocity) is exercised
@pytest.mark.parametrize("sort", [0, 5])
@pytest.mark.parametrize("va", ["a", "v"])
@pytest.mark.parametrize("scheme", ["USL", "USF"])
def test_fp64_2d_schemes(mpm, orc, scheme, va, sort):
    sc = _c2(mpm, constitutive="druckerprager", va=va, cohesion=200.0, scheme=scheme)
    sc[0].FLIP, sc[0].PIC = 0.7, 0.3
    _compare(mpm, orc, sc, 200, 1e-9, sort_interval=sort)


@pytest.mark.parametrize("sort", [0, 5])
@pytest.mark.parametrize("constitutive", ["linearelastic", "mohrcoulomb"])
@pytest.mark.parametrize("scheme", ["USL", "USF"])
def test_fp64_3d_schemes(mpm, orc, scheme, constitutive, sort):
    sc = _c3(mpm, constitutive=constitutive, cohesion=200.0, psi_deg=5.0, scheme=scheme)
    sc[0].FLIP, sc[0].PIC = 0.7, 0.3
    _compare(mpm, orc, sc, 100, 1e-9, sort_interval=sort)


@pytest.mark.parametrize("scheme", ["USL", "USF"])
def test_fp64_schemes_linear_basis(mpm, orc, scheme):
    # the linear basis (NIC 4 / 8) under both schedules, plain and tiled kernels
    for sort in (0, 5):
        _compare(mpm, orc, _c2(mpm, basis="linear", scheme=scheme), 100, 1e-9, sort_interval=sort)
        _compare(mpm, orc, _c3(mpm, basis="linear", scheme=scheme), 60, 1e-9, sort_interval=sort)


def test_fp64_schemes_mvl_and_ramp(mpm, orc):
    # USF runs MVL after the position update, on the stress of the step's head
    for scheme in ("USL", "USF"):
        _compare(mpm, orc, _c2(mpm, MVL=True, scheme=scheme, Te=0.004), 200, 1e-9)
        _compare(mpm, orc, _c3(mpm, MVL=True, scheme=scheme), 60, 1e-9, sort_interval=4)


@pytest.mark.parametrize("scheme", ["USL", "USF"])
def test_fp32_3d_usl_usf(mpm, orc, scheme):
    # USL/USF have no velocity re-map: the strain increment comes from
    # (p + f dt) / m, which is large at nodes that hold a sliver of mass, so the
    # step is far worse conditioned than MUSL's.  In the FP64 oracle a relative
    # perturbation of 6e-8 (FP32 rounding) of the initial positions of this
    # column changes σ after 100 steps by 1.0-1.4e-2 relative L2 (Drucker-Prager,
    # c = 0) and 8-9e-4 (linear elastic), against 7e-7 and 3.5e-7 under MUSL.
    # The FP32 bars below are about twice that conditioning.
    tol = dict(xi=2e-4, vs=1e-3, sigma=3e-2, eps_s=5e-4, eps_q=3e-2, omega=1e-5)
    _compare(mpm, orc, _c3(mpm, eps="FP32", scheme=scheme), 100, tol, sort_interval=10)
    tol = dict(xi=2e-4, vs=5e-4, sigma=3e-3, eps_s=5e-4, eps_q=3e-3, omega=1e-5)
    _compare(mpm, orc, _c3(mpm, eps="FP32", scheme=scheme, constitutive="linearelastic"), 100,
             tol, sort_interval=10)


def test_fp32_3d_mohrcoulomb(mpm, orc):
    tol = dict(xi=2e-4, vs=5e-4, sigma=2e-3, eps_s=5e-4, eps_q=2e-3, omega=1e-5)
    _compare(mpm, orc, _c3(mpm, eps="FP32", constitutive="mohrcoulomb", cohesion=200.0,
                           psi_deg=5.0), 100, tol, sort_interval=10)


def test_fp64_tiled_with_drift(mpm, orc):
    # sorted once at step 0 only: the block moves ~1.2 cells, so particles leave
    # their tile windows and take the direct-scatter fallback
    _compare(mpm, orc, _flying_block(mpm, 3, zeta=0.0, speed=4.0), 60, 1e-9, sort_interval=100000)
    _compare(mpm, orc, _flying_block(mpm, 2, zeta=0.0, speed=4.0), 100, 1e-9, sort_interval=100000)


def test_fp64_mvl(mpm, orc):
    # MVL: cell-averaged mean stress (vollock1/2)
    _compare(mpm, orc, _c2(mpm, MVL=True), 200, 1e-9)
    _compare(mpm, orc, _c3(mpm, MVL=True), 60, 1e-9, sort_interval=4)


def test_fp64_gravity_ramp_and_elastic_gate(mpm, orc):
    # Te > 0: gravity ramps with Ti / Te and plasticity starts at Ti >= Te
    _compare(mpm, orc, _c2(mpm, Te=0.004), 200, 1e-9)
    _compare(mpm, orc, _c3(mpm, Te=0.004), 60, 1e-9)


def _flying_block(mpm, dim, eps="FP64", zeta=0.1, speed=1.0):
    """A stressed elastic block in free flight far from every wall, with a uniform
    initial velocity: every nodal velocity component keeps a clear sign, so the
    damping term ζ|f| sign(v) of the grid solve is well defined to rounding."""
    if dim == 2:
        sc = mpm.scenes.column_collapse_2d(h=0.01, width=0.1, height=0.06, box_x=0.5, box_y=0.3,
                                           dt=5e-5, nsteps=100, eps=eps, zeta=zeta,
                                           constitutive="linearelastic")
    else:
        sc = mpm.scenes.column_collapse_3d(h=0.02, width=0.12, height=0.08, box_xy=0.6,
                                           box_z=0.4, nsteps=60, eps=eps, zeta=zeta,
                                           constitutive="linearelastic")
    args, grid, mp, attr, bc = sc
    shift = np.zeros(dim)
    shift[-1] = 0.1
    shift[0] = 0.1 if dim == 2 else 0.0
    mp.ξ[...] = mp.ξ + shift
    v0 = np.array([speed, 0.7] if dim == 2 else [speed, 0.7, 0.5])
    mp.vs[...] = v0
    mp.σij[:, 0] = -300.0
    mp.σij[:, 1] = -200.0
    mp.σij[:, 2] = -100.0
    return args, grid, mp, attr, bc


def test_fp64_damping(mpm, orc):
    # (A damped collapse from rest cannot be compared to rounding: its lateral
    # nodal velocities are zero up to rounding from the second step on, and the
    # reference's damping force ζ|f| sign(v) flips with their sign.)
    _compare(mpm, orc, _flying_block(mpm, 2), 100, 1e-9)
    _compare(mpm, orc, _flying_block(mpm, 3), 60, 1e-9)


def test_fp64_linear_basis(mpm, orc):
    _compare(mpm, orc, _c2(mpm, basis="linear"), 100, 1e-9)
    _compare(mpm, orc, _c3(mpm, basis="linear"), 40, 1e-9)


def test_fp64_sorted_matches_unsorted(mpm, orc):
    # the cell sort permutes particles; results are returned in caller order
    _compare(mpm, orc, _c3(mpm), 60, 1e-9, sort_interval=7)
    _compare(mpm, orc, _c2(mpm), 100, 1e-9, sort_interval=3)


def test_fp32_short_horizon(mpm, orc):
    # FP32 bulk vs the FP64 oracle.  Strain increments are differences of nearly
    # equal FP32 nodal velocities (|v| >> |Δv| across a cell early in a
    # collapse), which bounds the stress agreement at ~1e-3 after 200 steps.
    tol = dict(xi=2e-4, vs=5e-4, sigma=2e-3, eps_s=5e-4, eps_q=2e-3, omega=1e-5)
    _compare(mpm, orc, _c2(mpm, eps="FP32"), 200, tol, sort_interval=10)
    _compare(mpm, orc, _c3(mpm, eps="FP32"), 100, tol, sort_interval=10)


def test_fp32_flying_block(mpm, orc):
    # A 1 m/s uniform translation stored in FP32 particle velocities carries
    # ~1e-7 m/s rounding per step on top of ~4e-4 m/s velocity variations that
    # drive the strain, so stress / strain agree to ~2e-3 (Galilean precision
    # limit of FP32 velocities); positions and volumes stay at ~1e-7.
    tol = dict(xi=1e-6, vs=1e-5, sigma=5e-3, eps_s=5e-3, eps_q=1e-3, omega=1e-6)
    _compare(mpm, orc, _flying_block(mpm, 3, eps="FP32"), 60, tol, sort_interval=5)


def test_fp64_body_force_extension(mpm, orc):
    # gravity on the first axes (args.ext; not in the reference) in the parity mode
    sc = _c3(mpm, Te=0.004)
    sc[0].ext = {"gravity_x": 2.0, "gravity_y": -1.0}
    _compare(mpm, orc, sc, 60, 1e-9, sort_interval=5)
    sc2 = _c2(mpm)
    sc2[0].ext = {"gravity_x": 3.0}
    _compare(mpm, orc, sc2, 100, 1e-9)


@pytest.mark.parametrize("dim,basis", [(3, "uGIMP"), (3, "linear"), (2, "uGIMP"), (2, "linear")])
def test_fp32_tiled_kernels_match_plain_kernels(mpm, dim, basis):
    # A dense FP32 column right after the sort, where most rounds of the tile
    # scatters are full (64 particles on the 64 bases of their tile): the LDS
    # tile kernels and the plain global-atomic kernels sum the same terms, so the
    # two differ by FP32 summation order only (a missing or misplaced term would
    # show at 1e-2 and above).  Stresses come from differences of nearly equal
    # nodal velocities, which amplifies that rounding; the bounds leave 10x room.
    # (FP32 uGIMP windows interleave field pairs; linear-basis windows do not)
    if dim == 3:
        args, grid, mp, attr, bc = mpm.scenes.column_collapse_3d(
            h=0.01, width=0.16, height=0.08, box_xy=0.4, box_z=0.12, eps="FP32", nsteps=10,
            basis=basis)
    else:
        args, grid, mp, attr, bc = mpm.scenes.column_collapse_2d(
            h=0.005, width=0.2, height=0.1, box_x=0.5, box_y=0.14, eps="FP32", nsteps=10,
            basis=basis)
    dt = float(args.ΔT)
    x0 = np.asarray(mp.ξ, dtype=np.float64)
    tiled = mpm.HIPModel(grid, mp, attr, bc, sort_interval=10)
    plain = mpm.HIPModel(grid, mp, attr, bc, sort_interval=10)
    try:
        plain.set_option("tiles", False)
        sa = tiled.step_args(args)
        T = 0.0
        # measured: rel L2 ~1e-7 (x, v) and ~7e-7 (σ) after one step, on every case
        for nsteps, tol in ((1, 1e-5), (4, 1e-4)):
            tiled.run(sa, dt, T, nsteps)
            T = plain.run(sa, dt, T, nsteps)
            a, b = tiled.download(), plain.download()
            for k in ("xi", "vs", "sigma"):
                ga = a[k] - x0 if k == "xi" else a[k]
                gb = b[k] - x0 if k == "xi" else b[k]
                r = _rel(ga, gb)
                print(f"  after {nsteps} more step(s): {k} rel L2 {r:.2e}", flush=True)
                assert r <= tol, (k, nsteps, r)
    finally:
        tiled.close()
        plain.close()


# Particles whose stencil leaves the grid are frozen: every kernel skips them
# and mpm_sync reports them.  The sort places them after every tile, outside
# every tile segment, so they cost the tile kernels nothing; the tiled run must
# give the plain run's state, frozen particles included.  A small FP64 column
# without any boundary condition and a downward initial velocity falls out
# through the bottom of the grid part by part.
def test_frozen_particles_tiled_and_plain(mpm):
    states = {}
    for sort in (0, 5):
        args, grid, mp, attr, bc = _c3(mpm)
        for k in ("vx_s_idx", "vy_s_idx", "vz_s_idx"):
            getattr(bc, k)[:] = 0
        mp.vs[:, 2] = -5.0
        print(f"case: frozen particles, 3D FP64 np={mp.np} sort={sort}", flush=True)
        model = mpm.HIPModel(grid, mp, attr, bc, sort_interval=sort)
        try:
            sargs = model.step_args(args)
            dt = float(args.ΔT)
            Ti = model.run(sargs, dt, 0.0, 60)
            with pytest.raises(mpm._capi.MPMError, match="left the grid"):
                model.sync()
            mid = model.download()
            model.run(sargs, dt, Ti, 20)
            with pytest.raises(mpm._capi.MPMError):
                model.sync()
            end = model.download()
        finally:
            model.close()
        # the uGIMP stencil base floor(u - 1/2) is below the grid for z < z1 + dz / 2
        gone = mid["xi"][:, 2] < float(grid.z1) + 0.5 * float(grid.dz)
        # frozen particles keep their positions
        assert np.array_equal(end["xi"][gone], mid["xi"][gone])
        states[sort] = (gone, end)
        print(f"  {gone.sum()} of {mp.np} particles frozen after 60 steps", flush=True)
    (g0, e0), (g5, e5) = states[0], states[5]
    assert 0 < g0.sum() < len(g0) and np.array_equal(g0, g5)
    for key in ("xi", "vs"):
        d = np.linalg.norm(e5[key] - e0[key]) / np.linalg.norm(e0[key])
        print(f"  tiled vs plain {key}: rel L2 {d:.1e}", flush=True)
        assert d <= 1e-9
    # in free fall the stresses stay at rounding level: compare them against the
    # column's own scale rho g H instead
    scale = 2650.0 * 9.8 * 0.08
    ds = np.abs(e5["sigma"] - e0["sigma"]).max()
    print(f"  tiled vs plain sigma: max |diff| {ds:.1e} Pa (largest |sigma| "
          f"{np.abs(e0['sigma']).max():.1e} Pa, rho g H {scale:.0f} Pa)", flush=True)
    assert ds <= 1e-9 * scale
