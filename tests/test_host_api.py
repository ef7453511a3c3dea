"""Host-side mirror of the reference API, checked without a GPU: constructors
and their validation, node ordering, Julia ranges, and the exact C ABI call
sequence of materialpointsolver (with a recording stand-in for the library)."""
import numpy as np
import pytest


def test_julia_ranges(mpm):
    r = mpm.julia_range(0.0, 0.1, 1.0)
    assert len(r) == 11 and r[3] == 0.3          # Julia: (0:0.1:1)[4] == 0.3
    g = mpm.julia_range(-0.025, 0.0025, 0.82)
    assert len(g) == 339 and g[10] == 0.0
    assert len(mpm.julia_range(0.0, 0.3, 1.0)) == 4


def test_grid2d_node_order(mpm):
    g = mpm.UserGrid2D(x1=0.0, x2=0.3, y1=0.0, y2=0.2, dx=0.1, dy=0.1)
    assert (g.nnx, g.nny, g.ni, g.ncx, g.ncy, g.nc) == (4, 3, 12, 3, 2, 6)
    # node n = row + (col - 1) * nny, rows from the top (grid.jl:92-100)
    assert np.allclose(g.ξ[:3, 1], [0.2, 0.1, 0.0])
    assert np.allclose(g.ξ[:, 0], np.repeat([0.0, 0.1, 0.2, 0.3], 3))
    assert g.p2nD.shape == (16, 2) and g.p2nD[0, 0] == -3 and g.p2nD[15, 1] == -1


def test_grid3d_node_order(mpm):
    g = mpm.UserGrid3D(x1=0, x2=0.2, y1=0, y2=0.3, z1=0, z2=0.1, dx=0.1, dy=0.1, dz=0.1)
    assert (g.nnx, g.nny, g.nnz, g.ni) == (3, 4, 2, 24)
    n = lambda ix, iy, iz: iy + ix * g.nny + iz * g.nny * g.nnx  # 0-based (grid.jl:239-249)
    for ix in range(3):
        for iy in range(4):
            for iz in range(2):
                assert np.allclose(g.ξ[n(ix, iy, iz)], [0.1 * ix, 0.1 * iy, 0.1 * iz])
    assert g.p2nD.shape == (64, 4)


def test_meshbuilder_order(mpm):
    p = mpm.meshbuilder(np.array([0.0, 1.0]), np.array([5.0, 6.0, 7.0]))
    assert p.tolist() == [[0, 5], [0, 6], [0, 7], [1, 5], [1, 6], [1, 7]]
    q = mpm.meshbuilder(np.array([0.0, 1.0]), np.array([5.0, 6.0]), np.array([8.0, 9.0]))
    assert q[:4].tolist() == [[0, 5, 8], [0, 6, 8], [1, 5, 8], [1, 6, 8]]


def test_args_validation(mpm):
    base = dict(Ttol=1.0, ΔT=1e-3, constitutive="druckerprager", project_name="t",
                project_path="/tmp")
    mpm.UserArgs2D(**base)
    for bad, msg in [(dict(FLIP=0.5), r"FLIP\+PIC"), (dict(ζs=2), "ζs"), (dict(Te=3.0), "Elastic"),
                     (dict(basis="quad"), "basis"), (dict(va="x"), "velocity"),
                     (dict(device="TPU"), "device"), (dict(ΔT=2.0), "Time step")]:
        with pytest.raises(ValueError, match=msg):
            mpm.UserArgs3D(**{**base, **bad})
    a = mpm.UserArgs2D(**base, animation=True)
    assert a.hdf5


def test_particle_and_property(mpm):
    mp = mpm.UserParticle3D(dx=0.1, dy=0.1, dz=0.2, ξ=np.zeros((5, 3)), ρs=np.full(5, 2.0))
    assert mp.np == 5 and np.allclose(mp.Ω0, 0.002) and np.allclose(mp.F[:, [0, 4, 8]], 1)
    with pytest.raises(ValueError):
        mpm.UserParticle2D(dx=0.1, dy=0.1, ξ=np.zeros((5, 3)), ρs=[1.0])
    with pytest.raises(ValueError, match="nid layer"):
        mpm.UserProperty(nid=[1, 1, 2], ν=[0.3], Es=[1.0], Gs=[1.0], Ks=[1.0])
    at = mpm.UserProperty(nid=[1, 2, 2], ν=[0.3, 0.2], Es=[1, 2], Gs=[1, 2], Ks=[1, 2], ϕ=[0.5])
    assert at.nmat == 2 and np.allclose(at.ϕ, [0.5, 0.5])


def test_step_args_schemes(mpm):
    # procedure!'s Val(:MUSL) / Val(:USL) / Val(:USF) map onto mpm_args.scheme
    for name, code in (("MUSL", 0), ("USL", 1), ("USF", 2)):
        args = mpm.scenes.column_collapse_2d(h=0.02, nsteps=2, scheme=name)[0]
        assert mpm.HIPModel.step_args(args).scheme == code
    args.coupling = "TS"
    with pytest.raises(NotImplementedError, match="coupling"):
        mpm.HIPModel.step_args(args)
    args.coupling, args.scheme = "OS", "APIC"
    with pytest.raises(NotImplementedError):
        mpm.HIPModel.step_args(args)


def test_non_hip_device_raises(mpm):
    args, grid, mp, attr, bc = mpm.scenes.column_collapse_2d(h=0.02, nsteps=2)
    with pytest.raises(NotImplementedError, match="HIP"):
        mpm.host2device(grid, mp, attr, bc, "CPU")


class _Recorder:
    """Stand-in for libmpm_hip.so that records the ABI calls."""

    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        def f(*a):
            self.calls.append((name, a))
            if name == "mpm_last_error":
                return b"ok"
            if name == "mpm_abi_version":
                return 2
            return 0
        return f


def test_solver_call_sequence_and_clock(mpm, monkeypatch):
    rec = _Recorder()
    monkeypatch.setattr(mpm._capi, "_lib", rec)
    args, grid, mp, attr, bc = mpm.scenes.column_collapse_2d(h=0.02, dt=2e-5, Ttol=1e-3, eps="FP32")
    mpm.materialpointsolver(args, grid, mp, attr, bc)
    names = [c[0] for c in rec.calls]
    assert names[:4] == ["mpm_create", "mpm_upload_particles", "mpm_upload_materials", "mpm_upload_bc"]
    steps = [c for c in rec.calls if c[0] == "mpm_step"]
    # Ti accumulates in T2 = Float32, as in submit_work! (solver.jl:118-124)
    Ti, n = np.float32(0), 0
    while Ti < np.float32(1e-3):
        Ti = np.float32(Ti + np.float32(2e-5))
        n += 1
    assert len(steps) == n == args.iter_num
    assert [s[1][3] for s in steps[:3]] == [0.0, float(np.float32(2e-5)),
                                            float(np.float32(np.float32(2e-5) * 2))]
    assert names[-3:] == ["mpm_sync", "mpm_download_particles", "mpm_destroy"]


def test_autotune_picks_the_fastest_and_leaves_it_set(mpm):
    class FakeModel:
        """Step cost depends on the options; records the calls."""
        def __init__(self):
            self.opts, self.now, self.calls = {}, 0.0, []

        def set_option(self, k, v):
            self.opts[k] = v
            self.calls.append((k, v))

        def run(self, sargs, dt, Ti, n, graph=False):
            cost = {5: 3.0, 10: 1.0, 20: 1.5, 40: 2.0}[self.opts["sort_interval"]]
            form = {"owned": 1.0, "owned_layers": 0.9, "rounds": 1.3,
                    "rounds_layered": 1.1, "pairs": 0.95}[self.opts.get("scatter", "owned")]
            self.now += n * cost * (form if self.opts["tiles"] else 4.0)
            return Ti + n * dt

        def sync(self):
            pass

    m = FakeModel()
    best, table, Ti = mpm.autotune(m, None, 0.5, 0.0, steps=4, warm=1, clock=lambda: m.now)
    assert best == dict(sort_interval=10, tiles=True, scatter="rounds")
    assert m.opts == {"sort_interval": 10, "tiles": True, "scatter": "rounds",
                      "remap_scatter": None}
    assert len(table) == 8 and abs(min(t[3] for t in table) - 1.3) < 1e-12
    assert Ti == 8 * 5 * 0.5  # every trial advanced the clock
    # with every scatter form: 5 forms x 4 intervals with tiles + 4 without
    m = FakeModel()
    best, table, Ti = mpm.autotune(m, None, 0.5, 0.0, steps=4, warm=1, clock=lambda: m.now,
                                   scatters=mpm.SCATTER_FORMS)
    assert best == dict(sort_interval=10, tiles=True, scatter="owned_layers")
    assert m.opts == {"sort_interval": 10, "tiles": True, "scatter": "owned_layers",
                      "remap_scatter": None}
    assert len(table) == 24 and abs(min(t[3] for t in table) - 0.9) < 1e-12
    # a (P2G #1, re-map scatter) pair sets both forms
    m = FakeModel()
    best, table, Ti = mpm.autotune(m, None, 0.5, 0.0, steps=4, warm=1, clock=lambda: m.now,
                                   scatters=(("owned_layers", "rounds"), "rounds"),
                                   sort_intervals=(10,))
    assert best["scatter"] == ("owned_layers", "rounds")
    assert m.opts["scatter"] == "owned_layers" and m.opts["remap_scatter"] == "rounds"


def test_snapshot_layout_and_cadence(mpm, monkeypatch, tmp_path):
    # groups at iterations 0, hdf5_step, 2 hdf5_step, ... with the reference's
    # dataset names (src/solver.jl:53-115); the download precedes each group
    rec = _Recorder()
    monkeypatch.setattr(mpm._capi, "_lib", rec)
    args, grid, mp, attr, bc = mpm.scenes.column_collapse_2d(
        h=0.02, dt=2e-5, Ttol=1e-3, eps="FP32", project_name="snap", project_path=str(tmp_path))
    args.hdf5, args.hdf5_step = True, 7
    mpm.materialpointsolver(args, grid, mp, attr, bc)
    n = args.iter_num
    d = tmp_path / "snap" / "snap_snapshots"
    groups = sorted(p.name for p in d.glob("group*.npz"))
    want = len(range(0, n, 7))
    assert len(groups) == want
    g = np.load(d / "group1.npz")
    assert set(g.files) == {"stress", "strain_s", "eqstrain", "ekstrain", "eqrate", "coords",
                            "velocity_s", "volume", "mass_s", "time"}
    assert g["coords"].shape == (mp.np, 2) and float(g["time"]) == 0.0
    g2 = np.load(d / "group2.npz")
    assert float(g2["time"]) == pytest.approx(7 * 2e-5, rel=1e-5)
    meta = np.load(d / "meta.npz")
    assert int(meta["FILE_NUM"]) == want + 1
    assert meta["grid_coords"].shape == (grid.ni, 2) and meta["nid"].shape == (mp.np,)
    names = [c[0] for c in rec.calls]
    # one download per group plus the final one
    assert names.count("mpm_download_particles") == want + 1


def test_snapshot_state_for_resume(mpm):
    # the arrays a resume uploads from one snapshot group: F carries det F =
    # volume / Ω0 (the reference saves no F), ρs = mass / volume, ρs0 = mass / Ω0
    rng = np.random.default_rng(3)
    for dim in (2, 3):
        n, ns = 50, 4 if dim == 2 else 6
        g = dict(coords=rng.normal(size=(n, dim)), velocity_s=rng.normal(size=(n, dim)),
                 stress=rng.normal(size=(n, ns)), strain_s=rng.normal(size=(n, ns)),
                 eqstrain=rng.random(n), ekstrain=rng.random(n), eqrate=rng.random(n),
                 volume=rng.uniform(0.5, 2.0, n), mass_s=rng.uniform(1.0, 3.0, n),
                 time=np.asarray(0.125))
        vol0 = rng.uniform(0.5, 2.0, n)
        arrs, Ti = mpm.HIPModel.snapshot_state(g, vol0, np.ones(n))
        assert Ti == 0.125
        F = arrs["F"].reshape(n, dim, dim)
        np.testing.assert_allclose(np.linalg.det(F), g["volume"] / vol0, rtol=1e-14)
        np.testing.assert_allclose(arrs["rho"] * arrs["omega"], g["mass_s"], rtol=1e-15)
        np.testing.assert_allclose(arrs["rho0"] * arrs["omega0"], g["mass_s"], rtol=1e-15)
        for k, src in (("xi", "coords"), ("vs", "velocity_s"), ("sigma", "stress"),
                       ("eps_s", "strain_s"), ("eps_q", "eqstrain"), ("eps_k", "ekstrain")):
            np.testing.assert_array_equal(arrs[k], g[src])
            assert arrs[k].flags.f_contiguous


def test_particle_arrays_with_wrong_columns_are_rejected(mpm, monkeypatch):
    # the C ABI reads np x k host arrays; a wrong column count must stop on the host
    rec = _Recorder()
    monkeypatch.setattr(mpm._capi, "_lib", rec)
    args, grid, mp, attr, bc = mpm.scenes.column_collapse_3d(h=0.04, width=0.16, height=0.08,
                                                             box_xy=0.4, box_z=0.16, nsteps=2)
    mp.σij = np.asfortranarray(mp.σij[:, :4])
    with pytest.raises(ValueError, match="sigma must be np x 6"):
        mpm.HIPModel(grid, mp, attr, bc)
    assert "mpm_upload_particles" not in [c[0] for c in rec.calls]
