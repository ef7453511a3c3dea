"""Scene builders that the benchmarks and GPU tests use (no GPU needed)."""
import math

import numpy as np


def test_gaussian_random_field_statistics(mpm):
    grf = mpm.scenes.gaussian_random_field
    f = grf(400, 200, 0.5, 20.0, 2.0, 7)
    assert f.shape == (400, 200)
    assert abs(f.mean()) < 1e-12 and abs(f.std() - 2.0) < 1e-12
    # squared-exponential correlation exp(-r^2 / l^2): e^-1 at one correlation length,
    # e^-4 at two, along both axes (averaged over 24 seeded fields of 100 x 100 l)
    cs = np.zeros((2, 2))
    for seed in range(24):
        g = grf(1000, 1000, 0.5, 5.0, 1.0, 100 + seed)
        for i, lag in enumerate((10, 20)):
            cs[i, 0] += np.mean(g[:-lag] * g[lag:]) / g.var() / 24
            cs[i, 1] += np.mean(g[:, :-lag] * g[:, lag:]) / g.var() / 24
    assert np.all(np.abs(cs[0] - math.exp(-1.0)) < 0.03), cs
    assert np.all(np.abs(cs[1] - math.exp(-4.0)) < 0.03), cs
    # seeded: same seed, same field; other seed, other field
    assert np.array_equal(f, grf(400, 200, 0.5, 20.0, 2.0, 7))
    assert not np.allclose(f, grf(400, 200, 0.5, 20.0, 2.0, 8))


def test_slope_scene(mpm):
    args, grid, mp, attr, bc, lay = mpm.scenes.slope_3d(h=2.0, seed=3)
    xi = np.asarray(mp.ξ, dtype=np.float64)
    fx0, fy0, L, W = lay["footprint"]
    lp = 1.0
    assert mp.np == int(L / lp) * int(W / lp) * int(10.0 / lp)
    # every particle above the bed, the column 10 m deep
    zb = mpm.scenes._bilinear(lay["bed"], float(grid.x1), 2.0, float(grid.y1), 2.0, xi[:, 0], xi[:, 1])
    d = xi[:, 2] - zb
    assert d.min() > 0 and d.max() < 10.0
    # per-cell materials within the requested ranges
    phi = np.degrees(np.asarray(attr.ϕ, dtype=np.float64))
    assert phi.min() >= 25.0 and phi.max() <= 35.0 and len(phi) == lay["nmat"]
    coh = np.asarray(attr.c, dtype=np.float64)
    assert 3e3 < np.median(coh) < 8e3
    # gravity: |g| = 9.8 split over the slope frame at 35 degrees
    g = math.hypot(float(args.gravity), args.ext["gravity_x"])
    assert abs(g - 9.8) < 1e-6
    assert abs(math.degrees(math.atan2(args.ext["gravity_x"], -float(args.gravity))) - 35.0) < 1e-6
    # bed nodes fixed in all directions, none inside the mass
    vz = np.asarray(bc.vz_s_idx)
    assert vz.sum() > 0
    # the K0 phase is damped and ramps gravity; the runout phase is not damped
    assert float(args.ζs) > 0 and float(lay["runout_args"].ζs) == 0.0
    assert abs(float(args.Te) / (lay["k0_steps"] * lay["dt"]) - 1) < 1e-6  # T2 = Float32


def test_default_slope_is_twenty_million_particles(mpm):
    # particle count of the full-size scene from its defaults, without building it
    L, W, depth, lp = 250.0, 125.0, 10.0, 0.25
    assert int(L / lp) * int(W / lp) * int(depth / lp) == 20_000_000


def test_ensemble_layout(mpm):
    args, grid, mp, attr, bc, lay = mpm.scenes.ensemble_2d(n_instances=12, seed=5,
                                                          np_range=(500, 3000), row_width=2.0)
    r = lay["ranges"]
    assert len(r) == 12 and r[0][0] == 0 and r[-1][1] == mp.np
    n = np.array([b - a for a, b in r])
    assert n.min() >= 400 and n.max() <= 3200
    # boxes do not overlap and keep >= 4 empty cells between them
    h = lay["h"]
    boxes = [(ox, oy, ox + L, oy + H + 2 * h) for (ox, oy), L, H in
             zip(lay["origin"], lay["length"], lay["height"])]
    for i in range(len(boxes)):
        for j in range(i + 1, len(boxes)):
            a, b = boxes[i], boxes[j]
            gapx = max(b[0] - a[2], a[0] - b[2])
            gapy = max(b[1] - a[3], a[1] - b[3])
            assert max(gapx, gapy) >= 4 * h - 1e-9
    nid = np.asarray(attr.nid)
    for k, (a, b) in enumerate(r):
        assert np.all(nid[a:b] == k + 1)
