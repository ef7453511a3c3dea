"""Oracle pins: particle-to-cell ids, the candidate tables and the stored
per-slot basis, against brute force over all grid nodes."""
import itertools

import numpy as np
import pytest


def _scene(mpm, dim, pts, h=0.1, n=8, basis="uGIMP"):
    if dim == 2:
        args, grid, mp, attr, bc = mpm.scenes.column_collapse_2d(
            h=h, width=2 * h, height=2 * h, box_x=n * h, box_y=n * h, pad=0, basis=basis,
            nsteps=1, dt=1e-4)
        spacing = (mp.dx, mp.dy)
        mk = mpm.UserParticle2D
        mp = mk(NIC=mp.NIC, dx=spacing[0], dy=spacing[1], ξ=pts, ρs=np.full(len(pts), 1000.0))
    else:
        args, grid, mp, attr, bc = mpm.scenes.column_collapse_3d(
            h=h, width=2 * h, height=2 * h, box_xy=n * h, box_z=n * h, pad=0, eps="FP64",
            basis=basis, nsteps=1)
        mp = mpm.UserParticle3D(NIC=mp.NIC, dx=mp.dx, dy=mp.dy, dz=mp.dz, ξ=pts,
                                ρs=np.full(len(pts), 1000.0))
    attr = mpm.UserProperty(nid=np.ones(mp.np), ν=[0.3], Es=[1e6], Gs=[3.8e5], Ks=[8.3e5])
    return args, grid, mp, attr, bc


def _interior_points(rng, grid, dim, n):
    lo = np.array([grid.x1, grid.y1] + ([grid.z1] if dim == 3 else [])) + 2.0 * grid.dx
    hi = np.array([grid.x2, grid.y2] + ([grid.z2] if dim == 3 else [])) - 2.0 * grid.dx
    return rng.uniform(lo, hi, size=(n, dim))


@pytest.mark.parametrize("dim", [2, 3])
def test_p2c_is_the_containing_cell(mpm, orc, dim):
    rng = np.random.default_rng(dim)
    sc = _scene(mpm, dim, np.zeros((1, dim)))
    grid = sc[1]
    pts = _interior_points(rng, grid, dim, 300)
    o = orc.Oracle(*_scene(mpm, dim, pts))
    h = float(grid.dx)
    for p in pts:
        c = o.p2c(*p)
        # reference cell numbering: y row from the bottom (1-based, cld), then x
        # columns, then z layers (fld, 0-based) -- utils_OS.jl:325-328
        cx = int(np.floor((p[0] - grid.x1) / h))
        cy = int(np.ceil((p[1] - grid.y1) / h))
        cz = int(np.floor((p[2] - grid.z1) / h)) if dim == 3 else 0
        want = cy + cx * grid.ncy + cz * grid.ncy * grid.ncx
        assert c == want


@pytest.mark.parametrize("dim", [2, 3])
def test_candidate_table_covers_the_4x4x4_neighbourhood(mpm, orc, dim):
    sc = _scene(mpm, dim, np.zeros((1, dim)) + 0.35)
    grid = sc[1]
    o = orc.Oracle(*sc)
    h = float(grid.dx)
    ncand = 16 if dim == 2 else 64
    xi = np.asarray(grid.ξ)
    org = np.array([grid.x1, grid.y1, getattr(grid, "z1", 0.0)][:dim])
    for rel in ([3.5] * dim, [4.4] * dim, [3.1, 5.2, 4.7][:dim]):
        p = org + np.array(rel) * h
        c = o.p2c(*p)
        nodes = [o.p2n(c, iy) - 1 for iy in range(1, ncand + 1)]
        assert len(set(nodes)) == ncand
        base = np.floor((np.array(p) - np.array([grid.x1, grid.y1, getattr(grid, "z1", 0.0)][:dim])) / h)
        want = set()
        for off in itertools.product(range(-1, 3), repeat=dim):
            q = np.array([grid.x1, grid.y1, getattr(grid, "z1", 0.0)][:dim]) + (base + off) * h
            want.add(int(np.argmin(np.sum((xi - q) ** 2, axis=1))))
        assert set(nodes) == want


@pytest.mark.parametrize("dim", [2, 3])
def test_slots_hold_exactly_the_support(mpm, orc, dim):
    rng = np.random.default_rng(10 + dim)
    sc = _scene(mpm, dim, np.zeros((1, dim)))
    pts = _interior_points(rng, sc[1], dim, 200)
    args, grid, mp, attr, bc = _scene(mpm, dim, pts)
    o = orc.Oracle(args, grid, mp, attr, bc)
    o.initmp()
    xi = np.asarray(grid.ξ)
    h = float(grid.dx)
    lp = float(mp.dx)
    for i, p in enumerate(pts):
        d = np.abs(xi - p)
        support = set(np.nonzero(np.all(d < h + 0.5 * lp, axis=1))[0].tolist())
        slots = o.a["p2n"][i] - 1
        filled = o.a["Nij"][i] != 0
        assert set(slots[filled].tolist()) == support
        # weights of the slots equal the tensor product of 1D weights
        for k in np.nonzero(filled)[0]:
            n = slots[k]
            w = 1.0
            for a in range(dim):
                w *= orc.ugimp_basis(p[a] - xi[n, a], h, lp)[0]
            assert abs(o.a["Nij"][i, k] - w) < 1e-15
        assert abs(o.a["Nij"][i].sum() - 1.0) < 1e-13
        assert np.allclose(o.a["Nij"][i] @ xi[slots.clip(0)], p, atol=1e-13)


def test_linear_slots_are_cell_corners(mpm, orc):
    rng = np.random.default_rng(5)
    sc = _scene(mpm, 3, np.zeros((1, 3)), basis="linear")
    pts = _interior_points(rng, sc[1], 3, 50)
    args, grid, mp, attr, bc = _scene(mpm, 3, pts, basis="linear")
    o = orc.Oracle(args, grid, mp, attr, bc)
    o.initmp()
    xi = np.asarray(grid.ξ)
    for i, p in enumerate(pts):
        nodes = xi[o.a["p2n"][i] - 1]
        assert np.all(np.abs(nodes - p) <= grid.dx + 1e-15)
        assert abs(o.a["Nij"][i].sum() - 1) < 1e-14
        assert np.allclose(o.a["Nij"][i] @ nodes, p, atol=1e-14)
