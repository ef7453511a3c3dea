"""mpm_count_active_nodes_noalloc (node bitmap and counter in a file-scope
device array, no allocation) against mpm_count_active_nodes (allocates its
mark buffer per call), through the C ABI: two calls in a row, the second on a
smaller grid, then the allocating count of each; a grid above the array's
2**25-node maximum is rejected with MPM_E_RANGE."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _count(lib, ctx, fn):
    out = np.zeros(1, dtype=np.int64)
    rc = getattr(lib, fn)(ctx, out.ctypes.data_as(C.POINTER(C.c_int64)))
    return rc, int(out[0])


def _model(mpm, scene, steps):
    args, grid, mp, attr, bc = scene
    m = mpm.HIPModel(grid, mp, attr, bc, sort_interval=5)
    m.run(m.step_args(args), float(args.ΔT), 0.0, steps)
    m.sync()
    return m, grid


def test_noalloc_count_matches_the_allocating_count(mpm):
    lib = mpm._capi.load()
    big = mpm.scenes.column_collapse_3d(h=0.01, eps="FP32", nsteps=40)
    small = mpm.scenes.column_collapse_2d(h=0.01, width=0.2, height=0.1, box_x=0.5, box_y=0.12,
                                          dt=5e-5, nsteps=40, eps="FP64")
    mb, gb = _model(mpm, big, 40)
    ms, gs = _model(mpm, small, 40)
    try:
        print(f"case: noalloc count, grids of {gb.ni} then {gs.ni} nodes", flush=True)
        rb, nb = _count(lib, mb.ctx, "mpm_count_active_nodes_noalloc")
        rs, ns = _count(lib, ms.ctx, "mpm_count_active_nodes_noalloc")
        assert rb == 0 and rs == 0
        ab = _count(lib, mb.ctx, "mpm_count_active_nodes")
        as_ = _count(lib, ms.ctx, "mpm_count_active_nodes")
        print(f"3D: {nb} (noalloc) {ab[1]} (allocating); 2D: {ns} {as_[1]}", flush=True)
        assert ab == (0, nb) and as_ == (0, ns)
        assert 0 < ns < nb
        # the smaller grid again after the larger one: nothing left over
        assert _count(lib, mb.ctx, "mpm_count_active_nodes_noalloc") == (0, nb)
        assert _count(lib, ms.ctx, "mpm_count_active_nodes_noalloc") == (0, ns)
        assert mb.count_active_nodes(noalloc=True) == nb
    finally:
        mb.close()
        ms.close()


def test_noalloc_count_rejects_grids_above_its_maximum(mpm):
    lib = mpm._capi.load()
    # a 2D grid of 6,002 x 6,002 cells (one ghost cell per side): 36,036,009 nodes > 2**25
    sc = mpm.scenes.column_collapse_2d(h=1e-4, width=0.002, height=0.001, box_x=0.6, box_y=0.6,
                                       dt=1e-6, nsteps=1, eps="FP32", pad=1)
    args, grid, mp, attr, bc = sc
    assert grid.ni > 2 ** 25
    m = mpm.HIPModel(grid, mp, attr, bc)
    try:
        rc, _ = _count(lib, m.ctx, "mpm_count_active_nodes_noalloc")
        msg = lib.mpm_last_error(m.ctx).decode()
        print(f"case: {grid.ni} nodes: rc {rc}, {msg}", flush=True)
        assert rc == -6 and "maximum" in msg
        rc, n = _count(lib, m.ctx, "mpm_count_active_nodes")  # the allocating form still works
        assert rc == 0 and n > 0
    finally:
        m.close()
