"""The oracle against its committed golden vectors (regression pin), and -- on
the GPU -- the HIP FP64 path against the same vectors."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402

KEYS = ("xi", "vs", "sigma", "sigma_m", "omega", "eps_q", "eps_k", "eps_v", "eps_s", "ms")


def _cases(mpm):
    return list(make_golden.cases(mpm))


def _load(name):
    return np.load(os.path.join(HERE, "golden", name + ".npz"))


@pytest.mark.parametrize("idx", [0, 1])
def test_oracle_reproduces_golden(mpm, orc, idx):
    name, sc = _cases(mpm)[idx]
    g = _load(name)
    assert np.array_equal(np.asarray(sc[2].ξ, dtype=np.float64), g["x0"])
    o = orc.Oracle(*sc)
    dt = float(g["dt"])
    o.run(dt, 0.0, 1)
    s1 = o.state()
    o.run(dt, dt, 99)
    s100 = o.state()
    for k in KEYS:
        assert np.allclose(s1[k], g["step1_" + k], rtol=1e-12, atol=1e-13), k
        assert np.allclose(s100[k], g["step100_" + k], rtol=1e-10, atol=1e-12), k


@pytest.mark.gpu
@pytest.mark.parametrize("idx", [0, 1])
def test_hip_fp64_matches_golden(mpm, idx):
    name, sc = _cases(mpm)[idx]
    args, grid, mp, attr, bc = sc
    print(f"case: golden {name} np={mp.np}", flush=True)
    g = _load(name)
    model = mpm.HIPModel(grid, mp, attr, bc, sort_interval=0)
    try:
        sargs = model.step_args(args)
        dt = float(g["dt"])
        model.run(sargs, dt, 0.0, 1)
        s1 = model.download()
        model.run(sargs, dt, dt, 99)
        s100 = model.download()
    finally:
        model.close()
    for k in KEYS:
        ref1, ref100 = g["step1_" + k], g["step100_" + k]
        # relative L2 with an absolute floor for fields that are zero up to
        # rounding (e.g. ϵk ~ 1e-21 at the cohesionless apex)
        floor = 1e-14 * np.sqrt(ref1.size)
        assert np.linalg.norm(s1[k] - ref1) <= 1e-10 * np.linalg.norm(ref1) + floor, k
        assert np.linalg.norm(s100[k] - ref100) <= 1e-9 * np.linalg.norm(ref100) + floor, k
