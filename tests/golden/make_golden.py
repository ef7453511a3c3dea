"""Write the oracle golden vectors (run from the repository root).

Cases: a 2D block of 8x8 particles and a 3D block of 4x4x4 particles (Drucker-
Prager collapse geometry, FP64), state after 1 and after 100 steps.  Each npz
holds the inputs (initial positions, dt) and the device2host! field set.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def cases(mpm):
    yield "golden_2d_block", mpm.scenes.column_collapse_2d(
        h=0.01, width=0.04, height=0.04, box_x=0.12, box_y=0.06, dt=5e-5, nsteps=100)
    yield "golden_3d_block", mpm.scenes.column_collapse_3d(
        h=0.02, width=0.04, height=0.04, box_xy=0.16, box_z=0.08, eps="FP64", nsteps=100)


def main():
    import __graft_entry__ as ge
    import oracle

    mpm = ge.load_package()
    out_dir = os.path.dirname(os.path.abspath(__file__))
    for name, sc in cases(mpm):
        args, grid, mp, attr, bc = sc
        dt = float(args.ΔT)
        o = oracle.Oracle(*sc)
        o.run(dt, 0.0, 1)
        s1 = o.state()
        o.run(dt, dt, 99)
        s100 = o.state()
        data = {"x0": np.asarray(mp.ξ, dtype=np.float64), "dt": dt}
        for k, v in s1.items():
            data["step1_" + k] = v
        for k, v in s100.items():
            data["step100_" + k] = v
        np.savez_compressed(os.path.join(out_dir, name + ".npz"), **data)
        print(name, mp.np)


if __name__ == "__main__":
    main()
