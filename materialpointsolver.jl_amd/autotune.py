"""Autotuner for the step's runtime knobs: the sort interval and whether the
LDS tile kernels are used (the tile shape and window size are compile-time:
4x4x4 bases in 3D, 8x8 in 2D, windows without slack, two per scatter wavefront).

``autotune`` times candidate settings on the live model, in place: each
candidate runs ``warm`` untimed steps and ``steps`` timed steps of the actual
simulation (so the state advances through every trial), and the model is left
on the fastest setting.  Use it at the start of a run.
"""
from __future__ import annotations

import time


def autotune(model, sargs, dt, Ti, *, sort_intervals=(5, 10, 20, 40), tiles=(True, False),
             steps=40, warm=10, graph=False, clock=time.perf_counter):
    """Returns (best, table, Ti): best = dict(sort_interval=..., tiles=...), table =
    list of (sort_interval, tiles, seconds per step), Ti = time after the trials."""
    table = []
    for t in tiles:
        for k in sort_intervals:
            model.set_option("tiles", t)
            model.set_option("sort_interval", k)
            Ti = model.run(sargs, dt, Ti, warm, graph=graph)
            model.sync()
            t0 = clock()
            Ti = model.run(sargs, dt, Ti, steps, graph=graph)
            model.sync()
            table.append((k, bool(t), (clock() - t0) / steps))
    k, t, _ = min(table, key=lambda r: r[2])
    model.set_option("tiles", t)
    model.set_option("sort_interval", k)
    return dict(sort_interval=k, tiles=t), table, Ti
