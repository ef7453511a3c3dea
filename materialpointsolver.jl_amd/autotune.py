"""Autotuner for the step's runtime knobs: the sort interval, whether the LDS
tile kernels are used, and which tile-scatter form runs (P2G and the MUSL
re-map):

* ``"owned"``: each lane owns one stencil base of a tile and sums its dual
  cell's particles in registers (3D P2G: mass and momentum, then force; two
  passes, ~190 VGPRs, two wavefronts per SIMD);
* ``"owned_layers"``: the same with one z layer of the stencil per pass (six
  passes, ~130 VGPRs, three wavefronts per SIMD; more particle reloads);
* ``"rounds"``: rank rounds of distinct bases adding into two LDS windows per
  wavefront (12.6 KB of LDS per wavefront);
* ``"rounds_layered"``: rank rounds of one group adding into three layer
  windows, one per stencil z offset (the same 12.6 KB): three independent
  chains of LDS updates per round (FP32 3D uGIMP; elsewhere ``"rounds"``);
* ``"pairs"``: pair rounds: a lane takes two particles of one stencil base
  and does one LDS read-add-write per slot for both (half the window traffic
  of ``"rounds"``; FP32 3D uGIMP, elsewhere ``"owned"``).

The tile shape (4x4x4 bases in 3D, 8x8 in 2D) and the window size (no slack)
are compile-time.  The MUSL re-map's momentum scatter can take its own form
(``remap_scatter``; by default the base-owned one in FP32 3D uGIMP).

``autotune`` times candidate settings on the live model, in place: each
candidate runs ``warm`` untimed steps and ``steps`` timed steps of the actual
simulation (so the state advances through every trial), and the model is left
on the fastest setting.  Use it at the start of a run.
"""
from __future__ import annotations

import time

SCATTER_FORMS = ("owned", "owned_layers", "rounds", "rounds_layered", "pairs")


def _set_scatter(model, sc):
    """A candidate scatter form: "form" (P2G #1's form, the re-map's automatic)
    or (P2G #1's form, the re-map's form)."""
    if isinstance(sc, tuple):
        model.set_option("scatter", sc[0])
        model.set_option("remap_scatter", sc[1])
    else:
        model.set_option("scatter", sc)
        model.set_option("remap_scatter", None)


def autotune(model, sargs, dt, Ti, *, sort_intervals=(5, 10, 20, 40), tiles=(True, False),
             scatters=("rounds",), steps=40, warm=10, graph=False, clock=time.perf_counter):
    """Returns (best, table, Ti): best = dict(sort_interval=..., tiles=...,
    scatter=...), table = list of (sort_interval, tiles, scatter, seconds per
    step), Ti = time after the trials.  A scatter candidate is a form name
    (P2G #1's form; the re-map scatter's is automatic) or a pair (P2G #1's
    form, the re-map scatter's form).  The scatter form only matters with
    tiles; without tiles it is recorded as None."""
    table = []
    for t in tiles:
        for sc in (scatters if t else (None,)):
            for k in sort_intervals:
                model.set_option("tiles", t)
                if sc is not None:
                    _set_scatter(model, sc)
                model.set_option("sort_interval", k)
                Ti = model.run(sargs, dt, Ti, warm, graph=graph)
                model.sync()
                t0 = clock()
                Ti = model.run(sargs, dt, Ti, steps, graph=graph)
                model.sync()
                table.append((k, bool(t), sc, (clock() - t0) / steps))
    k, t, sc, _ = min(table, key=lambda r: r[3])
    model.set_option("tiles", t)
    if sc is not None:
        _set_scatter(model, sc)
    model.set_option("sort_interval", k)
    return dict(sort_interval=k, tiles=t, scatter=sc), table, Ti
