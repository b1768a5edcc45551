"""Tile-shape autotuner for the projector kernels.

Times candidate LDS band heights of the forward projector and tile heights
of the backprojection on scratch copies of the problems' state (so the real
iterates are untouched), through the batch plan's per-phase HIP-event
profile, and applies the fastest shapes to the geometries.  The forward band
only affects the forward phase and the tile height only the backprojection
phase, so the two are tuned independently.
"""
from __future__ import annotations

import json
import os

from .batch import BatchPlan

FWD_BANDS = (4, 8, 12, 16, 24, 32)
ADJ_TILES = (8, 16, 32)


# the forward kernel's LDS limit per work-group (bytes) and its line pitch
# (max(nx, ny) plus 4 zero candidates on each side, rounded up to odd)
LDS_LIMIT = 160 * 1024


def _slice_group(nslices):
    # slice groups of the projector kernels: 3 slices take a padded group of 4
    return 4 if nslices >= 3 else (2 if nslices == 2 else 1)


def _lds_ok(problem, band):
    K = _slice_group(problem.C * problem.Nz)
    pitch = max(problem.nx, problem.ny) + 8
    pitch |= 1
    return band * pitch * K * 4 <= LDS_LIMIT


def signature(problems):
    return ";".join(f"{p.dim}:{p.C}x{p.Nz}x{p.ny}x{p.nx}:{p.na}" for p in problems)


def autotune(problems, iters=3, plan=None, cache=None):
    """Tune tile shapes for ``problems`` (list of TgvProblem sharing nothing
    but the device).  Returns {"fwd_band": b, "adj_tile_y": t, "ms": {...}}
    and leaves every geometry set to the winner.  ``plan`` (a BatchPlan over
    the same problems) is retuned as well.  With ``cache`` (a JSON path) a
    previous result for the same signature is reused."""
    sig = signature(problems)
    if cache and os.path.exists(cache):
        with open(cache) as fh:
            known = json.load(fh)
        if sig in known:
            res = known[sig]
            _apply(problems, res["fwd_band"], res["adj_tile_y"], plan)
            return res
    scratch = [p.clone() for p in problems]
    splan = BatchPlan(scratch)
    geoms = {id(p.geometry): p.geometry for p in problems}
    fwd_ms, adj_ms = {}, {}
    splan.profile(1)  # warm-up
    for band in FWD_BANDS:
        if not all(_lds_ok(p, band) for p in problems):
            continue
        for g in geoms.values():
            g.set_tuning(fwd_band=band, adj_tile_y=16)
        splan.retune()
        fwd_ms[band] = splan.profile(iters)[1] / iters
    # no candidate fits (very wide images): the forward band stays as it is
    # (0 = the kernel's automatic choice)
    best_band = min(fwd_ms, key=fwd_ms.get) if fwd_ms else 0
    for tile in ADJ_TILES:
        for g in geoms.values():
            g.set_tuning(fwd_band=best_band, adj_tile_y=tile)
        splan.retune()
        prof = splan.profile(iters)
        adj_ms[tile] = (prof[2] + prof[3]) / iters
    best_tile = min(adj_ms, key=adj_ms.get)
    del splan, scratch
    _apply(problems, best_band, best_tile, plan)
    res = {"fwd_band": best_band, "adj_tile_y": best_tile,
           "ms": {"fwd": {str(k): v for k, v in fwd_ms.items()},
                  "adj": {str(k): v for k, v in adj_ms.items()}}}
    if cache:
        known = {}
        if os.path.exists(cache):
            with open(cache) as fh:
                known = json.load(fh)
        known[sig] = res
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        with open(cache, "w") as fh:
            json.dump(known, fh, indent=1)
    return res


def _apply(problems, band, tile, plan):
    for p in problems:
        p.geometry.set_tuning(fwd_band=band, adj_tile_y=tile)
    if plan is not None:
        plan.retune()
