"""Autotuner for the persistent decode launch.

The persistent decode kernel runs a fixed grid whose workgroups pull tiles
of 32 blocks (four wavefronts of eight) from per-XCD work counters. Two knobs shape it: the grid size
(workgroups, a multiple of 8 so every XCD gets the same count) and the tiles
taken per grab (a run of consecutive tiles per atomic). The tuner times
every candidate on the caller's own resident stream with HIP events, keeps
the fastest, installs it in the context and caches it per (device, shape of
the stream) in a JSON file.
"""
from __future__ import annotations

import json
import os

import numpy as np

from . import batch
from .device import Context, DeviceArray, Event

CACHE = os.environ.get("UTL_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "utlpfor_amd",
                                                        "tune.json"))


def _key(ctx: Context, s: batch.DeviceStream) -> str:
    name, cus = ctx.device_info()
    # blocks and bytes per block (log2 buckets) describe the work
    nb = max(s.nblocks, 1)
    return f"{name}|{cus}|u64={int(s.is64)}|nb={int(np.log2(nb))}|bpb={int(np.log2(max(s.nbytes / nb, 1)))}"


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(d):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        with open(CACHE, "w") as f:
            json.dump(d, f, indent=1)
    except OSError:
        pass


def time_decode(ctx: Context, s: batch.DeviceStream, out: DeviceArray, status: DeviceArray, persistent: bool,
                reps: int = 5) -> float:
    """Average decode time in ms over reps launches (after one warm-up)."""
    batch.decode(ctx, s, out=out, status=status, persistent=persistent)
    e0, e1 = Event(), Event()
    e0.record(ctx)
    for _ in range(reps):
        batch.decode(ctx, s, out=out, status=status, persistent=persistent)
    e1.record(ctx)
    ctx.sync()
    return e0.elapsed_ms(e1) / reps


def tune_decode(ctx: Context, s: batch.DeviceStream, out: DeviceArray | None = None,
                status: DeviceArray | None = None, use_cache: bool = True, reps: int = 5) -> dict:
    """Chooses the persistent decode's grid and grab size for stream s and
    installs it in ctx. Returns {"grid_wgs", "tiles_per_grab", "ms",
    "flat_ms", "candidates"}."""
    key = _key(ctx, s)
    cache = _load() if use_cache else {}
    if use_cache and key in cache:
        best = cache[key]
        batch.set_tune(ctx, 4, best["grid_wgs"], best["tiles_per_grab"])
        return dict(best, cached=True)
    own_out = out is None
    if out is None:
        out = DeviceArray(max(s.nvalues, 1), np.uint64 if s.is64 else np.uint32, ctx)
    if status is None:
        status = DeviceArray(max(s.nblocks, 1), np.int32, ctx)
    _, cus = ctx.device_info()
    results = []
    ntiles = (s.nblocks + 31) // 32
    for mult in (1, 2, 3, 4, 6, 8):
        grid = max(8, (cus * mult + 7) // 8 * 8)
        if grid > 8 * max(1, (ntiles + 7) // 8) and mult > 1:
            continue
        for grab in (1, 2, 4, 8, 16):
            batch.set_tune(ctx, 4, grid, grab)
            ms = time_decode(ctx, s, out, status, True, reps)
            results.append({"grid_wgs": grid, "tiles_per_grab": grab, "ms": ms})
    best = min(results, key=lambda r: r["ms"])
    flat_ms = time_decode(ctx, s, out, status, False, reps)
    batch.set_tune(ctx, 4, best["grid_wgs"], best["tiles_per_grab"])
    rec = {"grid_wgs": best["grid_wgs"], "tiles_per_grab": best["tiles_per_grab"], "ms": best["ms"],
           "flat_ms": flat_ms}
    if use_cache:
        cache[key] = rec
        _save(cache)
    if own_out:
        out.free()
    return dict(rec, candidates=results, cached=False)
