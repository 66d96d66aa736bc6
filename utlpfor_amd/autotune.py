"""Autotuner for the batched decode launch.

The decode kernels come in launch variants (utl_set_tune). Four knobs shape them:
  arena_bytes     the LDS bytes a wavefront stages its eight uint32 blocks in
                  (2560, 3328, 4352): a larger arena holds the eight blocks of
                  a wave in one staging round on streams of wide blocks, but
                  fewer waves fit a CU;
  waves_per_wg    wavefronts (of eight blocks each) per workgroup (2, 4),
                  which sets how finely LDS is divided among workgroups;
  grid_wgs        workgroups of the persistent form's fixed grid (a multiple
                  of 8, so every XCD gets the same count);
  tiles_per_grab  tiles the persistent form takes per work-queue atomic.
The tuner times every flat variant (arena x waves) and every persistent
variant (arena x waves x grid x grab) on the caller's own resident stream with
HIP events, times the fastest few again in turn (median), keeps the fastest,
installs its variant in the context and caches it per (device, shape of the
stream) in a JSON file. Each variant gives identical results, so the choice is
a matter of time only.
"""
from __future__ import annotations

import json
import os

import numpy as np

from . import batch
from .device import Context, DeviceArray, Event

FINALISTS = 6  # candidates timed again, interleaved
ROUNDS = 3

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
                reps: int = 5, wave: DeviceArray | None = None) -> float:
    """Average decode time in ms over reps launches (after one warm-up): the
    block form, or the meta form over the wave index `wave`."""
    if wave is None:
        def run():
            batch.decode(ctx, s, out=out, status=status, persistent=persistent)
    else:
        def run():
            batch.decode_meta(ctx, s, wave, out=out, status=status, persistent=persistent)
    run()
    e0, e1 = Event(), Event()
    e0.record(ctx)
    for _ in range(reps):
        run()
    e1.record(ctx)
    ctx.sync()
    return e0.elapsed_ms(e1) / reps


def candidates(ctx: Context, s: batch.DeviceStream):
    """Every launch variant the tuner times for stream s: dicts of the four
    knobs plus "persistent". uint64 streams have one arena size."""
    _, cus = ctx.device_info()
    arenas = (batch.ARENAS32[0],) if s.is64 else batch.ARENAS32
    out = []
    for arena in arenas:
        for waves in batch.WAVES_PER_WG:
            out.append({"arena_bytes": arena, "waves_per_wg": waves, "grid_wgs": cus, "tiles_per_grab": 1,
                        "persistent": False})
            ntiles = (s.nblocks + 8 * waves - 1) // (8 * waves)
            for mult in (1, 2, 3, 4, 6, 8):
                grid = max(8, (cus * mult * (4 // waves) + 7) // 8 * 8)
                if grid > 8 * max(1, (ntiles + 7) // 8) and mult > 1:
                    continue
                for grab in (1, 2, 4, 8, 16):
                    out.append({"arena_bytes": arena, "waves_per_wg": waves, "grid_wgs": grid,
                                "tiles_per_grab": grab, "persistent": True})
    return out


def _install(ctx: Context, c: dict) -> None:
    batch.set_tune(ctx, c["arena_bytes"], c["waves_per_wg"], c["grid_wgs"], c["tiles_per_grab"])


def tune_decode(ctx: Context, s: batch.DeviceStream, out: DeviceArray | None = None,
                status: DeviceArray | None = None, use_cache: bool = True, reps: int = 5,
                meta_form: bool = False) -> dict:
    """Chooses the decode's launch variant for stream s and installs it in
    ctx, timing the block form (utl_decode_u32/u64 over the block index) or,
    with meta_form, the meta form (utl_decode_meta over the wave index).
    Returns {"arena_bytes", "waves_per_wg", "grid_wgs", "tiles_per_grab",
    "persistent", "ms", "flat_ms", "persistent_ms", "candidates", "cached"}:
    the winner's knobs, its time, the best flat and best persistent times."""
    key = _key(ctx, s) + ("|meta" if meta_form else "")
    cache = _load() if use_cache else {}
    if use_cache and key in cache:
        best = cache[key]
        _install(ctx, best)
        return dict(best, cached=True)
    own_out = out is None
    own_status = status is None
    if own_out:
        out = DeviceArray(max(s.nvalues, 1), np.uint64 if s.is64 else np.uint32, ctx)
    if own_status:
        status = DeviceArray(max(s.nblocks, 1), np.int32, ctx)
    results = []
    wave = batch.build_wave_index(ctx, s) if meta_form else None
    try:
        for c in candidates(ctx, s):
            _install(ctx, c)
            results.append(dict(c, ms=time_decode(ctx, s, out, status, c["persistent"], reps, wave)))
        # Launch times drift by a few per cent over a sweep (clocks of a card
        # held at full HBM bandwidth): the best few are timed again, in turn,
        # and each keeps the median of its times.
        top = sorted(range(len(results)), key=lambda i: results[i]["ms"])[:FINALISTS]
        times = {i: [results[i]["ms"]] for i in top}
        for _ in range(ROUNDS):
            for i in top:
                _install(ctx, results[i])
                times[i].append(time_decode(ctx, s, out, status, results[i]["persistent"], reps, wave))
        for i in top:
            results[i]["ms"] = float(np.median(times[i]))
    finally:
        if own_out:
            out.free()
        if own_status:
            status.free()
        if wave is not None:
            wave.free()
    best = min(results, key=lambda r: r["ms"])
    flat_ms = min(r["ms"] for r in results if not r["persistent"])
    pers_ms = min(r["ms"] for r in results if r["persistent"])
    _install(ctx, best)
    rec = {k: best[k] for k in ("arena_bytes", "waves_per_wg", "grid_wgs", "tiles_per_grab", "persistent", "ms")}
    rec.update(flat_ms=flat_ms, persistent_ms=pers_ms)
    if use_cache:
        cache[key] = rec
        _save(cache)
    return dict(rec, candidates=results, cached=False)


# ---- encode form: slots + compaction, direct writes, or size / scan / write ----

def _encode_key(ctx: Context, nblocks: int, nvalues: int) -> str:
    name, cus = ctx.device_info()
    return f"encode|{name}|{cus}|nb={int(np.log2(max(nblocks, 1)))}|vpb={int(np.log2(max(nvalues / max(nblocks, 1), 1)))}"


def cached_encode_method(ctx: Context, nblocks: int, nvalues: int):
    """The encode form cached for this shape of work, or None."""
    rec = _load().get(_encode_key(ctx, nblocks, nvalues))
    return rec["method"] if rec else None


def tune_encode(ctx: Context, values: DeviceArray, stream_lengths=None, stream_flags=None, flag: int = 0,
                reps: int = 3, use_cache: bool = True) -> dict:
    """Times every uint32 encode form on the caller's resident values (the
    whole batch.encode: job table, kernels, index scan, allocations), in
    turn, keeps the fastest (median of reps) and caches the choice for
    batch.encode(method="auto"). Returns {"method", "<form>_ms" for each}."""
    import time
    times = {m: [] for m in batch.ENCODE_METHODS}
    nblocks = 0
    for _ in range(reps + 1):
        for m in batch.ENCODE_METHODS:
            ctx.sync()
            t0 = time.perf_counter()
            s = batch.encode(ctx, values, flag=flag, stream_lengths=stream_lengths, stream_flags=stream_flags,
                             method=m)
            times[m].append((time.perf_counter() - t0) * 1e3)
            nblocks = s.nblocks
            s.free()
    med = {m: float(np.median(t[1:])) for m, t in times.items()}  # the first round warms up
    rec = {"method": min(med, key=med.get)}
    rec.update({f"{m}_ms": t for m, t in med.items()})
    if use_cache:
        cache = _load()
        cache[_encode_key(ctx, nblocks, values.n)] = rec
        _save(cache)
    return rec
