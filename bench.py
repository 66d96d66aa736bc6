#!/usr/bin/env python3
"""Benchmark of the UTL-PFOR decode hot path on one MI355X.

Headline workload: 200,000 independent uint32 streams, lengths log-uniform
in [1, 200,000] values (partial final blocks), flags uniform over {0, Delta,
NoFOR, NoPatch, Delta|NoFOR}, values uniform in [0, 2^b) (b uniform 0..32) or
signed random walks with steps U[-2^10, 2^10], packed by the GPU encoder
into 25.7 M concatenated 128-value blocks (7.8 GB compressed, 13.1 GB
decoded) resident in HBM. One step = one launch of a HIP graph holding the
device prefix scan of the per-block lengths and one decode launch of every
block of every stream, in the faster of two graph forms on the same stream:
the wave index (the byte and value offsets of every eighth block; each
wavefront scans its eight blocks' lengths itself) + the meta-form decode,
or the block index + the block-form decode, each in the launch variant the
autotuner picks for it. value = decoded GB/s of the whole step.

Also reported: the decode kernel's roofline from HIP events on its own
stream, every launch variant the tuner timed, the persistent work-queue
form, the plain per-call drop-in (one C-ABI call per block with host
staging, p50/p99 latency), 10 M random Gets over the whole stream, the
64 Mi delta-postings stream as a second batched form, a copy kernel (the
practical HBM ceiling) and the oracle (the C restatement of the
reference's scalar path) timed on the host's cores over a bounded sample of
the same blocks.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "decoded GB/s & %HBM peak, batched vs plain per-call; p50/p99 latency, bit-exact"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_info() -> dict:
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity_cores": len(os.sched_getaffinity(0)), "cpu_model": model}


def cpu_baseline(stream_bytes, boff, ooff, nvalues, what: str, runs: int = 5, min_run_s: float = 1.0):
    """The oracle decoding blocks k = stream_bytes[boff[k]:boff[k+1]] into
    out[ooff[k]:] with one thread per host core this process may use: one
    warm-up, then `runs` timed runs of at least min_run_s each (repeating the
    decode within a run); the median rate."""
    from oracle import oracle
    info = cpu_info()
    threads = info["affinity_cores"]
    oracle.unpack_blocks(stream_bytes, boff, ooff, nvalues, False, nthreads=threads)  # warm-up
    rates = []
    for _ in range(runs):
        reps, t0 = 0, time.perf_counter()
        while True:
            oracle.unpack_blocks(stream_bytes, boff, ooff, nvalues, False, nthreads=threads)
            reps += 1
            el = time.perf_counter() - t0
            if el >= min_run_s:
                break
        rates.append(reps * nvalues * 4 / el / 1e9)
    rec = {"value": round(float(np.median(rates)), 3), "unit": "GB/s decoded", "cores": threads, "kind": "port",
           "sample": f"{what}: {boff.size - 1} blocks ({int(boff[-1])} compressed bytes, {nvalues} values) decoded "
                     f"by the C restatement of the reference scalar path on {threads} threads; median of {runs} "
                     f"runs of >= {min_run_s:.0f} s after a warm-up",
           "runs_GBps": [round(r, 3) for r in rates]}
    rec.update(info)
    return rec


def walk_cpu_baseline(stream_bytes, target_s: float = 3.0):
    """The oracle's serial BlockLength walk (example_test.go:171-187) over
    the same stream, one thread (the walk is serial)."""
    from oracle import oracle
    oracle.walk_stream(stream_bytes)  # warm-up
    reps, t0 = 0, time.perf_counter()
    while True:
        st, _, _, _ = oracle.walk_stream(stream_bytes)
        reps += 1
        el = time.perf_counter() - t0
        if el >= target_s:
            break
    return {"value": round(el / reps * 1e3, 3), "unit": "ms per walk", "cores": 1, "kind": "port",
            "sample": f"the whole {stream_bytes.size}-byte stream ({st.size} blocks) walked {reps}x by the C "
                      f"restatement of the reference's BlockLength walk on one thread"}


def per_call_form(ctx, blobs, calls: int = 2000):
    """One C-ABI UnpackUint32 call per block (host bytes in, host values out)."""
    from utlpfor_amd import codec
    for b in blobs[:50]:
        codec.UnpackUint32(b, ctx=ctx)
    lat = np.empty(len(blobs))
    t0 = time.perf_counter()
    for i, b in enumerate(blobs):
        t = time.perf_counter()
        _, _, err = codec.UnpackUint32(b, ctx=ctx)
        lat[i] = time.perf_counter() - t
        assert err is None
    el = time.perf_counter() - t0
    return {"GB/s": round(len(blobs) * 512 / el / 1e9, 6), "p50_us": round(float(np.percentile(lat, 50)) * 1e6, 1),
            "p99_us": round(float(np.percentile(lat, 99)) * 1e6, 1), "calls": len(blobs)}


def _events_time(ctx, fn, reps: int) -> float:
    from utlpfor_amd.device import Event
    e0, e1 = Event(), Event()
    e0.record(ctx)
    for _ in range(reps):
        fn()
    e1.record(ctx)
    ctx.sync()
    return e0.elapsed_ms(e1) / reps


def _timed_steps(ctx, step, steps: int):
    """steps launches of `step` with an event after each; returns (wall ms
    per step, per-step event times)."""
    from utlpfor_amd.device import Event
    ev = [Event() for _ in range(steps + 1)]
    ctx.sync()
    t0 = time.perf_counter()
    ev[0].record(ctx)
    for i in range(steps):
        step()
        ev[i + 1].record(ctx)
    ctx.sync()
    wall = time.perf_counter() - t0
    return wall * 1e3 / steps, [ev[i].elapsed_ms(ev[i + 1]) for i in range(steps)]


def _traffic(name: str, nblocks: int, kernel: str | None = None):
    """HBM bytes per decode dispatch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes over this workload (copy-kernel calibrated), when they
    were taken of the same kernel variant."""
    tpath = os.path.join(ROOT, "profiles", f"{name}_traffic.json")
    if not os.path.exists(tpath):
        return None
    with open(tpath) as f:
        t = json.load(f)
    if t.get("blocks") != nblocks or (kernel and not t.get("kernel", "").startswith(kernel.rstrip(">"))):
        return None
    return t.get("traffic_bytes")


def _copy_gbps(ctx, nbytes: int) -> float:
    from utlpfor_amd.device import DeviceArray, dev_copy
    src = DeviceArray(nbytes, np.uint8, ctx)
    dst = DeviceArray(nbytes, np.uint8, ctx)
    dev_copy(ctx, dst, src, nbytes)
    ms = _events_time(ctx, lambda: dev_copy(ctx, dst, src, nbytes), 10)
    src.free()
    dst.free()
    return round(2 * nbytes / (ms * 1e-3) / 1e9, 1)


def _stream_ceiling(ctx, src, read_bytes: int, dst, write_bytes: int) -> float:
    """GB/s of utl_dev_stream reading read_bytes of src and writing
    write_bytes of dst (the compressed and decoded bytes of a decode): the
    best grid of 1-16 workgroups per CU, each timed as the median of 3 runs
    of 5 launches."""
    from utlpfor_amd.device import dev_stream
    _, cus = ctx.device_info()
    rd = read_bytes // 16 * 16
    wr = write_bytes // 16 * 16
    best = float("inf")
    for per_cu in (1, 2, 4, 8, 16):
        g = per_cu * cus
        dev_stream(ctx, dst, wr, src, rd, g)
        ms = float(np.median([_events_time(ctx, lambda: dev_stream(ctx, dst, wr, src, rd, g), 5) for _ in range(3)]))
        best = min(best, ms)
    return round((rd + wr) / (best * 1e-3) / 1e9, 1)


def _copy_calibration(ctx) -> None:
    """A plain copy kernel of known size (256 MiB read + 256 MiB written),
    whose FETCH_SIZE/WRITE_SIZE scale the decode kernel's counters."""
    from utlpfor_amd.device import DeviceArray, dev_copy
    src = DeviceArray(256 << 20, np.uint8, ctx)
    dst = DeviceArray(256 << 20, np.uint8, ctx)
    for _ in range(3):
        dev_copy(ctx, dst, src, 256 << 20)
    ctx.sync()
    src.free()
    dst.free()


def _kernel_name(t: dict, meta: bool = False, persistent: bool = False) -> str:
    k = "k_decode_persistent" if persistent else "k_decode"
    return f"{k}<false, {str(meta).lower()}, {t['arena_bytes']}, {t['waves_per_wg']}>"


def _result(value, steps, warmup, ms_per_step, workload: dict, roofline: dict) -> dict:
    return {"metric": METRIC, "value": round(value, 2), "unit": "GB/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u32", "data": "synthetic", "config": workload, "roofline": roofline}


def _parse_tune(arg: str, cus: int) -> dict:
    """ARENA:WAVES (flat) or ARENA:WAVES:GRID:GRAB (persistent)."""
    f = [int(x) for x in arg.split(":")]
    if len(f) == 4:
        return {"arena_bytes": f[0], "waves_per_wg": f[1], "grid_wgs": f[2], "tiles_per_grab": f[3],
                "persistent": True}
    return {"arena_bytes": f[0], "waves_per_wg": f[1], "grid_wgs": cus * 2, "tiles_per_grab": 1, "persistent": False}


def headline_uneven(ctx, args) -> dict:
    """200,000 uneven streams, one graph launch (index scan + decode) per step."""
    from utlpfor_amd import _lib, autotune, batch, runbench, workloads
    from utlpfor_amd.device import DeviceArray, check
    lib = _lib.load()
    _, cus = ctx.device_info()
    t0 = time.time()
    v, lens, flags = workloads.uneven_streams(args.streams, seed=3)
    gen_s = time.time() - t0
    n = v.size
    log(f"{args.streams} streams, {n} values generated in {gen_s:.1f} s; GPU encode")
    s = batch.encode(ctx, v, stream_lengths=lens, stream_flags=flags)
    nb, b_in, d_out = s.nblocks, s.nbytes, n * 4
    log(f"{nb} blocks, {b_in} compressed bytes ({b_in / n:.3f} B/value)")
    out = DeviceArray(n, np.uint32, ctx)
    status = DeviceArray(nb, np.int32, ctx)
    # the two graph forms: the wave index (one (byte, value) offset pair per
    # eight blocks) + decode from the meta words, or the block index
    # (block_off / out_off) + decode from it; each in the launch variant
    # tuned for it
    forms_wanted = ("meta", "block") if args.form == "auto" else (args.form,)
    graphs, tunes = {}, {}
    for form in forms_wanted:
        meta_form = form == "meta"
        if args.tune:
            tune = _parse_tune(args.tune, cus)
            batch.set_tune(ctx, tune["arena_bytes"], tune["waves_per_wg"], tune["grid_wgs"], tune["tiles_per_grab"])
            tsum = dict(tune, fixed=True)
        else:
            log(f"autotuning the decode launch ({form} form)")
            tune = autotune.tune_decode(ctx, s, out, status, use_cache=False, reps=3, meta_form=meta_form)
            tsum = runbench.tune_summary(tune)
        log(f"{form} form launch variant: {tsum}")
        tunes[form] = (tune, tsum, batch.get_tune(ctx))
        graphs[form] = batch.DecodeGraph(ctx, s, out, status, persistent=bool(tune["persistent"]),
                                         block_index=not meta_form)
    # the faster graph (interleaved short runs of each), then the timed steps on it
    step_trial = {}
    for _ in range(3):
        for form, g in graphs.items():
            g.launch()
            ctx.sync()
            step_trial.setdefault(form, []).append(_events_time(ctx, g.launch, 5))
    step_trial = {f: float(np.median(t)) for f, t in step_trial.items()}
    form = min(step_trial, key=step_trial.get)
    tune, tsum, installed = tunes[form]
    batch.set_tune(ctx, installed["arena_bytes"], installed["waves_per_wg"], installed["grid_wgs"],
                   installed["tiles_per_grab"])
    graph = graphs[form]
    log(f"graph form: {form} (trial steps, ms: {step_trial})")
    wave = batch.build_wave_index(ctx, s)
    # one step, then the bit-exact round trip of every decoded value, and
    # every block OK
    graph.launch()
    ctx.sync()
    exact = bool(np.all(status.to_numpy() == 0))
    chunk = 1 << 28
    for a in range(0, n, chunk):
        m = min(chunk, n - a)
        got = np.empty(m, np.uint32)
        check(lib.utl_memcpy_d2h(ctx.handle, got.ctypes.data, out.ptr + 4 * a, 4 * m), "d2h")
        exact &= bool(np.array_equal(got, v[a:a + m]))
    log(f"round trip bit-exact: {exact}")
    # random access over the whole 7.8 GB stream: 10 M uniformly random value
    # positions, their expected values taken before the host copy goes
    grng = np.random.default_rng(11)
    get_pos = grng.integers(0, n, 10_000_000).astype(np.int64)
    get_want = v[get_pos]
    del v
    # the warm-up steps right before the timed ones (the check above leaves
    # the GPU idle for seconds)
    for _ in range(args.warmup):
        graph.launch()
    ctx.sync()
    ms_per_step, step_ms = _timed_steps(ctx, graph.launch, args.steps)
    value = d_out / (ms_per_step * 1e-3) / 1e9
    # dominant kernel alone: the graph's decode launch between two events
    persistent = bool(tune["persistent"])
    if form == "meta":
        decode_fn = lambda: batch.decode_meta(ctx, s, wave, out=out, status=status, persistent=persistent)  # noqa: E731
        # compressed bytes + decoded bytes + meta words + wave index entries
        algo_bytes = b_in + d_out + 4 * nb + 16 * ((nb + 7) // 8)
    else:
        decode_fn = lambda: batch.decode(ctx, s, out=out, status=status, persistent=persistent)  # noqa: E731
        # compressed bytes + decoded bytes + block index entries
        algo_bytes = b_in + d_out + 16 * nb
    dec_ms = _events_time(ctx, decode_fn, max(args.steps, 10))
    achieved = algo_bytes / (dec_ms * 1e-3) / 1e9
    # the practical ceiling for this traffic: a plain streaming kernel that
    # reads the compressed bytes and writes as many bytes as the decode
    ceiling = _stream_ceiling(ctx, s.data, b_in, out, d_out)
    t = batch.get_tune(ctx)
    kname = _kernel_name(t, meta=form == "meta", persistent=persistent)
    workload = {"workload": f"{args.streams} uneven uint32 streams, lengths logU[1,200000], flags U{{0, Delta, NoFOR, "
                            "NoPatch, Delta|NoFOR}, values U[0,2^b) b~U{0..32} or random walks (steps "
                            "U[-2^10,2^10]); one hipGraph launch (device prefix scan of the block lengths + decode of "
                            "every block) per step",
                "streams": args.streams, "values": n, "blocks": nb, "compressed_bytes": b_in, "bit_exact": exact,
                "seed": 3, "gen_seconds": round(gen_s, 1), "graph_form": form}
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": _traffic(f"decode_uneven200k_{form}", nb, kname) if args.streams == 200_000 else None,
            "kernel": kname, "kernel_ms": round(dec_ms, 5), "algo_bytes": algo_bytes,
            "stream_ceiling_GBps": ceiling, "frac_of_stream_ceiling": round(achieved / ceiling, 4)}
    res = _result(value, args.steps, args.warmup, ms_per_step, workload, roof)
    res["step_ms_events"] = {"p50": round(float(np.percentile(step_ms, 50)), 4),
                             "p99": round(float(np.percentile(step_ms, 99)), 4)}
    if args.copy_calibration:
        _copy_calibration(ctx)
    forms = {"graph_step_ms": round(ms_per_step, 4), "decode_kernel_ms": round(dec_ms, 5), "tune": tsum,
             "graph": ("wave index (2 passes) + decode in meta form" if form == "meta" else
                       "block index scan + decode from the block index"),
             "graph_trial_step_ms": {f: round(x, 5) for f, x in step_trial.items()},
             "tune_other_form": {f: tunes[f][1] for f in tunes if f != form}}
    if not args.no_extras:
        # each graph form's decode kernel alone, in the variant tuned for it
        for f, (tn, _, inst) in tunes.items():
            batch.set_tune(ctx, inst["arena_bytes"], inst["waves_per_wg"], inst["grid_wgs"], inst["tiles_per_grab"])
            p = bool(tn["persistent"])
            if f == "meta":
                fn = lambda: batch.decode_meta(ctx, s, wave, out=out, status=status, persistent=p)  # noqa: E731
            else:
                fn = lambda: batch.decode(ctx, s, out=out, status=status, persistent=p)  # noqa: E731
            forms[f"{f}_form_decode_ms"] = round(_events_time(ctx, fn, 5), 5)
            forms[f"{f}_form_ok"] = bool(np.all(status.to_numpy() == 0))
        batch.set_tune(ctx, installed["arena_bytes"], installed["waves_per_wg"], installed["grid_wgs"],
                       installed["tiles_per_grab"])
        forms["copy_kernel_GBps"] = _copy_gbps(ctx, 4 << 30)
        boff_h = s.block_off.to_numpy()
        rng = np.random.default_rng(0)
        ks = rng.integers(0, nb, 2000)
        blobs = [s.data.read(int(boff_h[k]), int(boff_h[k + 1] - boff_h[k])).tobytes() for k in ks]
        log("per-call form")
        forms["per_call"] = per_call_form(ctx, blobs)
        log("batched Get over the whole stream")
        forms["get_large_store"] = get_large_store(ctx, s, get_pos, get_want)
        if not args.no_cpu:
            # a bounded sample: the first 1/32 of the blocks (streams are in
            # random order, so the sample has the stream mix of the whole)
            k = max(1, nb // 32)
            ooff_h = s.out_off.read(0, k + 1)
            sb = s.data.read(0, int(boff_h[k]))
            log(f"cpu baseline (oracle) on {k} blocks")
            res["cpu_baseline"] = cpu_baseline(sb, boff_h[:k + 1], ooff_h, int(ooff_h[-1]),
                                               "the first 1/32 of the stream set's blocks")
    for g in graphs.values():
        g.close()
    for a in (wave, out, status):
        a.free()
    s.free()
    if not args.no_extras:
        log("second batched form: 64 Mi delta postings")
        d = headline_delta(ctx, args, cpu=False, extras=False)
        forms["delta64m"] = {k: d[k] for k in ("value", "ms_per_step")}
        forms["delta64m"].update({k: d["roofline"][k] for k in ("frac", "kernel_ms", "achieved")})
        forms["delta64m"]["bit_exact"] = d["config"]["bit_exact"]
        log("served form: a 16 GB resident store, two offered rates")
        forms["served"] = served_form(ctx)
    res["forms"] = forms
    return res


def get_large_store(ctx, s, pos, want) -> dict:
    """One batched GetUint32 launch of `pos` (stream value positions, each
    mapped to its block through the block index) over the resident stream,
    timed with HIP events (median of 3 runs of 5 launches), its values and
    statuses compared with `want`."""
    from utlpfor_amd import batch
    from utlpfor_amd.device import DeviceArray
    dpos = DeviceArray.from_numpy(pos, ctx)
    gout = DeviceArray(pos.size, np.uint32, ctx)
    gst = DeviceArray(pos.size, np.int32, ctx)
    fn = lambda: batch.get_batch(ctx, s, dpos, out=gout, status=gst)  # noqa: E731
    fn()
    ms = float(np.median([_events_time(ctx, fn, 5) for _ in range(3)]))
    ok = bool((gst.to_numpy() == 0).all() and np.array_equal(gout.to_numpy(), want))
    for a in (dpos, gout, gst):
        a.free()
    return {"positions": int(pos.size), "store_bytes": int(s.nbytes), "blocks": int(s.nblocks),
            "kernel_ms": round(ms, 4), "Gget_per_s": round(pos.size / (ms * 1e-3) / 1e9, 3),
            "ns_per_get": round(ms * 1e6 / pos.size, 4), "bit_exact": ok}


def served_form(ctx) -> dict:
    """The serving loop over a 16 GB resident store (60% decompress / 25%
    compress / 15% Get, sizes logU[128, 4Mi] values, mixed u32/u64): at 20,000
    requests/s with results left in HBM (near the p99 knee of the offered-load
    sweep) and at 2,000/s with results delivered to host memory."""
    from utlpfor_amd import runbench
    keep = ("offered_rate_per_s", "results", "requests", "p50_ms", "p99_ms", "GBps", "mean_batch", "max_batch",
            "ticks", "bit_exact_sampled")
    recs = list(runbench.serving_sweep(ctx, store_bytes=16 << 30, nreq=2000, rates=(20_000,), deliver_rates=(2_000,),
                                       progress=log))
    return {"store_bytes": recs[0]["store_bytes"], "points": [{k: r[k] for k in keep if k in r} for r in recs]}


def headline_delta(ctx, args, cpu: bool = True, extras: bool = True) -> dict:
    """64 Mi delta postings as one stream: index scan + decode per step."""
    from utlpfor_amd import batch, workloads
    from utlpfor_amd.device import DeviceArray
    lib = batch._lib.load()
    n = args.values
    log(f"generating {n} delta postings")
    v = workloads.delta_postings(n)
    s = batch.encode(ctx, v, flag=1)
    nb, b_in, d_out = s.nblocks, s.nbytes, n * 4
    log(f"{nb} blocks, {b_in} compressed bytes ({b_in / n:.3f} B/value)")
    out = DeviceArray(n, np.uint32, ctx)
    status = DeviceArray(nb, np.int32, ctx)
    boff, ooff = s.block_off, s.out_off
    saved = batch.get_tune(ctx)
    batch.set_tune(ctx, 2560, 4, saved["grid_wgs"], saved["tiles_per_grab"])

    def step():
        batch.check(lib.utl_index_build(ctx.handle, s.meta.ptr, nb, boff.ptr, ooff.ptr), "index")
        batch.decode(ctx, s, out=out, status=status)

    for _ in range(args.warmup):
        step()
    ctx.sync()
    exact = bool(np.all(status.to_numpy() == 0) and np.array_equal(out.to_numpy(), v))
    log(f"round trip bit-exact: {exact}")
    ms_per_step, step_ms = _timed_steps(ctx, step, args.steps)
    value = d_out / (ms_per_step * 1e-3) / 1e9
    dec_ms = _events_time(ctx, lambda: batch.decode(ctx, s, out=out, status=status), max(args.steps, 10))
    algo_bytes = b_in + d_out + 16 * nb
    achieved = algo_bytes / (dec_ms * 1e-3) / 1e9
    ceiling = _stream_ceiling(ctx, s.data, b_in, out, d_out)
    workload = {"workload": "delta postings: 64Mi uint32, gaps 1+Geom(1/8), 2% U[2^16,2^24), mod 2^32, flag Delta, "
                            "524288 blocks, whole-stream index scan + decode per step",
                "values": n, "blocks": nb, "compressed_bytes": b_in, "bit_exact": exact, "seed": 2}
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": _traffic("decode_delta64m", nb) if n == 64 << 20 else None,
            "kernel": _kernel_name(batch.get_tune(ctx)), "kernel_ms": round(dec_ms, 5), "algo_bytes": algo_bytes,
            "stream_ceiling_GBps": ceiling, "frac_of_stream_ceiling": round(achieved / ceiling, 4)}
    res = _result(value, args.steps, args.warmup, ms_per_step, workload, roof)
    res["step_ms_events"] = {"p50": round(float(np.percentile(step_ms, 50)), 4),
                             "p99": round(float(np.percentile(step_ms, 99)), 4)}
    if args.copy_calibration:
        _copy_calibration(ctx)
    if extras:
        pms = _events_time(ctx, lambda: batch.decode(ctx, s, out=out, status=status, persistent=True), 10)
        res["forms"] = {"batched_decode_kernel_GBps": round(d_out / (dec_ms * 1e-3) / 1e9, 1),
                        "persistent_decode_kernel_GBps": round(d_out / (pms * 1e-3) / 1e9, 1),
                        "persistent_ok": bool(np.all(status.to_numpy() == 0)), "tune": batch.get_tune(ctx),
                        "copy_kernel_GBps": _copy_gbps(ctx, d_out)}
        boff_h = boff.to_numpy()
        host_bytes = s.bytes_host()
        rng = np.random.default_rng(0)
        ks = rng.integers(0, nb, 2000)
        log("per-call form")
        res["forms"]["per_call"] = per_call_form(
            ctx, [host_bytes[int(boff_h[k]):int(boff_h[k + 1])].tobytes() for k in ks])
        if cpu:
            log("cpu baseline (oracle)")
            res["cpu_baseline"] = cpu_baseline(host_bytes, boff_h, ooff.to_numpy(), n, "the whole stream")
    batch.set_tune(ctx, saved["arena_bytes"], saved["waves_per_wg"], saved["grid_wgs"], saved["tiles_per_grab"])
    out.free()
    status.free()
    s.free()
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="uneven",
                    choices=["uneven", "delta64m", "plain1m", "u64", "serve", "serve-sweep", "walk", "encode",
                             "uneven-walk"],
                    help="uneven (200,000 streams) is the headline; the others print their own JSON lines")
    ap.add_argument("--streams", type=int, default=200_000, help="streams of the uneven workload")
    ap.add_argument("--values", type=int, default=64 << 20, help="values of the delta64m / walk workloads")
    ap.add_argument("--tune", default="", help="fixed decode variant ARENA:WAVES (flat) or ARENA:WAVES:GRID:GRAB "
                                               "(persistent); skips the autotuner")
    ap.add_argument("--form", default="auto", choices=["auto", "meta", "block"],
                    help="uneven workload: graph form (wave index + meta-form decode, or block index + decode); "
                         "auto times both and keeps the faster")
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle baseline")
    ap.add_argument("--no-extras", action="store_true", help="only the headline step (for profiling)")
    ap.add_argument("--store-gb", type=float, default=16.0, help="serve: compressed store size")
    ap.add_argument("--rate", type=float, default=2000.0, help="serve: offered requests per second")
    ap.add_argument("--requests", type=int, default=4000, help="serve: requests per run")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 256 MiB copy kernel (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        return 2

    from utlpfor_amd.device import Context

    ctx = Context(0)
    name, cus = ctx.device_info()
    log(f"device {name}, {cus} CUs")
    if args.workload == "uneven":
        print(json.dumps(headline_uneven(ctx, args)), flush=True)
        return 0
    if args.workload == "delta64m":
        print(json.dumps(headline_delta(ctx, args, cpu=not args.no_cpu, extras=not args.no_extras)), flush=True)
        return 0
    from utlpfor_amd import runbench
    if args.workload == "plain1m":
        print(json.dumps(runbench.plain_1m(ctx)), flush=True)
    elif args.workload == "walk":
        for i, rec in enumerate(runbench.walk_delta(ctx, args.values, progress=log)):
            if i == 0 and not args.no_cpu:
                log("CPU baseline: the oracle's serial walk of the same stream")
                rec["cpu_baseline"] = walk_cpu_baseline(rec.pop("stream_bytes"))
            rec.pop("stream_bytes", None)
            print(json.dumps(rec), flush=True)
    elif args.workload == "u64":
        for rec in runbench.u64_and_get(ctx):
            print(json.dumps(rec), flush=True)
    elif args.workload == "serve":
        print(json.dumps(runbench.serving(ctx, store_bytes=int(args.store_gb * (1 << 30)), rate=args.rate,
                                          nreq=args.requests, progress=log)), flush=True)
    elif args.workload == "encode":
        for rec in runbench.encode_forms(ctx, progress=log):
            print(json.dumps(rec), flush=True)
    elif args.workload == "uneven-walk":
        print(json.dumps(runbench.uneven(ctx, nstreams=args.streams, progress=log)), flush=True)
    elif args.workload == "serve-sweep":
        for rec in runbench.serving_sweep(ctx, store_bytes=int(args.store_gb * (1 << 30)), nreq=args.requests,
                                          progress=log):
            print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
