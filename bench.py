#!/usr/bin/env python3
"""Benchmark of the UTL-PFOR decode hot path on one MI355X.

Workload: 64 Mi uint32 delta postings (gaps 1 + Geometric(1/8), 2% of gaps
U[2^16, 2^24), accumulated mod 2^32), packed with flag Delta into 524,288
concatenated 128-value blocks by the GPU encoder. One step = decode of the
whole stream with inputs resident in HBM: the device index scan over the
per-block lengths (block byte offsets + value offsets) followed by one
decode launch. value = decoded GB/s of the whole step.

Also reported: the persistent work-queue form, the plain per-call drop-in
(one C-ABI call per block with host staging, p50/p99 latency), the decode
kernel's roofline from HIP events on its own stream, and the oracle (the C
restatement of the reference's scalar path) timed on the host's cores.
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


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(stream_bytes, boff, ooff, nvalues, target_s: float = 12.0):
    """The oracle decoding the same stream on the host's cores."""
    from oracle import oracle
    cores = len(os.sched_getaffinity(0))
    threads = max(1, min(cores, 16))
    out, st = oracle.unpack_blocks(stream_bytes, boff, ooff, nvalues, False, nthreads=threads)  # warm-up
    reps, t0 = 0, time.perf_counter()
    while True:
        oracle.unpack_blocks(stream_bytes, boff, ooff, nvalues, False, nthreads=threads)
        reps += 1
        el = time.perf_counter() - t0
        if el * threads >= target_s or el > target_s:
            break
    gbs = reps * nvalues * 4 / el / 1e9
    return {"value": round(gbs, 3), "unit": "GB/s decoded", "cores": threads, "kind": "port",
            "sample": f"the full {nvalues}-value stream decoded {reps}x by the C restatement of the "
                      f"reference scalar path on {threads} threads ({el:.1f} s wall)"}


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


def per_call_form(ctx, stream_bytes, boff, nblocks: int, calls: int = 2000):
    """One C-ABI UnpackUint32 call per block (host bytes in, host values out)."""
    from utlpfor_amd import codec
    rng = np.random.default_rng(0)
    ks = rng.integers(0, nblocks, calls)
    blobs = [stream_bytes[int(boff[k]):int(boff[k + 1])].tobytes() for k in ks]
    for b in blobs[:50]:
        codec.UnpackUint32(b, ctx=ctx)
    lat = np.empty(calls)
    t0 = time.perf_counter()
    for i, b in enumerate(blobs):
        t = time.perf_counter()
        _, _, err = codec.UnpackUint32(b, ctx=ctx)
        lat[i] = time.perf_counter() - t
        assert err is None
    el = time.perf_counter() - t0
    return {"GB/s": round(calls * 512 / el / 1e9, 6), "p50_us": round(float(np.percentile(lat, 50)) * 1e6, 1),
            "p99_us": round(float(np.percentile(lat, 99)) * 1e6, 1), "calls": calls}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--values", type=int, default=64 << 20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle baseline")
    ap.add_argument("--no-extras", action="store_true", help="only the headline step (for profiling)")
    ap.add_argument("--workload", default="delta64m", choices=["delta64m", "plain1m", "uneven", "u64", "serve", "walk"],
                    help="delta64m is the headline; the others print their own JSON lines")
    ap.add_argument("--store-gb", type=float, default=16.0, help="serve: compressed store size")
    ap.add_argument("--rate", type=float, default=2000.0, help="serve: offered requests per second")
    ap.add_argument("--requests", type=int, default=4000, help="serve: requests per run")
    ap.add_argument("--streams", type=int, default=200_000, help="streams of the uneven workload")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 256 MiB copy kernel (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        return 2

    from utlpfor_amd import batch, workloads
    from utlpfor_amd.device import Context, DeviceArray, Event, dev_copy

    ctx = Context(0)
    name, cus = ctx.device_info()
    log(f"device {name}, {cus} CUs")
    if args.workload != "delta64m":
        from utlpfor_amd import runbench
        if args.workload == "plain1m":
            print(json.dumps(runbench.plain_1m(ctx)), flush=True)
        elif args.workload == "uneven":
            print(json.dumps(runbench.uneven(ctx, nstreams=args.streams, progress=log)), flush=True)
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
        return 0
    n = args.values
    log(f"generating {n} delta postings")
    v = workloads.delta_postings(n)
    log("GPU encode (flag Delta)")
    s = batch.encode(ctx, v, flag=1)
    nb, b_in, d_out = s.nblocks, s.nbytes, n * 4
    log(f"{nb} blocks, {b_in} compressed bytes ({b_in / n:.3f} B/value)")

    out = DeviceArray(n, np.uint32, ctx)
    status = DeviceArray(nb, np.int32, ctx)
    boff, ooff = s.block_off, s.out_off

    lib = batch._lib.load()

    def step():
        batch.check(lib.utl_index_build(ctx.handle, s.meta.ptr, nb, boff.ptr, ooff.ptr), "index")
        batch.decode(ctx, s, out=out, status=status)

    for _ in range(args.warmup):
        step()
    ctx.sync()
    # bit-exact round trip of the decoded stream
    st = status.to_numpy()
    got = out.to_numpy()
    exact = bool(np.all(st == 0) and np.array_equal(got, v))
    log(f"round trip bit-exact: {exact}")
    del got

    ev = [Event() for _ in range(args.steps + 1)]
    ctx.sync()
    t0 = time.perf_counter()
    ev[0].record(ctx)
    for i in range(args.steps):
        step()
        ev[i + 1].record(ctx)
    ctx.sync()
    wall = time.perf_counter() - t0
    step_ms = [ev[i].elapsed_ms(ev[i + 1]) for i in range(args.steps)]
    ms_per_step = wall * 1e3 / args.steps
    value = d_out / (ms_per_step * 1e-3) / 1e9

    # dominant kernel alone: the decode launch between two events
    e0, e1 = Event(), Event()
    reps = max(args.steps, 10)
    e0.record(ctx)
    for _ in range(reps):
        batch.decode(ctx, s, out=out, status=status)
    e1.record(ctx)
    ctx.sync()
    dec_ms = e0.elapsed_ms(e1) / reps
    algo_bytes = b_in + d_out + 16 * nb  # + block_off/out_off reads
    achieved = algo_bytes / (dec_ms * 1e-3) / 1e9
    # HBM bytes per dispatch from the committed rocprofv3 FETCH_SIZE /
    # WRITE_SIZE passes over this workload (copy-kernel calibrated)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "decode_delta64m_traffic.json")
    if os.path.exists(tpath) and n == 64 << 20:
        with open(tpath) as f:
            traffic = json.load(f).get("traffic_bytes")
    result = {
        "metric": "decoded GB/s & %HBM peak, batched vs plain per-call; p50/p99 latency, bit-exact",
        "value": round(value, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": "delta postings: 64Mi uint32, gaps 1+Geom(1/8), 2% U[2^16,2^24), mod 2^32, "
                               "flag Delta, 524288 blocks, whole-stream index scan + decode per step",
                   "values": n, "blocks": nb, "compressed_bytes": b_in, "bit_exact": exact},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_decode<false>", "kernel_ms": round(dec_ms, 5), "algo_bytes": algo_bytes},
        "step_ms_events": {"p50": round(float(np.percentile(step_ms, 50)), 4),
                           "p99": round(float(np.percentile(step_ms, 99)), 4)},
    }
    if args.copy_calibration:
        # a plain copy kernel of known size (256 MiB read + 256 MiB written),
        # whose FETCH_SIZE/WRITE_SIZE scale the decode kernel's counters
        src = DeviceArray(256 << 20, np.uint8, ctx)
        dst = DeviceArray(256 << 20, np.uint8, ctx)
        for _ in range(3):
            dev_copy(ctx, dst, src, 256 << 20)
        ctx.sync()
        src.free()
        dst.free()
    if not args.no_extras:
        # persistent work-queue form
        e0.record(ctx)
        for _ in range(reps):
            batch.decode(ctx, s, out=out, status=status, persistent=True)
        e1.record(ctx)
        ctx.sync()
        pms = e0.elapsed_ms(e1) / reps
        pok = bool(np.all(status.to_numpy() == 0))
        result["forms"] = {"batched_decode_kernel_GBps": round(d_out / (dec_ms * 1e-3) / 1e9, 1),
                           "persistent_decode_kernel_GBps": round(d_out / (pms * 1e-3) / 1e9, 1),
                           "persistent_ok": pok, "tune": list(batch.get_tune(ctx))}
        # copy-kernel reference (what a plain copy of the same bytes achieves)
        src = DeviceArray(d_out, np.uint8, ctx)
        dst = DeviceArray(d_out, np.uint8, ctx)
        dev_copy(ctx, dst, src, d_out)
        e0.record(ctx)
        for _ in range(10):
            dev_copy(ctx, dst, src, d_out)
        e1.record(ctx)
        ctx.sync()
        cms = e0.elapsed_ms(e1) / 10
        result["forms"]["copy_kernel_GBps"] = round(2 * d_out / (cms * 1e-3) / 1e9, 1)
        src.free()
        dst.free()
        log("per-call form")
        host_bytes = s.bytes_host()
        boff_h = boff.to_numpy()
        result["forms"]["per_call"] = per_call_form(ctx, host_bytes, boff_h, nb)
        if not args.no_cpu:
            log("cpu baseline (oracle)")
            result["cpu_baseline"] = cpu_baseline(host_bytes, boff_h, ooff.to_numpy(), n)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
