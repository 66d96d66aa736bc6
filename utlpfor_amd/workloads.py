"""Seeded synthetic workloads of the benchmark configurations (numpy).

plain_1m        1,048,576 uint32, U[0, 2^12) with round(3%) positions (no
                replacement) replaced by U[2^20, 2^32); flag 0.
delta_postings  64 Mi uint32 postings: gaps 1 + Geometric(p=1/8) (numpy's
                geometric has support >= 1, so gaps >= 2, mean 9); each gap is
                replaced with probability 2% by U[2^16, 2^24). The sum exceeds
                2^32 after ~25k values, so the sequence is accumulated modulo
                2^32 (wrap points become zigzag blocks). Flag Delta.
uneven_streams  streams of log-uniform length in [1, max_len], flags uniform
                over {0, Delta, NoFOR, NoPatch, Delta|NoFOR}, values uniform
                in [0, 2^b) (b uniform 0..32) or a signed random walk with
                steps U[-2^10, 2^10] cast to uint32.
timestamps64    1.7e18 + cumsum(Geometric(1/1000)) uint64 (FOR64 blocks).
uniform64       uint64 uniform over the full range (two-block encoding).
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

STREAM_FLAGS = np.array([0, 1, 2, 4, 3], dtype=np.uint8)


def plain_1m(n: int = 1 << 20, seed: int = 1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    v = rng.integers(0, 1 << 12, n, dtype=np.uint64)
    idx = rng.choice(n, round(0.03 * n), replace=False)
    v[idx] = rng.integers(1 << 20, 1 << 32, idx.size, dtype=np.uint64)
    return v.astype(np.uint32)


def delta_postings(n: int = 64 << 20, seed: int = 2) -> np.ndarray:
    rng = np.random.default_rng(seed)
    gaps = 1 + rng.geometric(1 / 8, n).astype(np.uint64)
    m = rng.random(n) < 0.02
    gaps[m] = rng.integers(1 << 16, 1 << 24, int(m.sum()), dtype=np.uint64)
    return (np.cumsum(gaps, dtype=np.uint64) & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def uneven_streams(nstreams: int = 200_000, max_len: int = 200_000, seed: int = 3, progress=None,
                   threads: int | None = None):
    """Returns (values uint32, stream lengths, stream flags). The per-stream
    parameters come from one seeded generator; the values are drawn in
    chunks of consecutive streams (at most 2^23 values, or one stream), each
    from its own generator spawned from the seed, on a pool of threads
    (numpy's generators and most array operations release the GIL), with
    segmented cumulative sums for the random walks. The result depends only
    on the arguments, not on the thread count."""
    rng = np.random.default_rng(seed)
    lens = np.floor(np.exp(rng.uniform(0, np.log(max_len + 1), nstreams))).astype(np.int64)
    lens = np.clip(lens, 1, max_len)
    flags = STREAM_FLAGS[rng.integers(0, STREAM_FLAGS.size, nstreams)]
    walk = rng.random(nstreams) < 0.5
    bits = rng.integers(0, 33, nstreams)
    start = rng.integers(0, 1 << 31, nstreams)
    ends = np.cumsum(lens)
    total = int(ends[-1])
    out = np.empty(total, dtype=np.uint32)
    chunk = 1 << 23
    bounds = []
    s0 = 0
    while s0 < nstreams:
        a = int(ends[s0 - 1]) if s0 else 0
        s1 = max(int(np.searchsorted(ends, a + chunk, side="right")), s0 + 1)
        bounds.append((s0, s1))
        s0 = s1
    seeds = np.random.SeedSequence(seed).spawn(len(bounds))

    def fill(ci: int) -> int:
        s0, s1 = bounds[ci]
        crng = np.random.default_rng(seeds[ci])
        a = int(ends[s0 - 1]) if s0 else 0
        b = int(ends[s1 - 1])
        L = lens[s0:s1]
        # uniform in [0, 2^b): the top b bits of a uniform 32-bit draw
        raw = crng.integers(0, 1 << 32, b - a, dtype=np.uint64)
        raw >>= np.repeat((32 - bits[s0:s1]).astype(np.uint64), L)
        steps = crng.integers(-(1 << 10), (1 << 10) + 1, b - a, dtype=np.int32).astype(np.int64)
        cs = np.cumsum(steps)
        seg0 = np.concatenate([[0], np.cumsum(L)[:-1]])
        # the walk's value: its start plus the steps since the stream's first
        cs += np.repeat(start[s0:s1] - (cs[seg0] - steps[seg0]), L)
        out[a:b] = np.where(np.repeat(walk[s0:s1], L), cs.astype(np.uint32), raw.astype(np.uint32))
        return b

    nthreads = threads or max(1, min(16, len(os.sched_getaffinity(0))))
    with ThreadPoolExecutor(nthreads) as ex:
        done = 0
        for b in ex.map(fill, range(len(bounds))):
            done = max(done, b)
            if progress and done % (1 << 28) < chunk:
                progress(f"generated {done}/{total} values")
    return out, lens.astype(np.uint64), flags


def timestamps64(n: int = 32 << 20, seed: int = 4) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return (np.uint64(1_700_000_000_000_000_000) + np.cumsum(rng.geometric(1 / 1000, n).astype(np.uint64)))


def uniform64(n: int = 8 << 20, seed: int = 5) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.integers(0, 1 << 64, n, dtype=np.uint64, endpoint=False)
