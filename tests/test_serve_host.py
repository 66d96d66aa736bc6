"""Host logic of the serving loop (no GPU): admission in arrival order within
the heap, admission by the device's free memory (the heap grows while it
lasts), continuous batching of late arrivals into the next tick, latency
accounting, deliveries completed before serve() returns or raises, and the
request generator's mix and ranges."""
import numpy as np

from utlpfor_amd import serve


class _FakeStore:
    def __init__(self):
        self.streams = [serve.StoreStream(0, 100, 100 * 128, False, 0, 0),
                        serve.StoreStream(100, 50, 50 * 128 - 7, True, 1, 0)]


class _Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


def _engine(cap, clock, tick_s):
    eng = serve.Engine.__new__(serve.Engine)
    eng.store = _FakeStore()
    eng.heap = type("H", (), {"cap": cap})()
    eng.batches = []
    eng.growths = 0
    eng.admission_limit = lambda half: cap  # a fixed heap
    eng.fit_half = lambda half, need: None

    def launch(reqs, half):  # one tick's kernels: no deliveries
        eng.batches.append([id(r) for r in reqs])
        eng.halves.append(half)
        clock.t += tick_s
        return []

    eng.halves = []
    eng._launch = launch
    return eng


def test_admission_fifo_within_heap(monkeypatch):
    clock = _Clock()
    monkeypatch.setattr(serve.time, "sleep", lambda s: setattr(clock, "t", clock.t + s))
    reqs = [serve.Request(serve.DECOMPRESS, 0.0, stream=0, first=i, nblocks=10) for i in range(6)]
    cost = serve.Engine.cost(_engine(1, clock, 0), reqs[0])
    eng = _engine(cost * 2, clock, 0.01)
    # without delivery a request is done when its tick's kernels are
    stats = eng.serve(reqs, deliver=False, clock=clock)
    assert [len(b) for b in eng.batches] == [2, 2, 2]
    assert eng.batches[0] == [id(reqs[0]), id(reqs[1])]  # arrival order
    assert eng.halves == [0, 1, 0]  # ticks use the heap halves in turn
    lat = [r.latency() for r in reqs]
    assert np.allclose(lat, [0.01, 0.01, 0.02, 0.02, 0.03, 0.03])
    assert stats["ticks"] == 3 and stats["requests"] == 6


def test_delivery_marks_requests_done(monkeypatch):
    """With delivery, each tick's results are delivered by a worker thread
    (here nothing to copy) and every request gets a completion time no
    earlier than its tick's kernels, in tick order."""
    clock = _Clock()
    monkeypatch.setattr(serve.time, "sleep", lambda s: setattr(clock, "t", clock.t + s))
    reqs = [serve.Request(serve.DECOMPRESS, 0.0, stream=0, first=i, nblocks=10) for i in range(6)]
    cost = serve.Engine.cost(_engine(1, clock, 0), reqs[0])
    eng = _engine(cost * 2, clock, 0.01)
    stats = eng.serve(reqs, clock=clock)
    assert stats["ticks"] == 3 and [r.tick for r in reqs] == [0, 0, 1, 1, 2, 2]
    assert all(r.latency() >= 0.01 * (r.tick + 1) - 1e-12 for r in reqs)


def test_late_arrivals_join_next_tick(monkeypatch):
    clock = _Clock()
    monkeypatch.setattr(serve.time, "sleep", lambda s: setattr(clock, "t", clock.t + s))
    eng = _engine(1 << 40, clock, 0.005)
    reqs = [serve.Request(serve.GET, t, stream=0, positions=np.array([1, 2])) for t in (0.0, 0.001, 0.004, 0.02)]
    eng.serve(reqs, deliver=False, clock=clock)
    # the first tick starts at t=0 with one request; the two that arrive
    # during it share the next tick; the last arrives after an idle gap
    assert [len(b) for b in eng.batches] == [1, 2, 1]
    assert reqs[3].tick == 2 and reqs[3].latency() >= 0


def test_check_request_rejects_out_of_range():
    clock = _Clock()
    eng = _engine(1 << 40, clock, 0)
    bad = [serve.Request(serve.DECOMPRESS, stream=0, first=95, nblocks=10),
           serve.Request(serve.GET, stream=1, positions=np.array([50 * 128 - 7])),
           serve.Request(serve.COMPRESS, values=np.zeros(0, np.uint32))]
    for r in bad:
        try:
            eng.check_request(r)
        except ValueError:
            continue
        raise AssertionError(f"accepted {r}")
    eng.check_request(serve.Request(serve.GET, stream=1, positions=np.array([50 * 128 - 8])))


def test_cost_counts_device_bytes():
    clock = _Clock()
    eng = _engine(1 << 40, clock, 0)
    d = serve.Request(serve.DECOMPRESS, stream=1, first=49, nblocks=1)  # last, partial block of a u64 stream
    assert eng.cost(d) >= (128 - 7) * 8 + 2 * 16
    c = serve.Request(serve.COMPRESS, values=np.zeros(300, np.uint64))
    assert eng.cost(c) >= 300 * 8 + 3 * 2048


def test_make_requests_mix_and_ranges():
    store = _FakeStore()
    pool = [(np.arange(10000, dtype=np.uint32), 1), (np.arange(5000, dtype=np.uint64), 0)]
    reqs = serve.make_requests(store, pool, n=4000, rate=1000.0, seed=3, min_values=128, max_values=1 << 20)
    kinds = np.array([r.kind for r in reqs])
    frac = np.bincount(kinds, minlength=3) / kinds.size
    assert np.allclose(frac, [0.6, 0.25, 0.15], atol=0.03)
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) >= 0) and 3.0 < arr[-1] < 5.0  # ~n / rate seconds
    for r in reqs:
        if r.kind == serve.DECOMPRESS:
            st = store.streams[r.stream]
            assert 1 <= r.nblocks and r.first + r.nblocks <= st.nblocks
        elif r.kind == serve.GET:
            assert 1 <= r.positions.size <= 1024
            assert r.positions.max() < store.streams[r.stream].nvalues
        else:
            assert 1 <= r.values.size <= 10000


class _Half:
    def __init__(self, cap):
        self.cap = cap

    def grow(self, n):
        self.cap = max(self.cap, n)


def _growing_engine(cap, free, reserve, clock, tick_s):
    """An engine whose admission reads a fake free-memory figure."""
    eng = serve.Engine.__new__(serve.Engine)
    eng.store = _FakeStore()
    eng.heaps = [_Half(cap), _Half(cap)]
    eng.heap = eng.heaps[0]
    eng.grow, eng.reserve, eng.growths = True, reserve, 0
    eng.batches, eng.halves = [], []
    eng.free = free

    def launch(reqs, half):
        assert sum(eng.cost(r) for r in reqs) <= eng.heaps[half].cap  # fit_half ran first
        eng.batches.append(len(reqs))
        eng.halves.append(half)
        clock.t += tick_s
        return []

    eng._launch = launch
    return eng


def test_admission_by_live_free_memory(monkeypatch):
    """Admission follows the device's free memory at each tick: with room
    the heap half grows and one tick takes everything that arrived; when the
    free memory falls to the reserve, ticks are cut to the heap's size."""
    clock = _Clock()
    monkeypatch.setattr(serve.time, "sleep", lambda s: setattr(clock, "t", clock.t + s))
    reqs = [serve.Request(serve.DECOMPRESS, 0.0, stream=0, first=i, nblocks=10) for i in range(8)]
    one = serve.Engine.cost(_engine(1, clock, 0), reqs[0])
    state = {"free": 100 * one}
    monkeypatch.setattr(serve, "mem_info", lambda: (state["free"], 1 << 40))
    eng = _growing_engine(2 * one, 100 * one, 0, clock, 0.01)
    eng.serve(reqs, deliver=False, clock=clock)
    assert eng.batches == [8] and eng.growths == 1 and eng.heaps[0].cap >= 8 * one
    # no free memory beyond the reserve: the heap's own size bounds a tick
    state["free"] = 0
    eng = _growing_engine(2 * one, 0, 0, clock, 0.01)
    for r in reqs:
        r.done = float("nan")
    eng.serve(reqs, deliver=False, clock=clock)
    assert eng.batches == [2, 2, 2, 2] and eng.growths == 0


def test_serve_finishes_deliveries_when_a_tick_fails(monkeypatch):
    """A launch that raises on the second tick: serve() re-raises only after
    the first tick's delivery (running on the worker thread) has completed,
    so no later tick can reuse that tick's heap half while it is copied."""
    import threading
    clock = _Clock()
    monkeypatch.setattr(serve.time, "sleep", lambda s: setattr(clock, "t", clock.t + s))
    reqs = [serve.Request(serve.DECOMPRESS, 0.0, stream=0, first=i, nblocks=10) for i in range(4)]
    cost = serve.Engine.cost(_engine(1, clock, 0), reqs[0])
    eng = _engine(cost * 2, clock, 0.01)
    started, release, delivered = threading.Event(), threading.Event(), []

    def slow_delivery():
        started.set()
        release.wait(5)
        delivered.append(True)

    calls = []

    def launch(b, half):
        calls.append(half)
        if len(calls) == 2:
            assert started.wait(5)
            threading.Timer(0.2, release.set).start()  # the delivery is still copying
            raise MemoryError("serving heap exhausted")
        return [slow_delivery]

    eng._launch = launch
    try:
        eng.serve(reqs, clock=clock)
    except MemoryError:
        pass
    else:
        raise AssertionError("the launch error was swallowed")
    assert delivered == [True]  # the first tick's delivery finished before serve() raised
    assert reqs[0].tick == 0 and reqs[1].tick == 0
