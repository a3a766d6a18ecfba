"""Serving loop scheduling on the CPU with stub contexts (no GPU): admission
by HBM footprint with LRU eviction, rejection of oversize requests, stores
shared by requests of one database, micro-batches of the requests that wait
while the slots are busy, and slots that run at the same time."""

import threading
import time

import numpy as np

from cupami_amd.serve import Request, Server, latency_summary, row_stride_words


class StubStore:
    live = 0
    lock = threading.Lock()

    def __init__(self, key):
        self.key = key
        with StubStore.lock:
            StubStore.live += 1

    def close(self):
        with StubStore.lock:
            StubStore.live -= 1


class StubPatterns:
    # a one-pattern forest: column 0, support 1
    parent = np.array([0xFFFFFFFF], dtype=np.uint32)
    item = np.array([0], dtype=np.uint32)
    support = np.array([1], dtype=np.uint64)


class StubForest:
    def __init__(self, n):
        self.n = n

    def split_arrays(self, n_db):
        return [{("db", i): 1} for i in range(n_db)]


class StubCtx:
    """Records builds and searches; a search takes `mine_s` seconds."""
    mine_s = 0.0
    events = []
    active = 0
    active_max = 0
    lock = threading.Lock()

    def __init__(self, device):
        self.device = device

    def build_bits(self, rows, cols, n, items):
        StubCtx.events.append(("build", int(rows.size)))
        return StubStore(int(rows.size))

    def synchronize(self):
        pass

    def _search(self, n):
        with StubCtx.lock:
            StubCtx.active += 1
            StubCtx.active_max = max(StubCtx.active_max, StubCtx.active)
        time.sleep(StubCtx.mine_s)
        with StubCtx.lock:
            StubCtx.active -= 1
        StubCtx.events.append(("mine", n))

    def eclat_mine(self, bits, minsup):
        self._search(1)
        return StubPatterns()

    def eclat_mine_batch(self, stores, minsups):
        self._search(len(stores))
        return StubForest(len(stores))

    def close(self):
        pass


def reset(mine_s=0.0):
    StubCtx.mine_s = mine_s
    StubCtx.events = []
    StubCtx.active = StubCtx.active_max = 0


def req(rid, arrival, n_trans, items, nnz, key=None):
    return Request(rid, arrival, np.zeros(nnz, np.uint64), np.zeros(nnz, np.uint32), n_trans, items, 1, key=key)


def test_footprint():
    r = req(0, 0.0, 1000, 10, 50)
    assert r.footprint() == 10 * row_stride_words(1000) * 8 + 12 * 50
    assert row_stride_words(1) == 64 and row_stride_words(64 * 64 + 1) == 128


def test_admission_evicts_and_rejects():
    reset()
    small = req(0, 0.0, 64, 1, 10).footprint()
    budget = 2 * small + small // 2          # two stores fit, not three
    reqs = [req(0, 0.0, 64, 1, 10), req(1, 0.0, 64, 1, 10), req(2, 0.0, 64, 1, 10),
            req(3, 0.0, 10_000_000, 5000, 10),  # larger than the whole budget
            req(4, 0.01, 64, 1, 10)]
    server = Server(hbm_budget=budget, slots=2, contexts=StubCtx)
    out = server.run(reqs)
    assert [o.rid for o in out] == [0, 1, 2, 3, 4]
    assert out[3].rejected and not out[3].patterns
    assert all(not o.rejected and o.patterns for o in out if o.rid != 3)
    assert server.stats["builds"] == 4 and server.stats["evictions"] >= 2
    assert server.stats["resident_bytes"] <= budget and server.stats["resident_stores"] <= 2
    lat = latency_summary(out)
    assert lat["p50_ms"] is not None and lat["p99_ms"] >= lat["p50_ms"]
    server.close()


def test_requests_of_one_database_share_its_store():
    reset()
    reqs = [req(i, 0.001 * i, 64, 1, 10 + (i % 2), key=i % 2) for i in range(10)]
    server = Server(hbm_budget=1 << 30, slots=2, contexts=StubCtx)
    out = server.run(reqs)
    assert all(o.patterns for o in out)
    assert server.stats["builds"] == 2
    assert server.stats["resident_stores"] == 2
    # a second run finds both stores resident
    again = server.run([req(20 + i, 0.0, 64, 1, 10 + (i % 2), key=i % 2) for i in range(4)])
    assert all(o.patterns for o in again) and server.stats["builds"] == 0
    assert server.stats["resident_hits"] == 4
    server.close()


def test_waiting_requests_form_micro_batches():
    # one slot, searches of 50 ms: requests arriving meanwhile are mined together
    reset(0.05)
    reqs = [req(0, 0.0, 64, 1, 10, key=0)] + [req(i, 0.01, 64, 1, 10, key=0) for i in range(1, 6)]
    server = Server(hbm_budget=1 << 30, slots=1, contexts=StubCtx)
    out = server.run(reqs)
    assert all(o.patterns for o in out)
    sizes = [len(b[3]) for b in server.batches]
    assert sizes[0] == 1 and max(sizes) == 5 and sum(sizes) == 6
    assert out[1].batch == out[5].batch
    server.close()


def test_batch_entry_cap():
    reset(0.02)
    reqs = [req(0, 0.0, 64, 1, 10, key=9)] + [req(i, 0.005, 64, 1, 1000, key=i) for i in range(1, 5)]
    server = Server(hbm_budget=1 << 30, slots=1, max_batch_entries=2000, contexts=StubCtx)
    server.run(reqs)
    assert max(len(b[3]) for b in server.batches) <= 2
    server.close()


def test_slots_overlap():
    # three slots, searches of 40 ms, requests 5 ms apart: several run at once
    reset(0.04)
    reqs = [req(i, 0.005 * i, 64, 1, 10, key=i % 3) for i in range(9)]
    server = Server(hbm_budget=1 << 30, slots=3, max_batch=1, contexts=StubCtx)
    out = server.run(reqs)
    assert all(o.patterns for o in out)
    assert server.stats["busy_max"] >= 2 and StubCtx.active_max >= 2
    # no request waited for the ones before it to finish: the makespan is
    # well below nine sequential searches
    assert max(o.finish for o in out) < 9 * 0.04
    server.close()
