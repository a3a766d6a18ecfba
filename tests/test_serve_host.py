"""Serving loop admission and batching on the CPU, with a stub device
context (no GPU): oversize requests are rejected, requests that do not fit
the remaining budget wait for a later batch, and batches keep arrival order."""

import numpy as np

from cupami_amd.serve import SEARCH_RESERVE, Request, Server, latency_summary, row_stride_words


class StubForest:
    def __init__(self, n):
        self.n = n

    def split_arrays(self, n_db):
        return [{("db", i): 1} for i in range(n_db)]


class StubCtx:
    def __init__(self):
        self.batches = []

    def build_bits_batch(self, dbs):
        self.batches.append(len(dbs))
        return [StubStore() for _ in dbs]

    def eclat_mine_batch(self, stores, minsups):
        return StubForest(len(stores))

    def build_bits(self, rows, cols, n_trans, n_items):
        self.batches.append(1)
        return StubStore()

    def eclat_mine(self, bits, minsup):
        return StubPatterns()


class StubPatterns:
    # a one-pattern forest: column 0, support 1
    parent = np.array([0xFFFFFFFF], dtype=np.uint32)
    item = np.array([0], dtype=np.uint32)
    support = np.array([1], dtype=np.uint64)


class StubStore:
    closed = 0

    def close(self):
        StubStore.closed += 1


def req(rid, arrival, n_trans, items, nnz):
    return Request(rid, arrival, np.zeros(nnz, np.uint64), np.zeros(nnz, np.uint32), n_trans, items, 1)


def test_footprint():
    r = req(0, 0.0, 1000, 10, 50)
    assert r.footprint() == 10 * row_stride_words(1000) * 8 + 12 * 50 + SEARCH_RESERVE
    assert row_stride_words(1) == 64 and row_stride_words(64 * 64 + 1) == 128


def test_admission_rejects_defers_and_batches():
    small = req(0, 0.0, 64, 1, 10).footprint()
    budget = 2 * small + small // 2          # two small requests fit, not three
    reqs = [req(0, 0.0, 64, 1, 10), req(1, 0.0, 64, 1, 10), req(2, 0.0, 64, 1, 10),
            req(3, 0.0, 10_000_000, 5000, 10),  # larger than the whole budget
            req(4, 0.01, 64, 1, 10)]
    ctx = StubCtx()
    server = Server(ctx, hbm_budget=budget, max_batch=8, max_batch_pairs=1 << 20)
    out = server.run(reqs)
    assert [o.rid for o in out] == [0, 1, 2, 3, 4]
    assert out[3].rejected and not out[3].patterns
    assert all(not o.rejected and o.patterns for o in out if o.rid != 3)
    # never more than two requests in one batch; arrival order kept
    assert max(len(b[2]) for b in server.batches) <= 2
    order = [rid for b in server.batches for rid in b[2]]
    assert order == sorted(order)
    assert out[0].batch == out[1].batch and out[2].batch > out[1].batch
    assert StubStore.closed >= 4
    lat = latency_summary(out)
    assert lat["p50_ms"] is not None and lat["p99_ms"] >= lat["p50_ms"]


def test_work_cap_runs_large_requests_alone():
    reqs = [req(0, 0.0, 64, 1, 10), req(1, 0.0, 64, 1, 3000), req(2, 0.0, 64, 1, 10), req(3, 0.0, 64, 1, 10)]
    server = Server(StubCtx(), hbm_budget=1 << 40, max_batch=8, max_batch_pairs=2000)
    out = server.run(reqs)
    assert [b[2] for b in server.batches] == [[0], [1], [2, 3]]
    assert all(o.patterns for o in out)
