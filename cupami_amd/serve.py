"""Serving loop: many mining requests on one GPU, with continuous batching
and admission by HBM footprint.

Each request is one database (COO pairs in host memory) and an absolute
minimum support.  The loop keeps a FIFO queue of arrived requests.  Whenever
the device is free it admits, in arrival order, waiting requests while their
footprints fit the HBM budget and their pairs fit the batch work cap (a
request larger than the whole budget is rejected; one that does not fit
waits for the next turn), builds all admitted stores with one batch call,
mines them with one search (``cpm_eclat_mine_batch``) and hands each request
its own patterns.  Requests that arrive while a batch runs join the next
batch.  The work cap keeps batching where it pays: a request with millions
of pairs fills the device by itself and runs alone (packing it with others
would only make them wait for its deeper levels), while small requests,
which leave the device mostly idle one at a time, share a search.

``mine_one_at_a_time`` serves the same arrival stream one request at a time
(one store build and one search per request), the plain way to run the
reference's miner repeatedly; both return per-request latencies.  A
request's patterns are the same whichever way it is served.
"""

from __future__ import annotations

import time
from collections import deque
from collections.abc import Mapping
from dataclasses import dataclass, field

import numpy as np

from . import _hip

SEARCH_RESERVE = 64 << 20   # bytes per request for its share of the search buffers


def row_stride_words(n_transactions):
    words = (n_transactions + 63) // 64
    return (words + 63) // 64 * 64


@dataclass
class Request:
    """A database to mine, as pairs (``rows``, ``cols``: item columns in
    label string order) or in the BIN file layout (``offsets`` u64[T+1] and
    raw item ``ids`` u32 up to ``universe``, ``rows``/``cols`` None): the
    device builds the latter's store and labels itself, 4 bytes per entry
    and 8 per basket over PCIe instead of 12 per pair.  Its columns are the
    distinct ids in string order, the same numbering ``label_columns``
    gives the pairs."""
    rid: int
    arrival: float            # seconds after the start of the run
    rows: np.ndarray | None   # uint64 transaction index per pair
    cols: np.ndarray | None   # uint32 item column per pair
    n_transactions: int
    n_items: int
    minsup: int
    offsets: np.ndarray | None = None
    ids: np.ndarray | None = None
    universe: int = 0

    @property
    def entries(self):
        return int(self.ids.size if self.ids is not None else self.rows.size)

    def footprint(self):
        """Device bytes while it is being served: its store, its staged
        input, and a share of the search buffers."""
        store = self.n_items * row_stride_words(self.n_transactions) * 8
        staged = (4 * self.entries + 8 * (self.n_transactions + 1) if self.ids is not None
                  else 12 * self.entries)
        return store + staged + SEARCH_RESERVE


def _build(ctx, r):
    if r.ids is not None:
        return ctx.build_bits_csr(r.offsets, r.ids, r.universe, labels=False)[0]
    return ctx.build_bits(r.rows, r.cols, r.n_transactions, r.n_items)


@dataclass
class Outcome:
    rid: int
    arrival: float
    finish: float = 0.0
    patterns: Mapping = field(default_factory=dict)   # {column tuple: support}, built on first read
    batch: int = -1
    rejected: str = ""

    @property
    def latency(self):
        return self.finish - self.arrival


class Server:
    def __init__(self, ctx=None, hbm_budget=8 << 30, max_batch=256, max_batch_pairs=4 << 20):
        self.ctx = ctx or _hip.context(0)
        self.hbm_budget = int(hbm_budget)
        self.max_batch = int(max_batch)
        self.max_batch_pairs = int(max_batch_pairs)
        self.batches = []      # (start, end, request ids)

    def _admit(self, queue):
        """Admit in arrival order while the HBM budget and the work cap allow
        (the first request always fits: oversize ones were rejected)."""
        batch, used, pairs = [], 0, 0
        while queue and len(batch) < self.max_batch:
            fp = queue[0].footprint()
            n = queue[0].entries
            if batch and (used + fp > self.hbm_budget or pairs + n > self.max_batch_pairs):
                break
            batch.append(queue.popleft())
            used += fp
            pairs += n
        return batch

    def _mine(self, batch):
        if len(batch) == 1:  # nothing to share: the single-store path
            r = batch[0]
            bits = _build(self.ctx, r)
            try:
                return [_hip.PatternArrays.of(self.ctx.eclat_mine(bits, r.minsup))]
            finally:
                bits.close()
        if all(r.ids is None for r in batch):
            stores = self.ctx.build_bits_batch([(r.rows, r.cols, r.n_transactions, r.n_items) for r in batch])
        else:
            stores = []
            try:
                for r in batch:
                    stores.append(_build(self.ctx, r))
            except BaseException:
                for st in stores:
                    st.close()
                raise
        try:
            forest = self.ctx.eclat_mine_batch(stores, [r.minsup for r in batch])
        finally:
            for s in stores:
                s.close()
        return forest.split_arrays(len(batch))

    def run(self, requests):
        """Serve ``requests`` (any order; arrival times relative to now).
        Returns one ``Outcome`` per request, in request-id order."""
        pending = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        out = {r.rid: Outcome(r.rid, r.arrival) for r in requests}
        queue = deque()
        t0 = time.perf_counter()
        while pending or queue:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                r = pending.popleft()
                if r.footprint() > self.hbm_budget:
                    out[r.rid].rejected = f"footprint {r.footprint()} B exceeds the budget {self.hbm_budget} B"
                    out[r.rid].finish = now
                else:
                    queue.append(r)
            if not queue:
                if pending:
                    time.sleep(max(0.0, pending[0].arrival - (time.perf_counter() - t0)))
                continue
            batch = self._admit(queue)
            start = time.perf_counter() - t0
            results = self._mine(batch)
            end = time.perf_counter() - t0
            k = len(self.batches)
            self.batches.append((start, end, [r.rid for r in batch]))
            for r, res in zip(batch, results):
                o = out[r.rid]
                o.finish, o.patterns, o.batch = end, res, k
        return [out[rid] for rid in sorted(out)]


def mine_one_at_a_time(requests, ctx=None):
    """The same arrival stream, one request at a time in arrival order."""
    ctx = ctx or _hip.context(0)
    order = sorted(requests, key=lambda r: (r.arrival, r.rid))
    out = {}
    t0 = time.perf_counter()
    for r in order:
        now = time.perf_counter() - t0
        if r.arrival > now:
            time.sleep(r.arrival - now)
        bits = _build(ctx, r)
        try:
            pats = _hip.PatternArrays.of(ctx.eclat_mine(bits, r.minsup))
        finally:
            bits.close()
        o = Outcome(r.rid, r.arrival)
        o.finish = time.perf_counter() - t0
        o.patterns = pats
        out[r.rid] = o
    return [out[rid] for rid in sorted(out)]


def latency_summary(outcomes):
    lat = np.array([o.latency for o in outcomes if not o.rejected])
    if lat.size == 0:
        return {"p50_ms": None, "p99_ms": None, "mean_ms": None}
    return {"p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2),
            "mean_ms": round(float(lat.mean()) * 1e3, 2)}
