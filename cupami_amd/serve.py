"""Serving loop: mining requests on one GPU with continuous batching,
admission by HBM footprint and device-resident stores.

A request is one database (pairs in host memory, or a BIN file's arrays)
and an absolute minimum support.  Requests that name the same database
(``Request.key``) share its vertical store, which stays resident in HBM
after its first build.  The device work of a request is the reference's
miner on that store (``cuEclat.mine``, ``cuEclat.py:168-245``); a request's
itemsets are the same whichever way it is served.

Three kinds of worker overlap on the device, each with its own HIP stream
(streams of the default priority):

* the dispatcher (the caller's thread) moves requests in as they arrive.  A
  request whose store is resident (or being built) joins it at once; one
  that needs a new store is admitted, in arrival order, as soon as its
  footprint fits the HBM budget -- stores no request holds are evicted,
  least recently used first, to make room.  A request whose footprint
  exceeds the whole budget is rejected;
* a builder thread builds admitted stores on its own context (and frees
  evicted ones, as the context that built them);
* ``slots`` slot threads, each with its own context, take every ready
  request waiting (up to ``max_batch`` requests and ``max_batch_entries``
  store entries) as one micro-batch, mine it -- one search for one request,
  one batched search (``cpm_eclat_mine_batch``) for several -- and hand each
  request its result.

So a request that arrives while others are on the GPU does not wait for
them: it is admitted as memory allows, its store is built beside the
running searches, and it is mined by the next free slot, its work
overlapping the other slots' on the device.

``mine_one_at_a_time`` serves the same arrival stream the plain way -- one
request at a time, its store built, mined and freed -- and
``mine_sequential_resident`` one at a time on stores that stay resident
(the residency without the overlap).
"""

from __future__ import annotations

import threading
import time
from collections import OrderedDict, deque
from collections.abc import Mapping
from dataclasses import dataclass, field

import numpy as np

from . import _hip


def row_stride_words(n_transactions):
    words = (n_transactions + 63) // 64
    return (words + 63) // 64 * 64


@dataclass
class Request:
    """A database to mine, as pairs (``rows``, ``cols``: item columns in
    label string order) or in the BIN file layout (``offsets`` u64[T+1] and
    raw item ``ids`` u32 up to ``universe``, ``rows``/``cols`` None): the
    device builds the latter's store and labels itself.  Its columns are the
    distinct ids in string order, the numbering ``label_columns`` gives the
    pairs.  Requests with equal ``key`` (not None) mine the same database
    and share its resident store; ``key`` None is a database of its own."""
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
    key: object = None

    @property
    def entries(self):
        return int(self.ids.size if self.ids is not None else self.rows.size)

    @property
    def store_key(self):
        return ("request", self.rid) if self.key is None else ("db", self.key)

    def store_bytes(self):
        return self.n_items * row_stride_words(self.n_transactions) * 8

    def footprint(self):
        """Device bytes of its resident store plus the staging of its
        input while the store is built."""
        staged = (4 * self.entries + 8 * (self.n_transactions + 1) if self.ids is not None
                  else 12 * self.entries)
        return self.store_bytes() + staged


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


class _Store:
    """A resident (or building) store and the requests holding it."""

    def __init__(self, key, nbytes):
        self.key = key
        self.nbytes = nbytes
        self.bits = None
        self.ready = False
        self.error = None
        self.holders = 0          # admitted requests not yet finished
        self.waiting = []         # requests waiting for the build


class Server:
    """Continuous-batching server (see the module docstring)."""

    def __init__(self, device=0, hbm_budget=8 << 30, slots=3, max_batch=64, max_batch_entries=16 << 20,
                 contexts=None):
        """``contexts(device)`` makes each worker's context (default: a new
        ``_hip.Context`` -- its own HIP stream -- per worker)."""
        self.device = int(device)
        self.hbm_budget = int(hbm_budget)
        self.slots = max(1, int(slots))
        self.max_batch = max(1, int(max_batch))
        self.max_batch_entries = int(max_batch_entries)
        make = contexts or _hip.Context
        self.builder_ctx = make(self.device)
        self.slot_ctx = [make(self.device) for _ in range(self.slots)]
        self._stores = OrderedDict()   # key -> _Store, least recently used first
        self._used = 0
        self.batches = []              # (slot, start, end, request ids)
        self.stats = {}

    def evict_all(self):
        """Free every resident store (between runs)."""
        for st in self._stores.values():
            if st.bits is not None:
                st.bits.close()
        self._stores.clear()
        self._used = 0

    def close(self):
        """Free the resident stores and the contexts."""
        for st in self._stores.values():
            if st.bits is not None:
                st.bits.close()
        self._stores.clear()
        self._used = 0
        for c in [self.builder_ctx, *self.slot_ctx]:
            c.close()

    # ---- the run ---------------------------------------------------------------
    def run(self, requests):
        """Serve ``requests`` (arrival times relative to now).  Returns one
        ``Outcome`` per request, in request-id order.  Resident stores stay
        for the next run."""
        pending = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        out = {r.rid: Outcome(r.rid, r.arrival) for r in requests}
        cv = threading.Condition()
        admit = deque()          # arrived, needing a new store, in arrival order
        ready = deque()          # (request, store) ready to mine
        builds = deque()         # ("build", store, request) / ("free", bits)
        left = [len(requests)]
        stop = [False]
        failure = []
        counters = {"builds": 0, "evictions": 0, "resident_hits": 0, "busy_max": 0}
        busy = [0]
        self.batches = []
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        def finish(r, patterns=None, rejected=""):
            o = out[r.rid]
            o.finish = now()
            if patterns is not None:
                o.patterns = patterns
            o.rejected = rejected
            left[0] -= 1

        def release(st):
            st.holders -= 1

        def builder():
            while True:
                with cv:
                    while not builds and not stop[0]:
                        cv.wait()
                    if not builds:
                        return
                    job = builds.popleft()
                if job[0] == "free":
                    job[1].close()
                    continue
                _, st, r = job
                try:
                    bits = _build(self.builder_ctx, r)
                    self.builder_ctx.synchronize()
                except Exception as error:  # reported on the request
                    with cv:
                        st.error = error
                        waiting, st.waiting = st.waiting, []
                        for w in waiting:
                            release(st)
                            finish(w, rejected=f"store build failed: {error}")
                        self._stores.pop(st.key, None)
                        self._used -= st.nbytes
                        cv.notify_all()
                    continue
                with cv:
                    st.bits, st.ready = bits, True
                    for w in st.waiting:
                        ready.append((w, st))
                    st.waiting = []
                    counters["builds"] += 1
                    cv.notify_all()

        def slot(i):
            ctx = self.slot_ctx[i]
            while True:
                with cv:
                    while not ready and not stop[0]:
                        cv.wait()
                    if not ready:
                        return
                    batch = [ready.popleft()]
                    entries = batch[0][0].entries
                    while ready and len(batch) < self.max_batch and \
                            entries + ready[0][0].entries <= self.max_batch_entries:
                        entries += ready[0][0].entries
                        batch.append(ready.popleft())
                    busy[0] += 1
                    counters["busy_max"] = max(counters["busy_max"], busy[0])
                start = now()
                try:
                    if len(batch) == 1:
                        r, st = batch[0]
                        results = [_hip.PatternArrays.of(ctx.eclat_mine(st.bits, r.minsup))]
                    else:
                        res = ctx.eclat_mine_batch([st.bits for _, st in batch], [r.minsup for r, _ in batch])
                        results = res.split_arrays(len(batch))
                except Exception as error:
                    with cv:
                        failure.append(error)
                        stop[0] = True
                        busy[0] -= 1
                        cv.notify_all()
                    return
                end = now()
                with cv:
                    busy[0] -= 1
                    k = len(self.batches)
                    self.batches.append((i, start, end, [r.rid for r, _ in batch]))
                    for (r, st), pats in zip(batch, results):
                        finish(r, pats)
                        out[r.rid].batch = k
                        release(st)
                    cv.notify_all()

        def try_admit():
            """Admit waiting requests that need a new store, in arrival
            order, while the budget allows (evicting idle stores)."""
            while admit:
                r = admit[0]
                key = r.store_key
                st = self._stores.get(key)
                if st is not None:   # another request admitted it meanwhile
                    admit.popleft()
                    join(r, st)
                    continue
                need = r.footprint()
                if self._used + need > self.hbm_budget:
                    for k in list(self._stores):
                        if self._used + need <= self.hbm_budget:
                            break
                        old = self._stores[k]
                        if old.holders == 0 and old.ready:
                            del self._stores[k]
                            self._used -= old.nbytes
                            builds.append(("free", old.bits))
                            counters["evictions"] += 1
                    if self._used + need > self.hbm_budget:
                        return   # wait for running requests to release stores
                admit.popleft()
                st = _Store(key, need)
                self._stores[key] = st
                self._used += need
                st.holders = 1
                st.waiting.append(r)
                builds.append(("build", st, r))
                cv.notify_all()

        def join(r, st):
            self._stores.move_to_end(st.key)
            st.holders += 1
            if st.ready:
                counters["resident_hits"] += 1
                ready.append((r, st))
            else:
                st.waiting.append(r)
            cv.notify_all()

        threads = [threading.Thread(target=builder, daemon=True)]
        threads += [threading.Thread(target=slot, args=(i,), daemon=True) for i in range(self.slots)]
        for t in threads:
            t.start()
        try:
            with cv:
                while left[0] > 0 and not failure:
                    t = now()
                    while pending and pending[0].arrival <= t:
                        r = pending.popleft()
                        if r.footprint() > self.hbm_budget:
                            finish(r, rejected=f"footprint {r.footprint()} B exceeds the budget "
                                               f"{self.hbm_budget} B")
                            continue
                        st = self._stores.get(r.store_key)
                        if st is not None:   # resident or being built: needs no new memory
                            join(r, st)
                        else:
                            admit.append(r)
                    try_admit()
                    if left[0] == 0:
                        break
                    timeout = (pending[0].arrival - now()) if pending else None
                    if timeout is None or timeout > 0:
                        cv.wait(timeout=min(timeout, 0.05) if timeout is not None else 0.05)
        finally:
            with cv:
                stop[0] = True
                cv.notify_all()
            for t in threads:
                t.join()
            # the stores of this run stay resident, built and idle
            for key in [k for k, st in self._stores.items() if not st.ready]:
                del self._stores[key]
        if failure:
            raise failure[0]
        self.stats = dict(counters, resident_bytes=self._used, resident_stores=len(self._stores),
                          batches=len(self.batches),
                          max_batch=max((len(b[3]) for b in self.batches), default=0))
        return [out[rid] for rid in sorted(out)]


def mine_one_at_a_time(requests, ctx=None):
    """The same arrival stream the plain way: one request at a time in
    arrival order, its store built, mined and freed (what a new miner object
    per request does)."""
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


def mine_sequential_resident(requests, ctx=None):
    """One request at a time in arrival order on stores kept resident
    (built on first use, never evicted): the residency without the overlap."""
    ctx = ctx or _hip.context(0)
    order = sorted(requests, key=lambda r: (r.arrival, r.rid))
    stores, out = {}, {}
    t0 = time.perf_counter()
    try:
        for r in order:
            now = time.perf_counter() - t0
            if r.arrival > now:
                time.sleep(r.arrival - now)
            bits = stores.get(r.store_key)
            if bits is None:
                bits = stores[r.store_key] = _build(ctx, r)
            o = Outcome(r.rid, r.arrival)
            o.patterns = _hip.PatternArrays.of(ctx.eclat_mine(bits, r.minsup))
            o.finish = time.perf_counter() - t0
            out[r.rid] = o
    finally:
        for b in stores.values():
            b.close()
    return [out[rid] for rid in sorted(out)]


def latency_summary(outcomes):
    lat = np.array([o.latency for o in outcomes if not o.rejected])
    if lat.size == 0:
        return {"p50_ms": None, "p99_ms": None, "mean_ms": None}
    return {"p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2),
            "mean_ms": round(float(lat.mean()) * 1e3, 2)}
