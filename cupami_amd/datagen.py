"""Seeded synthetic transaction databases for the support-counting workloads.

Every generator draws from ``numpy.random.Generator(PCG64(seed))`` and returns
a COO pair list: ``rows`` (uint64 transaction index) and ``items`` (uint32 item
id, 1-based like the reference generators, ``TransactionalDatabase.py:104``).

Item popularity is Zipf over a seeded permutation of the ids
(``p(rank r) ~ r**-s``; ``s = 0`` is uniform).  A basket of length ``L`` is a
weighted sample *without replacement*: items are drawn one at a time from the
popularity distribution restricted to the items not yet chosen.  That is done
by rejection (redraw on a repeat), which has exactly the same law as successive
sampling (Plackett-Luce / Efraimidis-Spirakis), without touching all items per
basket.  ``L = max(1, Poisson(lam))``, capped at the item count.
"""

from __future__ import annotations

import hashlib
from dataclasses import dataclass

import numpy as np


@dataclass
class Database:
    rows: np.ndarray          # uint64[nnz], transaction index, non-decreasing
    items: np.ndarray         # uint32[nnz], item id in 1..n_items
    n_transactions: int
    n_items: int
    timestamps: np.ndarray | None = None   # int64[n_transactions], temporal only

    def sha256(self) -> str:
        h = hashlib.sha256()
        h.update(np.ascontiguousarray(self.rows, dtype="<u8").tobytes())
        h.update(np.ascontiguousarray(self.items, dtype="<u4").tobytes())
        if self.timestamps is not None:
            h.update(np.ascontiguousarray(self.timestamps, dtype="<i8").tobytes())
        return h.hexdigest()

    def baskets(self):
        """Python baskets (lists of int ids), for small databases only."""
        out = [[] for _ in range(self.n_transactions)]
        for r, i in zip(self.rows.tolist(), self.items.tolist()):
            out[r].append(i)
        return out


def _sample_baskets(rng, n, n_items, lam, zipf_s, chunk=200_000):
    ranks = np.arange(1, n_items + 1, dtype=np.float64)
    weights = ranks ** (-float(zipf_s))
    cdf = np.cumsum(weights / weights.sum())
    perm = rng.permutation(n_items).astype(np.uint32) + 1      # rank -> item id
    lengths = np.minimum(np.maximum(1, rng.poisson(lam, n)), n_items).astype(np.int64)
    row_parts, item_parts = [], []
    for start in range(0, n, chunk):
        stop = min(n, start + chunk)
        m = stop - start
        need = lengths[start:stop]
        width = int(need.max())
        out = np.full((m, width), -1, dtype=np.int64)   # -1: slot not filled yet
        count = np.zeros(m, dtype=np.int64)
        active = np.arange(m)
        while active.size:
            draw = np.searchsorted(cdf, rng.random(active.size), side="right")
            np.minimum(draw, n_items - 1, out=draw)
            # a draw is new unless the basket already holds it
            fresh = ~(out[active] == draw[:, None]).any(axis=1)
            a, d = active[fresh], draw[fresh]
            out[a, count[a]] = d
            count[a] += 1
            active = active[count[active] < need[active]]
        mask = np.arange(width)[None, :] < need[:, None]
        rr = np.broadcast_to(np.arange(start, stop, dtype=np.uint64)[:, None], (m, width))
        row_parts.append(rr[mask])
        item_parts.append(perm[out[mask]])
    rows = np.concatenate(row_parts) if row_parts else np.zeros(0, np.uint64)
    items = np.concatenate(item_parts) if item_parts else np.zeros(0, np.uint32)
    return rows.astype(np.uint64), items.astype(np.uint32)


def transactional(n, n_items, lam, zipf_s, seed) -> Database:
    """Binary baskets: ``n`` transactions, Poisson(lam) unique items each."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rows, items = _sample_baskets(rng, n, n_items, lam, zipf_s)
    return Database(rows, items, n, n_items)


def temporal(n, n_items, lam, zipf_s, seed, gap_low=1, gap_high=3) -> Database:
    """Temporal baskets with strictly increasing timestamps (gaps uniform in
    ``gap_low..gap_high``)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ts = np.cumsum(rng.integers(gap_low, gap_high + 1, n, dtype=np.int64))
    rows, items = _sample_baskets(rng, n, n_items, lam, zipf_s)
    return Database(rows, items, n, n_items, ts.astype(np.int64))


# Named workloads, with the seeds the measurement plan fixes.
def small_uniform(seed=101):
    """10,000 baskets over 50 items, Poisson(6), uniform popularity."""
    return transactional(10_000, 50, 6.0, 0.0, seed)


def eclat_1m_1k(seed=202):
    """1,000,000 baskets over 1,000 items, Poisson(10), Zipf s=1.1."""
    return transactional(1_000_000, 1000, 10.0, 1.1, seed)


def temporal_2m_500(seed=404):
    """2,000,000 timestamped baskets over 500 items, Poisson(8), Zipf s=1.05,
    timestamp gaps uniform in 1..3."""
    return temporal(2_000_000, 500, 8.0, 1.05, seed)


def batch_256(seed=303, n_db=256):
    """Many small databases for batched mining: ``n_db`` databases with
    T log-uniform in [10^3, 2*10^5] baskets, I uniform in [20, 500] items,
    Poisson(lam) lengths with lam uniform in [4, 12], Zipf s=1.0, and a
    relative minimum support log-uniform in [0.002, 0.05] (an absolute count
    ceil(minSup * T), the reference's rule for a fraction).  The parameters
    come from ``seed``; database i is drawn with seed ``seed * 1000 + i``.
    Returns a list of ``(Database, minsup_count)``."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = np.exp(rng.uniform(np.log(1e3), np.log(2e5), n_db)).round().astype(np.int64)
    items = rng.integers(20, 501, n_db)
    lam = rng.uniform(4.0, 12.0, n_db)
    frac = np.exp(rng.uniform(np.log(0.002), np.log(0.05), n_db))
    out = []
    for i in range(n_db):
        db = transactional(int(t[i]), int(items[i]), float(lam[i]), 1.0, seed * 1000 + i)
        out.append((db, int(np.ceil(frac[i] * t[i]))))
    return out


def serve_requests(n_requests, qps, seed=505):
    """Mining requests for the serving loop: T log-uniform in [10^4, 2*10^6]
    baskets, I uniform in [50, 2000] items, lam uniform in [4, 12], Zipf s
    uniform in [0.9, 1.2], relative minimum support log-uniform in
    [0.001, 0.05] (absolute ceil(minSup * T)), Poisson arrivals at ``qps``.
    The parameters and arrival times come from ``seed``; request i is drawn
    with seed ``seed * 1000 + i``.  Returns a list of
    ``(arrival_seconds, Database, minsup_count)``."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = np.exp(rng.uniform(np.log(1e4), np.log(2e6), n_requests)).round().astype(np.int64)
    items = rng.integers(50, 2001, n_requests)
    lam = rng.uniform(4.0, 12.0, n_requests)
    zipf = rng.uniform(0.9, 1.2, n_requests)
    frac = np.exp(rng.uniform(np.log(0.001), np.log(0.05), n_requests))
    arrival = np.cumsum(rng.exponential(1.0 / qps, n_requests))
    out = []
    for i in range(n_requests):
        db = transactional(int(t[i]), int(items[i]), float(lam[i]), float(zipf[i]), seed * 1000 + i)
        out.append((float(arrival[i]), db, int(np.ceil(frac[i] * t[i]))))
    return out


@dataclass
class DbSpec:
    """Parameters of one seeded database (generated by ``build``)."""
    key: int
    n_transactions: int
    n_items: int
    lam: float
    zipf_s: float
    seed: int

    def build(self) -> Database:
        return transactional(self.n_transactions, self.n_items, self.lam, self.zipf_s, self.seed)


def serve_pool(n_db=48, seed=505):
    """The databases served requests are drawn against: ``n_db`` specs with
    T log-uniform in [10^4, 2*10^6] baskets, I uniform in [50, 2000] items,
    lam uniform in [4, 12], Zipf s uniform in [0.9, 1.2]; database i is drawn
    with seed ``seed * 1000 + i``.  Generate them with ``build_pool``."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = np.exp(rng.uniform(np.log(1e4), np.log(2e6), n_db)).round().astype(np.int64)
    items = rng.integers(50, 2001, n_db)
    lam = rng.uniform(4.0, 12.0, n_db)
    zipf = rng.uniform(0.9, 1.2, n_db)
    return [DbSpec(i, int(t[i]), int(items[i]), float(lam[i]), float(zipf[i]), seed * 1000 + i) for i in range(n_db)]


def _build_spec(spec):
    return spec.build()


def build_pool(specs, workers=None):
    """Databases of ``specs`` (in order), generated in worker processes."""
    import multiprocessing as mp
    import os
    workers = workers or min(len(specs), max(1, min(16, os.cpu_count() or 1)))
    if workers <= 1 or len(specs) <= 1:
        return [s.build() for s in specs]
    # largest first, so the long ones do not finish last
    order = sorted(range(len(specs)), key=lambda i: -specs[i].n_transactions * specs[i].n_items)
    with mp.get_context("fork").Pool(workers) as pool:
        built = pool.map(_build_spec, [specs[i] for i in order], chunksize=1)
    out = [None] * len(specs)
    for i, db in zip(order, built):
        out[i] = db
    return out


def serve_stream(n_requests, qps, n_db, seed=505):
    """Served requests over a pool of ``n_db`` databases: Poisson arrivals at
    ``qps``; each request takes a database uniformly at random and a relative
    minimum support log-uniform in [0.001, 0.05].  Returns a list of
    ``(arrival_seconds, database_index, minSup_fraction)``."""
    rng = np.random.Generator(np.random.PCG64(seed * 7919 + 1))
    arrival = np.cumsum(rng.exponential(1.0 / qps, n_requests))
    db = rng.integers(0, n_db, n_requests)
    frac = np.exp(rng.uniform(np.log(0.001), np.log(0.05), n_requests))
    return [(float(a), int(d), float(f)) for a, d, f in zip(arrival, db, frac)]


def label_columns(item_ids: np.ndarray, n_items_hint=None):
    """Map numeric item ids to column indexes in *string* label order.

    The reference orders item labels as strings (``cuApriori.py:397``), so
    ``"10" < "2"``.  Returns ``(labels, columns)`` with ``labels`` the sorted
    distinct label strings and ``columns`` uint32 indexes into it.
    """
    unique = np.unique(item_ids)
    labels = sorted(str(int(v)) for v in unique)
    index = {lab: i for i, lab in enumerate(labels)}
    remap = np.array([index[str(int(v))] for v in unique], dtype=np.uint32)
    columns = remap[np.searchsorted(unique, item_ids)]
    return labels, columns.astype(np.uint32)
