"""Shared test helpers: stores and label maps built the reference's way."""

import json
from pathlib import Path

import numpy as np

import oracle

GOLDEN = Path(__file__).resolve().parent / "golden"


def kats():
    return json.loads((GOLDEN / "reference_kats.json").read_text())


def coo(rows):
    """Clean baskets like the reference, drop empty ones, order labels as
    strings (cuApriori.py:397-404); returns (labels, rows, cols, n)."""
    baskets = [oracle.clean_basket(r) for r in rows]
    baskets = [b for b in baskets if b]
    labels = sorted({i for b in baskets for i in b})
    index = {lab: i for i, lab in enumerate(labels)}
    rr, cc = [], []
    for r, b in enumerate(baskets):
        for item in b:
            rr.append(r)
            cc.append(index[item])
    return labels, np.asarray(rr, np.uint64), np.asarray(cc, np.uint32), len(baskets)


def temporal_coo(rows, end=None):
    """Temporal rows [ts, *items] -> (labels, rows, cols, n, ts, end); empty
    baskets are kept and rows are stably sorted by timestamp."""
    recs = [(int(r[0]), oracle.clean_basket(r[1:])) for r in rows if len(r)]
    recs.sort(key=lambda x: x[0])
    labels = sorted({i for _, b in recs for i in b})
    index = {lab: i for i, lab in enumerate(labels)}
    rr, cc = [], []
    for r, (_, b) in enumerate(recs):
        for item in b:
            rr.append(r)
            cc.append(index[item])
    ts = np.asarray([t for t, _ in recs], np.int64)
    last = int(ts[-1]) if len(ts) else 0
    return (labels, np.asarray(rr, np.uint64), np.asarray(cc, np.uint32), len(recs), ts,
            last if end is None else end)


def random_rows(rng, n, n_items, max_len):
    return [rng.sample(range(n_items), rng.randrange(max_len + 1)) for _ in range(n)]


def csr(rows):
    """Baskets (lists of int ids, repeats and empty baskets kept) -> the
    ``offsets`` u64[n+1] and ``ids`` u32[m] arrays of a CUPAMI01 file."""
    offsets = np.zeros(len(rows) + 1, np.uint64)
    offsets[1:] = np.cumsum([len(r) for r in rows], dtype=np.uint64) if rows else []
    ids = np.asarray([i for r in rows for i in r], np.uint32)
    return offsets, ids


def write_bin(path, rows, universe=None):
    """A CUPAMI01 file (docs/transaction-formats.md: magic, N, M, U, offsets,
    ids; little endian) holding the baskets as given."""
    import struct
    offsets, ids = csr(rows)
    if universe is None:
        universe = max(1, int(ids.max()) if ids.size else 1)
    Path(path).write_bytes(struct.pack("<8sQQQ", b"CUPAMI01", len(rows), ids.size, universe)
                           + offsets.astype("<u8").tobytes() + ids.astype("<u4").tobytes())


def write_temporal_bin(path, rows, universe=None):
    """A CUPAMIT1 file (timestamps i64[n], offsets u64[n+1], ids u32[m] after
    the CUPAMI01-style header) holding rows ``[ts, *items]`` as given."""
    import struct
    ts = np.asarray([r[0] for r in rows], "<i8")
    offsets, ids = csr([list(r[1:]) for r in rows])
    if universe is None:
        universe = max(1, int(ids.max()) if ids.size else 1)
    Path(path).write_bytes(struct.pack("<8sQQQ", b"CUPAMIT1", len(rows), ids.size, universe) + ts.tobytes()
                           + offsets.astype("<u8").tobytes() + ids.astype("<u4").tobytes())
