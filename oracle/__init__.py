"""CPU oracle for the support-counting hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product package (cupami_amd) never imports it.

Two layers:

* ``liboracle.so`` (``cpm_oracle.c``): vertical 64-bit bitset restatements of
  the reference kernels and searches -- see the header of that file for the
  reference lines each function follows.  Built by ``make -C oracle``.
* Pure-Python brute force (tiny inputs), restating the reference tests' own
  oracles: ``tests/test_cu_eclat.py:16-26`` (itemsets), and
  ``tests/test_periodic_miners.py:12-25`` (support + periodicity).
"""

from __future__ import annotations

import ctypes
import math
import os
import subprocess
from itertools import combinations
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None

_u64p = ctypes.POINTER(ctypes.c_uint64)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u32p = ctypes.POINTER(ctypes.c_uint32)


def build():
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = _HERE / "liboracle.so"
        if not path.exists():
            build()
        L = ctypes.CDLL(str(path))
        L.orc_bits_build.argtypes = [_u64p, _u32p, ctypes.c_uint64, ctypes.c_uint64,
                                     ctypes.c_uint64, _u64p]
        L.orc_count_supports.argtypes = [_u64p, ctypes.c_uint64, _u32p, ctypes.c_uint64,
                                         ctypes.c_uint32, _u64p]
        L.orc_eclat.restype = ctypes.c_void_p
        L.orc_eclat.argtypes = [_u64p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64]
        L.orc_apriori.restype = ctypes.c_void_p
        L.orc_apriori.argtypes = [_u64p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64,
                                  ctypes.c_uint64]
        L.orc_pfeclat.restype = ctypes.c_void_p
        L.orc_pfeclat.argtypes = [_u64p, ctypes.c_uint64, ctypes.c_uint32, _i64p,
                                  ctypes.c_uint64, ctypes.c_int64, ctypes.c_uint64,
                                  ctypes.c_int64]
        L.orc_periodic_eval.argtypes = [_u64p, ctypes.c_uint64, _i64p, ctypes.c_uint64,
                                        ctypes.c_int64, _u32p, ctypes.c_uint64,
                                        ctypes.c_uint32, _u64p, _i64p]
        L.orc_patterns_count.restype = ctypes.c_uint64
        L.orc_patterns_count.argtypes = [ctypes.c_void_p]
        L.orc_patterns_items.restype = ctypes.c_uint64
        L.orc_patterns_items.argtypes = [ctypes.c_void_p]
        L.orc_patterns_copy.argtypes = [ctypes.c_void_p, _u32p, _u64p, _u64p, _i64p]
        L.orc_patterns_free.argtypes = [ctypes.c_void_p]
        L.orc_threads.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t)


def words_for(n_transactions: int) -> int:
    return (int(n_transactions) + 63) // 64


def bits_build(rows, cols, n_transactions, n_items):
    """Vertical store ``uint64[n_items, ceil(N/64)]`` (``_bitsets.py:91-100``)."""
    rows = np.ascontiguousarray(rows, dtype=np.uint64)
    cols = np.ascontiguousarray(cols, dtype=np.uint32)
    words = words_for(n_transactions)
    bits = np.zeros((n_items, max(words, 0)), dtype=np.uint64)
    if n_items and words:
        lib().orc_bits_build(_p(rows, _u64p), _p(cols, _u32p), rows.size, n_items, words,
                             _p(bits, _u64p))
    return bits


def count_supports(bits, candidates):
    """Complete-intersection supports (``_bitsets.py:103-113``)."""
    cand = np.ascontiguousarray(candidates, dtype=np.uint32)
    if cand.size == 0:
        return np.zeros(len(cand), dtype=np.uint64)
    bits = np.ascontiguousarray(bits, dtype=np.uint64)
    out = np.zeros(cand.shape[0], dtype=np.uint64)
    lib().orc_count_supports(_p(bits, _u64p), bits.shape[1], _p(cand, _u32p), cand.shape[0],
                             cand.shape[1], _p(out, _u64p))
    return out


def _collect(handle, periodic=False):
    L = lib()
    n = L.orc_patterns_count(handle)
    m = L.orc_patterns_items(handle)
    items = np.zeros(max(m, 1), dtype=np.uint32)
    offsets = np.zeros(n + 1, dtype=np.uint64)
    sup = np.zeros(max(n, 1), dtype=np.uint64)
    per = np.zeros(max(n, 1), dtype=np.int64)
    L.orc_patterns_copy(handle, _p(items, _u32p), _p(offsets, _u64p), _p(sup, _u64p),
                        _p(per, _i64p) if periodic else None)
    L.orc_patterns_free(handle)
    out = {}
    for i in range(n):
        key = tuple(items[offsets[i]:offsets[i + 1]].tolist())
        out[key] = (int(sup[i]), int(per[i])) if periodic else int(sup[i])
    return out


def eclat(bits, minsup):
    """``{column tuple: support}`` of every itemset with support >= minsup
    (``cuEclat.py:168-245``)."""
    bits = np.ascontiguousarray(bits, dtype=np.uint64)
    h = lib().orc_eclat(_p(bits, _u64p), bits.shape[1], bits.shape[0], int(minsup))
    return _collect(h)


def apriori(bits, minsup, batch=4096):
    """clApriori's level-wise search (``clApriori.py:164-196``)."""
    bits = np.ascontiguousarray(bits, dtype=np.uint64)
    h = lib().orc_apriori(_p(bits, _u64p), bits.shape[1], bits.shape[0], int(minsup), int(batch))
    return _collect(h)


def periodic_eval(bits, timestamps, n_transactions, end, candidates):
    """``(supports, periods)`` per candidate (``_temporal.py:28-55``)."""
    cand = np.ascontiguousarray(candidates, dtype=np.uint32)
    bits = np.ascontiguousarray(bits, dtype=np.uint64)
    ts = np.ascontiguousarray(timestamps, dtype=np.int64)
    sup = np.zeros(cand.shape[0], dtype=np.uint64)
    per = np.zeros(cand.shape[0], dtype=np.int64)
    if cand.shape[0]:
        lib().orc_periodic_eval(_p(bits, _u64p), bits.shape[1], _p(ts, _i64p), int(n_transactions),
                                int(end), _p(cand, _u32p), cand.shape[0], cand.shape[1],
                                _p(sup, _u64p), _p(per, _i64p))
    return sup, per


def pfeclat(bits, timestamps, n_transactions, end, minsup, maxper_floor):
    """``{column tuple: (support, period)}`` (``cuPFEclat.py:146-175``)."""
    bits = np.ascontiguousarray(bits, dtype=np.uint64)
    ts = np.ascontiguousarray(timestamps, dtype=np.int64)
    if ts.size == 0:
        ts = np.zeros(1, dtype=np.int64)
    h = lib().orc_pfeclat(_p(bits, _u64p), bits.shape[1], bits.shape[0], _p(ts, _i64p),
                          int(n_transactions), int(end), int(minsup), int(maxper_floor))
    return _collect(h, periodic=True)


def threads():
    return lib().orc_threads()


def set_threads(n):
    lib().orc_set_threads(int(n))


# ---------------------------------------------------------------------------
# pure-Python brute force (tiny inputs only)

def clean_basket(row):
    """``cuApriori._cleanTransaction`` (``cuApriori.py:444-450``): strip, drop
    empty labels and repeats, keep first-seen order."""
    items = []
    for value in row:
        item = str(value).strip()
        if item and item not in items:
            items.append(item)
    return tuple(items)


def csr_to_coo(offsets, ids, keep_empty=False):
    """A basket file's arrays -> ``(labels, rows, cols, n_rows)``, the BIN
    branch of ``cuApriori._getDevicePairs`` (``cuApriori.py:381-396``): the
    distinct ids, their labels sorted as strings, each entry's basket renumbered
    over the non-empty baskets (``activeRows``, ``:391``) and its column from a
    search of the id among the distinct ones (``:394-395``).  ``keep_empty``
    keeps every basket as a row (the temporal readers, ``_temporal.py:125-139``).
    """
    offsets = np.asarray(offsets, dtype=np.uint64)
    ids = np.asarray(ids, dtype=np.uint32)
    unique = np.unique(ids)
    labels = sorted(str(int(v)) for v in unique)
    index = {label: i for i, label in enumerate(labels)}
    remap = np.array([index[str(int(v))] for v in unique], dtype=np.uint32)
    lengths = np.diff(offsets.astype(np.int64))
    kept = np.ones(lengths.size, bool) if keep_empty else lengths > 0
    row_of_basket = np.cumsum(kept) - 1
    rows = np.repeat(row_of_basket.astype(np.uint64), lengths)
    cols = remap[np.searchsorted(unique, ids)] if ids.size else np.zeros(0, np.uint32)
    return labels, rows, cols.astype(np.uint32), int(kept.sum())


def support_threshold(minSup, n):
    """``cuApriori._getSupportCount`` (``cuApriori.py:452-456``)."""
    if isinstance(minSup, float) and 0 < minSup <= 1:
        return math.ceil(minSup * n)
    return int(minSup)


def brute_force(rows, threshold):
    """Every itemset with support >= threshold, by subset tests over the
    baskets (``tests/test_cu_eclat.py:16-26``)."""
    baskets = [set(str(i).strip() for i in row if str(i).strip()) for row in rows]
    baskets = [b for b in baskets if b]
    items = sorted(set().union(*baskets)) if baskets else []
    out = {}
    for length in range(1, len(items) + 1):
        for pattern in combinations(items, length):
            count = sum(set(pattern).issubset(b) for b in baskets)
            if count >= threshold:
                out["\t".join(pattern)] = count
    return out


def brute_force_rows(rows, threshold):
    """Per-basket subset counting (``tests/test_transaction_formats.py:15-23``);
    independent of the item universe, cost ~ sum of 2**len(basket)."""
    counts = {}
    for row in rows:
        items = sorted({str(i).strip() for i in row if str(i).strip()})
        for size in range(1, len(items) + 1):
            for pattern in combinations(items, size):
                key = "\t".join(pattern)
                counts[key] = counts.get(key, 0) + 1
    return {k: c for k, c in counts.items() if c >= threshold}


def brute_periodic(rows, support, period, end=None):
    """Support and periodicity of every itemset, by enumeration
    (``tests/test_periodic_miners.py:12-25``): the period is the largest gap
    of ``[0, ts..., end]``."""
    labels = sorted({str(item) for row in rows for item in row[1:]})
    end = max((row[0] for row in rows), default=0) if end is None else end
    result = {}
    for width in range(1, len(labels) + 1):
        for pattern in combinations(labels, width):
            ts = sorted(row[0] for row in rows if set(pattern) <= set(map(str, row[1:])))
            if not ts:
                continue
            points = [0, *ts, end]
            gap = max(b - a for a, b in zip(points, points[1:]))
            if len(ts) >= support and gap <= period:
                result["\t".join(pattern)] = (len(ts), gap)
    return result


def labelled(patterns, labels):
    """Column-tuple keys -> ``"\\t".join(labels)`` keys."""
    return {"\t".join(labels[c] for c in key): v for key, v in patterns.items()}
