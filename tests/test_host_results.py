"""Host-side result handling without a device: a batched forest cut per
database (Patterns.split_arrays) against the dict-building split, the lazy
PatternArrays mapping, and the host pairs of a BIN file's arrays."""

import random

import numpy as np

import oracle
from cupami_amd import _hip
from cupami_amd.frequentPattern.cuApriori import binary_pairs
from helpers import csr


def synthetic_forest(seed, n, n_db, n_items):
    """A batched forest as the search returns it: parents precede children
    and belong to the same database; itemsets are unique per database."""
    rng = random.Random(seed)
    parent = np.full(n, _hip.NO_PARENT, np.uint32)
    item = np.zeros(n, np.uint32)
    db = np.zeros(n, np.uint32)
    sets, seen, i = [None] * n, set(), 0
    while i < n:
        d = rng.randrange(n_db)
        cands = [j for j in range(max(0, i - 40), i) if db[j] == d]
        p = rng.choice(cands) if cands and rng.random() < 0.8 else None
        c = rng.randrange(n_items)
        key = (d, (sets[p] if p is not None else ()) + (c,))
        if key in seen:
            continue
        seen.add(key)
        db[i], item[i] = d, c
        parent[i] = _hip.NO_PARENT if p is None else p
        sets[i] = key[1]
        i += 1
    P = _hip.Patterns.__new__(_hip.Patterns)
    P.n, P.periodic = n, False
    P.parent, P.item, P.db = parent, item, db
    P.support = np.asarray([rng.randrange(1, 1000) for _ in range(n)], np.uint64)
    return P


def test_split_arrays_matches_split():
    for seed in range(5):
        P = synthetic_forest(seed, 3000, 1 + seed * 3, 40)
        n_db = 1 + seed * 3
        views, dicts = P.split_arrays(n_db), P.split(n_db)
        assert [len(v) for v in views] == [len(d) for d in dicts]
        for v, d in zip(views, dicts):
            assert v == d and dict(v.items()) == d


def test_pattern_arrays_mapping():
    P = synthetic_forest(9, 200, 1, 12)
    m = _hip.PatternArrays.of(P)
    assert len(m) == 200 and bool(m)
    expect = P.to_dict()
    assert m == expect and sorted(m) == sorted(expect)
    key = next(iter(expect))
    assert m[key] == expect[key] and key in m
    assert m != {**expect, (99,): 1}
    assert _hip.PatternArrays(np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.uint64)) == {}


def test_binary_pairs_keep_empty():
    rows = [[3, 1], [], [10, 2, 2], [], [1]]
    offsets, ids = csr(rows)
    labels, rr, cc = binary_pairs(offsets, ids)
    l2, r2, c2, n2 = oracle.csr_to_coo(offsets, ids)
    assert labels == l2 == ["1", "10", "2", "3"]
    assert np.array_equal(rr, r2) and np.array_equal(cc, c2) and n2 == 3
    labels, rr, cc = binary_pairs(offsets, ids, keep_empty=True)
    l3, r3, c3, n3 = oracle.csr_to_coo(offsets, ids, keep_empty=True)
    assert labels == l3 and np.array_equal(rr, r3) and np.array_equal(cc, c3) and n3 == 5
