"""FP-tree oracle pinned: the numpy restatement of the reference's level-wise
build equals the trie built from the definition, the reference tests' known
answers hold, and FP-growth finds exactly the brute-force itemsets."""

import random

import numpy as np
import pytest

import oracle
from oracle import fptree
from helpers import coo


def same_tree(a, b):
    assert (a is None) == (b is None)
    if a is None:
        return
    assert a.parents.tolist() == b.parents.tolist()
    assert a.items.tolist() == b.items.tolist()
    assert a.counts.tolist() == b.counts.tolist()
    assert a.header == b.header
    assert a.single_path == b.single_path and a.depth == b.depth


@pytest.mark.parametrize("seed", range(8))
def test_levelwise_build_equals_trie(seed):
    rng = random.Random(seed)
    n, items = 40 + seed * 7, 9
    rows, cols = [], []
    for r in range(n):
        for c in rng.sample(range(items), rng.randrange(items)):
            rows.append(r)
            cols.append(c)
    weights = np.array([rng.randrange(1, 5) for _ in range(n)], np.uint64)
    for ms in (1, 3, 8, 25):
        same_tree(fptree.build_tree(rows, cols, weights, items, ms),
                  fptree.trie_tree(rows, cols, weights, items, ms))


def test_reference_kats():
    # tests/test_cu_fpgrowth.py:71-80: a one-item conditional base
    tree = fptree.build_tree([0, 0, 1], [0, 1499, 0], np.array([3, 2], np.uint64), 1500, 1)
    child = fptree.conditional_tree(tree, 1499, 1500, 1)
    assert child.items.tolist() == [-1, 0] and child.counts.tolist() == [0, 3]
    # tests/test_cu_fpgrowth.py:92-102: exact uint64 above 2**53
    big = 2**53 + 7
    tree = fptree.build_tree([0, 0, 1], [0, 1, 0], np.array([big, big + 1], np.uint64), 2, 1)
    assert dict(tree.header) == {0: 2 * big + 1, 1: big}
    cond = fptree.conditional_tree(tree, 1, 2, 1)
    assert dict(cond.header) == {0: big}


@pytest.mark.parametrize("seed", range(5))
def test_fpgrowth_equals_brute_force(seed):
    rng = random.Random(seed)
    rows = [rng.sample(range(7), rng.randrange(8)) for _ in range(33 + seed)]
    labels, rr, cc, n = coo(rows)
    for ms in (1, 3, 7, 20):
        got = fptree.fpgrowth(rr, cc, n, len(labels), ms)
        assert oracle.labelled(got, labels) == oracle.brute_force(rows, ms)


def test_unique_pairs():
    r, c = fptree.unique_pairs([3, 1, 1, 3, 0], [2, 5, 5, 2, 9])
    assert r.tolist() == [0, 1, 3] and c.tolist() == [9, 5, 2]
