"""Device FP-tree (build, conditional projection, FP-growth) against the
numpy restatement of the reference's arrays, bit-exact."""

import random

import numpy as np
import pytest

import oracle
from oracle import fptree
from cupami_amd import _hip, datagen
from cupami_amd.frequentPattern import _fptree
from helpers import coo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return _hip.context(0)


def same(dev, ref):
    assert (dev is None) == (ref is None)
    if dev is None:
        return
    assert dev.parents.tolist() == ref.parents.tolist()
    assert dev.items.tolist() == ref.items.tolist()
    assert dev.counts.tolist() == ref.counts.tolist()
    assert dev.header == ref.header
    assert dev.single_path == ref.single_path
    assert dev.depth == ref.depth


def test_unique_pairs(ctx):
    rng = np.random.default_rng(1)
    rows = rng.integers(0, 500, 20000).astype(np.uint64)
    cols = rng.integers(0, 40, 20000).astype(np.uint32)
    r, c = _fptree.unique_pairs(ctx, rows, cols)
    er, ec = fptree.unique_pairs(rows, cols)
    assert r.tolist() == er.tolist() and c.tolist() == ec.tolist()


@pytest.mark.parametrize("seed", range(6))
def test_build_and_conditionals(ctx, seed):
    rng = random.Random(seed)
    n, items = 60 + 50 * seed, 7 + seed
    rows, cols = [], []
    for r in range(n):
        for c in rng.sample(range(items), rng.randrange(items)):
            rows.append(r)
            cols.append(c)
    rows = np.array(rows, np.uint64)
    cols = np.array(cols, np.uint32)
    weights = np.array([rng.randrange(1, 9) for _ in range(n)], np.uint64)
    for ms in (1, 5, 40):
        ref = fptree.build_tree(rows, cols, weights, items, ms)
        dev = _fptree.build_tree(ctx, rows, cols, weights, items, ms)
        same(dev, ref)
        if ref is None:
            continue
        for item, _ in ref.header:
            same(_fptree.conditional_tree(ctx, dev, item, items, ms),
                 fptree.conditional_tree(ref, item, items, ms))


def test_reference_kats(ctx):
    tree = _fptree.build_tree(ctx, np.array([0, 0, 1], np.uint64), np.array([0, 1499, 0], np.uint32),
                              np.array([3, 2], np.uint64), 1500, 1)
    child = _fptree.conditional_tree(ctx, tree, 1499, 1500, 1)
    assert child.items.tolist() == [-1, 0] and child.counts.tolist() == [0, 3]
    big = 2**53 + 7
    tree = _fptree.build_tree(ctx, np.array([0, 0, 1], np.uint64), np.array([0, 1, 0], np.uint32),
                              np.array([big, big + 1], np.uint64), 2, 1)
    assert dict(tree.header) == {0: 2 * big + 1, 1: big}
    cond = _fptree.conditional_tree(ctx, tree, 1, 2, 1)
    assert dict(cond.header) == {0: big}
    assert _fptree.build_tree(ctx, np.zeros(0, np.uint64), np.zeros(0, np.uint32),
                              np.ones(3, np.uint64), 4, 1) is None


@pytest.mark.parametrize("seed", range(4))
def test_fpgrowth_mine_random(ctx, seed):
    rng = random.Random(seed)
    rows = [rng.sample(range(9), rng.randrange(10)) for _ in range(60 + 20 * seed)]
    labels, rr, cc, n = coo(rows)
    # duplicate pairs must not double count
    rr2 = np.concatenate([rr, rr[::3]])
    cc2 = np.concatenate([cc, cc[::3]])
    host = oracle.bits_build(rr, cc, n, len(labels))
    for ms in (1, 2, 6, 25):
        got = ctx.fpgrowth_mine(rr2, cc2, n, len(labels), ms)
        as_sets = {tuple(sorted(k)): v for k, v in got.to_dict().items()}
        assert as_sets == oracle.eclat(host, ms)
        assert len(as_sets) == len(got)


def test_fpgrowth_mine_zipf_medium(ctx):
    db = datagen.transactional(20_000, 60, 6.0, 1.0, seed=12)
    labels, cols = datagen.label_columns(db.items)
    host = oracle.bits_build(db.rows, cols, db.n_transactions, len(labels))
    got = ctx.fpgrowth_mine(db.rows, cols, db.n_transactions, len(labels), 60)
    as_sets = {tuple(sorted(k)): v for k, v in got.to_dict().items()}
    expect = oracle.eclat(host, 60)
    assert len(expect) > 500
    assert as_sets == expect


def _mine_sets(ctx, rows, minsup):
    labels, rr, cc, n = coo(rows)
    got = ctx.fpgrowth_mine(rr, cc, n, len(labels), minsup)
    as_sets = {tuple(sorted(k)): v for k, v in got.to_dict().items()}
    assert len(as_sets) == len(got)
    return as_sets, oracle.eclat(oracle.bits_build(rr, cc, n, len(labels)), minsup)


def test_fpgrowth_mine_wide_conditional(ctx):
    # item 0 (support 70) sits under 70 more frequent items (support 72),
    # each frequent with it (twice) while no two of them are: its
    # conditional tree has more items than one workgroup's mask holds and
    # takes the general projection
    ring = [[0, 1 + i, 1 + (i + 1) % 70] for i in range(70)]
    solo = [[1 + i] for i in range(70) for _ in range(70)]
    got, expect = _mine_sets(ctx, ring + solo, 2)
    assert any(len(k) == 2 for k in expect)
    assert got == expect


@pytest.mark.parametrize("seed", range(3))
def test_fpgrowth_mine_deep(ctx, seed):
    # dense baskets: patterns of up to ~9 items, conditional trees of
    # conditional trees several levels down, all built one per workgroup
    rng = random.Random(100 + seed)
    rows = [[i for i in range(14) if rng.random() < 0.72] for _ in range(400)]
    for minsup in (40, 90):
        got, expect = _mine_sets(ctx, rows, minsup)
        assert max(map(len, expect)) >= 5
        assert got == expect


def test_fpgrowth_mine_many_nodes_per_item(ctx):
    # 30000 random subsets of 16 items: the first tree holds thousands of
    # nodes of each low item (general projection), whose conditional trees
    # are searched in the compact layout further down
    rng = np.random.default_rng(5)
    dense = rng.random((30000, 16)) < 0.5
    rows = [np.flatnonzero(r).tolist() for r in dense]
    got, expect = _mine_sets(ctx, rows, 900)
    assert len(expect) > 1000
    assert got == expect
