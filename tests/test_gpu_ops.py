"""Operator parity: the HIP entry points against the CPU oracle, bit-exact."""

import random
from itertools import combinations

import numpy as np
import pytest

import oracle
from cupami_amd import _hip
from helpers import coo, random_rows

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return _hip.context(0)


def random_store(seed, n, n_items, density):
    rng = np.random.default_rng(seed)
    mask = rng.random((n, n_items)) < density
    r, c = np.nonzero(mask)
    return r.astype(np.uint64), c.astype(np.uint32)


@pytest.mark.parametrize("n", [0, 1, 31, 63, 64, 65, 129, 4096, 100_003])
@pytest.mark.parametrize("row_sorted", [True, False])
def test_bits_build(ctx, n, row_sorted):
    rows, cols = random_store(n, n, 7, 0.3)
    # duplicates must OR together
    rows = np.concatenate([rows, rows[:5]])
    cols = np.concatenate([cols, cols[:5]])
    if row_sorted:
        order = np.argsort(rows, kind="stable")
        rows, cols = rows[order], cols[order]
    bits = ctx.build_bits(rows, cols, n, 7)
    assert np.array_equal(bits.download(), oracle.bits_build(rows, cols, n, 7))


def test_bits_build_single_descent(ctx):
    # sorted rows with one descent: at the start, at the end, and where the
    # tile build's workgroups meet (512 rows apart for 7 items), so the
    # check that a workgroup's last pair precedes the next run's first is what
    # finds it; the store must come out as a scatter build gives it
    n = 100_003
    rows, cols = random_store(7, n, 7, 0.3)
    order = np.argsort(rows, kind="stable")
    rows, cols = rows[order], cols[order]
    starts = np.searchsorted(rows, np.arange(0, n, 512))
    picks = [0, rows.size - 2] + [int(b) - 1 for b in starts[1:40:7]] + [int(starts[100])]
    steps = np.flatnonzero(rows[1:] != rows[:-1])  # j with rows[j] < rows[j + 1]
    for i in picks:
        r, c = rows.copy(), cols.copy()
        j = int(steps[min(np.searchsorted(steps, i), steps.size - 1)])
        r[j], r[j + 1] = r[j + 1], r[j]
        c[j], c[j + 1] = c[j + 1], c[j]
        bits = ctx.build_bits(r, c, n, 7)
        assert np.array_equal(bits.download(), oracle.bits_build(r, c, n, 7)), i
        bits.close()


def test_bits_build_many_items_and_reuse(ctx):
    # item counts on both sides of the LDS tile limit, store buffers reused
    for items in (1, 1000, 4096, 4097, 7168, 7169, 9000):
        rows, cols = random_store(items, 3000, items, 0.002)
        for _ in range(2):
            bits = ctx.build_bits(rows, cols, 3000, items)
            assert np.array_equal(bits.download(), oracle.bits_build(rows, cols, 3000, items))
            bits.close()


def _device_coo(rows, cols):
    import torch
    d_rows = torch.from_numpy(np.ascontiguousarray(rows, np.uint64).view(np.int64)).to("cuda:0")
    d_cols = torch.from_numpy(np.ascontiguousarray(cols, np.uint32).view(np.int32)).to("cuda:0")
    torch.cuda.synchronize()
    return d_rows, d_cols


def test_bits_build_device_batch(ctx):
    """One batch call gives each database the store a single build gives."""
    cases = []
    for k, (n, items, sorted_rows) in enumerate([(1, 3, True), (65, 7, False), (4096, 40, True), (0, 5, True),
                                                 (3000, 9000, True), (100_003, 200, False), (777, 1, True)]):
        rows, cols = random_store(k, n, items, 0.05 if items < 1000 else 0.002) if n else (
            np.zeros(0, np.uint64), np.zeros(0, np.uint32))
        if sorted_rows and rows.size:
            order = np.argsort(rows, kind="stable")
            rows, cols = rows[order], cols[order]
        cases.append((rows, cols, n, items))
    dev = [_device_coo(r, c) for r, c, _, _ in cases]
    specs = [(dr.data_ptr(), dc.data_ptr(), r.size, n, items) for (dr, dc), (r, _, n, items) in zip(dev, cases)]
    for _ in range(2):  # the second batch reuses the first one's arena
        stores = ctx.build_bits_device_batch(specs)
        for bits, (rows, cols, n, items) in zip(stores, cases):
            assert np.array_equal(bits.download(), oracle.bits_build(rows, cols, n, items))
        for bits in stores[::2] + stores[1::2]:
            bits.close()


def test_bits_build_device_batch_rejects_and_mines(ctx):
    rng = random.Random(5)
    dbs = []
    for k in range(5):
        labels, rr, cc, n = coo(random_rows(rng, 300 + 50 * k, 12, 6))
        dbs.append((labels, rr, cc, n))
    dev = [_device_coo(rr, cc) for _, rr, cc, _ in dbs]
    specs = [(dr.data_ptr(), dc.data_ptr(), rr.size, n, len(labels))
             for (dr, dc), (labels, rr, _, n) in zip(dev, dbs)]
    # one bad column anywhere rejects the whole batch
    bad = list(specs)
    bad[3] = (bad[3][0], bad[3][1], bad[3][2], bad[3][3], 1)
    with pytest.raises(ValueError):
        ctx.build_bits_device_batch(bad)
    stores = ctx.build_bits_device_batch(specs)
    minsups = [3, 5, 8, 12, 20]
    parts = ctx.eclat_mine_batch(stores, minsups).split(len(stores))
    for part, bits, (labels, rr, cc, n), ms in zip(parts, stores, dbs, minsups):
        expect = oracle.eclat(oracle.bits_build(rr, cc, n, len(labels)), ms)
        assert part == expect
        assert ctx.eclat_mine(bits, ms).to_dict() == expect


def test_bits_build_rejects_out_of_range(ctx):
    for sorted_rows in ([0, 5], [5, 0]):
        with pytest.raises(ValueError):
            ctx.build_bits(np.array(sorted_rows, np.uint64), np.array([0, 3], np.uint32), 5, 4)
    for rows in ([0, 1], [1, 0]):
        with pytest.raises(ValueError):
            ctx.build_bits(np.array(rows, np.uint64), np.array([0, 4], np.uint32), 5, 4)


@pytest.mark.parametrize("n,width", [(70, 1), (70, 2), (5000, 3), (200_001, 4)])
def test_count_supports(ctx, n, width):
    rows, cols = random_store(width * 7 + n, n, 9, 0.6)
    bits = ctx.build_bits(rows, cols, n, 9)
    host = oracle.bits_build(rows, cols, n, 9)
    rng = np.random.default_rng(n)
    cand = np.stack([rng.choice(9, width, replace=False) for _ in range(300)]).astype(np.uint32)
    assert np.array_equal(ctx.count_supports(bits, cand), oracle.count_supports(host, cand))


def test_intersect_extensions(ctx):
    rng = np.random.default_rng(3)
    for words in (1, 2, 63, 64, 65, 3000):
        prefix = rng.integers(0, 2**63, words, dtype=np.uint64) | rng.integers(0, 2**63, words, dtype=np.uint64)
        ext = rng.integers(0, 2**63, (37, words), dtype=np.uint64)
        expect_rows = prefix[None, :] & ext
        expect_sup = np.array([sum(bin(int(x)).count("1") for x in r) for r in expect_rows])
        sup, rows, idx = ctx.intersect_extensions(prefix, ext, keep_min=0)
        assert np.array_equal(sup, expect_sup)
        assert np.array_equal(rows, expect_rows) and np.array_equal(idx, np.arange(37))
        thr = int(np.median(expect_sup))
        sup, rows, idx = ctx.intersect_extensions(prefix, ext, keep_min=thr)
        keep = np.nonzero(expect_sup >= thr)[0]
        assert np.array_equal(idx, keep) and np.array_equal(rows, expect_rows[keep])


# (30_000, 520) and (200_000, 600): blocks of 512+ items cut into K slices,
# summed through partial slots and the reduce launch; the others add their
# slices with atomics or write single-slice blocks directly
@pytest.mark.parametrize("n,items", [(64, 3), (1000, 17), (70_000, 65), (300_001, 130), (5000, 257),
                                     (30_000, 520), (200_000, 600)])
def test_pair_gram_equals_popcount(ctx, n, items):
    rows, cols = random_store(n + items, n, items, 0.35)
    bits = ctx.build_bits(rows, cols, n, items)
    host = oracle.bits_build(rows, cols, n, items)
    order = np.random.default_rng(items).permutation(items).astype(np.uint32)
    gram = ctx.pair_gram(bits, order)
    pairs = [(a, b) for a in range(items) for b in range(a, items)]
    cand = np.array([(order[a], order[b]) for a, b in pairs], np.uint32)
    expect = oracle.count_supports(host, cand)
    got = np.array([gram[a, b] for a, b in pairs], np.uint64)
    assert np.array_equal(got, expect)
    assert np.all(np.tril(gram, -1) == 0)


@pytest.mark.parametrize("n", [1, 64, 65, 5000, 100_001])
def test_periodic_eval(ctx, n):
    rows, cols = random_store(n, n, 6, 0.4)
    rng = np.random.default_rng(n + 1)
    ts = np.cumsum(rng.integers(0, 4, n)).astype(np.int64)   # repeats allowed
    end = int(ts[-1]) + 5 if n else 0
    bits = ctx.build_bits(rows, cols, n, 6)
    host = oracle.bits_build(rows, cols, n, 6)
    cand = np.array([c for w in (1, 2, 3) for c in combinations(range(6), w)][:40], dtype=object)
    for width in (1, 2, 3):
        cw = np.array([c for c in combinations(range(6), width)], np.uint32)
        s, p = ctx.periodic_eval(bits, ts, end, cw)
        es, ep = oracle.periodic_eval(host, ts, n, end, cw)
        assert np.array_equal(s, es) and np.array_equal(p, ep)


@pytest.mark.parametrize("seed", [3, 4])
def test_periodic_eval_inner_gaps(ctx, seed):
    """Dense rows whose largest gap lies between two occurrences of one
    64-row word: steps of 1 with a few jumps of up to 900.  The kernel walks
    a word's inner occurrences only when the word's span exceeds the largest
    gap already known; the periods must still be the oracle's."""
    n, items = 20_000, 6
    rng = np.random.default_rng(seed)
    dense = rng.random((n, items)) < np.array([0.9, 0.7, 0.5, 0.3, 0.1, 0.02])
    rows, cols = np.nonzero(dense)
    steps = np.ones(n, np.int64)
    steps[rng.choice(n, 12, replace=False)] = rng.integers(50, 900, 12)
    ts = np.cumsum(steps)
    end = int(ts[-1]) + 3
    bits = ctx.build_bits(rows.astype(np.uint64), cols.astype(np.uint32), n, items)
    host = oracle.bits_build(rows.astype(np.uint64), cols.astype(np.uint32), n, items)
    for width in (1, 2, 3):
        cw = np.array([c for c in combinations(range(items), width)], np.uint32)
        s, p = ctx.periodic_eval(bits, ts, end, cw)
        es, ep = oracle.periodic_eval(host, ts, n, end, cw)
        assert np.array_equal(s, es) and np.array_equal(p, ep), width
    for ms, per in [(1, 10_000), (200, 1000)]:
        assert ctx.pfeclat_mine(bits, ts, end, ms, per).to_dict() == oracle.pfeclat(host, ts, n, end, ms, per)


@pytest.mark.parametrize("seed", range(6))
def test_eclat_mine_random(ctx, seed):
    rng = random.Random(seed)
    rows = random_rows(rng, 40 + 37 * seed, 10, 9)
    labels, rr, cc, n = coo(rows)
    bits = ctx.build_bits(rr, cc, n, len(labels))
    host = oracle.bits_build(rr, cc, n, len(labels))
    for ms in (0, 1, 2, 7, n // 4, n, n + 1):
        expect = oracle.eclat(host, ms)
        for gram in (True, False):
            assert ctx.eclat_mine(bits, ms, use_gram=gram).to_dict() == expect, (ms, gram)


def test_eclat_mine_zipf_medium(ctx):
    from cupami_amd import datagen
    db = datagen.transactional(60_000, 120, 8.0, 1.0, seed=11)
    labels, cols = datagen.label_columns(db.items)
    bits = ctx.build_bits(db.rows, cols, db.n_transactions, len(labels))
    host = oracle.bits_build(db.rows, cols, db.n_transactions, len(labels))
    expect = oracle.eclat(host, 150)
    assert len(expect) > 1000
    assert ctx.eclat_mine(bits, 150, use_gram=True).to_dict() == expect
    assert ctx.eclat_mine(bits, 150, use_gram=False).to_dict() == expect


@pytest.mark.parametrize("seed", range(4))
def test_pfeclat_mine_random(ctx, seed):
    from helpers import temporal_coo
    rng = random.Random(seed)
    rows = [[rng.randrange(0, 400), *rng.sample(range(8), rng.randint(0, 8))] for _ in range(150)]
    labels, rr, cc, n, ts, end = temporal_coo(rows)
    bits = ctx.build_bits(rr, cc, n, len(labels))
    host = oracle.bits_build(rr, cc, n, len(labels))
    for ms, per in [(1, 400), (3, 60), (5, 30), (20, 200)]:
        expect = oracle.pfeclat(host, ts, n, end, ms, per)
        assert ctx.pfeclat_mine(bits, ts, end, ms, per).to_dict() == expect, (ms, per)


def test_pfeclat_large_result_streamed(ctx):
    """A periodic search of 2^19 - 1 patterns: the result streams to the host
    chunk by chunk (and the first attempt overflows its pattern capacity);
    the forest equals the oracle's."""
    rng = np.random.default_rng(19)
    n, items = 3000, 19
    dense = rng.random((n, items)) < 0.95
    rows, cols = np.nonzero(dense)
    ts = np.arange(1, n + 1, dtype=np.int64)
    bits = ctx.build_bits(rows.astype(np.uint64), cols.astype(np.uint32), n, items)
    host = oracle.bits_build(rows.astype(np.uint64), cols.astype(np.uint32), n, items)
    expect = oracle.pfeclat(host, ts, n, n, 1, n)
    assert len(expect) == 2 ** items - 1
    got = ctx.pfeclat_mine(bits, ts, n, 1, n).to_dict()
    assert got == expect


def test_pfeclat_mine_device_timestamps(ctx):
    """Device-resident timestamps give the host path's (and the oracle's) forest."""
    import torch
    from helpers import temporal_coo
    rng = random.Random(11)
    rows = [[rng.randrange(0, 2000), *rng.sample(range(12), rng.randint(0, 12))] for _ in range(900)]
    labels, rr, cc, n, ts, end = temporal_coo(rows)
    bits = ctx.build_bits(rr, cc, n, len(labels))
    host = oracle.bits_build(rr, cc, n, len(labels))
    d_ts = torch.from_numpy(np.ascontiguousarray(ts, dtype=np.int64)).to("cuda:0")
    torch.cuda.synchronize()
    for ms, per in [(2, 2000), (20, 300), (60, 120)]:
        expect = oracle.pfeclat(host, ts, n, end, ms, per)
        got = ctx.pfeclat_mine_device(bits, d_ts.data_ptr(), end, ms, per).to_dict()
        assert got == expect, (ms, per)
        assert ctx.pfeclat_mine(bits, ts, end, ms, per).to_dict() == expect, (ms, per)


@pytest.mark.parametrize("graph,prune,group,tile", [(True, True, 32, 2048), (False, True, 16, 1024),
                                                      (True, False, 64, 1024), (False, False, 8, 2048)])
def test_search_schedules_agree(ctx, graph, prune, group, tile):
    """Every schedule (graph replay or direct launches, subset test on or
    off, group and tile sizes) gives the oracle's forest, several times in a
    row on the same buffers."""
    from cupami_amd import datagen
    db = datagen.transactional(30_000, 80, 7.0, 1.0, seed=21)
    labels, cols = datagen.label_columns(db.items)
    bits = ctx.build_bits(db.rows, cols, db.n_transactions, len(labels))
    host = oracle.bits_build(db.rows, cols, db.n_transactions, len(labels))
    try:
        ctx.set_search_options(group, tile, graph, prune)
        for ms in (90, 300, 90):
            expect = oracle.eclat(host, ms)
            assert ctx.eclat_mine(bits, ms, use_gram=True).to_dict() == expect
            assert ctx.eclat_mine(bits, ms, use_gram=False).to_dict() == expect
    finally:
        ctx.set_search_options()


def test_batch_equals_separate_searches(ctx):
    from cupami_amd import datagen
    rng = np.random.default_rng(33)
    stores, hosts, mins = [], [], []
    for i in range(12):
        n = int(rng.integers(200, 20_000))
        items = int(rng.integers(3, 60))
        db = datagen.transactional(n, items, float(rng.uniform(2, 6)), 1.0, seed=1000 + i)
        labels, cols = datagen.label_columns(db.items)
        stores.append(ctx.build_bits(db.rows, cols, db.n_transactions, len(labels)))
        hosts.append(oracle.bits_build(db.rows, cols, db.n_transactions, len(labels)))
        mins.append(max(1, int(n * rng.uniform(0.005, 0.05))))
    got = ctx.eclat_mine_batch(stores, mins).split(len(stores))
    for i in range(len(stores)):
        assert got[i] == oracle.eclat(hosts[i], mins[i]), i
    # and again, reusing the captured graphs with other thresholds
    mins2 = [m + 3 for m in mins]
    got = ctx.eclat_mine_batch(stores, mins2, use_gram=False).split(len(stores))
    for i in range(len(stores)):
        assert got[i] == oracle.eclat(hosts[i], mins2[i]), i


def test_large_results_views_outlive_their_patterns(ctx):
    """Results of 262144+ patterns live in recycled host blocks and reach
    Python as views: a view keeps its result alive after the Patterns object
    is gone, later searches (which reuse freed blocks) do not disturb it, and
    their own results stay exact."""
    import gc
    rng = np.random.default_rng(808)
    n, items = 5000, 20
    dense = rng.random((n, items)) < 0.8
    rows, cols = np.nonzero(dense)
    rows = rows.astype(np.uint64)
    cols = cols.astype(np.uint32)
    bits = ctx.build_bits(rows, cols, n, items)
    host = oracle.bits_build(rows, cols, n, items)
    first = ctx.eclat_mine(bits, 600)
    assert len(first) > 262144
    expect_first = oracle.eclat(host, 600)
    assert first.to_dict() == expect_first
    support = first.support
    kept = support.copy()
    del first
    gc.collect()
    for minsup in (700, 650):
        again = ctx.eclat_mine(bits, minsup)
        assert again.to_dict() == oracle.eclat(host, minsup)
        del again
        gc.collect()
    assert np.array_equal(support, kept)
    del support
    gc.collect()
    last = ctx.eclat_mine_batch([bits, bits], [600, 800])
    parts = last.split(2)
    assert parts[0] == expect_first
    assert parts[1] == oracle.eclat(host, 800)


def test_roof_phases(ctx):
    """The phase-boundary microbenchmark: every wait ends (no time-out
    error), both costs are positive, and a grid that cannot be resident at
    once is refused before launch."""
    import math
    ub, ul = ctx.roof_phases(64, 4, 4096)
    assert 0 < ub < 1e5 and 0 < ul < 1e5 and math.isfinite(ub) and math.isfinite(ul)
    with pytest.raises(ValueError):
        ctx.roof_phases(1 << 20, 1, 4096)
    with pytest.raises(ValueError):
        ctx.roof_phases(64, 1, 100)   # not a multiple of 16 bytes
