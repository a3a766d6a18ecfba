"""The workload configurations end to end on the device, against the oracle:
the 10,000 x 50 uniform case through the public miners, all 256 databases of
the batched workload in one search, an input with 200,000 distinct ids of
which few are frequent, and the level-2 Gram on more than 2^24 transactions
(where its K slices must stay below 2^24 for FP32 sums to be exact)."""

import numpy as np
import pytest

import oracle
from cupami_amd import _hip, datagen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return _hip.context(0)


def test_small_uniform_public_miners(ctx):
    """10,000 baskets x 50 items, Poisson(6), uniform popularity, seed 101,
    minSup 0.01 (ceil -> 100 transactions): cuEclat and cuApriori's
    getPatterns() == the oracle's itemsets under the reference's labels."""
    from cupami_amd import cuApriori, cuEclat
    db = datagen.small_uniform(101)
    print(f"configs: small_uniform {db.n_transactions}x{db.n_items}", flush=True)
    labels, cols = datagen.label_columns(db.items)
    host = oracle.bits_build(db.rows, cols, db.n_transactions, len(labels))
    expect = {"\t".join(labels[c] for c in k): v for k, v in oracle.eclat(host, 100).items()}
    assert len(expect) > 50
    baskets = db.baskets()
    for cls in (cuEclat, cuApriori):
        m = cls(baskets, 0.01)
        m.mine()
        assert m.getPatterns() == expect, cls.__name__


def test_batch_256_equals_oracle(ctx):
    """All 256 databases of the batched workload (seed 303) in one device
    search; each database's itemsets == the oracle's."""
    import torch
    dbs = datagen.batch_256()
    print(f"configs: batch_256, {sum(db.rows.size for db, _ in dbs)} pairs", flush=True)
    cols_all, labels_all = [], []
    for db, _ in dbs:
        labels, cols = datagen.label_columns(db.items)
        labels_all.append(labels)
        cols_all.append(cols)
    d_rows = torch.from_numpy(np.concatenate([db.rows for db, _ in dbs]).view(np.int64)).to("cuda:0")
    d_cols = torch.from_numpy(np.concatenate(cols_all).view(np.int32)).to("cuda:0")
    torch.cuda.synchronize()
    specs, off = [], 0
    for (db, _), labels in zip(dbs, labels_all):
        k = db.rows.size
        specs.append((d_rows.data_ptr() + 8 * off, d_cols.data_ptr() + 4 * off, k, db.n_transactions,
                      len(labels)))
        off += k
    minsups = [ms for _, ms in dbs]
    stores = ctx.build_bits_device_batch(specs)
    try:
        result = ctx.eclat_mine_batch(stores, minsups)
    finally:
        for s in stores:
            s.close()
    parts = result.split(len(dbs))
    total = 0
    for i, ((db, ms), cols, labels, got) in enumerate(zip(dbs, cols_all, labels_all, parts)):
        host = oracle.bits_build(db.rows, cols, db.n_transactions, len(labels))
        expect = oracle.eclat(host, ms)
        assert got == expect, f"database {i}"
        total += len(expect)
    assert total == len(result) > 1_000_000


def test_many_distinct_few_frequent(ctx):
    """200,000 distinct ids, five of them frequent: level-2 buffers are sized
    by the frequent items, not by the square of the distinct ones."""
    rng = np.random.default_rng(17)
    n = 20_000
    pool = np.arange(1_000_000, 2**31, 7919, dtype=np.int64)
    cold = rng.choice(pool, size=n * 10, replace=False).reshape(n, 10)
    hot = rng.integers(1, 6, size=(n, 3))
    items = np.concatenate([hot, cold], axis=1)
    rows = np.repeat(np.arange(n, dtype=np.uint64), items.shape[1])
    ids = items.reshape(-1).astype(np.uint32)
    labels, cols = datagen.label_columns(ids)
    assert len(labels) == 200_005
    print(f"configs: {len(labels)} distinct ids", flush=True)
    bits = ctx.build_bits(rows, cols, n, len(labels))
    host = oracle.bits_build(rows, cols, n, len(labels))
    expect = oracle.eclat(host, 50)
    got = ctx.eclat_mine(bits, 50).to_dict()
    assert got == expect
    assert 5 <= len(expect) < 100
    again = ctx.eclat_mine(bits, 50, use_gram=False).to_dict()
    assert again == expect


def test_gram_beyond_2pow24_transactions(ctx):
    """3328 items (13 blocks of 256, so the planner's K slices are bounded by
    the 2^23-transaction cap, not by the work split) over 2^24 + 2^20 + 3
    transactions: item 0 is in every row (its count is not an FP32 value),
    item 1 in all but five; every sampled pair and the diagonal == popcount."""
    import torch
    n = (1 << 24) + (1 << 20) + 3
    f = 3328
    g = torch.Generator(device="cuda:0")
    g.manual_seed(23)
    m = 40_000_000
    r_rand = torch.randint(0, n, (m,), generator=g, device="cuda:0", dtype=torch.int64)
    c_rand = torch.randint(2, f, (m,), generator=g, device="cuda:0", dtype=torch.int32)
    all_rows = torch.arange(n, device="cuda:0", dtype=torch.int64)
    d_rows = torch.cat([r_rand, all_rows, all_rows[: n - 5]])
    d_cols = torch.cat([c_rand, torch.zeros(n, dtype=torch.int32, device="cuda:0"),
                        torch.ones(n - 5, dtype=torch.int32, device="cuda:0")])
    torch.cuda.synchronize()
    print(f"configs: gram {n} x {f}, {d_rows.numel()} pairs", flush=True)
    bits = ctx.build_bits_device(d_rows.data_ptr(), d_cols.data_ptr(), d_rows.numel(), n, f)
    del r_rand, c_rand, all_rows, d_rows, d_cols
    items = np.arange(f, dtype=np.uint32)
    gram = ctx.pair_gram(bits, items)
    assert int(gram[0, 0]) == n and int(gram[1, 1]) == n - 5 and int(gram[0, 1]) == n - 5
    diag = ctx.count_supports(bits, items[:, None])
    assert np.array_equal(np.diag(gram).astype(np.uint64), diag)
    assert np.array_equal(gram[0, 2:].astype(np.uint64), diag[2:])
    rng = np.random.default_rng(4)
    a = rng.integers(0, f, 3000)
    b = rng.integers(0, f, 3000)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    cand = np.stack([lo, hi], 1).astype(np.uint32)
    assert np.array_equal(gram[lo, hi].astype(np.uint64), ctx.count_supports(bits, cand))
    bits.close()
