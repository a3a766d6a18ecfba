"""Full-size parity (1,000,000 x 1,000 Zipf workload and the 2,000,000-row
temporal workload): device search against the oracle, and size-independent
properties (Gram == popcount, pair-table pruning on/off, repeatability)."""

import numpy as np
import pytest

import oracle
from cupami_amd import _hip, datagen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return _hip.context(0)


@pytest.fixture(scope="module")
def c2(ctx):
    db = datagen.eclat_1m_1k()
    labels, cols = datagen.label_columns(db.items)
    bits = ctx.build_bits(db.rows, cols, db.n_transactions, len(labels))
    return db, labels, cols, bits


def test_store_checksum(c2):
    db, labels, cols, bits = c2
    host = oracle.bits_build(db.rows, cols, db.n_transactions, len(labels))
    assert np.array_equal(bits.download(), host)


def test_gram_equals_popcount_at_full_size(ctx, c2):
    db, labels, cols, bits = c2
    items = np.arange(len(labels), dtype=np.uint32)
    gram = ctx.pair_gram(bits, items)
    rng = np.random.default_rng(5)
    a = rng.integers(0, len(labels), 4000)
    b = rng.integers(0, len(labels), 4000)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    cand = np.stack([lo, hi], 1).astype(np.uint32)
    assert np.array_equal(gram[lo, hi].astype(np.uint64), ctx.count_supports(bits, cand))
    # diagonal = item supports
    assert np.array_equal(np.diag(gram).astype(np.uint64),
                          ctx.count_supports(bits, items[:, None]))


def test_eclat_full_size_equals_oracle(ctx, c2):
    db, labels, cols, bits = c2
    host = oracle.bits_build(db.rows, cols, db.n_transactions, len(labels))
    expect = oracle.eclat(host, 1000)
    got = ctx.eclat_mine(bits, 1000, use_gram=True).to_dict()
    assert len(got) == len(expect) == 43645
    assert got == expect
    # the popcount level-2 path gives the same forest
    again = ctx.eclat_mine(bits, 1000, use_gram=False)
    assert again.to_dict() == expect


def test_pfeclat_full_size_equals_oracle(ctx):
    db = datagen.temporal_2m_500()
    labels, cols = datagen.label_columns(db.items)
    bits = ctx.build_bits(db.rows, cols, db.n_transactions, len(labels))
    host = oracle.bits_build(db.rows, cols, db.n_transactions, len(labels))
    end = int(db.timestamps[-1])
    expect = oracle.pfeclat(host, db.timestamps, db.n_transactions, end, 10_000, 2_000)
    got = ctx.pfeclat_mine(bits, db.timestamps, end, 10_000, 2_000).to_dict()
    assert len(got) == len(expect) == 879
    assert got == expect
