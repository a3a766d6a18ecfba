"""Store build straight from a basket file's arrays (cpm_bits_build_csr*),
bit-exact against the oracle's restatement of the reference's BIN path
(cuApriori.py:381-396), and the BIN cases of the reference's own tests
(test_transaction_formats.py:185-250)."""

import random

import numpy as np
import pytest

import oracle
from cupami_amd import _hip, cuApriori, cuEclat, cuPFEclat
from helpers import csr, write_bin, write_temporal_bin

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return _hip.context(0)


def check_against_oracle(ctx, offsets, ids, universe, keep_empty=False):
    bits, label_ids = ctx.build_bits_csr(offsets, ids, universe, keep_empty)
    try:
        labels, rr, cc, n = oracle.csr_to_coo(offsets, ids, keep_empty)
        assert [str(v) for v in label_ids.tolist()] == labels
        n_rows, n_items, _, _ = bits.shape()
        assert (n_rows, n_items) == (n, len(labels))
        assert np.array_equal(bits.download(), oracle.bits_build(rr, cc, n, len(labels)))
    finally:
        bits.close()


def random_csr(seed, n, universe, distinct, max_len, empty_share=0.1):
    rng = np.random.default_rng(seed)
    pool = np.unique(np.concatenate([rng.integers(0, universe + 1, distinct, dtype=np.uint64),
                                     [0, universe]])).astype(np.uint32)
    lengths = rng.integers(1, max_len + 1, n)
    lengths[rng.random(n) < empty_share] = 0
    offsets = np.zeros(n + 1, np.uint64)
    offsets[1:] = np.cumsum(lengths)
    ids = pool[rng.integers(0, pool.size, int(offsets[-1]))]   # repeats within a basket too
    return offsets, ids


@pytest.mark.parametrize("case", [
    # (seed, baskets, universe, distinct ids, max basket length, keep_empty)
    (1, 5000, 100, 60, 12, False),            # presence bitmap in LDS, LDS tile B = 8
    (2, 20000, 200_000, 3000, 20, False),     # presence bitmap in global memory, B = 2
    (3, 3000, 2**32 - 1, 40, 6, False),       # 2^32-bit bitmap, ids of every decimal length
    (4, 3000, 50_000, 9000, 30, False),       # > 7168 columns: global atomics
    (8, 3000, 100_000, 6000, 30, False),      # 4097..7168 columns: one-word tile, map in global memory
    (9, 6000, 4095, 4000, 12, False),         # largest map held in LDS
    (5, 4000, 1000, 700, 10, True),           # every basket a row
    (6, 129, 99, 99, 3, False),               # ragged: 129 baskets, a part-filled word
])
def test_csr_build_matches_oracle(ctx, case):
    seed, n, universe, distinct, max_len, keep_empty = case
    offsets, ids = random_csr(seed, n, universe, distinct, max_len)
    check_against_oracle(ctx, offsets, ids, universe, keep_empty)


def test_csr_build_edge_cases(ctx):
    for rows, universe in [([], 1), ([[], [], []], 8), ([[0]], 1), ([[5]] * 64, 5), ([[5]] * 65, 5),
                           ([[1, 1, 2], [], [0, 2], [4294967295, 2]], 4294967295)]:
        offsets, ids = csr(rows)
        for keep_empty in (False, True):
            check_against_oracle(ctx, offsets, ids, universe, keep_empty)


def test_csr_build_random_small(ctx):
    rng = random.Random(11)
    for trial in range(40):
        universe = rng.choice([1, 9, 10, 11, 99, 100, 101, 1000, 99999, 2**20])
        pool = sorted({rng.randrange(universe + 1) for _ in range(rng.randrange(1, 40))})
        rows = [[rng.choice(pool) for _ in range(rng.randrange(6))] for _ in range(rng.randrange(200))]
        offsets, ids = csr(rows)
        check_against_oracle(ctx, offsets, ids, universe, keep_empty=bool(trial & 1))


def test_csr_build_rejects_like_the_reference(ctx):
    """The array checks of _readBinaryFile (cuApriori.py:273-276): offsets
    first, then ids; no store comes back and the context stays usable."""
    offsets, ids = csr([[1, 2], [3]])
    bad_offsets = [np.array([1, 2, 3], np.uint64),      # does not start at 0
                   np.array([0, 3, 2], np.uint64),      # decreases
                   np.array([0, 2, 2], np.uint64)]      # does not end at m
    for off in bad_offsets:
        with pytest.raises(ValueError, match="invalid transaction offsets"):
            ctx.build_bits_csr(off, ids, 8)
    with pytest.raises(ValueError, match="outside configured universe"):
        ctx.build_bits_csr(offsets, ids, 2)
    with pytest.raises(ValueError, match="invalid transaction offsets"):   # both wrong: offsets win
        ctx.build_bits_csr(np.array([0, 3, 2], np.uint64), np.array([9, 9, 9], np.uint32), 2)
    check_against_oracle(ctx, offsets, ids, 3)


def test_csr_build_device_pointers(ctx):
    import torch
    offsets, ids = random_csr(7, 10000, 500, 400, 15)
    d_off = torch.from_numpy(offsets.view(np.int64)).cuda()
    d_ids = torch.from_numpy(ids.view(np.int32)).cuda()
    torch.cuda.synchronize()
    bits, label_ids = ctx.build_bits_csr_device(d_off.data_ptr(), offsets.size - 1, d_ids.data_ptr(),
                                                ids.size, 500)
    try:
        host_bits, host_labels = ctx.build_bits_csr(offsets, ids, 500)
        assert np.array_equal(label_ids, host_labels)
        assert np.array_equal(bits.download(), host_bits.download())
        host_bits.close()
    finally:
        bits.close()


def test_csr_build_unaligned_ids(ctx):
    """Ids at an address that is not 16-byte aligned (as in a mapped BIN
    file, where they follow 32 + 8(N+1) bytes) take the input scan's scalar
    path; the store is the same."""
    import torch
    offsets, ids = random_csr(12, 3000, 300, 250, 9)
    d_off = torch.from_numpy(offsets.view(np.int64)).cuda()
    padded = np.concatenate([np.zeros(1, np.uint32), ids])
    d_ids = torch.from_numpy(padded.view(np.int32)).cuda()
    torch.cuda.synchronize()
    addr = d_ids.data_ptr() + 4
    assert addr % 16 != 0
    bits, label_ids = ctx.build_bits_csr_device(d_off.data_ptr(), offsets.size - 1, addr, ids.size, 300)
    try:
        labels, rr, cc, n = oracle.csr_to_coo(offsets, ids)
        assert [str(v) for v in label_ids.tolist()] == labels
        assert np.array_equal(bits.download(), oracle.bits_build(rr, cc, n, len(labels)))
    finally:
        bits.close()


def test_csr_build_full_size_1m_x_1k(ctx):
    """The 1M-basket, 1000-item database of the Eclat benchmark, as a
    CUPAMI01 file would hold it: same store as the pairs path."""
    from cupami_amd import datagen
    db = datagen.eclat_1m_1k()
    offsets = np.searchsorted(db.rows, np.arange(db.n_transactions + 1, dtype=np.uint64)).astype(np.uint64)
    bits, label_ids = ctx.build_bits_csr(offsets, db.items, db.n_items)
    try:
        labels, cols = datagen.label_columns(db.items)
        assert [str(v) for v in label_ids.tolist()] == labels
        expect = oracle.bits_build(db.rows, cols, db.n_transactions, len(labels))
        assert np.array_equal(bits.download(), expect)
    finally:
        bits.close()


@pytest.mark.parametrize("rows", [[], [[0]], [[1, 1, 2], [], [0, 2], [4294967295, 2]]])
def test_bin_file_bulk_loading(tmp_path, rows):
    """test_transaction_formats.py:185-210: raw baskets (repeats, empty
    baskets) through the bulk BIN path give the per-basket subset counts."""
    path = tmp_path / "transactions.bin"
    write_bin(path, rows)
    miner = cuApriori(path, 1, use_gds=False, use_gpu_io=True)
    assert miner.getInputBackend()["backend"] == "numpy-hip"
    assert miner._Database is None
    miner.mine()
    assert miner.getPatterns() == oracle.brute_force_rows(rows, 1)


def test_bin_file_matches_csv_for_every_miner(tmp_path):
    import generator
    rows = generator.transactional(2000, 40, 5, 42)
    path = tmp_path / "t.bin"
    write_bin(path, rows, universe=40)
    csv = tmp_path / "t.csv"
    csv.write_text("\n".join(",".join(map(str, r)) for r in rows) + "\n")
    for miner_cls in (cuApriori, cuEclat):
        a, b = miner_cls(path, 0.02), miner_cls(str(csv), 0.02)
        a.mine()
        b.mine()
        assert a.getPatterns() == b.getPatterns() and len(a.getPatterns()) > 40


@pytest.mark.parametrize("sorted_ts", [True, False])
def test_temporal_bin_file_mines_like_the_rows(tmp_path, sorted_ts):
    """CUPAMIT1 through the device build (every row kept, rows ordered by
    timestamp) == the row-list input == brute force (test_periodic_miners.py:12-25)."""
    rng = random.Random(17)
    rows = []
    for t in range(1, 400):
        stamp = t if sorted_ts else rng.randrange(1, 200)
        rows.append([stamp, *[rng.randrange(1, 12) for _ in range(rng.randrange(6))]])
    path = tmp_path / "t.bin"
    write_temporal_bin(path, rows)
    a, b = cuPFEclat(path, 12, 40), cuPFEclat(rows, 12, 40)
    assert a.getInputBackend()["backend"] == "numpy-hip"
    a.mine()
    b.mine()
    got = a.getPatterns()
    assert got == b.getPatterns() == oracle.brute_periodic(rows, 12, 40)
    assert len(got) > 10
