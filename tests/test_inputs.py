"""Input readers on the CPU: Parquet files read as basket arrays give the
same store as the reference's row reading (cuApriori.py:406-426,
extras/_temporal_io.py:87-120), with its checks; BIN arrays and label rows
agree; label cleaning keeps the reference's rules."""

import random

import numpy as np
import pytest

import oracle
from cupami_amd import cuApriori, cuPFEclat
from cupami_amd.frequentPattern import _inputs

pa = pytest.importorskip("pyarrow")
pq = pytest.importorskip("pyarrow.parquet")


def _store(miner):
    labels, rows, cols = miner._getDevicePairs()
    return labels, oracle.bits_build(rows, cols, miner._transactionCount, len(labels))


def _write_items(path, rows, chunks=1):
    arr = pa.array(rows, type=pa.list_(pa.uint32()))
    table = pa.Table.from_arrays([arr], names=["items"])
    pq.write_table(table, path, row_group_size=max(1, len(rows) // chunks))


def test_clean_labels():
    assert _inputs.clean_labels([" a", "a", "", 3, "3 ", "b", "\t"]) == ("a", "3", "b")
    assert _inputs.kept_label_rows([[], [" "], [1, 1]]) == [("1",)]


def test_parquet_arrays_equal_rows(tmp_path):
    rng = random.Random(5)
    for trial in range(6):
        rows = [[rng.choice([0, 1, 2, 9, 10, 11, 100, 2**32 - 1]) for _ in range(rng.randrange(6))]
                for _ in range(rng.randrange(1, 200))]
        path = tmp_path / f"t{trial}.parquet"
        _write_items(path, rows, chunks=3)
        a = cuApriori(path, 1)
        b = cuApriori(path, 1, use_gpu_io=False)
        c = cuApriori(rows, 1)
        assert a._Database is None and a.getInputBackend()["backend"] == "numpy-hip"
        assert a._transactionCount == b._transactionCount == c._transactionCount
        la, sa = _store(a)
        lb, sb = _store(b)
        lc, sc = _store(c)
        assert la == lb == lc
        assert np.array_equal(sa, sb) and np.array_equal(sa, sc)


def test_parquet_checks(tmp_path):
    bad = tmp_path / "noitems.parquet"
    pq.write_table(pa.Table.from_arrays([pa.array([1, 2])], names=["x"]), bad)
    with pytest.raises(ValueError, match="missing items column"):
        cuApriori(bad, 1)
    wrong = tmp_path / "wrong.parquet"
    pq.write_table(pa.Table.from_arrays([pa.array([[1], [2]], type=pa.list_(pa.int64()))], names=["items"]), wrong)
    with pytest.raises(ValueError, match="list of uint32"):
        cuApriori(wrong, 1)
    nulls = tmp_path / "nulls.parquet"
    pq.write_table(pa.Table.from_arrays([pa.array([[1, None], [2]], type=pa.list_(pa.uint32()))],
                                        names=["items"]), nulls)
    with pytest.raises(ValueError, match="null"):
        cuApriori(nulls, 1)


def test_temporal_parquet_arrays_equal_rows(tmp_path):
    rng = random.Random(6)
    for trial in range(6):
        n = rng.randrange(1, 80)
        ts = [rng.choice([1, 2, 3, 5, 8]) if trial % 2 else i + 1 for i in range(n)]
        items = [[rng.choice([0, 3, 9, 10, 11]) for _ in range(rng.randrange(5))] for _ in range(n)]
        path = tmp_path / f"t{trial}.parquet"
        pq.write_table(pa.Table.from_arrays([pa.array(ts, type=pa.int64()),
                                             pa.array(items, type=pa.list_(pa.uint32()))],
                                            names=["timestamp", "items"]), path)
        a = cuPFEclat(path, 1, 10)
        b = cuPFEclat([[t, *r] for t, r in zip(ts, items)], 1, 10)
        assert a._Database is None and a._transactionCount == b._transactionCount == n
        assert list(a._timestamps) == b._timestamps and a._endTimestamp == b._endTimestamp
        la, sa = _store(a)
        lb, sb = _store(b)
        assert la == lb and np.array_equal(sa, sb)
    bad = tmp_path / "bad.parquet"
    pq.write_table(pa.Table.from_arrays([pa.array([1], type=pa.int32()),
                                         pa.array([[1]], type=pa.list_(pa.uint32()))],
                                        names=["timestamp", "items"]), bad)
    with pytest.raises(ValueError, match="int64"):
        cuPFEclat(bad, 1, 10)


def test_basket_array_pairs_string_order():
    offsets = np.array([0, 3, 3, 5], np.uint64)
    ids = np.array([2, 10, 1, 10, 3], np.uint32)
    labels, rows, cols = _inputs.basket_array_pairs(offsets, ids)
    assert labels == ["1", "10", "2", "3"]
    assert rows.tolist() == [0, 0, 0, 1, 1]
    assert [labels[c] for c in cols] == ["2", "10", "1", "10", "3"]
    labels, rows, cols = _inputs.basket_array_pairs(offsets, ids, keep_empty=True)
    assert rows.tolist() == [0, 0, 0, 2, 2]
