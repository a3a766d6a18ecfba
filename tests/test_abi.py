"""Host-side checks that need no GPU: the C ABI library loads and exports
every symbol its header declares, the package imports without a device, the
miners fail loudly without one, and the host input logic matches the
reference's."""

import ctypes
import math
import re
import struct
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "cupami_hip.h"
LIB = ROOT / "cupami_amd" / "libcupami_hip.so"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(cpm_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    from cupami_amd import _hip
    assert set(names) == set(_hip.SIGNATURES)


def test_package_imports_without_gpu():
    import cupami_amd
    assert {"cuApriori", "cuEclat", "cuFPGrowth", "cuPFEclat"} <= set(cupami_amd.__all__)
    assert cupami_amd.Apriori is cupami_amd.cuApriori
    assert cupami_amd.Eclat is cupami_amd.cuEclat


def test_missing_library_is_an_import_error(monkeypatch):
    from cupami_amd import _hip
    monkeypatch.setattr(_hip, "LIB_PATH", ROOT / "does-not-exist.so")
    monkeypatch.setattr(_hip, "_lib", None)
    with pytest.raises(ImportError):
        _hip.load()
    from cupami_amd import cuEclat
    miner = cuEclat([[1, 2]], 1)
    with pytest.raises(ImportError):
        miner.mine()


def test_thresholds_and_cleaning():
    from cupami_amd import cuApriori
    assert cuApriori._getSupportCount(0.01, 10000) == 100
    assert cuApriori._getSupportCount(1.0, 7) == 7
    assert cuApriori._getSupportCount(0.25, 33) == math.ceil(0.25 * 33)
    assert cuApriori._getSupportCount(2.5, 10) == 2
    assert cuApriori._getSupportCount(100, 10) == 100
    assert cuApriori._cleanTransaction([" a", "a", "", 3, "3 ", "b"]) == ("a", "3", "b")


def test_coo_follows_string_label_order():
    from cupami_amd import cuApriori
    miner = cuApriori([[2, 10, 1], [10], [], [3, 2]], 1)
    labels, rows, cols = miner._getDevicePairs()
    assert labels == ["1", "10", "2", "3"]
    assert miner._transactionCount == 3
    assert rows.tolist() == [0, 0, 0, 1, 2, 2]
    assert [labels[c] for c in cols] == ["2", "10", "1", "10", "3", "2"]


def test_bin_reader_bulk_and_cpu(tmp_path):
    from cupami_amd import cuApriori
    rows = [[0, 1, 2], [], [1, 2, 2], [3]]
    path = tmp_path / "t.bin"
    offsets = np.cumsum([0] + [len(r) for r in rows]).astype("<u8")
    ids = np.array([i for r in rows for i in r], "<u4")
    with path.open("wb") as f:
        f.write(struct.pack("<8sQQQ", b"CUPAMI01", len(rows), ids.size, 3))
        f.write(offsets.tobytes())
        f.write(ids.tobytes())
    bulk = cuApriori(path, 1)
    cpu = cuApriori(path, 1, use_gpu_io=False)
    assert bulk._transactionCount == cpu._transactionCount == 3
    lb, rb, cb = bulk._getDevicePairs()
    lc, rc, cc = cpu._getDevicePairs()
    assert lb == lc
    pairs_b = sorted(set(zip(rb.tolist(), cb.tolist())))
    pairs_c = sorted(set(zip(rc.tolist(), cc.tolist())))
    assert pairs_b == pairs_c
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"NOTMAGIC" + bytes(24))
    with pytest.raises(ValueError):
        cuApriori(bad, 1)


def test_csv_and_gpu_text_requirement(tmp_path):
    from cupami_amd import cuApriori
    path = tmp_path / "t.csv"
    path.write_text('"apple,pear"\tbanana\nbanana\tcherry\n')
    miner = cuApriori(path, 1, sep="\t")
    assert miner._Database == [("apple,pear", "banana"), ("banana", "cherry")]
    with pytest.raises(RuntimeError):
        cuApriori(path, 1, use_gpu_io=True)
    with pytest.raises(ValueError):
        cuApriori(path, 1, use_gds=True)


def test_temporal_reader_rules():
    from cupami_amd import cuPFEclat
    rows = [[5, "a"], [1, "b", "b"], [3], [1, "a"]]
    miner = cuPFEclat(rows, 1, 10)
    assert miner._timestamps == [1, 1, 3, 5]
    assert miner._Database == [("b",), ("a",), (), ("a",)]
    assert miner._transactionCount == 4
    with pytest.raises(ValueError):
        cuPFEclat([[1.5, "a"]], 1, 1)
    with pytest.raises(ValueError):
        cuPFEclat([[5, "a"]], 1, 1, endTime=4)
    with pytest.raises(ValueError):
        cuPFEclat([[5, "a"]], 0, 1)
    frac = cuPFEclat([[2, "a"], [10, "a"]], 0.4, 0.4)
    assert frac._maxPer == 0.4 * 10 and frac._maxPerFloor() == 4


def test_datagen_is_seeded_and_shaped():
    from cupami_amd import datagen
    a = datagen.transactional(2000, 50, 6.0, 1.1, seed=3)
    b = datagen.transactional(2000, 50, 6.0, 1.1, seed=3)
    assert a.sha256() == b.sha256()
    assert a.rows.dtype == np.uint64 and a.items.dtype == np.uint32
    assert a.items.min() >= 1 and a.items.max() <= 50
    # no repeated item in a basket
    pairs = set(zip(a.rows.tolist(), a.items.tolist()))
    assert len(pairs) == a.rows.size
    t = datagen.temporal(1000, 20, 4.0, 1.0, seed=4)
    assert np.all(np.diff(t.timestamps) >= 1) and np.all(np.diff(t.timestamps) <= 3)


def test_temporal_bin_arrays_match_the_row_reader(tmp_path):
    """A CUPAMIT1 file read as arrays (for the device build) gives the same
    timestamps, row order and store as the row reader (_temporal.py:125-139):
    timestamp-only rows kept, rows stably sorted by timestamp."""
    import random
    import oracle
    from cupami_amd import cuPFEclat
    from helpers import write_temporal_bin
    rng = random.Random(2)
    for trial in range(12):
        rows = [[rng.choice([1, 2, 3, 5, 8, 13, 21, 34]) if trial % 2 else t,
                 *[rng.choice([0, 3, 9, 10, 11, 100]) for _ in range(rng.randrange(5))]]
                for t in range(1, rng.randrange(2, 60))]
        path = tmp_path / "t.bin"
        write_temporal_bin(path, rows)
        a, b = cuPFEclat(path, 1, 10), cuPFEclat(rows, 1, 10)
        assert a._Database is None and a._transactionCount == b._transactionCount == len(rows)
        assert list(a._timestamps) == b._timestamps and a._endTimestamp == b._endTimestamp
        la, ra, ca = a._getDevicePairs()
        lb, rb, cb = b._getDevicePairs()
        assert la == lb
        n = len(rows)
        assert np.array_equal(oracle.bits_build(ra, ca, n, len(la)), oracle.bits_build(rb, cb, n, len(lb)))
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"CUPAMIT1" + struct.pack("<QQQ", 1, 1, 3) + struct.pack("<q2QI", 5, 0, 1, 7))
    with pytest.raises(ValueError):
        cuPFEclat(bad, 1, 10)
