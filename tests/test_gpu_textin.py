"""Delimited text parsed on the device (cpm_baskets_parse_text) against the
reference's host reader (csv + cleaning, cuApriori.py:428-450): for text of
decimal ids the device path gives the same transaction count, labels and
store; text it does not take goes to the host reader with the same result."""

import random

import numpy as np
import pytest

import oracle
from cupami_amd import cuApriori, cuEclat, datagen

pytestmark = pytest.mark.gpu


def _store(miner):
    labels, rows, cols = miner._getDevicePairs()
    return labels, oracle.bits_build(rows, cols, miner._transactionCount, len(labels))


def _random_text(rng, sep, n_lines):
    ids = [0, 1, 2, 9, 10, 11, 99, 100, 12345, 4294967295]
    lines = []
    for _ in range(n_lines):
        fields = []
        for _ in range(rng.randrange(0, 7)):
            f = str(rng.choice(ids))
            pad = rng.random()
            if pad < 0.2:
                f = " " + f + "  "
            elif pad < 0.25:
                f = ""           # an empty field
            fields.append(f)
        lines.append(sep.join(fields))
    end = rng.choice(["\n", "\r\n"])
    text = end.join(lines)
    if rng.random() < 0.5:
        text += end
    return text


@pytest.mark.parametrize("sep", [",", "\t", " "])
def test_device_text_equals_host_reader(tmp_path, sep):
    rng = random.Random(ord(sep))
    for trial in range(8):
        text = _random_text(rng, sep, rng.randrange(0, 300))
        path = tmp_path / f"t{trial}.txt"
        path.write_bytes(text.encode())
        # a comma is what sniffing finds when the first line has no tab; the
        # other separators are given
        kw = {} if sep == "," else {"sep": sep}
        dev = cuApriori(path, 1, **kw)
        host = cuApriori(path, 1, use_gpu_io=False, **kw)
        print(f"textin: sep {sep!r} trial {trial} {len(text)} bytes", flush=True)
        assert dev.getInputBackend()["backend"] == "hip-text"
        assert dev._transactionCount == host._transactionCount
        ld, sd = _store(dev)
        lh, sh = _store(host)
        assert ld == lh and np.array_equal(sd, sh)


def test_text_the_device_leaves_to_the_host(tmp_path):
    cases = ['1,"2,3"\n4\n', "1,007\n", "1,a\n2\n", "1\r2\n", "1,é2\n", "1,-2\n", "1,4294967296\n",
             "1," + " " * 70 + "2\n"]
    for i, text in enumerate(cases):
        path = tmp_path / f"c{i}.csv"
        path.write_bytes(text.encode())
        auto = cuApriori(path, 1)
        host = cuApriori(path, 1, use_gpu_io=False)
        assert auto.getInputBackend()["backend"] == "cpu", text
        assert auto._Database == host._Database
        with pytest.raises(RuntimeError):
            cuApriori(path, 1, use_gpu_io=True)


def test_device_text_mines_like_rows(tmp_path):
    db = datagen.transactional(20_000, 60, 6.0, 1.0, seed=31)
    baskets = db.baskets()
    path = tmp_path / "db.csv"
    path.write_text("\n".join(",".join(map(str, b)) for b in baskets) + "\n")
    dev = cuEclat(path, 0.01)
    rows = cuEclat(baskets, 0.01)
    assert dev.getInputBackend()["backend"] == "hip-text"
    dev.mine()
    rows.mine()
    assert dev.getPatterns() == rows.getPatterns() and len(dev.getPatterns()) > 100


def test_device_text_at_1m_1k(tmp_path):
    """The 1,000,000 x 1,000 workload as a text file (10 M ids): the store
    parsed and built on the device == the oracle's."""
    db = datagen.eclat_1m_1k()
    offsets = np.searchsorted(db.rows, np.arange(db.n_transactions + 1, dtype=np.uint64))
    lines = np.split(db.items.astype(str), offsets[1:-1])
    path = tmp_path / "c2.txt"
    path.write_text("\n".join(map("\t".join, lines)) + "\n")
    dev = cuApriori(path, 1)
    assert dev.getInputBackend()["backend"] == "hip-text" and dev._transactionCount == db.n_transactions
    labels, cols = datagen.label_columns(db.items)
    ctx = dev._getCudaModule()
    items, bits = dev._getDeviceStore(ctx)
    assert items == labels
    assert np.array_equal(bits.download(), oracle.bits_build(db.rows, cols, db.n_transactions, len(labels)))
