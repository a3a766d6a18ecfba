"""Pattern output rendered by the library (cpm_patterns_keys / _save) against
the reference's Python rendering: getPatterns keys "\\t".join(labels)
(cuApriori.py:537-539), save lines "p : s" sorted by label tuple
(cuApriori.py:548-552) and "p : s : per" for periodic results
(_temporal.py:178-181)."""

import random

import numpy as np
import pytest

import oracle
from cupami_amd import _hip, cuApriori, cuEclat, cuPFEclat
from helpers import coo, random_rows, temporal_coo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return _hip.context(0)


def python_keys(patterns, labels):
    return ["\t".join(labels[c] for c in sorted(s)) for s in patterns.itemsets()]


def python_save(patterns, labels, periodic=False):
    sets = [tuple(labels[c] for c in sorted(s)) for s in patterns.itemsets()]
    rows = list(zip(sets, patterns.support.tolist(),
                    patterns.period.tolist() if periodic else [None] * len(sets)))
    lines = []
    for key, sup, per in sorted(rows, key=lambda r: r[0]):
        lines.append(f"{chr(9).join(key)} : {sup}" + (f" : {per}" if periodic else "") + "\n")
    return "".join(lines)


def mined(ctx, rows, minsup, labels=None):
    lab, rr, cc, n = coo(rows)
    bits = ctx.build_bits(rr, cc, n, len(lab))
    try:
        return ctx.eclat_mine(bits, minsup), (labels or lab)
    finally:
        bits.close()


@pytest.mark.parametrize("seed", range(4))
def test_keys_and_save_match_python(ctx, tmp_path, seed):
    rng = random.Random(seed)
    rows = random_rows(rng, 300, 14, 9)
    patterns, labels = mined(ctx, rows, 12)
    assert len(patterns) > 50
    assert patterns.keys(labels) == python_keys(patterns, labels)
    out = tmp_path / "p.txt"
    patterns.save(out, labels)
    assert out.read_text(encoding="utf-8") == python_save(patterns, labels)


def test_unusual_labels(ctx, tmp_path):
    """Multi-byte UTF-8, and labels holding a tab, a newline and every
    separator the binding tries (it then renders in Python)."""
    rng = random.Random(7)
    rows = random_rows(rng, 200, 8, 6)
    patterns, _ = mined(ctx, rows, 5)
    for labels in (["é", "日本", "a\tb", "ß", "x\ny", "🙂", "z", "0"],
                   ["a\n", "b\x00", "c\x01", "d\x1f", "e", "f", "g", "h"]):
        labels = sorted(labels)
        assert patterns.keys(labels) == python_keys(patterns, labels)
        out = tmp_path / "u.txt"
        patterns.save(out, labels)
        assert out.read_text(encoding="utf-8") == python_save(patterns, labels)


def test_unordered_forest(ctx, tmp_path):
    """FP-growth's forest does not ascend: each key is sorted first."""
    rng = random.Random(3)
    rows = random_rows(rng, 250, 10, 7)
    labels, rr, cc, n = coo(rows)
    patterns = ctx.fpgrowth_mine(rr, cc, n, len(labels), 8)
    assert len(patterns) > 30
    assert patterns.keys(labels) == python_keys(patterns, labels)
    out = tmp_path / "f.txt"
    patterns.save(out, labels)
    assert out.read_text(encoding="utf-8") == python_save(patterns, labels)


def test_periodic_save(ctx, tmp_path):
    rng = random.Random(9)
    rows = [[t, *r] for t, r in enumerate(random_rows(rng, 300, 10, 6), start=1)]
    labels, rr, cc, n, ts, end = temporal_coo(rows)
    bits = ctx.build_bits(rr, cc, n, len(labels))
    try:
        patterns = ctx.pfeclat_mine(bits, ts, end, 10, 60)
    finally:
        bits.close()
    assert len(patterns) > 20
    out = tmp_path / "t.txt"
    patterns.save(out, labels)
    assert out.read_text(encoding="utf-8") == python_save(patterns, labels, periodic=True)


@pytest.mark.parametrize("miner_cls", [cuApriori, cuEclat])
def test_miner_outputs_match_the_tuple_dict(tmp_path, miner_cls):
    """getPatterns, the DataFrame and save from the library's rendering equal
    the same outputs built from _finalPatterns, the reference's dict."""
    rng = random.Random(21)
    rows = random_rows(rng, 400, 12, 8)
    a = miner_cls(rows, 15)
    a.mine()
    keys = a.getPatterns()
    frame = a.getPatternsAsDataFrame()
    a.save(tmp_path / "native.txt")
    tuples = a._finalPatterns            # materialises the reference's dict
    assert keys == {"\t".join(p): s for p, s in tuples.items()}
    assert keys == oracle.brute_force(rows, 15)
    assert list(frame.columns) == ["Patterns", "Support"] and frame["Support"].dtype == np.int64
    assert dict(zip(frame["Patterns"], frame["Support"].tolist())) == keys
    a.save(tmp_path / "python.txt")      # now through the dict
    assert (tmp_path / "native.txt").read_text() == (tmp_path / "python.txt").read_text()


def test_pfeclat_outputs_match_the_tuple_dict(tmp_path):
    rng = random.Random(4)
    rows = [[t, *r] for t, r in enumerate(random_rows(rng, 300, 9, 6), start=1)]
    a = cuPFEclat(rows, 10, 50)
    a.mine()
    got = a.getPatterns()
    frame = a.getPatternsAsDataFrame()
    a.save(tmp_path / "native.txt")
    tuples = a._finalPatterns
    assert got == {"\t".join(p): (s, a._periods[p]) for p, s in tuples.items()}
    assert got == oracle.brute_periodic(rows, 10, 50)
    assert list(frame.columns) == ["Patterns", "Support", "Periodicity"]
    a.save(tmp_path / "python.txt")
    assert (tmp_path / "native.txt").read_text() == (tmp_path / "python.txt").read_text()
