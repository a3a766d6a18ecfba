"""The CPU oracle against the reference's recorded outputs and against
brute-force enumeration (no GPU)."""

import random
from itertools import combinations

import numpy as np
import pytest

import generator
import oracle
from helpers import coo, csr, kats, temporal_coo, random_rows


K = kats()


def test_generator_reproduces_printed_rows():
    for name, v in K["generator_rows"].items():
        if name.startswith("binary"):
            rows = generator.transactional(*v["args"])
        else:
            rows = generator.temporal(*v["args"], repeat=v["repeat"], missing=v["missing"],
                                      merge=v["merge"])
        assert rows[:len(v["rows"])] == v["rows"], name


@pytest.mark.parametrize("name", ["quickstart_10000x50_avg6_seed42_minsup0.01",
                                  "notebook_100x8_avg3_seed42_minsup0.1"])
def test_eclat_and_apriori_match_recorded_outputs(name):
    v = K["frequent"][name]
    rows = generator.transactional(*v["data"]["args"])
    labels, rr, cc, n = coo(rows)
    bits = oracle.bits_build(rr, cc, n, len(labels))
    minsup = oracle.support_threshold(v["minSup"], n)
    eclat = oracle.labelled(oracle.eclat(bits, minsup), labels)
    assert len(eclat) == v["count"]
    for key, sup in v["some"].items():
        assert eclat[key] == sup
    assert oracle.labelled(oracle.apriori(bits, minsup), labels) == eclat


def test_singletons_only_case():
    v = K["frequent"]["quickstart_10000x50_minsup0.02_singletons_only"]
    rows = generator.transactional(*v["data"]["args"])
    labels, rr, cc, n = coo(rows)
    bits = oracle.bits_build(rr, cc, n, len(labels))
    pats = oracle.eclat(bits, oracle.support_threshold(v["minSup"], n))
    assert all(len(k) == 1 for k in pats)
    pair_max = max(oracle.count_supports(bits, list(combinations(range(len(labels)), 2))))
    assert pair_max < v["max_pair_support_below"]


def test_fpgrowth_paper_and_manual_and_dense():
    v = K["frequent"]["fpgrowth_paper_example"]
    labels, rr, cc, n = coo(v["rows"])
    bits = oracle.bits_build(rr, cc, n, len(labels))
    got = oracle.labelled(oracle.eclat(bits, v["minSup"]), labels)
    expected = dict(v["expected_base"])
    for length in range(2, 5):
        for p in combinations(v["expected_all_subsets_of"], length):
            expected["\t".join(p)] = v["expected_subset_support"]
    assert got == expected

    v = K["frequent"]["manual_counts"]
    labels, rr, cc, n = coo(v["rows"])
    bits = oracle.bits_build(rr, cc, n, len(labels))
    for ms in v["minSup"]:
        got = oracle.labelled(oracle.eclat(bits, oracle.support_threshold(ms, n)), labels)
        assert got == v["expected"]

    v = K["frequent"]["dense_8x65"]
    rows = [v["rows_repeat"][0]] * v["rows_repeat"][1]
    labels, rr, cc, n = coo(rows)
    bits = oracle.bits_build(rr, cc, n, len(labels))
    got = oracle.labelled(oracle.eclat(bits, v["minSup"]), labels)
    assert len(got) == v["count"] and v["contains"] in got


def test_periodic_kats():
    v = K["periodic"]["notebook_temporal_100x8_seed42"]
    rows = generator.temporal(*v["data"]["args"])
    labels, rr, cc, n, ts, end = temporal_coo(rows)
    bits = oracle.bits_build(rr, cc, n, len(labels))
    ms = oracle.support_threshold(v["minSup"], n)
    got = oracle.labelled(oracle.pfeclat(bits, ts, n, end, ms, v["maxPer"]), labels)
    assert len(got) == v["count"]
    for key, val in v["some"].items():
        assert got[key] == tuple(val)

    v = K["periodic"]["jss2016_paper_example"]
    rows = [[i + 1, *b] for i, b in enumerate(v["baskets"])]
    labels, rr, cc, n, ts, end = temporal_coo(rows)
    bits = oracle.bits_build(rr, cc, n, len(labels))
    got = oracle.labelled(oracle.pfeclat(bits, ts, n, end, v["minSup"], v["maxPer"]), labels)
    assert len(got) == v["count"] and got["a\tb"] == tuple(v["some"]["a\tb"])

    v = K["periodic"]["manual_temporal"]
    labels, rr, cc, n, ts, end = temporal_coo(v["rows"])
    bits = oracle.bits_build(rr, cc, n, len(labels))
    for ms in v["minSup"]:
        got = oracle.labelled(oracle.pfeclat(bits, ts, n, end, oracle.support_threshold(ms, n),
                                             v["maxPer"]), labels)
        assert got == {k: tuple(x) for k, x in v["expected"].items()}


@pytest.mark.parametrize("seed", range(5))
def test_eclat_and_apriori_against_brute_force(seed):
    rng = random.Random(seed)
    rows = random_rows(rng, 33 + seed, 8, 8)
    labels, rr, cc, n = coo(rows)
    bits = oracle.bits_build(rr, cc, n, len(labels))
    for support in (1, 7, 0.25, 1.0, 100):
        ms = oracle.support_threshold(support, n)
        expected = oracle.brute_force(rows, ms)
        assert oracle.labelled(oracle.eclat(bits, ms), labels) == expected
        # clApriori additionally requires support > 0; identical for ms >= 1
        assert oracle.labelled(oracle.apriori(bits, ms), labels) == expected
        assert oracle.brute_force_rows(rows, ms) == expected


@pytest.mark.parametrize("seed", range(4))
def test_periodic_against_brute_force(seed):
    rng = random.Random(seed)
    rows = [[i * 3, *rng.sample(range(6), rng.randint(0, 6))] for i in range(1, 19)]
    rng.shuffle(rows)
    labels, rr, cc, n, ts, end = temporal_coo(rows)
    bits = oracle.bits_build(rr, cc, n, len(labels))
    for support, period in [(1, 54), (3, 15), (5, 9)]:
        got = oracle.labelled(oracle.pfeclat(bits, ts, n, end, support, period), labels)
        assert got == oracle.brute_periodic(rows, support, period)


@pytest.mark.parametrize("count", [0, 1, 31, 32, 33, 63, 64, 65, 129])
def test_word_boundaries(count):
    rows = [["a", "a", "b", *(["c"] if i % 3 == 0 else [])] for i in range(count)] + [[]]
    labels, rr, cc, n = coo(rows)
    bits = oracle.bits_build(rr, cc, n, len(labels))
    assert oracle.labelled(oracle.eclat(bits, 1), labels) == oracle.brute_force(rows, 1)


def test_periodic_closed_form_and_plants():
    """period = max(diff([0, ts..., end])); a dropped start or end gap, or a
    gap measured in rows instead of timestamps, fails here."""
    rng = np.random.default_rng(7)
    n = 300
    ts = np.sort(rng.integers(0, 5000, n)).astype(np.int64)
    rows, cols = [], []
    occ = sorted(rng.choice(n, 40, replace=False).tolist())
    for r in occ:
        rows.append(r)
        cols.append(0)
    bits = oracle.bits_build(np.array(rows, np.uint64), np.array(cols, np.uint32), n, 1)
    end = int(ts[-1]) + 17
    sup, per = oracle.periodic_eval(bits, ts, n, end, np.array([[0]], np.uint32))
    pts = [0, *ts[occ].tolist(), end]
    assert sup[0] == 40
    assert per[0] == max(b - a for a, b in zip(pts, pts[1:]))
    # no occurrence: the whole span [0, end]
    bits0 = np.zeros_like(bits)
    sup, per = oracle.periodic_eval(bits0, ts, n, end, np.array([[0]], np.uint32))
    assert sup[0] == 0 and per[0] == end


def test_exact_above_2_53():
    rows = [[2**53 + 1, "x"], [2**53 + 3, "x"]]
    labels, rr, cc, n, ts, end = temporal_coo(rows)
    bits = oracle.bits_build(rr, cc, n, len(labels))
    got = oracle.labelled(oracle.pfeclat(bits, ts, n, end, 1, 2**53 + 1), labels)
    assert got == oracle.brute_periodic(rows, 1, 2**53 + 1)


def _raw_baskets(rng, universe, n):
    # ids across every decimal length up to the universe, repeats and empty
    # baskets included, as a BIN file may hold them
    pool = sorted({0, universe, *(rng.randrange(universe + 1) for _ in range(25)),
                   *(min(universe, 10 ** k) for k in range(10))})
    return [[rng.choice(pool) for _ in range(rng.randrange(7))] for _ in range(n)]


@pytest.mark.parametrize("universe", [9, 100, 12345, 2**32 - 1])
def test_csr_oracle_matches_the_basket_path(universe):
    """The BIN branch of _getDevicePairs (cuApriori.py:381-396) and the
    basket branch (_cleanTransaction + :397-404) must give the same labels,
    row count and store for the same baskets."""
    rng = random.Random(universe)
    for trial in range(30):
        rows = _raw_baskets(rng, universe, rng.randrange(80))
        labels, rr, cc, n = oracle.csr_to_coo(*csr(rows))
        l2, r2, c2, n2 = coo(rows)
        assert labels == l2 and n == n2
        assert np.array_equal(oracle.bits_build(rr, cc, n, len(labels)),
                              oracle.bits_build(r2, c2, n2, len(l2)))


def test_csr_oracle_keep_empty_matches_the_temporal_path():
    rng = random.Random(5)
    for trial in range(20):
        rows = _raw_baskets(rng, 1000, rng.randrange(1, 80))
        labels, rr, cc, n = oracle.csr_to_coo(*csr(rows), keep_empty=True)
        l2, r2, c2, n2, _, _ = temporal_coo([[t, *r] for t, r in enumerate(rows)])
        assert labels == l2 and n == n2 == len(rows)
        assert np.array_equal(oracle.bits_build(rr, cc, n, len(labels)),
                              oracle.bits_build(r2, c2, n2, len(l2)))


@pytest.mark.parametrize("rows", [[], [[0]], [[1, 1, 2], [], [0, 2], [4294967295, 2]]])
def test_csr_oracle_bulk_loading_cases(rows):
    """The bulk BIN cases of test_transaction_formats.py:185-210: mining the
    store built from the file's arrays gives the per-basket subset counts."""
    labels, rr, cc, n = oracle.csr_to_coo(*csr(rows))
    bits = oracle.bits_build(rr, cc, n, len(labels))
    got = oracle.labelled(oracle.eclat(bits, 1), labels) if labels else {}
    assert got == oracle.brute_force_rows(rows, 1)
