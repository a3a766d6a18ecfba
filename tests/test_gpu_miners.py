"""Miner parity through the public classes: restated reference tests
(tests/test_cu_eclat.py, test_cu_fpgrowth.py, test_periodic_miners.py,
test_miner_examples.py) against brute force and recorded outputs."""

import random
import subprocess
import sys
from itertools import combinations
from pathlib import Path

import pytest

import generator
import oracle
from cupami_amd import cuApriori, cuEclat, cuFPGrowth, cuPFEclat
from helpers import kats, random_rows

pytestmark = pytest.mark.gpu
K = kats()
MINERS = (cuApriori, cuEclat, cuFPGrowth)


class HostEclat(cuEclat):
    """The host-scheduled search (one operator call per batch)."""
    _DEVICE_SEARCH = False


@pytest.mark.parametrize("seed", range(5))
def test_randomized_against_brute_force(seed):
    rng = random.Random(seed)
    rows = [rng.sample(range(8), rng.randrange(9)) for _ in range(33 + seed)]
    for support in (1, 7, 0.25, 1.0, 100):
        for cls in (*MINERS, HostEclat):
            miner = cls(rows, support)
            miner.mine()
            assert miner.getPatterns() == oracle.brute_force(rows, miner._minSup), (cls, support)


def test_dense_and_tiny_batches():
    rows = [list(range(8))] * 65
    for cls in (*MINERS, HostEclat):
        miner = cls(rows, 1)
        miner._PAIR_BATCH_SIZE = 3
        miner._INTERSECTION_BATCH_BYTES = 12
        miner.mine()
        assert len(miner.getPatterns()) == 255
        assert miner.getPatterns() == oracle.brute_force(rows, 1)


def test_empty_duplicate_and_large_ids():
    for rows in ([], [[], []], [[0, 0, 1], [], [1, 4294967295], [0, 1, 4294967295]],
                 [["apple,pear", "banana"], ["banana", "cherry"], [" "]]):
        for cls in (*MINERS, HostEclat):
            miner = cls(rows, 1)
            miner.mine()
            assert miner.getPatterns() == oracle.brute_force(rows, 1)
    with pytest.raises(ValueError, match="positive minimum support"):
        cuFPGrowth([[1]], 0).mine()


def test_recorded_outputs():
    for name in ("quickstart_10000x50_avg6_seed42_minsup0.01",
                 "notebook_100x8_avg3_seed42_minsup0.1"):
        v = K["frequent"][name]
        rows = generator.transactional(*v["data"]["args"])
        results = []
        for cls in MINERS:
            miner = cls(rows, v["minSup"])
            miner.mine()
            pats = miner.getPatterns()
            assert len(pats) == v["count"]
            for key, sup in v["some"].items():
                assert pats[key] == sup
            results.append(pats)
        assert results[0] == results[1] == results[2]
    v = K["frequent"]["fpgrowth_paper_example"]
    expected = dict(v["expected_base"])
    for length in range(2, 5):
        for p in combinations(v["expected_all_subsets_of"], length):
            expected["\t".join(p)] = v["expected_subset_support"]
    for cls in MINERS:
        miner = cls(v["rows"], v["minSup"])
        miner.mine()
        assert miner.getPatterns() == expected


def test_singletons_only():
    v = K["frequent"]["quickstart_10000x50_minsup0.02_singletons_only"]
    rows = generator.transactional(*v["data"]["args"])
    miner = cuApriori(rows, v["minSup"])
    miner.mine()
    assert miner.getPatterns() == oracle.brute_force_rows(rows, miner._minSup)
    assert all("\t" not in k for k in miner.getPatterns())


def test_repeat_mining_and_outputs(tmp_path):
    miner = cuEclat([[1, 2], [2, 3], [1, 2, 3]], 1)
    miner.mine()
    first = miner.getPatterns()
    miner.mine()
    assert miner.getPatterns() == first
    frame = miner.getPatternsAsDataFrame()
    assert dict(zip(frame.Patterns, frame.Support)) == first
    out = tmp_path / "patterns.txt"
    miner.save(out)
    saved = dict((k, int(c)) for k, c in (l.rsplit(" : ", 1) for l in out.read_text().splitlines()))
    assert saved == first
    assert miner.getMiningTime() >= 0 and miner.getTotalRuntime() >= miner.getMiningTime()


def test_cli(tmp_path):
    v = K["frequent"]["cli_csv"]
    src = tmp_path / "transactions.csv"
    dst = tmp_path / "patterns.txt"
    src.write_text(v["csv"])
    root = Path(__file__).resolve().parents[1]
    run = subprocess.run([sys.executable, str(root / "cupami_amd/frequentPattern/cuEclat.py"),
                          str(src), str(dst), str(v["minSup"])], capture_output=True, text=True,
                         timeout=120)
    assert run.returncode == 0, run.stderr
    assert dst.read_text() == v["saved"]


def test_periodic_recorded_and_kats():
    v = K["periodic"]["notebook_temporal_100x8_seed42"]
    rows = generator.temporal(*v["data"]["args"])
    miner = cuPFEclat(rows, v["minSup"], v["maxPer"])
    miner.mine()
    pats = miner.getPatterns()
    assert len(pats) == v["count"]
    for key, val in v["some"].items():
        assert pats[key] == tuple(val)

    v = K["periodic"]["jss2016_paper_example"]
    rows = [[i + 1, *b] for i, b in enumerate(v["baskets"])]
    miner = cuPFEclat(rows, v["minSup"], v["maxPer"])
    miner.mine()
    assert miner.getPatterns() == oracle.brute_periodic(rows, 3, 3)
    assert len(miner.getPatterns()) == 9 and miner.getPatterns()["a\tb"] == (4, 3)

    v = K["periodic"]["manual_temporal"]
    for ms in v["minSup"]:
        miner = cuPFEclat(v["rows"], ms, maxPer=v["maxPer"])
        miner.mine()
        assert miner.getPatterns() == {k: tuple(x) for k, x in v["expected"].items()}

    v = K["periodic"]["fractional_thresholds"]
    miner = cuPFEclat(v["rows"], v["minSup"], v["maxPer"])
    miner.mine()
    assert miner.getPatterns() == {k: tuple(x) for k, x in v["expected"].items()}


def test_periodic_random_and_boundaries():
    for seed in range(4):
        rng = random.Random(seed)
        rows = [[i * 3, *rng.sample(range(6), rng.randint(0, 6))] for i in range(1, 19)]
        rng.shuffle(rows)
        for support, period in [(1, 54), (3, 15), (5, 9)]:
            miner = cuPFEclat(rows, support, period)
            miner.mine()
            assert miner.getPatterns() == oracle.brute_periodic(rows, support, period)
    cases = [([[9, "a"], [0, "a", "a", "0"], [4], [9, "a", "b"], [15]], 1, 10, None),
             ([[2, "x"], [3, "x"]], 1, 5, 20), ([], 1, 0, None), ([[0, "x"], [0, "y"]], 1, 0, None),
             ([[2**53 + 1, "x"], [2**53 + 3, "x"]], 1, 2**53 + 1, None),
             ([[i, "a", "b", *(["c"] if i % 2 else [])] for i in range(1, 66)], 1, 4, None)]
    for rows, support, period, end in cases:
        miner = cuPFEclat(rows, support, period, endTime=end)
        miner.mine()
        assert miner.getPatterns() == oracle.brute_periodic(rows, support, period, end)
        miner.mine()
        assert miner.getPatterns() == oracle.brute_periodic(rows, support, period, end)


def test_overridden_hooks_take_the_host_search():
    """A subclass that overrides cuFPGrowth._storePattern or cuPFEclat's
    _evaluate / _accept gets a search that calls them: pass-through overrides
    give the device search's result, and a False from _storePattern prunes
    that pattern's conditional tree (cuFPGrowth.py:205)."""
    from cupami_amd import cuFPGrowth, cuPFEclat
    rng = random.Random(12)
    rows = random_rows(rng, 400, 9, 6)

    class Recording(cuFPGrowth):
        def _storePattern(self, item_ids, support, labels):
            super()._storePattern(item_ids, support, labels)

    class Pruning(cuFPGrowth):
        def _storePattern(self, item_ids, support, labels):
            super()._storePattern(item_ids, support, labels)
            return False

    full = cuFPGrowth(rows, 5)
    full.mine()
    rec = Recording(rows, 5)
    assert not rec._usesDeviceSearch()
    rec.mine()
    assert rec.getPatterns() == full.getPatterns() and len(full.getPatterns()) > 20
    pruned = Pruning(rows, 5)
    pruned.mine()
    singles = {k: v for k, v in full.getPatterns().items() if "\t" not in k}
    assert pruned.getPatterns() == singles

    trows = [[rng.randrange(1, 60), *rng.sample(range(8), rng.randrange(4))] for _ in range(300)]

    class Evaluating(cuPFEclat):
        def _evaluate(self, ctx, candidates):
            return super()._evaluate(ctx, candidates)

    class Short(cuPFEclat):
        def _accept(self, pattern, support, period):
            return len(pattern) <= 2 and super()._accept(pattern, support, period)

    pf = cuPFEclat(trows, 10, 20)
    pf.mine()
    ev = Evaluating(trows, 10, 20)
    ev.mine()
    assert ev.getPatterns() == pf.getPatterns() and len(pf.getPatterns()) > 10
    sh = Short(trows, 10, 20)
    sh.mine()
    assert sh.getPatterns() == {k: v for k, v in pf.getPatterns().items() if k.count("\t") <= 1}
