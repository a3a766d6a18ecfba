"""cuPFEclat: periodic-frequent itemsets (support >= minSup, periodicity <= maxPer).

Surface of ``cuPFEclat.py:146-175``.  Both conditions are anti-monotone, so
the reference's depth-first search visits exactly the accepted itemsets;
``mine()`` finds the same set with the device search (``cpm_pfeclat_mine``),
whose support kernel walks the set bits of each AND-ed row tile, gathers the
occurrence timestamps and folds (first, last, largest gap, count) partials
in row order.
"""

from __future__ import annotations

import sys
import time
from pathlib import Path

if __name__ == "__main__" and not __package__:
    sys.path.insert(0, str(Path(__file__).resolve().parents[2]))
    __package__ = "cupami_amd.periodicFrequentPattern"

from ._temporal import TemporalInput


class cuPFEclat(TemporalInput):
    """Periodic-frequent Eclat (Kiran, Kitsuregawa and Reddy, JSS 2016)."""

    _BATCH_SIZE = 4096

    def _usesDeviceSearch(self):
        # a subclass with its own _evaluate or _accept (as the reference's
        # partial-periodic miner reuses this mine(), cuPPPEclat.py:139) gets
        # the host-scheduled search below, which calls them
        cls = type(self)
        return cls._evaluate is TemporalInput._evaluate and cls._accept is TemporalInput._accept

    def mine(self):
        self._startTime = time.perf_counter()
        ctx = self._getCudaModule()
        labels, _, _ = self._prepare(ctx)
        if labels and not self._usesDeviceSearch():
            try:
                self._mineHooked(ctx, labels)
            finally:
                self._temporalBits.close()
                self._temporalBits = None
        elif labels:
            try:
                patterns = ctx.pfeclat_mine(self._temporalBits, self._hostTimestamps,
                                            self._endTimestamp, self._minSup,
                                            self._maxPerFloor())
            finally:
                self._temporalBits.close()
                self._temporalBits = None
            self._keepDeviceResult(patterns, labels)
        ctx.synchronize()
        self._endTime = time.perf_counter()

    def _judge(self, ctx, candidates):
        """The candidates ``_accept`` keeps, evaluated in batches of
        ``_BATCH_SIZE`` (``cuPFEclat.py:151-168``)."""
        kept = []
        for lo in range(0, len(candidates), self._BATCH_SIZE):
            batch = candidates[lo:lo + self._BATCH_SIZE]
            kept.extend(p for p, (sup, per) in zip(batch, self._evaluate(ctx, batch)) if self._accept(p, sup, per))
        return kept

    def _mineHooked(self, ctx, labels):
        """Depth-first over prefix classes (``cuPFEclat.py:146-175``): a
        class holds the accepted extensions of its prefix; each member's child
        class tries the members after it."""
        frequent = [p[0] for p in self._judge(ctx, [(label,) for label in labels])]
        self._growClass(ctx, (), frequent)

    def _growClass(self, ctx, prefix, members):
        for i, head in enumerate(members[:-1]):
            child = prefix + (head,)
            kept = [p[-1] for p in self._judge(ctx, [child + (tail,) for tail in members[i + 1:]])]
            if len(kept) > 1:
                self._growClass(ctx, child, kept)


if __name__ == "__main__":
    if len(sys.argv) not in (5, 6):
        sys.exit("Usage: python cuPFEclat.py <inputFile> <outputFile> <minSup> <maxPer> [sep]")

    def _num(text):
        try:
            return int(text)
        except ValueError:
            return float(text)

    miner = cuPFEclat(sys.argv[1], _num(sys.argv[3]), _num(sys.argv[4]),
                      sys.argv[5] if len(sys.argv) == 6 else None)
    miner.mine()
    miner.save(sys.argv[2])
    miner.printResults()
