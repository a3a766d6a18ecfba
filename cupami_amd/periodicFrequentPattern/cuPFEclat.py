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

    def mine(self):
        self._startTime = time.perf_counter()
        ctx = self._getCudaModule()
        labels, _, _ = self._prepare(ctx)
        if labels:
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
