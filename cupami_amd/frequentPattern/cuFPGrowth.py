"""cuFPGrowth: frequent itemsets with the FP-growth surface.

Surface of ``cuFPGrowth.py:161-230`` (Han, Pei and Yin, SIGMOD 2000): same
constructor, ``mine()``, ``_storePattern`` hook and CLI; a positive minimum
support count is required on non-empty input (``cuFPGrowth.py:173-174``).
The itemsets and supports are the complete set of frequent itemsets, which
``mine()`` obtains from the device search.
"""

from __future__ import annotations

import sys
import time
from pathlib import Path

if __name__ == "__main__" and not __package__:
    sys.path.insert(0, str(Path(__file__).resolve().parents[2]))
    __package__ = "cupami_amd.frequentPattern"

from .cuApriori import cuApriori, _cli


class cuFPGrowth(cuApriori):
    """FP-growth surface over the device search."""

    def mine(self):
        self._startTime = time.perf_counter()
        ctx = self._getCudaModule()
        self._finalPatterns = {}
        items, rows, columns = self._getDevicePairs(ctx)
        if items and self._minSup < 1:
            raise ValueError("FP-growth requires a positive minimum support count")
        if items:
            bits = ctx.build_bits(rows, columns, self._transactionCount, len(items))
            try:
                patterns = self._deviceSearch(ctx, bits, items)
            finally:
                bits.close()
            for ids, support in zip(patterns.itemsets(), patterns.support.tolist()):
                self._storePattern(ids, support, items)
        ctx.synchronize()
        self._endTime = time.perf_counter()

    def _storePattern(self, item_ids, support, labels):
        pattern = tuple(sorted(labels[index] for index in item_ids))
        self._finalPatterns[pattern] = int(support)


FPGrowth = cuFPGrowth


if __name__ == "__main__":
    _cli(cuFPGrowth, sys.argv)
