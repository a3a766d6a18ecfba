"""cuFPGrowth: frequent itemsets with the FP-growth surface.

Surface of ``cuFPGrowth.py:161-230`` (Han, Pei and Yin, SIGMOD 2000): same
constructor, ``mine()``, ``_storePattern`` hook and CLI; a positive minimum
support count is required on non-empty input (``cuFPGrowth.py:173-174``).
``mine()`` runs ``cpm_fpgrowth_mine``: pair dedupe, the weighted FP-tree
build and every conditional tree are device kernels (sorts and segmented
group-bys, ``csrc/fptree.hip``); the depth-first order over header items is
the reference's.
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
    """FP-growth (Han, Pei and Yin, SIGMOD 2000) with device FP-trees."""

    def mine(self):
        self._startTime = time.perf_counter()
        ctx = self._getCudaModule()
        self._finalPatterns = {}
        items, rows, columns = self._getDevicePairs(ctx)
        if items and self._minSup < 1:
            raise ValueError("FP-growth requires a positive minimum support count")
        if items:
            patterns = ctx.fpgrowth_mine(rows, columns, self._transactionCount, len(items),
                                         self._minSup)
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
