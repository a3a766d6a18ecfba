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
from itertools import combinations
from pathlib import Path

import numpy as np

if __name__ == "__main__" and not __package__:
    sys.path.insert(0, str(Path(__file__).resolve().parents[2]))
    __package__ = "cupami_amd.frequentPattern"

from .cuApriori import cuApriori, _cli


class cuFPGrowth(cuApriori):
    """FP-growth (Han, Pei and Yin, SIGMOD 2000) with device FP-trees."""

    def _usesDeviceSearch(self):
        # a subclass that overrides _storePattern (it may prune by returning
        # False, cuFPGrowth.py:205) gets the host-scheduled search below
        return type(self)._storePattern is cuFPGrowth._storePattern

    def mine(self):
        self._startTime = time.perf_counter()
        ctx = self._getCudaModule()
        self._finalPatterns = {}
        items, rows, columns = self._getDevicePairs(ctx)
        if items and self._minSup < 1:
            raise ValueError("FP-growth requires a positive minimum support count")
        if items:
            if self._usesDeviceSearch():
                patterns = ctx.fpgrowth_mine(rows, columns, self._transactionCount, len(items),
                                             self._minSup)
                for ids, support in zip(patterns.itemsets(), patterns.support.tolist()):
                    self._storePattern(ids, support, items)
            else:
                self._mineHooked(ctx, items, rows, columns)
        ctx.synchronize()
        self._endTime = time.perf_counter()

    def _storePattern(self, item_ids, support, labels):
        pattern = tuple(sorted(labels[index] for index in item_ids))
        self._finalPatterns[pattern] = int(support)

    def _mineHooked(self, ctx, labels, rows, columns):
        """Depth-first over conditional FP-trees built on the device
        (``cuFPGrowth.py:187-213``), every pattern handed to ``_storePattern``
        as it is found; a False return skips that pattern's conditional tree."""
        rows, columns = ctx.unique_pairs(rows, columns)
        weights = np.ones(self._transactionCount, dtype=np.uint64)
        root = ctx.fptree_build(rows, columns, weights, len(labels), self._minSup)
        if root is not None:
            self._growFrom(ctx, root, (), labels)

    def _growFrom(self, ctx, tree, suffix, labels):
        """The patterns of ``tree`` extended by ``suffix``: a single path's
        every combination (support = its smallest count), else each header
        item in order, then its conditional tree, before the next item.
        Recursion depth is bounded by the longest pattern."""
        try:
            _, node_items, counts, header, single, _ = tree.arrays()
            if single:
                path = list(zip(node_items[1:].tolist(), counts[1:].tolist()))
                for size in range(1, len(path) + 1):
                    for chosen in combinations(path, size):
                        self._storePattern(suffix + tuple(i for i, _ in chosen),
                                           min(c for _, c in chosen), labels)
                return
            for item, support in header:
                grown = suffix + (item,)
                if self._storePattern(grown, support, labels) is False:
                    continue
                conditional = ctx.fptree_conditional(tree, item, self._minSup)
                if conditional is not None:
                    self._growFrom(ctx, conditional, grown, labels)
        finally:
            tree.close()


FPGrowth = cuFPGrowth


if __name__ == "__main__":
    _cli(cuFPGrowth, sys.argv)
