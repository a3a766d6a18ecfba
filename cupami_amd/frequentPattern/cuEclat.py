"""cuEclat: frequent itemsets by prefix classes over vertical bitsets.

Surface of ``cuEclat.py:81-262`` (Zaki, KDD 1997 / TKDE 2000): same
constructor, tunables (``_PAIR_BATCH_SIZE``, ``_INTERSECTION_BATCH_BYTES``),
hooks (``_candidatePairs``, ``_extensionBatch``, ``_searchClass``) and CLI.

``mine()`` runs the whole search on the device (``cpm_eclat_mine``): level-2
pair supports come from the Gram D^T D on the matrix cores (FP4/E2M1 0/1
operands with FP32 accumulation, cut into K slices below 2^24 transactions
so every partial sum is an exact integer), and the prefix
classes -- child class of K_j = the later kept siblings, exactly the
reference's ``_searchClass`` frames -- are generated, pair-pruned and counted
by device kernels.  A subclass that overrides ``_candidatePairs`` or
``_extensionBatch`` (as the reference's spatial miner does) gets the
host-scheduled search below, which calls the operator entry points once per
batch the way the reference does.
"""

from __future__ import annotations

import sys
import time
from itertools import combinations, islice
from pathlib import Path

import numpy as np

if __name__ == "__main__" and not __package__:
    sys.path.insert(0, str(Path(__file__).resolve().parents[2]))
    __package__ = "cupami_amd.frequentPattern"

from .cuApriori import cuApriori, _cli
from . import _bitsets


class cuEclat(cuApriori):
    """Eclat over vertical tidset bitsets (Zaki, TKDE 12(3), 2000)."""

    _PAIR_BATCH_SIZE = 65536
    _INTERSECTION_BATCH_BYTES = 64 * 1024 * 1024
    # Run the search on the device unless a subclass reshapes the classes.
    _DEVICE_SEARCH = True

    def _usesDeviceSearch(self):
        cls = type(self)
        return (self._DEVICE_SEARCH and cls._candidatePairs is cuEclat._candidatePairs
                and cls._extensionBatch is cuEclat._extensionBatch)

    def mine(self):
        if self._usesDeviceSearch():
            return super().mine()
        self._startTime = time.perf_counter()
        ctx = self._getCudaModule()
        self._finalPatterns = {}
        items, bits = self._getDeviceStore(ctx)
        if bits is not None:
            try:
                self._mineHost(ctx, items, bits)
            finally:
                bits.close()
        ctx.synchronize()
        self._endTime = time.perf_counter()

    # host-scheduled search (cuEclat.py:168-245) --------------------------------
    def _mineHost(self, ctx, items, bits):
        supports = _bitsets.count_supports(ctx, bits, [[i] for i in range(len(items))])
        frequent = []
        for index, support in enumerate(supports):
            if support >= self._minSup:
                frequent.append(index)
                self._finalPatterns[(items[index],)] = int(support)
        classes = {}
        pairs = self._candidatePairs(frequent, items)
        while True:
            batch = list(islice(pairs, self._PAIR_BATCH_SIZE))
            if not batch:
                break
            for (left, right), support in zip(batch, _bitsets.count_supports(ctx, bits, batch)):
                if support >= self._minSup:
                    self._finalPatterns[(items[left], items[right])] = int(support)
                    classes.setdefault(left, []).append(right)
        store = None
        for left, extensions in classes.items():
            if len(extensions) < 2:
                continue
            if store is None:
                store = bits.download()
            labels = [items[i] for i in extensions]
            vertical, _ = _bitsets.intersect_extensions(ctx, store[left], store[extensions])
            self._searchClass(ctx, (items[left],), labels, vertical)

    def _candidatePairs(self, frequent, items):
        return combinations(frequent, 2)

    def _extensionBatch(self, ctx, prefix, labels, bitsets, start, stop):
        return labels[start:stop], bitsets[start:stop]

    def _searchClass(self, ctx, prefix, labels, bitsets):
        stack = [[prefix, labels, bitsets, 0]]
        words32 = (self._transactionCount + 31) // 32
        batch = max(1, self._INTERSECTION_BATCH_BYTES // max(1, 4 * words32))
        while stack:
            frame = stack[-1]
            current, names_in, vertical, position = frame
            if position >= len(names_in) - 1:
                stack.pop()
                continue
            frame[3] += 1
            nextPrefix = current + (names_in[position],)
            names, chunks = [], []
            for start in range(position + 1, len(names_in), batch):
                stop = min(len(names_in), start + batch)
                batchLabels, batchBits = self._extensionBatch(
                    ctx, nextPrefix, names_in, vertical, start, stop)
                if not batchLabels:
                    continue
                sup, rows, kept = ctx.intersect_extensions(
                    vertical[position], batchBits, keep_min=self._minSup, want_rows=True)
                for offset in kept.tolist():
                    name = batchLabels[offset]
                    self._finalPatterns[nextPrefix + (name,)] = int(sup[offset])
                    names.append(name)
                if len(kept):
                    chunks.append(rows)
            if len(names) > 1:
                stack.append([nextPrefix, names, np.concatenate(chunks), 0])


Eclat = cuEclat


if __name__ == "__main__":
    _cli(cuEclat, sys.argv)
