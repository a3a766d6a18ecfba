"""cuApriori: all frequent itemsets, counted on the gfx950 support path.

Keeps the reference's public surface (``cuApriori.py:173-590``): constructor
``(iFile, minSup, sep=None, use_gds="auto", use_gpu_io="auto")``, ``mine()``,
``getPatterns()``, ``getPatternsAsDataFrame()``, ``save()``, the timing and
memory getters, ``printResults()``, ``getInputBackend()`` and the CLI.

``mine()`` hands the vertical store to the device search
(``cpm_eclat_mine``): candidates are generated, pruned and counted on the
device and only the pattern forest comes back.  The level-wise structure is
the reference's Apriori (``cuApriori.py:473-535``): level k+1 candidates join
frequent k-itemsets that share a k-1 prefix, a candidate whose 2-item subsets
are not all frequent is dropped before counting, and every candidate's support
is the popcount of the AND of its item rows (complete intersection,
``_bitsets.py:30-53``).  The pattern set and supports are the reference's.
"""

from __future__ import annotations

import math
import sys
import time
from itertools import combinations
from pathlib import Path
from typing import Dict, Tuple

import numpy as np

if __name__ == "__main__" and not __package__:
    sys.path.insert(0, str(Path(__file__).resolve().parents[2]))
    __package__ = "cupami_amd.frequentPattern"

from .. import _hip
from . import _inputs


def binary_pairs(offsets, ids, keep_empty=False):
    """Host ``(labels, rows, cols)`` of a BIN file's arrays
    (``cuApriori.py:381-396``); see ``_inputs.basket_array_pairs``."""
    return _inputs.basket_array_pairs(offsets, ids, keep_empty)


_BACKENDS = {
    "cpu": "Text or Parquet input",
    "numpy-hip": "Bulk numeric reads; the store is built on the device from the arrays",
    "hip-text": "Text parsed on the device into basket arrays; the store is built there",
}


class cuApriori:
    """Frequent itemsets with exact supports (Agrawal and Srikant, VLDB 1994).

    Parameters follow the reference: ``iFile`` is an iterable of baskets, a
    CSV/TSV path or http(s) URL, a ``CUPAMI01`` BIN file or a Parquet file with
    an ``items`` list<uint32> column; ``minSup`` is absolute (int, or float
    > 1) or a fraction of the transactions (float in (0, 1]).  ``use_gpu_io``
    chooses where input is parsed: ``"auto"`` reads BIN and Parquet arrays in
    bulk and parses text of decimal ids on the device (other text on the
    host), ``True`` requires that, ``False`` parses everything on the host.
    """

    def __init__(self, iFile, minSup, sep=None, use_gds="auto", use_gpu_io="auto"):
        started = time.perf_counter()
        for name, value in (("use_gds", use_gds), ("use_gpu_io", use_gpu_io)):
            if value not in ("auto", True, False):
                raise ValueError(f'{name} must be "auto", True, or False')
        self.iFile, self.minSup, self.sep = iFile, minSup, sep
        self.use_gds, self.use_gpu_io = use_gds, use_gpu_io
        # the input, in one of its forms: label rows (_Database), basket
        # arrays on the host (_binaryArrays) or on the device (_deviceBaskets)
        self._binaryArrays = None
        self._binaryUniverse = 0
        self._deviceBaskets = None
        self._transactionCount = 0
        self._setBackend("cpu")
        if use_gds is True and _inputs.source_kind(iFile) != "bin":
            raise ValueError("Required GDS input currently supports local BIN files only")
        self._Database = self._readTransactions(iFile)
        if self._Database is not None:
            self._transactionCount = len(self._Database)
        self._minSup = self._getSupportCount(minSup, self._transactionCount)
        self._finalPatterns: Dict[Tuple[str, ...], int] = {}
        self._startTime = self._endTime = 0.0
        self._loadingTime = time.perf_counter() - started

    def _setBackend(self, name, reason=None):
        self._inputBackend = {"backend": name, "reason": reason or _BACKENDS[name]}

    def getInputBackend(self):
        return dict(self._inputBackend)

    # ---- readers (cuApriori.py:202-450) ----------------------------------------
    def _readTransactions(self, source):
        """Label rows, or None when the input is kept as basket arrays."""
        kind = _inputs.source_kind(source)
        if kind == "rows":
            return self._cleanTransactions(source)
        if kind == "bin":
            return self._readBinaryFile(Path(source))
        if kind == "parquet":
            return self._readParquetFile(source)
        return self._readTextSource(source)

    def _readTextSource(self, source):
        data = _inputs.fetch_bytes(source)
        if self.use_gpu_io is not False:
            newline = data.find(b"\n")
            first = (data if newline < 0 else data[:newline]).decode("utf-8", "replace")
            sep = _inputs.separator_for(first, self.sep)
            try:
                ctx = self._getCudaModule()
            except (ImportError, RuntimeError) as error:
                if self.use_gpu_io is True:
                    raise RuntimeError(f"GPU input unavailable: {error}") from error
                self._setBackend("cpu", str(error))
                return self._readDelimitedFile(data.decode("utf-8"))
            baskets = ctx.parse_text(data, sep) if len(sep) == 1 else None
            if baskets is not None:
                self._deviceBaskets = baskets
                self._transactionCount = int(baskets.n_nonempty)
                self._setBackend("hip-text")
                return None
            if self.use_gpu_io is True:
                raise RuntimeError("GPU input unavailable: the device reader takes delimited decimal item ids "
                                   "(no quotes, ASCII only); this text needs the host reader")
            self._setBackend("cpu", "Text that the device reader does not take (labels, quotes or non-ASCII)")
        return self._readDelimitedFile(data.decode("utf-8"))

    def _readBinaryFile(self, path):
        """The generator's ``CUPAMI01`` format (``cuApriori.py:256-323``)."""
        _inputs.bin_header(path, _inputs.CUPAMI01)
        if self.use_gds is True:
            raise RuntimeError("GDS input requested but unavailable: no GPUDirect Storage "
                               "path on this platform")
        _, offsets, ids, universe = _inputs.bin_arrays(path, _inputs.CUPAMI01)
        if self.use_gpu_io is False:
            return _inputs.arrays_to_label_rows(offsets, ids)
        self._keepArrays(offsets, ids, universe)
        return None

    def _readParquetFile(self, path):
        """A Parquet file's ``items`` column (``cuApriori.py:406-426``), read as
        basket arrays (label rows with ``use_gpu_io=False``)."""
        offsets, ids, universe = _inputs.parquet_basket_arrays(path)
        if self.use_gpu_io is False:
            return _inputs.arrays_to_label_rows(offsets, ids)
        self._keepArrays(offsets, ids, universe)
        return None

    def _keepArrays(self, offsets, ids, universe):
        self._binaryArrays = (offsets, ids)
        self._binaryUniverse = int(universe)
        self._transactionCount = int(np.count_nonzero(np.diff(offsets.astype(np.int64))))
        self._setBackend("numpy-hip")

    def _readDelimitedFile(self, stream):
        """Label rows of delimited text (a str or a text stream)."""
        text = stream if isinstance(stream, str) else stream.read()
        return _inputs.delimited_label_rows(text, self.sep)

    def _cleanTransactions(self, rows):
        return _inputs.kept_label_rows(rows)

    @staticmethod
    def _cleanTransaction(transaction):
        """``str().strip()``, drop empty labels and repeats (``cuApriori.py:444-450``)."""
        return _inputs.clean_labels(transaction)

    @staticmethod
    def _getSupportCount(minSup, transactionCount):
        """``cuApriori.py:452-456``: a fraction in (0, 1] is rounded up, any
        other value is truncated."""
        fractional = isinstance(minSup, float) and 0.0 < minSup <= 1.0
        return math.ceil(minSup * transactionCount) if fractional else int(minSup)

    @staticmethod
    def _getCudaModule():
        """The device runtime (``cuApriori.py:458-471``): ImportError when the
        HIP library is missing, RuntimeError when no device is present."""
        try:
            lib = _hip.load()
        except ImportError as error:
            raise ImportError("GPU mining requires libcupami_hip.so and a HIP device.") from error
        if lib.cpm_device_count() <= 0:
            raise RuntimeError("GPU mining requires a HIP device; no HIP device was found.")
        return _hip.context(0)

    # ---- COO (cuApriori.py:378-404) -----------------------------------------------
    def _getDevicePairs(self, ctx=None):
        """Labels sorted as strings, and the (row, column) pairs of the store."""
        if self._deviceBaskets is not None:
            return binary_pairs(*self._deviceBaskets.arrays())
        if self._binaryArrays is not None:
            return binary_pairs(*self._binaryArrays)
        labels = sorted(set().union(*self._Database)) if self._Database else []
        column = {label: i for i, label in enumerate(labels)}
        sizes = np.fromiter(map(len, self._Database), dtype=np.int64, count=len(self._Database))
        rows = np.repeat(np.arange(sizes.size, dtype=np.uint64), sizes)
        cols = np.fromiter((column[label] for basket in self._Database for label in basket),
                           dtype=np.uint32, count=int(sizes.sum()))
        return labels, rows, cols

    def _getDeviceStore(self, ctx):
        """Labels and the device store (``None`` when there are no items).
        Basket arrays are built on the device as they are
        (``cpm_bits_build_csr``), label rows from their host pairs."""
        if self._deviceBaskets is not None:
            bits, ids = self._deviceBaskets.build_bits()
        elif self._binaryArrays is not None:
            bits, ids = ctx.build_bits_csr(*self._binaryArrays, self._binaryUniverse)
        else:
            labels, rows, cols = self._getDevicePairs(ctx)
            return labels, (ctx.build_bits(rows, cols, self._transactionCount, len(labels)) if labels else None)
        labels = [str(v) for v in ids.tolist()]
        if not labels:
            bits.close()
            bits = None
        return labels, bits

    # ---- reference hooks, kept for subclasses and operator-level use --------------
    def _generateCandidates(self, previousLevel):
        """Host join/prune over label tuples (``cuApriori.py:473-492``)."""
        if not previousLevel:
            return []
        size = len(previousLevel[0]) + 1
        previous = set(previousLevel)
        ordered = sorted(previousLevel)
        out = set()
        for i, left in enumerate(ordered):
            for right in ordered[i + 1:]:
                if left[:-1] != right[:-1]:
                    continue
                candidate = left + (right[-1],)
                if all(tuple(s) in previous for s in combinations(candidate, size - 1)):
                    out.add(candidate)
        return sorted(out)

    def _countCandidateSupports(self, ctx, bits, itemIndexes, candidates):
        from ._bitsets import count_supports
        return count_supports(ctx, bits, [[itemIndexes[i] for i in c] for c in candidates])

    # ---- mining ------------------------------------------------------------------
    def _deviceSearch(self, ctx, bits, items):
        return ctx.eclat_mine(bits, self._minSup, use_gram=True)

    def mine(self):
        self._startTime = time.perf_counter()
        ctx = self._getCudaModule()
        self._finalPatterns = {}
        items, bits = self._getDeviceStore(ctx)
        if bits is not None:
            try:
                patterns = self._deviceSearch(ctx, bits, items)
            finally:
                bits.close()
            self._keepDeviceResult(patterns, items)
        ctx.synchronize()
        self._endTime = time.perf_counter()

    # ---- results ---------------------------------------------------------------
    # A device search's result stays in the library's arrays: the outputs
    # render it there in bulk (cpm_patterns_keys / cpm_patterns_save), and the
    # reference's _finalPatterns dict (label tuple -> support) is built only
    # when something reads it.
    @property
    def _finalPatterns(self):
        pending = self.__dict__.get("_pendingResult")
        if pending is not None:
            self.__dict__["_pendingResult"] = None
            self.__dict__["_finalPatternStore"] = self._pendingToDict(*pending)
        return self.__dict__.setdefault("_finalPatternStore", {})

    @_finalPatterns.setter
    def _finalPatterns(self, value):
        self.__dict__["_pendingResult"] = None
        self.__dict__["_finalPatternStore"] = value

    def _keepDeviceResult(self, patterns, labels):
        self._finalPatterns = {}
        self.__dict__["_pendingResult"] = (patterns, labels)

    def _pendingToDict(self, patterns, labels):
        return patterns.to_dict(labels)

    def _pending(self):
        return self.__dict__.get("_pendingResult")

    # ---- outputs (cuApriori.py:537-590) -----------------------------------------------
    def getPatterns(self):
        pending = self._pending()
        if pending is not None:
            patterns, labels = pending
            return dict(zip(patterns.keys(labels), patterns.support.tolist()))
        return {"\t".join(p): s for p, s in self._finalPatterns.items()}

    def getPatternsAsDataFrame(self):
        import pandas as pd
        pending = self._pending()
        if pending is not None:
            patterns, labels = pending
            return pd.DataFrame({"Patterns": patterns.keys(labels),
                                 "Support": patterns.support.astype(np.int64)})
        rows = [("\t".join(p), s) for p, s in self._finalPatterns.items()]
        return pd.DataFrame(rows, columns=["Patterns", "Support"])

    def save(self, outputFile):
        pending = self._pending()
        if pending is not None:
            patterns, labels = pending
            patterns.save(outputFile, labels)
            return
        with Path(outputFile).open("w", encoding="utf-8") as handle:
            for pattern, support in sorted(self._finalPatterns.items()):
                handle.write(f"{chr(9).join(pattern)} : {support}\n")

    def getRuntime(self):
        """Mining time in seconds, excluding constructor input loading."""
        return self._endTime - self._startTime

    def getLoadingTime(self):
        return self._loadingTime

    def getMiningTime(self):
        """Wall time of the latest mine(), including device preparation."""
        return self.getRuntime()

    def getTotalRuntime(self):
        return self.getLoadingTime() + self.getMiningTime()

    @staticmethod
    def _memory(field):
        try:
            import psutil
        except ImportError:
            return 0
        process = psutil.Process()
        info = process.memory_full_info() if field == "uss" else process.memory_info()
        return int(getattr(info, field))

    def getMemoryRSS(self):
        return self._memory("rss")

    def getMemoryUSS(self):
        return self._memory("uss")

    def _patternCount(self):
        pending = self._pending()
        return len(pending[0]) if pending is not None else len(self._finalPatterns)

    def printResults(self):
        lines = (("Total number of Frequent Patterns:", self._patternCount()),
                 ("Total Memory in USS:", self.getMemoryUSS()),
                 ("Total Memory in RSS:", self.getMemoryRSS()),
                 ("Loading time (seconds):", self.getLoadingTime()),
                 ("Mining time (seconds):", self.getMiningTime()),
                 ("Total time (seconds):", self.getTotalRuntime()))
        for label, value in lines:
            print(label, value)


Apriori = cuApriori


def _cli(cls, argv):
    """``<inputFile> <outputFile> <minSup> [sep]`` (``cuApriori.py:596-609``)."""
    if len(argv) not in (4, 5):
        sys.exit(f"Usage: python {cls.__name__}.py <inputFile> <outputFile> <minSup> [sep]")
    try:
        minSup = int(argv[3])
    except ValueError:
        minSup = float(argv[3])
    sep = argv[4] if len(argv) == 5 else None
    miner = cls(argv[1], minSup, sep)
    miner.mine()
    miner.save(argv[2])
    miner.printResults()


if __name__ == "__main__":
    _cli(cuApriori, sys.argv)
