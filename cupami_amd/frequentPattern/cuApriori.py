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

import csv
import io
import math
import struct
import sys
import time
from itertools import chain, combinations
from pathlib import Path
from typing import Dict, Tuple
from urllib.parse import urlparse
from urllib.request import urlopen

import numpy as np

if __name__ == "__main__" and not __package__:
    sys.path.insert(0, str(Path(__file__).resolve().parents[2]))
    __package__ = "cupami_amd.frequentPattern"

from .. import _hip


def binary_pairs(offsets, ids, keep_empty=False):
    """Host (labels, rows, cols) of a BIN file's arrays (``cuApriori.py:381-396``):
    distinct ids as string-sorted labels, each entry's row among the non-empty
    baskets (every basket with ``keep_empty``) and its column.  The miners
    build the store on the device instead (``cpm_bits_build_csr``); this
    serves ``_getDevicePairs`` callers."""
    unique = np.unique(ids)
    items = sorted(str(int(v)) for v in unique)
    index = {label: i for i, label in enumerate(items)}
    remap = np.array([index[str(int(v))] for v in unique], dtype=np.uint32)
    lengths = np.diff(offsets).astype(np.int64)
    active = np.arange(lengths.size) if keep_empty else np.cumsum(lengths > 0) - 1
    rows = np.repeat(active.astype(np.uint64), lengths)
    cols = remap[np.searchsorted(unique, ids)] if ids.size else np.zeros(0, np.uint32)
    return items, rows, cols.astype(np.uint32)


class cuApriori:
    """Frequent itemsets with exact supports (Agrawal and Srikant, VLDB 1994).

    Parameters follow the reference: ``iFile`` is an iterable of baskets, a
    CSV/TSV path or http(s) URL, a ``CUPAMI01`` BIN file or a Parquet file with
    an ``items`` list<uint32> column; ``minSup`` is absolute (int, or float
    > 1) or a fraction of the transactions (float in (0, 1]).
    """

    def __init__(self, iFile, minSup, sep=None, use_gds="auto", use_gpu_io="auto"):
        loading_start = time.perf_counter()
        if use_gds not in ("auto", True, False):
            raise ValueError('use_gds must be "auto", True, or False')
        if use_gpu_io not in ("auto", True, False):
            raise ValueError('use_gpu_io must be "auto", True, or False')
        self.iFile = iFile
        self.minSup = minSup
        self.sep = sep
        self.use_gds = use_gds
        self.use_gpu_io = use_gpu_io
        self._binaryArrays = None
        self._binaryUniverse = 0
        self._transactionCount = 0
        self._inputBackend = {"backend": "cpu", "reason": "Text or Parquet input"}
        if use_gds is True and (not isinstance(iFile, (str, Path))
                                or Path(iFile).suffix.lower() != ".bin"
                                or urlparse(str(iFile)).scheme in ("http", "https")):
            raise ValueError("Required GDS input currently supports local BIN files only")
        self._Database = self._readTransactions(iFile)
        if self._Database is not None:
            self._transactionCount = len(self._Database)
        self._minSup = self._getSupportCount(minSup, self._transactionCount)
        self._finalPatterns: Dict[Tuple[str, ...], int] = {}
        self._startTime = 0.0
        self._endTime = 0.0
        self._loadingTime = time.perf_counter() - loading_start

    # ---- readers (cuApriori.py:202-450) ----------------------------------------
    def _readTransactions(self, source):
        if isinstance(source, str) and urlparse(source).scheme in ("http", "https"):
            with urlopen(source, timeout=30) as response:
                stream = io.TextIOWrapper(response, encoding="utf-8", newline="")
                return self._readDelimitedFile(stream)
        if isinstance(source, (str, Path)):
            suffix = Path(source).suffix.lower()
            if suffix == ".bin":
                return self._readBinaryFile(Path(source))
            if suffix == ".parquet":
                self._requireCpuText()
                return self._readParquetFile(source)
            self._requireCpuText()
            with Path(source).open("r", newline="", encoding="utf-8") as handle:
                return self._readDelimitedFile(handle)
        return self._cleanTransactions(source)

    def _requireCpuText(self):
        # Device-side text/Parquet parsing is not part of this build; the
        # reference raises the same way when GPU input is required but
        # unavailable (cuApriori.py:243-245).
        if self.use_gpu_io is True:
            raise RuntimeError("GPU input unavailable: text and Parquet inputs are parsed on "
                               "the host")
        self._inputBackend = {"backend": "cpu", "reason": "Text or Parquet input"}

    def _readBinaryFile(self, path):
        """The generator's ``CUPAMI01`` format (``cuApriori.py:256-323``)."""
        with path.open("rb") as source:
            header = source.read(32)
            if len(header) != 32:
                raise ValueError("Invalid BIN file: truncated header")
            magic, count, total, universe = struct.unpack("<8sQQQ", header)
            if magic != b"CUPAMI01":
                raise ValueError("Invalid BIN file: expected CUPAMI01 format")
            if not 1 <= universe <= 0xFFFFFFFF:
                raise ValueError("Invalid BIN file: item universe must fit in uint32")
            if path.stat().st_size != 32 + 8 * (count + 1) + 4 * total:
                raise ValueError("Invalid BIN file: size does not match header")
            if self.use_gds is True:
                raise RuntimeError("GDS input requested but unavailable: no GPUDirect Storage "
                                   "path on this platform")
            offsets = np.fromfile(source, dtype="<u8", count=count + 1)
            ids = np.fromfile(source, dtype="<u4", count=total)
        if offsets.size != count + 1 or ids.size != total:
            raise ValueError("Invalid BIN file: truncated arrays")
        if offsets[0] != 0 or offsets[-1] != total or np.any(offsets[1:] < offsets[:-1]):
            raise ValueError("Invalid BIN file: invalid transaction offsets")
        if np.any(ids > universe):
            raise ValueError("Invalid BIN file: item ID outside configured universe")
        if self.use_gpu_io is False:
            transactions = []
            for start, end in zip(offsets[:-1].tolist(), offsets[1:].tolist()):
                basket = self._cleanTransaction(ids[start:end].tolist())
                if basket:
                    transactions.append(basket)
            return transactions
        # Bulk numeric path: typed arrays go straight to the COO builder.
        self._transactionCount = int(np.count_nonzero(offsets[1:] > offsets[:-1]))
        self._binaryArrays = (offsets.astype(np.uint64), ids.astype(np.uint32))
        self._binaryUniverse = int(universe)
        self._inputBackend = {"backend": "numpy-hip",
                              "reason": "Bulk numeric BIN reads; the store is built on the device "
                                        "from the file's arrays"}
        return None

    def getInputBackend(self):
        return dict(self._inputBackend)

    def _readParquetFile(self, path):
        try:
            import pyarrow as pa
            import pyarrow.parquet as pq
        except ImportError as error:
            raise ImportError("Parquet input requires pyarrow: pip install pyarrow") from error
        parquet = pq.ParquetFile(path)
        schema = parquet.schema_arrow
        if "items" not in schema.names:
            raise ValueError("Invalid Parquet file: missing items column")
        item_type = schema.field("items").type
        if not pa.types.is_list(item_type) or not pa.types.is_uint32(item_type.value_type):
            raise ValueError("Invalid Parquet file: items must be a list of uint32 IDs")
        transactions = []
        for batch in parquet.iter_batches(columns=["items"]):
            rows = batch.column(0).to_pylist()
            if any(row is None or any(item is None for item in row) for row in rows):
                raise ValueError("Invalid Parquet file: null transactions or item IDs")
            transactions.extend(self._cleanTransactions(rows))
        return transactions

    def _readDelimitedFile(self, stream):
        first = stream.readline()
        delimiter = self.sep
        if delimiter is None:
            delimiter = "\t" if "\t" in first else ","
        return self._cleanTransactions(csv.reader(chain((first,), stream), delimiter=delimiter))

    def _cleanTransactions(self, rows):
        out = []
        for row in rows:
            basket = self._cleanTransaction(row)
            if basket:
                out.append(basket)
        return out

    @staticmethod
    def _cleanTransaction(transaction):
        """``str().strip()``, drop empty labels and repeats (``cuApriori.py:444-450``)."""
        items = []
        for value in transaction:
            item = str(value).strip()
            if item and item not in items:
                items.append(item)
        return tuple(items)

    @staticmethod
    def _getSupportCount(minSup, transactionCount):
        """``cuApriori.py:452-456``: fractions round up, other values truncate."""
        if isinstance(minSup, float) and 0 < minSup <= 1:
            return math.ceil(minSup * transactionCount)
        return int(minSup)

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
        if self._binaryArrays is not None:
            return binary_pairs(*self._binaryArrays)
        items = sorted({item for row in self._Database for item in row})
        index = {item: i for i, item in enumerate(items)}
        lengths = [len(row) for row in self._Database]
        rows = np.repeat(np.arange(len(self._Database), dtype=np.uint64),
                         np.asarray(lengths, dtype=np.int64))
        cols = np.fromiter((index[item] for row in self._Database for item in row),
                           dtype=np.uint32, count=int(sum(lengths)))
        return items, rows, cols

    def _getDeviceStore(self, ctx):
        """Labels and the device store (``None`` when there are no items).

        A BIN file's arrays go to the device as stored and the store is
        built there (``cpm_bits_build_csr``): the distinct ids, their string
        order, and each entry's row and column never exist on the host.
        Other inputs build it from the host pairs of ``_getDevicePairs``."""
        if self._binaryArrays is not None:
            offsets, ids = self._binaryArrays
            bits, labelIds = ctx.build_bits_csr(offsets, ids, self._binaryUniverse)
            items = [str(v) for v in labelIds.tolist()]
            if not items:
                bits.close()
                return items, None
            return items, bits
        items, rows, columns = self._getDevicePairs(ctx)
        if not items:
            return items, None
        return items, ctx.build_bits(rows, columns, self._transactionCount, len(items))

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

    def getMemoryRSS(self):
        try:
            import psutil
        except ImportError:
            return 0
        return int(psutil.Process().memory_info().rss)

    def getMemoryUSS(self):
        try:
            import psutil
        except ImportError:
            return 0
        return int(psutil.Process().memory_full_info().uss)

    def printResults(self):
        pending = self._pending()
        print("Total number of Frequent Patterns:",
              len(pending[0]) if pending is not None else len(self._finalPatterns))
        print("Total Memory in USS:", self.getMemoryUSS())
        print("Total Memory in RSS:", self.getMemoryRSS())
        print("Loading time (seconds):", self.getLoadingTime())
        print("Mining time (seconds):", self.getMiningTime())
        print("Total time (seconds):", self.getTotalRuntime())


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
