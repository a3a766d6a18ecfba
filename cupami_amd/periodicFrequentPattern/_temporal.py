"""Temporal input and the periodic evaluation operator.

``TemporalInput`` keeps the reference's temporal reader and acceptance rules
(``_temporal.py:65-181``): timestamp-first rows, digits-only non-negative int64
timestamps, rows stably sorted by timestamp, empty baskets kept (they count in
N), end = last timestamp or an explicit ``endTime`` (>= last), fractional
``maxPer`` in (0, 1] scaled by the end timestamp.  Periodicity of a pattern is
the largest gap of ``[0, ts of each occurrence in row order, end]``
(Kiran, Kitsuregawa and Reddy, JSS 2016).
"""

from __future__ import annotations

import csv
import io
import math
import numbers
import struct
import time
from pathlib import Path
from urllib.parse import urlparse
from urllib.request import urlopen

import numpy as np

from ..frequentPattern.cuApriori import binary_pairs, cuApriori

_INT64_MAX = np.iinfo(np.int64).max


def read_temporal_arrays(path):
    """The arrays of a ``CUPAMIT1`` file (``_temporal_io.py:48-72``), checked:
    ``(timestamps i64[n], offsets u64[n+1], items u32[m], universe)``."""
    header = struct.Struct("<8sQQQ")
    with Path(path).open("rb") as source:
        raw = source.read(header.size)
        if len(raw) != header.size:
            raise ValueError("Invalid temporal BIN header")
        magic, count, total, universe = header.unpack(raw)
        if magic != b"CUPAMIT1":
            raise ValueError("Expected a temporal BIN file with timestamp data")
        if not 1 <= universe <= 0xFFFFFFFF:
            raise ValueError("Invalid temporal BIN item range")
        if Path(path).stat().st_size != header.size + 8 * count + 8 * (count + 1) + 4 * total:
            raise ValueError("Invalid temporal BIN file size")
        ts = np.fromfile(source, dtype="<i8", count=count)
        offsets = np.fromfile(source, dtype="<u8", count=count + 1)
        items = np.fromfile(source, dtype="<u4", count=total)
    if len(ts) != count or len(offsets) != count + 1 or len(items) != total:
        raise ValueError("Truncated temporal BIN arrays")
    if np.any(ts < 0):
        raise ValueError("Temporal timestamps must be nonnegative")
    if offsets[0] != 0 or offsets[-1] != total or np.any(offsets[1:] < offsets[:-1]):
        raise ValueError("Invalid temporal BIN offsets")
    if np.any(items > universe):
        raise ValueError("Temporal BIN item ID exceeds header range")
    return ts, offsets, items, int(universe)


def read_temporal_binary(path):
    """Rows ``[ts, *items]`` of a ``CUPAMIT1`` file (``_temporal_io.py:48-72``)."""
    ts, offsets, items, _ = read_temporal_arrays(path)
    return [[int(t), *items[int(offsets[i]):int(offsets[i + 1])].tolist()]
            for i, t in enumerate(ts.tolist())]


def read_temporal_parquet(path):
    """Rows of a temporal Parquet file (``timestamp`` int64, ``items`` list<uint32>)."""
    try:
        import pyarrow.parquet as pq
    except ImportError as error:
        raise ImportError("Temporal Parquet requires pyarrow: pip install pyarrow") from error
    table = pq.read_table(path)
    if "timestamp" not in table.column_names or "items" not in table.column_names:
        raise ValueError("Temporal Parquet needs timestamp and items columns")
    ts = table.column("timestamp").to_pylist()
    items = table.column("items").to_pylist()
    return [[t, *(row or [])] for t, row in zip(ts, items)]


class TemporalInput(cuApriori):
    """Timestamp-first CSV, temporal BIN/Parquet, or iterable input."""

    def __init__(self, iFile, minSup, maxPer, sep=None, *, endTime=None):
        loading_start = time.perf_counter()
        self._timestamps = []
        self.endTime = endTime
        self.maxPer = maxPer
        super().__init__(iFile, minSup, sep, use_gds=False, use_gpu_io=False)
        if (isinstance(minSup, bool) or not isinstance(minSup, numbers.Real)
                or not math.isfinite(minSup) or minSup <= 0):
            raise ValueError("minSup must be a positive finite number")
        if (isinstance(maxPer, bool) or not isinstance(maxPer, numbers.Real)
                or not math.isfinite(maxPer) or maxPer < 0):
            raise ValueError("maxPer must be a nonnegative finite number")
        self._maxPer = (maxPer * self._endTimestamp
                        if isinstance(maxPer, float) and 0 < maxPer <= 1 else maxPer)
        self._periods = {}
        self._loadingTime = time.perf_counter() - loading_start

    @staticmethod
    def _timestamp(value):
        if isinstance(value, (bool, float)):
            raise ValueError("timestamps must be nonnegative int64 integers")
        text = str(value).strip()
        if not text or any(c not in "0123456789" for c in text):
            raise ValueError("timestamps must be nonnegative int64 integers")
        stamp = int(text)
        if stamp > _INT64_MAX:
            raise ValueError("timestamp exceeds int64")
        return stamp

    def _readTransactions(self, source):
        if isinstance(source, (str, Path)):
            suffix = Path(str(source)).suffix.lower()
            is_url = urlparse(str(source)).scheme in ("http", "https")
            if suffix in (".bin", ".parquet"):
                if is_url:
                    raise ValueError("Temporal BIN and Parquet input requires a local file")
                if suffix == ".bin":
                    return self._readTemporalArrays(source)
                return self._cleanTemporalRows(read_temporal_parquet(source))
            if is_url:
                with urlopen(str(source), timeout=30) as response:
                    text = response.read().decode("utf-8-sig")
            else:
                text = Path(source).read_text(encoding="utf-8-sig")
            separator = self.sep
            if separator is None:
                first = next((line for line in text.splitlines() if line.strip()), "")
                separator = "\t" if "\t" in first else "," if "," in first else None
            rows = (csv.reader(io.StringIO(text), delimiter=separator) if separator
                    else (line.split() for line in text.splitlines()))
        else:
            rows = source
        return self._cleanTemporalRows(rows)

    def _cleanTemporalRows(self, rows):
        records = []
        for row in rows:
            row = list(row)
            if not row or not any(str(value).strip() for value in row):
                continue
            records.append((self._timestamp(row[0]), self._cleanTransaction(row[1:])))
        records.sort(key=lambda record: record[0])
        self._timestamps = [record[0] for record in records]
        last = self._timestamps[-1] if records else 0
        self._endTimestamp = last if self.endTime is None else self._timestamp(self.endTime)
        if self._endTimestamp < last:
            raise ValueError("endTime must be at least the last timestamp")
        self._inputBackend = {"backend": "cpu", "reason": "Temporal reader; mining uses HIP"}
        return [record[1] for record in records]

    def _readTemporalArrays(self, source):
        """A ``CUPAMIT1`` file's arrays, kept for the device build: every row
        stays (timestamp-only rows too) and rows are ordered by timestamp,
        stably, as ``_cleanTemporalRows`` orders them (``_temporal.py:125-139``)."""
        ts, offsets, ids, universe = read_temporal_arrays(source)
        if ts.size > 1 and np.any(ts[1:] < ts[:-1]):
            order = np.argsort(ts, kind="stable")
            lengths = np.diff(offsets.astype(np.int64))[order]
            starts = offsets[:-1].astype(np.int64)[order]
            sorted_off = np.zeros(ts.size + 1, dtype=np.uint64)
            sorted_off[1:] = np.cumsum(lengths)
            base = np.repeat(starts - sorted_off[:-1].astype(np.int64), lengths)
            ids = ids[base + np.arange(ids.size, dtype=np.int64)]
            offsets, ts = sorted_off, ts[order]
        self._timestamps = ts.astype(np.int64)
        last = int(ts[-1]) if ts.size else 0
        self._endTimestamp = last if self.endTime is None else self._timestamp(self.endTime)
        if self._endTimestamp < last:
            raise ValueError("endTime must be at least the last timestamp")
        self._transactionCount = int(ts.size)
        self._binaryArrays = (offsets.astype(np.uint64), ids.astype(np.uint32))
        self._binaryUniverse = universe
        self._inputBackend = {"backend": "numpy-hip",
                              "reason": "Temporal BIN arrays; the store is built on the device"}
        return None

    def _getDevicePairs(self, ctx=None):
        if self._binaryArrays is not None:   # every row kept, timestamp-only rows too
            return binary_pairs(*self._binaryArrays, keep_empty=True)
        return super()._getDevicePairs(ctx)

    def _maxPerFloor(self):
        """Integer bound equivalent to ``period <= maxPer`` for integer periods."""
        value = math.floor(self._maxPer)
        return max(-1, min(int(value), int(_INT64_MAX)))

    def _prepare(self, ctx):
        if self._binaryArrays is not None:
            # the store straight from the file's arrays, every row kept
            offsets, ids = self._binaryArrays
            bits, label_ids = ctx.build_bits_csr(offsets, ids, self._binaryUniverse, keep_empty=True)
            labels, rows, columns = [str(v) for v in label_ids.tolist()], None, None
            if not labels:
                bits.close()
                bits = None
        else:
            labels, rows, columns = self._getDevicePairs(ctx)
            bits = (ctx.build_bits(rows, columns, self._transactionCount, len(labels))
                    if labels else None)
        self._labels = labels
        self._itemIndexes = {label: i for i, label in enumerate(labels)}
        self._temporalBits = bits
        self._hostTimestamps = np.asarray(self._timestamps, dtype=np.int64)
        self._finalPatterns = {}
        self._periods = {}
        return labels, rows, columns

    def _evaluate(self, ctx, candidates):
        """``[(support, period)]`` per candidate (``_temporal.py:151-161``)."""
        if not candidates:
            return []
        cand = np.asarray([[self._itemIndexes[i] for i in p] for p in candidates],
                          dtype=np.uint32)
        sup, per = ctx.periodic_eval(self._temporalBits, self._hostTimestamps,
                                     self._endTimestamp, cand)
        return list(zip(sup.tolist(), per.tolist()))

    def _accept(self, pattern, support, period):
        if support >= self._minSup and support > 0 and period <= self._maxPer:
            self._finalPatterns[pattern] = int(support)
            self._periods[pattern] = int(period)
            return True
        return False

    def _pendingToDict(self, patterns, labels):
        # support in _finalPatterns, period in _periods, as _accept fills them
        sets = [tuple(labels[c] for c in s) for s in patterns.itemsets()]
        self._periods = dict(zip(sets, patterns.period.tolist()))
        return dict(zip(sets, patterns.support.tolist()))

    def getPatterns(self):
        pending = self._pending()
        if pending is not None:
            patterns, labels = pending
            return dict(zip(patterns.keys(labels),
                            zip(patterns.support.tolist(), patterns.period.tolist())))
        return {"\t".join(p): (s, self._periods[p]) for p, s in self._finalPatterns.items()}

    def getPatternsAsDataFrame(self):
        import pandas as pd
        pending = self._pending()
        if pending is not None:
            patterns, labels = pending
            return pd.DataFrame({"Patterns": patterns.keys(labels),
                                 "Support": patterns.support.astype(np.int64),
                                 "Periodicity": patterns.period.astype(np.int64)})
        return pd.DataFrame([("\t".join(p), s, self._periods[p])
                             for p, s in self._finalPatterns.items()],
                            columns=["Patterns", "Support", "Periodicity"])

    def save(self, outputFile):
        pending = self._pending()
        if pending is not None:
            patterns, labels = pending
            patterns.save(outputFile, labels)
            return
        with Path(outputFile).open("w", encoding="utf-8") as handle:
            for pattern, support in sorted(self._finalPatterns.items()):
                handle.write("{} : {} : {}\n".format("\t".join(pattern), support,
                                                     self._periods[pattern]))
