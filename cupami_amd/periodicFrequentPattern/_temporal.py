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
import time
from pathlib import Path

import numpy as np

from ..frequentPattern import _inputs
from ..frequentPattern.cuApriori import binary_pairs, cuApriori

_INT64_MAX = np.iinfo(np.int64).max
_DIGITS = frozenset("0123456789")


def read_temporal_arrays(path):
    """The arrays of a ``CUPAMIT1`` file (``extras/_temporal_io.py:48-72``),
    checked: ``(timestamps i64[n], offsets u64[n+1], items u32[m], universe)``."""
    return _inputs.bin_arrays(path, _inputs.CUPAMIT1)


def read_temporal_binary(path):
    """Rows ``[ts, *items]`` of a ``CUPAMIT1`` file (``extras/_temporal_io.py:48-72``)."""
    ts, offsets, items, _ = read_temporal_arrays(path)
    bounds = offsets.tolist()
    return [[t, *items[bounds[i]:bounds[i + 1]].tolist()] for i, t in enumerate(ts.tolist())]


def _positive_real(value):
    return (not isinstance(value, bool) and isinstance(value, numbers.Real) and math.isfinite(value))


class TemporalInput(cuApriori):
    """Timestamp-first CSV, temporal BIN/Parquet, or iterable input.

    Rows keep their timestamps, equal ones included, and are ordered by
    timestamp (stably); a row with a timestamp and no item is a basket of its
    own and counts in N; the timeline starts at 0 and ends at the last
    timestamp or at ``endTime`` (``_temporal.py:65-139``)."""

    def __init__(self, iFile, minSup, maxPer, sep=None, *, endTime=None):
        started = time.perf_counter()
        self._timestamps = []
        self.endTime, self.maxPer = endTime, maxPer
        super().__init__(iFile, minSup, sep, use_gds=False, use_gpu_io=False)
        if not _positive_real(minSup) or minSup <= 0:
            raise ValueError("minSup must be a positive finite number")
        if not _positive_real(maxPer) or maxPer < 0:
            raise ValueError("maxPer must be a nonnegative finite number")
        # a fraction in (0, 1] is a share of the timeline, not of the rows
        fraction = isinstance(maxPer, float) and 0.0 < maxPer <= 1.0
        self._maxPer = maxPer * self._endTimestamp if fraction else maxPer
        self._periods = {}
        self._loadingTime = time.perf_counter() - started

    @staticmethod
    def _timestamp(value):
        """A digits-only non-negative int64 (``_temporal.py:89-98``)."""
        if isinstance(value, (bool, float)):
            raise ValueError("timestamps must be nonnegative int64 integers")
        text = str(value).strip()
        if not text or not _DIGITS.issuperset(text):
            raise ValueError("timestamps must be nonnegative int64 integers")
        stamp = int(text)
        if stamp > _INT64_MAX:
            raise ValueError("timestamp exceeds int64")
        return stamp

    def _readTransactions(self, source):
        kind = _inputs.source_kind(source)
        if kind == "rows":
            return self._cleanTemporalRows(source)
        if isinstance(source, (str, Path)) and Path(str(source)).suffix.lower() in (".bin", ".parquet"):
            if kind == "url":
                raise ValueError("Temporal BIN and Parquet input requires a local file")
            arrays = (read_temporal_arrays(source) if kind == "bin"
                      else _inputs.temporal_parquet_arrays(source))
            return self._keepTemporalArrays(*arrays)
        # a local file with universal newlines, a URL's body as sent
        text = (_inputs.fetch_bytes(source).decode("utf-8-sig") if kind == "url"
                else Path(source).read_text(encoding="utf-8-sig"))
        return self._cleanTemporalRows(self._delimitedRows(text))

    def _delimitedRows(self, text):
        """Rows of temporal text: the separator given, else a tab or a comma
        found on the first non-blank line, else runs of whitespace."""
        sep = self.sep
        if sep is None:
            first = next(filter(str.strip, text.splitlines()), "")
            sep = next((c for c in ("\t", ",") if c in first), None)
        if sep is None:
            return [line.split() for line in text.splitlines()]
        return csv.reader(io.StringIO(text), delimiter=sep)

    def _setTimeline(self, last):
        end = last if self.endTime is None else self._timestamp(self.endTime)
        if end < last:
            raise ValueError("endTime must be at least the last timestamp")
        self._endTimestamp = end

    def _cleanTemporalRows(self, rows):
        """``(timestamp, labels)`` of every row that is not blank, ordered by
        timestamp (stably); the label rows are returned, the timestamps kept."""
        records = []
        for row in map(list, rows):
            if any(str(v).strip() for v in row):
                records.append((self._timestamp(row[0]), _inputs.clean_labels(row[1:])))
        records.sort(key=lambda rec: rec[0])
        self._timestamps = [stamp for stamp, _ in records]
        self._setTimeline(self._timestamps[-1] if records else 0)
        self._setBackend("cpu", "Temporal reader; mining uses HIP")
        return [labels for _, labels in records]

    def _keepTemporalArrays(self, ts, offsets, ids, universe):
        """Temporal basket arrays kept for the device build: every row stays
        (timestamp-only rows too), rows ordered by timestamp (stably)."""
        ts, offsets, ids = _inputs.stable_by_timestamp(ts, offsets, ids)
        self._timestamps = ts.astype(np.int64)
        self._setTimeline(int(ts[-1]) if ts.size else 0)
        self._transactionCount = int(ts.size)
        self._binaryArrays = (offsets.astype(np.uint64), ids.astype(np.uint32))
        self._binaryUniverse = int(universe)
        self._setBackend("numpy-hip", "Temporal arrays; the store is built on the device")
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
