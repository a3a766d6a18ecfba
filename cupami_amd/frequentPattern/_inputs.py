"""Input sources of the miners.

The reference's readers (``cuApriori.py:202-450`` for baskets,
``_temporal.py:73-139`` and ``extras/_temporal_io.py:48-120`` for
timestamped rows) accept an iterable of rows, a delimited text file or
http(s) URL, a ``CUPAMI01`` / ``CUPAMIT1`` BIN file, or a Parquet file with
an ``items`` list<uint32> column.  Here every source ends in one of two
forms:

* label rows -- tuples of cleaned labels, one per basket (text that only
  the host reads, and Python rows);
* basket arrays -- ids as stored (``offsets`` u64[n+1], ``ids`` u32[m]) with
  the id universe, which go to the device store build untouched (BIN and
  Parquet files, and text of decimal ids parsed on the device).

Both forms give the same store: labels are ``str(value).strip()``, empty
ones dropped, each label once per basket, columns in string order of the
labels (``"10" < "2"``).
"""

from __future__ import annotations

import csv
import io
import struct
from itertools import chain
from pathlib import Path
from urllib.parse import urlparse
from urllib.request import urlopen

import numpy as np

_HEADER = struct.Struct("<8sQQQ")   # magic, baskets, ids, universe (little endian)


def clean_labels(values):
    """A basket's labels (``cuApriori.py:444-450``): ``str(v).strip()`` of each
    value, empty ones dropped, repeats dropped (first one kept)."""
    return tuple(dict.fromkeys(label for label in (str(v).strip() for v in values) if label))


def kept_label_rows(rows):
    """Cleaned rows, the empty ones left out (transactional readers do)."""
    return [basket for basket in map(clean_labels, rows) if basket]


def is_url(source):
    return isinstance(source, str) and urlparse(source).scheme in ("http", "https")


def source_kind(source):
    """``'url'``, ``'bin'``, ``'parquet'``, ``'text'`` or ``'rows'``."""
    if is_url(source):
        return "url"
    if isinstance(source, (str, Path)):
        return {".bin": "bin", ".parquet": "parquet"}.get(Path(source).suffix.lower(), "text")
    return "rows"


def fetch_bytes(source):
    """The raw bytes of a local file or an http(s) URL."""
    if is_url(source):
        with urlopen(source, timeout=30) as response:
            return response.read()
    return Path(source).read_bytes()


def separator_for(first_line, sep):
    """The reference's sniffing (``cuApriori.py:428-432``): the given ``sep``,
    else a tab when the first line holds one, else a comma."""
    if sep is not None:
        return sep
    return "\t" if "\t" in first_line else ","


def delimited_label_rows(text, sep):
    """Label rows of delimited text (a str), read with the csv module."""
    stream = io.StringIO(text, newline="")
    first = stream.readline()
    reader = csv.reader(chain((first,), stream), delimiter=separator_for(first, sep))
    return kept_label_rows(reader)


# ---- BIN files ---------------------------------------------------------------------

# (magic, has timestamps, message per check) of the two BIN formats
CUPAMI01 = (b"CUPAMI01", False, {
    "header": "Invalid BIN file: truncated header",
    "magic": "Invalid BIN file: expected CUPAMI01 format",
    "universe": "Invalid BIN file: item universe must fit in uint32",
    "size": "Invalid BIN file: size does not match header",
    "truncated": "Invalid BIN file: truncated arrays",
    "offsets": "Invalid BIN file: invalid transaction offsets",
    "ids": "Invalid BIN file: item ID outside configured universe",
})
CUPAMIT1 = (b"CUPAMIT1", True, {
    "header": "Invalid temporal BIN header",
    "magic": "Expected a temporal BIN file with timestamp data",
    "universe": "Invalid temporal BIN item range",
    "size": "Invalid temporal BIN file size",
    "truncated": "Truncated temporal BIN arrays",
    "timestamps": "Temporal timestamps must be nonnegative",
    "offsets": "Invalid temporal BIN offsets",
    "ids": "Temporal BIN item ID exceeds header range",
})


def bin_header(path, fmt):
    """Checked header ``(baskets, ids, universe)`` of a BIN file of format
    ``fmt`` (the checks that need no array)."""
    magic, timed, msg = fmt
    path = Path(path)
    with path.open("rb") as source:
        raw = source.read(_HEADER.size)
    if len(raw) != _HEADER.size:
        raise ValueError(msg["header"])
    found, count, total, universe = _HEADER.unpack(raw)
    if found != magic:
        raise ValueError(msg["magic"])
    if not 1 <= universe <= 0xFFFFFFFF:
        raise ValueError(msg["universe"])
    expected = _HEADER.size + (8 * count if timed else 0) + 8 * (count + 1) + 4 * total
    if path.stat().st_size != expected:
        raise ValueError(msg["size"])
    return count, total, universe


def bin_arrays(path, fmt):
    """Checked arrays of a BIN file: ``(timestamps or None, offsets, ids,
    universe)``; checks in the reference's order (header, arrays, timestamps,
    offsets, ids)."""
    magic, timed, msg = fmt
    count, total, universe = bin_header(path, fmt)
    with Path(path).open("rb") as source:
        source.seek(_HEADER.size)
        stamps = np.fromfile(source, dtype="<i8", count=count) if timed else None
        offsets = np.fromfile(source, dtype="<u8", count=count + 1)
        ids = np.fromfile(source, dtype="<u4", count=total)
    if offsets.size != count + 1 or ids.size != total or (timed and stamps.size != count):
        raise ValueError(msg["truncated"])
    if timed and stamps.size and stamps.min() < 0:
        raise ValueError(msg["timestamps"])
    rising = offsets.size < 2 or bool((offsets[1:] >= offsets[:-1]).all())
    if offsets[0] != 0 or offsets[-1] != total or not rising:
        raise ValueError(msg["offsets"])
    if ids.size and int(ids.max()) > universe:
        raise ValueError(msg["ids"])
    return stamps, offsets.astype(np.uint64), ids.astype(np.uint32), int(universe)


def arrays_to_label_rows(offsets, ids, keep_empty=False):
    """Label rows of basket arrays (the host reading of a BIN file)."""
    bounds = offsets.tolist()
    rows = (clean_labels(ids[a:b].tolist()) for a, b in zip(bounds[:-1], bounds[1:]))
    return list(rows) if keep_empty else [r for r in rows if r]


# ---- Parquet -------------------------------------------------------------------------

def _arrow(message):
    try:
        import pyarrow as pa
        import pyarrow.parquet as pq
    except ImportError as error:
        raise ImportError(message) from error
    return pa, pq


def _list_arrays(column):
    """``(offsets u64, ids u32)`` of a list<uint32> column, no Python loop."""
    lists = column.combine_chunks() if hasattr(column, "combine_chunks") else column
    raw = np.asarray(lists.offsets, dtype=np.int64)
    ids = np.asarray(lists.flatten(), dtype=np.uint32)
    return (raw - raw[0]).astype(np.uint64), ids


def parquet_basket_arrays(path):
    """Basket arrays of a Parquet file's ``items`` column, with the
    reference's checks (``cuApriori.py:406-426``)."""
    pa, pq = _arrow("Parquet input requires pyarrow: pip install pyarrow")
    schema = pq.read_schema(path)
    if "items" not in schema.names:
        raise ValueError("Invalid Parquet file: missing items column")
    kind = schema.field("items").type
    if not (pa.types.is_list(kind) and pa.types.is_uint32(kind.value_type)):
        raise ValueError("Invalid Parquet file: items must be a list of uint32 IDs")
    column = pq.read_table(path, columns=["items"]).column("items").combine_chunks()
    if column.null_count or column.flatten().null_count:
        raise ValueError("Invalid Parquet file: null transactions or item IDs")
    offsets, ids = _list_arrays(column)
    universe = max(1, int(ids.max()) if ids.size else 1)
    return offsets, ids, universe


def temporal_parquet_arrays(path):
    """``(timestamps i64, offsets, ids, universe)`` of a temporal Parquet
    file (``extras/_temporal_io.py:87-120``)."""
    pa, pq = _arrow("Temporal Parquet requires pyarrow: pip install pyarrow")
    table = pq.read_table(path)
    if not {"timestamp", "items"} <= set(table.column_names):
        raise ValueError("Temporal Parquet requires timestamp and items columns")
    stamps, items = table.column("timestamp").combine_chunks(), table.column("items").combine_chunks()
    if stamps.type != pa.int64() or not pa.types.is_list(items.type) or items.type.value_type != pa.uint32():
        raise ValueError("Temporal Parquet requires timestamp: int64 and items: list<uint32>")
    if stamps.null_count or items.null_count or items.flatten().null_count:
        raise ValueError("Temporal Parquet cannot contain null timestamps, baskets, or items")
    ts = np.asarray(stamps, dtype=np.int64)
    if ts.size and ts.min() < 0:
        raise ValueError("Temporal timestamps must be nonnegative")
    offsets, ids = _list_arrays(items)
    return ts, offsets, ids, max(1, int(ids.max()) if ids.size else 1)


# ---- pairs of basket arrays -------------------------------------------------------------

def label_order(ids):
    """Distinct ids of ``ids`` as labels in string order, and the column of
    each distinct id (``np.unique`` order)."""
    distinct = np.unique(ids)
    text = [str(v) for v in distinct.tolist()]
    order = sorted(range(len(text)), key=text.__getitem__)
    column = np.empty(len(text), dtype=np.uint32)
    column[order] = np.arange(len(text), dtype=np.uint32)
    return [text[i] for i in order], distinct, column


def basket_array_pairs(offsets, ids, keep_empty=False):
    """Host ``(labels, rows, cols)`` of basket arrays: rows number the
    non-empty baskets (every basket with ``keep_empty``), columns follow the
    labels' string order.  The device builds the same store from the arrays
    themselves (``cpm_bits_build_csr``); this serves ``_getDevicePairs``."""
    labels, distinct, column = label_order(ids)
    sizes = np.diff(np.asarray(offsets, dtype=np.int64))
    row_of_basket = np.arange(sizes.size) if keep_empty else np.cumsum(sizes > 0) - 1
    rows = np.repeat(row_of_basket.astype(np.uint64), sizes)
    cols = column[np.searchsorted(distinct, ids)] if ids.size else np.zeros(0, np.uint32)
    return labels, rows, cols.astype(np.uint32)


def stable_by_timestamp(ts, offsets, ids):
    """Rows reordered by timestamp, stably (``_temporal.py:134``), as arrays."""
    if ts.size < 2 or bool((ts[1:] >= ts[:-1]).all()):
        return ts, offsets, ids
    order = np.argsort(ts, kind="stable")
    sizes = np.diff(offsets.astype(np.int64))[order]
    starts = offsets[:-1].astype(np.int64)[order]
    new_off = np.zeros(ts.size + 1, dtype=np.uint64)
    new_off[1:] = np.cumsum(sizes)
    gather = np.repeat(starts - new_off[:-1].astype(np.int64), sizes) + np.arange(ids.size, dtype=np.int64)
    return ts[order], new_off, ids[gather]
