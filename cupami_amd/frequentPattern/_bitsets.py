"""Operator-level mirror of the reference's bitset module (``_bitsets.py:91-125``).

The reference's functions take the array module as their first argument; these
take the device context (``cupami_amd._hip.Context``) in that place, and
host arrays for rows.  Every call runs on the device through the C ABI.
"""

from __future__ import annotations

import numpy as np


def build_bitsets(ctx, rows, columns, transactions, items):
    """Vertical store of ``items`` rows over ``transactions`` bits
    (``_bitsets.py:91-100``); returns a device ``Bits`` handle."""
    return ctx.build_bits(rows, columns, transactions, items)


def count_supports(ctx, bits, candidates):
    """Complete-intersection support of each candidate (``_bitsets.py:103-113``)."""
    if not len(candidates):
        return []
    return ctx.count_supports(bits, np.asarray(candidates, dtype=np.uint32)).tolist()


def intersect_extensions(ctx, prefix, extensions):
    """``(prefix & extensions[k], supports)`` for every k (``_bitsets.py:116-125``).
    Rows are uint64 words (64 transactions per word)."""
    extensions = np.asarray(extensions, dtype=np.uint64)
    if not len(extensions):
        return np.zeros_like(extensions), []
    sup, rows, _ = ctx.intersect_extensions(prefix, extensions, keep_min=0, want_rows=True)
    return rows, sup.tolist()
