"""Operator-level mirror of the reference's FP-tree module (``_fptree.py:85-231``).

The functions take the device context where the reference takes the array
module, host arrays for inputs, and return ``FPTree`` records whose
``parents``/``items``/``counts`` arrays, ``header``, ``single_path`` and
``depth`` have the reference's meaning and numbering.  The tree itself stays
on the device (``FPTree.device``) for further conditional projections.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class FPTree:
    parents: np.ndarray
    items: np.ndarray
    counts: np.ndarray
    header: list
    single_path: bool
    depth: int
    device: object = field(default=None, repr=False)


def _wrap(handle):
    if handle is None:
        return None
    parents, items, counts, header, single, depth = handle.arrays()
    return FPTree(parents, items, counts, header, single, depth, handle)


def unique_pairs(ctx, rows, columns):
    """Distinct (row, column) pairs in (row, column) order (``_fptree.py:97-105``)."""
    return ctx.unique_pairs(rows, columns)


def build_tree(ctx, rows, columns, weights, item_count, min_support, *, supports=None):
    """Weighted FP-tree, or None when nothing is frequent (``_fptree.py:108-177``)."""
    if not len(rows):
        return None
    return _wrap(ctx.fptree_build(rows, columns, weights, item_count, min_support, supports))


def conditional_tree(ctx, tree, item, item_count, min_support):
    """Conditional tree of ``item``, or None (``_fptree.py:204-231``)."""
    return _wrap(ctx.fptree_conditional(tree.device, item, min_support))
