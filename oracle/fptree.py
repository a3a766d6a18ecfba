"""FP-tree oracle (numpy).  TEST INFRASTRUCTURE ONLY.

Restates ``_fptree.unique_pairs``, ``build_tree`` and ``conditional_tree``
(``_fptree.py:97-231``) and the FP-growth search of ``cuFPGrowth._search``
(``cuFPGrowth.py:187-213``) with numpy in place of CuPy.  Node numbering is
the reference's: nodes of each depth ordered by (parent, item).

``trie_tree`` builds the same arrays from the definition instead (a dict
trie over each basket's frequent items in (support desc, id asc) order,
numbered level by level in (parent, item) order); the tests check the two
against each other, and FP-growth's itemsets against brute force.
"""

from __future__ import annotations

from dataclasses import dataclass
from itertools import combinations

import numpy as np


@dataclass
class Tree:
    parents: np.ndarray
    items: np.ndarray
    counts: np.ndarray
    header: list
    single_path: bool
    depth: int


def unique_pairs(rows, cols):
    rows = np.asarray(rows, np.uint64)
    cols = np.asarray(cols, np.uint32)
    if not rows.size:
        return rows, cols
    order = np.lexsort((cols.astype(np.uint64), rows))
    rows, cols = rows[order], cols[order]
    keep = np.concatenate(([True], (rows[1:] != rows[:-1]) | (cols[1:] != cols[:-1])))
    return rows[keep], cols[keep]


def build_tree(rows, cols, weights, item_count, min_support, supports=None):
    rows = np.asarray(rows, np.int64)
    cols = np.asarray(cols, np.int64)
    weights = np.asarray(weights, np.uint64)
    if not rows.size:
        return None
    if supports is None:
        supports = np.zeros(item_count, np.uint64)
        np.add.at(supports, cols, weights[rows])
        frequent = supports >= np.uint64(min_support)
        keep = frequent[cols]
        rows, cols = rows[keep], cols[keep]
        if not rows.size:
            return None
    else:
        supports = np.asarray(supports, np.uint64)
        frequent = supports >= np.uint64(min_support)
    item_order = np.lexsort((np.arange(item_count, dtype=np.uint64), supports.max() - supports))
    rank = np.empty(item_count, np.int64)
    rank[item_order] = np.arange(item_count)
    order = np.lexsort((rank[cols], rows))
    rows, cols = rows[order], cols[order]
    lengths = np.bincount(rows, minlength=len(weights))
    offsets = np.concatenate(([0], np.cumsum(lengths)))
    parent_for_row = np.zeros(len(weights), np.int64)
    histogram = np.bincount(lengths)
    depth = len(histogram) - 1
    active_counts = np.cumsum(histogram[::-1])[::-1][1:]
    active_order = np.argsort(lengths, kind="stable")[::-1]
    parent_parts = [np.full(1, -1, np.int64)]
    item_parts = [np.full(1, -1, np.int64)]
    count_parts = [np.zeros(1, np.uint64)]
    node_count = 1
    single_path = True
    for level in range(depth):
        active = active_order[:int(active_counts[level])]
        level_items = cols[offsets[active] + level]
        level_parents = parent_for_row[active]
        level_order = np.lexsort((level_items, level_parents))
        parents = level_parents[level_order]
        items = level_items[level_order]
        first = np.concatenate(([True], (parents[1:] != parents[:-1]) | (items[1:] != items[:-1])))
        groups = np.cumsum(first) - 1
        group_count = int(groups[-1]) + 1
        counts = np.zeros(group_count, np.uint64)
        np.add.at(counts, groups, weights[active[level_order]])
        parent_parts.append(parents[first])
        item_parts.append(items[first])
        count_parts.append(counts)
        row_nodes = np.empty(len(active), np.int64)
        row_nodes[level_order] = groups + node_count
        parent_for_row[active] = row_nodes
        node_count += group_count
        single_path = single_path and group_count == 1
    header_items = item_order[frequent[item_order]][::-1]
    header = list(zip(header_items.tolist(), supports[header_items].tolist()))
    return Tree(np.concatenate(parent_parts), np.concatenate(item_parts),
                np.concatenate(count_parts), header, single_path, depth)


def ancestor_supports(tree, item, item_count):
    sup = np.zeros(item_count, np.uint64)
    for node in np.flatnonzero(tree.items == item):
        w = tree.counts[node]
        p = tree.parents[node]
        while p > 0:
            sup[tree.items[p]] += w
            p = tree.parents[p]
    return sup


def conditional_tree(tree, item, item_count, min_support):
    sup = ancestor_supports(tree, item, item_count)
    header = np.flatnonzero(sup >= np.uint64(min_support))
    if not header.size:
        return None
    if header.size == 1:
        sel = int(header[0])
        return Tree(np.array([-1, 0], np.int64), np.array([-1, sel], np.int64),
                    np.array([0, sup[sel]], np.uint64), [(sel, int(sup[sel]))], True, 1)
    frequent = sup >= np.uint64(min_support)
    nodes = np.flatnonzero(tree.items == item)
    rows, cols = [], []
    for r, node in enumerate(nodes):
        p = tree.parents[node]
        while p > 0:
            if frequent[tree.items[p]]:
                rows.append(r)
                cols.append(int(tree.items[p]))
            p = tree.parents[p]
    weights = tree.counts[nodes]
    return build_tree(np.array(rows, np.int64), np.array(cols, np.int64), weights, item_count,
                      min_support, supports=sup)


def fpgrowth(rows, cols, n_transactions, item_count, min_support):
    """``{item tuple (ascending ids): support}``."""
    rows, cols = unique_pairs(rows, cols)
    weights = np.ones(n_transactions, np.uint64)
    root = build_tree(rows, cols, weights, item_count, min_support)
    out = {}
    if root is None:
        return out
    stack = [[root, (), 0]]
    while stack:
        tree, suffix, position = stack[-1]
        if tree.single_path:
            path_items = tree.items[1:].tolist()
            path_counts = tree.counts[1:].tolist()
            for length in range(1, len(path_items) + 1):
                for sel in combinations(range(len(path_items)), length):
                    ids = suffix + tuple(path_items[i] for i in sel)
                    out[tuple(sorted(ids))] = int(min(path_counts[i] for i in sel))
            stack.pop()
            continue
        if position >= len(tree.header):
            stack.pop()
            continue
        stack[-1][2] += 1
        item, support = tree.header[position]
        nxt = suffix + (item,)
        out[tuple(sorted(nxt))] = int(support)
        cond = conditional_tree(tree, item, item_count, min_support)
        if cond is not None:
            stack.append([cond, nxt, 0])
    return out


def trie_tree(rows, cols, weights, item_count, min_support):
    """The tree from its definition: a trie of the frequent items of each
    basket in (weighted support desc, id asc) order, numbered by depth and
    then (parent id, item id)."""
    rows = np.asarray(rows, np.int64)
    cols = np.asarray(cols, np.int64)
    weights = np.asarray(weights, np.uint64)
    sup = {}
    for r, c in zip(rows.tolist(), cols.tolist()):
        sup[c] = sup.get(c, 0) + int(weights[r])
    freq = {c for c, s in sup.items() if s >= min_support}
    baskets = {}
    for r, c in zip(rows.tolist(), cols.tolist()):
        if c in freq:
            baskets.setdefault(r, set()).add(c)
    if not baskets:
        return None
    key = lambda c: (-sup[c], c)
    seqs = {r: sorted(b, key=key) for r, b in baskets.items()}
    # prefix -> count, then number level by level
    counts = {}
    for r, seq in seqs.items():
        for L in range(1, len(seq) + 1):
            p = tuple(seq[:L])
            counts[p] = counts.get(p, 0) + int(weights[r])
    node_of = {(): 0}
    parents, items, cnts = [-1], [-1], [0]
    depth = max(len(s) for s in seqs.values())
    single = True
    for L in range(1, depth + 1):
        level = [p for p in counts if len(p) == L]
        level.sort(key=lambda p: (node_of[p[:-1]], p[-1]))
        single = single and len(level) == 1
        for p in level:
            node_of[p] = len(parents)
            parents.append(node_of[p[:-1]])
            items.append(p[-1])
            cnts.append(counts[p])
    order = sorted(range(item_count), key=lambda c: (-sup.get(c, 0), c))
    header = [(c, sup[c]) for c in reversed(order) if c in freq]
    return Tree(np.array(parents, np.int64), np.array(items, np.int64), np.array(cnts, np.uint64),
                header, single, depth)
