"""Batched mining: many databases in one device search.

The reference mines one database per miner object (``cuEclat.mine``,
``cuEclat.py:168-245``), one kernel launch per prefix class.  Here a batch of
databases shares everything: one call builds all stores into one device
allocation (``cpm_bits_build_device_batch``), one search runs their levels
together (``cpm_eclat_mine_batch``: one level step, and one support launch per
level, serves every database), and one download returns the pattern forest,
each pattern tagged with its database.

``BatchMiner.tune`` picks the search schedule (entries per support work range,
words per tile) by timing each candidate on the batch itself; the choice is
cached per batch shape, in memory and optionally in a JSON file.
"""

from __future__ import annotations

import json
import math
import os
import time
from pathlib import Path

from . import _hip

# (entries per work range, words per tile); the search narrows the range on
# short levels by itself
DEFAULT_CONFIGS = ((16, 1024), (32, 1024), (16, 2048), (32, 2048), (64, 2048))


def shape_key(specs):
    """Batch shape for the tuning cache: database count and the log2 buckets
    of total transactions, total items and total pairs."""
    n = len(specs)
    t = sum(int(s[3]) for s in specs)
    i = sum(int(s[4]) for s in specs)
    p = sum(int(s[2]) for s in specs)
    b = lambda v: int(math.log2(v)) if v > 0 else -1
    return f"{n}:{b(t)}:{b(i)}:{b(p)}"


class BatchMiner:
    """Mine many databases together on one context.

    ``specs``: one ``(d_rows, d_cols, nnz, n_transactions, n_items)`` per
    database, with the COO pairs already in device memory (``d_rows`` and
    ``d_cols`` are device addresses); ``minsups`` are absolute counts.
    """

    def __init__(self, ctx=None, configs=DEFAULT_CONFIGS, cache_path=None):
        self.ctx = ctx or _hip.context(0)
        self.configs = tuple(configs)
        self.cache_path = Path(cache_path or os.environ.get("CUPAMI_AMD_AUTOTUNE_CACHE", "")) \
            if (cache_path or os.environ.get("CUPAMI_AMD_AUTOTUNE_CACHE")) else None
        self._cache = {}
        if self.cache_path and self.cache_path.exists():
            try:
                self._cache = {k: tuple(v) for k, v in json.loads(self.cache_path.read_text()).items()}
            except (OSError, ValueError):
                self._cache = {}
        self.config = (32, 1024)

    def build(self, specs):
        return self.ctx.build_bits_device_batch(specs)

    def mine(self, specs, minsups, use_gram=True):
        """Build, search and download; returns the forest (``Patterns``, with
        ``db`` per pattern).  The stores are released before returning."""
        stores = self.build(specs)
        try:
            return self.ctx.eclat_mine_batch(stores, minsups, use_gram=use_gram)
        finally:
            for s in stores:
                s.close()

    def use(self, config):
        group, tile = config
        self.ctx.set_search_options(group, tile, True, True)
        self.config = (group, tile)

    def tune(self, specs, minsups, repeats=1):
        """Time every candidate schedule on this batch (after one warm run)
        and keep the fastest; returns ``(config, {config: seconds})``."""
        key = shape_key(specs)
        if key in self._cache:
            self.use(self._cache[key])
            return self._cache[key], {}
        table = {}
        self.use(self.configs[0])
        self.mine(specs, minsups)  # warm: buffers and graphs
        for cfg in self.configs:
            self.use(cfg)
            self.mine(specs, minsups)  # graph capture for this schedule
            best = float("inf")
            for _ in range(repeats):
                self.ctx.synchronize()
                t = time.perf_counter()
                self.mine(specs, minsups)
                self.ctx.synchronize()
                best = min(best, time.perf_counter() - t)
            table[cfg] = best
        config = min(table, key=table.get)
        self.use(config)
        self._cache[key] = config
        if self.cache_path:
            try:
                self.cache_path.parent.mkdir(parents=True, exist_ok=True)
                self.cache_path.write_text(json.dumps({k: list(v) for k, v in self._cache.items()}))
            except OSError:
                pass
        return config, table
