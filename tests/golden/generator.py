"""Restatement of the reference's basket generators, for golden data only.

``TransactionalDatabase._generateTransactions``
(``cupami/extras/syntheticDataGenerator/TransactionalDatabase.py:103-114``):
a symmetric uniform length around ``avg_length`` with stochastic rounding,
then ``random.sample`` of the 1-based ids, sorted.  ``TemporalDatabase``
(``TemporalDatabase.py:40-67``) prefixes a timestamp that starts at 1 and
steps by ``randint(0 if repeat else 1, 2 if missing else 1)`` when either
flag is set, drawn after each basket; merging unions baskets that share a
timestamp.  Both consume one ``random.Random(seed)`` stream in the same order,
so the rows printed by the reference's notebooks pin this code
(``reference_kats.json``: ``generator_rows``).
"""

from __future__ import annotations

import math
import random


def _baskets(rnd, transactions, items, avg_length):
    if avg_length is None:
        avg_length = (items + 1) / 2
    avg = float(avg_length)
    ids = range(1, items + 1)
    radius = min(avg - 1, items - avg)
    for _ in range(transactions):
        target = rnd.uniform(avg - radius, avg + radius)
        length = math.floor(target)
        if rnd.random() < target - length:
            length += 1
        yield sorted(rnd.sample(ids, length))


def transactional(transactions, items, avg_length=None, seed=None):
    rnd = random.Random(seed)
    return list(_baskets(rnd, transactions, items, avg_length))


def temporal(transactions, items, avg_length=None, seed=None, repeat=False, missing=False,
             merge=False):
    rnd = random.Random(seed)
    out = []
    stamp = 1
    pending_stamp, pending = None, set()
    for index, basket in enumerate(_baskets(rnd, transactions, items, avg_length)):
        if index:
            step = 1
            if repeat or missing:
                step = rnd.randint(0 if repeat else 1, 2 if missing else 1)
            stamp += step
        if not merge:
            out.append([stamp, *basket])
            continue
        if pending_stamp is not None and pending_stamp != stamp:
            out.append([pending_stamp, *sorted(pending)])
            pending.clear()
        pending_stamp = stamp
        pending.update(basket)
    if pending_stamp is not None and merge:
        out.append([pending_stamp, *sorted(pending)])
    return out
