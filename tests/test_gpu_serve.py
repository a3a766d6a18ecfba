"""Batched mining and the serving loop on the device: every request's
itemsets are the same served (in a mix, with overlapping slots and resident
stores), alone, and by the oracle."""

import random

import numpy as np
import pytest

import oracle
from cupami_amd import _hip, datagen
from cupami_amd.batch import BatchMiner
from cupami_amd.serve import Request, Server, mine_one_at_a_time, mine_sequential_resident
from helpers import coo, csr, random_rows

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return _hip.context(0)


def small_requests(n, seed):
    rng = random.Random(seed)
    out = []
    for i in range(n):
        labels, rr, cc, m = coo(random_rows(rng, 200 + 97 * i, 8 + 3 * i, 7))
        out.append(Request(i, 0.002 * i, rr, cc, m, len(labels), 4 + i))
    return out


def test_build_bits_batch_host(ctx):
    reqs = small_requests(6, 1)
    stores = ctx.build_bits_batch([(r.rows, r.cols, r.n_transactions, r.n_items) for r in reqs])
    for bits, r in zip(stores, reqs):
        assert np.array_equal(bits.download(), oracle.bits_build(r.rows, r.cols, r.n_transactions, r.n_items))
        bits.close()


def test_server_results_equal_alone_and_oracle(ctx):
    reqs = small_requests(9, 2)
    server = Server(hbm_budget=4 * max(r.footprint() for r in reqs), max_batch=4)
    try:
        served = server.run(reqs)
    finally:
        server.close()
    alone = mine_one_at_a_time(reqs, ctx)
    budget = 4 * max(r.footprint() for r in reqs)
    assert server.stats["builds"] == 9 and server.stats["evictions"] >= 1
    assert server.stats["resident_bytes"] <= budget
    for s, a, r in zip(served, alone, reqs):
        expect = oracle.eclat(oracle.bits_build(r.rows, r.cols, r.n_transactions, r.n_items), r.minsup)
        assert not s.rejected
        assert s.patterns == expect and a.patterns == expect


def test_server_micro_batches_on_resident_stores(ctx):
    """Many requests on four resident databases, all arriving at once, one
    slot: they are mined in batched searches, each == the oracle."""
    base = small_requests(4, 5)
    reqs = []
    for i in range(24):
        b = base[i % 4]
        reqs.append(Request(i, 0.0, b.rows, b.cols, b.n_transactions, b.n_items, b.minsup + i // 4, key=i % 4))
    server = Server(hbm_budget=1 << 30, slots=1, max_batch=8)
    try:
        served = server.run(reqs)
    finally:
        server.close()
    assert server.stats["builds"] == 4
    assert server.stats["max_batch"] > 1
    for s, r in zip(served, reqs):
        assert s.patterns == oracle.eclat(oracle.bits_build(r.rows, r.cols, r.n_transactions, r.n_items), r.minsup)


def test_server_bin_layout_requests(ctx):
    """Requests in the BIN layout (offsets + raw ids, empty baskets and
    repeats included; the device builds store and labels), mixed in batches
    with pairs requests: each result == the oracle on the reference's COO of
    the same arrays (cuApriori.py:381-396)."""
    rng = random.Random(8)
    reqs, expect = [], []
    for i in range(8):
        pool = rng.sample(range(1, 400), 6 + 2 * i)
        rows = [[rng.choice(pool) for _ in range(rng.randrange(7))] for _ in range(150 + 60 * i)]
        offsets, ids = csr(rows)
        labels, rr, cc, n = oracle.csr_to_coo(offsets, ids)
        expect.append(oracle.eclat(oracle.bits_build(rr, cc, n, len(labels)), 3 + i))
        if i % 3 == 2:   # all arrive at once, so they queue and batch
            reqs.append(Request(i, 0.0, rr, cc, n, len(labels), 3 + i))
        else:
            reqs.append(Request(i, 0.0, None, None, n, len(labels), 3 + i,
                                offsets=offsets, ids=ids, universe=399))
    server = Server(hbm_budget=3 * max(r.footprint() for r in reqs), max_batch=3)
    try:
        served = server.run(reqs)
    finally:
        server.close()
    alone = mine_one_at_a_time(reqs, ctx)
    for s, a, e in zip(served, alone, expect):
        assert not s.rejected and len(e) > 0
        assert s.patterns == e and a.patterns == e


def test_batch_miner_tune_and_mine(ctx):
    import torch
    reqs = small_requests(5, 3)
    dev = []
    for r in reqs:
        d_rows = torch.from_numpy(np.ascontiguousarray(r.rows, np.uint64).view(np.int64)).to("cuda:0")
        d_cols = torch.from_numpy(np.ascontiguousarray(r.cols, np.uint32).view(np.int32)).to("cuda:0")
        dev.append((d_rows, d_cols))
    torch.cuda.synchronize()
    specs = [(dr.data_ptr(), dc.data_ptr(), r.rows.size, r.n_transactions, r.n_items)
             for (dr, dc), r in zip(dev, reqs)]
    minsups = [r.minsup for r in reqs]
    miner = BatchMiner(ctx, configs=((16, 1024), (32, 2048)))
    config, table = miner.tune(specs, minsups)
    assert config in table
    parts = miner.mine(specs, minsups).split(len(reqs))
    for part, r in zip(parts, reqs):
        assert part == oracle.eclat(oracle.bits_build(r.rows, r.cols, r.n_transactions, r.n_items), r.minsup)
    ctx.set_search_options()


def test_serve_full_size_range_mixed_alone_oracle(ctx):
    """Databases across the served workload's size range (seeded pool specs
    from 10^4 up to 2*10^6 baskets over 2000 items): 24 requests with their
    own minimum supports, arriving 2 ms apart, served by three slots under a
    budget that holds only part of the stores (so stores are evicted and
    rebuilt); every result == one-at-a-time == resident one-at-a-time ==
    the oracle."""
    pool = datagen.serve_pool(48)
    pick = sorted(pool, key=lambda s: s.n_transactions)
    specs = [pick[0], pick[12], pick[24], pick[36], pick[-2],
             datagen.DbSpec(48, 2_000_000, 2000, 8.0, 1.0, 505048)]
    print("serve: generating " + ", ".join(f"{s.n_transactions}x{s.n_items}" for s in specs), flush=True)
    dbs = datagen.build_pool(specs)
    data = []
    for db in dbs:
        labels, cols = datagen.label_columns(db.items)
        data.append((db, cols, len(labels)))
    rng = np.random.default_rng(77)
    reqs = []
    for i in range(24):
        k = int(rng.integers(0, len(specs)))
        db, cols, items = data[k]
        frac = float(np.exp(rng.uniform(np.log(0.002 if k == 5 else 0.001), np.log(0.05))))
        minsup = int(np.ceil(frac * db.n_transactions))
        reqs.append(Request(i, 0.002 * i, db.rows, cols, db.n_transactions, items, minsup, key=k))
    # the 2M x 2000 request and two more are in the mix
    reqs[3] = Request(3, 0.006, data[5][0].rows, data[5][1], 2_000_000, data[5][2], 10_000, key=5)
    # the two largest stores never fit together
    fps = sorted({r.key: r.footprint() for r in reqs}.values(), reverse=True)
    budget = fps[0] + fps[1] - 1
    server = Server(hbm_budget=budget, slots=3)
    try:
        served = server.run(reqs)
    finally:
        server.close()
    print(f"serve: {server.stats}", flush=True)
    assert server.stats["evictions"] >= 1
    alone = mine_one_at_a_time(reqs, ctx)
    resident = mine_sequential_resident(reqs, ctx)
    hosts, expect = {}, {}
    for r in reqs:
        if (r.key, r.minsup) not in expect:
            if r.key not in hosts:
                hosts[r.key] = oracle.bits_build(r.rows, r.cols, r.n_transactions, r.n_items)
            expect[(r.key, r.minsup)] = oracle.eclat(hosts[r.key], r.minsup)
    for s, a, q, r in zip(served, alone, resident, reqs):
        e = expect[(r.key, r.minsup)]
        assert not s.rejected and len(e) > 0
        assert s.patterns == e, f"request {r.rid} served"
        assert a.patterns == e and q.patterns == e, f"request {r.rid} alone"
