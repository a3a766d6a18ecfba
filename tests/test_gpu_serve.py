"""Batched mining and the serving loop on the device: every request's
itemsets are the same served in a batch, alone, and by the oracle."""

import random

import numpy as np
import pytest

import oracle
from cupami_amd import _hip
from cupami_amd.batch import BatchMiner
from cupami_amd.serve import Request, Server, mine_one_at_a_time
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
    small = min(r.footprint() for r in reqs)
    server = Server(ctx, hbm_budget=4 * max(r.footprint() for r in reqs), max_batch=4)
    served = server.run(reqs)
    alone = mine_one_at_a_time(reqs, ctx)
    assert len(server.batches) >= 3  # at most four per batch
    for s, a, r in zip(served, alone, reqs):
        expect = oracle.eclat(oracle.bits_build(r.rows, r.cols, r.n_transactions, r.n_items), r.minsup)
        assert not s.rejected
        assert s.patterns == expect and a.patterns == expect
    assert small > 0


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
    server = Server(ctx, hbm_budget=3 * max(r.footprint() for r in reqs), max_batch=3)
    served = server.run(reqs)
    alone = mine_one_at_a_time(reqs, ctx)
    assert any(len(b[2]) > 1 for b in server.batches)
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
