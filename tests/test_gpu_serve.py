"""Batched mining and the serving loop on the device: every request's
itemsets are the same served in a batch, alone, and by the oracle."""

import random

import numpy as np
import pytest

import oracle
from cupami_amd import _hip
from cupami_amd.batch import BatchMiner
from cupami_amd.serve import Request, Server, mine_one_at_a_time
from helpers import coo, random_rows

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
