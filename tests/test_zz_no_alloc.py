"""cpm_count_supports_noalloc (the entry point that allocates no device
memory) against cpm_count_supports and the oracle: two calls in a row on the
same device scratch, the second smaller, bit-exact."""

import numpy as np
import pytest

import oracle
from cupami_amd import _hip

pytestmark = pytest.mark.gpu


def _store(seed, n, n_items, density):
    rng = np.random.default_rng(seed)
    mask = rng.random((n, n_items)) < density
    r, c = np.nonzero(mask)
    return r.astype(np.uint64), c.astype(np.uint32)


def _candidates(rng, n_items, n_cand, width):
    """Sorted candidates with runs of shared prefixes (as a level's
    extensions come), plus a few singletons of other prefixes."""
    cand = np.sort(rng.integers(0, n_items, size=(n_cand, width)), axis=1).astype(np.uint32)
    cand = cand[np.lexsort(cand.T[::-1])]
    return np.ascontiguousarray(cand)


def _noalloc(ctx, bits, cand):
    import torch
    d_cand = torch.from_numpy(cand.view(np.int32).reshape(-1).copy()).to("cuda:0")
    d_sup = torch.full((cand.shape[0],), -1, dtype=torch.int64, device="cuda:0")
    torch.cuda.synchronize()
    ctx.count_supports_noalloc(bits, d_cand.data_ptr(), cand.shape[0], cand.shape[1], d_sup.data_ptr())
    return d_sup.cpu().numpy().view(np.uint64)


def test_noalloc_equals_count_supports_twice():
    ctx = _hip.context(0)
    rng = np.random.default_rng(9)
    rows, cols = _store(1, 100_003, 40, 0.2)
    bits = ctx.build_bits(rows, cols, 100_003, 40)
    host = oracle.bits_build(rows, cols, 100_003, 40)
    # first call: 5000 candidates of width 3; second: 1200 of width 2
    for n_cand, width in ((5000, 3), (1200, 2)):
        cand = _candidates(rng, 40, n_cand, width)
        print(f"noalloc: {n_cand} x {width}", flush=True)
        got = _noalloc(ctx, bits, cand)
        want = ctx.count_supports(bits, cand)
        assert np.array_equal(got, want)
        assert np.array_equal(got, np.asarray(oracle.count_supports(host, cand), np.uint64))


def test_noalloc_limits_and_bad_column():
    import torch
    ctx = _hip.context(0)
    rows, cols = _store(2, 5000, 10, 0.3)
    bits = ctx.build_bits(rows, cols, 5000, 10)
    # the largest call allowed, then one candidate more (rejected on the host)
    cand = _candidates(np.random.default_rng(3), 10, 65536, 2)
    got = _noalloc(ctx, bits, cand)
    assert np.array_equal(got, ctx.count_supports(bits, cand))
    d = torch.zeros(65537 * 2, dtype=torch.int32, device="cuda:0")
    s = torch.zeros(65537, dtype=torch.int64, device="cuda:0")
    with pytest.raises(ValueError, match="at most 65536"):
        ctx.count_supports_noalloc(bits, d.data_ptr(), 65537, 2, s.data_ptr())
    # a column beyond the store: the table kernel flags it, no support work runs
    bad = np.array([[0, 3], [1, 10]], np.uint32)
    with pytest.raises(ValueError, match="candidate column"):
        _noalloc(ctx, bits, bad)
    # the scratch is usable again afterwards
    ok = np.array([[0, 3], [1, 9]], np.uint32)
    assert np.array_equal(_noalloc(ctx, bits, ok), ctx.count_supports(bits, ok))
