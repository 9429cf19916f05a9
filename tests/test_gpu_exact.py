"""The engine's exact brute-force search (puckgpu_exact_search: the ground
truth of recall and the probe's labels) against float64 brute force in
numpy: unit-norm float rows (distances within 1e-12 relative, ids equal
wherever the float64 distances are not within that of each other), integer
rows stored as uint8 with many exact ties (ids and distances equal, ties by
id), and point counts that end in a partial 65,536-point chunk."""
import numpy as np
import pytest

from puck_amd import datasets as D

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from puck_amd.engine import Context
    return Context(0)


def _brute(x, q, k):
    xd = x.astype(np.float64)
    ids, dist = [], []
    for r in range(q.shape[0]):
        d = ((xd - q[r].astype(np.float64)) ** 2).sum(1)
        o = np.lexsort((np.arange(x.shape[0]), d))[:k]  # by (distance, id)
        ids.append(o)
        dist.append(d[o])
    return np.array(ids), np.array(dist)


def test_exact_search_float_rows(ctx):
    x = D.unit_norm_mixture(150_000, 96, 500, seed=40, means_seed=41)
    q = D.unit_norm_mixture(64, 96, 500, seed=42, means_seed=41)
    ids, dist = ctx.exact_search(x, q, 10)
    want_i, want_d = _brute(x, q, 10)
    assert np.allclose(dist, want_d, rtol=1e-12, atol=0)
    for r in range(q.shape[0]):
        for j in range(10):
            if ids[r, j] != want_i[r, j]:
                # only a float64 near-tie may swap places
                assert abs(want_d[r, j] - dist[r, j]) <= 1e-12 * want_d[r, j]


@pytest.mark.parametrize("n,k", [(70_000, 10), (131_072, 37), (5_000, 100)])
def test_exact_search_uint8_rows_with_ties(ctx, n, k):
    rng = np.random.default_rng(n)
    x = rng.integers(0, 4, size=(n, 128)).astype(np.uint8)  # tiny range: many exactly tied distances
    q = rng.integers(0, 4, size=(40, 128)).astype(np.float32)
    ids, dist = ctx.exact_search(x, q, k)
    want_i, want_d = _brute(x, q, k)
    assert np.array_equal(dist, want_d)
    assert np.array_equal(ids.astype(np.int64), want_i)


def test_exact_search_device_tensors_match_host(ctx):
    import torch
    x = D.mixture_uint8_grid(80_000, 128, 100, seed=43, means_seed=44)
    q = D.mixture_uint8_grid(32, 128, 100, seed=45, means_seed=44)
    ih, dh = ctx.exact_search(x, q, 10)
    dev = torch.device("cuda:0")
    idv, ddv = ctx.exact_search(torch.as_tensor(x, device=dev), torch.as_tensor(q, device=dev), 10)
    assert np.array_equal(ih, idv.cpu().numpy().view(np.uint32)) and np.array_equal(dh, ddv.cpu().numpy())
    i8, d8 = ctx.exact_search(x.astype(np.uint8), q, 10)
    assert np.array_equal(ih, i8) and np.array_equal(dh, d8)
