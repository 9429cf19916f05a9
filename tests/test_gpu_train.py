"""GPU k-means and coarse / fine training (puckgpu_kmeans,
puckgpu_train_centers) against the oracle's restatement of Kmeans::kmeans
and HierarchicalClusterIndex::train (oracle/train_oracle.cpp): with the same
seed every centroid, assignment, error, iteration count and split matches bit
for bit -- both centroid-update modes, k-means++ and uniform seeding, short
(filter-PQ codebook) and long vectors, runs with empty clusters -- plus the
Lloyd invariants in float64 on the GPU's own output."""
import numpy as np
import pytest

from oracle import oracle as O
from puck_amd import datasets as D

pytestmark = pytest.mark.gpu
F64 = np.float64


@pytest.fixture(scope="module")
def ctx():
    from puck_amd.engine import Context
    return Context(0)


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("dim,k,init_pp,mean_mode", [(16, 24, True, 0), (16, 24, False, 1), (96, 40, True, 1),
                                                     (2, 256, True, 0), (4, 256, False, 1), (1, 64, True, 0)])
def test_kmeans_bitexact(ctx, dim, k, init_pp, mean_mode):
    x = D.mixture_uint8_grid(12000, max(dim, 4), 50, seed=150 + dim, means_seed=151)[:, :dim].copy()
    x = (x + np.random.default_rng(dim).normal(0, 0.3, x.shape)).astype(np.float32)
    g = ctx.kmeans(x, k, niter=25, init_pp=init_pp, seed=7, mean_mode=mean_mode)
    o = O.kmeans(x, k, niter=25, init_pp=init_pp, seed=7, mean_mode=mean_mode)
    assert g[3] == o[3] and g[4] == o[4]          # iterations, splits
    assert np.array_equal(g[1], o[1])             # assignment
    assert np.array_equal(_bits(g[0]), _bits(o[0]))
    assert np.float32(g[2]) == np.float32(o[2])


def test_kmeans_with_empty_clusters_bitexact(ctx):
    # 40 distinct rows, each 50 times: uniform seeding picks repeats, whose
    # clusters stay empty and are split off populated ones
    base = D.mixture_uint8_grid(40, 16, 10, seed=152, means_seed=153)
    x = np.repeat(base, 50, axis=0).astype(np.float32)
    for mean_mode in (0, 1):
        g = ctx.kmeans(x, 30, niter=10, init_pp=False, seed=3, mean_mode=mean_mode)
        o = O.kmeans(x, 30, niter=10, init_pp=False, seed=3, mean_mode=mean_mode)
        assert o[4] > 0
        assert (g[3], g[4]) == (o[3], o[4])
        assert np.array_equal(g[1], o[1]) and np.array_equal(_bits(g[0]), _bits(o[0]))


@pytest.mark.parametrize("mean_mode", [0, 1])
def test_train_centers_bitexact(ctx, mean_mode):
    x = D.mixture_uint8_grid(6000, 24, 40, seed=154, means_seed=155)
    g = ctx.train_centers(x, 10, 12, iterations=5, seed=9, mean_mode=mean_mode)
    o = O.train_centers(x, 10, 12, iterations=5, seed=9, mean_mode=mean_mode)
    assert (g[3], g[4]) == (o[3], o[4])
    assert np.array_equal(_bits(g[2]), _bits(o[2]))
    assert np.array_equal(_bits(g[0]), _bits(o[0])) and np.array_equal(_bits(g[1]), _bits(o[1]))


def test_kmeans_at_scale_bitexact(ctx):
    x = D.unit_norm_mixture(100000, 32, 400, seed=156, means_seed=157)
    g = ctx.kmeans(x, 256, niter=8, init_pp=True, seed=11, mean_mode=0)
    o = O.kmeans(x, 256, niter=8, init_pp=True, seed=11, mean_mode=0)
    assert np.array_equal(g[1], o[1]) and np.array_equal(_bits(g[0]), _bits(o[0]))


def test_gpu_lloyd_invariants_fp64(ctx):
    # plain mean mode: each centroid is the float64 mean of the points of the
    # last assignment (up to float accumulation), and the error does not
    # increase from one iteration to the next
    x = D.unit_norm_mixture(20000, 48, 100, seed=158, means_seed=159)
    k = 64
    errs = []
    for it in range(1, 8):
        cent, a, err, iters, sp = ctx.kmeans(x, k, niter=it, init_pp=True, seed=5, mean_mode=1)
        errs.append(err)
    s = np.zeros((k, x.shape[1]), F64)
    np.add.at(s, a, x.astype(F64))
    cnt = np.bincount(a, minlength=k)
    ok = cnt > 0
    assert np.allclose(cent[ok], s[ok] / cnt[ok, None], rtol=1e-5, atol=1e-6)
    e = np.array(errs, F64)
    assert np.all(np.diff(e) <= 1e-6 * e[:-1])
