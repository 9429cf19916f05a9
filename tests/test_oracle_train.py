"""The oracle's k-means and training loop (oracle/train_oracle.cpp, restating
puck/hierarchical_cluster/kmeans.cpp and hierarchical_cluster_index.cpp:
882-985) against what the mathematics fixes, in float64:

* mean_mode = 1 (plain Lloyd): every centroid is the mean of the points the
  last assignment gave it, and the error sum taken at each assignment does
  not increase;
* mean_mode = 0 (the reference: only the first 4 bytes of the centroids are
  cleared before accumulating, kmeans.cpp:275): every centroid is
  (previous centroid + sum of its points) / count, element [0][0] from 0;
* an empty cluster is split off a populated one: the pair is symmetric about
  the donor's old centroid and a tiny distance apart (kmeans.cpp:112-147);
* k-means++ seeding draws by squared distance: two far blobs give one seed
  in each;
* the assignments are nearest_center's (the oracle's own, pinned elsewhere);
* the training loop's round count follows the keep-best / niter rule, re-run
  here in Python on its error trace, and fine centers come out by descending
  norm.
"""
import numpy as np
import pytest

from oracle import oracle as O
from puck_amd import datasets as D

F64 = np.float64


@pytest.fixture(scope="module")
def data():
    return D.mixture_uint8_grid(4000, 16, 25, seed=140, means_seed=141)


def _members_mean(x, a, k):
    out = np.zeros((k, x.shape[1]), F64)
    cnt = np.bincount(a, minlength=k)
    np.add.at(out, a, x.astype(F64))
    return out, cnt


def test_mean_mode_centroids_are_member_means(data):
    x = data
    k = 24
    init = x[np.random.default_rng(1).choice(x.shape[0], k, replace=False)]
    cent, a, dis, q, it, sp, tr = O.kmeans_core(x, init, 12, seed=2, mean_mode=1)
    s, cnt = _members_mean(x, a, k)
    ok = cnt > 0
    want = s[ok] / cnt[ok, None]
    # float32 sequential accumulation of up to a few hundred points
    assert np.allclose(cent[ok], want, rtol=1e-5, atol=1e-4 * np.abs(x).max())
    # the error at each assignment does not increase (Lloyd), up to rounding
    assert np.all(np.diff(tr) <= 1e-6 * tr[:-1] + 1e-3)
    assert it >= 2


def test_reference_mode_accumulates_onto_previous_centroids(data):
    x = data
    k = 16
    init = x[np.random.default_rng(3).choice(x.shape[0], k, replace=False)].copy()
    cent, a, dis, q, it, sp, tr = O.kmeans_core(x, init, 1, seed=4, mean_mode=0)
    assert it == 1 and sp == 0
    s, cnt = _members_mean(x, a, k)
    prev = init.astype(F64)
    prev[0, 0] = 0.0
    want = (prev + s) / np.maximum(cnt, 1)[:, None]
    assert np.allclose(cent, want, rtol=1e-5, atol=1e-4 * np.abs(x).max())
    # differs from the plain mean
    assert np.abs(cent - s / np.maximum(cnt, 1)[:, None]).max() > 1e-2


@pytest.mark.parametrize("mean_mode", [0, 1])
def test_empty_cluster_is_split_off_a_populated_one(data, mean_mode):
    x = data
    k = 10
    init = x[np.random.default_rng(5).choice(x.shape[0], k, replace=False)].copy()
    init[7] = init[3]          # a duplicate: strict '<' gives every tie to 3, 7 stays empty
    init[9] = 1e6              # far away: no point
    cent, a, dis, q, it, sp, tr = O.kmeans_core(x, init, 1, seed=6, mean_mode=mean_mode)
    assert sp == 2
    cnt = np.bincount(a, minlength=k)
    assert cnt[7] == 0 and cnt[9] == 0
    # each empty one sits symmetric to some populated donor, a tiny step apart
    for e in (7, 9):
        gap = np.abs(cent[e][None, :] - cent).max(1)
        gap[e] = np.inf
        j = int(np.argmin(gap))
        assert cnt[j] >= 2
        scale = np.sqrt((cent[j].astype(F64) ** 2).sum())
        assert 0 < gap[j] <= 1e-5 * scale


def test_kmeanspp_seeds_by_squared_distance():
    rng = np.random.default_rng(7)
    a = rng.normal(0, 0.01, (500, 8)).astype(np.float32)
    b = (rng.normal(0, 0.01, (500, 8)) + 100).astype(np.float32)
    x = np.concatenate([a, b])
    for seed in range(8):
        cent, _, _, it, _ = O.kmeans(x, 2, niter=0, init_pp=True, seed=seed)
        assert it == 0
        assert sorted((cent[:, 0] > 50).tolist()) == [False, True], seed


def test_kmeans_assignment_is_nearest_center(data):
    cent, a, err, it, sp = O.kmeans(data, 20, niter=5, init_pp=False, seed=8)
    # the returned assignment is the last one, taken before the final update:
    # it is nearest_center of the points against the previous centroids, so
    # recompute from a core run of one more step
    c2, a2, dis2, q2, _, _, _ = O.kmeans_core(data, cent, 1, seed=9, mean_mode=0)
    a_np = ((data[:, None, :].astype(F64) - cent[None].astype(F64)) ** 2).sum(-1).argmin(1)
    d = ((data[:, None, :].astype(F64) - cent[None].astype(F64)) ** 2).sum(-1)
    best = d.min(1)
    assert np.all(d[np.arange(data.shape[0]), a2] <= best + 1e-3 * (1 + best))
    assert np.mean(a2 == a_np) > 0.99


def _simulate_rounds(trace, iterations):
    min_err, niter, rounds = np.float32(np.finfo(np.float32).max), 30, 0
    for ite in range(iterations):
        rounds += 1
        err = np.float32(trace[ite, 1]) if ite < len(trace) else None
        if err is None:
            break
        if float(np.float32(min_err - err)) >= 1e-4:
            min_err = err
        else:
            if niter > 80:
                break
            niter += 10
    return rounds, niter


@pytest.mark.parametrize("mean_mode", [0, 1])
def test_training_rounds_follow_the_keep_best_rule(data, mean_mode):
    co, fi, tr, rounds, niter = O.train_centers(data, 8, 8, iterations=12, seed=10, mean_mode=mean_mode)
    assert (rounds, niter) == _simulate_rounds(tr, 12)
    key = np.array([0 - np.sqrt(np.float32(O.dot(f, f)) / np.float32(2)) for f in fi], np.float32)
    assert np.all(np.diff(key) >= 0)  # descending norm
    assert np.all(np.isfinite(co)) and np.all(np.isfinite(fi))


def test_training_is_reproducible_and_seeded(data):
    r1 = O.train_centers(data, 6, 6, iterations=3, seed=11)
    r2 = O.train_centers(data, 6, 6, iterations=3, seed=11)
    r3 = O.train_centers(data, 6, 6, iterations=3, seed=12)
    assert np.array_equal(r1[0], r2[0]) and np.array_equal(r1[1], r2[1])
    assert not np.array_equal(r1[0], r3[0])
