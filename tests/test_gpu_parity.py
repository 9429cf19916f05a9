"""GPU parity: libpuckgpu.so (through the C ABI) against the CPU oracle on the
same seeded inputs.  Integer/byte/index outputs and the reference-order float
arithmetic are compared bit for bit."""
import numpy as np
import pytest

from oracle import oracle as O
from puck_amd import datasets as D
import cpu_index as CI

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from puck_amd.engine import Context
    return Context(0)


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def idx_l2():
    x = D.mixture_uint8_grid(12000, 128, 200, seed=10, means_seed=11)
    q = D.mixture_uint8_grid(64, 128, 200, seed=12, means_seed=11)
    ix = CI.build_index(x, 48, 48, 32, whether_norm=False, train_n=6000)
    return ix, q


@pytest.fixture(scope="module")
def idx_cos():
    x = D.unit_norm_mixture(15000, 96, 300, seed=20, means_seed=21)
    q = D.unit_norm_mixture(64, 96, 300, seed=22, means_seed=21)
    ix = CI.build_index(x, 40, 40, 48, whether_norm=True, train_n=6000)
    return ix, q


def _gpu_index(ctx, ix):
    from puck_amd.engine import GpuIndex
    return GpuIndex(ctx, ix["coarse"], ix["fine"], ix["codebook"].reshape(-1), ix["cell_assign"],
                    ix["records"], ix["raw"], ix["whether_norm"])


@pytest.mark.parametrize("lsq", [1, 2, 4])
def test_lut_bitexact(ctx, lsq):
    rng = np.random.default_rng(lsq)
    nsq = 16
    q = (rng.standard_normal((5, nsq * lsq)) * 30).astype(np.float32)
    cb = (rng.standard_normal(nsq * 256 * lsq) * 30).astype(np.float32)
    g = ctx.lut(q, cb, nsq)
    for i in range(5):
        assert np.array_equal(_bits(g[i]), _bits(O.lut(q[i], cb, nsq, lsq)))


@pytest.mark.parametrize("which", ["l2", "cos"])
def test_index_derivation_bitexact(ctx, which, idx_l2, idx_cos):
    ix, _ = idx_l2 if which == "l2" else idx_cos
    g = _gpu_index(ctx, ix).export()
    o = CI.oracle_index(ix).export()
    assert np.array_equal(_bits(g["coarse_norms"]), _bits(o["coarse_norms"]))
    assert np.array_equal(_bits(g["stationary"]), _bits(o["stationary"]))
    assert np.array_equal(g["cell_start"], o["cell_start"])
    assert np.array_equal(_bits(g["min_dist_offset"]), _bits(o["min_dist_offset"]))
    assert np.array_equal(g["mem2local"], o["mem2local"])


def _sorted_pairs(v, i):
    order = np.lexsort((i, v))
    return v[order], i[order]


@pytest.mark.parametrize("which,scc,ftk,topk", [("l2", 4, 100, 10), ("l2", 16, 500, 10), ("l2", 48, 300, 100),
                                                ("cos", 5, 200, 10), ("cos", 20, 1000, 10), ("cos", 40, 64, 1)])
def test_search_bitexact(ctx, which, scc, ftk, topk, idx_l2, idx_cos):
    ix, q = idx_l2 if which == "l2" else idx_cos
    gi = _gpu_index(ctx, ix)
    oi = CI.oracle_index(ix)
    Dg, Ig, Rg = gi.search(q, topk, scc, ftk)
    ft, fm, fn = gi.last_filter(q.shape[0], ftk)
    for k in range(q.shape[0]):
        o = oi.search_one(q[k], topk, scc, ftk)
        assert o["rc"] == 0
        # filter-stage set: same (distance, memory id) multiset
        assert fn[k] == o["filt_n"], f"query {k}: filter count {fn[k]} vs {o['filt_n']}"
        gv, gid = _sorted_pairs(ft[k, :fn[k]], fm[k, :fn[k]])
        ov, oid = _sorted_pairs(o["filt_val"][:o["filt_n"]], o["filt_mem"][:o["filt_n"]])
        assert np.array_equal(_bits(gv), _bits(ov)), f"query {k}: filter distances differ"
        assert np.array_equal(gid, oid), f"query {k}: filter ids differ"
        # final result
        assert Rg[k] == o["n"]
        assert np.array_equal(_bits(Dg[k, :o["n"]]), _bits(o["dist"][:o["n"]])), f"query {k}"
        assert np.array_equal(Ig[k, :o["n"]], o["ids"][:o["n"]]), f"query {k}"


@pytest.mark.parametrize("which,rr,replay", [("l2", 1.5, False), ("cos", 2.0, False), ("cos", 1.25, True)])
def test_search_radius_rate_bitexact(ctx, which, rr, replay, idx_l2, idx_cos):
    # radius_rate > 1 shrinks the pivot (puck_index.cpp:517, :544) and prunes
    # more cells; a small filter_topk makes the pivot move early
    ix, q = idx_l2 if which == "l2" else idx_cos
    gi = _gpu_index(ctx, ix)
    oi = CI.oracle_index(ix)
    for scc, ftk, topk in [(12, 40, 10), (30, 200, 10)]:
        Dg, Ig, Rg = gi.search(q, topk, scc, ftk, radius_rate=rr, replay_all=replay)
        ft, fm, fn = gi.last_filter(q.shape[0], ftk)
        for k in range(q.shape[0]):
            o = oi.search_one(q[k], topk, scc, ftk, radius_rate=rr)
            assert fn[k] == o["filt_n"], f"query {k}"
            gv, gid = _sorted_pairs(ft[k, :fn[k]], fm[k, :fn[k]])
            ov, oid = _sorted_pairs(o["filt_val"][:o["filt_n"]], o["filt_mem"][:o["filt_n"]])
            assert np.array_equal(_bits(gv), _bits(ov)), f"query {k}: filter distances differ"
            assert np.array_equal(gid, oid), f"query {k}: filter ids differ"
            assert Rg[k] == o["n"]
            assert np.array_equal(Ig[k, :o["n"]], o["ids"][:o["n"]]), f"query {k}"
            assert np.array_equal(_bits(Dg[k, :o["n"]]), _bits(o["dist"][:o["n"]])), f"query {k}"


@pytest.mark.parametrize("which", ["l2", "cos"])
def test_search_replay_path_bitexact(ctx, which, idx_l2, idx_cos):
    # every query through the reference-heap replay (the path taken when the
    # result depends on the order of exactly equal distances)
    ix, q = idx_l2 if which == "l2" else idx_cos
    gi = _gpu_index(ctx, ix)
    oi = CI.oracle_index(ix)
    for scc, ftk, topk in [(6, 300, 10), (2, 1000, 50)]:
        Dg, Ig, Rg = gi.search(q, topk, scc, ftk, replay_all=True)
        ft, fm, fn = gi.last_filter(q.shape[0], ftk)
        for k in range(q.shape[0]):
            o = oi.search_one(q[k], topk, scc, ftk)
            # replay returns the filter heap in the reference's array order
            assert fn[k] == o["filt_n"]
            assert np.array_equal(_bits(ft[k, :fn[k]]), _bits(o["filt_val"][:fn[k]]))
            assert np.array_equal(fm[k, :fn[k]], o["filt_mem"][:fn[k]])
            assert Rg[k] == o["n"]
            assert np.array_equal(Ig[k, :o["n"]], o["ids"][:o["n"]])
            assert np.array_equal(_bits(Dg[k, :o["n"]]), _bits(o["dist"][:o["n"]]))


def test_search_not_full_and_clamps(ctx, idx_l2):
    ix, q = idx_l2
    gi = _gpu_index(ctx, ix)
    oi = CI.oracle_index(ix)
    # scc > C is clamped; a tiny scc leaves the filter heap not full
    for scc, ftk, topk in [(1000, 200, 10), (1, 1024, 50)]:
        Dg, Ig, Rg = gi.search(q[:8], topk, scc, ftk, topk_max=100)
        for k in range(8):
            o = oi.search_one(q[k], topk, scc, ftk, max_topk=100)
            assert Rg[k] == o["n"]
            assert np.array_equal(Ig[k, :o["n"]], o["ids"][:o["n"]])
            assert np.array_equal(_bits(Dg[k, :o["n"]]), _bits(o["dist"][:o["n"]]))


@pytest.mark.parametrize("which,replay", [("l2", False), ("cos", False), ("cos", True)])
def test_search_batch_per_query_bitexact(ctx, which, replay, idx_l2, idx_cos):
    # per-query topk / search_coarse_count / fine budget, and an invalid topk
    ix, q = idx_l2 if which == "l2" else idx_cos
    gi = _gpu_index(ctx, ix)
    oi = CI.oracle_index(ix)
    n, F = q.shape[0], ix["fine"].shape[0]
    rng = np.random.default_rng(31)
    topk = rng.choice([1, 10, 100], n).astype(np.uint32)
    topk[7] = 0
    topk[9] = 101
    scc = rng.integers(1, 24, n).astype(np.uint32)
    nf = rng.integers(0, F + 4, n).astype(np.uint32)
    nf[:8] = [1, 2, 3, 5, F - 1, F, 0, 4]
    Dg, Ig, Rg = gi.search_batch(q, 100, 20, 1100, topk=topk, scc=scc, fine_budget=nf, replay_all=replay)
    ft, fm, fn = gi.last_filter(n, 1100)
    Do, Io, Ro = oi.search_batch_pq(q, 100, 20, 1100, topk=topk, scc_arr=scc, fine_budget=nf)
    assert np.array_equal(Rg, Ro)
    for k in range(n):
        if not (0 < nf[k] < F) or topk[k] == 0 or topk[k] > 100:
            continue
        # budget scan: the filter set (unordered rows) is the oracle's
        o = oi.search_one(q[k], int(topk[k]), int(min(scc[k], 20)), 1100, max_topk=100, fine_budget=int(nf[k]))
        assert fn[k] == o["filt_n"], k
        gv, gid = _sorted_pairs(ft[k, :fn[k]], fm[k, :fn[k]])
        assert np.array_equal(_bits(gv), _bits(o["filt_val"][:fn[k]])) and np.array_equal(gid, o["filt_mem"][:fn[k]])
    for k in range(n):
        m = int(Ro[k])
        assert np.array_equal(Ig[k, :m], Io[k, :m]), f"query {k}"
        assert np.array_equal(_bits(Dg[k, :m]), _bits(Do[k, :m])), f"query {k}"
        assert np.all(Ig[k, m:] == 0xffffffff) and np.all(_bits(Dg[k, m:]) == 0x7f7f7f7f), f"query {k}"


@pytest.fixture(scope="module")
def idx_wide():
    # F above 2048 (cells taken in two segments), nsq not a multiple of 16,
    # lists of ~12k points: budgets of hundreds of cells score thousands of points
    x = D.unit_norm_mixture(24000, 32, 400, seed=40, means_seed=41)
    q = D.unit_norm_mixture(16, 32, 400, seed=42, means_seed=41)
    ix = CI.build_index(x, 2, 2100, 8, whether_norm=True, train_n=24000)
    return ix, q


def test_budget_scan_wide_bitexact(ctx, idx_wide):
    # budget scan over several batches of selected cells and candidate-buffer
    # cuts, against the oracle's filter sets and results
    ix, q = idx_wide
    gi = _gpu_index(ctx, ix)
    oi = CI.oracle_index(ix)
    n, F = q.shape[0], ix["fine"].shape[0]
    assert F > 2048
    nf = np.array([1, 7, 300, 513, 600, 1100, 1500, F - 1, 0, F, 2000, 900, 64, 2048, 2049, 1024], np.uint32)[:n]
    scc = np.array([1, 2] * (n // 2), np.uint32)
    topk = np.full(n, 10, np.uint32)
    for ftk in (300, 1000):
        Dg, Ig, Rg = gi.search_batch(q, 10, 2, ftk, topk=topk, scc=scc, fine_budget=nf)
        ft, fm, fn = gi.last_filter(n, ftk)
        Do, Io, Ro = oi.search_batch_pq(q, 10, 2, ftk, topk=topk, scc_arr=scc, fine_budget=nf)
        assert np.array_equal(Rg, Ro)
        for k in range(n):
            if 0 < nf[k] < F:
                o = oi.search_one(q[k], 10, int(scc[k]), ftk, max_topk=10, fine_budget=int(nf[k]))
                assert fn[k] == o["filt_n"], (ftk, k)
                gv, gid = _sorted_pairs(ft[k, :fn[k]], fm[k, :fn[k]])
                assert np.array_equal(_bits(gv), _bits(o["filt_val"][:fn[k]])), (ftk, k)
                assert np.array_equal(gid, o["filt_mem"][:fn[k]]), (ftk, k)
            m = int(Ro[k])
            assert np.array_equal(Ig[k, :m], Io[k, :m]), (ftk, k)
            assert np.array_equal(_bits(Dg[k, :m]), _bits(Do[k, :m])), (ftk, k)


@pytest.fixture(scope="module")
def idx_dense():
    # few large cells (C = 4, F = 8, 40k points): a budget of most cells over
    # every list gives more candidates than the split scan's key buffer (8192),
    # a budget of 1 cell far fewer, so one batch mixes the split scan and the
    # per-query kernel it leaves queries to
    x = D.unit_norm_mixture(40000, 32, 64, seed=70, means_seed=71)
    q = D.unit_norm_mixture(48, 32, 64, seed=72, means_seed=71)
    ix = CI.build_index(x, 4, 8, 16, whether_norm=True, train_n=20000)
    return ix, q


@pytest.mark.parametrize("ftk", [64, 1000, 2048])
def test_budget_scan_split_and_overflow_bitexact(ctx, idx_dense, ftk):
    ix, q = idx_dense
    gi = _gpu_index(ctx, ix)
    oi = CI.oracle_index(ix)
    n, F = q.shape[0], ix["fine"].shape[0]
    nf = np.array([1, 2, 7, 6, 3, 5, 4, 7] * (n // 8), np.uint32)
    scc = np.array([4, 1, 4, 2, 3, 4, 1, 4] * (n // 8), np.uint32)
    topk = np.full(n, 10, np.uint32)
    Dg, Ig, Rg = gi.search_batch(q, 10, 4, ftk, topk=topk, scc=scc, fine_budget=nf)
    ft, fm, fn = gi.last_filter(n, ftk)
    ms, cnt = gi.last_stage_ms()
    Do, Io, Ro = oi.search_batch_pq(q, 10, 4, ftk, topk=topk, scc_arr=scc, fine_budget=nf)
    assert np.array_equal(Rg, Ro)
    big = 0
    for k in range(n):
        o = oi.search_one(q[k], 10, int(scc[k]), ftk, max_topk=10, fine_budget=int(nf[k]))
        big += o["stats"]["points_offered"] > 8192
        assert fn[k] == o["filt_n"], k
        gv, gid = _sorted_pairs(ft[k, :fn[k]], fm[k, :fn[k]])
        assert np.array_equal(_bits(gv), _bits(o["filt_val"][:fn[k]])), k
        assert np.array_equal(gid, o["filt_mem"][:fn[k]]), k
        m = int(Ro[k])
        assert np.array_equal(Ig[k, :m], Io[k, :m]), k
        assert np.array_equal(_bits(Dg[k, :m]), _bits(Do[k, :m])), k
    # both paths ran: some queries above the key buffer, most below
    assert 0 < big < n
    # the scan's counters (cells scored, points scored) match the oracle's
    assert int(cnt[3]) == sum(oi.search_one(q[k], 10, int(scc[k]), ftk, max_topk=10,
                                            fine_budget=int(nf[k]))["stats"]["points_offered"] for k in range(n))


@pytest.fixture(scope="module")
def idx_ties():
    # Integer data and centers, all distances exact integers.  Fine centers come
    # in pairs that differ by swapping coordinates 0 and 1 (coarse centers
    # too), and every query has q[0] == q[1]: both members of a pair are at
    # exactly the same distance from every query, with both cells populated.
    rng = np.random.default_rng(60)
    d = 16
    g = rng.integers(-6, 7, (2, d))
    g[:, 1] = g[:, 0] + rng.integers(3, 9, 2)
    coarse = np.concatenate([g, g[:, [1, 0] + list(range(2, d))]]) * 4 + 40
    h = rng.integers(-3, 4, (8, d))
    h[:, 1] = h[:, 0] + rng.integers(1, 4, 8)
    fine = np.concatenate([h, h[:, [1, 0] + list(range(2, d))]])
    n = 6000
    x = coarse[rng.integers(0, 4, n)] + fine[rng.integers(0, 16, n)] + rng.integers(-1, 2, (n, d))
    q = coarse[rng.integers(0, 4, 48)] + rng.integers(-4, 5, (48, d))
    q[:, 1] = q[:, 0]
    ix = CI.build_index_with_centers(x.astype(np.float32), coarse, fine, 8)
    return ix, q.astype(np.float32)


@pytest.mark.parametrize("scc,ftk,topk", [(1, 50, 10), (2, 200, 10), (4, 100, 20)])
def test_search_tied_centers_bitexact(ctx, idx_ties, scc, ftk, topk):
    # coarse and fine distance ties: replays (reference coarse / fine heap
    # orders, then the walk) must give the reference's result
    ix, q = idx_ties
    gi = _gpu_index(ctx, ix)
    oi = CI.oracle_index(ix)
    Dg, Ig, Rg = gi.search(q, topk, scc, ftk)
    _, cnt = gi.last_stage_ms()
    assert cnt[4] > 0 and cnt[8] > 0  # replays, some for a fine-tie decision
    for k in range(q.shape[0]):
        o = oi.search_one(q[k], topk, scc, ftk)
        assert Rg[k] == o["n"], k
        assert np.array_equal(Ig[k, :o["n"]], o["ids"][:o["n"]]), k
        assert np.array_equal(_bits(Dg[k, :o["n"]]), _bits(o["dist"][:o["n"]])), k


@pytest.mark.parametrize("topk", [1, 10, 40])
def test_budget_scan_exact_distance_ties_bitexact(ctx, idx_ties, topk):
    # integer distances: many survivors share an exact distance, so the
    # rerank's (distance, local id) order decides which ids are returned
    ix, q = idx_ties
    gi = _gpu_index(ctx, ix)
    oi = CI.oracle_index(ix)
    n = q.shape[0]
    nf = (np.arange(n) % 5 + 1).astype(np.uint32)
    Dg, Ig, Rg = gi.search_batch(q, topk, 2, 200, fine_budget=nf)
    Do, Io, Ro = oi.search_batch_pq(q, topk, 2, 200, fine_budget=nf)
    assert np.array_equal(Rg, Ro)
    ties = 0
    for k in range(n):
        m = int(Ro[k])
        assert np.array_equal(Ig[k, :m], Io[k, :m]), k
        assert np.array_equal(_bits(Dg[k, :m]), _bits(Do[k, :m])), k
        ties += int(np.any(np.diff(Dg[k, :m]) == 0)) if m > 1 else 0
    assert topk == 1 or ties > 0


def test_large_batch_query_order(ctx, idx_cos):
    # batches of 2048 queries or more run the per-query kernels in an order
    # grouped by nearest coarse list: results equal the same queries in small
    # batches (natural order) and the oracle
    ix, _ = idx_cos
    gi = _gpu_index(ctx, ix)
    q = D.unit_norm_mixture(2500, 96, 300, seed=23, means_seed=21)
    rng = np.random.default_rng(41)
    nf = rng.integers(0, 44, q.shape[0]).astype(np.uint32)
    for budgets in (None, nf):
        Dg, Ig, Rg = gi.search_batch(q, 10, 4, 300, fine_budget=budgets)
        for s0 in range(0, q.shape[0], 500):
            Ds, Is, Rs = gi.search_batch(q[s0:s0 + 500], 10, 4, 300,
                                         fine_budget=None if budgets is None else budgets[s0:s0 + 500])
            assert np.array_equal(Rg[s0:s0 + 500], Rs)
            assert np.array_equal(Ig[s0:s0 + 500], Is)
            assert np.array_equal(_bits(Dg[s0:s0 + 500]), _bits(Ds))
    oi = CI.oracle_index(ix)
    Dg, Ig, Rg = gi.search(q, 10, 4, 300)
    for k in range(0, q.shape[0], 97):
        o = oi.search_one(q[k], 10, 4, 300)
        assert Rg[k] == o["n"] and np.array_equal(Ig[k, :o["n"]], o["ids"][:o["n"]]), k
        assert np.array_equal(_bits(Dg[k, :o["n"]]), _bits(o["dist"][:o["n"]])), k


def test_search_batch_device_arrays(ctx, idx_cos):
    import torch
    ix, q = idx_cos
    gi = _gpu_index(ctx, ix)
    n = q.shape[0]
    rng = np.random.default_rng(5)
    topk = rng.choice([1, 10], n).astype(np.uint32)
    nf = rng.integers(1, 40, n).astype(np.uint32)
    Dh, Ih, Rh = gi.search_batch(q, 10, 8, 300, topk=topk, fine_budget=nf)
    qd = torch.from_numpy(q).cuda()
    Dd, Id, Rd = gi.search_batch(qd, 10, 8, 300, topk=topk, fine_budget=nf)
    assert np.array_equal(Rd.cpu().numpy().view(np.uint32), Rh)
    assert np.array_equal(Id.cpu().numpy().view(np.uint32), Ih)
    assert np.array_equal(_bits(Dd.cpu().numpy()), _bits(Dh))


def test_search_errors(ctx, idx_cos):
    from puck_amd.engine import PuckGpuError
    ix, q = idx_cos
    gi = _gpu_index(ctx, ix)
    with pytest.raises(PuckGpuError):
        gi.search(q[:2], 20, 5, 100, topk_max=10)   # topk > FLAGS_topk
    with pytest.raises(PuckGpuError):
        gi.search(q[:2], 10, 5, 5)                 # filter_topk < topk
    qz = q[:3].copy()
    qz[1] = 0.0                                    # zero-norm query under whether_norm
    Dg, Ig, Rg = gi.search(qz, 10, 5, 100)
    assert Rg[1] == 0 and Rg[0] == 10 and Rg[2] == 10


def test_search_max_filter_topk(ctx, idx_wide):
    # filter_topk at the engine's maximum on lists of ~12k points: the filter
    # set fills on the fast walk, on the reference-heap replay and in the
    # budget scan; one above the maximum is refused
    from puck_amd.engine import PuckGpuError
    ix, q = idx_wide
    gi = _gpu_index(ctx, ix)
    oi = CI.oracle_index(ix)
    ftk = int(ctx.L.puckgpu_filter_topk_max())
    assert ftk == 2048
    n = q.shape[0]
    for replay in (False, True):
        Dg, Ig, Rg = gi.search(q, 100, 2, ftk, replay_all=replay)
        ft, fm, fn = gi.last_filter(n, ftk)
        for k in range(n):
            o = oi.search_one(q[k], 100, 2, ftk)
            assert o["filt_n"] == ftk and fn[k] == ftk, (replay, k)
            if replay:  # the reference heap's array order
                assert np.array_equal(_bits(ft[k]), _bits(o["filt_val"][:ftk])), (replay, k)
                assert np.array_equal(fm[k], o["filt_mem"][:ftk]), (replay, k)
            else:
                gv, gid = _sorted_pairs(ft[k], fm[k])
                ov, oid = _sorted_pairs(o["filt_val"][:ftk], o["filt_mem"][:ftk])
                assert np.array_equal(_bits(gv), _bits(ov)) and np.array_equal(gid, oid), (replay, k)
            assert Rg[k] == o["n"] == 100
            assert np.array_equal(Ig[k], o["ids"][:100]), (replay, k)
            assert np.array_equal(_bits(Dg[k]), _bits(o["dist"][:100])), (replay, k)
    nf = np.array([1, 300, 2000, 0] * (n // 4), np.uint32)
    Dg, Ig, Rg = gi.search_batch(q, 100, 2, ftk, fine_budget=nf)
    ft, fm, fn = gi.last_filter(n, ftk)
    for k in range(n):
        o = oi.search_one(q[k], 100, 2, ftk, fine_budget=int(nf[k]))
        assert fn[k] == o["filt_n"], k
        gv, gid = _sorted_pairs(ft[k, :fn[k]], fm[k, :fn[k]])
        ov, oid = _sorted_pairs(o["filt_val"][:fn[k]], o["filt_mem"][:fn[k]])
        assert np.array_equal(_bits(gv), _bits(ov)) and np.array_equal(gid, oid), k
        assert Rg[k] == o["n"]
        assert np.array_equal(Ig[k, :o["n"]], o["ids"][:o["n"]]), k
        assert np.array_equal(_bits(Dg[k, :o["n"]]), _bits(o["dist"][:o["n"]])), k
    with pytest.raises(PuckGpuError):
        gi.search(q[:2], 10, 2, ftk + 1)


def test_empty_batches(ctx, idx_cos):
    # a batch of no queries returns at once from every search entry point
    from puck_amd.engine import Server
    ix, q = idx_cos
    gi = _gpu_index(ctx, ix)
    q0 = np.ascontiguousarray(q[:0])
    D0, I0, R0 = gi.search(q0, 10, 5, 100)
    assert D0.shape == (0, 10) and I0.shape == (0, 10) and R0.shape == (0,)
    D0, I0, R0 = gi.search_batch(q0, 10, 5, 100, topk=np.zeros(0, np.uint32), fine_budget=np.zeros(0, np.uint32))
    assert D0.shape == (0, 10) and R0.shape == (0,)
    sv = Server(gi, 10, 5, 100, max_batch=8)
    D0, I0, R0 = sv.search(q0)
    assert D0.shape == (0, 10) and R0.shape == (0,)
    Dg, Ig, Rg = sv.search(q[:3])  # the server still serves after an empty batch
    Dw, Iw, Rw = gi.search(q[:3], 10, 5, 100)
    assert np.array_equal(Ig, Iw) and np.array_equal(_bits(Dg), _bits(Dw)) and np.array_equal(Rg, Rw)
    sv.free()


def test_assign_cells_bitexact(ctx, idx_l2, idx_cos):
    rng = np.random.default_rng(9)
    for ix, _ in (idx_l2, idx_cos):
        x = ix["raw"][:3000]
        # fine centers in stored (descending-norm) order, then in a shuffled order
        for fine in (ix["fine"], np.ascontiguousarray(ix["fine"][rng.permutation(ix["fine"].shape[0])])):
            cg, dg = ctx.assign_cells(ix["coarse"], fine, x, with_dist=True)
            co, do = O.assign_cells(ix["coarse"], fine, x)
            assert np.array_equal(cg, co)
            assert np.array_equal(_bits(dg), _bits(do))


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_assign_modes_bitexact(ctx, idx_l2, idx_cos, mode):
    """The bounded assignment (mode 0), the sequential walk (1) and the bounded
    kernel with every point walked again (2) all give the oracle's cells and
    scores, also with exact ties between coarse clusters (duplicated coarse
    centers), between fine centers (duplicated fine centers), and with fine
    centers not stored by descending norm."""
    rng = np.random.default_rng(19)
    try:
        ctx.set_assign_mode(mode)
        for ix, _ in (idx_l2, idx_cos):
            x = ix["raw"][:4000]
            coarse = ix["coarse"].copy()
            coarse[5] = coarse[3]
            coarse[9] = coarse[1]
            fine = ix["fine"].copy()
            fine[7] = fine[2]
            for fn in (fine, np.ascontiguousarray(fine[rng.permutation(fine.shape[0])])):
                p0, s0 = ctx.assign_stats()
                cg, dg = ctx.assign_cells(coarse, fn, x, with_dist=True)
                p1, s1 = ctx.assign_stats()
                co, do = O.assign_cells(coarse, fn, x)
                assert np.array_equal(cg, co)
                assert np.array_equal(_bits(dg), _bits(do))
                assert p1 - p0 == x.shape[0]
                if mode == 0:
                    print(f"mode 0: {s1 - s0} of {x.shape[0]} points walked sequentially", flush=True)
                    assert s1 - s0 <= x.shape[0]
                else:
                    assert s1 - s0 == x.shape[0]
    finally:
        ctx.set_assign_mode(0)


def test_encode_filter_bitexact(ctx, idx_l2, idx_cos):
    for ix, _ in (idx_l2, idx_cos):
        x = ix["raw"][:3000]
        cell = ix["cell_assign"][:3000]
        rg = ctx.encode_filter(ix["coarse"], ix["fine"], ix["codebook"].reshape(-1), ix["nsq"], x, cell)
        ro = O.encode_filter(ix["coarse"], ix["fine"], ix["codebook"], ix["nsq"], x, cell)
        assert np.array_equal(rg, ro)


@pytest.mark.parametrize("lsq,nsq", [(1, 32), (2, 16), (4, 8)])
def test_encode_filter_near_ties_bitexact(ctx, lsq, nsq):
    # codes decided by the order of the short dot products (MKL's order for
    # 1, 2 or 4 floats): distances that cancel to rounding noise
    coarse, fine, cb, x, cell = CI.near_tie_encode_input(31, nsq, lsq, 3000)
    rg = ctx.encode_filter(coarse, fine, cb.reshape(-1), nsq, x, cell)
    ro = O.encode_filter(coarse, fine, cb, nsq, x, cell)
    assert np.array_equal(rg, ro)


def test_kmeans_assign_bitexact(ctx):
    rng = np.random.default_rng(3)
    for d, k in [(128, 300), (96, 64), (2, 256), (4, 256)]:
        x = (rng.standard_normal((4000, d)) * 10).astype(np.float32)
        c = (rng.standard_normal((k, d)) * 10).astype(np.float32)
        ag, dg = ctx.kmeans_assign(c, x, with_dist=True)
        ao, do = O.nearest_center(c, x)
        assert np.array_equal(ag, ao)
        assert np.array_equal(_bits(dg), _bits(do))
    # sub-vector k-means on distances that cancel to rounding noise: the order
    # of the short dot products (MKL's for 1, 2 or 4 floats) decides the argmin
    for dim in (1, 2, 4):
        _, _, cb, xs, _ = CI.near_tie_encode_input(41, 1, dim, 5000)
        ag, dg = ctx.kmeans_assign(cb, xs, with_dist=True)
        ao, do = O.nearest_center(cb, xs)
        assert np.array_equal(ag, ao), dim
        assert np.array_equal(_bits(dg), _bits(do)), dim


@pytest.mark.parametrize("which", ["l2", "cos"])
def test_server_matches_search_batch(ctx, which, idx_l2, idx_cos):
    # graph-captured buckets (padded batches) give the same bits as the plain
    # batch call, before and after autotuning; uneven batch sizes
    from puck_amd.engine import Server
    ix, q = idx_l2 if which == "l2" else idx_cos
    gi = _gpu_index(ctx, ix)
    n = q.shape[0]
    rng = np.random.default_rng(12)
    topk = rng.choice([1, 10, 100], n).astype(np.uint32)
    scc = rng.integers(1, 24, n).astype(np.uint32)
    Dw, Iw, Rw = gi.search_batch(q, 100, 20, 1100, topk=topk, scc=scc)
    sv = Server(gi, 100, 20, 1100, max_batch=64)
    assert sv.buckets == 7
    for tuned in (False, True):
        if tuned:
            shapes, times = sv.tune(np.repeat(q, 1, axis=0)[:64], reps=1)
            assert shapes.shape == (7,) and np.all((shapes >= 0) & (shapes < 4)) and np.all(times > 0)
        s0 = 0
        for bs in (1, 3, 17, 2, 41):
            e = min(n, s0 + bs)
            D, I, R = sv.search(q[s0:e], topk=topk[s0:e], scc=scc[s0:e])
            assert np.array_equal(R, Rw[s0:e])
            assert np.array_equal(I, Iw[s0:e]) and np.array_equal(_bits(D), _bits(Dw[s0:e]))
            s0 = e
            if s0 >= n:
                break


def test_index_build_reproducible(ctx):
    # two builds of the same data give the same centers, codebook, cells and records
    import torch
    from puck_amd import train
    x = torch.from_numpy(D.unit_norm_mixture(20000, 32, 50, seed=80, means_seed=81)).cuda()
    b1 = train.build(ctx, x, 32, 32, 16, whether_norm=True, train_n=8000, niter=4, rounds=2)
    b2 = train.build(ctx, x, 32, 32, 16, whether_norm=True, train_n=8000, niter=4, rounds=2)
    for k in ("coarse", "fine", "codebook", "cell_assign", "records"):
        assert torch.equal(b1[k], b2[k]), k
    assert torch.isfinite(b1["coarse"]).all() and torch.isfinite(b1["fine"]).all()
