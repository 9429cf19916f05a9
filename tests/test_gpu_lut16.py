"""GPU tests of the FP16 filter table (PUCKGPU_LUT_FP16): the engine's search
with the table rounded to FP16 returns the oracle's FP16-table search bit for
bit -- the reference walk (fast path and the reference-heap replay), the
split budget scan and the per-query kernel it leaves the large queries to --
and the returned distances stay the exact FP32 rerank distances."""
import numpy as np
import pytest

from oracle import oracle as O
from puck_amd import datasets as D
import cpu_index as CI

pytestmark = pytest.mark.gpu

LUT_FP16 = 16


@pytest.fixture(scope="module")
def ctx():
    from puck_amd.engine import Context
    return Context(0)


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _sorted_pairs(v, i):
    o = np.lexsort((i, v))
    return v[o], i[o]


@pytest.fixture(scope="module")
def idx_l2():
    x = D.mixture_uint8_grid(12000, 128, 200, seed=90, means_seed=91)
    q = D.mixture_uint8_grid(48, 128, 200, seed=92, means_seed=91)
    return CI.build_index(x, 40, 40, 32, whether_norm=False, train_n=6000), q


@pytest.fixture(scope="module")
def idx_cos():
    x = D.unit_norm_mixture(15000, 96, 300, seed=93, means_seed=94)
    q = D.unit_norm_mixture(48, 96, 300, seed=95, means_seed=94)
    return CI.build_index(x, 40, 40, 48, whether_norm=True, train_n=6000), q


@pytest.fixture(scope="module")
def idx_dense():
    # few large cells: large budgets overflow the split scan's 8192-key buffer
    x = D.unit_norm_mixture(40000, 32, 64, seed=96, means_seed=97)
    q = D.unit_norm_mixture(48, 32, 64, seed=98, means_seed=97)
    return CI.build_index(x, 4, 8, 16, whether_norm=True, train_n=20000), q


def _gpu_index(ctx, ix):
    from puck_amd.engine import GpuIndex
    return GpuIndex(ctx, ix["coarse"], ix["fine"], ix["codebook"].reshape(-1), ix["cell_assign"],
                    ix["records"], ix["raw"], ix["whether_norm"])


def _oracle16(ix):
    oi = CI.oracle_index(ix)
    oi.set_lut16(True)
    return oi


@pytest.mark.parametrize("which", ["l2", "cos"])
@pytest.mark.parametrize("replay", [False, True])
def test_lut16_walk_bitexact(ctx, which, replay, idx_l2, idx_cos):
    ix, q = idx_l2 if which == "l2" else idx_cos
    gi = _gpu_index(ctx, ix)
    oi = _oracle16(ix)
    for scc, ftk, topk in [(6, 300, 10), (2, 1000, 50)]:
        Dg, Ig, Rg = gi.search(q, topk, scc, ftk, replay_all=replay, precision=LUT_FP16)
        ft, fm, fn = gi.last_filter(q.shape[0], ftk)
        for k in range(q.shape[0]):
            o = oi.search_one(q[k], topk, scc, ftk)
            assert fn[k] == o["filt_n"], k
            if replay:  # the reference heap's array order
                assert np.array_equal(_bits(ft[k, :fn[k]]), _bits(o["filt_val"][:fn[k]])), k
                assert np.array_equal(fm[k, :fn[k]], o["filt_mem"][:fn[k]]), k
            else:
                gv, gid = _sorted_pairs(ft[k, :fn[k]], fm[k, :fn[k]])
                ov, oid = _sorted_pairs(o["filt_val"][:fn[k]], o["filt_mem"][:fn[k]])
                assert np.array_equal(_bits(gv), _bits(ov)), k
                assert np.array_equal(gid, oid), k
            assert Rg[k] == o["n"], k
            assert np.array_equal(Ig[k, :o["n"]], o["ids"][:o["n"]]), k
            assert np.array_equal(_bits(Dg[k, :o["n"]]), _bits(o["dist"][:o["n"]])), k


@pytest.mark.parametrize("which", ["l2", "cos"])
def test_lut16_budget_scan_bitexact(ctx, which, idx_l2, idx_cos):
    ix, q = idx_l2 if which == "l2" else idx_cos
    gi = _gpu_index(ctx, ix)
    oi = _oracle16(ix)
    n = q.shape[0]
    nf = (np.arange(n, dtype=np.uint32) % 9) + 1
    Dg, Ig, Rg = gi.search_batch(q, 10, 6, 500, fine_budget=nf, precision=LUT_FP16)
    ft, fm, fn = gi.last_filter(n, 500)
    Do, Io, Ro = oi.search_batch_pq(q, 10, 6, 500, fine_budget=nf)
    assert np.array_equal(Rg, Ro)
    for k in range(n):
        m = int(Ro[k])
        assert np.array_equal(Ig[k, :m], Io[k, :m]), k
        assert np.array_equal(_bits(Dg[k, :m]), _bits(Do[k, :m])), k
        o = oi.search_one(q[k], 10, 6, 500, fine_budget=int(nf[k]))
        gv, gid = _sorted_pairs(ft[k, :fn[k]], fm[k, :fn[k]])
        assert np.array_equal(_bits(gv), _bits(o["filt_val"][:fn[k]])), k
        assert np.array_equal(gid, o["filt_mem"][:fn[k]]), k
    # the flag took effect: some filter distances differ from the FP32 table's
    gi.search_batch(q, 10, 6, 500, fine_budget=nf)
    ft32, fm32, fn32 = gi.last_filter(n, 500)
    assert any(not np.array_equal(_bits(np.sort(ft32[k, :fn32[k]])), _bits(np.sort(ft[k, :fn[k]]))) for k in range(n))


@pytest.mark.parametrize("ftk", [64, 1000])
def test_lut16_split_and_overflow_bitexact(ctx, idx_dense, ftk):
    ix, q = idx_dense
    gi = _gpu_index(ctx, ix)
    oi = _oracle16(ix)
    n = q.shape[0]
    nf = np.array([1, 2, 7, 6, 3, 5, 4, 7] * (n // 8), np.uint32)
    scc = np.array([4, 1, 4, 2, 3, 4, 1, 4] * (n // 8), np.uint32)
    topk = np.full(n, 10, np.uint32)
    Dg, Ig, Rg = gi.search_batch(q, 10, 4, ftk, topk=topk, scc=scc, fine_budget=nf, precision=LUT_FP16)
    ft, fm, fn = gi.last_filter(n, ftk)
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
    assert 0 < big < n  # both kernels ran


def test_lut16_with_bf16_centers_exact_distances(ctx, idx_cos):
    # FP16 table with BF16 center distances: returned distances are the
    # exact FP32 rerank of the returned ids, sorted
    ix, q = idx_cos
    gi = _gpu_index(ctx, ix)
    Dp, Ip, Rp = gi.search(q, 10, 8, 500, precision=1 | LUT_FP16)
    for k in range(q.shape[0]):
        qn = O.normalize(q[k])
        for j in range(int(Rp[k])):
            assert _bits(Dp[k, j]) == _bits(np.float32(O.rerank_dist(ix["raw"][Ip[k, j]], qn)))
        assert np.all(np.diff(Dp[k, :Rp[k]]) >= 0)


def test_lut16_refused_where_unsupported(ctx, idx_cos):
    from puck_amd.engine import PuckGpuError
    ix, q = idx_cos
    gi = _gpu_index(ctx, ix)
    with pytest.raises(PuckGpuError):
        gi.search(q, 10, 8, 500, precision=32)
    with pytest.raises(PuckGpuError):
        gi.search(q, 10, 8, 500, precision=3 | LUT_FP16)
