"""GPU parity of the second-stage PQ (whether_pq = true) against the oracle:
the two-quantizer encode (puck_index.cpp:602-709) bit for bit, and searches of
an index that holds no raw vectors and ranks its filter survivors by the PQ
distance (rank_topN_points, puck_index.cpp:402-427): the walk, the
reference-heap replay and the budget scan."""
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


@pytest.fixture(scope="module", params=["cos", "l2"])
def pq_index(request):
    if request.param == "cos":
        x = D.unit_norm_mixture(12000, 32, 120, seed=91, means_seed=92)
        q = D.unit_norm_mixture(48, 32, 120, seed=93, means_seed=92)
        ix = CI.build_index_pq(x, 24, 24, 16, 32, whether_norm=True, train_n=8000)
    else:
        x = D.mixture_uint8_grid(10000, 64, 150, seed=94, means_seed=95)
        q = D.mixture_uint8_grid(48, 64, 150, seed=96, means_seed=95)
        ix = CI.build_index_pq(x, 20, 20, 16, 64, whether_norm=False, train_n=8000)
    return ix, q


def _gpu_index(ctx, ix):
    from puck_amd.engine import GpuIndex
    return GpuIndex(ctx, ix["coarse"], ix["fine"], ix["codebook"].reshape(-1), ix["cell_assign"], ix["records"],
                    None, ix["whether_norm"], pq_codebook=ix["pq_codebook"].reshape(-1), pq_records=ix["pq_records"])


def test_encode_pq_bitexact(ctx, pq_index):
    ix, _ = pq_index
    nsq, pnsq = ix["nsq"], ix["pq_nsq"]
    x = ix["raw"]
    rec = ctx.encode_pq(ix["coarse"], ix["fine"], ix["codebook"].reshape(-1), nsq, ix["records"],
                        ix["pq_codebook"].reshape(-1), pnsq, x, ix["cell_assign"])
    assert np.array_equal(rec, ix["pq_records"])


def _sorted_pairs(v, i):
    order = np.lexsort((i, v))
    return v[order], i[order]


@pytest.mark.parametrize("scc,ftk,topk,replay", [(3, 200, 10, False), (8, 600, 10, False), (24, 64, 1, False),
                                                 (6, 300, 10, True), (4, 1000, 50, True)])
def test_pq_search_bitexact(ctx, pq_index, scc, ftk, topk, replay):
    ix, q = pq_index
    gi = _gpu_index(ctx, ix)
    oi = CI.oracle_index(ix)
    Dg, Ig, Rg = gi.search(q, topk, scc, ftk, replay_all=replay)
    ft, fm, fn = gi.last_filter(q.shape[0], ftk)
    for k in range(q.shape[0]):
        o = oi.search_one(q[k], topk, scc, ftk)
        assert fn[k] == o["filt_n"], k
        if replay:
            assert np.array_equal(_bits(ft[k, :fn[k]]), _bits(o["filt_val"][:fn[k]])), k
            assert np.array_equal(fm[k, :fn[k]], o["filt_mem"][:fn[k]]), k
        else:
            gv, gid = _sorted_pairs(ft[k, :fn[k]], fm[k, :fn[k]])
            ov, oid = _sorted_pairs(o["filt_val"][:fn[k]], o["filt_mem"][:fn[k]])
            assert np.array_equal(_bits(gv), _bits(ov)) and np.array_equal(gid, oid), k
        assert Rg[k] == o["n"], k
        assert np.array_equal(Ig[k, :o["n"]], o["ids"][:o["n"]]), f"query {k}"
        assert np.array_equal(_bits(Dg[k, :o["n"]]), _bits(o["dist"][:o["n"]])), f"query {k}"


def test_pq_budget_scan_bitexact(ctx, pq_index):
    ix, q = pq_index
    gi = _gpu_index(ctx, ix)
    oi = CI.oracle_index(ix)
    n, F = q.shape[0], ix["fine"].shape[0]
    nf = (np.arange(n) % 7 + 1).astype(np.uint32)
    nf[::9] = 0  # some queries take the walk in the same batch
    Dg, Ig, Rg = gi.search_batch(q, 10, 6, 400, fine_budget=nf)
    Do, Io, Ro = oi.search_batch_pq(q, 10, 6, 400, fine_budget=nf)
    assert np.array_equal(Rg, Ro)
    for k in range(n):
        m = int(Ro[k])
        assert np.array_equal(Ig[k, :m], Io[k, :m]), k
        assert np.array_equal(_bits(Dg[k, :m]), _bits(Do[k, :m])), k


def test_index_needs_raw_vectors_or_pq(ctx, pq_index):
    # an index with neither raw vectors nor a PQ cannot be searched
    from puck_amd.engine import GpuIndex, PuckGpuError
    ix, q = pq_index
    with pytest.raises(ValueError):
        GpuIndex(ctx, ix["coarse"], ix["fine"], ix["codebook"].reshape(-1), ix["cell_assign"], ix["records"], None,
                 ix["whether_norm"])
