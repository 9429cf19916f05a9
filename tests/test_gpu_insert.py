"""GPU parity of the realtime insert (RealtimeInsertPuckIndex,
realtime_insert_puck_index.cpp:273-541) against the oracle's overlay: ids,
cells, records and stored features of the inserted points bit for bit, the
merged memory layout (each cell: loaded points, then inserted points newest
first), and searches after inserts -- the walk (fast and replay paths, with
exact ties on integer data), the budget scan, the second-stage PQ rerank,
inserts into an empty index and a server whose graphs predate the insert."""
import numpy as np
import pytest

from oracle import oracle as O
from puck_amd import datasets as D
import cpu_index as CI

pytestmark = pytest.mark.gpu
NONE = 0xffffffff


@pytest.fixture(scope="module")
def ctx():
    from puck_amd.engine import Context
    return Context(0)


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _gpu_index(ctx, ix):
    from puck_amd.engine import GpuIndex
    pq = "pq_records" in ix
    return GpuIndex(ctx, ix["coarse"], ix["fine"], ix["codebook"].reshape(-1), ix["cell_assign"], ix["records"],
                    None if pq else ix["raw"], ix["whether_norm"],
                    pq_codebook=ix["pq_codebook"].reshape(-1) if pq else None,
                    pq_records=ix["pq_records"] if pq else None)


def _case(kind):
    if kind == "l2_int":  # integer data: exact distance ties
        x = D.mixture_uint8_grid(8000, 32, 60, seed=121, means_seed=122)
        ix = CI.build_index(x[:6000], 12, 12, 16, whether_norm=False, train_n=6000)
        q = D.mixture_uint8_grid(40, 32, 60, seed=123, means_seed=122)
    elif kind == "cos":
        x = D.unit_norm_mixture(8000, 64, 80, seed=124, means_seed=125)
        x[6000:] *= np.linspace(0.5, 3.0, 2000, dtype=np.float32)[:, None]  # inserts are normalised
        ix = CI.build_index(x[:6000], 16, 16, 32, whether_norm=True, train_n=6000)
        q = D.unit_norm_mixture(40, 64, 80, seed=126, means_seed=125)
    else:  # whether_pq
        x = D.mixture_uint8_grid(8000, 32, 60, seed=127, means_seed=128)
        ix = CI.build_index_pq(x[:6000], 12, 12, 8, 16, whether_norm=False, train_n=6000)
        q = D.mixture_uint8_grid(40, 32, 60, seed=129, means_seed=128)
    return ix, x[6000:], q


@pytest.fixture(scope="module", params=["l2_int", "cos", "pq"])
def inserted(request, ctx):
    ix, extra, q = _case(request.param)
    gi = _gpu_index(ctx, ix)
    oi = CI.oracle_index(ix)
    outs = []
    s = 0
    for n in (1, 7, 300, extra.shape[0] - 308):
        g = gi.insert(extra[s:s + n])
        o = oi.insert(extra[s:s + n])
        outs.append((g, o))
        s += n
    return ix, gi, oi, q, outs


def test_insert_outputs_bitexact(inserted):
    ix, gi, oi, q, outs = inserted
    for g, o in outs:
        assert np.array_equal(g["ids"], o["ids"])
        assert np.array_equal(g["cells"], o["cells"])
        assert np.array_equal(g["records"], o["records"])
        assert np.array_equal(_bits(g["features"]), _bits(o["features"]))
        if "pq_records" in o:
            assert np.array_equal(g["pq_records"], o["pq_records"])
    assert gi.N == ix["raw"].shape[0] + oi.inserted()
    assert gi.version() == len(outs)


def test_merged_layout_matches_oracle_positions(inserted):
    ix, gi, oi, q, outs = inserted
    ex = gi.export()
    N0 = ix["raw"].shape[0]
    total = N0 + oi.inserted()
    m2l = ex["mem2local"]
    assert m2l.shape[0] == total
    base_m2l = oi.export()["mem2local"]
    for tag in range(total):
        want = int(base_m2l[tag]) if tag < N0 else tag
        assert m2l[oi.position(tag)] == want, tag
    # stationary values (including cells opened by an insert) and the layout's cell starts
    oe = oi.export()
    assert np.array_equal(_bits(ex["stationary"]), _bits(oe["stationary"]))
    assert np.array_equal(_bits(ex["min_dist_offset"]), _bits(oe["min_dist_offset"]))
    cnt = np.bincount(np.concatenate([ix["cell_assign"], np.concatenate([o["cells"] for _, o in outs])]),
                      minlength=ix["coarse"].shape[0] * ix["fine"].shape[0])
    assert np.array_equal(ex["cell_start"], np.concatenate([[0], np.cumsum(cnt)]).astype(np.uint32))


@pytest.mark.parametrize("scc,ftk,topk,replay", [(2, 100, 10, False), (5, 400, 20, False), (12, 2000, 50, False),
                                                 (4, 300, 10, True)])
def test_walk_after_insert_bitexact(inserted, scc, ftk, topk, replay):
    ix, gi, oi, q, _ = inserted
    scc = min(scc, ix["coarse"].shape[0])
    Dg, Ig, Rg = gi.search(q, topk, scc, ftk, replay_all=replay)
    hits = 0
    for k in range(q.shape[0]):
        o = oi.search_one(q[k], topk, scc, ftk)
        assert Rg[k] == o["n"], k
        assert np.array_equal(Ig[k, :o["n"]], o["ids"][:o["n"]]), k
        assert np.array_equal(_bits(Dg[k, :o["n"]]), _bits(o["dist"][:o["n"]])), k
        hits += int(np.any(Ig[k, :o["n"]] >= ix["raw"].shape[0]))
    assert hits > 0


def test_budget_scan_after_insert_bitexact(inserted):
    ix, gi, oi, q, _ = inserted
    n = q.shape[0]
    nf = (np.arange(n) % 5 + 1).astype(np.uint32)
    nf[::7] = 0  # the walk in the same batch
    Dg, Ig, Rg = gi.search_batch(q, 10, 4, 300, fine_budget=nf)
    Do, Io, Ro = oi.search_batch_pq(q, 10, 4, 300, fine_budget=nf)
    assert np.array_equal(Rg, Ro)
    for k in range(n):
        m = int(Ro[k])
        assert np.array_equal(Ig[k, :m], Io[k, :m]), k
        assert np.array_equal(_bits(Dg[k, :m]), _bits(Do[k, :m])), k


def test_zero_norm_insert_is_refused(ctx):
    ix, extra, q = _case("cos")
    gi = _gpu_index(ctx, ix)
    oi = CI.oracle_index(ix)
    x = np.stack([extra[0], np.zeros_like(extra[0]), extra[1]]).astype(np.float32)
    g, o = gi.insert(x), oi.insert(x)
    N = ix["raw"].shape[0]
    assert list(g["ids"]) == [N, NONE, N + 1] == list(o["ids"])
    assert np.array_equal(g["cells"], o["cells"]) and np.array_equal(g["records"], o["records"])


def test_inserts_into_an_empty_index(ctx):
    x = D.mixture_uint8_grid(3000, 16, 20, seed=130, means_seed=131)
    ix = CI.build_index(x, 4, 4, 8, whether_norm=False)
    empty = dict(ix, cell_assign=np.zeros(0, np.uint32), records=np.zeros((0, 12), np.uint8),
                 raw=np.zeros((0, 16), np.float32))
    gi = _gpu_index(ctx, empty)
    oi = CI.oracle_index(empty)
    for s0, s1 in ((0, 1), (1, 40), (40, 200)):
        g, o = gi.insert(x[s0:s1]), oi.insert(x[s0:s1])
        assert np.array_equal(g["ids"], o["ids"]) and np.array_equal(g["cells"], o["cells"])
    assert np.array_equal(_bits(gi.export()["stationary"]), _bits(oi.export()["stationary"]))
    Dg, Ig, Rg = gi.search(x[:20], 5, 4, 60)
    for k in range(20):
        r = oi.search_one(x[k], 5, 4, 60)
        assert Rg[k] == r["n"] and np.array_equal(Ig[k, :r["n"]], r["ids"][:r["n"]]), k


def test_server_after_insert_matches_eager(ctx):
    from puck_amd.engine import Server
    ix, extra, q = _case("l2_int")
    gi = _gpu_index(ctx, ix)
    sv = Server(gi, topk_max=10, search_coarse_count=4, filter_topk=300, max_batch=64)
    D0, I0, R0 = sv.search(q)
    gi.insert(extra[:500])
    D1, I1, R1 = sv.search(q)  # the graph predates the insert: captured again
    De, Ie, Re = gi.search(q, 10, 4, 300)
    assert np.array_equal(I1, Ie) and np.array_equal(_bits(D1), _bits(De)) and np.array_equal(R1, Re)
    assert not np.array_equal(I0, I1)
