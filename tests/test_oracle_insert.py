"""The oracle's realtime-insert overlay (RealtimeInsertPuckIndex,
puck/puck/realtime_insert_puck_index.cpp) against what the mathematics and an
independent construction fix:

* an inserted point's cell is the populated cell nearest to it: the
  assignment score |c|^2/2 - <c,x> + |x|^2/2 - <f,x> + stationary equals
  ||x - c - f||^2 / 2, and empty cells hold the sentinel (checked in float64);
* its records are the build's encoding of the stored point in that cell;
* whether_norm stores the normalised point, and a zero-norm point is refused
  without taking an insert id (single_build, hierarchical_cluster_index.cpp:
  1463-1470);
* a search over the overlay (loaded points, then each cell's inserted points
  newest first, :398-453 and :455-541) returns what a search returns over an
  index built from scratch with the same points, cells and records -- the
  walk only keeps the smallest filter distances of the cells it scans, so the
  order of a cell's points matters for exact ties only (checked on float
  data without ties), and the same holds for the budget scan and the
  second-stage PQ rerank;
* a cell that had no point takes the stationary value sdot(c,f) - sdot(f,f)
  (:326-338), so inserts into an empty index are found by a search;
* the memory positions of the overlay layout are a permutation, cell-major,
  loaded points first, inserted points newest first.
"""
import numpy as np
import pytest

from oracle import oracle as O
from puck_amd import datasets as D
import cpu_index as CI

F64 = np.float64
NONE = 0xffffffff


def _rebuilt(ix, ins):
    """The same points as one index built from scratch (local ids kept)."""
    out = dict(ix)
    out["cell_assign"] = np.concatenate([ix["cell_assign"], ins["cells"]]).astype(np.uint32)
    out["records"] = np.concatenate([ix["records"], ins["records"]])
    out["raw"] = np.concatenate([ix["raw"], ins["features"]]).astype(np.float32)
    if "pq_records" in ix:
        out["pq_records"] = np.concatenate([ix["pq_records"], ins["pq_records"]])
    return out


def _untie(x, seed):
    # integer data has exact distance ties, whose order follows arrival order
    return (x + np.random.default_rng(seed).uniform(0, 1e-2, x.shape)).astype(np.float32)


@pytest.fixture(scope="module")
def l2_case():
    x = _untie(D.mixture_uint8_grid(7000, 32, 60, seed=101, means_seed=102), 1)
    base, extra = x[:5000], x[5000:]
    ix = CI.build_index(base, 12, 12, 16, whether_norm=False, train_n=5000)
    q = D.mixture_uint8_grid(30, 32, 60, seed=103, means_seed=102)
    return ix, extra, q


@pytest.fixture(scope="module")
def cos_case():
    x = D.unit_norm_mixture(6000, 32, 50, seed=104, means_seed=105)
    base, extra = x[:4500], x[4500:]
    ix = CI.build_index(base, 10, 10, 16, whether_norm=True, train_n=4500)
    q = D.unit_norm_mixture(30, 32, 50, seed=106, means_seed=105)
    return ix, extra, q


def _insert_in_chunks(oi, extra, chunks=(1, 7, 300)):
    outs, s = [], 0
    sizes = list(chunks) + [extra.shape[0] - sum(chunks)]
    for n in sizes:
        outs.append(oi.insert(extra[s:s + n]))
        s += n
    cat = {k: np.concatenate([o[k] for o in outs]) for k in outs[0]}
    return cat


def test_inserted_cell_is_the_nearest_populated_cell(l2_case):
    ix, extra, _ = l2_case
    oi = CI.oracle_index(ix)
    ins = oi.insert(extra[:400])
    F = ix["fine"].shape[0]
    cf = (ix["coarse"][:, None, :].astype(F64) + ix["fine"][None, :, :].astype(F64)).reshape(-1, extra.shape[1])
    populated = np.bincount(ix["cell_assign"], minlength=cf.shape[0]) > 0
    assert np.array_equal(ins["ids"], np.arange(ix["raw"].shape[0], ix["raw"].shape[0] + 400))
    for i in range(400):
        d = ((cf - extra[i].astype(F64)) ** 2).sum(1) / 2
        d[~populated] = np.inf
        got = d[ins["cells"][i]]
        assert populated[ins["cells"][i]]
        assert got <= d.min() + 1e-4 * (1.0 + d.min()), i
    assert np.all(populated[ins["cells"] % (cf.shape[0])])


@pytest.mark.parametrize("case", ["l2_case", "cos_case"])
def test_inserted_records_are_the_build_encoding(case, request):
    ix, extra, _ = request.getfixturevalue(case)
    oi = CI.oracle_index(ix)
    ins = _insert_in_chunks(oi, extra[:500], chunks=(3, 97))
    rec = O.encode_filter(ix["coarse"], ix["fine"], ix["codebook"], ix["nsq"], ins["features"], ins["cells"])
    assert np.array_equal(rec, ins["records"])
    if ix["whether_norm"]:
        n = np.sqrt((ins["features"].astype(F64) ** 2).sum(1))
        assert np.allclose(n, 1.0, atol=1e-6)


def test_zero_norm_point_is_refused_without_an_id(cos_case):
    ix, extra, _ = cos_case
    oi = CI.oracle_index(ix)
    x = np.stack([extra[0] * 3.0, np.zeros_like(extra[0]), extra[1] * 0.5]).astype(np.float32)
    ins = oi.insert(x)
    N = ix["raw"].shape[0]
    assert list(ins["ids"]) == [N, NONE, N + 1]
    assert len(ins["cells"]) == 2
    # scaling does not change the stored (normalised) point
    assert np.allclose(ins["features"][0], O.normalize(extra[0]), atol=1e-7)


@pytest.mark.parametrize("case", ["l2_case", "cos_case"])
@pytest.mark.parametrize("scc,ftk,topk", [(2, 100, 10), (5, 400, 20), (12, 3000, 50)])
def test_overlay_walk_equals_rebuilt_index(case, scc, ftk, topk, request):
    ix, extra, q = request.getfixturevalue(case)
    C = ix["coarse"].shape[0]
    scc = min(scc, C)
    oi = CI.oracle_index(ix)
    ins = _insert_in_chunks(oi, extra)
    populated = np.bincount(ix["cell_assign"], minlength=C * ix["fine"].shape[0]) > 0
    assert np.all(populated[ins["cells"]])  # no insert opened an empty cell
    rb = CI.oracle_index(_rebuilt(ix, ins))
    hits = 0
    for k in range(q.shape[0]):
        a = oi.search_one(q[k], topk, scc, ftk)
        b = rb.search_one(q[k], topk, scc, ftk)
        assert a["n"] == b["n"], k
        assert np.array_equal(a["ids"][:a["n"]], b["ids"][:b["n"]]), k
        assert np.array_equal(a["dist"][:a["n"]].view(np.uint32), b["dist"][:b["n"]].view(np.uint32)), k
        # the filter sets agree as (value, local id) pairs
        fa = sorted(zip(a["filt_val"][:a["filt_n"]].tolist(), [oi.local_of_tag(t) for t in a["filt_mem"][:a["filt_n"]]]))
        fb = sorted(zip(b["filt_val"][:b["filt_n"]].tolist(), [rb.local_of_tag(t) for t in b["filt_mem"][:b["filt_n"]]]))
        assert fa == fb, k
        hits += int(np.any(a["ids"][:a["n"]] >= ix["raw"].shape[0]))
    assert hits > 0  # inserted points are returned


@pytest.mark.parametrize("budget", [1, 3])
def test_overlay_budget_scan_equals_rebuilt_index(l2_case, budget):
    ix, extra, q = l2_case
    oi = CI.oracle_index(ix)
    ins = _insert_in_chunks(oi, extra)
    rb = CI.oracle_index(_rebuilt(ix, ins))
    for k in range(q.shape[0]):
        a = oi.search_one(q[k], 10, 4, 300, fine_budget=budget)
        b = rb.search_one(q[k], 10, 4, 300, fine_budget=budget)
        assert a["n"] == b["n"] and np.array_equal(a["ids"][:a["n"]], b["ids"][:b["n"]]), k
        assert np.array_equal(a["dist"][:a["n"]].view(np.uint32), b["dist"][:b["n"]].view(np.uint32)), k
        assert a["stats"]["points_offered"] == b["stats"]["points_offered"], k


def test_overlay_pq_rerank_equals_rebuilt_index():
    x = _untie(D.mixture_uint8_grid(6000, 32, 50, seed=107, means_seed=108), 2)
    ix = CI.build_index_pq(x[:4500], 10, 10, 8, 16, whether_norm=False, train_n=4500)
    q = D.mixture_uint8_grid(20, 32, 50, seed=109, means_seed=108)
    oi = CI.oracle_index(ix)
    ins = _insert_in_chunks(oi, x[4500:], chunks=(5, 200))
    frec, prec = O.encode_puck(ix["coarse"], ix["fine"], ix["codebook"], 8, ix["pq_codebook"], 16, ins["features"],
                               ins["cells"])
    assert np.array_equal(frec, ins["records"]) and np.array_equal(prec, ins["pq_records"])
    rb = CI.oracle_index(_rebuilt(ix, ins))
    for k in range(q.shape[0]):
        for kw in (dict(), dict(fine_budget=2)):
            a = oi.search_one(q[k], 10, 4, 250, **kw)
            b = rb.search_one(q[k], 10, 4, 250, **kw)
            assert a["n"] == b["n"] and np.array_equal(a["ids"][:a["n"]], b["ids"][:b["n"]]), (k, kw)
            assert np.array_equal(a["dist"][:a["n"]].view(np.uint32), b["dist"][:b["n"]].view(np.uint32)), (k, kw)


def test_inserts_into_an_empty_index():
    # every cell holds the sentinel: the first insert takes the first cell
    # visited, which then gets sdot(c,f) - sdot(f,f) and draws the later ones
    x = D.mixture_uint8_grid(3000, 16, 20, seed=110, means_seed=111)
    ix = CI.build_index(x, 4, 4, 8, whether_norm=False)
    empty = dict(ix, cell_assign=np.zeros(0, np.uint32), records=np.zeros((0, 12), np.uint8),
                 raw=np.zeros((0, 16), np.float32))
    oi = CI.oracle_index(empty)
    st0 = oi.export()["stationary"].copy()
    assert np.all(st0 == O.empty_cell_stationary())
    ins = oi.insert(x[:50])
    cell = int(ins["cells"][0])
    assert np.all(ins["cells"] == cell)
    st = oi.export()["stationary"]
    c, f = ix["coarse"][cell // 4], ix["fine"][cell % 4]
    want = np.float32(O.dot(c, f)) - np.float32(O.dot(f, f))
    assert st[cell] == want
    assert np.sum(st != st0) == 1
    # min_dist_offset keeps its loaded value
    assert np.all(oi.export()["min_dist_offset"] == O.empty_cell_stationary())
    r = oi.search_one(x[3], 5, 4, 50)
    assert r["n"] == 5 and 3 in r["ids"][:5].tolist()


def test_overlay_positions_are_cell_major(l2_case):
    ix, extra, _ = l2_case
    oi = CI.oracle_index(ix)
    ins = _insert_in_chunks(oi, extra[:600], chunks=(1, 50))
    N = ix["raw"].shape[0]
    n = oi.inserted()
    pos = np.array([oi.position(t) for t in range(N + n)])
    assert np.array_equal(np.sort(pos), np.arange(N + n))
    # cell of every position, in position order: non-decreasing
    ex = oi.export()
    cs = ex["cell_start"]
    base_cell = np.searchsorted(cs, np.arange(N), side="right") - 1
    cell = np.concatenate([base_cell, ins["cells"]])
    by_pos = np.empty(N + n, np.int64)
    by_pos[pos] = cell
    assert np.all(np.diff(by_pos) >= 0)
    # within a cell: loaded points first, then inserted points newest first
    tag_at = np.empty(N + n, np.int64)
    tag_at[pos] = np.arange(N + n)
    for c in np.unique(ins["cells"])[:40]:
        tags = tag_at[by_pos == c]
        nb = int((tags < N).sum())
        assert np.all(tags[:nb] < N) and np.all(np.diff(tags[:nb]) > 0)
        assert np.all(np.diff(tags[nb:]) < 0)
