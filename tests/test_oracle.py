"""CPU tests of the oracle (the CPU restatement of PuckIndex::search): each
function is pinned by a known answer derived from the reference's algorithm
text, a closed form, brute force, or the reference sources themselves built
in oracle/_ref (skipped when the reference checkout is absent)."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from puck_amd import datasets as D
from puck_amd import metrics as PM
import cpu_index as CI

F32 = np.float32


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


# ---------------------------------------------------------------- MaxHeap
def test_heap_known_answer():
    # Hand-traced from max_heap.cpp:52-100: inserts fill from slot
    # _default_point_cnt downwards, replacements sift down from the root.
    n, v, t = O.heap_run(3, [3, 1, 2], reorder=False)
    assert n == 3 and v.tolist() == [3, 1, 2] and t.tolist() == [0, 1, 2]
    n, v, t = O.heap_run(3, [3, 1, 2, 0.5], reorder=False)
    assert v.tolist() == [2, 1, 0.5] and t.tolist() == [2, 1, 3]
    n, v, t = O.heap_run(3, [3, 1, 2, 0.5], reorder=True)
    assert v.tolist() == [0.5, 1, 2] and t.tolist() == [3, 1, 2]


def test_heap_is_k_smallest():
    rng = np.random.default_rng(0)
    for cap in (1, 2, 7, 64, 500):
        vals = rng.standard_normal(1000).astype(F32)
        n, v, t = O.heap_run(cap, vals)
        assert n == cap
        assert np.array_equal(v, np.sort(vals)[:cap])
        assert np.array_equal(vals[t], v)


def test_heap_not_full_and_sentinel():
    n, v, t = O.heap_run(5, [4.0, -1.0])
    assert n == 2
    assert v[:2].tolist() == [-1.0, 4.0]
    assert np.all(bits(v[2:]) == 0x7F7F7F7F)


def test_heap_ties_first_come_at_boundary():
    # once full, a value equal to the top is rejected (puck max_heap.cpp:56)
    n, v, t = O.heap_run(2, [1.0, 2.0, 2.0, 2.0], tags=[10, 11, 12, 13])
    assert v.tolist() == [1.0, 2.0] and t.tolist() == [10, 11]


# ---------------------------------------------------------------- LUT
def _lut_closed_form(q, cb, nsq, lsq):
    out = np.zeros(nsq * 256, F32)
    for m in range(nsq):
        for k in range(256):
            dd = (q[m * lsq:(m + 1) * lsq] - cb[(m * 256 + k) * lsq:(m * 256 + k + 1) * lsq]).astype(F32)
            p = (dd * dd).astype(F32)
            if lsq == 1:
                out[m * 256 + k] = p[0]
            elif lsq == 2:
                out[m * 256 + k] = F32(p[0] + p[1])
            else:
                out[m * 256 + k] = F32(F32(p[0] + p[1]) + F32(p[2] + p[3]))
    return out


@pytest.mark.parametrize("lsq", [1, 2, 4])
def test_lut_closed_form(lsq):
    rng = np.random.default_rng(lsq)
    nsq = 6
    q = (rng.standard_normal(nsq * lsq) * 40).astype(F32)
    cb = (rng.standard_normal(nsq * 256 * lsq) * 40).astype(F32)
    assert np.array_equal(bits(O.lut(q, cb, nsq, lsq)), bits(_lut_closed_form(q, cb, nsq, lsq)))


def test_lut_order_is_discriminating():
    # the pairwise order of D4 differs from a left-to-right sum on some inputs
    rng = np.random.default_rng(5)
    diff = 0
    for _ in range(200):
        p = (rng.standard_normal(4) * 1000).astype(F32) ** 2
        a = F32(F32(p[0] + p[1]) + F32(p[2] + p[3]))
        b = F32(F32(F32(p[0] + p[1]) + p[2]) + p[3])
        diff += int(a != b)
    assert diff > 0


@pytest.mark.parametrize("lsq", [1, 2, 4])
def test_lut_matches_reference_build(lsq):
    R = O.ref_lib()
    if R is None:
        pytest.skip("reference checkout not present: oracle/_ref not built")
    rng = np.random.default_rng(10 + lsq)
    nsq = 8
    for _ in range(20):
        q = (rng.standard_normal(nsq * lsq) * rng.uniform(0.1, 100)).astype(F32)
        cb = (rng.standard_normal(nsq * 256 * lsq) * rng.uniform(0.1, 100)).astype(F32)
        ref = np.zeros(nsq * 256, F32)
        R.ref_dist_table(q, cb, nsq, lsq, ref)
        assert np.array_equal(bits(O.lut(q, cb, nsq, lsq)), bits(ref))


# ---------------------------------------------------------------- lookup
def _lookup_closed_form(codes, lut, nsq):
    s = [F32(0)] * 4
    for m in range(nsq):
        s[m % 4] = F32(s[m % 4] + lut[m * 256 + int(codes[m])])
    return F32(F32(s[0] + s[1]) + F32(s[2] + s[3]))


@pytest.mark.parametrize("nsq", [1, 3, 4, 8, 12, 32, 48, 50])
def test_lookup_four_lane_order(nsq):
    rng = np.random.default_rng(nsq)
    lut = (rng.standard_normal(nsq * 256) ** 2 * 100).astype(F32)
    for _ in range(50):
        codes = rng.integers(0, 256, nsq).astype(np.uint8)
        assert bits(O.lookup(codes, lut, nsq)) == bits(_lookup_closed_form(codes, lut, nsq))


# ---------------------------------------------------------------- rerank
@pytest.mark.parametrize("dim", [96, 128, 100, 37, 16, 5])
def test_rerank_distance(dim):
    rng = np.random.default_rng(dim)
    R = O.ref_lib()
    for _ in range(200):
        x = (rng.standard_normal(dim) * 30).astype(F32)
        q = (rng.standard_normal(dim) * 30).astype(F32)
        d = O.rerank_dist(x, q)
        exact = float(((x.astype(np.float64) - q) ** 2).sum())
        assert abs(d - exact) <= 1e-5 * exact
        if R is not None:
            assert bits(F32(d)) == bits(F32(R.ref_rerank_dist(x, q, dim)))


# ---------------------------------------------------------------- dot / normalise
def test_dot_exact_on_integers():
    rng = np.random.default_rng(1)
    a = rng.integers(-50, 50, 128).astype(F32)
    b = rng.integers(-50, 50, 128).astype(F32)
    assert O.dot(a, b) == float(np.dot(a.astype(np.int64), b.astype(np.int64)))


def test_normalize():
    q = np.array([3.0, 4.0], F32)
    out = O.normalize(q)
    alpha = F32(1.0 / np.float64(np.sqrt(F32(25.0))))
    assert np.array_equal(bits(out), bits(q * alpha))
    assert O.normalize(np.zeros(4, F32)) is None


# ---------------------------------------------------------------- index + search
@pytest.fixture(scope="module")
def small_l2():
    x = D.mixture_uint8_grid(3000, 32, 40, seed=1, means_seed=2)
    q = D.mixture_uint8_grid(20, 32, 40, seed=3, means_seed=2)
    ix = CI.build_index(x, 8, 8, 8, whether_norm=False, train_n=3000)
    return ix, q


def test_index_derivation(small_l2):
    ix, _ = small_l2
    e = CI.oracle_index(ix).export()
    C, F = ix["coarse"].shape[0], ix["fine"].shape[0]
    counts = np.bincount(ix["cell_assign"], minlength=C * F)
    assert np.array_equal(np.diff(e["cell_start"].astype(np.int64)), counts)
    m2l = e["mem2local"]
    off = ix["records"][:, :4].copy().view(F32).reshape(-1)
    key = list(zip(ix["cell_assign"][m2l], off[m2l], m2l))
    assert key == sorted(key)
    empty = counts == 0
    assert np.all(e["stationary"][empty] == np.sqrt(F32(np.finfo(F32).max)))
    st = e["stationary"].reshape(C, F)
    assert np.array_equal(e["min_dist_offset"], st.min(1))


def test_filter_distance_identity(small_l2):
    # 2*cell_dist + offset + LUT = |q - (c + f + r)|^2 (closed form)
    ix, q = small_l2
    oi = CI.oracle_index(ix)
    C, F = ix["coarse"].shape[0], ix["fine"].shape[0]
    e = oi.export()
    m2l = e["mem2local"]
    nsq, d = ix["nsq"], ix["coarse"].shape[1]
    lsq = d // nsq
    cb = ix["codebook"].reshape(nsq, 256, lsq)
    for k in range(5):
        o = oi.search_one(q[k], 10, C, 200)
        for v, mem in zip(o["filt_val"][:o["filt_n"]], o["filt_mem"][:o["filt_n"]]):
            loc = m2l[mem]
            cell = ix["cell_assign"][loc]
            codes = ix["records"][loc, 4:]
            recon = ix["coarse"][cell // F] + ix["fine"][cell % F] + np.concatenate(
                [cb[m, codes[m]] for m in range(nsq)])
            want = ((q[k].astype(np.float64) - recon) ** 2).sum()
            assert abs(v - want) <= 1e-4 * max(1.0, abs(want))


def test_search_full_scan_is_brute_force(small_l2):
    # scc = C and filter_topk >= N: the heap never fills, the pivot stays at
    # the sentinel, every cell is scanned and the rerank is exhaustive.
    ix, q = small_l2
    oi = CI.oracle_index(ix)
    N = ix["raw"].shape[0]
    gt, gd = O.exact_topk(ix["raw"], q, 10)
    for k in range(q.shape[0]):
        o = oi.search_one(q[k], 10, 1000, N)
        exact = ((ix["raw"][o["ids"]].astype(np.float64) - q[k]) ** 2).sum(1)
        assert np.allclose(exact, gd[k], rtol=1e-6, atol=1e-3)
        assert np.allclose(o["dist"], gd[k], rtol=1e-4)


def test_search_results_sorted_and_exact(small_l2):
    ix, q = small_l2
    oi = CI.oracle_index(ix)
    for k in range(q.shape[0]):
        o = oi.search_one(q[k], 10, 3, 64)
        assert o["n"] == 10
        assert np.all(np.diff(o["dist"]) >= 0)
        exact = ((ix["raw"][o["ids"]].astype(np.float64) - q[k]) ** 2).sum(1)
        assert np.allclose(o["dist"], exact, rtol=1e-4)


def test_search_error_codes(small_l2):
    ix, q = small_l2
    oi = CI.oracle_index(ix)
    assert oi.search_one(q[0], 20, 3, 64, max_topk=10)["rc"] == -1


def test_search_monotone_in_scc(small_l2):
    ix, q = small_l2
    oi = CI.oracle_index(ix)
    gt, gd = O.exact_topk(ix["raw"], q, 10)
    rec = []
    for scc in (1, 2, 4, 8):
        _, I, _, _ = oi.search_batch(q, 10, scc, 100)
        rec.append(O.recall_without_ties(gt, I.astype(np.int64), 10)[0])
    assert rec[-1] >= rec[0]


# ---------------------------------------------------------------- build path
def test_assign_cells_is_argmin():
    rng = np.random.default_rng(4)
    x = (rng.standard_normal((500, 16)) * 5).astype(F32)
    coarse = (rng.standard_normal((6, 16)) * 5).astype(F32)
    fine = (rng.standard_normal((7, 16))).astype(F32)
    norms = (fine.astype(np.float64) ** 2).sum(1)
    fine = fine[np.argsort(-norms, kind="stable")]
    cell, dist = O.assign_cells(coarse, fine, x)
    full = ((x[:, None, None, :].astype(np.float64) - coarse[None, :, None, :] - fine[None, None, :, :]) ** 2).sum(-1)
    best = full.reshape(500, -1).min(1)
    got = full.reshape(500, -1)[np.arange(500), cell]
    assert np.allclose(got, best, rtol=1e-5, atol=1e-4)
    assert np.allclose(dist, best / 2, rtol=1e-4, atol=1e-3)


def test_nearest_center_is_argmin():
    rng = np.random.default_rng(6)
    x = (rng.standard_normal((400, 4)) * 3).astype(F32)
    c = (rng.standard_normal((256, 4)) * 3).astype(F32)
    a, dis = O.nearest_center(c, x)
    full = ((x[:, None, :].astype(np.float64) - c[None]) ** 2).sum(-1)
    assert np.allclose(full[np.arange(400), a], full.min(1), rtol=1e-5, atol=1e-5)


def test_encode_filter_reconstruction(small_l2):
    ix, _ = small_l2
    F = ix["fine"].shape[0]
    nsq, d = ix["nsq"], ix["coarse"].shape[1]
    lsq = d // nsq
    cb = ix["codebook"].reshape(nsq, 256, lsq)
    rec = ix["records"][:200]
    for i in range(200):
        cell = ix["cell_assign"][i]
        v = ix["coarse"][cell // F].astype(np.float64) + ix["fine"][cell % F]
        dec = np.concatenate([cb[m, rec[i, 4 + m]] for m in range(nsq)])
        off = rec[i, :4].copy().view(F32)[0]
        assert abs(off - 2 * np.dot(dec, v)) <= 1e-4 * max(1.0, abs(off))
        r = ix["raw"][i] - ix["coarse"][cell // F] - ix["fine"][cell % F]
        for m in range(nsq):
            dd = ((cb[m].astype(np.float64) - r[m * lsq:(m + 1) * lsq]) ** 2).sum(1)
            assert dd[rec[i, 4 + m]] <= dd.min() + 1e-4


# ---------------------------------------------------------------- recall KATs
with open(os.path.join(os.path.dirname(__file__), "golden", "recall_kats.json")) as _f:
    KATS = json.load(_f)


@pytest.mark.parametrize("kat", KATS, ids=[k["name"] for k in KATS])
def test_recall_known_answers(kat):
    ti = [np.array(r) for r in kat["true_ids"]]
    td = [np.array(r, dtype=np.float64) for r in kat["true_dists"]]
    ri = [np.array(r) for r in kat["run_ids"]]
    c = kat["count"]
    r_ties, _ = PM.recall_with_ties(ti, td, ri, c)
    assert r_ties == pytest.approx(kat["expected_with_ties"], abs=1e-12)
    r_ties_o, _, _ = O.recall_with_ties(ti, td, ri, c)
    assert r_ties_o == pytest.approx(kat["expected_with_ties"], abs=1e-12)
    r_plain, _ = O.recall_without_ties(ti, ri, c)
    assert r_plain == pytest.approx(kat["expected_no_ties"], abs=1e-12)


def _f2ord(v):
    u = np.ascontiguousarray(v, np.float32).view(np.uint32)
    return np.where(u & 0x80000000, ~u, u | 0x80000000).astype(np.uint32)


def test_fine_budget_full_scan_is_restricted_brute_force(small_l2):
    # scc = C and filter_topk >= N with fine budget b: nothing prunes, so the
    # candidates are exactly the points of the cells whose cell distance
    # (cd + stationary) + fine distance is among the b smallest of their coarse
    # list, and the result is their exact top-k.  A budget of F (or 0) is the
    # reference's walk.
    ix, q = small_l2
    oi = CI.oracle_index(ix)
    ex = oi.export()
    raw, coarse, fine = ix["raw"], ix["coarse"], ix["fine"]
    N, C, F = raw.shape[0], coarse.shape[0], fine.shape[0]
    st = ex["stationary"].reshape(C, F)
    for k in range(8):
        b = [1, 2, 3, 5][k % 4]
        o = oi.search_one(q[k], 10, 1000, N, fine_budget=b)
        cd = np.array([ex["coarse_norms"][c] - O.dot(coarse[c], q[k]) for c in range(C)], np.float32)
        fd = np.array([-O.dot(fine[f], q[k]) for f in range(F)], np.float32)
        keys = _f2ord((cd[:, None] + st) + fd[None, :])          # float32 adds, reference order
        kcut = np.sort(keys, axis=1)[:, b - 1]
        ok = keys <= kcut[:, None]
        cand = np.nonzero(ok.reshape(-1)[ix["cell_assign"]])[0]
        assert o["filt_n"] == cand.size
        got = np.sort(o["ids"][:o["n"]])
        d = ((raw[cand].astype(np.float64) - q[k]) ** 2).sum(1)
        want = np.sort(cand[np.argsort(d, kind="stable")[:min(10, cand.size)]])
        assert np.array_equal(got, want)
        full = oi.search_one(q[k], 10, 4, 200)
        for bb in (0, F, F + 5):
            o2 = oi.search_one(q[k], 10, 4, 200, fine_budget=bb)
            assert np.array_equal(o2["ids"], full["ids"]) and np.array_equal(o2["dist"], full["dist"])


def test_per_query_batch_matches_single_calls(small_l2):
    ix, q = small_l2
    oi = CI.oracle_index(ix)
    n = q.shape[0]
    rng = np.random.default_rng(7)
    topk = rng.choice([1, 5, 10], n).astype(np.uint32)
    topk[3] = 0
    topk[5] = 11  # > max_topk: that query fails
    scc = rng.integers(1, 6, n).astype(np.uint32)
    nf = rng.integers(0, ix["fine"].shape[0] + 3, n).astype(np.uint32)
    D, I, R = oi.search_batch_pq(q, 10, 4, 64, topk=topk, scc_arr=scc, fine_budget=nf)
    sent = np.frombuffer(np.uint32(0x7f7f7f7f).tobytes(), np.float32)[0]
    for i in range(n):
        if topk[i] == 0 or topk[i] > 10:
            assert R[i] == 0 and np.all(I[i] == 0xffffffff) and np.all(D[i] == sent)
            continue
        o = oi.search_one(q[i], int(topk[i]), int(min(scc[i], 4)), 64, max_topk=10, fine_budget=int(nf[i]))
        k = int(topk[i])
        assert R[i] == o["n"]
        assert np.array_equal(D[i, :o["n"]].view(np.uint32), o["dist"][:o["n"]].view(np.uint32))
        assert np.array_equal(I[i, :o["n"]], o["ids"][:o["n"]])
        assert np.all(I[i, o["n"]:] == 0xffffffff) and np.all(D[i, k:] == sent)


# ------------------------------------------------- MKL's order for short sdot
MKL_PATH = "/opt/conda/lib/libmkl_rt.so"
needs_mkl = pytest.mark.skipif(not os.path.exists(MKL_PATH), reason="MKL runtime not installed")


def _mkl_sdot():
    import ctypes
    mkl = ctypes.CDLL(MKL_PATH)
    # sequential layer: no second OpenMP runtime beside the oracle's libgomp
    mkl.MKL_Set_Threading_Layer.argtypes = [ctypes.c_int]
    mkl.MKL_Set_Threading_Layer(1)
    mkl.cblas_sdot.restype = ctypes.c_float
    mkl.cblas_sdot.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]

    def sdot(a, b):
        a, b = np.ascontiguousarray(a, F32), np.ascontiguousarray(b, F32)
        return F32(mkl.cblas_sdot(a.size, a.ctypes.data, 1, b.ctypes.data, 1))
    return sdot


def test_sdot_short_closed_form():
    # rounded products, then p0 + p1 / (p0 + p1) + (p2 + p3); an FMA chain differs on these inputs
    rng = np.random.default_rng(11)
    diff = 0
    for n in (1, 2, 4):
        for _ in range(500):
            a = (rng.standard_normal(n) * 10 ** rng.uniform(-2, 2, n)).astype(F32)
            b = (rng.standard_normal(n) * 10 ** rng.uniform(-2, 2, n)).astype(F32)
            p = a * b
            want = p[0] if n == 1 else (p[0] + p[1] if n == 2 else (p[0] + p[1]) + (p[2] + p[3]))
            got = O.sdot_short(a, b)
            assert got.view(np.uint32) == F32(want).view(np.uint32), (a, b)
            diff += got != F32(O.dot(a, b))
    assert diff > 0


@needs_mkl
def test_sdot_short_matches_mkl():
    # the order was read off MKL's results: every sample must agree bit for bit
    sdot = _mkl_sdot()
    rng = np.random.default_rng(12)
    for n in (1, 2, 4):
        for _ in range(3000):
            a = (rng.standard_normal(n) * 10 ** rng.uniform(-3, 3, n)).astype(F32)
            b = (rng.standard_normal(n) * 10 ** rng.uniform(-3, 3, n)).astype(F32)
            assert O.sdot_short(a, b).view(np.uint32) == sdot(a, b).view(np.uint32), (n, a, b)


@needs_mkl
@pytest.mark.parametrize("lsq,nsq", [(1, 32), (2, 16), (4, 8)])
def test_encode_codes_match_nearest_center_with_mkl(lsq, nsq):
    # the filter-PQ codes the oracle writes are the ones nearest_center
    # (kmeans.cpp:221-258) picks with MKL's own cblas_sdot: float pn + cn,
    # minus 2.0 * sdot in double, strict '<' from FLT_MAX.  The input makes
    # the distances cancel to rounding noise, so the order of the short dot
    # products decides the codes (an FMA chain picks other codes for lsq 2, 4).
    sdot = _mkl_sdot()
    coarse, fine, cb, x, cell = CI.near_tie_encode_input(31, nsq, lsq, 12)
    rec = O.encode_filter(coarse, fine, cb, nsq, x, cell)
    cbm = cb.reshape(nsq, 256, lsq)

    def nearest(dotf, sub, m):
        pn = F32(dotf(sub, sub))
        best, arg = F32(np.finfo(F32).max), 0
        for k in range(256):
            cn = F32(dotf(cbm[m, k], cbm[m, k]))
            dist = F32(float(F32(pn + cn)) - 2.0 * float(F32(dotf(sub, cbm[m, k]))))
            if dist < best:
                best, arg = dist, k
        return arg

    fma_differs = 0
    for i in range(x.shape[0]):
        for m in range(nsq):
            sub = x[i, m * lsq:(m + 1) * lsq]
            want = nearest(sdot, sub, m)
            assert rec[i, 4 + m] == want, (i, m)
            fma_differs += int(nearest(O.dot, sub, m) != want)
    assert rec[:, :4].view(F32).ravel().tolist() == [0.0] * x.shape[0]  # offsets against zero centers
    assert fma_differs > 0 if lsq > 1 else fma_differs == 0


@needs_mkl
@pytest.mark.parametrize("dim", [1, 2, 4])
def test_nearest_center_short_dims_match_mkl(dim):
    # the codebook k-means' assignment (nearest_center on sub-vectors) equals
    # nearest_center evaluated with MKL's cblas_sdot, on points and centroids
    # whose distances cancel to rounding noise
    sdot = _mkl_sdot()
    _, _, cb, x, _ = CI.near_tie_encode_input(41, 1, dim, 12)
    a, dis = O.nearest_center(cb, x)
    cn = [sdot(c, c) for c in cb]
    fma_differs = 0
    for i in range(x.shape[0]):
        for dotf, is_mkl in ((sdot, True), (O.dot, False)):
            pn = F32(dotf(x[i], x[i]))
            best, arg = F32(np.finfo(F32).max), 0
            for k in range(cb.shape[0]):
                c_n = cn[k] if is_mkl else F32(O.dot(cb[k], cb[k]))
                dist = F32(float(F32(pn + c_n)) - 2.0 * float(F32(dotf(x[i], cb[k]))))
                if dist < best:
                    best, arg = dist, k
            if is_mkl:
                assert a[i] == arg and dis[i].view(np.uint32) == best.view(np.uint32), i
            else:
                fma_differs += int(arg != a[i])
    assert fma_differs > 0 if dim > 1 else fma_differs == 0


def test_oracle_clean_under_sanitizers(tmp_path):
    # the checker itself: walk, budget scan, batch, heap, assignment, encoding
    # and nearest_center under AddressSanitizer / UBSan / LeakSanitizer
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / "oracle_sanitize")
    cc = subprocess.run(["gcc", "-O1", "-g", "-fopenmp", "-ffp-contract=off", "-fsanitize=address,undefined",
                         "-fno-omit-frame-pointer", "-o", exe, os.path.join(root, "tests", "cpp", "oracle_sanitize_main.c"),
                         os.path.join(root, "oracle", "puck_oracle.c"), "-lm"], capture_output=True, text=True,
                        timeout=300)
    if cc.returncode != 0:
        pytest.skip("gcc sanitizer runtime unavailable: " + cc.stderr[-200:])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1", OMP_NUM_THREADS="2"))
    assert r.returncode == 0 and "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-2000:]
    assert "oracle driver done" in r.stdout


def test_reference_build_keeps_subnormals():
    # oracle/_ref is compiled with the reference's -Ofast but linked without it:
    # an -Ofast link adds crtfastmath.o, whose constructor would switch this
    # process (and every later oracle and numpy computation) to flush-to-zero
    if O.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    tiny = np.float32(1e-39)
    assert tiny * np.float32(1.0) == tiny and tiny != 0


def test_fast_build_is_bit_identical(small_l2):
    # the CPU baseline's -O3 -march=x86-64-v3 build gives the checker's bits
    ix, q = small_l2
    ref = CI.oracle_index(ix)
    fast = O.OracleIndex(ix["coarse"], ix["fine"], ix["codebook"], ix["cell_assign"], ix["records"], ix["raw"],
                         ix["whether_norm"], L=O.lib_fast())
    for scc, ftk in [(3, 100), (12, 500)]:
        a = ref.search_batch(q, 10, scc, ftk)
        b = fast.search_batch(q, 10, scc, ftk)
        assert np.array_equal(a[1], b[1]) and np.array_equal(bits(a[0]), bits(b[0])) and np.array_equal(a[2], b[2])
        nf = (np.arange(q.shape[0]) % 5 + 1).astype(np.uint32)
        a = ref.search_batch_pq(q, 10, scc, ftk, fine_budget=nf)
        b = fast.search_batch_pq(q, 10, scc, ftk, fine_budget=nf)
        assert np.array_equal(a[1], b[1]) and np.array_equal(bits(a[0]), bits(b[0]))
