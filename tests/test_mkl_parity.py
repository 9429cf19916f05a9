"""Search results against an emulated MKL build of the reference.

The reference takes its center inner products (sgemm_ through
matrix_multiplication, hierarchical_cluster_index.cpp:78-90, :706;
puck_index.cpp:494), the
coarse and fine half-norms and the query norm (cblas_sdot, :279, :305,
puck_index.cpp:515), stationary (fine half-norm + sgemm_, :311-320), the
query normalisation (cblas_snrm2 + cblas_sscal, :1524-1531) and the filter
offsets (2 * cblas_sdot, quantization.cpp:251) from MKL, whose summation order
is unspecified; the engine and the oracle use sequential FMA chains there.
These tests take all of those values from the installed MKL (same index cells
and codes), run the oracle's walk and rerank on them, and compare the results
with the oracle's own, which the engine reproduces bit for bit.  They skip
where MKL is not installed.
"""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O
from puck_amd import datasets as D
import cpu_index as CI

MKL_PATH = "/opt/conda/lib/libmkl_rt.so"
pytestmark = pytest.mark.skipif(not os.path.exists(MKL_PATH), reason="MKL runtime not installed")
F32 = np.float32


@pytest.fixture(scope="module")
def mkl():
    L = ctypes.CDLL(MKL_PATH)
    L.MKL_Set_Threading_Layer.argtypes = [ctypes.c_int]
    L.MKL_Set_Threading_Layer(1)  # sequential: no second OpenMP runtime beside the oracle's
    L.cblas_sdot.restype = ctypes.c_float
    L.cblas_sdot.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    L.cblas_snrm2.restype = ctypes.c_float
    L.cblas_snrm2.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    return L


def _sdot(L, a, b):
    a, b = np.ascontiguousarray(a, F32), np.ascontiguousarray(b, F32)
    return F32(L.cblas_sdot(a.size, a.ctypes.data, 1, b.ctypes.data, 1))


def _sgemm_tn(L, left, right, m, n, k):
    """matrix_multiplication(left, right, m, n, k, "TN", result): column-major m x n."""
    left, right = np.ascontiguousarray(left, F32), np.ascontiguousarray(right, F32)
    out = np.zeros(m * n, F32)
    M, N, K = ctypes.c_int(m), ctypes.c_int(n), ctypes.c_int(k)
    lda, ldb, ldc = ctypes.c_int(k), ctypes.c_int(k), ctypes.c_int(m)
    alpha, beta = ctypes.c_float(1.0), ctypes.c_float(0.0)
    L.sgemm_(ctypes.c_char_p(b"T"), ctypes.c_char_p(b"N"), ctypes.byref(M), ctypes.byref(N), ctypes.byref(K),
             ctypes.byref(alpha), left.ctypes.data_as(ctypes.c_void_p), ctypes.byref(lda),
             right.ctypes.data_as(ctypes.c_void_p), ctypes.byref(ldb), ctypes.byref(beta),
             out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(ldc))
    return out


def _mkl_index(L, ix):
    """The oracle index an MKL build would load: offsets, half-norms and
    stationary from MKL, same cells and codes."""
    coarse, fine = ix["coarse"], ix["fine"]
    C, d = coarse.shape
    F = fine.shape[0]
    nsq = ix["records"].shape[1] - 4
    cb = ix["codebook"].reshape(nsq, 256, d // nsq)
    rec = ix["records"].copy()
    for i in range(rec.shape[0]):
        cell = int(ix["cell_assign"][i])
        v = (coarse[cell // F] + fine[cell % F]).astype(F32)          # copy of c, saxpy(1.0, f)
        dec = np.concatenate([cb[m, rec[i, 4 + m]] for m in range(nsq)])
        rec[i, :4] = np.frombuffer(F32(2 * _sdot(L, dec, v)).tobytes(), np.uint8)
    oi = O.OracleIndex(coarse, fine, ix["codebook"], ix["cell_assign"], rec, ix["raw"], ix["whether_norm"])
    cn = np.array([_sdot(L, c, c) / F32(2) for c in coarse], F32)
    fn = np.array([_sdot(L, g, g) / F32(2) for g in fine], F32)
    st = (fn[None, :] + _sgemm_tn(L, fine, coarse, F, C, d).reshape(C, F)).astype(F32)
    oi.set_derived(cn, st.reshape(-1))
    return oi, st, cn


def _mkl_query(L, q, whether_norm):
    q = np.ascontiguousarray(q, F32)
    if not whether_norm:
        return q
    alpha = F32(1.0 / float(F32(L.cblas_snrm2(q.size, q.ctypes.data, 1))))
    return (alpha * q).astype(F32)


def test_search_one_given_reproduces_search_one():
    # the hook with the oracle's own inner products is the oracle's walk
    x = D.unit_norm_mixture(4000, 32, 40, seed=81, means_seed=82)
    q = D.unit_norm_mixture(20, 32, 40, seed=83, means_seed=82)
    ix = CI.build_index(x, 8, 8, 16, whether_norm=True, train_n=4000)
    oi = CI.oracle_index(ix)
    for k in range(q.shape[0]):
        qn = O.normalize(q[k])
        cip = np.array([O.dot(c, qn) for c in ix["coarse"]], F32)
        fip = np.array([O.dot(g, qn) for g in ix["fine"]], F32)
        Dg, Ig, ng = oi.search_one_given(qn, cip, fip, O.dot(qn, qn), 10, 3, 200)
        o = oi.search_one(q[k], 10, 3, 200)
        assert ng == o["n"]
        assert np.array_equal(Ig, o["ids"]) and np.array_equal(Dg.view(np.uint32), o["dist"].view(np.uint32))


@pytest.mark.parametrize("l2", [False, True])
def test_results_match_emulated_mkl_build(mkl, l2):
    # every query returns the same ids at the same distances (bit for bit on
    # L2 data, within 1e-4 relative after the cosine normalisation) as under
    # MKL's arithmetic; ids may swap only between exactly equal distances (the
    # filter heap's arrival order follows the filter distances' last bits)
    if l2:
        x = D.mixture_uint8_grid(20000, 64, 200, seed=91, means_seed=92)
        q = D.mixture_uint8_grid(200, 64, 200, seed=93, means_seed=92)
        ix = CI.build_index(x, 48, 48, 16, whether_norm=False, train_n=20000)
        scc, ftk = 12, 400
    else:
        x = D.unit_norm_mixture(20000, 96, 200, seed=94, means_seed=95)
        q = D.unit_norm_mixture(200, 96, 200, seed=96, means_seed=95)
        ix = CI.build_index(x, 48, 48, 48, whether_norm=True, train_n=20000)
        scc, ftk = 6, 500
    oi = CI.oracle_index(ix)
    oim, st_mkl, _ = _mkl_index(mkl, ix)
    st_own = oi.export()["stationary"].reshape(st_mkl.shape)
    live = st_own < np.sqrt(np.finfo(F32).max)
    assert np.mean(st_mkl[live].view(np.uint32) != st_own[live].view(np.uint32)) > 0.3   # MKL's order differs
    C, F = st_mkl.shape
    same = total = 0
    for k in range(q.shape[0]):
        qm = _mkl_query(mkl, q[k], not l2)
        cip = _sgemm_tn(mkl, ix["coarse"], qm, C, 1, qm.size)
        fip = _sgemm_tn(mkl, ix["fine"], qm, F, 1, qm.size)
        Dm, Im, nm = oim.search_one_given(qm, cip, fip, _sdot(mkl, qm, qm), 10, scc, ftk)
        o = oi.search_one(q[k], 10, scc, ftk)
        assert nm == o["n"], k
        if l2:  # the query is used as given: the rerank distances are the same bits
            assert np.array_equal(Dm[:nm].view(np.uint32), o["dist"][:nm].view(np.uint32)), k
        else:   # snrm2's last bit can differ from the oracle's norm: within 1e-4 relative
            rel = np.abs(Dm[:nm].astype(np.float64) - o["dist"][:nm]) / np.maximum(np.abs(o["dist"][:nm]), 1e-30)
            assert np.all(rel <= 1e-4), k
        assert set(Im[:nm].tolist()) == set(o["ids"][:nm].tolist()), k
        for j in np.nonzero(Im[:nm] != o["ids"][:nm])[0]:
            assert np.sum(o["dist"][:nm] == o["dist"][j]) > 1, (k, j)   # a swap inside a tie group
        same += int(np.sum(Im[:nm] == o["ids"][:nm]))
        total += nm
    assert same / total >= 0.999


@pytest.mark.parametrize("l2", [False, True])
def test_cell_assignment_matches_emulated_mkl_build(mkl, l2):
    # nearest_cell_assign (hierarchical_cluster_index.cpp:1042-1138) on MKL's
    # arithmetic -- sgemm_ inner products per chunk (the sequential FMA chain,
    # see below), cblas_sdot half-norms (another order), stationary and the
    # fine bound from read_coodbooks / :1186-1194 -- gives the oracle's cells
    # (which the engine reproduces bit for bit); a point may differ only where
    # its two cells are equally far in FP64
    gen, d = (D.mixture_uint8_grid, 64) if l2 else (D.unit_norm_mixture, 96)
    x = gen(20000, d, 200, seed=97, means_seed=98)
    ix = CI.build_index(x, 48, 48, 16, whether_norm=not l2, train_n=20000)
    coarse, fine = ix["coarse"], ix["fine"]
    C, F = coarse.shape[0], fine.shape[0]
    cell_own, _ = O.assign_cells(coarse, fine, x)
    cn = np.array([_sdot(mkl, c, c) / F32(2) for c in coarse], F32)
    fn = np.array([_sdot(mkl, g, g) / F32(2) for g in fine], F32)
    fbound = (F32(0) - np.sqrt(fn)).astype(F32)
    st = (fn[None, :] + _sgemm_tn(mkl, fine, coarse, F, C, d).reshape(C, F)).astype(F32)
    xs = np.ascontiguousarray(x, F32)
    cip = _sgemm_tn(mkl, coarse, xs, C, xs.shape[0], d).reshape(-1, C)
    fip = _sgemm_tn(mkl, fine, xs, F, xs.shape[0], d).reshape(-1, F)
    xn = np.array([_sdot(mkl, v, v) / F32(2) for v in xs], F32)
    own_n = np.array([F32(O.dot(v, v)) / F32(2) for v in np.concatenate([coarse, fine])], F32)
    assert np.mean(np.concatenate([cn, fn]).view(np.uint32) != own_n.view(np.uint32)) > 0.3   # MKL's sdot differs
    cell_mkl, _ = O.assign_cells_given(cn, st, fbound, cip, fip, xn)
    diff = np.nonzero(cell_mkl != cell_own)[0]
    assert diff.size <= 0.001 * x.shape[0]
    for i in diff:
        a, b = int(cell_mkl[i]), int(cell_own[i])
        da = np.sum((xs[i].astype(np.float64) - coarse[a // F] - fine[a % F]) ** 2)
        db = np.sum((xs[i].astype(np.float64) - coarse[b // F] - fine[b % F]) ** 2)
        assert abs(da - db) <= 1e-6 * max(da, db), i


def test_multi_column_sgemm_is_the_fma_chain(mkl):
    # MKL's sgemm_ "TN" with six or more columns accumulates every entry as a
    # sequential fused multiply-add chain from +0 -- the engine's and the
    # oracle's order -- on this host; with one to five columns (the search's
    # single-query products) it takes another order.  The reference's
    # multi-column calls are stationary's product (C columns) and the build
    # assignment's chunks of points.
    rng = np.random.default_rng(99)
    for n, k in ((6, 96), (8, 128), (64, 96), (1000, 64), (256, 128)):
        A = rng.standard_normal((300, k)).astype(F32)
        B = rng.standard_normal((n, k)).astype(F32)
        R = _sgemm_tn(mkl, A, B, 300, n, k).reshape(n, 300)
        for j in rng.choice(n, min(n, 8), replace=False):
            own = np.array([O.dot(A[i], B[j]) for i in range(300)], F32)
            assert np.array_equal(R[j].view(np.uint32), own.view(np.uint32)), (n, k, j)
    A = rng.standard_normal((300, 96)).astype(F32)
    b = rng.standard_normal((1, 96)).astype(F32)
    one = _sgemm_tn(mkl, A, b, 300, 1, 96)
    own = np.array([O.dot(A[i], b[0]) for i in range(300)], F32)
    assert np.mean(one.view(np.uint32) != own.view(np.uint32)) > 0.3
