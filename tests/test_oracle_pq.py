"""The oracle's second-stage PQ (whether_pq = true) against what the
mathematics fixes.

PuckIndex::puck_assign with both quantizers (puck/puck/puck_index.cpp:602-709)
encodes r = x - c - f with the filter codebook, subtracts the chosen
codewords, then encodes what is left with the PQ codebook; the offsets are
2<dec_f, c + f> and 2<dec_pq, c + f + dec_f> (quantization.cpp:244-258).
rank_topN_points (puck_index.cpp:402-427) ranks the filter survivors by
(filter distance - |q|^2) + PQ offset + PQ LUT sum, which equals
||q - (c + f + dec_f + dec_pq)||^2.  Checked in float64:
* every PQ code is the nearest codeword of the residual left by the filter
  codewords (a code chosen on r itself, or with the filter residual not
  subtracted, fails);
* the offsets are those inner products;
* returned distances are the squared distances to the two-stage
  reconstructions, and a full scan returns their brute-force top-k.
"""
import numpy as np
import pytest

from oracle import oracle as O
from puck_amd import datasets as D
import cpu_index as CI

F64 = np.float64


@pytest.fixture(scope="module")
def pq_cos():
    x = D.unit_norm_mixture(6000, 32, 60, seed=81, means_seed=82)
    q = D.unit_norm_mixture(24, 32, 60, seed=83, means_seed=82)
    ix = CI.build_index_pq(x, 6, 6, 16, 32, whether_norm=True, train_n=6000)
    return ix, q


@pytest.fixture(scope="module")
def pq_l2():
    x = D.mixture_uint8_grid(5000, 16, 40, seed=84, means_seed=85)
    q = D.mixture_uint8_grid(16, 16, 40, seed=86, means_seed=85)
    ix = CI.build_index_pq(x, 5, 5, 4, 8, whether_norm=False, train_n=5000)
    return ix, q


def _decode(cb, codes):
    nsq = cb.shape[0]
    return np.concatenate([cb[m][codes[:, m]] for m in range(nsq)], axis=1)


def _recon(ix):
    F = ix["fine"].shape[0]
    cell = ix["cell_assign"]
    cf = ix["coarse"][cell // F].astype(F64) + ix["fine"][cell % F].astype(F64)
    df = _decode(ix["codebook"], ix["records"][:, 4:].astype(np.int64)).astype(F64)
    dp = _decode(ix["pq_codebook"], ix["pq_records"][:, 4:].astype(np.int64)).astype(F64)
    return cf, df, dp


@pytest.mark.parametrize("which", ["pq_cos", "pq_l2"])
def test_pq_codes_are_nearest_on_the_filter_residual(which, request):
    ix, _ = request.getfixturevalue(which)
    x = ix["raw"].astype(F64)
    cf, df, dp = _recon(ix)
    r1 = x - cf - df                                   # what the filter codewords leave
    pcb = ix["pq_codebook"].astype(F64)
    pnsq, plsq = pcb.shape[0], pcb.shape[2]
    codes = ix["pq_records"][:, 4:].astype(np.int64)
    worse_on_r0 = 0
    for m in range(pnsq):
        sub = r1[:, m * plsq:(m + 1) * plsq]
        d = ((sub[:, None, :] - pcb[m][None, :, :]) ** 2).sum(-1)
        best = d.min(1)
        got = d[np.arange(len(codes)), codes[:, m]]
        # the chosen codeword is the nearest up to float32 rounding of the scores
        assert np.all(got <= best + 1e-4 * (1.0 + np.abs(best))), f"subspace {m}"
        sub0 = (x - cf)[:, m * plsq:(m + 1) * plsq]
        worse_on_r0 += int((((sub0[:, None, :] - pcb[m][None]) ** 2).sum(-1).argmin(1) != codes[:, m]).sum())
    assert worse_on_r0 > 0  # encoding r without the filter codewords subtracted picks other codes


@pytest.mark.parametrize("which", ["pq_cos", "pq_l2"])
def test_pq_offsets(which, request):
    ix, _ = request.getfixturevalue(which)
    cf, df, dp = _recon(ix)
    off_f = ix["records"][:, :4].copy().view(np.float32).reshape(-1).astype(F64)
    off_p = ix["pq_records"][:, :4].copy().view(np.float32).reshape(-1).astype(F64)
    want_f = 2 * (df * cf).sum(1)
    want_p = 2 * (dp * (cf + df)).sum(1)
    scale = np.abs(cf).sum(1) * (np.abs(df).sum(1) + np.abs(dp).sum(1)) + 1e-6
    assert np.all(np.abs(off_f - want_f) <= 1e-5 * scale)
    assert np.all(np.abs(off_p - want_p) <= 1e-5 * scale)
    # the PQ offset includes the filter decode (2<dec_pq, c + f> alone differs)
    assert np.abs(off_p - 2 * (dp * cf).sum(1)).max() > 1e-3 * np.abs(off_p).max()


@pytest.mark.parametrize("which", ["pq_cos", "pq_l2"])
def test_pq_distances_are_two_stage_reconstruction_distances(which, request):
    ix, q = request.getfixturevalue(which)
    oi = CI.oracle_index(ix)
    cf, df, dp = _recon(ix)
    recon = cf + df + dp
    C = ix["coarse"].shape[0]
    for k in range(q.shape[0]):
        qq = O.normalize(q[k]) if ix["whether_norm"] else q[k]
        for scc, ftk in [(2, 200), (C, 10 ** 6)]:
            o = oi.search_one(q[k], 10, scc, ftk)
            n = o["n"]
            ids = o["ids"][:n].astype(np.int64)
            want = ((recon[ids] - qq.astype(F64)) ** 2).sum(1)
            scale = (np.abs(recon[ids]) + np.abs(qq)).sum(1) ** 2 + 1e-6
            assert np.all(np.abs(o["dist"][:n] - want) <= 1e-5 * scale), (k, scc)
            assert np.all(np.diff(o["dist"][:n]) >= 0)
            if ftk >= ix["raw"].shape[0]:
                # full scan: the brute-force top-k of the reconstruction distances
                full = ((recon - qq.astype(F64)) ** 2).sum(1)
                kth = np.sort(full)[9]
                margin = 1e-5 * float(scale.max())
                assert np.all(full[ids] <= kth + margin)
                assert np.sum(full < kth - margin) <= n


def test_pq_budget_scan_ranks_by_pq_distance(pq_cos):
    ix, q = pq_cos
    oi = CI.oracle_index(ix)
    cf, df, dp = _recon(ix)
    recon = cf + df + dp
    for k in range(8):
        qq = O.normalize(q[k]).astype(F64)
        o = oi.search_one(q[k], 10, 3, 300, fine_budget=3)
        n = o["n"]
        ids = o["ids"][:n].astype(np.int64)
        want = ((recon[ids] - qq) ** 2).sum(1)
        assert np.allclose(o["dist"][:n], want, rtol=1e-4, atol=1e-6)
        assert np.all(np.diff(o["dist"][:n]) >= 0)


def test_pq_search_without_raw_vectors(pq_cos):
    # a whether_pq index ranks by the PQ distance: replacing the raw vectors
    # changes nothing (the reference does not load them, puck_index.cpp:198-227)
    ix, q = pq_cos
    a = CI.oracle_index(ix)
    ix2 = dict(ix, raw=np.zeros_like(ix["raw"]))
    b = CI.oracle_index(ix2)
    for k in range(6):
        ra, rb = a.search_one(q[k], 10, 4, 200), b.search_one(q[k], 10, 4, 200)
        assert np.array_equal(ra["ids"], rb["ids"]) and np.array_equal(ra["dist"].view(np.uint32),
                                                                       rb["dist"].view(np.uint32))
