"""GPU tests of the reduced-precision forms (BF16 / FP8 center distances on
the matrix cores, FP32 rescoring of the coarse boundary, exact rerank)."""
import numpy as np
import pytest

from oracle import oracle as O
from puck_amd import datasets as DS
import cpu_index as CI

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from puck_amd.engine import Context
    return Context(0)


@pytest.fixture(scope="module")
def idx_cos():
    x = DS.unit_norm_mixture(20000, 96, 300, seed=40, means_seed=41)
    q = DS.unit_norm_mixture(128, 96, 300, seed=42, means_seed=41)
    ix = CI.build_index(x, 48, 48, 48, whether_norm=True, train_n=8000)
    return ix, q


@pytest.fixture(scope="module")
def idx_l2():
    x = DS.mixture_uint8_grid(12000, 128, 200, seed=43, means_seed=44)
    q = DS.mixture_uint8_grid(64, 128, 200, seed=45, means_seed=44)
    ix = CI.build_index(x, 32, 32, 32, whether_norm=False, train_n=6000)
    return ix, q


def _gpu_index(ctx, ix):
    from puck_amd.engine import GpuIndex
    return GpuIndex(ctx, ix["coarse"], ix["fine"], ix["codebook"].reshape(-1), ix["cell_assign"],
                    ix["records"], ix["raw"], ix["whether_norm"])


@pytest.mark.parametrize("which", ["cos", "l2"])
@pytest.mark.parametrize("prec,rel", [(1, 2.0 ** -7), (2, 2.0 ** -3)])
def test_center_distance_error_vs_fp64(ctx, which, prec, rel, idx_cos, idx_l2):
    # |approx - exact| <= rel * sum_k |q_k c_k| (+ a tiny absolute term): the
    # bound of rounding each operand to 8 (BF16) or 4 (FP8 e4m3) significant
    # bits, with per-row scales for FP8
    ix, q = idx_cos if which == "cos" else idx_l2
    gi = _gpu_index(ctx, ix)
    co, fo = gi.center_distances(q, prec)
    ce, fe = gi.center_distances(q, 0)
    qn = np.stack([O.normalize(v) for v in q]) if ix["whether_norm"] else q.astype(np.float32)
    q64 = qn.astype(np.float64)
    for cent, got, ex, is_coarse in ((ix["coarse"], co, ce, True), (ix["fine"], fo, fe, False)):
        c64 = cent.astype(np.float64)
        dot = q64 @ c64.T
        mag = np.abs(q64) @ np.abs(c64).T
        want = (0.5 * (c64 ** 2).sum(1))[None, :] - dot if is_coarse else -dot
        # exact form: the FP32 engine order, close to FP64
        assert np.all(np.abs(ex - want) <= 1e-5 * mag + 1e-5 * np.abs(want) + 1e-6)
        err = np.abs(got.astype(np.float64) - want)
        assert np.all(err <= rel * mag + 1e-4 * np.abs(want) + 1e-3), (prec, float((err / (mag + 1e-9)).max()))


@pytest.mark.parametrize("prec", [1, 2])
def test_precision_search_exact_distances_and_recall(ctx, prec, idx_cos):
    ix, q = idx_cos
    gi = _gpu_index(ctx, ix)
    gt, gd = O.exact_topk(ix["raw"], np.stack([O.normalize(v) for v in q]), 10)
    De, Ie, Re = gi.search(q, 10, 8, 500)
    Dp, Ip, Rp = gi.search(q, 10, 8, 500, precision=prec)
    rec_e = O.recall_with_ties(gt, gd, Ie.astype(np.int64), 10)[0]
    rec_p = O.recall_with_ties(gt, gd, Ip.astype(np.int64), 10)[0]
    assert rec_p >= rec_e - (0.02 if prec == 1 else 0.06), (rec_e, rec_p)
    # the rerank is exact: every returned distance is the FP32 rerank distance
    for k in range(q.shape[0]):
        qn = O.normalize(q[k])
        for j in range(int(Rp[k])):
            assert np.float32(Dp[k, j]).view(np.uint32) == np.float32(O.rerank_dist(ix["raw"][Ip[k, j]], qn)).view(
                np.uint32)
        assert np.all(np.diff(Dp[k, :Rp[k]]) >= 0)
