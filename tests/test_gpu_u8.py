"""The uint8 raw-vector storage form (puckgpu_index_attach_raw_u8) on integer
data: every search form returns the bits of the same index holding float
rows, and the oracle's results; the misuse cases are refused.

Data: the 1M x 128 generator's mixture (integers 0..255 stored as float) at
12,000 x 128, 48 x 48 cells, filter PQ nsq 32, built by the CPU restatement.
"""
import numpy as np
import pytest

from oracle import oracle as O
from puck_amd import datasets as D
import cpu_index as CI

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def env():
    from puck_amd.engine import Context, GpuIndex
    ctx = Context(0)
    x = D.mixture_uint8_grid(12000, 128, 200, seed=30, means_seed=31)
    q = D.mixture_uint8_grid(96, 128, 200, seed=32, means_seed=31)
    ix = CI.build_index(x, 48, 48, 32, whether_norm=False, train_n=6000)
    assert np.array_equal(x, np.rint(x)) and x.min() >= 0 and x.max() <= 255
    cb = ix["codebook"].reshape(-1)
    gf = GpuIndex(ctx, ix["coarse"], ix["fine"], cb, ix["cell_assign"], ix["records"], ix["raw"], False)
    g8 = GpuIndex(ctx, ix["coarse"], ix["fine"], cb, ix["cell_assign"], ix["records"], x.astype(np.uint8), False)
    return ctx, ix, q, gf, g8


@pytest.mark.parametrize("scc,ftk,replay", [(6, 200, False), (48, 1000, False), (8, 300, True)])
def test_u8_plain_search_equals_float_rows_and_oracle(env, scc, ftk, replay):
    ctx, ix, q, gf, g8 = env
    D8, I8, R8 = g8.search(q, 10, scc, ftk, replay_all=replay)
    Df, If, Rf = gf.search(q, 10, scc, ftk, replay_all=replay)
    assert np.array_equal(R8, Rf) and np.array_equal(I8, If) and np.array_equal(_bits(D8), _bits(Df))
    Do, Io, Ro, _ = CI.oracle_index(ix).search_batch(q, 10, scc, ftk)
    assert np.array_equal(R8, Ro) and np.array_equal(I8.view(np.uint32), Io)
    assert np.array_equal(_bits(D8), _bits(Do))


def test_u8_budget_scan_equals_float_rows(env):
    ctx, ix, q, gf, g8 = env
    nf = (np.arange(q.shape[0], dtype=np.uint32) % 9) + 1
    D8, I8, R8 = g8.search_batch(q, 10, 12, 400, fine_budget=nf)
    Df, If, Rf = gf.search_batch(q, 10, 12, 400, fine_budget=nf)
    assert np.array_equal(R8, Rf) and np.array_equal(I8, If) and np.array_equal(_bits(D8), _bits(Df))
    Do, Io, Ro = CI.oracle_index(ix).search_batch_pq(q, 10, 12, 400, fine_budget=nf)
    assert np.array_equal(I8.view(np.uint32), Io) and np.array_equal(_bits(D8), _bits(Do))


def test_u8_refusals(env):
    from puck_amd.engine import GpuIndex
    ctx, ix, q, gf, g8 = env
    x8 = ix["raw"].astype(np.uint8)
    with pytest.raises(ValueError):
        GpuIndex(ctx, ix["coarse"], ix["fine"], ix["codebook"].reshape(-1), ix["cell_assign"], ix["records"], x8,
                 True)
    with pytest.raises(Exception):
        g8.insert(q[:2])
