"""PySearcher (the reference's Python entry points) on the GPU engine,
checked against the oracle on the same index directory."""
import numpy as np
import pytest

from oracle import oracle as O
from puck_amd import datasets as D
from puck_amd import flags as FL
from puck_amd import index_io
from puck_amd import py_puck_api as api
import cpu_index as CI

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _reset_flags():
    FL.reset()
    yield
    FL.reset()


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_searcher_matches_oracle(tmp_path):
    x = D.mixture_uint8_grid(6000, 64, 100, seed=30, means_seed=31)
    q = D.mixture_uint8_grid(40, 64, 100, seed=32, means_seed=31)
    ix = CI.build_index(x, 24, 24, 16, whether_norm=False, train_n=4000)
    index_io.write_index(str(tmp_path), FL, ix["coarse"], ix["fine"], ix["codebook"], ix["cell_assign"],
                         ix["records"], ix["raw"], False)
    for k, v in [("index_path", str(tmp_path)), ("topk", 20), ("search_coarse_count", 6), ("filter_topk", 300)]:
        api.update_gflag(k, v)
    s = api.PySearcher()
    assert s.init() == 0
    Dm = np.zeros((40, 10), np.float32)
    Im = np.zeros((40, 10), np.uint32)
    s.search(40, q, 10, Dm, Im)
    oi = CI.oracle_index(ix)
    Do, Io, _, _ = oi.search_batch(q, 10, 6, 300, max_topk=20)
    assert np.array_equal(Im, Io)
    assert np.array_equal(_bits(Dm), _bits(Do))
    # topk > FLAGS_topk: outputs untouched (puck_index.cpp:560-563)
    D2 = np.full((40, 30), 7.0, np.float32)
    I2 = np.full((40, 30), 7, np.uint32)
    s.search(40, q, 30, D2, I2)
    assert np.all(D2 == 7.0) and np.all(I2 == 7)
    # radius_rate < 1 or filter_topk < topk fail init (index_conf.cpp:187-195)
    api.update_gflag("radius_rate", "0.5")
    assert s.init() == -1
    api.update_gflag("radius_rate", "1.0")
    api.update_gflag("filter_topk", "5")
    assert s.init() == -1


@pytest.mark.parametrize("whether_pq", [False, True])
def test_searcher_build_then_search(tmp_path, whether_pq):
    # whether_pq defaults to true (puck_gflags.cpp): the build then trains and
    # writes the second-stage PQ and the index ranks by it
    x = D.unit_norm_mixture(8000, 32, 64, seed=40, means_seed=41)
    q = D.unit_norm_mixture(16, 32, 64, seed=42, means_seed=41)
    index_io.write_fvecs(str(tmp_path / "all_data.feat.bin"), x)
    for k, v in [("index_path", str(tmp_path)), ("feature_dim", 32), ("coarse_cluster_count", 16),
                 ("fine_cluster_count", 16), ("filter_nsq", 16), ("train_points_count", 4000),
                 ("kmeans_iterations_count", 4), ("topk", 10), ("search_coarse_count", 4), ("filter_topk", 100)]:
        api.update_gflag(k, v)
    if not whether_pq:
        api.update_gflag("whether_pq", "false")
    s = api.PySearcher()
    assert s.build(8000) == 0
    assert s.init() == 0
    Dm = np.zeros((16, 10), np.float32)
    Im = np.zeros((16, 10), np.uint32)
    s.search(16, q, 10, Dm, Im)
    conf, arr = index_io.read_index(str(tmp_path), FL)
    assert bool(conf["whether_pq"]) == whether_pq and (arr["features"] is None) == whether_pq
    oi = O.OracleIndex(arr["coarse"], arr["fine"], arr["codebook"], arr["cell_assign"], arr["records"],
                       x if whether_pq else arr["features"], conf["whether_norm"])
    if whether_pq:
        assert conf["nsq"] == 32  # one subspace per dimension by default (index_conf.cpp:39-49)
        oi.set_pq(arr["pq_codebook"], arr["pq_records"])
        # the GPU's PQ records equal the oracle's encoding of the same
        # (normalised, as the build normalises them) points
        alpha = (1.0 / np.sqrt((x.astype(np.float64) ** 2).sum(1))).astype(np.float32)
        xn = (x * alpha[:, None]).astype(np.float32)
        h = O.encode_puck(arr["coarse"], arr["fine"], arr["codebook"], 16, arr["pq_codebook"], 32, xn,
                          arr["cell_assign"])[1]
        assert np.array_equal(h, arr["pq_records"])
    Do, Io, _, _ = oi.search_batch(q, 10, 4, 100)
    assert np.array_equal(Im, Io)
    assert np.array_equal(_bits(Dm), _bits(Do))


def test_searcher_serves_pq_index(tmp_path):
    # a whether_pq = true directory written from the CPU build: no raw vectors
    # on disk, searched by the engine and the oracle alike
    x = D.mixture_uint8_grid(6000, 64, 100, seed=33, means_seed=34)
    q = D.mixture_uint8_grid(32, 64, 100, seed=35, means_seed=34)
    ix = CI.build_index_pq(x, 24, 24, 16, 64, whether_norm=False, train_n=4000)
    index_io.write_index(str(tmp_path), FL, ix["coarse"], ix["fine"], ix["codebook"], ix["cell_assign"],
                         ix["records"], None, False, pq_codebook=ix["pq_codebook"], pq_records=ix["pq_records"])
    assert not (tmp_path / "all_data.feat.bin").exists()
    for k, v in [("index_path", str(tmp_path)), ("topk", 20), ("search_coarse_count", 6), ("filter_topk", 300)]:
        api.update_gflag(k, v)
    s = api.PySearcher()
    assert s.init() == 0
    Dm = np.zeros((32, 10), np.float32)
    Im = np.zeros((32, 10), np.uint32)
    s.search(32, q, 10, Dm, Im)
    Do, Io, _, _ = CI.oracle_index(ix).search_batch(q, 10, 6, 300, max_topk=20)
    assert np.array_equal(Im, Io)
    assert np.array_equal(_bits(Dm), _bits(Do))
