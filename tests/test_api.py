"""Host-side API: flags, reference on-disk formats, PySearcher error paths
(no GPU needed)."""
import os
import struct

import numpy as np
import pytest

from puck_amd import flags as FL
from puck_amd import index_io
from puck_amd import py_puck_api as api


@pytest.fixture(autouse=True)
def _reset_flags():
    FL.reset()
    yield
    FL.reset()


def test_flag_defaults_and_update():
    assert FL.get("whether_norm") is True          # puck_gflags.cpp:42
    assert FL.get("filter_topk") == 1100            # FLAGS_topk * 11 at static init
    assert FL.get("filter_nsq") == 64               # default feature_dim / 4
    assert FL.update_gflag("topk", "10") and FL.get("topk") == 10
    assert FL.update_gflag("whether_norm", "false") and FL.get("whether_norm") is False
    assert FL.update_gflag("radius_rate", "1.5") and FL.get("radius_rate") == 1.5
    assert not FL.update_gflag("no_such_flag", "1")
    assert not FL.update_gflag("topk", "ten")
    assert FL.is_default("filter_topk") and not FL.is_default("topk")


def test_index_conf_layout(tmp_path):
    p = tmp_path / "index.dat"
    index_io.write_index_conf(str(p), 96, True, 1234, 48, 2000, 2000)
    raw = p.read_bytes()
    assert len(raw) == 8 + 43
    assert struct.unpack_from("<Q", raw)[0] == 43
    # packed fields (hierarchical_cluster_index.cpp:522-543)
    assert struct.unpack_from("<iI", raw, 8) == (1, 96)
    assert raw[16] == 1
    c = index_io.read_index_conf(str(p))
    assert (c["feature_dim"], c["whether_norm"], c["total_point_count"], c["filter_nsq"],
            c["coarse_cluster_count"], c["fine_cluster_count"], c["whether_pq"]) == (96, True, 1234, 48, 2000, 2000, False)


def test_fvecs_roundtrip_and_checks(tmp_path):
    x = np.random.default_rng(0).standard_normal((7, 5)).astype(np.float32)
    p = str(tmp_path / "a.dat")
    index_io.write_fvecs(p, x)
    assert os.path.getsize(p) == 7 * (4 + 20)
    assert np.array_equal(index_io.read_fvecs(p, 5, 7), x)
    with pytest.raises(ValueError):
        index_io.read_fvecs(p, 5, 8)
    with pytest.raises(ValueError):
        index_io.read_fvecs(p, 4)


def test_index_dir_roundtrip(tmp_path):
    rng = np.random.default_rng(1)
    C, F, d, nsq, N = 3, 4, 8, 4, 50
    arrs = dict(coarse=rng.standard_normal((C, d)).astype(np.float32),
                fine=rng.standard_normal((F, d)).astype(np.float32),
                codebook=rng.standard_normal(nsq * 256 * 2).astype(np.float32),
                cell_assign=rng.integers(0, C * F, N).astype(np.uint32),
                records=rng.integers(0, 256, (N, 4 + nsq)).astype(np.uint8),
                features=rng.standard_normal((N, d)).astype(np.float32))
    index_io.write_index(str(tmp_path), FL, arrs["coarse"], arrs["fine"], arrs["codebook"], arrs["cell_assign"],
                         arrs["records"], arrs["features"], False)
    conf, back = index_io.read_index(str(tmp_path), FL)
    assert conf["index_type"] == index_io.INDEX_PUCK and conf["filter_nsq"] == nsq
    for k in arrs:
        assert np.array_equal(back[k], arrs[k]), k
    # a truncated file fails the length check
    p = index_io.paths(str(tmp_path), FL)["filter_data"]
    open(p, "r+b").truncate(10)
    with pytest.raises(ValueError):
        index_io.read_index(str(tmp_path), FL)


def test_searcher_init_errors(tmp_path):
    s = api.PySearcher()
    api.update_gflag("index_path", str(tmp_path / "missing"))
    assert s.init() == -1
    d = tmp_path / "tinker"
    d.mkdir()
    index_io.write_index_conf(str(d / "index.dat"), 8, False, 0, 4, 2, 2, index_type=index_io.INDEX_TINKER)
    api.update_gflag("index_path", str(d))
    assert s.init() == -1
    e = tmp_path / "pq"  # a whether_pq index whose codebook files are missing
    e.mkdir()
    index_io.write_index_conf(str(e / "index.dat"), 8, False, 0, 4, 2, 2, whether_pq=True)
    api.update_gflag("index_path", str(e))
    assert s.init() == -1
    f = tmp_path / "ip2cos"
    f.mkdir()
    index_io.write_index_conf(str(f / "index.dat"), 9, False, 0, 3, 2, 2, ip2cos=1)
    api.update_gflag("index_path", str(f))
    assert s.init() == -1
    # a search before a successful init() is refused; the reference's wrapper
    # returns 0 from every search it runs
    assert s.search(1, np.zeros(9, np.float32), 1, np.zeros(1, np.float32), np.zeros(1, np.uint32)) == -1


def test_index_io_pq_round_trip(tmp_path):
    # whether_pq = true: learn_codebooks.dat and pq_data.dat written and read
    # back in the reference layouts; no raw feature file
    rng = np.random.default_rng(3)
    C, F, d, nsq, pnsq, N = 3, 4, 8, 4, 8, 50
    coarse, fine = rng.normal(size=(C, d)).astype(np.float32), rng.normal(size=(F, d)).astype(np.float32)
    cb = rng.normal(size=(nsq, 256, d // nsq)).astype(np.float32)
    pcb = rng.normal(size=(pnsq, 256, d // pnsq)).astype(np.float32)
    cell = rng.integers(0, C * F, N).astype(np.uint32)
    rec = rng.integers(0, 256, (N, 4 + nsq)).astype(np.uint8)
    prec = rng.integers(0, 256, (N, 4 + pnsq)).astype(np.uint8)
    index_io.write_index(str(tmp_path), FL, coarse, fine, cb, cell, rec, None, True, pq_codebook=pcb, pq_records=prec)
    assert not (tmp_path / FL.get("feature_file_name")).exists()
    assert (tmp_path / FL.get("pq_codebook_file_name")).stat().st_size == pnsq * 256 * (4 + 4 * (d // pnsq))
    conf, arr = index_io.read_index(str(tmp_path), FL)
    assert conf["whether_pq"] and conf["nsq"] == pnsq and arr["features"] is None
    assert np.array_equal(arr["pq_codebook"], pcb.reshape(-1)) and np.array_equal(arr["pq_records"], prec)
    assert np.array_equal(arr["records"], rec) and np.array_equal(arr["codebook"], cb.reshape(-1))
