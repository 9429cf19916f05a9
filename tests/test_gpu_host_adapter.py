"""The C++ puck::Index adapter (puck_amd/host) on the GPU: Request/Response
searches, -1 on topk > FLAGS_topk, and build() of cell_assign / filter_data,
checked against the oracle on the same index files."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from puck_amd import datasets as D
from puck_amd import flags as FL
from puck_amd import index_io
import cpu_index as CI

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "host_adapter_test")


def _run(index_path, q, topk, flags_topk, scc, ftk, tmp, build=False, threads=0, log=None):
    qf = str(tmp / "q.f32")
    np.ascontiguousarray(q, np.float32).tofile(qf)
    out = str(tmp / "out.bin")
    args = [BIN, index_path, qf, str(q.shape[0]), str(q.shape[1]), str(topk), str(flags_topk), str(scc), str(ftk), out]
    if build:
        args.append("build")
    if threads:
        args.append(f"threads={threads}")
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    if log is not None:
        log.append(r.stdout)
    raw = np.fromfile(out, dtype=np.uint32)
    n = q.shape[0]
    Dm = raw[:n * topk].view(np.float32).reshape(n, topk)
    Im = raw[n * topk:2 * n * topk].reshape(n, topk)
    R = raw[2 * n * topk:]
    return Dm, Im, R


def test_cpp_adapter_search(tmp_path):
    x = D.mixture_uint8_grid(5000, 32, 60, seed=50, means_seed=51)
    q = D.mixture_uint8_grid(24, 32, 60, seed=52, means_seed=51)
    ix = CI.build_index(x, 16, 16, 8, whether_norm=False, train_n=4000)
    d = tmp_path / "idx"
    index_io.write_index(str(d), FL, ix["coarse"], ix["fine"], ix["codebook"], ix["cell_assign"], ix["records"],
                         ix["raw"], False)
    Dm, Im, R = _run(str(d), q, 10, 50, 5, 200, tmp_path)
    Do, Io, Ro, _ = CI.oracle_index(ix).search_batch(q, 10, 5, 200, max_topk=50)
    assert np.array_equal(R, Ro)
    assert np.array_equal(Im, Io)
    assert np.array_equal(Dm.view(np.uint32), Do.view(np.uint32))


def test_cpp_adapter_concurrent_search(tmp_path):
    """16 threads call Index::search at once (ParallelFor's pattern,
    py_api_wrapper.cpp:111-152): the calls are coalesced into device batches,
    and every query's result equals the oracle's and the one-at-a-time run."""
    x = D.mixture_uint8_grid(8000, 32, 60, seed=53, means_seed=54)
    q = D.mixture_uint8_grid(800, 32, 60, seed=55, means_seed=54)
    ix = CI.build_index(x, 16, 16, 8, whether_norm=False, train_n=4000)
    d = tmp_path / "idx"
    index_io.write_index(str(d), FL, ix["coarse"], ix["fine"], ix["codebook"], ix["cell_assign"], ix["records"],
                         ix["raw"], False)
    log = []
    Dc, Ic, Rc = _run(str(d), q, 10, 50, 6, 200, tmp_path, threads=16, log=log)
    Ds, Is, Rs = _run(str(d), q, 10, 50, 6, 200, tmp_path)
    Do, Io, Ro, _ = CI.oracle_index(ix).search_batch(q, 10, 6, 200, max_topk=50)
    print(log[0].strip(), flush=True)
    f = log[0].split()
    batches, served = int(f[1]), int(f[3])
    assert served == q.shape[0] and 1 <= batches <= q.shape[0]
    for Dm, Im, R in ((Dc, Ic, Rc), (Ds, Is, Rs)):
        assert np.array_equal(R, Ro)
        assert np.array_equal(Im, Io)
        assert np.array_equal(Dm.view(np.uint32), Do.view(np.uint32))


def test_cpp_adapter_build(tmp_path):
    x = D.unit_norm_mixture(6000, 32, 40, seed=60, means_seed=61)
    q = D.unit_norm_mixture(16, 32, 40, seed=62, means_seed=61)
    ix = CI.build_index(x, 12, 12, 16, whether_norm=True, train_n=4000)
    d = tmp_path / "idx"
    p = index_io.paths(str(d), FL)
    os.makedirs(str(d))
    index_io.write_index_conf(p["index"], 32, True, x.shape[0], 16, 12, 12)
    index_io.write_fvecs(p["coarse"], ix["coarse"])
    index_io.write_fvecs(p["fine"], ix["fine"])
    index_io.write_fvecs(p["filter_codebook"], ix["codebook"].reshape(16 * 256, 2))
    index_io.write_fvecs(p["feature"], x)
    Dm, Im, R = _run(str(d), q, 10, 10, 4, 100, tmp_path, build=True)
    conf, arr = index_io.read_index(str(d), FL)
    assert np.array_equal(arr["cell_assign"], ix["cell_assign"])       # GPU assignment == oracle assignment
    assert np.array_equal(arr["records"], ix["records"])               # GPU encoding == oracle encoding
    Do, Io, _, _ = CI.oracle_index(ix).search_batch(q, 10, 4, 100)
    assert np.array_equal(Im, Io)
    assert np.array_equal(Dm.view(np.uint32), Do.view(np.uint32))


def test_cpp_adapter_serves_and_builds_pq(tmp_path):
    # whether_pq = true: init reads learn_codebooks.dat + pq_data.dat in place
    # of the raw vectors; build() writes pq_data.dat with the GPU encoder
    x = D.mixture_uint8_grid(6000, 32, 50, seed=63, means_seed=64)
    q = D.mixture_uint8_grid(20, 32, 50, seed=65, means_seed=64)
    ix = CI.build_index_pq(x, 12, 12, 8, 16, whether_norm=False, train_n=5000)
    d = tmp_path / "idx"
    index_io.write_index(str(d), FL, ix["coarse"], ix["fine"], ix["codebook"], ix["cell_assign"], ix["records"],
                         None, False, pq_codebook=ix["pq_codebook"], pq_records=ix["pq_records"])
    Dm, Im, R = _run(str(d), q, 10, 10, 4, 120, tmp_path)
    Do, Io, Ro, _ = CI.oracle_index(ix).search_batch(q, 10, 4, 120)
    assert np.array_equal(R, Ro)
    assert np.array_equal(Im, Io)
    assert np.array_equal(Dm.view(np.uint32), Do.view(np.uint32))
    # build from the codebooks and the raw vectors alone
    b = tmp_path / "built"
    p = index_io.paths(str(b), FL)
    os.makedirs(str(b))
    index_io.write_index_conf(p["index"], 32, False, x.shape[0], 8, 12, 12, whether_pq=True, nsq=16)
    index_io.write_fvecs(p["coarse"], ix["coarse"])
    index_io.write_fvecs(p["fine"], ix["fine"])
    index_io.write_fvecs(p["filter_codebook"], ix["codebook"].reshape(8 * 256, 4))
    index_io.write_fvecs(p["pq_codebook"], ix["pq_codebook"].reshape(16 * 256, 2))
    index_io.write_fvecs(p["feature"], x)
    Db, Ib, Rb = _run(str(b), q, 10, 10, 4, 120, tmp_path, build=True)
    conf, arr = index_io.read_index(str(b), FL)
    assert conf["whether_pq"]
    assert np.array_equal(arr["cell_assign"], ix["cell_assign"])
    assert np.array_equal(arr["records"], ix["records"])
    assert np.array_equal(arr["pq_records"], ix["pq_records"])
    assert np.array_equal(Ib, Io) and np.array_equal(Db.view(np.uint32), Do.view(np.uint32))


INS_BIN = os.path.join(ROOT, "tests", "cpp", "insert_adapter_test")


def _run_insert(index_path, x_ins, q, topk, scc, ftk, tmp):
    xf, qf, out = str(tmp / "ins.f32"), str(tmp / "q.f32"), str(tmp / "out.bin")
    np.ascontiguousarray(x_ins, np.float32).tofile(xf)
    np.ascontiguousarray(q, np.float32).tofile(qf)
    args = [INS_BIN, index_path, xf, str(x_ins.shape[0]), str(q.shape[1]), qf, str(q.shape[0]), str(topk), str(topk),
            str(scc), str(ftk), out]
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, dtype=np.uint32)
    n = q.shape[0]
    Dm = raw[:n * topk].view(np.float32).reshape(n, topk)
    Im = raw[n * topk:2 * n * topk].reshape(n, topk)
    R = raw[2 * n * topk:]
    labels = open(out + ".labels").read().split("\n")[:-1]
    return Dm, Im, R, labels


@pytest.mark.parametrize("pq", [False, True])
def test_cpp_realtime_insert_then_reload(tmp_path, pq):
    # inserts through GpuRealtimeInsertPuckIndex equal the oracle's overlay;
    # after a restart the appended files load as an ordinary index
    x = D.mixture_uint8_grid(7000, 32, 50, seed=66, means_seed=67)
    q = D.mixture_uint8_grid(24, 32, 50, seed=68, means_seed=67)
    base, extra = x[:5500], x[5500:5800]
    if pq:
        ix = CI.build_index_pq(base, 12, 12, 8, 16, whether_norm=False, train_n=5000)
    else:
        ix = CI.build_index(base, 12, 12, 16, whether_norm=False, train_n=5000)
    d = tmp_path / "idx"
    index_io.write_index(str(d), FL, ix["coarse"], ix["fine"], ix["codebook"], ix["cell_assign"], ix["records"],
                         None if pq else ix["raw"], False, pq_codebook=ix.get("pq_codebook"),
                         pq_records=ix.get("pq_records"))
    N = base.shape[0]
    with open(d / FL.get("label_file_name"), "w") as f:
        f.write("".join(f"base_{i}\n" for i in range(N)))
    Dm, Im, R, labels = _run_insert(str(d), extra, q, 10, 4, 150, tmp_path)
    oi = CI.oracle_index(ix)
    oi.insert(extra)
    Do, Io, Ro, _ = oi.search_batch(q, 10, 4, 150)
    assert np.array_equal(R, Ro) and np.array_equal(Im, Io)
    assert np.array_equal(Dm.view(np.uint32), Do.view(np.uint32))
    want = [f"base_{i}" if i < N else f"ins_{i - N}" for k in range(q.shape[0]) for i in Io[k, :Ro[k]]]
    assert labels == want
    assert any(l.startswith("ins_") for l in labels)
    # restart: the files now hold every point (index.dat keeps the loaded
    # count until init takes it from the label file); the index is rebuilt
    # from them
    assert index_io.read_index_conf(str(d / FL.get("index_file_name")))["total_point_count"] == N
    Dr, Ir, Rr, labels_r = _run_insert(str(d), extra[:0], q, 10, 4, 150, tmp_path)
    conf, arr = index_io.read_index(str(d), FL)
    assert conf["total_point_count"] == N + extra.shape[0]
    rb = dict(ix, cell_assign=arr["cell_assign"], records=arr["records"],
              raw=arr["features"] if not pq else np.zeros((N + extra.shape[0], 32), np.float32))
    if pq:
        rb["pq_records"] = arr["pq_records"]
    Db, Ib, Rb, _ = CI.oracle_index(rb).search_batch(q, 10, 4, 150)
    assert np.array_equal(Rr, Rb) and np.array_equal(Ir, Ib)
    assert np.array_equal(Dr.view(np.uint32), Db.view(np.uint32))
