"""The C++ puck::Index adapter (puck_amd/host) refuses bad search parameters
and index types in init() before it creates a device context, so these checks
need no GPU: an ip2cos index, filter_topk below topk (index_conf.cpp:192-195) and, beyond the reference,
a filter_topk above the engine's maximum (the unset default is 11 * topk,
index_conf.cpp:178-180, which passes 2048 from topk 187 on)."""
import os
import subprocess

import numpy as np
import pytest

from puck_amd import datasets as D
from puck_amd import flags as FL
from puck_amd import index_io
import cpu_index as CI

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "host_adapter_test")


@pytest.fixture(scope="module")
def index_dir(tmp_path_factory):
    x = D.mixture_uint8_grid(600, 16, 8, seed=70, means_seed=71)
    ix = CI.build_index(x, 4, 4, 8, whether_norm=False)
    d = tmp_path_factory.mktemp("adapter") / "idx"
    index_io.write_index(str(d), FL, ix["coarse"], ix["fine"], ix["codebook"], ix["cell_assign"], ix["records"],
                         ix["raw"], False)
    qf = d / "q.f32"
    np.ascontiguousarray(x[:2], np.float32).tofile(str(qf))
    return d, qf


@pytest.mark.parametrize("flags_topk, ftk, msg", [
    (200, -1, "filter_topk above the engine's maximum"),  # unset: 11 * 200 = 2200
    (100, 3000, "filter_topk above the engine's maximum"),
    (10, 5, "filter_topk should >= topk"),
])
def test_init_rejects_search_params(index_dir, tmp_path, flags_topk, ftk, msg):
    d, qf = index_dir
    out = str(tmp_path / "out.bin")
    args = [BIN, str(d), str(qf), "2", "16", "10", str(flags_topk), "2", str(ftk), out]
    r = subprocess.run(args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 4, (r.returncode, r.stderr)  # init() returned -1
    assert msg in r.stderr
    assert not os.path.exists(out)


def test_init_accepts_default_filter_topk(index_dir, tmp_path):
    # 11 * 100 = 1100 passes the checks; without a device init then stops at the context
    d, qf = index_dir
    out = str(tmp_path / "out.bin")
    args = [BIN, str(d), str(qf), "2", "16", "10", "100", "2", "-1", out]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        assert r.returncode == 4 and "no GPU context" in r.stderr, (r.returncode, r.stderr)


def test_init_rejects_ip2cos_index(tmp_path):
    d = tmp_path / "ip2cos"
    d.mkdir()
    index_io.write_index_conf(str(d / "index.dat"), 17, False, 0, 17, 4, 4, ip2cos=1)
    out = str(tmp_path / "out.bin")
    r = subprocess.run([BIN, str(d), str(d / "q.f32"), "1", "16", "1", "10", "2", "-1", out], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 4 and "ip2cos" in r.stderr, (r.returncode, r.stderr)


def test_adapter_host_paths_clean_under_sanitizers(index_dir, tmp_path):
    # the adapter's file reading and refusals (host code only, built with g++)
    # under AddressSanitizer / UndefinedBehaviorSanitizer / LeakSanitizer:
    # valid files up to the missing device, refusals, truncated files
    exe = str(tmp_path / "adapter_asan")
    cc = subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
                         "-o", exe, os.path.join(ROOT, "tests", "cpp", "host_adapter_main.cpp"),
                         os.path.join(ROOT, "puck_amd", "host", "gpu_puck_index.cpp"),
                         "-L" + os.path.join(ROOT, "puck_amd"), "-lpuckgpu",
                         "-Wl,-rpath," + os.path.join(ROOT, "puck_amd")], capture_output=True, text=True, timeout=300)
    if cc.returncode != 0:
        pytest.skip("g++ sanitizer runtime unavailable: " + cc.stderr[-200:])
    d, qf = index_dir
    cut = {}
    for name, victim, keep in (("data", FL.get("filter_data_file_name"), -7), ("conf", FL.get("index_file_name"), 20)):
        cut[name] = tmp_path / ("trunc_" + name)
        cut[name].mkdir()
        for f in os.listdir(d):
            b = open(os.path.join(d, f), "rb").read()
            open(os.path.join(cut[name], f), "wb").write(b[:keep] if f == victim else b)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    out = str(tmp_path / "out.bin")
    for path, flags_topk, ftk in ((d, "200", "-1"), (d, "10", "5"), (d, "100", "-1"), (cut["data"], "100", "-1"),
                                  (cut["conf"], "100", "-1")):
        r = subprocess.run([exe, str(path), str(qf), "2", "16", "10", flags_topk, "2", ftk, out], capture_output=True,
                           text=True, timeout=120, env=env)
        assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-2000:]
        assert r.returncode in (0, 4), (path, r.returncode, r.stderr[-500:])


@pytest.mark.parametrize("index_type", [0, 1, 2])
def test_load_index_type(tmp_path, index_type):
    # load_index_type reads the type field of index.dat
    # (hierarchical_cluster_index.cpp:1544-1549); no device is touched
    d = tmp_path / f"t{index_type}"
    d.mkdir()
    index_io.write_index_conf(str(d / "index.dat"), 16, False, 10, 8, 2, 2, index_type=index_type)
    r = subprocess.run([BIN, str(d), "type"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert int(r.stdout.strip()) == index_type


def test_load_index_type_missing_file(tmp_path):
    r = subprocess.run([BIN, str(tmp_path / "absent"), "type"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and int(r.stdout.strip()) == -1


def test_init_reports_missing_pq_files(tmp_path):
    # a whether_pq index needs learn_codebooks.dat and pq_data.dat, not the
    # raw vectors; without them init fails before any device call
    x = D.mixture_uint8_grid(600, 16, 8, seed=72, means_seed=73)
    ix = CI.build_index(x, 4, 4, 8, whether_norm=False)
    d = tmp_path / "pq"
    index_io.write_index(str(d), FL, ix["coarse"], ix["fine"], ix["codebook"], ix["cell_assign"], ix["records"],
                         ix["raw"], False)
    index_io.write_index_conf(str(d / "index.dat"), 16, False, 600, 8, 4, 4, whether_pq=True, nsq=8)
    out = str(tmp_path / "out.bin")
    r = subprocess.run([BIN, str(d), str(d / "q.f32"), "1", "16", "1", "10", "2", "-1", out], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 4 and "check file length" in r.stderr, (r.returncode, r.stderr)


INS_BIN = os.path.join(ROOT, "tests", "cpp", "insert_adapter_test")


def test_realtime_init_cuts_files_to_the_label_count(tmp_path):
    # RealtimeInsertPuckIndex::init takes the point count from the label file
    # and cuts longer index files to it (realtime_insert_puck_index.cpp:52-129),
    # before any device call
    x = D.mixture_uint8_grid(700, 16, 8, seed=74, means_seed=75)
    ix = CI.build_index(x, 4, 4, 8, whether_norm=False)
    d = tmp_path / "rt"
    index_io.write_index(str(d), FL, ix["coarse"], ix["fine"], ix["codebook"], ix["cell_assign"], ix["records"],
                         ix["raw"], False)
    keep = 650
    with open(d / FL.get("label_file_name"), "w") as f:
        f.write("".join(f"p{i}\n" for i in range(keep)))
    q = tmp_path / "q.f32"
    x[:2].astype(np.float32).tofile(str(q))
    r = subprocess.run([INS_BIN, str(d), str(q), "0", "16", str(q), "2", "10", "100", "2", "-1",
                        str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        assert r.returncode == 4 and "no GPU context" in r.stderr, (r.returncode, r.stderr)
    assert os.path.getsize(d / FL.get("filter_data_file_name")) == keep * 12
    assert os.path.getsize(d / FL.get("cell_assign_file_name")) == keep * 4
    assert os.path.getsize(d / FL.get("feature_file_name")) == keep * (4 + 16 * 4)
    conf, arr = index_io.read_index(str(d), FL)
    assert conf["total_point_count"] == keep
    assert np.array_equal(arr["cell_assign"], ix["cell_assign"][:keep])


def test_realtime_init_refuses_short_files(tmp_path):
    x = D.mixture_uint8_grid(300, 16, 8, seed=76, means_seed=77)
    ix = CI.build_index(x, 4, 4, 8, whether_norm=False)
    d = tmp_path / "rt"
    index_io.write_index(str(d), FL, ix["coarse"], ix["fine"], ix["codebook"], ix["cell_assign"], ix["records"],
                         ix["raw"], False)
    with open(d / FL.get("label_file_name"), "w") as f:
        f.write("".join(f"p{i}\n" for i in range(301)))
    r = subprocess.run([INS_BIN, str(d), str(d / "none"), "0", "16", str(d / "none"), "0", "10", "100", "2", "-1",
                        str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 4 and "shorter than the label file" in r.stderr, (r.returncode, r.stderr)


def test_group_commit_under_thread_sanitizer(tmp_path):
    """The coalescer of concurrent Index::search calls (group_commit.h) under
    ThreadSanitizer: every item served exactly once, by one batch at a time of
    at most max_batch items, before its submit() returns."""
    exe = str(tmp_path / "gct")
    src = os.path.join(ROOT, "tests", "cpp", "group_commit_main.cpp")
    subprocess.check_call(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=thread", "-o", exe, src, "-lpthread"])
    for args in (("16", "1500", "7"), ("4", "3000", "1"), ("24", "400", "4096")):
        r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, (args, r.stdout, r.stderr)
        assert "ThreadSanitizer" not in r.stderr, r.stderr
        f = r.stdout.split()
        assert int(f[3]) == int(args[0]) * int(args[1])
        assert int(f[5]) <= int(args[2])
