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
