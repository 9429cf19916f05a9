"""The C-ABI library loads and exports every symbol include/puck_gpu.h
declares (no compute call: no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "puck_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(puckgpu_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for need in ("puckgpu_create", "puckgpu_index_load", "puckgpu_search", "puckgpu_assign_cells",
                 "puckgpu_encode_filter", "puckgpu_kmeans_assign", "puckgpu_last_error"):
        assert need in syms


def test_library_exports_every_declared_symbol():
    from puck_amd import _lib
    path = _lib.LIB_PATH
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (run python puck_amd/build.py)")
    L = ctypes.CDLL(path)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert set(_lib.EXPORTS) == set(declared_symbols())


def test_library_rejects_null_context():
    from puck_amd import _lib
    L = _lib.load()
    assert L.puckgpu_filter_topk_max() >= 1000
    assert L.puckgpu_create(0, None) != 0
    assert L.puckgpu_last_error(None) == b"null context"


def test_bench_refuses_more_than_one_gpu():
    # the bench measures one GPU; --gpus N > 1 exits with a message before touching a device
    import json
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert msg["n_gpus"] == 2 and "one GPU" in msg["error"]


def test_binding_argument_checks_before_any_device_call():
    # the Python binding refuses bad arrays before it reaches the library
    import types
    import numpy as np
    from puck_amd import engine as E
    with pytest.raises(TypeError):
        E._ptr(np.zeros(3, np.float64), np.float32)
    with pytest.raises(ValueError):
        E._ptr(np.zeros((4, 4), np.float32)[:, ::2])
    assert E._mode(np.zeros(2, np.float32), None) == 0
    dev = types.SimpleNamespace(is_cuda=True)            # stands in for a device tensor
    assert E._mode(dev, None, dev) == 1
    with pytest.raises(ValueError):
        E._mode(np.zeros(2, np.float32), dev)           # host and device arrays mixed
    sv = object.__new__(E.Server)                       # no device: only the shape check runs
    sv.max_batch, sv.index = 8, types.SimpleNamespace(dim=4)
    with pytest.raises(ValueError):
        sv.tune(np.zeros((4, 4), np.float32))           # the largest bucket reads max_batch rows


def test_binding_refuses_bad_search_arrays_before_any_device_call():
    # dtype, shape, output size and device are checked before the library is
    # reached; stand-in objects play the index and device tensors (no GPU here)
    import types
    import numpy as np
    from puck_amd import engine as E
    ix = object.__new__(E.GpuIndex)
    ix.ctx = types.SimpleNamespace(device=0)
    ix.dim, ix.C, ix.F = 8, 4, 4
    q = np.zeros((3, 8), np.float32)
    D, I, R = np.zeros((3, 5), np.float32), np.zeros((3, 5), np.uint32), np.zeros(3, np.uint32)
    ix._check_io(q, 3, D, I, R, 5)                                   # accepted
    with pytest.raises(TypeError):
        ix._check_io(q.astype(np.float64), 3, D, I, R, 5)            # float64 queries
    with pytest.raises(ValueError):
        ix._check_io(np.zeros((3, 7), np.float32), 3, D, I, R, 5)    # wrong dim
    with pytest.raises(ValueError):
        ix._check_io(q, 3, np.zeros((2, 5), np.float32), I, R, 5)    # too few output rows
    with pytest.raises(ValueError):
        ix._check_io(q, 3, D, np.zeros((3, 4), np.uint32), R, 5)     # rows narrower than the stride
    with pytest.raises(TypeError):
        ix._check_io(q, 3, D, I.astype(np.int64), R, 5)              # 64-bit labels
    with pytest.raises(ValueError):
        ix._check_io(q, 3, D, I, R, 5, (("topk", np.zeros(2, np.uint32)),))   # per-query array too short

    def dev_tensor(shape, dtype, index):
        return types.SimpleNamespace(is_cuda=True, dtype=dtype, shape=shape, device=types.SimpleNamespace(index=index),
                                     is_contiguous=lambda: True)
    ok = dev_tensor((3, 8), "torch.float32", 0)
    E._check(ok, "queries", E._F32, shape=(3, 8), device=0)
    with pytest.raises(ValueError):
        E._check(dev_tensor((3, 8), "torch.float32", 1), "queries", E._F32, shape=(3, 8), device=0)
    with pytest.raises(TypeError):
        E._check(dev_tensor((3, 8), "torch.float64", 0), "queries", E._F32, shape=(3, 8), device=0)
    sv = object.__new__(E.Server)
    sv.max_batch, sv.index, sv.topk_max, sv.ctx = 2, ix, 5, ix.ctx
    with pytest.raises(ValueError):
        sv.search(q, out=(D, I, R))                                  # more queries than max_batch
