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
