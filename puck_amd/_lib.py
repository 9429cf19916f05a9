"""ctypes binding of libpuckgpu.so (include/puck_gpu.h).

The engine has no CPU fallback: if the library is missing or cannot be loaded
this module raises, and every search / build entry fails loudly.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpuckgpu.so")

# Every symbol declared in include/puck_gpu.h.
EXPORTS = [
    "puckgpu_create", "puckgpu_destroy", "puckgpu_last_error", "puckgpu_filter_topk_max",
    "puckgpu_index_load", "puckgpu_index_free", "puckgpu_index_export", "puckgpu_search",
    "puckgpu_last_filter", "puckgpu_last_stage_ms", "puckgpu_lut", "puckgpu_assign_cells",
    "puckgpu_encode_filter", "puckgpu_kmeans_assign", "puckgpu_search_batch",
    "puckgpu_probe_create", "puckgpu_probe_free", "puckgpu_probe_param_count", "puckgpu_probe_set_params",
    "puckgpu_probe_get_params", "puckgpu_probe_features", "puckgpu_probe_train_step", "puckgpu_probe_predict",
    "puckgpu_search_learned", "puckgpu_budget_ranks", "puckgpu_server_create", "puckgpu_server_free",
    "puckgpu_server_buckets", "puckgpu_server_search", "puckgpu_server_tune", "puckgpu_last_scan_ms",
    "puckgpu_center_distances", "puckgpu_index_attach_pq", "puckgpu_encode_pq", "puckgpu_search_noalloc",
    "puckgpu_index_insert", "puckgpu_index_version", "puckgpu_kmeans", "puckgpu_train_centers",
    "puckgpu_set_assign_mode", "puckgpu_assign_stats", "puckgpu_index_attach_raw_u8", "puckgpu_set_progress",
    "puckgpu_exact_search", "puckgpu_server_tuned", "puckgpu_set_rerank_rows",
]

# void (*)(const char* message, void* user)
PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.c_char_p, ctypes.c_void_p)


class IndexParams(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_uint32), ("coarse_cluster_count", ctypes.c_uint32),
                ("fine_cluster_count", ctypes.c_uint32), ("filter_nsq", ctypes.c_uint32),
                ("total_point_count", ctypes.c_uint32), ("whether_norm", ctypes.c_int32)]


class SearchParams(ctypes.Structure):
    _fields_ = [("topk_max", ctypes.c_uint32), ("search_coarse_count", ctypes.c_uint32),
                ("filter_topk", ctypes.c_uint32), ("radius_rate", ctypes.c_double),
                ("replay_all", ctypes.c_uint32), ("precision", ctypes.c_uint32)]


class QueryParams(ctypes.Structure):
    _fields_ = [("topk", ctypes.c_void_p), ("search_coarse_count", ctypes.c_void_p),
                ("fine_budget", ctypes.c_void_p)]


class ProbeShape(ctypes.Structure):
    _fields_ = [("coarse_features", ctypes.c_uint32), ("fine_features", ctypes.c_uint32),
                ("hidden", ctypes.c_uint32)]


class AdamParams(ctypes.Structure):
    _fields_ = [("lr", ctypes.c_float), ("beta1", ctypes.c_float), ("beta2", ctypes.c_float),
                ("eps", ctypes.c_float)]


class BudgetParams(ctypes.Structure):
    _fields_ = [("margin", ctypes.c_float), ("min_budget", ctypes.c_uint32), ("fallback_frac", ctypes.c_float)]


_lib = None


def load():
    """Load libpuckgpu.so (built by puck_amd/build.py); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python puck_amd/build.py` "
                           "(the engine has no CPU fallback)")
    # torch ships its own HIP/HSA runtime under the same sonames.  Loading it
    # first makes libpuckgpu.so bind to it; loading /opt/rocm's first would put
    # two HSA runtimes in the process, and torch's would then find no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, u32, i32, c_int = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32, ctypes.c_int
    L.puckgpu_create.argtypes = [c_int, ctypes.POINTER(vp)]
    L.puckgpu_create.restype = c_int
    L.puckgpu_destroy.argtypes = [vp]
    L.puckgpu_destroy.restype = None
    L.puckgpu_last_error.argtypes = [vp]
    L.puckgpu_last_error.restype = ctypes.c_char_p
    L.puckgpu_filter_topk_max.argtypes = []
    L.puckgpu_filter_topk_max.restype = u32
    L.puckgpu_index_load.argtypes = [vp, ctypes.POINTER(IndexParams), vp, vp, vp, vp, vp, vp, c_int,
                                     ctypes.POINTER(vp)]
    L.puckgpu_index_load.restype = c_int
    L.puckgpu_index_free.argtypes = [vp]
    L.puckgpu_index_free.restype = None
    L.puckgpu_index_export.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.puckgpu_index_export.restype = c_int
    L.puckgpu_search.argtypes = [vp, vp, ctypes.POINTER(SearchParams), u32, vp, u32, vp, vp, vp, c_int]
    L.puckgpu_search.restype = c_int
    L.puckgpu_search_batch.argtypes = [vp, vp, ctypes.POINTER(SearchParams), ctypes.POINTER(QueryParams), u32,
                                       vp, vp, vp, vp, c_int]
    L.puckgpu_search_batch.restype = c_int
    L.puckgpu_last_filter.argtypes = [vp, u32, u32, vp, vp, vp]
    L.puckgpu_last_filter.restype = c_int
    L.puckgpu_last_stage_ms.argtypes = [vp, vp, vp]
    L.puckgpu_last_stage_ms.restype = c_int
    L.puckgpu_lut.argtypes = [vp, u32, u32, vp, vp, u32, vp, c_int]
    L.puckgpu_lut.restype = c_int
    L.puckgpu_assign_cells.argtypes = [vp, vp, vp, u32, u32, u32, u32, vp, vp, vp, c_int]
    L.puckgpu_assign_cells.restype = c_int
    L.puckgpu_index_attach_raw_u8.argtypes = [vp, vp, vp, c_int]
    L.puckgpu_index_attach_raw_u8.restype = c_int
    L.puckgpu_server_tuned.argtypes = [vp, u32, vp, vp]
    L.puckgpu_server_tuned.restype = c_int
    L.puckgpu_set_rerank_rows.argtypes = [vp, u32]
    L.puckgpu_set_rerank_rows.restype = c_int
    L.puckgpu_exact_search.argtypes = [vp, vp, c_int, ctypes.c_uint64, u32, vp, u32, u32, u32, vp, vp, c_int]
    L.puckgpu_exact_search.restype = c_int
    L.puckgpu_set_progress.argtypes = [vp, PROGRESS_FN, vp]
    L.puckgpu_set_progress.restype = c_int
    L.puckgpu_set_assign_mode.argtypes = [vp, c_int]
    L.puckgpu_set_assign_mode.restype = c_int
    L.puckgpu_assign_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
    L.puckgpu_assign_stats.restype = c_int
    L.puckgpu_encode_filter.argtypes = [vp, vp, vp, u32, u32, u32, vp, u32, u32, vp, vp, vp, c_int]
    L.puckgpu_encode_filter.restype = c_int
    L.puckgpu_search_noalloc.argtypes = [vp, vp, ctypes.POINTER(SearchParams), u32, vp, u32, vp, vp, vp]
    L.puckgpu_search_noalloc.restype = c_int
    L.puckgpu_index_attach_pq.argtypes = [vp, vp, u32, vp, vp, c_int]
    L.puckgpu_index_attach_pq.restype = c_int
    L.puckgpu_encode_pq.argtypes = [vp, vp, vp, u32, u32, u32, vp, u32, vp, vp, u32, u32, vp, vp, vp, c_int]
    L.puckgpu_encode_pq.restype = c_int
    L.puckgpu_index_insert.argtypes = [vp, vp, u32, vp, vp, vp, vp, vp, vp, c_int]
    L.puckgpu_index_insert.restype = c_int
    L.puckgpu_kmeans.argtypes = [vp, u32, u32, u32, vp, c_int, c_int, u32, c_int, vp, vp, vp, vp, vp, c_int]
    L.puckgpu_kmeans.restype = c_int
    L.puckgpu_train_centers.argtypes = [vp, u32, u32, vp, u32, u32, c_int, c_int, c_int, u32, c_int, vp, vp, vp, vp,
                                        vp, c_int]
    L.puckgpu_train_centers.restype = c_int
    L.puckgpu_index_version.argtypes = [vp]
    L.puckgpu_index_version.restype = ctypes.c_uint64
    L.puckgpu_kmeans_assign.argtypes = [vp, u32, vp, u32, u32, vp, vp, vp, c_int]
    L.puckgpu_kmeans_assign.restype = c_int
    L.puckgpu_probe_create.argtypes = [vp, u32, ctypes.POINTER(ProbeShape), ctypes.POINTER(vp)]
    L.puckgpu_probe_create.restype = c_int
    L.puckgpu_probe_free.argtypes = [vp]
    L.puckgpu_probe_free.restype = None
    L.puckgpu_probe_param_count.argtypes = [vp]
    L.puckgpu_probe_param_count.restype = u32
    L.puckgpu_probe_set_params.argtypes = [vp, vp, vp, c_int]
    L.puckgpu_probe_set_params.restype = c_int
    L.puckgpu_probe_get_params.argtypes = [vp, vp, vp, c_int]
    L.puckgpu_probe_get_params.restype = c_int
    L.puckgpu_probe_features.argtypes = [vp, vp, vp, u32, vp, vp, vp, c_int]
    L.puckgpu_probe_features.restype = c_int
    L.puckgpu_probe_train_step.argtypes = [vp, vp, u32, vp, vp, ctypes.POINTER(AdamParams),
                                           ctypes.POINTER(ctypes.c_float), vp, c_int]
    L.puckgpu_probe_train_step.restype = c_int
    L.puckgpu_probe_predict.argtypes = [vp, vp, u32, vp, vp, c_int]
    L.puckgpu_probe_predict.restype = c_int
    L.puckgpu_search_learned.argtypes = [vp, vp, vp, ctypes.POINTER(SearchParams), ctypes.POINTER(BudgetParams),
                                         ctypes.POINTER(QueryParams), u32, vp, vp, vp, vp, vp, c_int]
    L.puckgpu_search_learned.restype = c_int
    L.puckgpu_budget_ranks.argtypes = [vp, vp, u32, u32, vp, u32, vp, vp, c_int]
    L.puckgpu_budget_ranks.restype = c_int
    L.puckgpu_server_create.argtypes = [vp, vp, vp, ctypes.POINTER(SearchParams), ctypes.POINTER(BudgetParams), u32,
                                        ctypes.POINTER(vp)]
    L.puckgpu_server_create.restype = c_int
    L.puckgpu_server_free.argtypes = [vp]
    L.puckgpu_server_free.restype = None
    L.puckgpu_server_buckets.argtypes = [vp]
    L.puckgpu_server_buckets.restype = u32
    L.puckgpu_server_search.argtypes = [vp, u32, vp, vp, vp, vp, vp, vp, c_int]
    L.puckgpu_server_search.restype = c_int
    L.puckgpu_server_tune.argtypes = [vp, vp, u32, vp, vp]
    L.puckgpu_server_tune.restype = c_int
    L.puckgpu_last_scan_ms.argtypes = [vp, vp]
    L.puckgpu_last_scan_ms.restype = c_int
    L.puckgpu_center_distances.argtypes = [vp, vp, u32, u32, vp, vp, vp, c_int]
    L.puckgpu_center_distances.restype = c_int
    del i32
    _lib = L
    return L
