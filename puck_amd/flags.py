"""Process-global search / build knobs with the reference's gflag names and
defaults (puck/gflags/puck_gflags.cpp:26-81, puck/hierarchical_cluster/
hierarchical_cluster_index.cpp:70-76, puck/puck/puck_index.cpp:40-41).

`update_gflag(key, value)` mirrors py_puck_api.update_gflag
(pyapi_wrapper/py_api_wrapper.cpp:33-35): values are strings parsed to the
flag's type; unknown keys return False, as SetCommandLineOption does.
"""
import os

_NPROC = os.cpu_count() or 1

# name -> (type, default); defaults computed at static-init time in the
# reference use the *default* feature_dim (256) and topk (100).
_SPEC = {
    "index_path": (str, "puck_index"),
    "index_file_name": (str, "index.dat"),
    "feature_file_name": (str, "all_data.feat.bin"),
    "coarse_codebook_file_name": (str, "coarse.dat"),
    "fine_codebook_file_name": (str, "fine.dat"),
    "cell_assign_file_name": (str, "cell_assign.dat"),
    "label_file_name": (str, "all_data.url"),
    "filter_codebook_file_name": (str, "filter_codebook.dat"),
    "filter_data_file_name": (str, "filter_data.dat"),
    "pq_codebook_file_name": (str, "learn_codebooks.dat"),
    "pq_data_file_name": (str, "pq_data.dat"),
    "feature_dim": (int, 256),
    "whether_norm": (bool, True),
    "coarse_cluster_count": (int, 2000),
    "fine_cluster_count": (int, 2000),
    "threads_count": (int, _NPROC),
    "ip2cos": (int, 0),
    "whether_pq": (bool, True),
    "nsq": (int, 256),
    "filter_nsq": (int, 256 // 4),
    "context_initial_pool_size": (int, _NPROC),
    "search_coarse_count": (int, 200),
    "topk": (int, 100),
    "filter_topk": (int, 100 * 11),
    "radius_rate": (float, 1.0),
    "kmeans_iterations_count": (int, 10),
    "train_points_count": (int, 5000000),
    "pq_train_points_count": (int, 1000000),
    "index_type": (int, 1),
}

_values = {k: v for k, (_, v) in _SPEC.items()}
_set = set()


def _parse(t, s):
    if t is bool:
        v = str(s).strip().lower()
        if v in ("1", "true", "t", "yes", "y"):
            return True
        if v in ("0", "false", "f", "no", "n"):
            return False
        raise ValueError(s)
    return t(s)


def update_gflag(key, value):
    """SetCommandLineOption: returns True when the flag exists and parses."""
    if key not in _SPEC:
        return False
    try:
        _values[key] = _parse(_SPEC[key][0], value)
    except (TypeError, ValueError):
        return False
    _set.add(key)
    return True


def get(key):
    return _values[key]


def is_default(key):
    return key not in _set


def reset():
    _values.clear()
    _values.update({k: v for k, (_, v) in _SPEC.items()})
    _set.clear()
