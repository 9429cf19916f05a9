"""Reader / writer of the reference's on-disk index directory.

Formats (all little-endian):
* index.dat -- size_t length, then the packed IndexConf fields written by
  HierarchicalClusterIndex::save_model_config
  (hierarchical_cluster_index.cpp:522-543): int32 index_type, u32 dim, bool
  whether_norm, u32 N, bool whether_filter, u32 filter_nsq, u32 ks, bool
  whether_pq, u32 nsq, u32 ks, u32 ip2cos, u32 C, u32 F (43 bytes).
* fvecs (coarse.dat, fine.dat, filter_codebook.dat, all_data.feat.bin) --
  per row an int32 dimension then the floats (:109-176).
* cell_assign.dat -- N x u32 cell id (coarse * F + fine), local order.
* filter_data.dat -- N x (f32 offset + filter_nsq code bytes), local order
  (quantization.cpp:276-294).
* whether_pq = true: learn_codebooks.dat (the second-stage codebook, fvecs
  rows of dim / nsq floats) and pq_data.dat (N x (f32 offset + nsq code
  bytes), local order); all_data.feat.bin is not read (the reference does not
  load raw vectors for such an index, puck_index.cpp:198-227).
"""
import os
import struct

import numpy as np

INDEX_HIERARCHICAL_CLUSTER, INDEX_PUCK, INDEX_TINKER = 0, 1, 2

_CONF = struct.Struct("<iI?I?II?IIIII")
assert _CONF.size == 43


def write_index_conf(path, dim, whether_norm, N, filter_nsq, C, F, whether_pq=False, nsq=None, ip2cos=0,
                     index_type=INDEX_PUCK, whether_filter=True, ks=256):
    nsq = dim if nsq is None else nsq
    body = _CONF.pack(index_type, dim, bool(whether_norm), N, bool(whether_filter), filter_nsq, ks,
                      bool(whether_pq), nsq, ks, ip2cos, C, F)
    with open(path, "wb") as f:
        f.write(struct.pack("<Q", len(body)))
        f.write(body)


def read_index_conf(path):
    data = open(path, "rb").read()
    if len(data) < 8 + _CONF.size:
        raise ValueError(f"{path}: too short for an index configuration")
    (n,) = struct.unpack_from("<Q", data, 0)
    f = _CONF.unpack_from(data, 8)
    keys = ["index_type", "feature_dim", "whether_norm", "total_point_count", "whether_filter", "filter_nsq",
            "ks", "whether_pq", "nsq", "ks2", "ip2cos", "coarse_cluster_count", "fine_cluster_count"]
    conf = dict(zip(keys, f))
    conf["part_size"] = n
    return conf


def write_fvecs(path, x):
    x = np.ascontiguousarray(x, np.float32)
    n, d = x.shape
    rows = np.empty((n, d + 1), np.float32)
    rows[:, 0] = np.frombuffer(np.int32(d).tobytes(), np.float32)[0]
    rows[:, 1:] = x
    rows.tofile(path)


def read_fvecs(path, dim, n=None):
    """Rows of `dim` floats; checks every row's length prefix (read_fvec_format)."""
    raw = np.fromfile(path, dtype=np.int32)
    if raw.size % (dim + 1):
        raise ValueError(f"{path}: size is not a multiple of (1 + {dim}) words")
    rows = raw.reshape(-1, dim + 1)
    if n is not None and rows.shape[0] != n:
        raise ValueError(f"{path}: {rows.shape[0]} rows, expected {n}")
    if not np.all(rows[:, 0] == dim):
        raise ValueError(f"{path}: row dimension differs from {dim}")
    return np.ascontiguousarray(rows[:, 1:]).view(np.float32)


def paths(index_path, flags):
    j = lambda k: os.path.join(index_path, flags.get(k))
    return dict(index=j("index_file_name"), coarse=j("coarse_codebook_file_name"),
                fine=j("fine_codebook_file_name"), cell_assign=j("cell_assign_file_name"),
                filter_codebook=j("filter_codebook_file_name"), filter_data=j("filter_data_file_name"),
                feature=j("feature_file_name"), pq_codebook=j("pq_codebook_file_name"),
                pq_data=j("pq_data_file_name"))


def write_index(index_path, flags, coarse, fine, codebook, cell_assign, records, features, whether_norm,
                pq_codebook=None, pq_records=None):
    """Writes a PUCK index directory the reference can load: whether_pq =
    false with the raw features, or whether_pq = true with the second-stage
    codebook and records (features may then be None)."""
    os.makedirs(index_path, exist_ok=True)
    p = paths(index_path, flags)
    C, dim = coarse.shape
    F = fine.shape[0]
    N = cell_assign.shape[0]
    nsq = records.shape[1] - 4
    pq = pq_records is not None
    pnsq = pq_records.shape[1] - 4 if pq else dim
    write_index_conf(p["index"], dim, whether_norm, N, nsq, C, F, whether_pq=pq, nsq=pnsq)
    write_fvecs(p["coarse"], coarse)
    write_fvecs(p["fine"], fine)
    write_fvecs(p["filter_codebook"], np.asarray(codebook, np.float32).reshape(nsq * 256, dim // nsq))
    np.ascontiguousarray(cell_assign, np.uint32).tofile(p["cell_assign"])
    np.ascontiguousarray(records, np.uint8).tofile(p["filter_data"])
    if pq:
        write_fvecs(p["pq_codebook"], np.asarray(pq_codebook, np.float32).reshape(pnsq * 256, dim // pnsq))
        np.ascontiguousarray(pq_records, np.uint8).tofile(p["pq_data"])
    if features is not None:
        write_fvecs(p["feature"], features)


def read_index(index_path, flags):
    """Returns (conf, arrays) of a PUCK index directory; raises ValueError on
    length mismatches (check_file_length_info, hierarchical_cluster_index.cpp:93-106)."""
    p = paths(index_path, flags)
    conf = read_index_conf(p["index"])
    d, C, F = conf["feature_dim"], conf["coarse_cluster_count"], conf["fine_cluster_count"]
    N, nsq = conf["total_point_count"], conf["filter_nsq"]
    if nsq == 0 or d % nsq:
        raise ValueError(f"filter_nsq {nsq} does not divide dim {d}")
    lsq = d // nsq
    coarse = read_fvecs(p["coarse"], d, C)
    fine = read_fvecs(p["fine"], d, F)
    codebook = read_fvecs(p["filter_codebook"], lsq, nsq * 256).reshape(-1)
    cell = np.fromfile(p["cell_assign"], dtype=np.uint32)
    if cell.size != N:
        raise ValueError(f"{p['cell_assign']}: {cell.size} entries, expected {N}")
    rec = np.fromfile(p["filter_data"], dtype=np.uint8)
    if rec.size != N * (4 + nsq):
        raise ValueError(f"{p['filter_data']}: {rec.size} bytes, expected {N * (4 + nsq)}")
    out = dict(coarse=coarse, fine=fine, codebook=codebook, cell_assign=cell, records=rec.reshape(N, 4 + nsq),
               features=None, pq_codebook=None, pq_records=None)
    if conf["whether_pq"]:
        pnsq = conf["nsq"]
        if pnsq == 0 or d % pnsq:
            raise ValueError(f"nsq {pnsq} does not divide dim {d}")
        out["pq_codebook"] = read_fvecs(p["pq_codebook"], d // pnsq, pnsq * 256).reshape(-1)
        pr = np.fromfile(p["pq_data"], dtype=np.uint8)
        if pr.size != N * (4 + pnsq):
            raise ValueError(f"{p['pq_data']}: {pr.size} bytes, expected {N * (4 + pnsq)}")
        out["pq_records"] = pr.reshape(N, 4 + pnsq)
    else:
        out["features"] = read_fvecs(p["feature"], d, N)
    return conf, out
