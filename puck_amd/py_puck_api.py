"""The reference's Python entry points, served by the MI355X engine.

Mirrors pyapi_wrapper/py_puck_api.i + py_api_wrapper.{h,cpp}:

    update_gflag(key, value)                 py_api_wrapper.cpp:33-35
    PySearcher().init() -> 0 / -1            py_api_wrapper.cpp:78-108
    PySearcher().build(total_cnt) -> 0 / -1  py_api_wrapper.cpp:43-76
    PySearcher().search(n, query, topk, distance, labels) -> 0
                                             py_api_wrapper.cpp:154-171

`search` writes into caller-owned arrays with row stride `topk` (numpy
float32 / uint32, or CUDA tensors); a request the reference would reject
(topk > FLAGS_topk) leaves them untouched, as PuckIndex::search returns -1
before touching its response (puck_index.cpp:560-563) and PySearcher ignores
per-query return codes.  PUCK indexes with ip2cos = 0 are served, ranked by
exact L2 over the raw vectors (whether_pq = false) or by the second-stage PQ
(whether_pq = true); other indexes make init() return -1.
"""
import sys

import numpy as np

from . import flags as FL
from . import index_io


def update_gflag(key, value):
    FL.update_gflag(key, str(value))


def _err(msg):
    print(f"[puck_amd] ERROR {msg}", file=sys.stderr, flush=True)


class PySearcher:
    def __init__(self):
        self._ctx = None
        self._index = None
        self._loaded_path = None
        self._conf = None
        self._dim = 0
        self._params = None

    # -- IndexConf + adaptive_search_param (index_conf.cpp:28-95, :152-198)
    def _search_params(self):
        topk = FL.get("topk")
        filter_topk = FL.get("filter_topk") if not FL.is_default("filter_topk") else topk * 11
        rr = FL.get("radius_rate")
        if rr < 1.0:
            _err("radius_rate should >= 1.0, should update the search params")
            return None
        if filter_topk < topk:
            _err("filter_topk should >= topk, should update the search params")
            return None
        if filter_topk > self._ctx.L.puckgpu_filter_topk_max():
            _err(f"filter_topk {filter_topk} above the engine's maximum {self._ctx.L.puckgpu_filter_topk_max()}")
            return None
        scc = min(FL.get("search_coarse_count"), self._conf["coarse_cluster_count"])
        return dict(topk=topk, filter_topk=filter_topk, radius_rate=rr, search_coarse_count=scc)

    def init(self):
        from .engine import Context, GpuIndex, PuckGpuError
        path = FL.get("index_path")
        try:
            conf = index_io.read_index_conf(index_io.paths(path, FL)["index"])
        except (OSError, ValueError) as e:
            _err(f"model file read error: {e}")
            return -1
        if conf["index_type"] != index_io.INDEX_PUCK:
            _err(f"index_type {conf['index_type']} is not served by the GPU engine (PUCK only)")
            return -1
        if conf["ip2cos"]:
            _err("ip2cos indexes (inner product stored as an extra dimension) are not served by the GPU engine")
            return -1
        self._conf = conf
        if self._loaded_path != path or self._index is None:
            try:
                conf, arr = index_io.read_index(path, FL)
            except (OSError, ValueError) as e:
                _err(f"load index failed: {e}")
                return -1
            if self._ctx is None:
                self._ctx = Context(0)
            try:
                self._index = GpuIndex(self._ctx, arr["coarse"], arr["fine"], arr["codebook"], arr["cell_assign"],
                                       arr["records"], arr["features"], conf["whether_norm"],
                                       pq_codebook=arr["pq_codebook"], pq_records=arr["pq_records"])
            except PuckGpuError as e:
                _err(str(e))
                return -1
            self._loaded_path = path
        self._params = self._search_params()
        if self._params is None:
            return -1
        self._dim = conf["feature_dim"]
        return 0

    def build(self, total_cnt):
        """Train (coarse, fine, filter codebook), assign and encode the
        all_data.feat.bin of FLAGS_index_path on the GPU; write the index files."""
        import torch
        from . import train
        from .engine import Context
        path = FL.get("index_path")
        p = index_io.paths(path, FL)
        dim = FL.get("feature_dim")
        try:
            x = index_io.read_fvecs(p["feature"], dim, total_cnt)
        except (OSError, ValueError) as e:
            _err(f"read features failed: {e}")
            return -1
        nsq = FL.get("filter_nsq")
        if nsq == 0 or dim % nsq or dim // nsq not in (1, 2, 4):
            _err(f"filter_nsq {nsq} must divide feature_dim {dim} into sub-vectors of 1, 2 or 4")
            return -1
        whether_norm = FL.get("whether_norm")
        xd = torch.from_numpy(x).cuda()
        if whether_norm:
            # the reference's callers write normalised features (the Python
            # wrapper divides by the norm before writing all_data.feat.bin);
            # here every row is scaled by float32(1 / norm), the norm taken in
            # float64
            alpha = (1.0 / xd.double().square().sum(1).sqrt().clamp_min(1e-30)).float()
            xd = xd * alpha[:, None]
        if self._ctx is None:
            self._ctx = Context(0)
        pq_nsq = None
        if FL.get("whether_pq"):
            # IndexConf: nsq left at its default means one subspace per dimension (index_conf.cpp:39-49)
            pq_nsq = dim if FL.is_default("nsq") else FL.get("nsq")
            if pq_nsq == 0 or dim % pq_nsq or dim // pq_nsq not in (1, 2, 4) or pq_nsq > 128:
                _err(f"nsq {pq_nsq} must divide feature_dim {dim} into sub-vectors of 1, 2 or 4 (at most 128)")
                return -1
        ix = train.build(self._ctx, xd.contiguous(), FL.get("coarse_cluster_count"), FL.get("fine_cluster_count"),
                         nsq, whether_norm, train_n=min(FL.get("train_points_count"), total_cnt),
                         rounds=FL.get("kmeans_iterations_count"), pq_nsq=pq_nsq)
        h = {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in ix.items()}
        index_io.write_index(path, FL, h["coarse"], h["fine"], h["codebook"], h["cell_assign"].view(np.uint32),
                             h["records"], None if pq_nsq else (x if not whether_norm else h["raw"]), whether_norm,
                             pq_codebook=h.get("pq_codebook"), pq_records=h.get("pq_records"))
        return 0

    def search(self, n, query_fea, topk, distance, labels):
        """0 as PySearcher::search returns (py_api_wrapper.cpp:169), also when
        every query is refused; -1 only for a search before init(), where the
        reference has no index to call."""
        if self._index is None or self._params is None:
            _err("search before a successful init()")
            return -1
        if topk > self._params["topk"] or topk == 0:
            _err(f"topk should <= topk, topk = {self._params['topk']}")
            return 0
        if n == 0:
            return 0
        q = query_fea
        if isinstance(q, np.ndarray):
            q = np.ascontiguousarray(q, np.float32).reshape(n, self._conf["feature_dim"])
            D = np.empty((n, topk), np.float32)
            I = np.empty((n, topk), np.uint32)
            R = np.empty(n, np.uint32)
            self._index.search(q, topk, self._params["search_coarse_count"], self._params["filter_topk"],
                               topk_max=self._params["topk"], radius_rate=self._params["radius_rate"],
                               out=(D, I, R))
            distance.reshape(-1)[: n * topk] = D.reshape(-1)
            labels.reshape(-1)[: n * topk] = I.reshape(-1)
        else:
            # as the numpy path: the queries are read as float32; the outputs
            # must already be float32 / 32-bit integer tensors (checked by the index)
            import torch
            q = q.reshape(n, self._conf["feature_dim"]).to(torch.float32).contiguous()
            self._index.search(q, topk, self._params["search_coarse_count"], self._params["filter_topk"],
                               topk_max=self._params["topk"], radius_rate=self._params["radius_rate"],
                               out=(distance.view(n, topk), labels.view(n, topk), None))
        return 0

    def show(self):
        print("1")
