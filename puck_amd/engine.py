"""Python handles over the C ABI: a device context and a loaded index.

Arrays may be numpy arrays (host; copied in and out by the library) or torch
CUDA tensors (device; used in place).  Every call goes through libpuckgpu.so.
"""
import ctypes

import numpy as np

from . import _lib


class PuckGpuError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"libpuckgpu error {rc}: {msg}")
        self.rc = rc


def _is_dev(a):
    return a is not None and not isinstance(a, np.ndarray) and getattr(a, "is_cuda", False)


def _ptr(a, dtype=None):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if dtype is not None and a.dtype != dtype:
            raise TypeError(f"expected {dtype}, got {a.dtype}")
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data_as(ctypes.c_void_p)
    if not a.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return ctypes.c_void_p(a.data_ptr())


_F32 = ("float32",)
_U32 = ("uint32", "int32")
_U8 = ("uint8",)


def _dtype_name(a):
    return str(a.dtype).rsplit(".", 1)[-1]


def _check(a, name, kinds, shape=None, min_shape=None, device=None):
    """Refuses an argument before any library call: its dtype is one of
    `kinds`, it is contiguous, its shape equals `shape` (None entries are
    free) or is at least `min_shape` in every dimension, and a device tensor
    lives on `device` (the context's ordinal).  A None argument passes."""
    if a is None:
        return
    dn = _dtype_name(a)
    if dn not in kinds:
        raise TypeError(f"{name}: expected dtype {' or '.join(kinds)}, got {dn}")
    sh = tuple(int(v) for v in a.shape)
    if shape is not None:
        if len(sh) != len(shape) or any(w is not None and v != w for v, w in zip(sh, shape)):
            raise ValueError(f"{name}: expected shape {list(shape)}, got {list(sh)}")
    if min_shape is not None:
        if len(sh) != len(min_shape) or any(v < w for v, w in zip(sh, min_shape)):
            raise ValueError(f"{name}: expected at least shape {list(min_shape)}, got {list(sh)}")
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError(f"{name}: array must be C-contiguous")
    else:
        if not a.is_contiguous():
            raise ValueError(f"{name}: tensor must be contiguous")
        if _is_dev(a) and device is not None:
            idx = getattr(a.device, "index", None)
            if idx is not None and int(idx) != int(device):
                raise ValueError(f"{name}: tensor on device {idx}, context on device {device}")


def _mode(*arrays):
    devs = [_is_dev(a) for a in arrays if a is not None]
    if any(devs) and not all(devs):
        raise ValueError("mix of host and device arrays in one call")
    return 1 if devs and devs[0] else 0


class Context:
    def __init__(self, device=0):
        self.L = _lib.load()
        h = ctypes.c_void_p()
        rc = self.L.puckgpu_create(int(device), ctypes.byref(h))
        if rc != 0:
            raise PuckGpuError(rc, f"puckgpu_create(device={device}) failed")
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            self.L.puckgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc != 0:
            msg = self.L.puckgpu_last_error(self.h)
            raise PuckGpuError(rc, msg.decode() if msg else "")
        return rc

    def sync_torch(self, *arrays):
        if any(_is_dev(a) for a in arrays if a is not None):
            import torch
            torch.cuda.synchronize()

    # ---- test hooks / build entry points -------------------------------
    def lut(self, queries, codebook, nsq):
        n, dim = queries.shape
        _check(queries, "queries", _F32, device=self.device)
        _check(codebook, "codebook", _F32, shape=(nsq * 256 * (dim // max(nsq, 1)),), device=self.device)
        out = np.zeros((n, nsq * 256), np.float32) if not _is_dev(queries) else None
        if out is None:
            import torch
            out = torch.empty((n, nsq * 256), dtype=torch.float32, device=queries.device)
        m = _mode(queries, codebook, out)
        self.sync_torch(queries)
        self.check(self.L.puckgpu_lut(self.h, n, dim, _ptr(queries), _ptr(codebook), nsq, _ptr(out), m))
        return out

    def assign_cells(self, coarse, fine, x, with_dist=False):
        C, dim = coarse.shape
        F = fine.shape[0]
        n = x.shape[0]
        _check(coarse, "coarse", _F32, device=self.device)
        _check(fine, "fine", _F32, shape=(F, dim), device=self.device)
        _check(x, "x", _F32, shape=(n, dim), device=self.device)
        if _is_dev(x):
            import torch
            cell = torch.empty(n, dtype=torch.int32, device=x.device)
            dist = torch.empty(n, dtype=torch.float32, device=x.device) if with_dist else None
        else:
            cell = np.zeros(n, np.uint32)
            dist = np.zeros(n, np.float32) if with_dist else None
        m = _mode(coarse, fine, x, cell)
        self.sync_torch(x, coarse)
        self.check(self.L.puckgpu_assign_cells(self.h, _ptr(coarse), _ptr(fine), C, F, dim, n, _ptr(x),
                                               _ptr(cell), _ptr(dist), m))
        return (cell, dist) if with_dist else cell

    def exact_search(self, x, q, k, extra=8):
        """Exact squared-L2 top-k of q [nq, d] float32 in x [N, d] (float32, or
        uint8 integer features), ordered by (distance, id): (ids uint32
        [nq, k], dists float64 [nq, k]); torch CUDA tensors or numpy arrays."""
        N, d = int(x.shape[0]), int(x.shape[1])
        nq = int(q.shape[0])
        u8 = _dtype_name(x) == "uint8"
        _check(x, "x", _U8 if u8 else _F32, device=self.device)
        _check(q, "q", _F32, shape=(nq, d), device=self.device)
        if _is_dev(q):
            import torch
            ids = torch.empty((nq, k), dtype=torch.int32, device=q.device)
            dist = torch.empty((nq, k), dtype=torch.float64, device=q.device)
        else:
            ids = np.zeros((nq, k), np.uint32)
            dist = np.zeros((nq, k), np.float64)
        m = _mode(x, q, ids)
        self.sync_torch(x, q)
        self.check(self.L.puckgpu_exact_search(self.h, _ptr(x), int(u8), N, d, _ptr(q), nq, int(k), int(extra),
                                               _ptr(ids), _ptr(dist), m))
        return ids, dist

    def set_progress(self, fn):
        """fn(message: str) receives the library's progress lines (None: off)."""
        if fn is None:
            self._progress = None
            self.check(self.L.puckgpu_set_progress(self.h, _lib.PROGRESS_FN(), None))
            return
        self._progress = _lib.PROGRESS_FN(lambda m, u: fn(m.decode()))
        self.check(self.L.puckgpu_set_progress(self.h, self._progress, None))

    def set_rerank_rows(self, rows):
        """Rows in flight per 8-lane group of the exact rerank (1, 2 or 4)."""
        self.check(self.L.puckgpu_set_rerank_rows(self.h, int(rows)))

    def set_assign_mode(self, mode):
        """0: bounded assignment with certificate (default), 1: the sequential
        walk for every point, 2: bounded, then every point walked again."""
        self.check(self.L.puckgpu_set_assign_mode(self.h, int(mode)))

    def assign_stats(self):
        """(points assigned, points walked sequentially) since creation."""
        a, b = ctypes.c_uint64(0), ctypes.c_uint64(0)
        self.check(self.L.puckgpu_assign_stats(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def encode_filter(self, coarse, fine, codebook, nsq, x, cell):
        C, dim = coarse.shape
        F = fine.shape[0]
        n = x.shape[0]
        _check(coarse, "coarse", _F32, device=self.device)
        _check(fine, "fine", _F32, shape=(F, dim), device=self.device)
        if nsq == 0 or dim % nsq:
            raise ValueError(f"filter_nsq {nsq} must divide dim {dim}")
        _check(codebook, "codebook", _F32, device=self.device)
        if int(np.prod(codebook.shape)) != nsq * 256 * (dim // nsq):
            raise ValueError(f"codebook: expected {nsq * 256 * (dim // nsq)} floats")
        _check(x, "x", _F32, shape=(n, dim), device=self.device)
        _check(cell, "cell", _U32, shape=(n,), device=self.device)
        if _is_dev(x):
            import torch
            rec = torch.empty((n, 4 + nsq), dtype=torch.uint8, device=x.device)
        else:
            rec = np.zeros((n, 4 + nsq), np.uint8)
        m = _mode(coarse, fine, codebook, x, cell, rec)
        self.sync_torch(x, cell)
        self.check(self.L.puckgpu_encode_filter(self.h, _ptr(coarse), _ptr(fine), C, F, dim, _ptr(codebook),
                                                nsq, n, _ptr(x), _ptr(cell), _ptr(rec), m))
        return rec

    def encode_pq(self, coarse, fine, codebook, nsq, filter_records, pq_codebook, pq_nsq, x, cell):
        """Second-stage PQ records [n, 4 + pq_nsq] (whether_pq = true) of x,
        given its cells and filter records (include/puck_gpu.h)."""
        C, dim = coarse.shape
        F = fine.shape[0]
        n = x.shape[0]
        _check(coarse, "coarse", _F32, device=self.device)
        _check(fine, "fine", _F32, shape=(F, dim), device=self.device)
        if nsq == 0 or dim % nsq or pq_nsq == 0 or dim % pq_nsq:
            raise ValueError(f"filter_nsq {nsq} and pq_nsq {pq_nsq} must divide dim {dim}")
        _check(codebook, "codebook", _F32, device=self.device)
        _check(pq_codebook, "pq_codebook", _F32, device=self.device)
        if int(np.prod(pq_codebook.shape)) != pq_nsq * 256 * (dim // pq_nsq):
            raise ValueError(f"pq_codebook: expected {pq_nsq * 256 * (dim // pq_nsq)} floats")
        _check(filter_records, "filter_records", _U8, shape=(n, 4 + nsq), device=self.device)
        _check(x, "x", _F32, shape=(n, dim), device=self.device)
        _check(cell, "cell", _U32, shape=(n,), device=self.device)
        if _is_dev(x):
            import torch
            rec = torch.empty((n, 4 + pq_nsq), dtype=torch.uint8, device=x.device)
        else:
            rec = np.zeros((n, 4 + pq_nsq), np.uint8)
        m = _mode(coarse, fine, codebook, filter_records, pq_codebook, x, cell, rec)
        self.sync_torch(x, cell, filter_records)
        self.check(self.L.puckgpu_encode_pq(self.h, _ptr(coarse), _ptr(fine), C, F, dim, _ptr(codebook), nsq,
                                            _ptr(filter_records), _ptr(pq_codebook), pq_nsq, n, _ptr(x), _ptr(cell),
                                            _ptr(rec), m))
        return rec

    def kmeans_assign(self, centroids, x, with_dist=False):
        k, dim = centroids.shape
        n = x.shape[0]
        _check(centroids, "centroids", _F32, device=self.device)
        _check(x, "x", _F32, shape=(n, dim), device=self.device)
        if _is_dev(x):
            import torch
            a = torch.empty(n, dtype=torch.int32, device=x.device)
            dd = torch.empty(n, dtype=torch.float32, device=x.device) if with_dist else None
        else:
            a = np.zeros(n, np.int32)
            dd = np.zeros(n, np.float32) if with_dist else None
        m = _mode(centroids, x, a)
        self.sync_torch(x, centroids)
        self.check(self.L.puckgpu_kmeans_assign(self.h, dim, _ptr(centroids), k, n, _ptr(x), _ptr(a), _ptr(dd), m))
        return (a, dd) if with_dist else a

    def kmeans(self, x, k, niter=30, init_pp=True, seed=0, mean_mode=0):
        """Kmeans::kmeans on the device (puckgpu_kmeans).  Returns
        (centroids [k, d], assign [n], mean error, iterations, splits); host
        arrays for host x, device tensors for device x."""
        n, dim = x.shape
        _check(x, "x", _F32, device=self.device)
        if _is_dev(x):
            import torch
            cent = torch.empty((k, dim), dtype=torch.float32, device=x.device)
            a = torch.empty(n, dtype=torch.int32, device=x.device)
        else:
            cent = np.zeros((k, dim), np.float32)
            a = np.zeros(n, np.int32)
        err = ctypes.c_float(0)
        it, sp = ctypes.c_int(0), ctypes.c_int(0)
        self.sync_torch(x)
        self.check(self.L.puckgpu_kmeans(self.h, dim, n, k, _ptr(x), int(niter), int(bool(init_pp)), int(seed),
                                         int(mean_mode), _ptr(cent), _ptr(a), ctypes.byref(err), ctypes.byref(it),
                                         ctypes.byref(sp), _mode(x, cent, a)))
        return cent, a, float(err.value), it.value, sp.value

    def train_centers(self, x, C, F, iterations=10, init_pp=True, seed=0, mean_mode=0, niter=0):
        """HierarchicalClusterIndex::train on the device
        (puckgpu_train_centers).  Returns (coarse [C, d], fine [F, d] by
        descending norm, errors [rounds, 2], rounds, final Lloyd limit)."""
        n, dim = x.shape
        _check(x, "x", _F32, device=self.device)
        if _is_dev(x):
            import torch
            co = torch.empty((C, dim), dtype=torch.float32, device=x.device)
            fi = torch.empty((F, dim), dtype=torch.float32, device=x.device)
        else:
            co = np.zeros((C, dim), np.float32)
            fi = np.zeros((F, dim), np.float32)
        tr = np.zeros(2 * max(int(iterations), 1), np.float32)
        r, ni = ctypes.c_int(0), ctypes.c_int(0)
        self.sync_torch(x)
        self.check(self.L.puckgpu_train_centers(self.h, dim, n, _ptr(x), C, F, int(iterations), int(niter),
                                                int(bool(init_pp)),
                                                int(seed), int(mean_mode), _ptr(co), _ptr(fi), _ptr(tr),
                                                ctypes.byref(r), ctypes.byref(ni), _mode(x, co, fi)))
        return co, fi, tr[:2 * r.value].reshape(-1, 2), r.value, ni.value


class GpuIndex:
    """A Puck index resident in HBM (memory order, derived arrays on device)."""

    def __init__(self, ctx, coarse, fine, codebook, cell_assign, records, features, whether_norm, pq_codebook=None,
                 pq_records=None):
        """features may be None for an index ranked by a second-stage PQ
        (pq_codebook [pq_nsq, 256, dim / pq_nsq] and pq_records [N, 4 + pq_nsq],
        whether_pq = true)."""
        self.ctx = ctx
        C, dim = coarse.shape
        F = fine.shape[0]
        N = cell_assign.shape[0]
        nsq = records.shape[1] - 4
        self.dim, self.C, self.F, self.nsq, self.N = dim, C, F, nsq, N
        self.whether_norm = bool(whether_norm)
        dv = ctx.device
        _check(coarse, "coarse", _F32, device=dv)
        _check(fine, "fine", _F32, shape=(F, dim), device=dv)
        _check(codebook, "codebook", _F32, device=dv)
        if nsq <= 0 or dim % nsq or int(np.prod(codebook.shape)) != nsq * 256 * (dim // nsq):
            raise ValueError(f"codebook: expected {nsq} x 256 x {dim // max(nsq, 1)} floats")
        _check(cell_assign, "cell_assign", _U32, shape=(N,), device=dv)
        _check(records, "records", _U8, shape=(N, 4 + nsq), device=dv)
        raw_u8 = features is not None and _dtype_name(features) == "uint8"
        if raw_u8:
            # integer features stored as uint8 (puckgpu_index_attach_raw_u8)
            if whether_norm:
                raise ValueError("uint8 features need whether_norm = False")
            _check(features, "features", _U8, shape=(N, dim), device=dv)
        else:
            _check(features, "features", _F32, shape=(N, dim), device=dv)
        if features is None and pq_records is None:
            raise ValueError("an index needs its raw features or a second-stage PQ to rank by")
        self.pq_nsq = 0
        if pq_records is not None:
            pq_nsq = int(pq_records.shape[1]) - 4
            _check(pq_records, "pq_records", _U8, shape=(N, 4 + pq_nsq), device=dv)
            _check(pq_codebook, "pq_codebook", _F32, device=dv)
            if pq_nsq <= 0 or dim % pq_nsq or int(np.prod(pq_codebook.shape)) != pq_nsq * 256 * (dim // pq_nsq):
                raise ValueError(f"pq_codebook: expected {pq_nsq} x 256 x {dim // max(pq_nsq, 1)} floats")
        p = _lib.IndexParams(dim, C, F, nsq, N, int(bool(whether_norm)))
        h = ctypes.c_void_p()
        m = _mode(coarse, fine, codebook, cell_assign, records, features, pq_codebook, pq_records)
        ctx.sync_torch(coarse, features, pq_records)
        ctx.check(ctx.L.puckgpu_index_load(ctx.h, ctypes.byref(p), _ptr(coarse), _ptr(fine), _ptr(codebook),
                                           _ptr(cell_assign), _ptr(records), None if raw_u8 else _ptr(features), m,
                                           ctypes.byref(h)))
        self.h = h
        self.raw_u8 = raw_u8
        if raw_u8:
            ctx.check(ctx.L.puckgpu_index_attach_raw_u8(ctx.h, self.h, _ptr(features), m))
        if pq_records is not None:
            ctx.check(ctx.L.puckgpu_index_attach_pq(ctx.h, self.h, pq_nsq, _ptr(pq_codebook), _ptr(pq_records), m))
            self.pq_nsq = pq_nsq

    def free(self):
        if getattr(self, "h", None):
            self.ctx.L.puckgpu_index_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def insert(self, x):
        """Realtime insert of the rows of x [n, dim] float32, in order
        (puckgpu_index_insert: RealtimeInsertPuckIndex::insert).  Returns
        dict(ids [n]: the local id given, or 0xffffffff for a refused
        zero-norm row; for the m accepted rows: cells [m], records
        [m, 4 + nsq], pq_records [m, 4 + pq_nsq] (whether_pq), features
        [m, dim] (the stored rows)).  Host arrays in, host arrays out; device
        tensors in, device tensors out (sized for every row, the first m
        filled)."""
        _check(x, "x", _F32, shape=(None, self.dim), device=self.ctx.device)
        n = int(x.shape[0])
        dev = _is_dev(x)
        pn = self.pq_nsq
        if dev:
            import torch
            mk = lambda shape, dt: torch.empty(shape, dtype=dt, device=x.device)
            ids, cells = mk((n,), torch.int32), mk((n,), torch.int32)
            rec, prec = mk((n, 4 + self.nsq), torch.uint8), mk((n, 4 + pn), torch.uint8)
            feats = mk((n, self.dim), torch.float32)
            self.ctx.sync_torch(x)
        else:
            ids, cells = np.zeros(max(n, 1), np.uint32), np.zeros(max(n, 1), np.uint32)
            rec, prec = np.zeros((max(n, 1), 4 + self.nsq), np.uint8), np.zeros((max(n, 1), 4 + pn), np.uint8)
            feats = np.zeros((max(n, 1), self.dim), np.float32)
        self.ctx.check(self.ctx.L.puckgpu_index_insert(self.ctx.h, self.h, n, _ptr(x), _ptr(ids), _ptr(cells),
                                                       _ptr(rec), _ptr(prec) if pn else None, _ptr(feats),
                                                       1 if dev else 0))
        if dev:
            m = int((ids != -1).sum().item())
            out = dict(ids=ids, cells=cells[:m], records=rec[:m], features=feats[:m])
        else:
            ids = ids[:n]
            m = int((ids != 0xffffffff).sum())
            out = dict(ids=ids, cells=cells[:m], records=rec[:m], features=feats[:m])
        if pn:
            out["pq_records"] = prec[:m]
        self.N += m
        return out

    def version(self):
        return int(self.ctx.L.puckgpu_index_version(self.h))

    def export(self):
        C, F, N = self.C, self.F, self.N
        out = dict(coarse_norms=np.zeros(C, np.float32), stationary=np.zeros(C * F, np.float32),
                   cell_start=np.zeros(C * F + 1, np.uint32), min_dist_offset=np.zeros(C, np.float32),
                   mem2local=np.zeros(max(N, 1), np.uint32))
        self.ctx.check(self.ctx.L.puckgpu_index_export(self.ctx.h, self.h, _ptr(out["coarse_norms"]),
                                                       _ptr(out["stationary"]), _ptr(out["cell_start"]),
                                                       _ptr(out["min_dist_offset"]), _ptr(out["mem2local"])))
        out["mem2local"] = out["mem2local"][:N]
        return out

    def _check_io(self, queries, n, D, I, R, width, per_query=()):
        """Queries [n, dim] float32, distances [>= n, >= width] float32,
        labels [>= n, >= width] and counts [>= n] 32-bit integers, per-query
        arrays [n] 32-bit integers, all on the context's device."""
        dv = self.ctx.device
        _check(queries, "queries", _F32, shape=(n, self.dim), device=dv)
        _check(D, "distance", _F32, min_shape=(n, width), device=dv)
        _check(I, "labels", _U32, min_shape=(n, width), device=dv)
        _check(R, "result_num", _U32, min_shape=(n,), device=dv)
        for nm, a in per_query:
            _check(a, nm, _U32, shape=(n,), device=dv)
        for a in (D, I):
            if a is not None and int(a.shape[1]) != width:
                raise ValueError(f"output rows must have exactly {width} columns (the library's row stride)")

    def search(self, queries, topk, search_coarse_count, filter_topk, topk_max=None, radius_rate=1.0,
               out=None, replay_all=False, precision=0):
        """Returns (distance [n, topk], labels [n, topk], result_num [n])."""
        n = queries.shape[0]
        topk_max = topk if topk_max is None else topk_max
        sp = _lib.SearchParams(topk_max, search_coarse_count, filter_topk, radius_rate, int(bool(replay_all)),
                               int(precision))
        if out is None:
            if _is_dev(queries):
                import torch
                dv = queries.device
                out = (torch.empty((n, topk), dtype=torch.float32, device=dv),
                       torch.empty((n, topk), dtype=torch.int32, device=dv),
                       torch.empty(n, dtype=torch.int32, device=dv))
            else:
                out = (np.zeros((n, topk), np.float32), np.zeros((n, topk), np.uint32), np.zeros(n, np.uint32))
        D, I, R = out
        self._check_io(queries, n, D, I, R, topk)
        m = _mode(queries, D, I, R)
        self.ctx.sync_torch(queries)
        self.ctx.check(self.ctx.L.puckgpu_search(self.ctx.h, self.h, ctypes.byref(sp), n, _ptr(queries), topk,
                                                 _ptr(D), _ptr(I), _ptr(R), m))
        return D, I, R

    def search_batch(self, queries, topk_max, search_coarse_count, filter_topk, topk=None, scc=None,
                     fine_budget=None, radius_rate=1.0, out=None, replay_all=False, precision=0):
        """Per-query topk / search_coarse_count / fine_budget arrays (uint32,
        n entries, or None for the batch-wide value).  search_coarse_count is
        the batch maximum.  Returns (distance [n, topk_max], labels, result_num)."""
        n = queries.shape[0]
        sp = _lib.SearchParams(topk_max, search_coarse_count, filter_topk, radius_rate, int(bool(replay_all)),
                               int(precision))
        dev = _is_dev(queries)

        def arr(a):
            if a is None:
                return None
            if dev:
                import torch
                if not _is_dev(a):
                    a = torch.as_tensor(np.asarray(a, np.uint32).view(np.int32), device=queries.device)
                return a.to(torch.int32).contiguous()
            return np.ascontiguousarray(a, np.uint32)

        qt, qs, qf = arr(topk), arr(scc), arr(fine_budget)
        if out is None:
            if dev:
                import torch
                dv = queries.device
                out = (torch.empty((n, topk_max), dtype=torch.float32, device=dv),
                       torch.empty((n, topk_max), dtype=torch.int32, device=dv),
                       torch.empty(n, dtype=torch.int32, device=dv))
            else:
                out = (np.zeros((n, topk_max), np.float32), np.zeros((n, topk_max), np.uint32),
                       np.zeros(n, np.uint32))
        D, I, R = out
        self._check_io(queries, n, D, I, R, topk_max, (("topk", qt), ("search_coarse_count", qs),
                                                       ("fine_budget", qf)))
        m = _mode(queries, D, I, R, qt, qs, qf)
        qp = _lib.QueryParams(_ptr(qt), _ptr(qs), _ptr(qf))
        self.ctx.sync_torch(queries, qt, qs, qf)
        self.ctx.check(self.ctx.L.puckgpu_search_batch(self.ctx.h, self.h, ctypes.byref(sp), ctypes.byref(qp), n,
                                                       _ptr(queries), _ptr(D), _ptr(I), _ptr(R), m))
        return D, I, R

    def center_distances(self, queries, precision=0):
        """(coarse [n, C], fine [n, F]) as the search computes them."""
        n = queries.shape[0]
        if _is_dev(queries):
            import torch
            co = torch.empty((n, self.C), dtype=torch.float32, device=queries.device)
            fo = torch.empty((n, self.F), dtype=torch.float32, device=queries.device)
        else:
            co = np.zeros((n, self.C), np.float32)
            fo = np.zeros((n, self.F), np.float32)
        _check(queries, "queries", _F32, shape=(n, self.dim), device=self.ctx.device)
        self.ctx.sync_torch(queries)
        self.ctx.check(self.ctx.L.puckgpu_center_distances(self.ctx.h, self.h, int(precision), n, _ptr(queries),
                                                           _ptr(co), _ptr(fo), _mode(queries, co, fo)))
        return co, fo

    def budget_ranks(self, queries, cells, search_coarse_count):
        """Smallest fine budget scoring each given cell (include/puck_gpu.h)."""
        n, k = cells.shape
        if _is_dev(queries):
            import torch
            r = torch.empty((n, k), dtype=torch.int32, device=queries.device)
        else:
            r = np.zeros((n, k), np.uint32)
            cells = np.ascontiguousarray(cells, np.uint32)
        _check(queries, "queries", _F32, shape=(n, self.dim), device=self.ctx.device)
        _check(cells, "cells", _U32, shape=(n, k), device=self.ctx.device)
        self.ctx.sync_torch(queries, cells)
        self.ctx.check(self.ctx.L.puckgpu_budget_ranks(self.ctx.h, self.h, search_coarse_count, n, _ptr(queries), k,
                                                       _ptr(cells), _ptr(r), _mode(queries, cells, r)))
        return r

    def last_filter(self, n, filter_topk):
        t = np.zeros((n, filter_topk), np.float32)
        i = np.zeros((n, filter_topk), np.uint32)
        c = np.zeros(n, np.uint32)
        self.ctx.check(self.ctx.L.puckgpu_last_filter(self.ctx.h, n, filter_topk, _ptr(t), _ptr(i), _ptr(c)))
        return t, i, c

    def last_scan_ms(self):
        v = ctypes.c_float(0.0)
        self.ctx.check(self.ctx.L.puckgpu_last_scan_ms(self.ctx.h, ctypes.byref(v)))
        return v.value

    def last_stage_ms(self):
        ms = np.zeros(5, np.float32)
        cnt = np.zeros(13, np.uint64)
        self.ctx.check(self.ctx.L.puckgpu_last_stage_ms(self.ctx.h, _ptr(ms), _ptr(cnt)))
        return ms, cnt


class Probe:
    """The learned fine-budget predictor resident on the device (BF16 MLP,
    FP32 master weights and Adam state); see include/puck_gpu.h."""

    def __init__(self, ctx, dim, coarse_features=64, fine_features=64, hidden=256):
        self.ctx = ctx
        self.dim, self.nc, self.nf, self.hidden = dim, coarse_features, fine_features, hidden
        self.D = dim + coarse_features + fine_features
        h = ctypes.c_void_p()
        shape = _lib.ProbeShape(coarse_features, fine_features, hidden)
        ctx.check(ctx.L.puckgpu_probe_create(ctx.h, dim, ctypes.byref(shape), ctypes.byref(h)))
        self.h = h
        self.count = ctx.L.puckgpu_probe_param_count(h)

    def free(self):
        if getattr(self, "h", None):
            self.ctx.L.puckgpu_probe_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def set_params(self, p):
        if not _is_dev(p):
            p = np.ascontiguousarray(p, np.float32)
        if p.shape[0] != self.count:
            raise ValueError(f"expected {self.count} parameters, got {p.shape[0]}")
        self.ctx.sync_torch(p)
        self.ctx.check(self.ctx.L.puckgpu_probe_set_params(self.ctx.h, self.h, _ptr(p), _mode(p)))

    def get_params(self):
        p = np.zeros(self.count, np.float32)
        self.ctx.check(self.ctx.L.puckgpu_probe_get_params(self.ctx.h, self.h, _ptr(p), 0))
        return p

    def features(self, index, queries, fine_order=False):
        """Raw features [n, D] (and fine order [n, F]) like the queries' arrays
        (numpy in, numpy out; CUDA tensor in, CUDA tensors out)."""
        n = queries.shape[0]
        if _is_dev(queries):
            import torch
            f = torch.empty((n, self.D), dtype=torch.float32, device=queries.device)
            fo = torch.empty((n, index.F), dtype=torch.int32, device=queries.device) if fine_order else None
        else:
            f = np.zeros((n, self.D), np.float32)
            fo = np.zeros((n, index.F), np.uint32) if fine_order else None
        _check(queries, "queries", _F32, shape=(n, self.dim), device=self.ctx.device)
        self.ctx.sync_torch(queries)
        self.ctx.check(self.ctx.L.puckgpu_probe_features(self.ctx.h, index.h, self.h, n, _ptr(queries), _ptr(f),
                                                         _ptr(fo), _mode(queries, f, fo)))
        return (f, fo) if fine_order else f

    def train_step(self, features, target, lr=None, betas=(0.9, 0.999), eps=1e-8, want_loss=True,
                   want_grads=False):
        n = features.shape[0]
        _check(features, "features", _F32, shape=(n, self.D), device=self.ctx.device)
        _check(target, "target", _F32, shape=(n,), device=self.ctx.device)
        adam = _lib.AdamParams(lr, betas[0], betas[1], eps) if lr is not None else None
        g = None
        if want_grads:
            k = self.count - 2 * self.D
            if _is_dev(features):
                import torch
                g = torch.empty(k, dtype=torch.float32, device=features.device)
            else:
                g = np.zeros(k, np.float32)
        loss = ctypes.c_float(0.0)
        self.ctx.sync_torch(features, target)
        self.ctx.check(self.ctx.L.puckgpu_probe_train_step(
            self.ctx.h, self.h, n, _ptr(features), _ptr(target), ctypes.byref(adam) if adam else None,
            ctypes.byref(loss) if want_loss else None, _ptr(g), _mode(features, target, g)))
        return (loss.value, g) if want_grads else loss.value

    def predict(self, features):
        n = features.shape[0]
        _check(features, "features", _F32, shape=(n, self.D), device=self.ctx.device)
        if _is_dev(features):
            import torch
            y = torch.empty(n, dtype=torch.float32, device=features.device)
        else:
            y = np.zeros(n, np.float32)
        self.ctx.sync_torch(features)
        self.ctx.check(self.ctx.L.puckgpu_probe_predict(self.ctx.h, self.h, n, _ptr(features), _ptr(y),
                                                        _mode(features, y)))
        return y


def search_learned(index, probe, queries, topk_max, search_coarse_count, filter_topk, margin=0.0, min_budget=1,
                   fallback_frac=1.0, topk=None, scc=None, radius_rate=1.0, replay_all=False, out=None, precision=0):
    """Batched search whose per-query fine budgets come from the probe, in the
    same pipeline.  Returns (distance, labels, result_num, budgets)."""
    ctx = index.ctx
    n = queries.shape[0]
    sp = _lib.SearchParams(topk_max, search_coarse_count, filter_topk, radius_rate, int(bool(replay_all)),
                           int(precision))
    bp = _lib.BudgetParams(margin, min_budget, fallback_frac)
    dev = _is_dev(queries)

    def arr(a):
        if a is None:
            return None
        if dev:
            import torch
            if not _is_dev(a):
                a = torch.as_tensor(np.asarray(a, np.uint32).view(np.int32), device=queries.device)
            return a.to(torch.int32).contiguous()
        return np.ascontiguousarray(a, np.uint32)

    qt, qs = arr(topk), arr(scc)
    if out is None:
        if dev:
            import torch
            dv = queries.device
            out = (torch.empty((n, topk_max), dtype=torch.float32, device=dv),
                   torch.empty((n, topk_max), dtype=torch.int32, device=dv),
                   torch.empty(n, dtype=torch.int32, device=dv), torch.empty(n, dtype=torch.int32, device=dv))
        else:
            out = (np.zeros((n, topk_max), np.float32), np.zeros((n, topk_max), np.uint32), np.zeros(n, np.uint32),
                   np.zeros(n, np.uint32))
    D, I, R, B = out
    index._check_io(queries, n, D, I, R, topk_max, (("topk", qt), ("search_coarse_count", qs)))
    _check(B, "budget_out", _U32, min_shape=(n,), device=ctx.device)
    if probe.dim != index.dim:
        raise ValueError(f"probe built for dim {probe.dim}, index has dim {index.dim}")
    m = _mode(queries, D, I, R, B, qt, qs)
    qp = _lib.QueryParams(_ptr(qt), _ptr(qs), None)
    ctx.sync_torch(queries, qt, qs)
    ctx.check(ctx.L.puckgpu_search_learned(ctx.h, index.h, probe.h, ctypes.byref(sp), ctypes.byref(bp),
                                           ctypes.byref(qp), n, _ptr(queries), _ptr(D), _ptr(I), _ptr(R), _ptr(B),
                                           m))
    return D, I, R, B


class Server:
    """Batched serving (include/puck_gpu.h, puckgpu_server_*): per-bucket HIP
    graphs of the whole pipeline, optional learned budgets, autotuned GEMM
    tiles."""

    def __init__(self, index, topk_max, search_coarse_count, filter_topk, max_batch=4096, radius_rate=1.0,
                 probe=None, margin=0.0, min_budget=1, fallback_frac=1.0, precision=0):
        self.index, self.ctx = index, index.ctx
        self.topk_max = topk_max
        sp = _lib.SearchParams(topk_max, search_coarse_count, filter_topk, radius_rate, 0, int(precision))
        bp = _lib.BudgetParams(margin, min_budget, fallback_frac)
        h = ctypes.c_void_p()
        self.ctx.check(self.ctx.L.puckgpu_server_create(self.ctx.h, index.h, probe.h if probe else None,
                                                        ctypes.byref(sp), ctypes.byref(bp), max_batch,
                                                        ctypes.byref(h)))
        self.h = h
        self.probe = probe
        self.max_batch = max_batch
        self.buckets = self.ctx.L.puckgpu_server_buckets(h)

    def free(self):
        if getattr(self, "h", None):
            self.ctx.L.puckgpu_server_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def search(self, queries, topk=None, scc=None, out=None):
        n = queries.shape[0]
        dev = _is_dev(queries)

        def arr(a):
            if a is None:
                return None
            if dev:
                import torch
                if not _is_dev(a):
                    a = torch.as_tensor(np.asarray(a, np.uint32).view(np.int32), device=queries.device)
                return a.to(torch.int32).contiguous()
            return np.ascontiguousarray(a, np.uint32)

        qt, qs = arr(topk), arr(scc)
        if out is None:
            if dev:
                import torch
                dv = queries.device
                out = (torch.empty((n, self.topk_max), dtype=torch.float32, device=dv),
                       torch.empty((n, self.topk_max), dtype=torch.int32, device=dv),
                       torch.empty(n, dtype=torch.int32, device=dv))
            else:
                out = (np.zeros((n, self.topk_max), np.float32), np.zeros((n, self.topk_max), np.uint32),
                       np.zeros(n, np.uint32))
        D, I, R = out
        if n > self.max_batch:
            raise ValueError(f"batch of {n} queries exceeds the server's max_batch {self.max_batch}")
        self.index._check_io(queries, n, D, I, R, self.topk_max, (("topk", qt), ("search_coarse_count", qs)))
        m = _mode(queries, D, I, R, qt, qs)
        self.ctx.sync_torch(queries, qt, qs)
        self.ctx.check(self.ctx.L.puckgpu_server_search(self.h, n, _ptr(queries), _ptr(qt), _ptr(qs), _ptr(D),
                                                        _ptr(I), _ptr(R), m))
        return D, I, R

    def tune(self, sample, reps=3):
        """sample: host float32 [max_batch, dim].  Returns (shapes [buckets],
        times [buckets, shapes] ms)."""
        sample = np.ascontiguousarray(sample, np.float32)
        if sample.shape != (self.max_batch, self.index.dim):
            # the largest bucket reads max_batch rows of the sample
            raise ValueError(f"sample must be [{self.max_batch}, {self.index.dim}], got {list(sample.shape)}")
        nsh = 4
        shapes = np.zeros(self.buckets, np.int32)
        times = np.zeros((self.buckets, nsh), np.float32)
        self.ctx.check(self.ctx.L.puckgpu_server_tune(self.h, _ptr(sample), reps, _ptr(shapes), _ptr(times)))
        return shapes, times

    def tuned(self):
        """Per bucket: (GEMM tile shape, rerank rows per lane group) chosen."""
        out = []
        for j in range(self.buckets):
            sh, rows = ctypes.c_int32(0), ctypes.c_uint32(0)
            self.ctx.check(self.ctx.L.puckgpu_server_tuned(self.h, j, ctypes.byref(sh), ctypes.byref(rows)))
            out.append((sh.value, rows.value))
        return out
