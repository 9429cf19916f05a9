"""ctypes wrapper over oracle/liboracle.so (the C restatement of Puck's search
path) plus numpy helpers: brute-force exact top-k and recall with ties.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The engine (puck_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(HERE, "liboracle.so")
_REF_PATH = os.path.join(HERE, "_ref", "libpuckref.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class Stats(ctypes.Structure):
    _fields_ = [("cells_considered", ctypes.c_uint64), ("cells_scanned", ctypes.c_uint64),
                ("points_offered", ctypes.c_uint64), ("heap_updates", ctypes.c_uint64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


def build(force=False):
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE, "all"])
    if os.path.isdir("/root/reference") and (force or not os.path.exists(_REF_PATH)):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


class WalkStep(ctypes.Structure):
    """po_walk_step: one considered cell of the walk (decision 0 break,
    1 continue, 2 scanned)."""
    _fields_ = [("l", ctypes.c_uint32), ("idx", ctypes.c_uint32), ("k", ctypes.c_uint32),
                ("coarse", ctypes.c_uint32), ("decision", ctypes.c_int32), ("pivot_before", ctypes.c_float),
                ("top_before", ctypes.c_float), ("max_st_before", ctypes.c_float), ("pivot_after", ctypes.c_float),
                ("max_st_after", ctypes.c_float), ("offered", ctypes.c_uint32), ("updated", ctypes.c_uint32)]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    _lib = load_variant(_LIB_PATH)
    return _lib


_fast = None


def lib_fast():
    """liboracle_fast.so: the same source built -O3 -march=x86-64-v3 (the
    bench's CPU baseline); results are bit-identical to lib()'s."""
    global _fast
    if _fast is None:
        path = os.path.join(HERE, "liboracle_fast.so")
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", HERE, "all"])
        _fast = load_variant(path)
    return _fast


_train = None


def lib_train():
    """liboracle_train.so: the k-means and coarse / fine training restatement."""
    global _train
    if _train is None:
        path = os.path.join(HERE, "liboracle_train.so")
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", HERE, "all"])
        L = ctypes.CDLL(path)
        ip = ctypes.POINTER(ctypes.c_int)
        dp = ctypes.POINTER(ctypes.c_double)
        L.pot_kmeans.restype = ctypes.c_float
        L.pot_kmeans.argtypes = [ctypes.c_uint32] * 3 + [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                                         ctypes.c_int, _f32p, _i32p, ip, ip]
        L.pot_kmeans_core.restype = ctypes.c_int
        L.pot_kmeans_core.argtypes = [ctypes.c_uint32] * 3 + [_f32p, ctypes.c_int, ctypes.c_uint32, ctypes.c_int,
                                                              _f32p, _i32p, _f32p, dp, ip, ip, _f64p]
        L.pot_train.restype = ctypes.c_int
        L.pot_train.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _f32p, ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_int, _f32p, _f32p,
                                _f32p, ip, ip]
        _train = L
    return _train


def kmeans(x, k, niter=30, init_pp=True, seed=0, mean_mode=0):
    """Kmeans::kmeans (kmeans.cpp:149-219) with a generator seeded `seed`.
    Returns (centroids [k, d], assign [n], mean error, iterations, splits)."""
    x = c32(x)
    n, d = x.shape
    cent = np.zeros((k, d), np.float32)
    a = np.zeros(n, np.int32)
    it, sp = ctypes.c_int(0), ctypes.c_int(0)
    err = lib_train().pot_kmeans(d, n, k, x.reshape(-1), niter, int(bool(init_pp)), seed, mean_mode,
                                 cent.reshape(-1), a, ctypes.byref(it), ctypes.byref(sp))
    if err < 0:
        raise ValueError("pot_kmeans failed")
    return cent, a, float(err), it.value, sp.value


def kmeans_core(x, centroids, niter, seed=0, mean_mode=0):
    """kmeans_core (kmeans.cpp:260-329) from given centroids.  Returns
    (centroids, assign, dis, qerr, iterations, splits, qerr per iteration)."""
    x = c32(x)
    n, d = x.shape
    cent = np.array(centroids, np.float32, copy=True).reshape(-1, d)
    k = cent.shape[0]
    a = np.zeros(n, np.int32)
    dis = np.zeros(n, np.float32)
    q = ctypes.c_double(0)
    it, sp = ctypes.c_int(0), ctypes.c_int(0)
    tr = np.zeros(max(niter, 1), np.float64)
    rc = lib_train().pot_kmeans_core(d, n, k, x.reshape(-1), niter, seed, mean_mode, cent.reshape(-1), a, dis,
                                     ctypes.byref(q), ctypes.byref(it), ctypes.byref(sp), tr)
    if rc != 0:
        raise ValueError("pot_kmeans_core failed")
    return cent, a, dis, q.value, it.value, sp.value, tr[:it.value]


def train_centers(x, C, F, iterations=10, init_pp=True, seed=0, mean_mode=0, niter=0):
    """HierarchicalClusterIndex::train (hierarchical_cluster_index.cpp:882-985).
    Returns (coarse [C, d], fine [F, d] by descending norm, errors [rounds, 2],
    rounds run, final niter)."""
    x = c32(x)
    n, d = x.shape
    co = np.zeros((C, d), np.float32)
    fi = np.zeros((F, d), np.float32)
    tr = np.zeros(2 * max(iterations, 1), np.float32)
    r, ni = ctypes.c_int(0), ctypes.c_int(0)
    rc = lib_train().pot_train(d, n, x.reshape(-1), C, F, iterations, niter, int(bool(init_pp)), seed, mean_mode,
                               co.reshape(-1), fi.reshape(-1), tr, ctypes.byref(r), ctypes.byref(ni))
    if rc != 0:
        raise ValueError("pot_train failed")
    return co, fi, tr[:2 * r.value].reshape(-1, 2), r.value, ni.value


def load_variant(path):
    """Load a build of puck_oracle.c at `path` (liboracle.so, or a variant
    compiled by a test) with its ctypes signatures."""
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    L.po_sentinel_value.restype = ctypes.c_float
    L.po_dot.restype = ctypes.c_float
    L.po_dot.argtypes = [_f32p, _f32p, ctypes.c_uint32]
    L.po_sdot_short.restype = ctypes.c_float
    L.po_sdot_short.argtypes = [_f32p, _f32p, ctypes.c_uint32]
    L.po_heap_run.restype = ctypes.c_uint32
    L.po_heap_run.argtypes = [ctypes.c_uint32, _f32p, _u32p, ctypes.c_uint32, ctypes.c_int, _f32p, _u32p]
    L.po_lut.restype = ctypes.c_int
    L.po_lut.argtypes = [_f32p, _f32p, ctypes.c_uint32, ctypes.c_uint32, _f32p]
    L.po_lookup.restype = ctypes.c_float
    L.po_lookup.argtypes = [_u8p, _f32p, ctypes.c_uint32]
    L.po_rerank_dist.restype = ctypes.c_float
    L.po_rerank_dist.argtypes = [_f32p, _f32p, ctypes.c_uint32]
    L.po_normalize.restype = ctypes.c_int
    L.po_normalize.argtypes = [_f32p, ctypes.c_uint32, _f32p]
    L.po_index_create.restype = vp
    L.po_index_create.argtypes = [ctypes.c_uint32] * 5 + [ctypes.c_int, _f32p, _f32p, _f32p, _u32p, _u8p, _f32p]
    L.po_index_destroy.argtypes = [vp]
    L.po_index_export.argtypes = [vp, _f32p, _f32p, _u32p, _f32p, _u32p]
    L.po_search_one.restype = ctypes.c_int
    L.po_search_one.argtypes = [vp, _f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.c_uint32, ctypes.c_double, ctypes.c_uint32, _f32p, _u32p,
                                ctypes.POINTER(ctypes.c_uint32), _f32p, _u32p,
                                ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(Stats), _i32p, _i32p]
    L.po_search_one_given.restype = ctypes.c_int
    L.po_search_one_given.argtypes = [vp, _f32p, _f32p, _f32p, ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, _f32p, _u32p,
                                      ctypes.POINTER(ctypes.c_uint32)]
    L.po_index_set_derived.restype = None
    L.po_index_set_derived.argtypes = [vp, _f32p, _f32p]
    L.po_assign_cells_given.restype = ctypes.c_int
    L.po_assign_cells_given.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _f32p, _f32p, _f32p, ctypes.c_uint32,
                                        _f32p, _f32p, _f32p, _u32p, _f32p, ctypes.c_int]
    L.po_search_batch.restype = ctypes.c_int
    L.po_search_batch.argtypes = [vp, ctypes.c_uint32, _f32p, ctypes.c_uint32, ctypes.c_uint32,
                                  ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, _f32p, _u32p,
                                  _u32p, ctypes.c_int, ctypes.POINTER(Stats)]
    L.po_search_batch_pq.restype = ctypes.c_int
    L.po_search_batch_pq.argtypes = [vp, ctypes.c_uint32, _f32p, vp, ctypes.c_uint32, vp, ctypes.c_uint32, vp,
                                     ctypes.c_uint32, ctypes.c_double, _f32p, _u32p, _u32p, ctypes.c_int]
    L.po_assign_cells.restype = ctypes.c_int
    L.po_assign_cells.argtypes = [_f32p, _f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                  _f32p, ctypes.c_uint32, _u32p, _f32p, ctypes.c_int]
    L.po_nearest_center.restype = ctypes.c_int
    L.po_nearest_center.argtypes = [ctypes.c_uint32, _f32p, ctypes.c_uint32, _f32p, ctypes.c_uint32,
                                    _i32p, _f32p, ctypes.c_int]
    L.po_encode_filter.restype = ctypes.c_int
    L.po_encode_filter.argtypes = [_f32p, _f32p, ctypes.c_uint32, ctypes.c_uint32, _f32p,
                                   ctypes.c_uint32, _f32p, _u32p, ctypes.c_uint32, _u8p, ctypes.c_int]
    L.po_index_set_pq.restype = ctypes.c_int
    L.po_index_set_pq.argtypes = [vp, ctypes.c_uint32, _f32p, _u8p]
    L.po_encode_puck.restype = ctypes.c_int
    L.po_encode_puck.argtypes = [_f32p, _f32p, ctypes.c_uint32, ctypes.c_uint32, _f32p, ctypes.c_uint32, _f32p,
                                 ctypes.c_uint32, _f32p, _u32p, ctypes.c_uint32, _u8p, _u8p, ctypes.c_int]
    L.po_index_insert.restype = ctypes.c_int
    L.po_index_insert.argtypes = [vp, ctypes.c_uint32, _f32p, _u32p, vp, vp, vp, vp]
    L.po_index_position.restype = ctypes.c_uint32
    L.po_index_position.argtypes = [vp, ctypes.c_uint32]
    L.po_index_inserted.restype = ctypes.c_uint32
    L.po_index_inserted.argtypes = [vp]
    L.po_half_round.restype = ctypes.c_float
    L.po_half_round.argtypes = [ctypes.c_float]
    L.po_lut16_exp.restype = ctypes.c_int
    L.po_lut16_exp.argtypes = [ctypes.c_float, ctypes.c_float]
    L.po_lut16.restype = None
    L.po_lut16.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_float, ctypes.c_float]
    L.po_index_set_lut16.restype = None
    L.po_index_set_lut16.argtypes = [vp, ctypes.c_int]
    L.po_walk_trace.restype = ctypes.c_int
    L.po_walk_trace.argtypes = [vp, _f32p, _f32p, _f32p, ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.c_double, ctypes.POINTER(WalkStep), ctypes.c_uint32,
                                ctypes.POINTER(ctypes.c_uint32), _f32p, _u32p, ctypes.POINTER(ctypes.c_uint32),
                                ctypes.POINTER(Stats)]
    return L


def ref_lib():
    """oracle/_ref/libpuckref.so (reference sources built as-is) or None."""
    if not os.path.exists(_REF_PATH):
        if os.path.isdir("/root/reference"):
            build()
        if not os.path.exists(_REF_PATH):
            return None
    L = ctypes.CDLL(_REF_PATH)
    L.ref_dist_table.argtypes = [_f32p, _f32p, ctypes.c_uint, ctypes.c_uint, _f32p]
    L.ref_rerank_dist.restype = ctypes.c_float
    L.ref_rerank_dist.argtypes = [_f32p, _f32p, ctypes.c_uint]
    return L


def c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


SENTINEL = np.frombuffer(np.uint32(0x7F7F7F7F).tobytes(), dtype=np.float32)[0]


def dot(a, b):
    a, b = c32(a), c32(b)
    return float(lib().po_dot(a, b, a.size))


def sdot_short(a, b):
    """cblas_sdot of 1, 2 or 4 floats in MKL's order (float32 result)."""
    a, b = c32(a), c32(b)
    return np.float32(lib().po_sdot_short(a, b, a.size))


def heap_run(cap, vals, tags=None, reorder=True):
    vals = c32(vals)
    tags = np.arange(vals.size, dtype=np.uint32) if tags is None else np.ascontiguousarray(tags, np.uint32)
    ov = np.zeros(max(cap, 1), np.float32)
    ot = np.zeros(max(cap, 1), np.uint32)
    n = lib().po_heap_run(cap, vals, tags, vals.size, int(reorder), ov, ot)
    return n, ov[:cap], ot[:cap]


def lut(q, codebook, nsq, lsq):
    out = np.zeros(nsq * 256, np.float32)
    rc = lib().po_lut(c32(q), c32(codebook), nsq, lsq, out)
    assert rc == 0
    return out


def lookup(codes, lut_, nsq):
    return float(lib().po_lookup(np.ascontiguousarray(codes, np.uint8), c32(lut_), nsq))


def half_round(v):
    """FP16 round-to-nearest-even of one float, returned as float32."""
    return np.float32(lib().po_half_round(ctypes.c_float(float(v))))


def lut16_exp(qn, cbn):
    """Scale exponent of the FP16 filter table for |q|^2 = qn, codeword norm bound cbn."""
    return int(lib().po_lut16_exp(ctypes.c_float(float(qn)), ctypes.c_float(float(cbn))))


def lut16(lut_, qn, cbn):
    """The FP16 filter table's values in float32: half(v * 2^e) * 2^-e."""
    out = c32(lut_).copy()
    lib().po_lut16(out, out.size, ctypes.c_float(float(qn)), ctypes.c_float(float(cbn)))
    return out


def rerank_dist(x, q):
    x, q = c32(x), c32(q)
    return float(lib().po_rerank_dist(x, q, x.size))


def empty_cell_stationary():
    """sqrtf(FLT_MAX), the stationary value of an empty cell
    (puck_index.cpp:287-306)."""
    return np.sqrt(np.float32(np.finfo(np.float32).max), dtype=np.float32)


def normalize(q):
    q = c32(q)
    out = np.zeros_like(q)
    rc = lib().po_normalize(q, q.size, out)
    return out if rc == 0 else None


class OracleIndex:
    """CPU restatement of a loaded PuckIndex (memory order, derived arrays)."""

    def __init__(self, coarse, fine, codebook, cell_assign, records, raw, whether_norm, L=None):
        self._L = lib() if L is None else L
        self.coarse, self.fine, self.codebook = c32(coarse), c32(fine), c32(codebook)
        self.C, self.dim = self.coarse.shape
        self.F = self.fine.shape[0]
        self.N = int(cell_assign.shape[0])
        self.nsq = records.shape[1] - 4 if records.ndim == 2 else None
        self.whether_norm = bool(whether_norm)
        self.pq_nsq = 0
        rec = np.ascontiguousarray(records, np.uint8).reshape(-1)
        self._h = self._L.po_index_create(self.dim, self.C, self.F, self.nsq, self.N,
                                        int(whether_norm), self.coarse, self.fine, self.codebook,
                                        np.ascontiguousarray(cell_assign, np.uint32), rec,
                                        c32(raw).reshape(-1))
        if not self._h:
            raise ValueError("oracle index creation failed")

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.po_index_destroy(self._h)
            self._h = None

    def set_lut16(self, on=True):
        """Searches use the FP16 filter table (the engine's PUCKGPU_LUT_FP16 form)."""
        self._L.po_index_set_lut16(self._h, 1 if on else 0)

    def set_pq(self, pq_codebook, pq_records):
        """Second-stage PQ (whether_pq = true): codebook [pnsq, 256, dim/pnsq]
        and records [N, 4 + pnsq] in local order; the rerank then ranks by
        the PQ distance instead of the raw vectors."""
        pq_records = np.ascontiguousarray(pq_records, np.uint8)
        rc = self._L.po_index_set_pq(self._h, pq_records.shape[1] - 4, c32(pq_codebook).reshape(-1),
                                     pq_records.reshape(-1))
        if rc != 0:
            raise ValueError("po_index_set_pq refused its arguments")
        self.pq_nsq = pq_records.shape[1] - 4

    def insert(self, x):
        """Realtime insert of the rows of x in order (po_index_insert).
        Returns ids (local id, or 0xffffffff for a rejected row) and, for the
        accepted rows, their cells, filter records, second-stage PQ records
        (whether_pq) and stored features."""
        x = c32(x)
        n = x.shape[0]
        ids = np.zeros(max(n, 1), np.uint32)
        cells = np.zeros(max(n, 1), np.uint32)
        rec = np.zeros((max(n, 1), 4 + self.nsq), np.uint8)
        pn = self.pq_nsq if self.pq_nsq else 0
        prec = np.zeros((max(n, 1), 4 + pn), np.uint8)
        xo = np.zeros((max(n, 1), self.dim), np.float32)
        ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        acc = self._L.po_index_insert(self._h, n, x.reshape(-1), ids, ptr(cells), ptr(rec),
                                      ptr(prec) if pn else None, ptr(xo))
        if acc < 0:
            raise ValueError("po_index_insert failed")
        out = dict(ids=ids[:n], cells=cells[:acc], records=rec[:acc], features=xo[:acc])
        if pn:
            out["pq_records"] = prec[:acc]
        return out

    def inserted(self):
        return int(self._L.po_index_inserted(self._h))

    def position(self, tag):
        """Memory position of a heap tag (memory id, or N + insert ordinal)."""
        return int(self._L.po_index_position(self._h, int(tag)))

    def local_of_tag(self, tag):
        tag = int(tag)
        if tag >= self.N:
            return tag
        if getattr(self, "_m2l", None) is None:
            self._m2l = self.export()["mem2local"]
        return int(self._m2l[tag])

    def export(self):
        C, F, N = self.C, self.F, self.N
        cn = np.zeros(C, np.float32)
        st = np.zeros(C * F, np.float32)
        cs = np.zeros(C * F + 1, np.uint32)
        mdo = np.zeros(C, np.float32)
        m2l = np.zeros(max(N, 1), np.uint32)
        self._L.po_index_export(self._h, cn, st, cs, mdo, m2l)
        return dict(coarse_norms=cn, stationary=st, cell_start=cs, min_dist_offset=mdo,
                    mem2local=m2l[:N])

    def search_one(self, q, topk, scc, ftk, max_topk=None, radius_rate=1.0, trace=False, fine_budget=0):
        max_topk = topk if max_topk is None else max_topk
        D = np.zeros(max(topk, 1), np.float32)
        I = np.zeros(max(topk, 1), np.uint32)
        n = ctypes.c_uint32(0)
        fv = np.zeros(max(ftk, 1), np.float32)
        ft = np.zeros(max(ftk, 1), np.uint32)
        fn = ctypes.c_uint32(0)
        st = Stats()
        co = np.zeros(max(scc, 1), np.int32)
        fo = np.zeros(self.F, np.int32)
        rc = self._L.po_search_one(self._h, c32(q), topk, max_topk, scc, ftk, radius_rate, fine_budget, D, I,
                                 ctypes.byref(n), fv, ft, ctypes.byref(fn), ctypes.byref(st), co, fo)
        out = dict(rc=rc, dist=D[:topk], ids=I[:topk], n=n.value, filt_val=fv[:ftk],
                   filt_mem=ft[:ftk], filt_n=fn.value, stats=st.as_dict(),
                   coarse=co[:min(scc, self.C)], fine=fo)
        return out

    def set_derived(self, coarse_norms, stationary):
        """Replace the coarse half-norms and stationary values (e.g. with a
        BLAS's results); empty cells and min_dist_offset are re-derived."""
        self._L.po_index_set_derived(self._h, c32(coarse_norms).reshape(-1), c32(stationary).reshape(-1))

    def walk_trace(self, q, coarse_ip, fine_ip, qn, scc, ftk, radius_rate=1.0, cap=100000):
        """The walk of a prepared query with supplied center inner products and
        squared norm, with its trace.  Returns (steps as a list of dicts, filter
        values, filter memory ids, filter count, stats)."""
        tr = (WalkStep * cap)()
        tn = ctypes.c_uint32(0)
        fv = np.zeros(max(ftk, 1), np.float32)
        ft = np.zeros(max(ftk, 1), np.uint32)
        fn = ctypes.c_uint32(0)
        st = Stats()
        rc = self._L.po_walk_trace(self._h, c32(q), c32(coarse_ip), c32(fine_ip), float(qn), scc, ftk, radius_rate,
                                   tr, cap, ctypes.byref(tn), fv, ft, ctypes.byref(fn), ctypes.byref(st))
        if rc != 0:
            raise ValueError("po_walk_trace refused its arguments")
        steps = [{k: getattr(tr[i], k) for k, _ in WalkStep._fields_} for i in range(min(tn.value, cap))]
        return steps, fv[:ftk], ft[:ftk], fn.value, st.as_dict()

    def search_one_given(self, q, coarse_ip, fine_ip, qn, topk, scc, ftk, max_topk=None, radius_rate=1.0):
        """The walk and rerank for a prepared query with its center inner
        products and squared norm supplied.  Returns (dist, ids, n)."""
        max_topk = topk if max_topk is None else max_topk
        D = np.zeros(max(topk, 1), np.float32)
        I = np.zeros(max(topk, 1), np.uint32)
        n = ctypes.c_uint32(0)
        rc = self._L.po_search_one_given(self._h, c32(q), c32(coarse_ip), c32(fine_ip), float(qn), topk, max_topk, scc,
                                       ftk, radius_rate, D, I, ctypes.byref(n))
        if rc != 0:
            raise ValueError("po_search_one_given refused its arguments")
        return D[:topk], I[:topk], n.value

    def search_batch(self, Q, topk, scc, ftk, max_topk=None, radius_rate=1.0, nthreads=0):
        Q = c32(Q)
        nq = Q.shape[0]
        max_topk = topk if max_topk is None else max_topk
        D = np.zeros((nq, topk), np.float32)
        I = np.zeros((nq, topk), np.uint32)
        nres = np.zeros(nq, np.uint32)
        st = Stats()
        self._L.po_search_batch(self._h, nq, Q.reshape(-1), topk, max_topk, scc, ftk, radius_rate,
                              D.reshape(-1), I.reshape(-1), nres, nthreads, ctypes.byref(st))
        return D, I, nres, st.as_dict()

    def search_batch_pq(self, Q, max_topk, scc, ftk, topk=None, scc_arr=None, fine_budget=None,
                        radius_rate=1.0, nthreads=0):
        """Per-query topk / search_coarse_count / fine budget (None: scalar);
        rows of stride max_topk."""
        Q = c32(Q)
        nq = Q.shape[0]
        keep = []

        def arr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, np.uint32)
            keep.append(a)
            return a.ctypes.data_as(ctypes.c_void_p)

        D = np.zeros((nq, max_topk), np.float32)
        I = np.zeros((nq, max_topk), np.uint32)
        nres = np.zeros(nq, np.uint32)
        self._L.po_search_batch_pq(self._h, nq, Q.reshape(-1), arr(topk), max_topk, arr(scc_arr), scc,
                                 arr(fine_budget), ftk, radius_rate, D.reshape(-1), I.reshape(-1), nres, nthreads)
        return D, I, nres


def assign_cells(coarse, fine, X, nthreads=0):
    coarse, fine, X = c32(coarse), c32(fine), c32(X)
    C, d = coarse.shape
    F = fine.shape[0]
    n = X.shape[0]
    cell = np.zeros(n, np.uint32)
    dist = np.zeros(n, np.float32)
    lib().po_assign_cells(coarse.reshape(-1), fine.reshape(-1), C, F, d, X.reshape(-1), n, cell, dist,
                          nthreads)
    return cell, dist


def assign_cells_given(cn, st, fbound, cip, fip, xn, nthreads=0):
    """po_assign_cells on supplied arithmetic: coarse half-norms cn [C],
    stationary st [C*F], fbound [F], inner products cip [n, C], fip [n, F] and
    half squared norms xn [n].  Returns (cell, dist)."""
    cip, fip = c32(cip), c32(fip)
    n, C = cip.shape
    F = fip.shape[1]
    cell = np.zeros(n, np.uint32)
    dist = np.zeros(n, np.float32)
    lib().po_assign_cells_given(C, F, c32(cn), c32(st).reshape(-1), c32(fbound), n, cip.reshape(-1),
                                fip.reshape(-1), c32(xn), cell, dist, nthreads)
    return cell, dist


def nearest_center(centroids, X, nthreads=0):
    centroids, X = c32(centroids), c32(X)
    k, d = centroids.shape
    n = X.shape[0]
    a = np.zeros(n, np.int32)
    dis = np.zeros(n, np.float32)
    lib().po_nearest_center(d, centroids.reshape(-1), k, X.reshape(-1), n, a, dis, nthreads)
    return a, dis


def encode_puck(coarse, fine, filter_codebook, fnsq, pq_codebook, pnsq, X, cell, nthreads=0):
    """Filter and second-stage PQ records of X (puck_assign with both
    quantizers): returns (filter records [n, 4 + fnsq], pq records [n, 4 + pnsq])."""
    coarse, fine, X = c32(coarse), c32(fine), c32(X)
    d = coarse.shape[1]
    F = fine.shape[0]
    n = X.shape[0]
    fo = np.zeros((n, 4 + fnsq), np.uint8)
    po = np.zeros((n, 4 + pnsq), np.uint8)
    rc = lib().po_encode_puck(coarse.reshape(-1), fine.reshape(-1), F, d, c32(filter_codebook).reshape(-1), fnsq,
                              c32(pq_codebook).reshape(-1), pnsq, X.reshape(-1), np.ascontiguousarray(cell, np.uint32),
                              n, fo.reshape(-1), po.reshape(-1), nthreads)
    assert rc == 0
    return fo, po


def encode_filter(coarse, fine, codebook, nsq, X, cell, nthreads=0):
    coarse, fine, codebook, X = c32(coarse), c32(fine), c32(codebook), c32(X)
    d = coarse.shape[1]
    F = fine.shape[0]
    n = X.shape[0]
    out = np.zeros((n, 4 + nsq), np.uint8)
    rc = lib().po_encode_filter(coarse.reshape(-1), fine.reshape(-1), F, d, codebook.reshape(-1), nsq,
                                X.reshape(-1), np.ascontiguousarray(cell, np.uint32), n,
                                out.reshape(-1), nthreads)
    assert rc == 0
    return out


# ---------------------------------------------------------------------------
# Exact brute force and recall (ann-benchmarks measurement spec).

def exact_topk(X, Q, k, recheck=8):
    """Exact squared-L2 top-k: FP32 GEMM candidate pass, then an FP64 recheck
    of the top (k + recheck) candidates so that ranks near the boundary are
    decided in double precision.  Returns (ids int64 [nq,k], dists f64 [nq,k])."""
    X = c32(X)
    Q = c32(Q)
    xn = np.einsum("ij,ij->i", X, X, dtype=np.float64).astype(np.float32)
    kk = min(k + recheck, X.shape[0])
    ids = np.zeros((Q.shape[0], k), np.int64)
    dd = np.zeros((Q.shape[0], k), np.float64)
    bs = 1024
    for s in range(0, Q.shape[0], bs):
        q = Q[s:s + bs]
        d = xn[None, :] - 2.0 * (q @ X.T)
        cand = np.argpartition(d, kk - 1, axis=1)[:, :kk]
        xc = X[cand].astype(np.float64)
        exact = ((xc - q.astype(np.float64)[:, None, :]) ** 2).sum(-1)
        order = np.lexsort((cand, exact), axis=1) if False else np.argsort(exact, axis=1, kind="stable")
        ids[s:s + bs] = np.take_along_axis(cand, order, 1)[:, :k]
        dd[s:s + bs] = np.take_along_axis(exact, order, 1)[:, :k]
    return ids, dd


def recall_with_ties(true_ids, true_dists, run_ids, count):
    """benchmark/plotting/metrics.py:15-40 semantics, returns (mean, per-query, n_ties)."""
    nq = len(run_ids)
    rec = np.zeros(nq)
    ties = 0
    for i in range(nq):
        tid, tdist = true_ids[i], true_dists[i]
        gt = len(tdist)
        if gt == count:
            r = len(set(np.asarray(tid[:count]).tolist()) & set(np.asarray(run_ids[i]).tolist()))
        else:
            anchor = tdist[count - 1]
            end = gt
            for j in range(count, gt):
                if not abs(anchor - tdist[j]) < 1e-6:
                    end = j
                    break
            if end > count:
                ties += 1
            r = len(set(np.asarray(tid[:end]).tolist()) & set(np.asarray(run_ids[i]).tolist()))
        rec[i] = r
    return float(np.mean(rec) / count), rec, ties


def recall_without_ties(true_ids, run_ids, count):
    rec = np.array([len(set(np.asarray(true_ids[i][:count]).tolist()) & set(np.asarray(run_ids[i]).tolist()))
                    for i in range(len(run_ids))], dtype=np.float64)
    return float(np.mean(rec) / count), rec
