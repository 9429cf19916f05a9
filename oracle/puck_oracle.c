/*
 * puck_oracle.c -- CPU restatement of Puck's batched search path
 * (PuckIndex::search with whether_pq=false), of its index-load derivations
 * and of its build-time assignment / PQ encoding.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP engine is
 * compared against.  It is compiled into oracle/liboracle.so and is used only
 * by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.
 * The product library (libpuckgpu.so) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Paths are
 * relative to the reference checkout (heaoxiang/puck).
 *
 * Floating-point conventions (the parts the reference leaves to MKL):
 *   - po_dot() is a sequential fused-multiply-add chain starting from +0.0f.
 *     It stands in for cblas_sdot and for every entry of sgemm_; the HIP
 *     engine reproduces it bit for bit.  MKL's sgemm_ with six or more
 *     columns computes exactly this chain (checked by calling it); its
 *     single-column sgemm_ and cblas_sdot of long vectors take other,
 *     unspecified orders, so those values are unpinned against an MKL build.
 *   - The one exception: nearest_center's dot products of 1, 2 or 4 floats
 *     (the filter-PQ codes and the codebook k-means on sub-vectors) follow
 *     MKL's own order for those lengths (po_sdot_short, checked against MKL
 *     where it is installed).
 *   - Everything else (LUT, LUT lookup, heap, cell walk, rerank) follows the
 *     reference's own written order, which the reference fixes in source.
 *
 * Build: gcc -O2 -fopenmp -ffp-contract=off -fno-fast-math -shared -fPIC
 * (contraction is OFF so that every fmaf() below is explicit and every other
 * product is rounded on its own, as the reference's SSE intrinsics do).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PO_KS 256u

/* 0x7f7f7f7f as float: the value MaxHeap's memset(0x7f) leaves in each slot
 * (puck/hierarchical_cluster/max_heap.cpp:30-32). */
static float po_sentinel(void) {
    uint32_t b = 0x7f7f7f7fu;
    float f;
    memcpy(&f, &b, 4);
    return f;
}

float po_sentinel_value(void) { return po_sentinel(); }

/* ------------------------------------------------------------------------ */
/* Dot product: stands in for cblas_sdot / sgemm_ entries
 * (hierarchical_cluster_index.cpp:78-90, :279, :305, :1125; puck_index.cpp:398,
 * :515; quantization.cpp:251; kmeans.cpp:230-245).  Sequential FMA chain. */
float po_dot(const float* a, const float* b, uint32_t n) {
    float acc = 0.0f;
    for (uint32_t i = 0; i < n; ++i) acc = fmaf(a[i], b[i], acc);
    return acc;
}

/* cblas_sdot of 1, 2 or 4 floats, in the order Intel MKL computes it: every
 * product rounded, then p0 + p1 (n = 2) or (p0 + p1) + (p2 + p3) (n = 4), no
 * fused multiply-add.  Determined by calling MKL 2021.4's cblas_sdot on an
 * x86-64 AVX-512 host and comparing result bits with candidate orders (every
 * sample agreed with this one); tests/test_oracle.py repeats that comparison
 * wherever libmkl_rt.so is installed.  nearest_center (kmeans.cpp:221-258)
 * takes these dot products with n = lsq when it picks the filter-PQ codes
 * (puck_index.cpp:667-668).  Other n: po_dot. */
float po_sdot_short(const float* a, const float* b, uint32_t n) {
    if (n == 1) return a[0] * b[0];
    if (n == 2) {
        float p0 = a[0] * b[0], p1 = a[1] * b[1];
        return p0 + p1;
    }
    if (n == 4) {
        float p0 = a[0] * b[0], p1 = a[1] * b[1], p2 = a[2] * b[2], p3 = a[3] * b[3];
        float s01 = p0 + p1, s23 = p2 + p3;
        return s01 + s23;
    }
    return po_dot(a, b, n);
}

/* ------------------------------------------------------------------------ */
/* MaxHeap (puck/hierarchical_cluster/max_heap.cpp:28-102, max_heap.h:28-70).
 * 1-based arrays over caller memory, filled with the 0x7f7f7f7f sentinel,
 * non-full inserts start sifting at slot _default_point_cnt--. */
typedef struct {
    float* val;      /* 1-based view: val[1] is caller[0] */
    uint32_t* tag;
    uint32_t size;
    uint32_t default_cnt;
} po_heap;

static void po_heap_init(po_heap* h, uint32_t size, float* val, uint32_t* tag) {
    /* max_heap.cpp:28-35 */
    memset(val, 0x7f, sizeof(float) * size);
    h->val = val - 1;
    h->tag = tag - 1;
    h->size = size;
    h->default_cnt = size;
}

static void po_heap_insert(po_heap* h, uint32_t heap_size, uint32_t father, float v, uint32_t t) {
    /* max_heap.cpp:67-100 */
    for (;;) {
        uint32_t l = father << 1;
        if (l > heap_size) break;
        uint32_t r = l + 1;
        if (r > heap_size || h->val[l] > h->val[r]) {
            if (h->val[l] < v) break;
            h->val[father] = h->val[l];
            h->tag[father] = h->tag[l];
            father = l;
        } else {
            if (h->val[r] < v) break;
            h->val[father] = h->val[r];
            h->tag[father] = h->tag[r];
            father = r;
        }
    }
    h->val[father] = v;
    h->tag[father] = t;
}

static void po_heap_update(po_heap* h, float v, uint32_t t) {
    /* max_heap.cpp:52-65 */
    uint32_t father = 1;
    if (v >= h->val[father]) return;
    if (h->default_cnt > 0) father = h->default_cnt--;
    po_heap_insert(h, h->size, father, v, t);
}

static void po_heap_reorder(po_heap* h) {
    /* max_heap.cpp:41-50 (heap sort to ascending) */
    for (uint32_t i = 0; i < h->size; i++) {
        float vb = h->val[1];
        uint32_t tb = h->tag[1];
        uint32_t hs = h->size - i;
        po_heap_insert(h, hs - 1, 1, h->val[hs], h->tag[hs]);
        h->val[hs] = vb;
        h->tag[hs] = tb;
    }
}

static uint32_t po_heap_count(const po_heap* h) { return h->size - h->default_cnt; }

/* Exported heap driver for tests: feed n values, return the reordered arrays. */
uint32_t po_heap_run(uint32_t cap, const float* vals, const uint32_t* tags, uint32_t n,
                     int reorder, float* out_val, uint32_t* out_tag) {
    po_heap h;
    po_heap_init(&h, cap, out_val, out_tag);
    for (uint32_t i = 0; i < n; ++i) po_heap_update(&h, vals[i], tags[i]);
    if (reorder) po_heap_reorder(&h);
    return po_heap_count(&h);
}

/* ------------------------------------------------------------------------ */
/* LUT: Quantization::get_dist_table (quantization.cpp:260-270) ->
 * faiss::fvec_L2sqr_ny_D{1,2,4} (distances_simd.cpp:72-133).
 * Each product (x-y)^2 is rounded, then the lanes are combined by the
 * hadd order of the matching SSE kernel.  Sub-vector lengths other than
 * 1, 2 and 4 are rejected (the engine supports the same three; the D8/D12
 * kernels are contracted differently by -Ofast and are not restated). */
static float po_sq(float x) { return x * x; }

static float po_lut_entry(const float* x, const float* y, uint32_t lsq) {
    switch (lsq) {
    case 1:  /* D1 :72-96 */
        return po_sq(x[0] - y[0]);
    case 2:  /* D2 :98-118 : hadd -> p0+p1 */
        return po_sq(x[0] - y[0]) + po_sq(x[1] - y[1]);
    case 4: { /* D4 :120-133 : (p0+p1)+(p2+p3) */
        float p0 = po_sq(x[0] - y[0]), p1 = po_sq(x[1] - y[1]);
        float p2 = po_sq(x[2] - y[2]), p3 = po_sq(x[3] - y[3]);
        return (p0 + p1) + (p2 + p3);
    }
    default:
        return NAN;
    }
}

int po_lut(const float* q, const float* codebook, uint32_t nsq, uint32_t lsq, float* lut) {
    if (!(lsq == 1 || lsq == 2 || lsq == 4)) return -1;
    for (uint32_t m = 0; m < nsq; ++m) {
        const float* cb = codebook + (size_t)m * PO_KS * lsq;
        for (uint32_t k = 0; k < PO_KS; ++k)
            lut[m * PO_KS + k] = po_lut_entry(q + m * lsq, cb + (size_t)k * lsq, lsq);
    }
    return 0;
}

/* LUT lookup sum: lookup_dist_table (puck_index.cpp:75-113).  Four SSE lanes,
 * lane j accumulates subspaces m = j (mod 4) in order from +0, then two hadds
 * give (s0+s1)+(s2+s3).  A partial last group adds zeros in the unused lanes. */
float po_lookup(const uint8_t* codes, const float* lut, uint32_t nsq) {
    float s[4] = {0.f, 0.f, 0.f, 0.f};
    uint32_t m = 0;
    for (; m + 4 <= nsq; m += 4)
        for (int j = 0; j < 4; ++j) s[j] = s[j] + lut[(m + j) * PO_KS + codes[m + j]];
    for (uint32_t j = 0; m + j < nsq; ++j) s[j] = s[j] + lut[(m + j) * PO_KS + codes[m + j]];
    return (s[0] + s[1]) + (s[2] + s[3]);
}

/* ------------------------------------------------------------------------ */
/* FP16 filter table (a precision form of the engine; no counterpart in the
 * reference, whose table is FP32).  po_half_round rounds a float to the
 * nearest IEEE binary16 value (ties to even; subnormals below 2^-14 in steps
 * of 2^-24; magnitudes from 65520 up become infinite) and returns it as a
 * float.  po_lut16_exp(qn, cbn) is the table's scale exponent:
 * floor(log2(65504 / (sqrt(qn) + cbn)^2)) clamped to [-30, 30], qn = |q|^2 and
 * cbn the largest codeword norm, so that no entry ||q_m - c||^2 <=
 * (|q| + cbn)^2 overflows.  po_lut16 replaces each entry v by
 * half(v * 2^e) * 2^-e (both scalings exact). */
float po_half_round(float v) {
    if (!(v == v)) return v;
    float a = fabsf(v);
    if (a >= 65520.0f) return copysignf(INFINITY, v);
    int e = 0;
    (void)frexpf(a, &e);          /* a = m * 2^e, m in [0.5, 1) */
    int lead = e - 1;             /* exponent of the leading bit */
    if (lead < -14) lead = -14;   /* binary16 subnormals: quantum 2^-24 */
    float quantum = ldexpf(1.0f, lead - 10);
    float r = rintf(a / quantum) * quantum;  /* a / quantum is exact; rintf: ties to even */
    return copysignf(r, v);
}

int po_lut16_exp(float qn, float cbn) {
    float r = sqrtf(qn) + cbn;
    float B = r * r;
    float x = 65504.0f / B;
    int e = 30;
    if (x < INFINITY) {
        int ex = 0;
        (void)frexpf(x, &ex);
        e = ex - 1;
    }
    return e < -30 ? -30 : (e > 30 ? 30 : e);
}

void po_lut16(float* lut, uint32_t n, float qn, float cbn) {
    int e = po_lut16_exp(qn, cbn);
    float sc = ldexpf(1.0f, e), inv = ldexpf(1.0f, -e);
    for (uint32_t i = 0; i < n; ++i) lut[i] = po_half_round(lut[i] * sc) * inv;
}

/* Largest codeword norm: sqrt of the largest sum of rounded squares, summed in
 * order over each codeword's lsq floats. */
static float po_codebook_max_norm(const float* cb, uint32_t n, uint32_t lsq) {
    float mx = 0.f;
    for (uint32_t k = 0; k < n; ++k) {
        float sq = 0.f;
        for (uint32_t j = 0; j < lsq; ++j) sq = sq + cb[(size_t)k * lsq + j] * cb[(size_t)k * lsq + j];
        if (sq > mx) mx = sq;
    }
    return sqrtf(mx);
}

/* ------------------------------------------------------------------------ */
/* Exact squared L2 used by the rerank: similarity::L2Sqr16Ext and L2SqrExt
 * (puck/tinker/method/hnsw_distfunc_opt_impl_inline.h:41-122), called in
 * blocks of 16 by rank_topN_points (puck_index.cpp:428-453).
 * The reference is built with -Ofast -march=native (CMakeLists.txt:16-22),
 * which lets the compiler contract and reassociate.  The order below is the
 * one GCC 11.4 emits for that loop on x86-64 with AVX (checked bit for bit
 * against the reference sources compiled with those flags, oracle/_ref):
 *   per 16-block: a_j = fma(d_j, d_j, d_{8+j}^2)          (j = 0..7)
 *   t <- (((a0+a1)+(a2+a3)) + ((a4+a5)+a6)) + (a7 + t)
 * The running total t enters the lane reduction: -Ofast reassociates the
 * `temp_dist +=` of the caller into the inlined TmpRes sum. */
static float po_l2sqr16_acc(const float* a, const float* b, float t) {
    float l[8];
    for (int j = 0; j < 8; ++j) {
        float d0 = a[j] - b[j];
        float d1 = a[8 + j] - b[8 + j];
        l[j] = fmaf(d0, d0, d1 * d1);
    }
    return (((l[0] + l[1]) + (l[2] + l[3])) + ((l[4] + l[5]) + l[6])) + (l[7] + t);
}

static float po_l2sqr_tail(const float* a, const float* b, uint32_t n) {
    /* L2SqrExt on n < 16 elements: no 16-blocks, 4-lane blocks, then scalars. */
    float acc[4] = {0, 0, 0, 0};
    uint32_t i = 0;
    for (; i + 4 <= n; i += 4)
        for (int j = 0; j < 4; ++j) {
            float d = a[i + j] - b[i + j];
            acc[j] = fmaf(d, d, acc[j]);
        }
    float s = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    for (; i < n; ++i) {
        float d = a[i] - b[i];
        s = fmaf(d, d, s);
    }
    return s;
}

float po_rerank_dist(const float* x, const float* q, uint32_t dim) {
    float t = 0.0f;
    uint32_t m = 0;
    for (; m + 16 <= dim; m += 16) t = po_l2sqr16_acc(x + m, q + m, t);
    if (dim - m > 0) t = t + po_l2sqr_tail(x + m, q + m, dim - m);
    return t;
}

/* ------------------------------------------------------------------------ */
/* Index in memory order: HierarchicalClusterIndex::read_coodbooks
 * (hierarchical_cluster_index.cpp:257-324), convert_local_to_memory_idx
 * (puck_index.cpp:229-309), read_feature_index (puck_index.cpp:198-227,
 * hierarchical_cluster_index.cpp:367-406), Quantization::load
 * (quantization.cpp:152-226). */
typedef struct {
    uint32_t dim, C, F, nsq, lsq, N;
    int whether_norm;
    float* coarse;        /* C x dim */
    float* fine;          /* F x dim (stored order) */
    float* codebook;      /* nsq x 256 x lsq */
    float* coarse_norms;  /* C : dot(c,c)/2 */
    float* stationary;    /* C*F : dot(f,f)/2 + dot(f,c); empty cell -> sqrt(FLT_MAX) */
    uint32_t* cell_start; /* C*F+1 */
    float* min_dist_offset;/* C */
    uint32_t* mem2local;  /* N */
    uint8_t* records;     /* N x (4+nsq), memory order: [f32 offset][nsq codes] */
    float* raw;           /* N x dim, memory order */
    /* second-stage PQ (whether_pq = true), or pnsq = 0 */
    uint32_t pnsq, plsq;
    float* pq_codebook;   /* pnsq x 256 x plsq */
    uint8_t* pq_records;  /* N x (4+pnsq), memory order: [f32 offset][pnsq codes] */
    /* realtime-insert overlay (RealtimeInsertPuckIndex,
     * realtime_insert_puck_index.h:99-140): points inserted after the load,
     * kept apart from the N loaded ones.  Inserted point j has heap tag and
     * local id N + j (insert_id + _labels.size(), :408, :517); each cell keeps
     * its inserted points in a list, newest first (InsertFineCluster, :312-318). */
    uint32_t ins_n, ins_cap;
    uint32_t* ins_cell;   /* cell of inserted point j */
    uint32_t* ins_next;   /* the next older inserted point of the same cell, or PO_NONE */
    uint32_t* ins_ord;    /* j's ordinal among its cell's inserted points (0 = oldest) */
    uint8_t* ins_rec;     /* filter records (4 + nsq) of the inserted points */
    float* ins_raw;       /* stored (normalised) features, whether_pq = false */
    uint8_t* ins_pq;      /* second-stage PQ records (4 + pnsq), whether_pq = true */
    uint32_t* ins_head;   /* C*F: newest inserted point of the cell, or PO_NONE */
    uint32_t* ins_cnt;    /* C*F: inserted points of the cell */
    uint32_t* ins_before; /* C*F: inserted points in the cells before (memory positions) */
    float* fbound;        /* F: -sqrt(|f|^2/2), the assignment's fine bound */
    /* FP16 filter table (the engine's precision form; the reference's table
     * is FP32): non-zero lut16 quantizes every table entry as po_lut16 does */
    int lut16;
    float cb_nmax;        /* largest codeword norm of the filter codebook */
} po_index;

#define PO_NONE 0xffffffffu

/* Filter record, raw row, PQ record and local id of heap tag `tag` (a memory
 * id below N, an inserted point N + j above). */
static const uint8_t* po_frec(const po_index* ix, uint32_t tag) {
    size_t rec = 4 + ix->nsq;
    return tag < ix->N ? ix->records + (size_t)tag * rec : ix->ins_rec + (size_t)(tag - ix->N) * rec;
}
static const float* po_rawp(const po_index* ix, uint32_t tag) {
    return tag < ix->N ? ix->raw + (size_t)tag * ix->dim : ix->ins_raw + (size_t)(tag - ix->N) * ix->dim;
}
static const uint8_t* po_pqrec(const po_index* ix, uint32_t tag) {
    size_t rec = 4 + ix->pnsq;
    return tag < ix->N ? ix->pq_records + (size_t)tag * rec : ix->ins_pq + (size_t)(tag - ix->N) * rec;
}
static uint32_t po_local(const po_index* ix, uint32_t tag) { return tag < ix->N ? ix->mem2local[tag] : tag; }

typedef struct { uint32_t cell; float off; uint32_t local; } po_order;

static int po_order_cmp(const void* pa, const void* pb) {
    /* std::pair<uint32_t, std::pair<float, uint32_t>> ordering, with the local
     * id as last key: std::stable_sort over (cell, offset) with ties kept in
     * local order (puck_index.cpp:248-272). */
    const po_order* a = (const po_order*)pa;
    const po_order* b = (const po_order*)pb;
    if (a->cell != b->cell) return a->cell < b->cell ? -1 : 1;
    if (a->off < b->off) return -1;
    if (a->off > b->off) return 1;
    if (a->local != b->local) return a->local < b->local ? -1 : 1;
    return 0;
}

void* po_index_create(uint32_t dim, uint32_t C, uint32_t F, uint32_t nsq, uint32_t N,
                      int whether_norm, const float* coarse, const float* fine,
                      const float* codebook, const uint32_t* cell_assign,
                      const uint8_t* records_local, const float* raw_local) {
    if (nsq == 0 || dim % nsq != 0) return NULL;
    if (!(dim / nsq == 1 || dim / nsq == 2 || dim / nsq == 4)) return NULL;
    po_index* ix = (po_index*)calloc(1, sizeof(po_index));
    ix->dim = dim; ix->C = C; ix->F = F; ix->nsq = nsq; ix->lsq = dim / nsq; ix->N = N;
    ix->whether_norm = whether_norm;
    size_t rec = 4 + nsq;
    ix->coarse = (float*)malloc(sizeof(float) * C * dim);
    ix->fine = (float*)malloc(sizeof(float) * F * dim);
    ix->codebook = (float*)malloc(sizeof(float) * nsq * PO_KS * ix->lsq);
    ix->coarse_norms = (float*)malloc(sizeof(float) * C);
    ix->stationary = (float*)malloc(sizeof(float) * (size_t)C * F);
    ix->cell_start = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)C * F + 1));
    ix->min_dist_offset = (float*)malloc(sizeof(float) * C);
    ix->mem2local = (uint32_t*)malloc(sizeof(uint32_t) * (N ? N : 1));
    ix->records = (uint8_t*)malloc(rec * (N ? N : 1));
    ix->raw = (float*)malloc(sizeof(float) * (size_t)(N ? N : 1) * dim);
    memcpy(ix->coarse, coarse, sizeof(float) * C * dim);
    memcpy(ix->fine, fine, sizeof(float) * F * dim);
    memcpy(ix->codebook, codebook, sizeof(float) * nsq * PO_KS * ix->lsq);

    /* read_coodbooks :278-321 */
    for (uint32_t c = 0; c < C; ++c)
        ix->coarse_norms[c] = po_dot(coarse + (size_t)c * dim, coarse + (size_t)c * dim, dim) / 2;
    float* fnorm = (float*)malloc(sizeof(float) * F);
    for (uint32_t f = 0; f < F; ++f)
        fnorm[f] = po_dot(fine + (size_t)f * dim, fine + (size_t)f * dim, dim) / 2;
    for (uint32_t c = 0; c < C; ++c)
        for (uint32_t f = 0; f < F; ++f)
            ix->stationary[(size_t)c * F + f] =
                fnorm[f] + po_dot(fine + (size_t)f * dim, coarse + (size_t)c * dim, dim);
    free(fnorm);

    /* convert_local_to_memory_idx (puck_index.cpp:229-309); cell_start is
     * pre-filled with N (hierarchical_cluster_index.cpp:352) */
    uint32_t cells = C * F;
    for (uint32_t i = 0; i <= cells; ++i) ix->cell_start[i] = N;
    po_order* ord = (po_order*)malloc(sizeof(po_order) * (N ? N : 1));
    for (uint32_t i = 0; i < N; ++i) {
        if (cell_assign[i] >= cells) { free(ord); return NULL; }
        ord[i].cell = cell_assign[i];
        memcpy(&ord[i].off, records_local + (size_t)i * rec, 4);
        ord[i].local = i;
    }
    qsort(ord, N, sizeof(po_order), po_order_cmp);
    uint32_t* local2mem = (uint32_t*)malloc(sizeof(uint32_t) * (N ? N : 1));
    uint32_t cc = 0;
    ix->cell_start[0] = 0;
    for (uint32_t i = 0; i < N; ++i) {
        local2mem[ord[i].local] = i;
        ix->mem2local[i] = ord[i].local;
        while (cc + 1 <= ord[i].cell) { ++cc; ix->cell_start[cc] = i; }
    }
    free(ord);
    for (uint32_t c = 0; c < C; ++c) ix->min_dist_offset[c] = FLT_MAX;
    for (uint32_t i = 0; i < cells; ++i) {
        uint32_t c = i / F;
        if (ix->cell_start[i] == ix->cell_start[i + 1])
            ix->stationary[i] = sqrtf(FLT_MAX);
        if (ix->stationary[i] < ix->min_dist_offset[c]) ix->min_dist_offset[c] = ix->stationary[i];
    }
    /* records and raw features permuted to memory order */
    for (uint32_t i = 0; i < N; ++i) {
        uint32_t m = local2mem[i];
        memcpy(ix->records + (size_t)m * rec, records_local + (size_t)i * rec, rec);
        memcpy(ix->raw + (size_t)m * dim, raw_local + (size_t)i * dim, sizeof(float) * dim);
    }
    free(local2mem);
    return ix;
}

/* Switches the index's searches to the FP16 filter table (on != 0) or back. */
void po_index_set_lut16(void* p, int on) {
    po_index* ix = (po_index*)p;
    ix->lut16 = on;
    ix->cb_nmax = po_codebook_max_norm(ix->codebook, ix->nsq * PO_KS, ix->lsq);
}

void po_index_destroy(void* p) {
    po_index* ix = (po_index*)p;
    if (!ix) return;
    free(ix->coarse); free(ix->fine); free(ix->codebook); free(ix->coarse_norms);
    free(ix->stationary); free(ix->cell_start); free(ix->min_dist_offset);
    free(ix->mem2local); free(ix->records); free(ix->raw);
    free(ix->pq_codebook); free(ix->pq_records);
    free(ix->ins_cell); free(ix->ins_next); free(ix->ins_ord); free(ix->ins_rec); free(ix->ins_raw);
    free(ix->ins_pq); free(ix->ins_head); free(ix->ins_cnt); free(ix->ins_before); free(ix->fbound);
    free(ix);
}

/* Export the derived arrays so tests can compare them with the engine's. */
void po_index_export(void* p, float* coarse_norms, float* stationary, uint32_t* cell_start,
                     float* min_dist_offset, uint32_t* mem2local) {
    po_index* ix = (po_index*)p;
    size_t cells = (size_t)ix->C * ix->F;
    if (coarse_norms) memcpy(coarse_norms, ix->coarse_norms, sizeof(float) * ix->C);
    if (stationary) memcpy(stationary, ix->stationary, sizeof(float) * cells);
    if (cell_start) memcpy(cell_start, ix->cell_start, sizeof(uint32_t) * (cells + 1));
    if (min_dist_offset) memcpy(min_dist_offset, ix->min_dist_offset, sizeof(float) * ix->C);
    if (mem2local) memcpy(mem2local, ix->mem2local, sizeof(uint32_t) * ix->N);
}

/* Second-stage PQ of a whether_pq = true index (PuckIndex::read_coodbooks /
 * read_feature_index, puck_index.cpp:173-227; Quantization::load,
 * quantization.cpp:152-226): pnsq subspaces of dim / pnsq floats (1, 2 or
 * 4), codebook [pnsq x 256 x plsq], records [N x (4 + pnsq)] in local order
 * (pq_data.dat), permuted to memory order here.  The rerank then uses them
 * in place of the raw vectors. */
int po_index_set_pq(void* p, uint32_t pnsq, const float* pq_codebook, const uint8_t* pq_records_local) {
    po_index* ix = (po_index*)p;
    if (pnsq == 0 || ix->dim % pnsq != 0 || ix->ins_n) return -1;
    uint32_t plsq = ix->dim / pnsq;
    if (!(plsq == 1 || plsq == 2 || plsq == 4)) return -1;
    size_t rec = 4 + pnsq;
    free(ix->pq_codebook); free(ix->pq_records);
    ix->pnsq = pnsq; ix->plsq = plsq;
    ix->pq_codebook = (float*)malloc(sizeof(float) * pnsq * PO_KS * plsq);
    memcpy(ix->pq_codebook, pq_codebook, sizeof(float) * pnsq * PO_KS * plsq);
    ix->pq_records = (uint8_t*)malloc(rec * (ix->N ? ix->N : 1));
    for (uint32_t i = 0; i < ix->N; ++i)
        memcpy(ix->pq_records + (size_t)i * rec, pq_records_local + (size_t)ix->mem2local[i] * rec, rec);
    return 0;
}

/* Second-stage distance of filter survivor `mid` (memory id) with filter
 * distance fd: rank_topN_points' whether_pq branch (puck_index.cpp:402-427),
 * (fd - |q|^2) + offset, then + the LUT sum of its PQ codes in the 4-lane
 * order of lookup_dist_table (:75-113). */
static float po_pq_dist(const po_index* ix, uint32_t mid, float fd, float qn, const float* plut) {
    const uint8_t* r = po_pqrec(ix, mid);
    float off;
    memcpy(&off, r, 4);
    float t = fd - qn + off;
    return t + po_lookup(r + 4, plut, ix->pnsq);
}

/* Test hook: replace the index's coarse half-norms and stationary values with
 * ones computed elsewhere (e.g. by a BLAS in its own summation order), then
 * re-apply the empty-cell sentinel and min_dist_offset as po_index_create
 * does (puck_index.cpp:287-306). */
void po_index_set_derived(void* p, const float* coarse_norms, const float* stationary) {
    po_index* ix = (po_index*)p;
    uint32_t cells = ix->C * ix->F;
    memcpy(ix->coarse_norms, coarse_norms, sizeof(float) * ix->C);
    memcpy(ix->stationary, stationary, sizeof(float) * cells);
    for (uint32_t c = 0; c < ix->C; ++c) ix->min_dist_offset[c] = FLT_MAX;
    for (uint32_t i = 0; i < cells; ++i) {
        uint32_t c = i / ix->F;
        if (ix->cell_start[i] == ix->cell_start[i + 1]) ix->stationary[i] = sqrtf(FLT_MAX);
        if (ix->stationary[i] < ix->min_dist_offset[c]) ix->min_dist_offset[c] = ix->stationary[i];
    }
}

/* ------------------------------------------------------------------------ */
/* Query normalisation: HierarchicalClusterIndex::normalization
 * (hierarchical_cluster_index.cpp:1513-1536), whether_norm branch:
 * norm = snrm2(q); if norm < 1e-6 -> error; q *= (float)(1.0/norm). */
int po_normalize(const float* q, uint32_t dim, float* out) {
    float norm = sqrtf(po_dot(q, q, dim));
    if (norm < 1e-6f) return -1;
    float alpha = (float)(1.0 / (double)norm);
    for (uint32_t i = 0; i < dim; ++i) out[i] = alpha * q[i];
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Search statistics returned for tests and for sizing GPU work. */
typedef struct {
    uint64_t cells_considered;  /* break-checked cells */
    uint64_t cells_scanned;     /* cells passed to compute_quantized_distance */
    uint64_t points_offered;    /* points with t_partial < top (LUT summed) */
    uint64_t heap_updates;      /* max_heap_update calls that inserted */
} po_stats;

/* Order-preserving map of float bits to uint32 (ascending floats give
 * ascending keys; -0 sorts just below +0). */
static uint32_t po_f2ord(float v) {
    uint32_t u;
    memcpy(&u, &v, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
static int po_u32_cmp(const void* a, const void* b) {
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return x < y ? -1 : (x > y);
}

static int po_u64_cmp(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : (x > y);
}

/* A budget-scan candidate: key (filter distance, memory position), heap tag. */
typedef struct { uint64_t key; uint32_t tag; } po_cand;
static int po_cand_cmp(const void* a, const void* b) {
    uint64_t x = ((const po_cand*)a)->key, y = ((const po_cand*)b)->key;
    return x < y ? -1 : (x > y);
}

/* Budget scan (no counterpart in the reference; the learned probe's fine
 * budget).  q is the prepared query.
 *   coarse lists: the scc smallest |c|^2/2 - <c,q> by (distance, id);
 *   cells of a list: the b smallest cell distances (cd + stationary) + fine
 *     distance, compared as order-preserving keys, ties at the b-th included;
 *   filter set: the ftk smallest (filter distance, memory position) over
 *     every point of those cells, the filter distance computed as the
 *     reference's compute_quantized_distance does (puck_index.cpp:346-390)
 *     without its early exits (which drop only points that cannot be among
 *     the smallest).  The memory position is the memory id while nothing is
 *     inserted; with inserted points it is the position in a layout where
 *     every cell holds its loaded points, then its inserted points newest
 *     first;
 *   result: the topk smallest (exact distance, local id) of the filter set.
 * filt_val / filt_tag receive the filter set ascending by (distance, id). */
static int po_search_budget(const po_index* ix, const float* q, uint32_t topk, uint32_t scc, uint32_t ftk,
                            uint32_t b, float* out_dist, uint32_t* out_ids, uint32_t* out_n, float* filt_val,
                            uint32_t* filt_tag, uint32_t* filt_n, po_stats* st) {
    uint32_t dim = ix->dim, C = ix->C, F = ix->F, nsq = ix->nsq;
    size_t rec = 4 + nsq;
    uint64_t* ck = (uint64_t*)malloc(sizeof(uint64_t) * C);
    float* cdv = (float*)malloc(sizeof(float) * C);
    for (uint32_t c = 0; c < C; ++c) {
        cdv[c] = ix->coarse_norms[c] - po_dot(ix->coarse + (size_t)c * dim, q, dim);
        ck[c] = ((uint64_t)po_f2ord(cdv[c]) << 32) | c;
    }
    qsort(ck, C, sizeof(uint64_t), po_u64_cmp);
    float* fd = (float*)malloc(sizeof(float) * F);
    for (uint32_t k = 0; k < F; ++k) fd[k] = -po_dot(ix->fine + (size_t)k * dim, q, dim);
    float* lut = (float*)malloc(sizeof(float) * nsq * PO_KS);
    po_lut(q, ix->codebook, nsq, ix->lsq, lut);
    if (ix->lut16) po_lut16(lut, nsq * PO_KS, po_dot(q, q, dim), ix->cb_nmax);
    uint32_t* keys = (uint32_t*)malloc(sizeof(uint32_t) * F);
    uint32_t* sorted = (uint32_t*)malloc(sizeof(uint32_t) * F);
    size_t cap = 1024, n = 0;
    po_cand* cand = (po_cand*)malloc(sizeof(po_cand) * cap);
    po_stats s = {0, 0, 0, 0};
    for (uint32_t l = 0; l < scc; ++l) {
        uint32_t c = (uint32_t)(ck[l] & 0xffffffffu);
        float cd = cdv[c];
        const float* cst = ix->stationary + (size_t)c * F;
        for (uint32_t k = 0; k < F; ++k) keys[k] = po_f2ord(cd + cst[k] + fd[k]);
        memcpy(sorted, keys, sizeof(uint32_t) * F);
        qsort(sorted, F, sizeof(uint32_t), po_u32_cmp);
        uint32_t kcut = sorted[b - 1];
        for (uint32_t k = 0; k < F; ++k) {
            s.cells_considered++;
            if (keys[k] > kcut) continue;
            s.cells_scanned++;
            float cell_dist = cd + cst[k] + fd[k];
            const size_t cell = (size_t)c * F + k;
            uint32_t b0 = ix->cell_start[cell], e0 = ix->cell_start[cell + 1];
            /* the cell's loaded points, then its inserted points newest first;
             * a point's memory position counts the inserted points of the
             * cells before it and, for an inserted point, the cell's loaded
             * points and newer inserted points */
            const uint32_t shift = ix->ins_before ? ix->ins_before[cell] : 0u;
            uint32_t j = ix->ins_head ? ix->ins_head[cell] : PO_NONE;
            for (uint32_t i = b0, slot = 0;; ++slot) {
                uint32_t tag, pos;
                if (i < e0) { tag = i; pos = i + shift; ++i; }
                else if (j != PO_NONE) { tag = ix->N + j; pos = e0 + shift + (slot - (e0 - b0)); j = ix->ins_next[j]; }
                else break;
                const uint8_t* r = po_frec(ix, tag);
                float off;
                memcpy(&off, r, 4);
                float t = (float)(2.0 * (double)cell_dist + (double)off);
                t = t + po_lookup(r + 4, lut, nsq);
                if (n == cap) {
                    cap *= 2;
                    cand = (po_cand*)realloc(cand, sizeof(po_cand) * cap);
                }
                cand[n].key = ((uint64_t)po_f2ord(t) << 32) | pos;
                cand[n].tag = tag;
                ++n;
                s.points_offered++;
            }
        }
    }
    qsort(cand, n, sizeof(po_cand), po_cand_cmp);
    uint32_t nf = n < ftk ? (uint32_t)n : ftk;
    uint64_t* rk = (uint64_t*)malloc(sizeof(uint64_t) * (nf ? nf : 1));
    float* plut = NULL;
    float qn = 0.f;
    if (ix->pnsq) {
        plut = (float*)malloc(sizeof(float) * ix->pnsq * PO_KS);
        po_lut(q, ix->pq_codebook, ix->pnsq, ix->plsq, plut);
        qn = po_dot(q, q, dim);
    }
    for (uint32_t i = 0; i < nf; ++i) {
        uint32_t mid = cand[i].tag;
        uint32_t tb = (uint32_t)(cand[i].key >> 32);
        tb = (tb & 0x80000000u) ? (tb & 0x7fffffffu) : ~tb;
        float fv;
        memcpy(&fv, &tb, 4);
        if (filt_val) filt_val[i] = fv;
        if (filt_tag) filt_tag[i] = mid;
        float d = plut ? po_pq_dist(ix, mid, fv, qn, plut) : po_rerank_dist(po_rawp(ix, mid), q, dim);
        rk[i] = ((uint64_t)po_f2ord(d) << 32) | po_local(ix, mid);
    }
    free(plut);
    if (filt_n) *filt_n = nf;
    qsort(rk, nf, sizeof(uint64_t), po_u64_cmp);
    uint32_t kout = nf < topk ? nf : topk;
    for (uint32_t i = 0; i < topk; ++i) {
        uint32_t db = 0x7f7f7f7fu, id = 0xffffffffu;
        if (i < kout) {
            db = (uint32_t)(rk[i] >> 32);
            db = (db & 0x80000000u) ? (db & 0x7fffffffu) : ~db;
            id = (uint32_t)(rk[i] & 0xffffffffu);
        }
        memcpy(out_dist + i, &db, 4);
        out_ids[i] = id;
    }
    *out_n = kout;
    s.heap_updates = nf;
    if (st) *st = s;
    free(ck); free(cdv); free(fd); free(lut); free(keys); free(sorted); free(cand); free(rk);
    return 0;
}

/* One considered cell of the walk, for the trace mode of po_walk_query:
 * position (coarse rank l, fine rank idx, fine id k, coarse id), the state
 * before the cell's break / continue tests, the decision (0 break, 1
 * continue, 2 scanned), the state after it, and for a scanned cell the
 * points whose LUT sum was taken and the heap updates. */
typedef struct {
    uint32_t l, idx, k, coarse;
    int32_t decision;
    float pivot_before, top_before, max_st_before;
    float pivot_after, max_st_after;
    uint32_t offered, updated;
} po_walk_step;

/* The walk's decision rules (puck_index.cpp:514-547).  Each is a macro so
 * that tests/test_walk_pinning.py can compile deliberately broken variants
 * of this file (-DPO_MUT_...) and check that the hand-traced known answers
 * reject every one of them; liboracle.so defines none of them.
 *   pivot      (top - |q|^2) / radius_rate / 2.0: float subtraction, then
 *              double divisions, rounded to float (:517, :544)
 *   min_dist   min_dist_offset[c] + coarse_distance[l] (:522)
 *   max_st     pivot - coarse_distance[l] - fine_distance[0] at the list's
 *              start (:523-524), lowered after every scanned cell (:547)
 *   break      fine_distance[idx] + min_dist >= pivot (:527)
 *   continue   stationary >= max_stationary_dist (:534)
 *   point      2.0 * cell_dist + offset >= heap top ends the cell (:362-364) */
#if defined(PO_MUT_PIVOT_NO_HALF)
#define PO_PIVOT(top, qn, rr) ((float)((double)((top) - (qn)) / (rr)))
#elif defined(PO_MUT_PIVOT_NO_RR)
#define PO_PIVOT(top, qn, rr) ((float)((double)((top) - (qn)) / 2.0))
#elif defined(PO_MUT_PIVOT_NO_QN)
#define PO_PIVOT(top, qn, rr) ((float)((double)(top) / (rr) / 2.0))
#else
#define PO_PIVOT(top, qn, rr) ((float)((double)((top) - (qn)) / (rr) / 2.0))
#endif
#ifdef PO_MUT_MIN_DIST_NO_MDO
#define PO_MIN_DIST(mdo, cd) (cd)
#else
#define PO_MIN_DIST(mdo, cd) ((mdo) + (cd))
#endif
#ifdef PO_MUT_MAXST_INIT_NO_FD0
#define PO_MAX_ST_INIT(p, cd, fd0) ((p) - (cd))
#else
#define PO_MAX_ST_INIT(p, cd, fd0) ((p) - (cd) - (fd0))
#endif
#ifdef PO_MUT_BREAK_GT
#define PO_BREAK_TEST(x, p) ((x) > (p))
#else
#define PO_BREAK_TEST(x, p) ((x) >= (p))
#endif
#ifdef PO_MUT_CONT_GT
#define PO_CONT_TEST(st, m) ((st) > (m))
#else
#define PO_CONT_TEST(st, m) ((st) >= (m))
#endif
#ifdef PO_MUT_POINT_GT
#define PO_POINT_BREAK_TEST(t, top) ((t) > (top))
#else
#define PO_POINT_BREAK_TEST(t, top) ((t) >= (top))
#endif

/* The reference walk and rerank for one prepared query (puck_index.cpp:485-600,
 * hierarchical_cluster_index.cpp:700-727).  coarse_ip / fine_ip / qn_in: the
 * query's inner products with every coarse and fine center and its squared
 * norm, or NULL to take them with po_dot. */
static void po_walk_query(const po_index* ix, const float* q, uint32_t topk, uint32_t scc, uint32_t ftk,
                          double radius_rate, const float* coarse_ip, const float* fine_ip, const float* qn_in,
                          float* out_dist, uint32_t* out_ids, uint32_t* out_n,
                          float* filt_val, uint32_t* filt_tag, uint32_t* filt_n,
                          po_stats* st, int* coarse_out, int* fine_out,
                          po_walk_step* trace, uint32_t trace_cap, uint32_t* trace_n) {
    uint32_t dim = ix->dim, C = ix->C, F = ix->F, nsq = ix->nsq;
    /* search_nearest_coarse_cluster (hierarchical_cluster_index.cpp:700-727) */
    float* coarse_d = (float*)malloc(sizeof(float) * (scc ? scc : 1));
    uint32_t* coarse_t = (uint32_t*)malloc(sizeof(uint32_t) * (scc ? scc : 1));
    {
        po_heap h;
        po_heap_init(&h, scc, coarse_d, coarse_t);
        for (uint32_t c = 0; c < C; ++c) {
            float ip = coarse_ip ? coarse_ip[c] : po_dot(ix->coarse + (size_t)c * dim, q, dim);
            float t = ix->coarse_norms[c] - ip;
            if (scc > 0 && t < coarse_d[0]) po_heap_update(&h, t, c);
        }
        po_heap_reorder(&h);
    }
    if (coarse_out) for (uint32_t l = 0; l < scc; ++l) coarse_out[l] = (int)coarse_t[l];
    /* pre_filter_search -> get_dist_table (puck_index.cpp:459-483) */
    float* lut = (float*)malloc(sizeof(float) * nsq * PO_KS);
    po_lut(q, ix->codebook, nsq, ix->lsq, lut);
    if (ix->lut16) po_lut16(lut, nsq * PO_KS, qn_in ? *qn_in : po_dot(q, q, dim), ix->cb_nmax);
    /* fine GEMV + full heap sort (puck_index.cpp:494-504) */
    float* fine_d = (float*)malloc(sizeof(float) * F);
    uint32_t* fine_t = (uint32_t*)malloc(sizeof(uint32_t) * F);
    {
        po_heap h;
        po_heap_init(&h, F, fine_d, fine_t);
        for (uint32_t k = 0; k < F; ++k)
            po_heap_update(&h, -(fine_ip ? fine_ip[k] : po_dot(ix->fine + (size_t)k * dim, q, dim)), k);
        po_heap_reorder(&h);
    }
    if (fine_out) for (uint32_t k = 0; k < F; ++k) fine_out[k] = (int)fine_t[k];
    /* filter heap + cell walk (puck_index.cpp:506-557) */
    float* fval = (float*)malloc(sizeof(float) * ftk);
    uint32_t* ftag = (uint32_t*)malloc(sizeof(uint32_t) * ftk);
    po_heap fh;
    po_heap_init(&fh, ftk, fval, ftag);
    float qn = qn_in ? *qn_in : po_dot(q, q, dim);
    float pivot = PO_PIVOT(fval[0], qn, radius_rate);
    size_t rec = 4 + nsq;
    po_stats s = {0, 0, 0, 0};
    uint32_t nt = 0;

    for (uint32_t l = 0; l < scc; ++l) {
        uint32_t c = coarse_t[l];
        const float* cst = ix->stationary + (size_t)c * F;
        const uint32_t* cstart = ix->cell_start + (size_t)c * F;
        float cd = coarse_d[l];
        float min_dist = PO_MIN_DIST(ix->min_dist_offset[c], cd);
        float max_st = PO_MAX_ST_INIT(pivot, cd, fine_d[0]);
        for (uint32_t idx = 0; idx < F; ++idx) {
            s.cells_considered++;
            po_walk_step* ts = (trace && nt < trace_cap) ? trace + nt : NULL;
            ++nt;
            if (ts) {
                ts->l = l; ts->idx = idx; ts->k = fine_t[idx]; ts->coarse = c;
                ts->pivot_before = pivot; ts->top_before = fval[0]; ts->max_st_before = max_st;
                ts->offered = 0; ts->updated = 0;
            }
            if (PO_BREAK_TEST(fine_d[idx] + min_dist, pivot)) {
                if (ts) { ts->decision = 0; ts->pivot_after = pivot; ts->max_st_after = max_st; }
                break;
            }
            uint32_t k = fine_t[idx];
            if (PO_CONT_TEST(cst[k], max_st)) {
#ifdef PO_MUT_MAXST_ON_CONTINUE
                { float cand = pivot - cd - fine_d[idx]; if (cand < max_st) max_st = cand; }
#endif
                if (ts) { ts->decision = 1; ts->pivot_after = pivot; ts->max_st_after = max_st; }
                continue;
            }
            float cell_dist = cd + cst[k] + fine_d[idx];
            s.cells_scanned++;
            /* compute_quantized_distance (puck_index.cpp:346-390) */
            uint32_t b = cstart[k], e = cstart[k + 1];
            int updated = 0;
            uint32_t offered = 0;
            for (uint32_t i = b; i < e; ++i) {
                const uint8_t* r = ix->records + (size_t)i * rec;
                float off;
                memcpy(&off, r, 4);
                float t = (float)(2.0 * (double)cell_dist + (double)off);
                if (PO_POINT_BREAK_TEST(t, fval[0])) break;
                s.points_offered++;
                ++offered;
                t = t + po_lookup(r + 4, lut, nsq);
                if (t < fval[0]) {
                    po_heap_update(&fh, t, i);
                    ++updated;
                    s.heap_updates++;
                }
            }
            /* the cell's inserted points, newest first, with no early exit;
             * each one counts towards the returned count, updated or not
             * (RealtimeInsertPuckIndex::compute_quantized_distance,
             * realtime_insert_puck_index.cpp:398-453) */
            for (uint32_t j = ix->ins_head ? ix->ins_head[(size_t)c * F + k] : PO_NONE; j != PO_NONE;
                 j = ix->ins_next[j]) {
                const uint8_t* r = ix->ins_rec + (size_t)j * rec;
                float off;
                memcpy(&off, r, 4);
                float t = (float)(2.0 * (double)cell_dist + (double)off);
                s.points_offered++;
                ++offered;
                ++updated;
                t = t + po_lookup(r + 4, lut, nsq);
                if (t < fval[0]) {
                    po_heap_update(&fh, t, ix->N + j);
                    s.heap_updates++;
                }
            }
            if (updated > 0) pivot = PO_PIVOT(fval[0], qn, radius_rate);
#ifndef PO_MUT_NO_MAXST_UPDATE
            float cand = pivot - cd - fine_d[idx];
            if (cand < max_st) max_st = cand;
#endif
            if (ts) {
                ts->decision = 2; ts->pivot_after = pivot; ts->max_st_after = max_st;
                ts->offered = offered; ts->updated = (uint32_t)updated;
            }
        }
    }
    if (trace_n) *trace_n = nt;
    uint32_t nf = po_heap_count(&fh);
    if (nf < ftk) po_heap_reorder(&fh);
    if (filt_val) memcpy(filt_val, fval, sizeof(float) * ftk);
    if (filt_tag) memcpy(filt_tag, ftag, sizeof(uint32_t) * ftk);
    if (filt_n) *filt_n = nf;
    /* rank_topN_points (puck_index.cpp:392-457): exact L2 over the raw
     * vectors (whether_pq = false, :428-453) or the second-stage PQ distance
     * (whether_pq = true, :402-427), both over the filter heap in array order */
    {
        po_heap rh;
        po_heap_init(&rh, topk, out_dist, out_ids);
        float* plut = NULL;
        if (ix->pnsq) {
            plut = (float*)malloc(sizeof(float) * ix->pnsq * PO_KS);
            po_lut(q, ix->pq_codebook, ix->pnsq, ix->plsq, plut);
        }
        for (uint32_t i = 0; i < nf; ++i) {
            /* (RealtimeInsertPuckIndex::rank_topN_points reads an inserted
             * point's row from the insert memory, :455-541) */
            float t = plut ? po_pq_dist(ix, ftag[i], fval[i], qn, plut) : po_rerank_dist(po_rawp(ix, ftag[i]), q, dim);
            if (t < out_dist[0]) po_heap_update(&rh, t, po_local(ix, ftag[i]));
        }
        free(plut);
        po_heap_reorder(&rh);
        *out_n = po_heap_count(&rh);
    }
    if (st) *st = s;
    free(coarse_d); free(coarse_t); free(lut); free(fine_d); free(fine_t);
    free(fval); free(ftag);
}

/* One query: PuckIndex::search (puck_index.cpp:559-600) with whether_pq=false.
 * Returns 0, or -1 for topk > max_topk / null feature / zero-norm query.
 * filt_val/filt_tag (ftk each) receive the filter heap as rank_topN_points
 * reads it (heap order when full, ascending when not full).
 * fine_budget b: 0 or >= F is the reference's walk; 0 < b < F is the budget
 * scan of po_search_budget (an extension with no counterpart in the
 * reference). */
int po_search_one(const void* p, const float* feature, uint32_t topk, uint32_t max_topk,
                  uint32_t scc, uint32_t ftk, double radius_rate, uint32_t fine_budget,
                  float* out_dist, uint32_t* out_ids, uint32_t* out_n,
                  float* filt_val, uint32_t* filt_tag, uint32_t* filt_n,
                  po_stats* st, int* coarse_out, int* fine_out) {
    const po_index* ix = (const po_index*)p;
    uint32_t dim = ix->dim, C = ix->C, F = ix->F;
    /* puck_index.cpp:560-563 */
    if (topk > max_topk || feature == NULL) return -1;
    if (scc > C) scc = C;   /* index_conf.cpp:64 search_coarse_count = min(flag, C) */
    float* q = (float*)malloc(sizeof(float) * dim);
    if (ix->whether_norm) {
        if (po_normalize(feature, dim, q) != 0) { free(q); return -1; }
    } else {
        memcpy(q, feature, sizeof(float) * dim);
    }
    if (fine_budget != 0 && fine_budget < F) {
        int rc = po_search_budget(ix, q, topk, scc, ftk, fine_budget, out_dist, out_ids, out_n, filt_val,
                                  filt_tag, filt_n, st);
        free(q);
        return rc;
    }
    po_walk_query(ix, q, topk, scc, ftk, radius_rate, NULL, NULL, NULL, out_dist, out_ids, out_n, filt_val,
                  filt_tag, filt_n, st, coarse_out, fine_out, NULL, 0, NULL);
    free(q);
    return 0;
}

/* Test hook: po_search_one's walk and rerank for a prepared (already
 * normalised) query whose center inner products and squared norm are given,
 * e.g. computed by a BLAS in its own summation order. */
int po_search_one_given(const void* p, const float* q, const float* coarse_ip, const float* fine_ip, float qn,
                        uint32_t topk, uint32_t max_topk, uint32_t scc, uint32_t ftk, double radius_rate,
                        float* out_dist, uint32_t* out_ids, uint32_t* out_n) {
    const po_index* ix = (const po_index*)p;
    if (topk > max_topk || q == NULL || coarse_ip == NULL || fine_ip == NULL) return -1;
    if (scc > ix->C) scc = ix->C;
    po_walk_query(ix, q, topk, scc, ftk, radius_rate, coarse_ip, fine_ip, &qn, out_dist, out_ids, out_n, NULL,
                  NULL, NULL, NULL, NULL, NULL, NULL, 0, NULL);
    return 0;
}

/* Test hook: the walk of po_search_one_given with its trace.  trace
 * receives up to trace_cap considered cells in visiting order (*trace_n =
 * the number considered); filt_val / filt_tag / filt_n the filter heap as
 * po_search_one gives it; st the walk's counters. */
int po_walk_trace(const void* p, const float* q, const float* coarse_ip, const float* fine_ip, float qn,
                  uint32_t scc, uint32_t ftk, double radius_rate, po_walk_step* trace, uint32_t trace_cap,
                  uint32_t* trace_n, float* filt_val, uint32_t* filt_tag, uint32_t* filt_n, po_stats* st) {
    const po_index* ix = (const po_index*)p;
    if (q == NULL || coarse_ip == NULL || fine_ip == NULL || ftk == 0) return -1;
    if (scc > ix->C) scc = ix->C;
    float od[1];
    uint32_t oi[1], on = 0;
    po_walk_query(ix, q, 1, scc, ftk, radius_rate, coarse_ip, fine_ip, &qn, od, oi, &on, filt_val, filt_tag,
                  filt_n, st, NULL, NULL, trace, trace_cap, trace_n);
    return 0;
}

/* Batched entry: PySearcher::search + ParallelFor (pyapi_wrapper/py_api_wrapper.cpp:111-171):
 * one query per thread, per-query return codes ignored (output rows of a failed
 * query are left as the heap left them).  Row stride is topk. */
int po_search_batch(const void* p, uint32_t nq, const float* queries, uint32_t topk,
                    uint32_t max_topk, uint32_t scc, uint32_t ftk, double radius_rate,
                    float* dist, uint32_t* ids, uint32_t* nres, int nthreads,
                    po_stats* stats_out) {
    const po_index* ix = (const po_index*)p;
    uint64_t a = 0, b = 0, c = 0, d = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads) reduction(+ : a, b, c, d)
#endif
    for (int64_t i = 0; i < (int64_t)nq; ++i) {
        po_stats s = {0, 0, 0, 0};
        uint32_t n = 0;
        int rc = po_search_one(ix, queries + (size_t)i * ix->dim, topk, max_topk, scc, ftk,
                               radius_rate, 0, dist + (size_t)i * topk, ids + (size_t)i * topk, &n,
                               NULL, NULL, NULL, &s, NULL, NULL);
        if (nres) nres[i] = rc == 0 ? n : 0;
        a += s.cells_considered; b += s.cells_scanned; c += s.points_offered; d += s.heap_updates;
    }
    if (stats_out) {
        stats_out->cells_considered = a; stats_out->cells_scanned = b;
        stats_out->points_offered = c; stats_out->heap_updates = d;
    }
    return 0;
}

/* Batch with per-query topk / search_coarse_count / fine budget (arrays may
 * be NULL: the scalar applies).  Rows have stride max_topk; slots past a
 * query's topk, and every slot of a query whose search fails (topk 0 or
 * > max_topk, zero-norm query), hold distance 0x7f7f7f7f and id 0xffffffff,
 * with nres 0 for a failed query. */
int po_search_batch_pq(const void* p, uint32_t nq, const float* queries, const uint32_t* topk_arr,
                       uint32_t max_topk, const uint32_t* scc_arr, uint32_t scc, const uint32_t* nf_arr,
                       uint32_t ftk, double radius_rate, float* dist, uint32_t* ids, uint32_t* nres,
                       int nthreads) {
    const po_index* ix = (const po_index*)p;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
    for (int64_t i = 0; i < (int64_t)nq; ++i) {
        uint32_t k = topk_arr ? topk_arr[i] : max_topk;
        uint32_t sc = scc_arr ? (scc_arr[i] < scc ? scc_arr[i] : scc) : scc;
        uint32_t nf = nf_arr ? nf_arr[i] : 0;
        float* D = dist + (size_t)i * max_topk;
        uint32_t* I = ids + (size_t)i * max_topk;
        uint32_t n = 0;
        int rc = -1;
        if (k >= 1 && k <= max_topk)
            rc = po_search_one(ix, queries + (size_t)i * ix->dim, k, max_topk, sc, ftk, radius_rate, nf, D, I,
                               &n, NULL, NULL, NULL, NULL, NULL, NULL);
        uint32_t first = rc == 0 ? k : 0;
        for (uint32_t r = first; r < max_topk; ++r) {
            uint32_t sent = 0x7f7f7f7fu;
            memcpy(D + r, &sent, 4);
            I[r] = 0xffffffffu;
        }
        if (rc == 0)
            for (uint32_t r = n; r < k; ++r) I[r] = 0xffffffffu;
        if (nres) nres[i] = rc == 0 ? n : 0;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Build-time cell assignment: HierarchicalClusterIndex::assign +
 * nearest_cell_assign (hierarchical_cluster_index.cpp:1042-1138), with the
 * fine-norm bound array of batch_assign (:1186-1195). */
typedef struct { float term; int idx; } po_pair;
static int po_pair_cmp(const void* pa, const void* pb) {
    const po_pair* a = (const po_pair*)pa;
    const po_pair* b = (const po_pair*)pb;
    if (a->term < b->term) return -1;
    if (a->term > b->term) return 1;
    return a->idx < b->idx ? -1 : (a->idx > b->idx);
}

/* nearest_cell_assign (hierarchical_cluster_index.cpp:1042-1085) for one point
 * from its inner products with every coarse (cip) and fine (fip) center and
 * its half squared norm xn; cn are the coarse half-norms, st stationary,
 * fbound = -sqrt(fine half-norm). */
static void po_assign_one(uint32_t C, uint32_t F, const float* cn, const float* st, const float* fbound,
                          const float* cip, const float* fip, float xn, po_pair* pairs, uint32_t* cell_out,
                          float* dist_out) {
    for (uint32_t c = 0; c < C; ++c) {
        pairs[c].idx = (int)c;
        pairs[c].term = cn[c] - cip[c] + xn;
    }
    qsort(pairs, C, sizeof(po_pair), po_pair_cmp);
    int best_cell = -1;
    float best = FLT_MAX;
    for (uint32_t j = 0; j < C; ++j) {
        int c = pairs[j].idx;
        float ct = pairs[j].term;
        for (uint32_t f = 0; f < F; ++f) {
            if (sqrtf(best) - fbound[f] < sqrtf(ct)) break;
            float score = ct - fip[f] + st[(size_t)c * F + f];
            if (score < best) { best = score; best_cell = c * (int)F + (int)f; }
        }
    }
    *cell_out = (uint32_t)best_cell;
    if (dist_out) *dist_out = best;
}

int po_assign_cells(const float* coarse, const float* fine, uint32_t C, uint32_t F, uint32_t dim,
                    const float* x, uint32_t n, uint32_t* cell_out, float* dist_out, int nthreads) {
    float* cn = (float*)malloc(sizeof(float) * C);
    float* fnorm = (float*)malloc(sizeof(float) * F);
    float* fbound = (float*)malloc(sizeof(float) * F);
    float* st = (float*)malloc(sizeof(float) * (size_t)C * F);
    for (uint32_t c = 0; c < C; ++c) cn[c] = po_dot(coarse + (size_t)c * dim, coarse + (size_t)c * dim, dim) / 2;
    for (uint32_t f = 0; f < F; ++f) {
        fnorm[f] = po_dot(fine + (size_t)f * dim, fine + (size_t)f * dim, dim) / 2;
        fbound[f] = 0 - sqrtf(fnorm[f]);
    }
    for (uint32_t c = 0; c < C; ++c)
        for (uint32_t f = 0; f < F; ++f)
            st[(size_t)c * F + f] = fnorm[f] + po_dot(fine + (size_t)f * dim, coarse + (size_t)c * dim, dim);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        po_pair* pairs = (po_pair*)malloc(sizeof(po_pair) * C);
        float* cip = (float*)malloc(sizeof(float) * C);
        float* fdist = (float*)malloc(sizeof(float) * F);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            const float* xi = x + (size_t)i * dim;
            float xn = po_dot(xi, xi, dim) / 2;
            for (uint32_t c = 0; c < C; ++c) cip[c] = po_dot(coarse + (size_t)c * dim, xi, dim);
            for (uint32_t f = 0; f < F; ++f) fdist[f] = po_dot(fine + (size_t)f * dim, xi, dim);
            po_assign_one(C, F, cn, st, fbound, cip, fdist, xn, pairs, cell_out + i, dist_out ? dist_out + i : NULL);
        }
        free(pairs); free(cip); free(fdist);
    }
    free(cn); free(fnorm); free(fbound); free(st);
    return 0;
}

/* Test hook: po_assign_cells with every arithmetic input supplied, e.g.
 * computed by a BLAS in its own summation order: coarse half-norms cn [C],
 * stationary st [C*F], fbound [F], per point the inner products cip [n*C] and
 * fip [n*F] (row per point) and the half squared norm xn [n]. */
int po_assign_cells_given(uint32_t C, uint32_t F, const float* cn, const float* st, const float* fbound, uint32_t n,
                          const float* cip, const float* fip, const float* xn, uint32_t* cell_out, float* dist_out,
                          int nthreads) {
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        po_pair* pairs = (po_pair*)malloc(sizeof(po_pair) * C);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int64_t i = 0; i < (int64_t)n; ++i)
            po_assign_one(C, F, cn, st, fbound, cip + (size_t)i * C, fip + (size_t)i * F, xn[i], pairs, cell_out + i,
                          dist_out ? dist_out + i : NULL);
        free(pairs);
    }
    return 0;
}

/* nearest_center (puck/hierarchical_cluster/kmeans.cpp:221-258):
 * dist = (float)((double)(pn + cn) - 2.0 * (double)sdot(x, c)), strict '<'
 * argmin starting from FLT_MAX; sdot is po_sdot_short (MKL's order for 1, 2
 * or 4 floats, po_dot for longer vectors). */
int po_nearest_center(uint32_t dim, const float* centroids, uint32_t k, const float* x,
                      uint32_t n, int32_t* assign, float* dis, int nthreads) {
    float* cnorm = (float*)malloc(sizeof(float) * k);
    for (uint32_t c = 0; c < k; ++c)
        cnorm[c] = po_sdot_short(centroids + (size_t)c * dim, centroids + (size_t)c * dim, dim);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (int64_t j = 0; j < (int64_t)n; ++j) {
        const float* xj = x + (size_t)j * dim;
        float pn = po_sdot_short(xj, xj, dim);
        float best = FLT_MAX;
        uint32_t arg = 0;
        for (uint32_t c = 0; c < k; ++c) {
            float d = (float)((double)(pn + cnorm[c]) -
                              2.0 * (double)po_sdot_short(xj, centroids + (size_t)c * dim, dim));
            if (d < best) { best = d; arg = c; }
        }
        assign[j] = (int32_t)arg;
        if (dis) dis[j] = best;
    }
    free(cnorm);
    return 0;
}

/* Filter-PQ and second-stage PQ encoding (whether_pq = true):
 * PuckIndex::puck_assign (puck_index.cpp:602-709) with quantizations =
 * {filter, pq}.  r = (x - c) - f; per filter subspace the nearest codeword
 * (nearest_center, kmeans.cpp:221-258) is chosen and subtracted from r
 * (saxpy with alpha -1, :672-675); the PQ codes are then chosen on that
 * residual and subtracted the same way.  Offsets
 * (Quantization::set_static_value_of_formula, quantization.cpp:244-258, in
 * the quantizers' order): v = c + f; filter offset = 2 sdot(dec_f, v);
 * v += dec_f; PQ offset = 2 sdot(dec_pq, v).  Both sub-vector lengths must be
 * 1, 2 or 4.  frec_out [n x (4 + fnsq)], prec_out [n x (4 + pnsq)]. */
int po_encode_puck(const float* coarse, const float* fine, uint32_t F, uint32_t dim, const float* fcb,
                   uint32_t fnsq, const float* pcb, uint32_t pnsq, const float* x, const uint32_t* cell, uint32_t n,
                   uint8_t* frec_out, uint8_t* prec_out, int nthreads) {
    if (fnsq == 0 || pnsq == 0 || dim % fnsq != 0 || dim % pnsq != 0) return -1;
    const uint32_t nq2[2] = {fnsq, pnsq}, lq2[2] = {dim / fnsq, dim / pnsq};
    const float* cb2[2] = {fcb, pcb};
    uint8_t* out2[2] = {frec_out, prec_out};
    float* cbn[2];
    for (int z = 0; z < 2; ++z) {
        if (!(lq2[z] == 1 || lq2[z] == 2 || lq2[z] == 4)) return -1;
        cbn[z] = (float*)malloc(sizeof(float) * nq2[z] * PO_KS);
        for (uint32_t e = 0; e < nq2[z] * PO_KS; ++e)
            cbn[z][e] = po_sdot_short(cb2[z] + (size_t)e * lq2[z], cb2[z] + (size_t)e * lq2[z], lq2[z]);
    }
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        float* r = (float*)malloc(sizeof(float) * dim);
        float* dec = (float*)malloc(sizeof(float) * dim * 2);
        float* v = (float*)malloc(sizeof(float) * dim);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            uint32_t c = cell[i] / F, f = cell[i] % F;
            const float* cc = coarse + (size_t)c * dim;
            const float* ff = fine + (size_t)f * dim;
            for (uint32_t j = 0; j < dim; ++j) {
                float t = fmaf(-1.0f, cc[j], x[(size_t)i * dim + j]);
                r[j] = fmaf(-1.0f, ff[j], t);
            }
            for (int z = 0; z < 2; ++z) {
                const uint32_t nsq = nq2[z], lsq = lq2[z];
                uint8_t* out = out2[z] + (size_t)i * (4 + nsq);
                for (uint32_t m = 0; m < nsq; ++m) {
                    float* sub = r + m * lsq;
                    const float* cb = cb2[z] + (size_t)m * PO_KS * lsq;
                    float pn = po_sdot_short(sub, sub, lsq);
                    float best = FLT_MAX;
                    uint32_t arg = 0;
                    for (uint32_t k = 0; k < PO_KS; ++k) {
                        float d = (float)((double)(pn + cbn[z][m * PO_KS + k]) -
                                          2.0 * (double)po_sdot_short(sub, cb + (size_t)k * lsq, lsq));
                        if (d < best) { best = d; arg = k; }
                    }
                    out[4 + m] = (uint8_t)arg;
                    for (uint32_t j = 0; j < lsq; ++j) {
                        dec[z * dim + m * lsq + j] = cb[(size_t)arg * lsq + j];
                        sub[j] = fmaf(-1.0f, cb[(size_t)arg * lsq + j], sub[j]);
                    }
                }
            }
            for (uint32_t j = 0; j < dim; ++j) v[j] = fmaf(1.0f, ff[j], cc[j]);
            float off = 2 * po_dot(dec, v, dim);
            memcpy(out2[0] + (size_t)i * (4 + fnsq), &off, 4);
            for (uint32_t j = 0; j < dim; ++j) v[j] = v[j] + dec[j];
            off = 2 * po_dot(dec + dim, v, dim);
            memcpy(out2[1] + (size_t)i * (4 + pnsq), &off, 4);
        }
        free(r); free(dec); free(v);
    }
    free(cbn[0]); free(cbn[1]);
    return 0;
}

/* Filter-PQ encoding of one chunk: PuckIndex::puck_assign
 * (puck_index.cpp:602-709) with quantizations = {filter} (whether_pq=false).
 * residual r = (x - c) - f (two saxpy with alpha -1); per subspace the
 * nearest codeword by nearest_center; then offset = 2*sdot(decode, c+f)
 * (Quantization::set_static_value_of_formula, quantization.cpp:244-258).
 * records_out: n x (4+nsq) in the filter_data.dat layout. */
int po_encode_filter(const float* coarse, const float* fine, uint32_t F, uint32_t dim,
                     const float* codebook, uint32_t nsq, const float* x, const uint32_t* cell,
                     uint32_t n, uint8_t* records_out, int nthreads) {
    if (nsq == 0 || dim % nsq != 0) return -1;
    uint32_t lsq = dim / nsq;
    size_t rec = 4 + nsq;
    float* cbn = (float*)malloc(sizeof(float) * nsq * PO_KS);
    for (uint32_t m = 0; m < nsq; ++m)
        for (uint32_t k = 0; k < PO_KS; ++k) {
            const float* cw = codebook + ((size_t)m * PO_KS + k) * lsq;
            cbn[m * PO_KS + k] = po_sdot_short(cw, cw, lsq);
        }
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        float* r = (float*)malloc(sizeof(float) * dim);
        float* dec = (float*)malloc(sizeof(float) * dim);
        float* v = (float*)malloc(sizeof(float) * dim);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            uint32_t c = cell[i] / F, f = cell[i] % F;
            const float* cc = coarse + (size_t)c * dim;
            const float* ff = fine + (size_t)f * dim;
            for (uint32_t j = 0; j < dim; ++j) {
                float t = fmaf(-1.0f, cc[j], x[(size_t)i * dim + j]);
                r[j] = fmaf(-1.0f, ff[j], t);
            }
            uint8_t* out = records_out + (size_t)i * rec;
            for (uint32_t m = 0; m < nsq; ++m) {
                const float* sub = r + m * lsq;
                const float* cb = codebook + (size_t)m * PO_KS * lsq;
                float pn = po_sdot_short(sub, sub, lsq);
                float best = FLT_MAX;
                uint32_t arg = 0;
                for (uint32_t k = 0; k < PO_KS; ++k) {
                    float d = (float)((double)(pn + cbn[m * PO_KS + k]) -
                                      2.0 * (double)po_sdot_short(sub, cb + (size_t)k * lsq, lsq));
                    if (d < best) { best = d; arg = k; }
                }
                out[4 + m] = (uint8_t)arg;
                for (uint32_t j = 0; j < lsq; ++j) dec[m * lsq + j] = cb[(size_t)arg * lsq + j];
            }
            for (uint32_t j = 0; j < dim; ++j) v[j] = fmaf(1.0f, ff[j], cc[j]);
            float off = 2 * po_dot(dec, v, dim);
            memcpy(out, &off, 4);
        }
        free(r); free(dec); free(v);
    }
    free(cbn);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* RealtimeInsertPuckIndex::insert (realtime_insert_puck_index.cpp:273-345)
 * for n points, one after another:
 *   single_build (hierarchical_cluster_index.cpp:1455-1494): normalisation
 *     (whether_norm; a norm below 1e-6 fails and takes no insert id), then
 *     nearest_cell_assign (:1042-1085) over the index's live stationary
 *     values -- an empty cell holds the sentinel and is not chosen while a
 *     populated cell is in reach;
 *   puck_single_assign (puck_index.cpp:938-988): filter codes (and the
 *     second-stage PQ codes of a whether_pq index) and offsets, as the build;
 *   the point goes to the front of its cell's list; a cell that had no
 *     point takes the stationary value sdot(c, f) - sdot(f, f) (:326-338),
 *     and min_dist_offset keeps its loaded value.
 * ids_out[i]: the local id N + insert ordinal, or PO_NONE for a rejected
 * point.  cells_out, rec_out [4 + nsq], pq_out [4 + pnsq] and x_out [dim]
 * (any may be NULL) receive the accepted points in order.  Returns the number
 * accepted. */
int po_index_insert(void* p, uint32_t n, const float* x, uint32_t* ids_out, uint32_t* cells_out, uint8_t* rec_out,
                    uint8_t* pq_out, float* x_out) {
    po_index* ix = (po_index*)p;
    const uint32_t dim = ix->dim, C = ix->C, F = ix->F, nsq = ix->nsq, pnsq = ix->pnsq;
    const size_t cells = (size_t)C * F, rec = 4 + nsq, prec = 4 + pnsq;
    if (!ix->ins_head) {
        ix->ins_head = (uint32_t*)malloc(sizeof(uint32_t) * cells);
        for (size_t i = 0; i < cells; ++i) ix->ins_head[i] = PO_NONE;
        ix->ins_cnt = (uint32_t*)calloc(cells, sizeof(uint32_t));
        ix->ins_before = (uint32_t*)calloc(cells, sizeof(uint32_t));
        /* _fine_norms of init (:245-254): -sqrt(sdot(f, f) / 2) */
        ix->fbound = (float*)malloc(sizeof(float) * F);
        for (uint32_t f = 0; f < F; ++f) {
            const float* ff = ix->fine + (size_t)f * dim;
            ix->fbound[f] = 0 - sqrtf(po_dot(ff, ff, dim) / 2);
        }
    }
    float* xq = (float*)malloc(sizeof(float) * dim);
    float* cip = (float*)malloc(sizeof(float) * C);
    float* fip = (float*)malloc(sizeof(float) * F);
    po_pair* pairs = (po_pair*)malloc(sizeof(po_pair) * C);
    uint32_t acc = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (ids_out) ids_out[i] = PO_NONE;
        const float* xi = x + (size_t)i * dim;
        if (ix->whether_norm) {
            if (po_normalize(xi, dim, xq) != 0) continue;
        } else {
            memcpy(xq, xi, sizeof(float) * dim);
        }
        const float xn = po_dot(xq, xq, dim) / 2;
        for (uint32_t c = 0; c < C; ++c) cip[c] = po_dot(ix->coarse + (size_t)c * dim, xq, dim);
        for (uint32_t f = 0; f < F; ++f) fip[f] = po_dot(ix->fine + (size_t)f * dim, xq, dim);
        uint32_t cell = PO_NONE;
        po_assign_one(C, F, ix->coarse_norms, ix->stationary, ix->fbound, cip, fip, xn, pairs, &cell, NULL);
        if (cell >= cells) continue;
        if (ix->ins_n == ix->ins_cap) {
            uint32_t cap = ix->ins_cap ? 2 * ix->ins_cap : 64;
            ix->ins_cell = (uint32_t*)realloc(ix->ins_cell, sizeof(uint32_t) * cap);
            ix->ins_next = (uint32_t*)realloc(ix->ins_next, sizeof(uint32_t) * cap);
            ix->ins_ord = (uint32_t*)realloc(ix->ins_ord, sizeof(uint32_t) * cap);
            ix->ins_rec = (uint8_t*)realloc(ix->ins_rec, rec * cap);
            if (pnsq) ix->ins_pq = (uint8_t*)realloc(ix->ins_pq, prec * cap);
            else ix->ins_raw = (float*)realloc(ix->ins_raw, sizeof(float) * dim * cap);
            ix->ins_cap = cap;
        }
        const uint32_t j = ix->ins_n++;
        if (pnsq)
            po_encode_puck(ix->coarse, ix->fine, F, dim, ix->codebook, nsq, ix->pq_codebook, pnsq, xq, &cell, 1,
                           ix->ins_rec + (size_t)j * rec, ix->ins_pq + (size_t)j * prec, 1);
        else
            po_encode_filter(ix->coarse, ix->fine, F, dim, ix->codebook, nsq, xq, &cell, 1,
                             ix->ins_rec + (size_t)j * rec, 1);
        if (!pnsq) memcpy(ix->ins_raw + (size_t)j * dim, xq, sizeof(float) * dim);
        ix->ins_cell[j] = cell;
        ix->ins_ord[j] = ix->ins_cnt[cell];
        ix->ins_next[j] = ix->ins_head[cell];
        ix->ins_head[cell] = j;
        ix->ins_cnt[cell]++;
        if (ix->cell_start[cell + 1] - ix->cell_start[cell] + ix->ins_cnt[cell] == 1) {
            const float* cc = ix->coarse + (size_t)(cell / F) * dim;
            const float* ff = ix->fine + (size_t)(cell % F) * dim;
            float sv = po_dot(cc, ff, dim);
            sv -= po_dot(ff, ff, dim);
            ix->stationary[cell] = sv;
        }
        if (ids_out) ids_out[i] = ix->N + j;
        if (cells_out) cells_out[acc] = cell;
        if (rec_out) memcpy(rec_out + (size_t)acc * rec, ix->ins_rec + (size_t)j * rec, rec);
        if (pq_out && pnsq) memcpy(pq_out + (size_t)acc * prec, ix->ins_pq + (size_t)j * prec, prec);
        if (x_out) memcpy(x_out + (size_t)acc * dim, xq, sizeof(float) * dim);
        ++acc;
    }
    uint32_t run = 0;
    for (size_t c = 0; c < cells; ++c) {
        ix->ins_before[c] = run;
        run += ix->ins_cnt[c];
    }
    free(xq); free(cip); free(fip); free(pairs);
    return (int)acc;
}

/* Memory position of heap tag `tag` in the layout where every cell holds its
 * loaded points, then its inserted points newest first (the budget scan's
 * tie order; the memory id itself while nothing is inserted). */
uint32_t po_index_position(const void* p, uint32_t tag) {
    const po_index* ix = (const po_index*)p;
    if (!ix->ins_head) return tag;
    if (tag >= ix->N) {
        const uint32_t j = tag - ix->N, cell = ix->ins_cell[j];
        return ix->cell_start[cell + 1] + ix->ins_before[cell] + (ix->ins_cnt[cell] - 1 - ix->ins_ord[j]);
    }
    /* the cell holding memory id tag: the last cell starting at or before it
     * among the non-empty ones */
    size_t lo = 0, hi = (size_t)ix->C * ix->F;
    while (hi - lo > 1) {
        size_t mid = (lo + hi) / 2;
        if (ix->cell_start[mid] <= tag) lo = mid; else hi = mid;
    }
    while (lo > 0 && ix->cell_start[lo] == ix->cell_start[lo + 1]) --lo;
    return tag + ix->ins_before[lo];
}

uint32_t po_index_inserted(const void* p) { return ((const po_index*)p)->ins_n; }
