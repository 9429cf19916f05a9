// capi.hip -- the C ABI of libpuckgpu.so (declared in include/puck_gpu.h).
// Host-side orchestration of the gfx950 kernels: context, index load, batched
// search, build-time assignment and encoding.
#include "common.h"
#include "launch_args.h"
#include "../../include/puck_gpu.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <limits>
#include <random>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

// ---- kernel launchers (defined in gemm/sort/search/build.hip)
extern "C" {
hipError_t pk_launch_gemm_exact(const float*, const float*, float*, uint32_t, uint32_t, uint32_t, uint64_t, int,
                                const float*, hipStream_t);
hipError_t pk_launch_row_norms(const float*, uint32_t, uint32_t, int, float*, hipStream_t);
hipError_t pk_launch_sort_rows(const float*, uint32_t, uint32_t, uint64_t, uint32_t, float*, uint32_t*, uint8_t*,
                               uint32_t, hipStream_t);
hipError_t pk_launch_sort_rows_list(const float*, uint32_t, uint32_t, uint64_t, uint32_t, float*, uint32_t*,
                                    const uint32_t*, const uint32_t*, hipStream_t);
hipError_t pk_launch_topk_rows(const float*, uint32_t, uint32_t, uint64_t, uint32_t, float*, uint32_t*, uint8_t*,
                               uint32_t, hipStream_t);
hipError_t pk_launch_cell_topk(const float*, const uint32_t*, uint32_t, const float*, const float*, uint32_t, uint32_t,
                               uint32_t, float*, uint32_t*, hipStream_t);
hipError_t pk_launch_exact_center_order(const float*, uint32_t, const float*, uint32_t, uint32_t, uint32_t,
                                        const uint32_t*, const uint32_t*, uint32_t, const uint8_t*, int, float*,
                                        uint32_t*, float*, uint32_t*, hipStream_t);
hipError_t pk_launch_exact_center_order_dev(const float*, uint32_t, const float*, uint32_t, uint32_t, uint32_t,
                                            const uint32_t*, const uint8_t*, uint32_t, uint32_t, uint32_t*, uint32_t*,
                                            float*, uint32_t*, float*, uint32_t*, hipStream_t);
hipError_t pk_launch_prep_queries(const float*, uint32_t, uint32_t, int, float*, float*, int32_t*, hipStream_t);
hipError_t pk_launch_prep_queries_ex(const float*, uint32_t, uint32_t, int, float*, float*, int32_t*, const uint32_t*,
                                     uint8_t*, hipStream_t);
hipError_t pk_launch_gemm_exact_shape(const float*, const float*, float*, uint32_t, uint32_t, uint32_t, uint64_t, int,
                                      const float*, int, hipStream_t);
uint32_t pk_gemm_shapes(void);
uint32_t pkc_ld(uint32_t);
hipError_t pkc_launch_convert(const float*, uint32_t, uint32_t, int, uint16_t*, uint8_t*, float*, hipStream_t);
hipError_t pkc_launch_center(const void*, const float*, const void*, const float*, uint32_t, uint32_t, uint32_t, int,
                             int, const float*, float*, uint64_t, hipStream_t);
hipError_t pkc_launch_rescore_coarse(const float*, uint32_t, const float*, const float*, const uint32_t*, uint32_t,
                                     uint32_t, uint32_t, float*, uint32_t*, uint8_t*, uint32_t, hipStream_t);
hipError_t pk_launch_lut(const float*, uint32_t, uint32_t, const float*, uint32_t, uint32_t, float*, hipStream_t);
hipError_t pk_launch_walk(const PkWalkArgs*, uint32_t, hipStream_t, hipEvent_t, hipStream_t, hipEvent_t, hipEvent_t);
hipError_t pk_launch_rerank(const PkRerankArgs*, uint32_t, hipStream_t);
uint32_t pk_walk_ftk_max(void);
uint32_t pk_probe_trained_count(uint32_t, uint32_t);
uint32_t pk_probe_param_count(uint32_t, uint32_t);
size_t pk_probe_work_bytes(uint32_t, uint32_t, uint32_t);
hipError_t pk_launch_probe_raw_features(const float*, uint32_t, const float*, uint32_t, uint32_t, const float*,
                                        uint32_t, uint32_t, float*, hipStream_t);
hipError_t pk_probe_pass(const float*, const uint16_t*, uint32_t, uint32_t, uint32_t, const float*, const float*,
                         void*, float*, float*, float*, hipStream_t);
hipError_t pk_probe_adam_step(float*, float*, float*, const float*, uint16_t*, uint32_t, uint32_t, float, float,
                              float, float, uint32_t, hipStream_t);
hipError_t pk_launch_probe_budget(const float*, uint32_t, float, uint32_t, float, uint32_t, uint32_t*,
                                  hipStream_t);
hipError_t pk_launch_to_bf16(const float*, uint32_t, uint16_t*, hipStream_t);
hipError_t pk_launch_build_query_order(const uint32_t*, uint32_t, uint32_t, uint32_t, uint32_t*, hipStream_t);
hipError_t pk_launch_build_walk_list(const uint32_t*, uint32_t, uint32_t, uint32_t*, uint32_t*, hipStream_t);
hipError_t pk_launch_budget_rank(const float*, const uint32_t*, uint32_t, uint32_t, const float*, const float*,
                                 uint32_t, uint32_t, const uint32_t*, uint32_t, uint32_t, uint32_t*, hipStream_t);
hipError_t pk_launch_assign_cells(const float*, const float*, const float*, const float*, const float*,
                                  const float*, uint32_t, uint32_t, uint32_t, int, const uint32_t*, uint32_t*,
                                  float*, hipStream_t);
hipError_t pk_launch_st_blocks(const float*, uint32_t, uint32_t, float*, float*, uint32_t*, hipStream_t);
hipError_t pk_launch_u8_to_f32(const uint8_t*, uint64_t, float*, hipStream_t);
hipError_t pk_launch_exact_append(const float*, const uint32_t*, uint32_t, uint32_t, uint32_t, float*, uint32_t*,
                                  uint32_t, uint32_t, hipStream_t);
hipError_t pk_launch_exact_recheck(const void*, int, uint32_t, const float*, uint32_t, const uint32_t*,
                                   const uint32_t*, uint32_t, uint32_t, uint32_t, uint32_t*, double*, hipStream_t);
hipError_t pk_launch_assign_bounded(const float*, const float*, const float*, const float*, const float*, const float*,
                                    const float*, const float*, const float*, const uint32_t*, uint32_t, uint32_t,
                                    uint32_t, int, int, uint32_t*, float*, uint32_t*, uint32_t*, hipStream_t);
hipError_t pk_launch_nearest_coarse(const float*, const float*, const float*, uint32_t, uint32_t, uint32_t*,
                                    hipStream_t);
hipError_t pk_launch_fine_bound(const float*, uint32_t, uint32_t, float*, hipStream_t);
hipError_t pk_launch_encode_pq(const float*, const uint32_t*, uint32_t, const float*, const float*, uint32_t, uint32_t,
                               const float*, uint32_t, const uint8_t*, uint64_t, const float*, uint32_t, uint8_t*,
                               hipStream_t);
hipError_t pk_launch_encode_filter(const float*, const uint32_t*, uint32_t, const float*, const float*, uint32_t,
                                   uint32_t, const float*, uint32_t, float*, uint8_t*, uint64_t, uint64_t,
                                   hipStream_t);
hipError_t pk_launch_kmeans_argmin(const float*, const float*, const float*, uint32_t, uint32_t, int32_t*, float*,
                                   hipStream_t);
hipError_t pk_launch_kmeans_short(const float*, uint32_t, uint32_t, const float*, uint32_t, int32_t*, float*,
                                  hipStream_t);
hipError_t pk_launch_finalize_cells(float*, const uint32_t*, uint32_t, uint32_t, float*, hipStream_t);
hipError_t pk_launch_gather_rows(const uint8_t*, uint64_t, uint64_t, uint8_t*, uint64_t, uint32_t,
                                 const uint32_t*, uint32_t, hipStream_t);
hipError_t pk_launch_gather_f32(const float*, uint32_t, float*, const uint32_t*, uint32_t, hipStream_t);
uint64_t pk_exscan_work(uint64_t);
hipError_t pk_launch_insert_prepare(const uint32_t*, uint32_t, uint32_t, uint32_t*, uint32_t*, uint32_t*, hipStream_t);
hipError_t pk_launch_insert_dest(const uint32_t*, uint32_t, const uint32_t*, const uint32_t*, const uint32_t*, uint32_t,
                                 uint32_t*, const uint32_t*, const uint32_t*, uint32_t, uint32_t*, uint32_t*, uint32_t*,
                                 hipStream_t);
hipError_t pk_launch_cell_sizes(const uint32_t*, uint32_t, uint32_t*, hipStream_t);
hipError_t pk_launch_insert_stationary(const float*, const float*, uint32_t, uint32_t, const uint32_t*, uint32_t, float*,
                                       hipStream_t);
hipError_t pk_launch_scatter_rows(const void*, uint64_t, uint64_t, void*, uint64_t, uint32_t, const uint32_t*, uint32_t,
                                  hipStream_t);
uint64_t pk_ks_hist_words(uint32_t, uint32_t);
hipError_t pk_launch_sort_by_cluster(const int32_t*, uint32_t, uint32_t, uint32_t*, uint32_t*, uint32_t*, uint32_t*,
                                     uint32_t*, hipStream_t);
hipError_t pk_launch_kmeans_update(const float*, uint32_t, const uint32_t*, const uint32_t*, uint32_t, float*, int,
                                   hipStream_t);
hipError_t pk_launch_sub_assigned(const float*, uint32_t, uint32_t, const float*, const int32_t*, float*, hipStream_t);
hipError_t pk_launch_norms_sdot(const float*, uint32_t, uint32_t, float*, hipStream_t);
hipError_t pk_launch_kmeanspp_step(const float*, const float*, uint32_t, uint32_t, uint32_t, float*, hipStream_t);
}

enum { PK_EPI_RAW = 0, PK_EPI_SUB_FROM_BIAS = 1, PK_EPI_NEG = 2, PK_EPI_ADD_BIAS = 3 };


namespace {

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

}  // namespace

struct puckgpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;        // the fallback walk beside the budget scan
    hipEvent_t fork_ev = nullptr, join_ev = nullptr, order_ev = nullptr;
    hipEvent_t eorder_ev = nullptr;  // exact center orders of the batch's tied queries done (second stream)
    uint8_t* hpin = nullptr;  // pinned host staging of per-query flags and walk counters
    size_t hpin_cap = 0;
    std::string err;
    std::vector<DevBuf> bufs;          // named scratch slots
    hipEvent_t ev[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    float stage_ms[5] = {0, 0, 0, 0, 0};
    float scan_ms = 0;  // budget-scan launch alone (inside stage 2)
    uint64_t counters[13] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t last_n = 0, last_ftk = 0;
    unsigned char* noalloc = nullptr;  // this device's copy of `scratch` (puckgpu_search_noalloc)
    int assign_mode = 0;                // puckgpu_set_assign_mode
    uint32_t rerank_rows = 2;           // puckgpu_set_rerank_rows
    uint64_t assign_points = 0, assign_sequential = 0;
    puckgpu_progress_fn progress = nullptr;  // puckgpu_set_progress
    void* progress_user = nullptr;
};

// A progress line to the caller's callback (long calls: training).
void report(puckgpu_ctx* ctx, const char* fmt, ...) {
    if (!ctx || !ctx->progress) return;
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    ctx->progress(buf, ctx->progress_user);
}

struct puckgpu_index {
    puckgpu_index_params p{};
    uint32_t lsq = 0;
    float* coarse = nullptr;
    float* fine = nullptr;
    float* codebook = nullptr;
    float cb_nmax = 0.f;  // largest filter codeword norm (scale of the FP16 table)
    float* coarse_norms = nullptr;
    float* stationary = nullptr;
    uint32_t* cell_start = nullptr;
    float* mdo = nullptr;
    uint32_t* mem2local = nullptr;
    float* rec_off = nullptr;
    uint8_t* rec_codes = nullptr;
    float* raw = nullptr;  // null when the index holds no raw vectors (whether_pq = true)
    uint8_t* raw8 = nullptr;  // uint8 raw rows (puckgpu_index_attach_raw_u8) in place of raw
    // second-stage PQ (whether_pq = true), attached by puckgpu_index_attach_pq
    uint32_t pnsq = 0, plsq = 0;
    float* pq_codebook = nullptr;
    float* pq_off = nullptr;
    uint8_t* pq_codes = nullptr;
    // reduced-precision copies of the centers (built on first use)
    mutable uint16_t* coarse16 = nullptr;
    mutable uint16_t* fine16 = nullptr;
    mutable uint8_t* coarse8 = nullptr;
    mutable uint8_t* fine8 = nullptr;
    mutable float* coarse8s = nullptr;
    mutable float* fine8s = nullptr;
    // realtime insert (puckgpu_index_insert), set up by the first insert:
    // the loaded points per cell, the host count of points per cell, the
    // assignment's fine bound; version counts the layout changes
    uint32_t* nbase = nullptr;
    std::vector<uint32_t> h_cnt;
    float* fbound = nullptr;
    int fb_sorted = 1;
    uint64_t version = 0;
};

struct puckgpu_probe {
    int device = 0;
    uint32_t dim = 0, nc = 0, nfine = 0, D = 0, H = 0;
    uint32_t count = 0, trained = 0, step = 0;
    float* params = nullptr;  // packed FP32 (trained block, then mu, 1/sigma)
    float* m = nullptr;       // Adam moments (trained block)
    float* v = nullptr;
    float* grads = nullptr;
    float* loss = nullptr;
    uint16_t* wbf = nullptr;  // BF16 copy of the trained block
};

struct PkBucket;

// Pass buffers of puckgpu_search_noalloc, carved from a file-scope device
// array: each device holds its own copy for the life of the process.  Host
// code reaches it only through the address hipGetSymbolAddress gives on that
// device (puckgpu_create keeps it in the context).  Calls on one device share
// it; calls that overlap on one device race.
#define PK_NOALLOC_BYTES (64ull << 20)
#define PK_NOALLOC_MAXN 4096u
static __device__ __align__(256) unsigned char scratch[PK_NOALLOC_BYTES];
// Host staging of the same calls (flags, counters, replay lists), per device.
static uint32_t pk_noalloc_hstats[PK_MAX_DEVICES][4 * PK_NOALLOC_MAXN];
static uint8_t pk_noalloc_hflags[PK_MAX_DEVICES][PK_NOALLOC_MAXN];
static uint32_t pk_noalloc_hlist[PK_MAX_DEVICES][2 * PK_NOALLOC_MAXN];

struct puckgpu_server {
    puckgpu_ctx* ctx = nullptr;
    const puckgpu_index* ix = nullptr;
    puckgpu_search_params sp{};
    const puckgpu_probe* probe = nullptr;
    puckgpu_budget_params bp{};
    uint32_t max_batch = 0;
    std::vector<PkBucket*> buckets;
};

namespace {

int set_err(puckgpu_ctx* ctx, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

#define PK_CHECK(expr)                                                                                   \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            return set_err(ctx, PUCKGPU_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),   \
                           __FILE__, __LINE__);                                                          \
    } while (0)

// scratch slot i, at least `bytes` long
void* scratch_slot(puckgpu_ctx* ctx, size_t slot, size_t bytes, int* rc) {
    if (ctx->bufs.size() <= slot) ctx->bufs.resize(slot + 1);
    DevBuf& b = ctx->bufs[slot];
    if (b.cap < bytes) {
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
        size_t want = bytes < 256 ? 256 : bytes;
        if (hipMalloc(&b.p, want) != hipSuccess) {
            *rc = set_err(ctx, PUCKGPU_ENOMEM, "hipMalloc(%zu) failed for scratch slot %zu", want, slot);
            return nullptr;
        }
        b.cap = want;
    }
    return b.p;
}

enum Slot {
    S_Q = 0, S_QN, S_QST, S_CD, S_FD, S_CSD, S_CSI, S_FSD, S_FSI, S_FT, S_FI, S_FN, S_OD, S_OI, S_ON, S_STATS,
    S_ELIST, S_IN_Q, S_A0, S_A1, S_A2, S_A3, S_A4, S_A5, S_A6, S_A7, S_FLAGS, S_QLIST, S_TIEID, S_TIEN, S_QTOPK, S_QSCC,
    S_QNF, S_PFEAT, S_PWORK, S_PNF, S_PY, S_PT, S_PCV, S_PCID, S_LP16, S_LP8, S_LPS, S_CAND_D, S_CAND_I, S_WLIST, S_SCI,
    S_SCM, S_INS_X, S_INS_Y, S_INS_CELL, S_INS_REC, S_INS_PQ, S_INS_RANK, S_INS_ADD, S_INS_SHIFT, S_INS_WORK,
    S_INS_DOLD, S_INS_DNEW, S_INS_ACC, S_KM_IP, S_KM_PN, S_KM_CN, S_KM_ASSIGN, S_KM_DIS, S_KM_PERM, S_KM_STARTS,
    S_KM_HIST, S_KM_OFF, S_KM_WORK, S_KM_BEST, S_KM_IDS, S_KM_X, S_KM_TV, S_KM_C0, S_KM_F0, S_KM_CB, S_KM_FB,
    S_KM_FA, S_KM_CENT, S_AB_STMIN, S_AB_ROW, S_AB_ABS, S_AB_FBPM, S_AB_LIST, S_AB_CNT, S_EX_X, S_EX_Q, S_EX_D,
    S_EX_XF, S_EX_XN, S_EX_TV, S_EX_TI, S_EX_CV, S_EX_CI, S_EX_MV, S_EX_MP, S_EX_OI, S_EX_OD, S_COUNT
};

// Device view of an input array: the pointer itself, or a scratch copy of it.
const void* dev_in(puckgpu_ctx* ctx, size_t slot, const void* p, size_t bytes, int device_ptrs, int* rc) {
    if (device_ptrs || bytes == 0) return p;
    void* d = scratch_slot(ctx, slot, bytes, rc);
    if (!d) return nullptr;
    if (hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
        *rc = set_err(ctx, PUCKGPU_EHIP, "host-to-device copy of %zu bytes failed", bytes);
        return nullptr;
    }
    return d;
}

template <typename T>
int dalloc(puckgpu_ctx* ctx, T** p, size_t count) {
    size_t bytes = sizeof(T) * (count ? count : 1);
    if (hipMalloc((void**)p, bytes) != hipSuccess)
        return set_err(ctx, PUCKGPU_ENOMEM, "hipMalloc(%zu) failed", bytes);
    return 0;
}

void index_release(puckgpu_index* ix) {
    if (!ix) return;
    void* ptrs[] = {ix->coarse, ix->fine, ix->codebook, ix->coarse_norms, ix->stationary, ix->cell_start,
                    ix->mdo, ix->mem2local, ix->rec_off, ix->rec_codes, ix->raw, ix->coarse16, ix->fine16,
                    ix->coarse8, ix->fine8, ix->coarse8s, ix->fine8s, ix->pq_codebook, ix->pq_off, ix->pq_codes,
                    ix->nbase, ix->fbound, ix->raw8};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete ix;
}

}  // namespace

extern "C" {

uint32_t puckgpu_filter_topk_max(void) { return pk_walk_ftk_max(); }

int puckgpu_create(int device, puckgpu_ctx** out) {
    if (!out) return PUCKGPU_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return PUCKGPU_EHIP;
    if (device < 0 || device >= n) return PUCKGPU_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return PUCKGPU_EHIP;
    puckgpu_ctx* ctx = new puckgpu_ctx();
    ctx->device = device;
    bool ok = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) == hipSuccess;
    for (auto& e : ctx->ev) ok = ok && hipEventCreate(&e) == hipSuccess;
    ok = ok && hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&ctx->join_ev, hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&ctx->order_ev, hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&ctx->eorder_ev, hipEventDisableTiming) == hipSuccess;
    // this device's copy of the no-alloc scratch array (its name is a host
    // stand-in; only this address may be handed to kernels or copies)
    ok = ok && hipGetSymbolAddress((void**)&ctx->noalloc, HIP_SYMBOL(scratch)) == hipSuccess && ctx->noalloc;
    if (!ok) {
        puckgpu_destroy(ctx);  // releases whatever was created
        return PUCKGPU_EHIP;
    }
    *out = ctx;
    return PUCKGPU_OK;
}

void puckgpu_destroy(puckgpu_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& b : ctx->bufs)
        if (b.p) (void)hipFree(b.p);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->hpin) (void)hipHostFree(ctx->hpin);
    if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
    if (ctx->join_ev) (void)hipEventDestroy(ctx->join_ev);
    if (ctx->order_ev) (void)hipEventDestroy(ctx->order_ev);
    if (ctx->eorder_ev) (void)hipEventDestroy(ctx->eorder_ev);
    if (ctx->stream2) {
        (void)hipStreamSynchronize(ctx->stream2);
        (void)hipStreamDestroy(ctx->stream2);
    }
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* puckgpu_last_error(const puckgpu_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

// ---------------------------------------------------------------------------
int puckgpu_index_load(puckgpu_ctx* ctx, const puckgpu_index_params* params, const float* coarse, const float* fine,
                       const float* filter_codebook, const uint32_t* cell_assign, const uint8_t* filter_records,
                       const float* features, int device_ptrs, puckgpu_index** out) {
    if (!ctx || !params || !out) return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    *out = nullptr;
    const puckgpu_index_params p = *params;
    if (p.dim == 0 || p.coarse_cluster_count == 0 || p.fine_cluster_count == 0 || p.filter_nsq == 0)
        return set_err(ctx, PUCKGPU_EINVAL, "index geometry must be non-zero");
    if (p.dim % p.filter_nsq != 0 || !(p.dim / p.filter_nsq == 1 || p.dim / p.filter_nsq == 2 ||
                                       p.dim / p.filter_nsq == 4))
        return set_err(ctx, PUCKGPU_EINVAL, "dim / filter_nsq must be 1, 2 or 4 (dim=%u nsq=%u)", p.dim,
                       p.filter_nsq);
    const uint64_t cells = (uint64_t)p.coarse_cluster_count * p.fine_cluster_count;
    if (cells >= 0xffffffffull) return set_err(ctx, PUCKGPU_EINVAL, "C*F too large");
    if (p.fine_cluster_count > 8192 || p.coarse_cluster_count > 8192)
        return set_err(ctx, PUCKGPU_EINVAL, "C and F must be <= 8192");
    if (!coarse || !fine || !filter_codebook || (p.total_point_count && (!cell_assign || !filter_records)))
        return set_err(ctx, PUCKGPU_EINVAL, "null array");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    const uint32_t C = p.coarse_cluster_count, F = p.fine_cluster_count, d = p.dim, nsq = p.filter_nsq,
                   N = p.total_point_count;
    const uint32_t lsq = d / nsq;
    const size_t rec = 4 + nsq;
    hipStream_t s = ctx->stream;
    int rc = 0;

    puckgpu_index* ix = new puckgpu_index();
    ix->p = p;
    ix->lsq = lsq;
#define PK_TRY(x)                  \
    do {                           \
        int r_ = (x);              \
        if (r_) {                  \
            index_release(ix);     \
            return r_;             \
        }                          \
    } while (0)
#define PK_TRYH(x)                                                                                  \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) {                                                                     \
            index_release(ix);                                                                      \
            return set_err(ctx, PUCKGPU_EHIP, "%s failed: %s", #x, hipGetErrorString(e_));          \
        }                                                                                           \
    } while (0)
    PK_TRY(dalloc(ctx, &ix->coarse, (size_t)C * d));
    PK_TRY(dalloc(ctx, &ix->fine, (size_t)F * d));
    PK_TRY(dalloc(ctx, &ix->codebook, (size_t)nsq * PK_KS * lsq));
    PK_TRY(dalloc(ctx, &ix->coarse_norms, C));
    PK_TRY(dalloc(ctx, &ix->stationary, cells));
    PK_TRY(dalloc(ctx, &ix->cell_start, cells + 1));
    PK_TRY(dalloc(ctx, &ix->mdo, C));
    PK_TRY(dalloc(ctx, &ix->mem2local, N));
    PK_TRY(dalloc(ctx, &ix->rec_off, N));
    PK_TRY(dalloc(ctx, &ix->rec_codes, (size_t)N * nsq));
    if (features) PK_TRY(dalloc(ctx, &ix->raw, (size_t)N * d));
    const hipMemcpyKind kind = device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    PK_TRYH(hipMemcpyAsync(ix->coarse, coarse, sizeof(float) * C * d, kind, s));
    PK_TRYH(hipMemcpyAsync(ix->fine, fine, sizeof(float) * F * d, kind, s));
    PK_TRYH(hipMemcpyAsync(ix->codebook, filter_codebook, sizeof(float) * nsq * PK_KS * lsq, kind, s));
    {
        // largest codeword norm: sqrt of the largest sum of rounded squares,
        // summed in order (the FP16 table's scale bound)
        std::vector<float> hcb((size_t)nsq * PK_KS * lsq);
        PK_TRYH(hipMemcpyAsync(hcb.data(), ix->codebook, sizeof(float) * hcb.size(), hipMemcpyDeviceToHost, s));
        PK_TRYH(hipStreamSynchronize(s));
        float mx = 0.f;
        for (size_t k = 0; k < (size_t)nsq * PK_KS; ++k) {
            float sq = 0.f;
            for (uint32_t j = 0; j < lsq; ++j) sq = sq + hcb[k * lsq + j] * hcb[k * lsq + j];
            if (sq > mx) mx = sq;
        }
        ix->cb_nmax = sqrtf(mx);
    }

    // |c|^2/2, |f|^2/2 and stationary[c][f] = |f|^2/2 + <f,c> (read_coodbooks)
    float* fnorm = (float*)scratch_slot(ctx, S_A0, sizeof(float) * F, &rc);
    if (!fnorm) { index_release(ix); return rc; }
    PK_TRYH(pk_launch_row_norms(ix->coarse, C, d, 1, ix->coarse_norms, s));
    PK_TRYH(pk_launch_row_norms(ix->fine, F, d, 1, fnorm, s));
    PK_TRYH(pk_launch_gemm_exact(ix->coarse, ix->fine, ix->stationary, C, F, d, F, PK_EPI_ADD_BIAS, fnorm, s));

    // memory order: cells ascending, then offset, then local id
    // (convert_local_to_memory_idx, puck_index.cpp:248-282)
    std::vector<uint32_t> h_cell(N);
    std::vector<float> h_off(N);
    const uint8_t* rec_dev = (const uint8_t*)dev_in(ctx, S_A1, filter_records, rec * N, device_ptrs, &rc);
    if (N && !rec_dev) { index_release(ix); return rc; }
    const float* feat_dev =
        features ? (const float*)dev_in(ctx, S_A2, features, sizeof(float) * N * d, device_ptrs, &rc) : nullptr;
    if (N && features && !feat_dev) { index_release(ix); return rc; }
    if (N) {
        if (device_ptrs) {
            PK_TRYH(hipMemcpyAsync(h_cell.data(), cell_assign, sizeof(uint32_t) * N, hipMemcpyDeviceToHost, s));
            // offsets: first 4 bytes of each record
            float* offs = (float*)scratch_slot(ctx, S_A3, sizeof(float) * N, &rc);
            if (!offs) { index_release(ix); return rc; }
            std::vector<uint32_t> ident;  // gather with identity map via a device iota
            uint32_t* iota = (uint32_t*)scratch_slot(ctx, S_A4, sizeof(uint32_t) * N, &rc);
            if (!iota) { index_release(ix); return rc; }
            ident.resize(N);
            for (uint32_t i = 0; i < N; ++i) ident[i] = i;
            PK_TRYH(hipMemcpyAsync(iota, ident.data(), sizeof(uint32_t) * N, hipMemcpyHostToDevice, s));
            PK_TRYH(pk_launch_gather_rows(rec_dev, rec, 0, (uint8_t*)offs, 4, 4, iota, N, s));
            PK_TRYH(hipMemcpyAsync(h_off.data(), offs, sizeof(float) * N, hipMemcpyDeviceToHost, s));
            PK_TRYH(hipStreamSynchronize(s));
        } else {
            std::memcpy(h_cell.data(), cell_assign, sizeof(uint32_t) * N);
            for (uint32_t i = 0; i < N; ++i) std::memcpy(&h_off[i], filter_records + (size_t)i * rec, 4);
        }
    }
    std::vector<uint32_t> h_cs(cells + 1, 0);
    for (uint32_t i = 0; i < N; ++i) {
        if (h_cell[i] >= cells) {
            index_release(ix);
            return set_err(ctx, PUCKGPU_EINVAL, "cell_assign[%u] = %u out of range", i, h_cell[i]);
        }
        h_cs[h_cell[i] + 1]++;
    }
    for (uint64_t c = 0; c < cells; ++c) h_cs[c + 1] += h_cs[c];
    std::vector<uint32_t> h_m2l(N);
    {
        std::vector<uint32_t> fill(h_cs.begin(), h_cs.end() - 1);
        for (uint32_t i = 0; i < N; ++i) h_m2l[fill[h_cell[i]]++] = i;  // stable by local id
        for (uint64_t c = 0; c < cells; ++c) {
            uint32_t b = h_cs[c], e = h_cs[c + 1];
            if (e - b > 1)
                std::stable_sort(h_m2l.begin() + b, h_m2l.begin() + e,
                                 [&](uint32_t x, uint32_t y) { return h_off[x] < h_off[y]; });
        }
    }
    PK_TRYH(hipMemcpyAsync(ix->cell_start, h_cs.data(), sizeof(uint32_t) * (cells + 1), hipMemcpyHostToDevice, s));
    if (N) PK_TRYH(hipMemcpyAsync(ix->mem2local, h_m2l.data(), sizeof(uint32_t) * N, hipMemcpyHostToDevice, s));
    PK_TRYH(pk_launch_finalize_cells(ix->stationary, ix->cell_start, C, F, ix->mdo, s));
    if (N) {
        PK_TRYH(pk_launch_gather_rows(rec_dev, rec, 0, (uint8_t*)ix->rec_off, 4, 4, ix->mem2local, N, s));
        PK_TRYH(pk_launch_gather_rows(rec_dev, rec, 4, ix->rec_codes, nsq, nsq, ix->mem2local, N, s));
        if (feat_dev) PK_TRYH(pk_launch_gather_f32(feat_dev, d, ix->raw, ix->mem2local, N, s));
    }
    PK_TRYH(hipStreamSynchronize(s));
    // host staging copies of records / features are no longer needed
    if (!device_ptrs)
        for (size_t slot : {(size_t)S_A1, (size_t)S_A2})
            if (slot < ctx->bufs.size() && ctx->bufs[slot].p) {
                (void)hipFree(ctx->bufs[slot].p);
                ctx->bufs[slot] = DevBuf();
            }
    *out = ix;
    return PUCKGPU_OK;
#undef PK_TRY
#undef PK_TRYH
}

void puckgpu_index_free(puckgpu_index* index) { index_release(index); }

int puckgpu_index_attach_raw_u8(puckgpu_ctx* ctx, puckgpu_index* ix, const uint8_t* features, int device_ptrs) {
    if (!ctx || !ix || (ix->p.total_point_count && !features)) return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    if (ix->raw || ix->raw8 || ix->pq_codes)
        return set_err(ctx, PUCKGPU_EINVAL, "the index already holds raw vectors or a second-stage PQ");
    if (ix->p.whether_norm)
        return set_err(ctx, PUCKGPU_EINVAL, "uint8 raw vectors need whether_norm = 0 (normalised rows are not integers)");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    hipStream_t s = ctx->stream;
    const uint32_t d = ix->p.dim, N = ix->p.total_point_count;
    int rc = 0;
    uint8_t* r8 = nullptr;
    if ((rc = dalloc(ctx, &r8, (size_t)N * d))) return rc;
    const uint8_t* src = (const uint8_t*)dev_in(ctx, S_A2, features, (size_t)N * d, device_ptrs, &rc);
    hipError_t e = hipSuccess;
    if (N && !src) e = hipErrorOutOfMemory;
    // rows into memory order with the records' permutation
    if (e == hipSuccess && N) e = pk_launch_gather_rows(src, d, 0, r8, d, d, ix->mem2local, N, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (!device_ptrs && S_A2 < ctx->bufs.size() && ctx->bufs[S_A2].p) {
        (void)hipFree(ctx->bufs[S_A2].p);
        ctx->bufs[S_A2] = DevBuf();
    }
    if (e != hipSuccess) {
        (void)hipFree(r8);
        return rc ? rc : set_err(ctx, PUCKGPU_EHIP, "attaching uint8 raw vectors failed: %s", hipGetErrorString(e));
    }
    ix->raw8 = r8;
    return PUCKGPU_OK;
}

int puckgpu_index_attach_pq(puckgpu_ctx* ctx, puckgpu_index* ix, uint32_t pq_nsq, const float* pq_codebook,
                            const uint8_t* pq_records, int device_ptrs) {
    if (!ctx || !ix || !pq_codebook || (ix->p.total_point_count && !pq_records))
        return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    const uint32_t d = ix->p.dim, N = ix->p.total_point_count;
    if (pq_nsq == 0 || d % pq_nsq || !(d / pq_nsq == 1 || d / pq_nsq == 2 || d / pq_nsq == 4))
        return set_err(ctx, PUCKGPU_EINVAL, "pq_nsq %u must divide dim %u into sub-vectors of 1, 2 or 4", pq_nsq, d);
    if (pq_nsq > 128) return set_err(ctx, PUCKGPU_EINVAL, "pq_nsq must be <= 128 (its lookup table lives in LDS)");
    if (ix->pq_codes) return set_err(ctx, PUCKGPU_EINVAL, "the index already has a second-stage PQ");
    if (ix->raw8) return set_err(ctx, PUCKGPU_EINVAL, "the index ranks by its uint8 raw vectors");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    hipStream_t s = ctx->stream;
    const uint32_t plsq = d / pq_nsq;
    const size_t rec = 4 + pq_nsq;
    int rc = 0;
    float* cb = nullptr;
    float* off = nullptr;
    uint8_t* codes = nullptr;
    if ((rc = dalloc(ctx, &cb, (size_t)pq_nsq * PK_KS * plsq)) || (rc = dalloc(ctx, &off, N)) ||
        (rc = dalloc(ctx, &codes, (size_t)N * pq_nsq))) {
        for (void* p : {(void*)cb, (void*)off, (void*)codes})
            if (p) (void)hipFree(p);
        return rc;
    }
    const uint8_t* rec_dev = (const uint8_t*)dev_in(ctx, S_A1, pq_records, rec * N, device_ptrs, &rc);
    hipError_t e = hipSuccess;
    if (N && !rec_dev) e = hipErrorOutOfMemory;
    if (e == hipSuccess)
        e = hipMemcpyAsync(cb, pq_codebook, sizeof(float) * pq_nsq * PK_KS * plsq,
                           device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s);
    // records into memory order with the filter records' permutation
    if (e == hipSuccess && N) e = pk_launch_gather_rows(rec_dev, rec, 0, (uint8_t*)off, 4, 4, ix->mem2local, N, s);
    if (e == hipSuccess && N) e = pk_launch_gather_rows(rec_dev, rec, 4, codes, pq_nsq, pq_nsq, ix->mem2local, N, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (!device_ptrs && S_A1 < ctx->bufs.size() && ctx->bufs[S_A1].p) {
        (void)hipFree(ctx->bufs[S_A1].p);
        ctx->bufs[S_A1] = DevBuf();
    }
    if (e != hipSuccess) {
        for (void* p : {(void*)cb, (void*)off, (void*)codes})
            if (p) (void)hipFree(p);
        return rc ? rc : set_err(ctx, PUCKGPU_EHIP, "attaching the second-stage PQ failed: %s", hipGetErrorString(e));
    }
    ix->pnsq = pq_nsq;
    ix->plsq = plsq;
    ix->pq_codebook = cb;
    ix->pq_off = off;
    ix->pq_codes = codes;
    return PUCKGPU_OK;
}

int puckgpu_index_export(puckgpu_ctx* ctx, const puckgpu_index* ix, float* coarse_norms, float* stationary,
                         uint32_t* cell_start, float* min_dist_offset, uint32_t* mem2local) {
    if (!ctx || !ix) return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    const uint64_t cells = (uint64_t)ix->p.coarse_cluster_count * ix->p.fine_cluster_count;
    if (coarse_norms)
        PK_CHECK(hipMemcpy(coarse_norms, ix->coarse_norms, 4 * ix->p.coarse_cluster_count, hipMemcpyDeviceToHost));
    if (stationary) PK_CHECK(hipMemcpy(stationary, ix->stationary, 4 * cells, hipMemcpyDeviceToHost));
    if (cell_start) PK_CHECK(hipMemcpy(cell_start, ix->cell_start, 4 * (cells + 1), hipMemcpyDeviceToHost));
    if (min_dist_offset)
        PK_CHECK(hipMemcpy(min_dist_offset, ix->mdo, 4 * ix->p.coarse_cluster_count, hipMemcpyDeviceToHost));
    if (mem2local && ix->p.total_point_count)
        PK_CHECK(hipMemcpy(mem2local, ix->mem2local, 4ull * ix->p.total_point_count, hipMemcpyDeviceToHost));
    return PUCKGPU_OK;
}

// ---------------------------------------------------------------------------
namespace {

// Pinned host memory of at least `bytes` (grown on demand, kept by the context).
uint8_t* host_pinned(puckgpu_ctx* ctx, size_t bytes) {
    if (bytes <= ctx->hpin_cap) return ctx->hpin;
    if (ctx->hpin) (void)hipHostFree(ctx->hpin);
    ctx->hpin = nullptr;
    ctx->hpin_cap = 0;
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
    ctx->hpin = (uint8_t*)p;
    ctx->hpin_cap = bytes;
    return ctx->hpin;
}

// Batched search.  `topk` is the largest per-query top-k (the rerank
// capacity) and `ostride` the output row stride; q_topk / q_scc / q_nf are
// optional per-query arrays (host or device as device_ptrs says).
// The `keep` smallest coarse distances of m rows, ascending by (distance, id):
// a select-based top-k for short prefixes, a full row sort otherwise.
hipError_t coarse_select(const float* cdist, uint32_t m, uint32_t C, uint32_t keep, float* csd, uint32_t* csi,
                         uint8_t* flags, uint32_t flag_lim, hipStream_t s) {
    if (keep == 0) return hipSuccess;
    if (keep <= 256) return pk_launch_topk_rows(cdist, m, C, C, keep, csd, csi, flags, flag_lim, s);
    return pk_launch_sort_rows(cdist, m, C, C, keep, csd, csi, flags, flag_lim, s);
}

// Probe features of m prepared queries: cell-distance row of the nearest
// coarse list, its nfine smallest values, then the raw feature rows.
// cv: m x nfine scratch; cid (may be NULL): their fine ids.
int probe_features_chunk(puckgpu_ctx* ctx, const puckgpu_index* ix, const puckgpu_probe* p, uint32_t m,
                         const float* q, const float* csd, const uint32_t* csi, uint32_t cstride, const float* fdist,
                         float* cv, uint32_t* cid, float* feat, hipStream_t s) {
    const uint32_t F = ix->p.fine_cluster_count;
    if (p->nfine) PK_CHECK(pk_launch_cell_topk(csd, csi, cstride, fdist, ix->stationary, F, p->nfine, m, cv, cid, s));
    PK_CHECK(pk_launch_probe_raw_features(q, ix->p.dim, csd, cstride, p->nc, cv, p->nfine, m, feat, s));
    return PUCKGPU_OK;
}

struct ProbeUse {
    const puckgpu_probe* probe = nullptr;
    puckgpu_budget_params bp{};
    uint32_t* budget_out = nullptr;
};

// Device buffers of one pipeline pass over up to `cap` queries.
struct PassBufs {
    uint32_t* elist = nullptr;   // queries with a coarse or fine tie (device list) and its count
    uint32_t* ecount = nullptr;
    float* q = nullptr;
    float* qn = nullptr;
    int32_t* qst = nullptr;
    float* cdist = nullptr;
    float* fdist = nullptr;
    float* csd = nullptr;
    uint32_t* csi = nullptr;
    float* fsd = nullptr;
    uint32_t* fsi = nullptr;
    uint8_t* flags = nullptr;
    uint32_t* tie_ids = nullptr;
    uint8_t* tie_n = nullptr;
    uint32_t* stats = nullptr;
    uint32_t* qlist = nullptr;
    float* pfeat = nullptr;
    void* pwork = nullptr;
    uint32_t* pnf = nullptr;
    float* py = nullptr;
    float* pcv = nullptr;  // probe cell features, cap x nfine
    uint32_t* pcid = nullptr;  // their fine ids (the budget scan's cut of the nearest list)
    uint16_t* q16 = nullptr;   // reduced-precision queries (precision 1 / 2)
    uint8_t* q8 = nullptr;
    float* qs = nullptr;
    float* cand_d = nullptr;   // approximate coarse candidates before rescoring
    uint32_t* cand_i = nullptr;
    uint32_t* wlist = nullptr;   // queries taking the walk when budgets are given
    uint32_t* qorder = nullptr;  // workgroup -> query order of the per-query kernels
    uint32_t* wcount = nullptr;
    // split budget scan: per-query selected cells, per-query counts (points,
    // cells), the list of queries left to the per-query kernel and its length
    uint4* sc_cells = nullptr;
    uint32_t* sc_meta = nullptr;
    uint32_t* sc_olist = nullptr;
    uint32_t* sc_ocount = nullptr;
};

// Candidate points and cells per query of the split budget scan; a query
// with more is scanned by the per-query kernel.
constexpr uint32_t PK_SC_CAPP = 8192;  // points per query in the split scan (keys in registers)
constexpr uint32_t PK_SC_CELLS = 512;  // cells per query in the split scan (its LDS list)

// Per-pass configuration.
struct PassCfg {
    uint32_t scc = 0, keep = 0, ftk = 0, topk = 0, ostride = 0;
    int precision = 0;
    int lut16 = 0;    // FP16 filter table
    uint32_t kc = 0;  // coarse candidates rescored in FP32 (precision 1 / 2)
    double rr = 1.0;
    int gemm_shape = 0;
    uint32_t rerank_rows = 2;  // pk_rerank rows per lane group (autotuned per server bucket)
    const ProbeUse* pu = nullptr;
    bool budget = false;  // fine budgets may be given (explicit or predicted): the split scan's buffers
    // plain exact search on context buffers: the exact center orders of the
    // queries with a coarse or fine tie are computed on the second stream
    // beside the rerank (the replay's first stage then skips them)
    bool early_order = false;
};

// Enqueues the device stages of one pass over m queries: no host
// synchronisation and no memset, so the sequence can be captured in a graph.
// ev (5 events, may be null) brackets the stages.  nvalid (device, may be
// null): rows at or past it are padding.  wa / ra receive the launch arguments
// (for a replay).
int enqueue_pass(puckgpu_ctx* ctx, const puckgpu_index* ix, const PassCfg& cfg, const PassBufs& b, uint32_t m,
                 const float* qin, const uint32_t* nvalid, const uint32_t* q_topk, const uint32_t* q_scc,
                 const uint32_t* q_nf, float* ft, uint32_t* fi, uint32_t* fn, float* od, uint32_t* oi, uint32_t* on,
                 uint32_t* budget_out, int budget_out_host, hipEvent_t* ev, PkWalkArgs* wa_out, PkRerankArgs* ra_out,
                 hipStream_t s) {
    const uint32_t d = ix->p.dim, C = ix->p.coarse_cluster_count, F = ix->p.fine_cluster_count;
    const puckgpu_probe* pr = cfg.pu ? cfg.pu->probe : nullptr;
    int rc = 0;
    if (ev) PK_CHECK(hipEventRecord(ev[0], s));
    PK_CHECK(pk_launch_prep_queries_ex(qin, m, d, ix->p.whether_norm, b.q, b.qn, b.qst, nvalid, b.flags, s));
    if (cfg.precision == 0) {
        PK_CHECK(pk_launch_gemm_exact_shape(b.q, ix->coarse, b.cdist, m, C, d, C, PK_EPI_SUB_FROM_BIAS,
                                            ix->coarse_norms, cfg.gemm_shape, s));
        PK_CHECK(pk_launch_gemm_exact_shape(b.q, ix->fine, b.fdist, m, F, d, F, PK_EPI_NEG, nullptr, cfg.gemm_shape,
                                            s));
    } else {
        const bool bf = cfg.precision == 1;
        PK_CHECK(pkc_launch_convert(b.q, m, d, cfg.precision, b.q16, b.q8, b.qs, s));
        const void* qa = bf ? (const void*)b.q16 : (const void*)b.q8;
        PK_CHECK(pkc_launch_center(qa, b.qs, bf ? (const void*)ix->coarse16 : (const void*)ix->coarse8, ix->coarse8s,
                                   m, C, d, cfg.precision, PK_EPI_SUB_FROM_BIAS, ix->coarse_norms, b.cdist, C, s));
        PK_CHECK(pkc_launch_center(qa, b.qs, bf ? (const void*)ix->fine16 : (const void*)ix->fine8, ix->fine8s, m, F,
                                   d, cfg.precision, PK_EPI_NEG, nullptr, b.fdist, F, s));
    }
    if (ev) PK_CHECK(hipEventRecord(ev[1], s));
    if (cfg.precision == 0) {
        PK_CHECK(coarse_select(b.cdist, m, C, cfg.keep, b.csd, b.csi, b.flags, cfg.scc, s));
    } else if (cfg.keep) {
        // approximate candidates, then the exact FP32 order decides the kept lists
        PK_CHECK(coarse_select(b.cdist, m, C, cfg.kc, b.cand_d, b.cand_i, nullptr, 0, s));
        PK_CHECK(pkc_launch_rescore_coarse(b.q, d, ix->coarse, ix->coarse_norms, b.cand_i, cfg.kc, cfg.keep, m, b.csd,
                                           b.csi, b.flags, cfg.scc, s));
    }
    // large batches: the per-query kernels take the queries grouped by nearest
    // coarse list, dealt to the XCDs in slices (shared records stay in one L2)
    // (on the second stream, beside the probe or the fine sort; joined before the walk)
    const uint32_t* order = nullptr;
    if (m >= 2048 && cfg.keep) {
        PK_CHECK(hipEventRecord(ctx->fork_ev, s));
        PK_CHECK(hipStreamWaitEvent(ctx->stream2, ctx->fork_ev, 0));
        PK_CHECK(pk_launch_build_query_order(b.csi, cfg.keep, m, C, b.qorder, ctx->stream2));
        PK_CHECK(hipEventRecord(ctx->order_ev, ctx->stream2));
        order = b.qorder;
    }
    const uint32_t* nf = q_nf;
    if (pr) {
        // learned fine budget: features -> MLP -> budget, on this stream
        if ((rc = probe_features_chunk(ctx, ix, pr, m, b.q, b.csd, b.csi, cfg.keep, b.fdist, b.pcv, b.pcid,
                                       b.pfeat, s)))
            return rc;
        PK_CHECK(pk_probe_pass(pr->params, pr->wbf, pr->D, pr->H, m, b.pfeat, nullptr, b.pwork, nullptr, nullptr,
                               b.py, s));
        PK_CHECK(pk_launch_probe_budget(b.py, m, cfg.pu->bp.margin, cfg.pu->bp.min_budget, cfg.pu->bp.fallback_frac,
                                        F, b.pnf, s));
        if (budget_out)
            PK_CHECK(hipMemcpyAsync(budget_out, b.pnf, sizeof(uint32_t) * m,
                                    budget_out_host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, s));
        nf = b.pnf;
    }
    // fine order for the walk: every row, or with budgets only the rows of the
    // queries that take the walk
    if (nf) {
        PK_CHECK(pk_launch_build_walk_list(nf, m, F, b.wlist, b.wcount, s));
        // on the second stream, ahead of the walk that reads them (the walk's
        // fork below orders it after this sort and joins it back)
        PK_CHECK(hipEventRecord(ctx->fork_ev, s));
        PK_CHECK(hipStreamWaitEvent(ctx->stream2, ctx->fork_ev, 0));
        PK_CHECK(pk_launch_sort_rows_list(b.fdist, m, F, F, F, b.fsd, b.fsi, b.wlist, b.wcount, ctx->stream2));
    } else {
        PK_CHECK(pk_launch_sort_rows(b.fdist, m, F, F, F, b.fsd, b.fsi, nullptr, 0, s));
    }
    if (order) PK_CHECK(hipStreamWaitEvent(s, ctx->order_ev, 0));
    if (ev) PK_CHECK(hipEventRecord(ev[2], s));
    PkWalkArgs wa{};
    wa.q = b.q; wa.qn = b.qn; wa.qstatus = b.qst; wa.csel_d = b.csd; wa.csel_id = b.csi; wa.fsort_d = b.fsd;
    wa.fsort_id = b.fsi; wa.codebook = ix->codebook; wa.stationary = ix->stationary;
    wa.cell_start = ix->cell_start; wa.mdo = ix->mdo; wa.rec_off = ix->rec_off; wa.rec_codes = ix->rec_codes;
    wa.dim = d; wa.C = C; wa.F = F; wa.nsq = ix->p.filter_nsq; wa.lsq = ix->lsq; wa.scc = cfg.scc; wa.ftk = cfg.ftk;
    wa.rr = cfg.rr; wa.out_t = ft; wa.out_id = fi; wa.out_n = fn; wa.stats = b.stats; wa.qlist = order;
    wa.flags = b.flags; wa.exact = 0; wa.tie_ids = b.tie_ids; wa.tie_n = b.tie_n;
    wa.q_scc = q_scc; wa.q_nf = nf; wa.cstride = cfg.keep ? cfg.keep : 1; wa.fdist = b.fdist;
    wa.wlist = nullptr; wa.wcount = nullptr;
    wa.sc_cells = b.sc_cells; wa.sc_meta = b.sc_meta; wa.sc_olist = b.sc_olist;
    wa.sc_ocount = b.sc_ocount; wa.sc_capp = PK_SC_CAPP;
    wa.lut16 = cfg.lut16; wa.cb_nmax = ix->cb_nmax;
    if (pr && pr->nfine) {
        wa.top_d = b.pcv;
        wa.top_f = b.pcid;
        wa.top_k = pr->nfine;
    }
    if (nf) {
        // the walk runs only over the queries without a budget below F (the list built above)
        wa.wlist = b.wlist;
        wa.wcount = b.wcount;
    }
    PK_CHECK(pk_launch_walk(&wa, m, s, ev ? ev[6] : nullptr, ctx->stream2, ctx->fork_ev, ctx->join_ev));
    if (cfg.early_order) {
        // flags 1 and 8 are final after the walk; the rerank below reads none
        // of the center orders these kernels rewrite
        PK_CHECK(hipEventRecord(ctx->fork_ev, s));
        PK_CHECK(hipStreamWaitEvent(ctx->stream2, ctx->fork_ev, 0));
        PK_CHECK(pk_launch_exact_center_order_dev(b.cdist, C, b.fdist, F, cfg.keep ? cfg.keep : 1, cfg.scc, q_scc,
                                                  b.flags, 9u, m, b.elist, b.ecount, b.csd, b.csi, b.fsd, b.fsi,
                                                  ctx->stream2));
        PK_CHECK(hipEventRecord(ctx->eorder_ev, ctx->stream2));
    }
    wa.wlist = nullptr;  // replays address queries through qlist
    wa.wcount = nullptr;
    if (ev) PK_CHECK(hipEventRecord(ev[3], s));
    PkRerankArgs ra{};
    ra.rows = cfg.rerank_rows;
    ra.q = b.q; ra.qstatus = b.qst; ra.filt_t = ft; ra.filt_id = fi; ra.filt_n = fn;
    ra.raw = ix->raw; ra.raw8 = ix->raw8; ra.mem2local = ix->mem2local; ra.dim = d; ra.ftk = cfg.ftk; ra.topk = cfg.topk;
    ra.out_d = od; ra.out_id = oi; ra.out_n = on;
    ra.qlist = order; ra.flags = b.flags; ra.exact = 0; ra.tie_ids = b.tie_ids; ra.tie_n = b.tie_n;
    ra.q_topk = q_topk; ra.out_stride = cfg.ostride; ra.q_nf = nf; ra.F = F;
    ra.qn = b.qn; ra.pq_off = ix->pq_off; ra.pq_codes = ix->pq_codes; ra.pq_codebook = ix->pq_codebook;
    ra.pnsq = ix->pnsq; ra.plsq = ix->plsq;
    PK_CHECK(pk_launch_rerank(&ra, m, s));
    if (ev) PK_CHECK(hipEventRecord(ev[4], s));
    if (cfg.early_order) PK_CHECK(hipStreamWaitEvent(s, ctx->eorder_ev, 0));
    if (wa_out) *wa_out = wa;
    if (ra_out) *ra_out = ra;
    return PUCKGPU_OK;
}

// Calls f(i) for each i < m with flags[i] != 0, testing eight flags per
// 64-bit word (flags are almost all zero: the scan after every pass is on
// the host, between the GPU's batches).
extern "C++" template <typename Fn>
void for_each_flagged(const uint8_t* flags, uint32_t m, Fn f) {
    uint32_t i = 0;
    for (; i + 8 <= m; i += 8) {
        uint64_t w;
        std::memcpy(&w, flags + i, 8);
        if (!w) continue;
        for (uint32_t j = 0; j < 8; ++j)
            if (flags[i + j]) f(i + j);
    }
    for (; i < m; ++i)
        if (flags[i]) f(i);
}

// After the pass has completed (host flags copied in): queries whose result
// depends on the order among exactly equal distances are replayed with the
// reference heaps.  flags: 1 coarse-distance tie in the kept range, 2
// filter-boundary tie at the final top, 4 exact-distance tie in the top-k, 8 a
// decision on a fine-distance tie group.
//   stage 1: queries flagged only 1 and/or 8 get the reference's coarse order
//            (1) and fine order (8) and run the fast walk and rerank again;
//            the walk visits the cells of a fine row given in the reference's
//            order in that order, so fine ties need no flag there, and any
//            other remaining tie is flagged anew;
//   stage 2: queries still flagged run the walk and rerank with the reference
//            heaps (the reference's fine order first when flagged 8; orders
//            computed in stage 1 are already in place).
// Synchronises the stream.  Returns the number of queries replayed, or < 0.
// hlist, stage1: host arrays of m entries each (the replay lists).
int replay_pass(puckgpu_ctx* ctx, const puckgpu_index* ix, const PassCfg& cfg, const PassBufs& b, uint32_t m,
                uint8_t* hflags, int replay_all, PkWalkArgs wa, PkRerankArgs ra, uint32_t* hlist, uint32_t* stage1,
                hipStream_t s) {
    const uint32_t C = ix->p.coarse_cluster_count, F = ix->p.fine_cluster_count;
    uint32_t n1 = 0, nl = 0;
    if (replay_all) {
        for (uint32_t i = 0; i < m; ++i) hlist[nl++] = i;
    } else {
        for_each_flagged(hflags, m, [&](uint32_t i) {
            if ((hflags[i] & 15u) != 0u && (hflags[i] & 6u) == 0u) stage1[n1++] = i;
            else if (hflags[i] & 15u) hlist[nl++] = i;
        });
    }
    if (n1) {
        PK_CHECK(hipMemcpyAsync(b.qlist, stage1, sizeof(uint32_t) * n1, hipMemcpyHostToDevice, s));
        // every stage-1 query has flag 1 or 8, so with early_order its exact
        // center order was computed during the pass
        if (!cfg.early_order)
            PK_CHECK(pk_launch_exact_center_order(b.cdist, C, b.fdist, F, wa.cstride, cfg.scc, wa.q_scc, b.qlist, n1,
                                                  b.flags, 0, b.csd, b.csi, b.fsd, b.fsi, s));
        PkWalkArgs w1 = wa;
        w1.qlist = b.qlist;
        w1.fine_given = 1;
        PK_CHECK(pk_launch_walk(&w1, n1, s, nullptr, nullptr, nullptr, nullptr));
        PkRerankArgs r1 = ra;
        r1.qlist = b.qlist;
        PK_CHECK(pk_launch_rerank(&r1, n1, s));
        PK_CHECK(hipMemcpyAsync(hflags, b.flags, m, hipMemcpyDeviceToHost, s));
        PK_CHECK(hipStreamSynchronize(s));
        for (uint32_t j = 0; j < n1; ++j)
            if (hflags[stage1[j]] & 14u) hlist[nl++] = stage1[j];
    }
    if (nl) {
        PK_CHECK(hipMemcpyAsync(b.qlist, hlist, sizeof(uint32_t) * nl, hipMemcpyHostToDevice, s));
        PK_CHECK(pk_launch_exact_center_order(b.cdist, C, b.fdist, F, wa.cstride, cfg.scc, wa.q_scc, b.qlist, nl,
                                              b.flags, replay_all, b.csd, b.csi, b.fsd, b.fsi, s));
        wa.qlist = b.qlist; wa.flags = nullptr; wa.exact = 1; wa.tie_ids = nullptr; wa.tie_n = nullptr;
        PK_CHECK(pk_launch_walk(&wa, nl, s, nullptr, nullptr, nullptr, nullptr));
        ra.qlist = b.qlist; ra.flags = nullptr; ra.exact = 1; ra.tie_ids = nullptr; ra.tie_n = nullptr;
        PK_CHECK(pk_launch_rerank(&ra, nl, s));
    }
    return (int)(n1 + nl);
}

// Allocates pass buffers for `cap` queries from context scratch slots.
int scratch_pass(puckgpu_ctx* ctx, const puckgpu_index* ix, const PassCfg& cfg, uint32_t cap, PassBufs& b) {
    const uint32_t d = ix->p.dim, C = ix->p.coarse_cluster_count, F = ix->p.fine_cluster_count;
    const puckgpu_probe* pr = cfg.pu ? cfg.pu->probe : nullptr;
    const uint32_t keep = cfg.keep ? cfg.keep : 1;
    int rc = 0;
    b.q = (float*)scratch_slot(ctx, S_Q, sizeof(float) * cap * d, &rc);
    b.qn = (float*)scratch_slot(ctx, S_QN, sizeof(float) * cap, &rc);
    b.qst = (int32_t*)scratch_slot(ctx, S_QST, sizeof(int32_t) * cap, &rc);
    b.cdist = (float*)scratch_slot(ctx, S_CD, sizeof(float) * cap * C, &rc);
    b.fdist = (float*)scratch_slot(ctx, S_FD, sizeof(float) * cap * F, &rc);
    b.csd = (float*)scratch_slot(ctx, S_CSD, sizeof(float) * cap * keep, &rc);
    b.csi = (uint32_t*)scratch_slot(ctx, S_CSI, sizeof(uint32_t) * cap * keep, &rc);
    b.fsd = (float*)scratch_slot(ctx, S_FSD, sizeof(float) * cap * F, &rc);
    b.fsi = (uint32_t*)scratch_slot(ctx, S_FSI, sizeof(uint32_t) * cap * F, &rc);
    b.flags = (uint8_t*)scratch_slot(ctx, S_FLAGS, cap, &rc);
    b.qlist = (uint32_t*)scratch_slot(ctx, S_QLIST, sizeof(uint32_t) * cap, &rc);
    b.elist = (uint32_t*)scratch_slot(ctx, S_ELIST, sizeof(uint32_t) * (cap + 1), &rc);
    b.ecount = b.elist ? b.elist + cap : nullptr;
    b.tie_ids = (uint32_t*)scratch_slot(ctx, S_TIEID, sizeof(uint32_t) * 16 * cap, &rc);
    b.tie_n = (uint8_t*)scratch_slot(ctx, S_TIEN, cap, &rc);
    b.stats = (uint32_t*)scratch_slot(ctx, S_STATS, sizeof(uint32_t) * 4 * cap, &rc);
    b.wlist = (uint32_t*)scratch_slot(ctx, S_WLIST, sizeof(uint32_t) * (2 * (size_t)cap + 1), &rc);
    b.qorder = b.wlist ? b.wlist + cap : nullptr;
    b.wcount = b.wlist ? b.wlist + 2 * (size_t)cap : nullptr;
    if (cfg.budget) {
        b.sc_cells = (uint4*)scratch_slot(ctx, S_SCI, sizeof(uint4) * (size_t)cap * PK_SC_CELLS, &rc);
        b.sc_meta = (uint32_t*)scratch_slot(ctx, S_SCM, sizeof(uint32_t) * (3 * (size_t)cap + 1), &rc);
        b.sc_olist = b.sc_meta ? b.sc_meta + 2 * (size_t)cap : nullptr;
        b.sc_ocount = b.sc_meta ? b.sc_meta + 3 * (size_t)cap : nullptr;
    }
    if (pr) {
        b.pfeat = (float*)scratch_slot(ctx, S_PFEAT, sizeof(float) * cap * pr->D, &rc);
        b.pwork = scratch_slot(ctx, S_PWORK, pk_probe_work_bytes(cap, pr->D, pr->H), &rc);
        b.pnf = (uint32_t*)scratch_slot(ctx, S_PNF, sizeof(uint32_t) * cap, &rc);
        b.py = (float*)scratch_slot(ctx, S_PY, sizeof(float) * cap, &rc);
        b.pcv = (float*)scratch_slot(ctx, S_PCV, sizeof(float) * cap * (pr->nfine ? pr->nfine : 1), &rc);
        b.pcid = (uint32_t*)scratch_slot(ctx, S_PCID, sizeof(uint32_t) * cap * (pr->nfine ? pr->nfine : 1), &rc);
    }
    if (cfg.precision) {
        const uint32_t ld = pkc_ld(d), kc = cfg.kc ? cfg.kc : 1;
        b.q16 = (uint16_t*)scratch_slot(ctx, S_LP16, sizeof(uint16_t) * cap * ld, &rc);
        b.q8 = (uint8_t*)scratch_slot(ctx, S_LP8, (size_t)cap * ld, &rc);
        b.qs = (float*)scratch_slot(ctx, S_LPS, sizeof(float) * cap, &rc);
        b.cand_d = (float*)scratch_slot(ctx, S_CAND_D, sizeof(float) * cap * kc, &rc);
        b.cand_i = (uint32_t*)scratch_slot(ctx, S_CAND_I, sizeof(uint32_t) * cap * kc, &rc);
    }
    return rc;
}

// Reduced-precision copies of the index centers (once per precision).
int ensure_lowp(puckgpu_ctx* ctx, const puckgpu_index* ix, int prec) {
    const uint32_t d = ix->p.dim, C = ix->p.coarse_cluster_count, F = ix->p.fine_cluster_count, ld = pkc_ld(d);
    hipStream_t s = ctx->stream;
    int rc = 0;
    if (prec == 1 && !ix->coarse16) {
        if ((rc = dalloc(ctx, &ix->coarse16, (size_t)C * ld)) || (rc = dalloc(ctx, &ix->fine16, (size_t)F * ld)))
            return rc;
        PK_CHECK(pkc_launch_convert(ix->coarse, C, d, 1, ix->coarse16, nullptr, nullptr, s));
        PK_CHECK(pkc_launch_convert(ix->fine, F, d, 1, ix->fine16, nullptr, nullptr, s));
        PK_CHECK(hipStreamSynchronize(s));
    }
    if (prec == 2 && !ix->coarse8) {
        if ((rc = dalloc(ctx, &ix->coarse8, (size_t)C * ld)) || (rc = dalloc(ctx, &ix->fine8, (size_t)F * ld)) ||
            (rc = dalloc(ctx, &ix->coarse8s, C)) || (rc = dalloc(ctx, &ix->fine8s, F)))
            return rc;
        PK_CHECK(pkc_launch_convert(ix->coarse, C, d, 2, nullptr, ix->coarse8, ix->coarse8s, s));
        PK_CHECK(pkc_launch_convert(ix->fine, F, d, 2, nullptr, ix->fine8, ix->fine8s, s));
        PK_CHECK(hipStreamSynchronize(s));
    }
    return PUCKGPU_OK;
}

// Validates the search knobs and derives the pass configuration.
int make_cfg(puckgpu_ctx* ctx, const puckgpu_index* ix, const puckgpu_search_params* sp, uint32_t topk,
             uint32_t ostride, const ProbeUse* pu, PassCfg& cfg) {
    const uint32_t d = ix->p.dim, C = ix->p.coarse_cluster_count, F = ix->p.fine_cluster_count;
    const uint32_t ftk = sp->filter_topk;
    if (ftk < topk || ftk == 0 || ftk > pk_walk_ftk_max())
        return set_err(ctx, PUCKGPU_EINVAL, "filter_topk %u must be in [topk=%u, %u]", ftk, topk, pk_walk_ftk_max());
    if (!(sp->radius_rate >= 1.0)) return set_err(ctx, PUCKGPU_EINVAL, "radius_rate should >= 1.0");
    const puckgpu_probe* pr = pu ? pu->probe : nullptr;
    if (pr && (pr->dim != d || pr->nc > C || pr->nfine > F))
        return set_err(ctx, PUCKGPU_EINVAL, "probe shape (dim %u, %u coarse, %u fine features) does not fit the index",
                       pr->dim, pr->nc, pr->nfine);
    cfg.scc = sp->search_coarse_count < C ? sp->search_coarse_count : C;
    cfg.rerank_rows = ctx->rerank_rows;
    // sorted coarse entries kept per query: the walk's scc, or the probe's features
    cfg.keep = pr && pr->nc > cfg.scc ? pr->nc : cfg.scc;
    if (pr && cfg.keep == 0) cfg.keep = 1;  // the probe's cell features need the nearest coarse list
    cfg.ftk = ftk;
    cfg.topk = topk;
    cfg.ostride = ostride;
    cfg.rr = sp->radius_rate;
    cfg.pu = pu;
    if (!ix->raw && !ix->raw8 && !ix->pq_codes && ix->p.total_point_count)
        return set_err(ctx, PUCKGPU_EINVAL, "the index holds neither raw vectors nor a second-stage PQ to rank by");
    if ((sp->precision & ~(15u | PUCKGPU_LUT_FP16)) || (sp->precision & 15u) > 2)
        return set_err(ctx, PUCKGPU_EINVAL, "precision must be 0, 1 or 2, optionally with PUCKGPU_LUT_FP16");
    cfg.precision = (int)(sp->precision & 15u);
    cfg.lut16 = (sp->precision & PUCKGPU_LUT_FP16) ? 1 : 0;
    if (cfg.precision) {
        if (pkc_ld(d) > 256) return set_err(ctx, PUCKGPU_EINVAL, "reduced precision needs dim <= 256");
        // margin of 32 coarse candidates (at most 512 and C) rescored in FP32
        uint32_t kc = cfg.keep + 32;
        if (kc > C) kc = C;
        if (kc > 512) kc = 512;
        if (cfg.keep > kc) return set_err(ctx, PUCKGPU_EINVAL, "reduced precision supports search_coarse_count <= 480");
        cfg.kc = kc;
        int rc = ensure_lowp(ctx, ix, cfg.precision);
        if (rc) return rc;
    }
    return PUCKGPU_OK;
}

// One device pass over m queries and its exact-tie replay, on the pass
// buffers b: enqueue_pass, the flags and counters copied to the host
// (hstats: 4 m words, hflags: m bytes), replay_pass (hl, h1: m words each),
// stage times and counters accumulated on the context.
int run_pass(puckgpu_ctx* ctx, const puckgpu_index* ix, const PassCfg& cfg, const puckgpu_search_params* sp,
             const PassBufs& b, uint32_t m, const float* qin, const uint32_t* q_topk, const uint32_t* q_scc,
             const uint32_t* q_nf, float* ft, uint32_t* fi, uint32_t* fn, float* od, uint32_t* oi, uint32_t* on,
             uint32_t* budget_out, int budget_out_host, uint32_t* hstats, uint8_t* hflags, uint32_t* hl, uint32_t* h1,
             hipStream_t s) {
    PkWalkArgs wa;
    PkRerankArgs ra;
    int rc = enqueue_pass(ctx, ix, cfg, b, m, qin, nullptr, q_topk, q_scc, q_nf, ft, fi, fn, od, oi, on, budget_out,
                          budget_out_host, ctx->ev, &wa, &ra, s);
    if (rc) return rc;
    PK_CHECK(hipMemcpyAsync(hstats, b.stats, sizeof(uint32_t) * 4 * m, hipMemcpyDeviceToHost, s));
    PK_CHECK(hipMemcpyAsync(hflags, b.flags, m, hipMemcpyDeviceToHost, s));
    PK_CHECK(hipStreamSynchronize(s));
    if (cfg.precision) std::fill(hflags, hflags + m, (uint8_t)0);  // no replay in these forms
    for_each_flagged(hflags, m, [&](uint32_t i) {  // first-pass replay causes
        for (int bb = 0; bb < 8; ++bb)
            if (hflags[i] & (1u << bb)) ctx->counters[5 + bb]++;
    });
    const int nrep = replay_pass(ctx, ix, cfg, b, m, hflags, cfg.precision ? 0 : sp->replay_all, wa, ra, hl, h1, s);
    if (nrep < 0) return nrep;
    // stages 0-3 are complete (the flags copy was waited for); the replay
    // stage is timed only when something was replayed
    for (int k = 0; k < 4; ++k) {
        float ms = 0;
        PK_CHECK(hipEventElapsedTime(&ms, ctx->ev[k], ctx->ev[k + 1]));
        ctx->stage_ms[k] += ms;
    }
    if (nrep > 0) {
        PK_CHECK(hipEventRecord(ctx->ev[5], s));
        PK_CHECK(hipEventSynchronize(ctx->ev[5]));
        float ms = 0;
        PK_CHECK(hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]));
        ctx->stage_ms[4] += ms;
    }
    {
        float ms = 0;
        PK_CHECK(hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[6]));
        ctx->scan_ms += ms;
    }
    uint64_t acc[4] = {0, 0, 0, 0};  // local sums: a tight loop over the host copy
    for (uint32_t i = 0; i < m; ++i)
        for (int k = 0; k < 4; ++k) acc[k] += hstats[4 * (size_t)i + k];
    for (int k = 0; k < 4; ++k) ctx->counters[k] += acc[k];
    ctx->counters[4] += (uint64_t)nrep;
    return PUCKGPU_OK;
}

// Batched search.  `topk` is the largest per-query top-k (the rerank
// capacity) and `ostride` the output row stride; q_topk / q_scc / q_nf are
// optional per-query arrays (host or device as device_ptrs says).
int search_impl(puckgpu_ctx* ctx, const puckgpu_index* ix, const puckgpu_search_params* sp,
                const uint32_t* q_topk_in, const uint32_t* q_scc_in, const uint32_t* q_nf_in, uint32_t n,
                const float* queries, uint32_t topk, uint32_t ostride, float* distance, uint32_t* labels,
                uint32_t* result_num, int device_ptrs, const ProbeUse* pu = nullptr) {
    if (n == 0) return PUCKGPU_OK;
    if (!queries || !distance || !labels) return set_err(ctx, PUCKGPU_EINVAL, "null array");
    PassCfg cfg;
    int rc = make_cfg(ctx, ix, sp, topk, ostride, pu, cfg);
    if (rc) return rc;
    cfg.budget = q_nf_in != nullptr || (pu && pu->probe);
    cfg.early_order = !cfg.budget && cfg.precision == 0;
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    const uint32_t d = ix->p.dim, ftk = cfg.ftk;
    hipStream_t s = ctx->stream;
    const uint32_t QB = 16384;  // queries per device pass
    const uint32_t nb = n < QB ? n : QB;
    PassBufs b;
    if ((rc = scratch_pass(ctx, ix, cfg, nb, b))) return rc;
    float* ft = (float*)scratch_slot(ctx, S_FT, sizeof(float) * (size_t)n * ftk, &rc);
    uint32_t* fi = (uint32_t*)scratch_slot(ctx, S_FI, sizeof(uint32_t) * (size_t)n * ftk, &rc);
    uint32_t* fn = (uint32_t*)scratch_slot(ctx, S_FN, sizeof(uint32_t) * n, &rc);
    if (rc) return rc;
    float* od;
    uint32_t *oi, *on;
    if (device_ptrs) {
        od = distance;
        oi = labels;
        on = result_num ? result_num : (uint32_t*)scratch_slot(ctx, S_ON, sizeof(uint32_t) * n, &rc);
    } else {
        od = (float*)scratch_slot(ctx, S_OD, sizeof(float) * (size_t)n * ostride, &rc);
        oi = (uint32_t*)scratch_slot(ctx, S_OI, sizeof(uint32_t) * (size_t)n * ostride, &rc);
        on = (uint32_t*)scratch_slot(ctx, S_ON, sizeof(uint32_t) * n, &rc);
    }
    if (rc) return rc;
    const float* qin = (const float*)dev_in(ctx, S_IN_Q, queries, sizeof(float) * (size_t)n * d, device_ptrs, &rc);
    if (!qin) return rc;
    const uint32_t* q_topk = (const uint32_t*)dev_in(ctx, S_QTOPK, q_topk_in, q_topk_in ? 4ull * n : 0, device_ptrs, &rc);
    const uint32_t* q_scc = (const uint32_t*)dev_in(ctx, S_QSCC, q_scc_in, q_scc_in ? 4ull * n : 0, device_ptrs, &rc);
    const uint32_t* q_nf = (const uint32_t*)dev_in(ctx, S_QNF, q_nf_in, q_nf_in ? 4ull * n : 0, device_ptrs, &rc);
    if (rc) return rc;
    for (int k = 0; k < 5; ++k) ctx->stage_ms[k] = 0;
    ctx->scan_ms = 0;
    for (int k = 0; k < 13; ++k) ctx->counters[k] = 0;
    uint8_t* hpin = host_pinned(ctx, 17 * (size_t)nb);
    if (!hpin) return set_err(ctx, PUCKGPU_ENOMEM, "pinned host allocation failed");
    uint32_t* hstats = (uint32_t*)hpin;
    uint8_t* hflags = hpin + 16 * (size_t)nb;
    std::vector<uint32_t> hl(2 * (size_t)nb);
    for (uint32_t q0 = 0; q0 < n; q0 += QB) {
        const uint32_t m = (n - q0) < QB ? (n - q0) : QB;
        rc = run_pass(ctx, ix, cfg, sp, b, m, qin + (size_t)q0 * d, q_topk ? q_topk + q0 : nullptr,
                      q_scc ? q_scc + q0 : nullptr, q_nf ? q_nf + q0 : nullptr, ft + (size_t)q0 * ftk,
                      fi + (size_t)q0 * ftk, fn + q0, od + (size_t)q0 * ostride, oi + (size_t)q0 * ostride, on + q0,
                      pu && pu->budget_out ? pu->budget_out + q0 : nullptr, !device_ptrs, hstats, hflags, hl.data(),
                      hl.data() + nb, s);
        if (rc) return rc;
    }
    if (!q_nf && !(pu && pu->probe)) ctx->scan_ms = -1.f;  // no budget scan was launched
    if (!device_ptrs) {
        PK_CHECK(hipMemcpyAsync(distance, od, sizeof(float) * (size_t)n * ostride, hipMemcpyDeviceToHost, s));
        PK_CHECK(hipMemcpyAsync(labels, oi, sizeof(uint32_t) * (size_t)n * ostride, hipMemcpyDeviceToHost, s));
        if (result_num) PK_CHECK(hipMemcpyAsync(result_num, on, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
        PK_CHECK(hipStreamSynchronize(s));
    }
    ctx->last_n = n;
    ctx->last_ftk = ftk;
    return PUCKGPU_OK;
}

}  // namespace

int puckgpu_search(puckgpu_ctx* ctx, const puckgpu_index* ix, const puckgpu_search_params* sp, uint32_t n,
                   const float* queries, uint32_t topk, float* distance, uint32_t* labels, uint32_t* result_num,
                   int device_ptrs) {
    if (!ctx || !ix || !sp) return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    if (topk == 0 || topk > sp->topk_max)
        return set_err(ctx, PUCKGPU_EINVAL, "topk should <= topk, topk = %u, requested %u", sp->topk_max, topk);
    return search_impl(ctx, ix, sp, nullptr, nullptr, nullptr, n, queries, topk, topk, distance, labels,
                       result_num, device_ptrs);
}

namespace {

// Pass buffers of m queries (and the filter rows and counts, and result
// counts when out_n is wanted) carved from base in 256-byte aligned pieces.
// Returns the bytes they take; nothing is carved when base is null.
size_t carve_pass(unsigned char* base, const puckgpu_index* ix, const PassCfg& cfg, uint32_t m, PassBufs& b,
                  float** ft, uint32_t** fi, uint32_t** fn, uint32_t** on) {
    const size_t d = ix->p.dim, C = ix->p.coarse_cluster_count, F = ix->p.fine_cluster_count;
    const size_t keep = cfg.keep ? cfg.keep : 1, ftk = cfg.ftk;
    size_t off = 0;
    auto take = [&](auto** p, size_t bytes) {
        if (base) *p = (std::remove_reference_t<decltype(**p)>*)(base + off);
        off += (bytes + 255) & ~(size_t)255;
    };
    take(&b.q, 4 * m * d);
    take(&b.qn, 4 * (size_t)m);
    take(&b.qst, 4 * (size_t)m);
    take(&b.cdist, 4 * m * C);
    take(&b.fdist, 4 * m * F);
    take(&b.csd, 4 * m * keep);
    take(&b.csi, 4 * m * keep);
    take(&b.fsd, 4 * m * F);
    take(&b.fsi, 4 * m * F);
    take(&b.flags, (size_t)m);
    take(&b.qlist, 4 * (size_t)m);
    take(&b.tie_ids, 64 * (size_t)m);
    take(&b.tie_n, (size_t)m);
    take(&b.stats, 16 * (size_t)m);
    take(&b.wlist, 4 * (2 * (size_t)m + 1));
    take(ft, 4 * m * ftk);
    take(fi, 4 * m * ftk);
    take(fn, 4 * (size_t)m);
    take(on, 4 * (size_t)m);
    if (base) {
        b.qorder = b.wlist + m;
        b.wcount = b.wlist + 2 * (size_t)m;
    }
    return off;
}

}  // namespace

int puckgpu_search_noalloc(puckgpu_ctx* ctx, const puckgpu_index* ix, const puckgpu_search_params* sp, uint32_t n,
                           const float* queries, uint32_t topk, float* distance, uint32_t* labels,
                           uint32_t* result_num) {
    if (!ctx || !ix || !sp) return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    if (topk == 0 || topk > sp->topk_max)
        return set_err(ctx, PUCKGPU_EINVAL, "topk should <= topk, topk = %u, requested %u", sp->topk_max, topk);
    if (n == 0) return PUCKGPU_OK;
    if (!queries || !distance || !labels) return set_err(ctx, PUCKGPU_EINVAL, "null array");
    if (n > PK_NOALLOC_MAXN)
        return set_err(ctx, PUCKGPU_EINVAL, "no-alloc search takes at most %u queries per call, got %u",
                       PK_NOALLOC_MAXN, n);
    if (sp->precision != 0) return set_err(ctx, PUCKGPU_EINVAL, "no-alloc search runs the exact FP32 form only");
    if (ctx->device < 0 || ctx->device >= PK_MAX_DEVICES || !ctx->noalloc)
        return set_err(ctx, PUCKGPU_EINVAL, "no scratch array on this device");
    PassCfg cfg;
    int rc = make_cfg(ctx, ix, sp, topk, topk, nullptr, cfg);
    if (rc) return rc;
    PassBufs b;
    float* ft = nullptr;
    uint32_t *fi = nullptr, *fn = nullptr, *on = nullptr;
    const size_t need = carve_pass(nullptr, ix, cfg, n, b, &ft, &fi, &fn, &on);
    if (need > PK_NOALLOC_BYTES)
        return set_err(ctx, PUCKGPU_EINVAL,
                       "no-alloc search of %u queries needs %zu bytes of scratch, the maximum is %llu", n, need,
                       (unsigned long long)PK_NOALLOC_BYTES);
    carve_pass(ctx->noalloc, ix, cfg, n, b, &ft, &fi, &fn, &on);
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    for (int k = 0; k < 5; ++k) ctx->stage_ms[k] = 0;
    ctx->scan_ms = 0;
    for (int k = 0; k < 13; ++k) ctx->counters[k] = 0;
    const int dv = ctx->device;
    rc = run_pass(ctx, ix, cfg, sp, b, n, queries, nullptr, nullptr, nullptr, ft, fi, fn, distance, labels,
                  result_num ? result_num : on, nullptr, 0, pk_noalloc_hstats[dv], pk_noalloc_hflags[dv],
                  pk_noalloc_hlist[dv], pk_noalloc_hlist[dv] + PK_NOALLOC_MAXN, ctx->stream);
    if (rc) return rc;
    PK_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->scan_ms = -1.f;
    ctx->last_n = 0;  // the filter rows of this call live in the shared array, not the context
    return PUCKGPU_OK;
}

int puckgpu_search_batch(puckgpu_ctx* ctx, const puckgpu_index* ix, const puckgpu_search_params* sp,
                         const puckgpu_query_params* qp, uint32_t n, const float* queries, float* distance,
                         uint32_t* labels, uint32_t* result_num, int device_ptrs) {
    if (!ctx || !ix || !sp) return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    if (sp->topk_max == 0) return set_err(ctx, PUCKGPU_EINVAL, "topk_max must be positive");
    const uint32_t* qt = qp ? qp->topk : nullptr;
    const uint32_t* qs = qp ? qp->search_coarse_count : nullptr;
    const uint32_t* qf = qp ? qp->fine_budget : nullptr;
    return search_impl(ctx, ix, sp, qt, qs, qf, n, queries, sp->topk_max, sp->topk_max, distance, labels,
                       result_num, device_ptrs);
}

int puckgpu_search_learned(puckgpu_ctx* ctx, const puckgpu_index* ix, const puckgpu_probe* probe,
                           const puckgpu_search_params* sp, const puckgpu_budget_params* bp,
                           const puckgpu_query_params* qp, uint32_t n, const float* queries, float* distance,
                           uint32_t* labels, uint32_t* result_num, uint32_t* budget_out, int device_ptrs) {
    if (!ctx || !ix || !sp || !probe || !bp) return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    if (sp->topk_max == 0) return set_err(ctx, PUCKGPU_EINVAL, "topk_max must be positive");
    if (!(bp->fallback_frac > 0.f)) return set_err(ctx, PUCKGPU_EINVAL, "fallback_frac must be positive");
    ProbeUse pu;
    pu.probe = probe;
    pu.bp = *bp;
    pu.budget_out = budget_out;
    const uint32_t* qt = qp ? qp->topk : nullptr;
    const uint32_t* qs = qp ? qp->search_coarse_count : nullptr;
    return search_impl(ctx, ix, sp, qt, qs, nullptr, n, queries, sp->topk_max, sp->topk_max, distance, labels,
                       result_num, device_ptrs, &pu);
}

int puckgpu_budget_ranks(puckgpu_ctx* ctx, const puckgpu_index* ix, uint32_t search_coarse_count, uint32_t n,
                         const float* queries, uint32_t k, const uint32_t* cells, uint32_t* ranks, int device_ptrs) {
    if (!ctx || !ix || (n && (!queries || !cells || !ranks))) return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    if (n == 0 || k == 0) return PUCKGPU_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    const uint32_t d = ix->p.dim, C = ix->p.coarse_cluster_count, F = ix->p.fine_cluster_count;
    const uint32_t scc = search_coarse_count < C ? search_coarse_count : C;
    hipStream_t s = ctx->stream;
    int rc = 0;
    const uint32_t QB = 16384;
    const uint32_t nb = n < QB ? n : QB;
    const uint32_t keep = scc ? scc : 1;
    float* q = (float*)scratch_slot(ctx, S_Q, sizeof(float) * nb * d, &rc);
    float* qn = (float*)scratch_slot(ctx, S_QN, sizeof(float) * nb, &rc);
    int32_t* qst = (int32_t*)scratch_slot(ctx, S_QST, sizeof(int32_t) * nb, &rc);
    float* cdist = (float*)scratch_slot(ctx, S_CD, sizeof(float) * nb * C, &rc);
    float* fdist = (float*)scratch_slot(ctx, S_FD, sizeof(float) * nb * F, &rc);
    float* csd = (float*)scratch_slot(ctx, S_CSD, sizeof(float) * nb * keep, &rc);
    uint32_t* csi = (uint32_t*)scratch_slot(ctx, S_CSI, sizeof(uint32_t) * nb * keep, &rc);
    if (rc) return rc;
    const float* qin = (const float*)dev_in(ctx, S_IN_Q, queries, sizeof(float) * (size_t)n * d, device_ptrs, &rc);
    const uint32_t* cin = (const uint32_t*)dev_in(ctx, S_A0, cells, sizeof(uint32_t) * (size_t)n * k, device_ptrs, &rc);
    uint32_t* rk = device_ptrs ? ranks : (uint32_t*)scratch_slot(ctx, S_A1, sizeof(uint32_t) * (size_t)n * k, &rc);
    if (rc) return rc;
    for (uint32_t q0 = 0; q0 < n; q0 += QB) {
        const uint32_t m = (n - q0) < QB ? (n - q0) : QB;
        PK_CHECK(pk_launch_prep_queries(qin + (size_t)q0 * d, m, d, ix->p.whether_norm, q, qn, qst, s));
        PK_CHECK(pk_launch_gemm_exact(q, ix->coarse, cdist, m, C, d, C, PK_EPI_SUB_FROM_BIAS, ix->coarse_norms, s));
        PK_CHECK(pk_launch_gemm_exact(q, ix->fine, fdist, m, F, d, F, PK_EPI_NEG, nullptr, s));
        PK_CHECK(coarse_select(cdist, m, C, scc, csd, csi, nullptr, 0, s));
        PK_CHECK(pk_launch_budget_rank(csd, csi, keep, scc, fdist, ix->stationary, C, F, cin + (size_t)q0 * k, k, m,
                                       rk + (size_t)q0 * k, s));
    }
    if (!device_ptrs) PK_CHECK(hipMemcpyAsync(ranks, rk, sizeof(uint32_t) * (size_t)n * k, hipMemcpyDeviceToHost, s));
    PK_CHECK(hipStreamSynchronize(s));
    return PUCKGPU_OK;
}

// ---------------------------------------------------------------------------
// Learned probe
int puckgpu_probe_create(puckgpu_ctx* ctx, uint32_t dim, const puckgpu_probe_shape* shape, puckgpu_probe** out) {
    if (!ctx || !shape || !out) return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    *out = nullptr;
    if (dim == 0 || dim > 4096 || shape->hidden == 0 || shape->hidden > 4096)
        return set_err(ctx, PUCKGPU_EINVAL, "probe: dim and hidden must be in 1..4096");
    if (shape->coarse_features > 8192 || shape->fine_features > 8192)
        return set_err(ctx, PUCKGPU_EINVAL, "probe: coarse_features and fine_features must be at most 8192");
    {
        // the parameter count in 64 bits (the kernels index parameters in 32;
        // with the bounds above it stays near 1e8)
        const uint64_t D = (uint64_t)dim + shape->coarse_features + shape->fine_features, H = shape->hidden;
        const uint64_t cnt = H * D + H + H * H + H + H + 1 + 2 * D;
        if (cnt > UINT32_MAX)
            return set_err(ctx, PUCKGPU_EINVAL, "probe: parameter count exceeds 32-bit indexing");
    }
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    puckgpu_probe* p = new puckgpu_probe();
    p->device = ctx->device;
    p->dim = dim;
    p->nc = shape->coarse_features;
    p->nfine = shape->fine_features;
    p->D = dim + p->nc + p->nfine;
    p->H = shape->hidden;
    p->count = pk_probe_param_count(p->D, p->H);
    p->trained = pk_probe_trained_count(p->D, p->H);
    int rc = 0;
    if ((rc = dalloc(ctx, &p->params, p->count)) || (rc = dalloc(ctx, &p->m, p->trained)) ||
        (rc = dalloc(ctx, &p->v, p->trained)) || (rc = dalloc(ctx, &p->grads, p->trained)) ||
        (rc = dalloc(ctx, &p->loss, 4)) || (rc = dalloc(ctx, &p->wbf, p->trained))) {
        puckgpu_probe_free(p);
        return rc;
    }
    if (hipMemset(p->params, 0, sizeof(float) * p->count) != hipSuccess ||
        hipMemset(p->m, 0, sizeof(float) * p->trained) != hipSuccess ||
        hipMemset(p->v, 0, sizeof(float) * p->trained) != hipSuccess ||
        hipMemset(p->wbf, 0, sizeof(uint16_t) * p->trained) != hipSuccess) {
        puckgpu_probe_free(p);
        return set_err(ctx, PUCKGPU_EHIP, "probe: clearing parameters failed");
    }
    *out = p;
    return PUCKGPU_OK;
}

void puckgpu_probe_free(puckgpu_probe* p) {
    if (!p) return;
    (void)hipSetDevice(p->device);
    void* ptrs[] = {p->params, p->m, p->v, p->grads, p->loss, p->wbf};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    delete p;
}

uint32_t puckgpu_probe_param_count(const puckgpu_probe* p) { return p ? p->count : 0u; }

int puckgpu_probe_set_params(puckgpu_ctx* ctx, puckgpu_probe* p, const float* params, int device_ptrs) {
    if (!ctx || !p || !params) return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    hipStream_t s = ctx->stream;
    PK_CHECK(hipMemcpyAsync(p->params, params, sizeof(float) * p->count,
                            device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    PK_CHECK(hipMemsetAsync(p->m, 0, sizeof(float) * p->trained, s));
    PK_CHECK(hipMemsetAsync(p->v, 0, sizeof(float) * p->trained, s));
    PK_CHECK(pk_launch_to_bf16(p->params, p->trained, p->wbf, s));
    PK_CHECK(hipStreamSynchronize(s));
    p->step = 0;
    return PUCKGPU_OK;
}

int puckgpu_probe_get_params(puckgpu_ctx* ctx, const puckgpu_probe* p, float* params, int device_ptrs) {
    if (!ctx || !p || !params) return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    PK_CHECK(hipMemcpyAsync(params, p->params, sizeof(float) * p->count,
                            device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx->stream));
    PK_CHECK(hipStreamSynchronize(ctx->stream));
    return PUCKGPU_OK;
}

int puckgpu_probe_features(puckgpu_ctx* ctx, const puckgpu_index* ix, const puckgpu_probe* p, uint32_t n,
                           const float* queries, float* features, uint32_t* fine_order, int device_ptrs) {
    if (!ctx || !ix || !p || !features || (n && !queries)) return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    const uint32_t d = ix->p.dim, C = ix->p.coarse_cluster_count, F = ix->p.fine_cluster_count;
    if (p->dim != d || p->nc > C || p->nfine > F)
        return set_err(ctx, PUCKGPU_EINVAL, "probe shape does not fit the index");
    if (n == 0) return PUCKGPU_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    hipStream_t s = ctx->stream;
    int rc = 0;
    const uint32_t QB = 16384;
    const uint32_t nb = n < QB ? n : QB;
    const uint32_t keep = p->nc ? p->nc : 1;  // the nearest coarse list is always needed
    float* q = (float*)scratch_slot(ctx, S_Q, sizeof(float) * nb * d, &rc);
    float* qn = (float*)scratch_slot(ctx, S_QN, sizeof(float) * nb, &rc);
    int32_t* qst = (int32_t*)scratch_slot(ctx, S_QST, sizeof(int32_t) * nb, &rc);
    float* cdist = (float*)scratch_slot(ctx, S_CD, sizeof(float) * nb * C, &rc);
    float* fdist = (float*)scratch_slot(ctx, S_FD, sizeof(float) * nb * F, &rc);
    float* csd = (float*)scratch_slot(ctx, S_CSD, sizeof(float) * nb * keep, &rc);
    uint32_t* csi = (uint32_t*)scratch_slot(ctx, S_CSI, sizeof(uint32_t) * nb * keep, &rc);
    float* fsd = (float*)scratch_slot(ctx, S_FSD, sizeof(float) * nb * F, &rc);
    uint32_t* fsi = (uint32_t*)scratch_slot(ctx, S_FSI, sizeof(uint32_t) * nb * F, &rc);
    float* feat = device_ptrs ? features : (float*)scratch_slot(ctx, S_PFEAT, sizeof(float) * nb * p->D, &rc);
    if (rc) return rc;
    const float* qin = (const float*)dev_in(ctx, S_IN_Q, queries, sizeof(float) * (size_t)n * d, device_ptrs, &rc);
    if (!qin) return rc;
    for (uint32_t q0 = 0; q0 < n; q0 += QB) {
        const uint32_t m = (n - q0) < QB ? (n - q0) : QB;
        PK_CHECK(pk_launch_prep_queries(qin + (size_t)q0 * d, m, d, ix->p.whether_norm, q, qn, qst, s));
        PK_CHECK(pk_launch_gemm_exact(q, ix->coarse, cdist, m, C, d, C, PK_EPI_SUB_FROM_BIAS, ix->coarse_norms, s));
        PK_CHECK(pk_launch_gemm_exact(q, ix->fine, fdist, m, F, d, F, PK_EPI_NEG, nullptr, s));
        PK_CHECK(coarse_select(cdist, m, C, keep, csd, csi, nullptr, 0, s));
        if (fine_order) PK_CHECK(pk_launch_sort_rows(fdist, m, F, F, F, fsd, fsi, nullptr, 0, s));
        float* fo = device_ptrs ? feat + (size_t)q0 * p->D : feat;
        float* cv = (float*)scratch_slot(ctx, S_PCV, sizeof(float) * (size_t)m * (p->nfine ? p->nfine : 1), &rc);
        if (rc) return rc;
        if ((rc = probe_features_chunk(ctx, ix, p, m, q, csd, csi, keep, fdist, cv, nullptr, fo, s))) return rc;
        if (!device_ptrs)
            PK_CHECK(hipMemcpyAsync(features + (size_t)q0 * p->D, fo, sizeof(float) * m * p->D,
                                    hipMemcpyDeviceToHost, s));
        if (fine_order)
            PK_CHECK(hipMemcpyAsync(fine_order + (size_t)q0 * F, fsi, sizeof(uint32_t) * m * F,
                                    device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
        PK_CHECK(hipStreamSynchronize(s));
    }
    return PUCKGPU_OK;
}

int puckgpu_probe_train_step(puckgpu_ctx* ctx, puckgpu_probe* p, uint32_t n, const float* features,
                             const float* target, const puckgpu_adam_params* adam, float* loss, float* grads,
                             int device_ptrs) {
    if (!ctx || !p || (n && (!features || !target))) return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    if (n == 0) return set_err(ctx, PUCKGPU_EINVAL, "empty batch");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    hipStream_t s = ctx->stream;
    int rc = 0;
    const float* f = (const float*)dev_in(ctx, S_A0, features, sizeof(float) * (size_t)n * p->D, device_ptrs, &rc);
    const float* t = (const float*)dev_in(ctx, S_PT, target, sizeof(float) * n, device_ptrs, &rc);
    void* work = scratch_slot(ctx, S_PWORK, pk_probe_work_bytes(n, p->D, p->H), &rc);
    if (rc) return rc;
    PK_CHECK(pk_probe_pass(p->params, p->wbf, p->D, p->H, n, f, t, work, p->grads, p->loss, nullptr, s));
    if (grads)
        PK_CHECK(hipMemcpyAsync(grads, p->grads, sizeof(float) * p->trained,
                                device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    if (adam) {
        p->step++;
        PK_CHECK(pk_probe_adam_step(p->params, p->m, p->v, p->grads, p->wbf, p->D, p->H, adam->lr, adam->beta1,
                                    adam->beta2, adam->eps, p->step, s));
    }
    float ls = 0.f;
    if (loss) PK_CHECK(hipMemcpyAsync(&ls, p->loss, sizeof(float), hipMemcpyDeviceToHost, s));
    PK_CHECK(hipStreamSynchronize(s));
    if (loss) *loss = ls / (float)n;
    return PUCKGPU_OK;
}

int puckgpu_probe_predict(puckgpu_ctx* ctx, const puckgpu_probe* p, uint32_t n, const float* features, float* y,
                          int device_ptrs) {
    if (!ctx || !p || (n && (!features || !y))) return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    if (n == 0) return PUCKGPU_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    hipStream_t s = ctx->stream;
    int rc = 0;
    const uint32_t QB = 65536;
    const uint32_t nb = n < QB ? n : QB;
    const float* f = (const float*)dev_in(ctx, S_A0, features, sizeof(float) * (size_t)n * p->D, device_ptrs, &rc);
    void* work = scratch_slot(ctx, S_PWORK, pk_probe_work_bytes(nb, p->D, p->H), &rc);
    float* yb = device_ptrs ? y : (float*)scratch_slot(ctx, S_PY, sizeof(float) * n, &rc);
    if (rc) return rc;
    for (uint32_t q0 = 0; q0 < n; q0 += QB) {
        const uint32_t m = (n - q0) < QB ? (n - q0) : QB;
        PK_CHECK(pk_probe_pass(p->params, p->wbf, p->D, p->H, m, f + (size_t)q0 * p->D, nullptr, work, nullptr,
                               nullptr, yb + q0, s));
    }
    if (!device_ptrs) PK_CHECK(hipMemcpyAsync(y, yb, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    PK_CHECK(hipStreamSynchronize(s));
    return PUCKGPU_OK;
}

// ---------------------------------------------------------------------------
// Batched server
}  // extern "C"

struct PkBucket {
    uint32_t cap = 0;
    int shape = 0;
    uint32_t rerank_rows = 2;
    PassBufs b;
    float* qin = nullptr;
    uint32_t* qtopk = nullptr;
    uint32_t* qscc = nullptr;
    uint32_t* nvalid = nullptr;
    float* ft = nullptr;
    uint32_t* fi = nullptr;
    uint32_t* fn = nullptr;
    float* od = nullptr;
    uint32_t* oi = nullptr;
    uint32_t* on = nullptr;
    std::vector<void*> allocs;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t version = 0;  // the index layout the graph was captured on
    PkWalkArgs wa{};
    PkRerankArgs ra{};
};

namespace {

template <typename T>
int balloc(puckgpu_ctx* ctx, PkBucket* bk, T** p, size_t count) {
    int rc = dalloc(ctx, p, count);
    if (!rc) bk->allocs.push_back((void*)*p);
    return rc;
}

void bucket_release(PkBucket* bk) {
    if (!bk) return;
    if (bk->exec) (void)hipGraphExecDestroy(bk->exec);
    if (bk->graph) (void)hipGraphDestroy(bk->graph);
    for (void* p : bk->allocs) (void)hipFree(p);
    delete bk;
}

int bucket_create(puckgpu_server* sv, uint32_t cap, const PassCfg& cfg, PkBucket** out) {
    puckgpu_ctx* ctx = sv->ctx;
    const puckgpu_index* ix = sv->ix;
    const uint32_t d = ix->p.dim, C = ix->p.coarse_cluster_count, F = ix->p.fine_cluster_count;
    const uint32_t keep = cfg.keep ? cfg.keep : 1, ftk = cfg.ftk, os = cfg.ostride;
    PkBucket* bk = new PkBucket();
    bk->cap = cap;
    PassBufs& b = bk->b;
    int rc = 0;
    if ((rc = balloc(ctx, bk, &b.q, (size_t)cap * d)) || (rc = balloc(ctx, bk, &b.qn, cap)) ||
        (rc = balloc(ctx, bk, &b.qst, cap)) || (rc = balloc(ctx, bk, &b.cdist, (size_t)cap * C)) ||
        (rc = balloc(ctx, bk, &b.fdist, (size_t)cap * F)) || (rc = balloc(ctx, bk, &b.csd, (size_t)cap * keep)) ||
        (rc = balloc(ctx, bk, &b.csi, (size_t)cap * keep)) || (rc = balloc(ctx, bk, &b.fsd, (size_t)cap * F)) ||
        (rc = balloc(ctx, bk, &b.fsi, (size_t)cap * F)) || (rc = balloc(ctx, bk, &b.flags, cap)) ||
        (rc = balloc(ctx, bk, &b.tie_ids, (size_t)cap * 16)) || (rc = balloc(ctx, bk, &b.tie_n, cap)) ||
        (rc = balloc(ctx, bk, &b.stats, (size_t)cap * 4)) || (rc = balloc(ctx, bk, &b.qlist, cap)) ||
        (rc = balloc(ctx, bk, &b.wlist, 2 * (size_t)cap + 1)) ||
        (rc = balloc(ctx, bk, &bk->qin, (size_t)cap * d)) || (rc = balloc(ctx, bk, &bk->qtopk, cap)) ||
        (rc = balloc(ctx, bk, &bk->qscc, cap)) || (rc = balloc(ctx, bk, &bk->nvalid, 1)) ||
        (rc = balloc(ctx, bk, &bk->ft, (size_t)cap * ftk)) || (rc = balloc(ctx, bk, &bk->fi, (size_t)cap * ftk)) ||
        (rc = balloc(ctx, bk, &bk->fn, cap)) || (rc = balloc(ctx, bk, &bk->od, (size_t)cap * os)) ||
        (rc = balloc(ctx, bk, &bk->oi, (size_t)cap * os)) || (rc = balloc(ctx, bk, &bk->on, cap))) {
        bucket_release(bk);
        return rc;
    }
    if (cfg.precision) {
        const uint32_t ld = pkc_ld(d);
        if ((rc = balloc(ctx, bk, &b.q16, (size_t)cap * ld)) || (rc = balloc(ctx, bk, &b.q8, (size_t)cap * ld)) ||
            (rc = balloc(ctx, bk, &b.qs, cap)) || (rc = balloc(ctx, bk, &b.cand_d, (size_t)cap * cfg.kc)) ||
            (rc = balloc(ctx, bk, &b.cand_i, (size_t)cap * cfg.kc))) {
            bucket_release(bk);
            return rc;
        }
    }
    b.qorder = b.wlist + cap;
    b.wcount = b.wlist + 2 * (size_t)cap;
    if (sv->probe) {
        const puckgpu_probe* pr = sv->probe;
        void* w = nullptr;
        if ((rc = balloc(ctx, bk, &b.pfeat, (size_t)cap * pr->D)) || (rc = balloc(ctx, bk, &b.pnf, cap)) ||
            (rc = balloc(ctx, bk, &b.py, cap)) || (rc = balloc(ctx, bk, &b.pcv, (size_t)cap * (pr->nfine ? pr->nfine : 1))) ||
            (rc = balloc(ctx, bk, &b.pcid, (size_t)cap * (pr->nfine ? pr->nfine : 1))) ||
            (rc = balloc(ctx, bk, (char**)&w, pk_probe_work_bytes(cap, pr->D, pr->H)))) {
            bucket_release(bk);
            return rc;
        }
        b.pwork = w;
    }
    if (cfg.budget) {
        if ((rc = balloc(ctx, bk, &b.sc_cells, (size_t)cap * PK_SC_CELLS)) ||
            (rc = balloc(ctx, bk, &b.sc_meta, 3 * (size_t)cap + 1))) {
            bucket_release(bk);
            return rc;
        }
        b.sc_olist = b.sc_meta + 2 * (size_t)cap;
        b.sc_ocount = b.sc_meta + 3 * (size_t)cap;
    }
    *out = bk;
    return PUCKGPU_OK;
}

// The bucket's pass on its own buffers (eager, or under capture).
int bucket_enqueue(puckgpu_server* sv, PkBucket* bk, const PassCfg& cfg, hipStream_t s) {
    return enqueue_pass(sv->ctx, sv->ix, cfg, bk->b, bk->cap, bk->qin, bk->nvalid, bk->qtopk, bk->qscc, nullptr,
                        bk->ft, bk->fi, bk->fn, bk->od, bk->oi, bk->on, nullptr, 0, nullptr, &bk->wa, &bk->ra, s);
}

int bucket_capture(puckgpu_server* sv, PkBucket* bk, const PassCfg& cfg) {
    puckgpu_ctx* ctx = sv->ctx;
    hipStream_t s = ctx->stream;
    if (bk->exec) (void)hipGraphExecDestroy(bk->exec);
    if (bk->graph) (void)hipGraphDestroy(bk->graph);
    bk->exec = nullptr;
    bk->graph = nullptr;
    PK_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = bucket_enqueue(sv, bk, cfg, s);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "graph capture failed: %s", hipGetErrorString(e));
    bk->graph = g;
    PK_CHECK(hipGraphInstantiate(&bk->exec, g, nullptr, nullptr, 0));
    bk->version = sv->ix->version;
    return PUCKGPU_OK;
}

int server_cfg(puckgpu_server* sv, PassCfg& cfg, ProbeUse& pu) {
    pu.probe = sv->probe;
    pu.bp = sv->bp;
    pu.budget_out = nullptr;
    const int rc = make_cfg(sv->ctx, sv->ix, &sv->sp, sv->sp.topk_max, sv->sp.topk_max, sv->probe ? &pu : nullptr,
                            cfg);
    cfg.budget = sv->probe != nullptr;
    return rc;
}

}  // namespace

namespace {

#define PK_ASSIGN_CHUNK 32768u

// Is the fine bound -sqrt(|f|^2/2) non-decreasing (fine centers stored by
// descending norm)?  The assignment kernel may then stop a list early.
// fbpm (may be NULL) receives the prefix maxima max(fb[0..f]) on the host.
int fine_bound_sorted(puckgpu_ctx* ctx, const float* fb, uint32_t F, int* sorted, std::vector<float>* fbpm = nullptr) {
    std::vector<float> fb_h(F);
    PK_CHECK(hipMemcpyAsync(fb_h.data(), fb, 4ull * F, hipMemcpyDeviceToHost, ctx->stream));
    PK_CHECK(hipStreamSynchronize(ctx->stream));
    *sorted = 1;
    for (uint32_t f = 1; f < F; ++f)
        if (!(fb_h[f - 1] <= fb_h[f])) *sorted = 0;
    if (fbpm) {
        fbpm->resize(F);
        float m = fb_h[0];
        for (uint32_t f = 0; f < F; ++f) {
            m = fb_h[f] > m ? fb_h[f] : m;
            (*fbpm)[f] = m;
        }
    }
    return PUCKGPU_OK;
}

// Block lower bounds of a stationary table for pk_assign_bounded.
struct AssignBounds {
    const float* stmin = nullptr;
    const float* stminrow = nullptr;
    const uint32_t* stabs = nullptr;
    const float* fbpm = nullptr;
};

// nearest_cell_assign (hierarchical_cluster_index.cpp:1042-1085) of m <=
// PK_ASSIGN_CHUNK points on the device, over the coarse half-norms cn, the
// stationary values st and the fine bound fb; cell_out / dist_out (may be
// NULL) on the device.  With bounds (and assign mode 0 or 2) the bounded
// kernel assigns every point it can certify and the sequential kernel the
// rest; without, the sequential kernel assigns every point.
int assign_chunk(puckgpu_ctx* ctx, const float* cd, const float* fd, const float* cn, const float* st, const float* fb,
                 int fb_sorted, uint32_t C, uint32_t F, uint32_t dim, uint32_t m, const float* xp, uint32_t* cell_out,
                 float* dist_out, hipStream_t s, const AssignBounds* ab = nullptr) {
    int rc = 0;
    float* ipc = (float*)scratch_slot(ctx, S_A6, 4ull * m * C, &rc);
    float* ipf = (float*)scratch_slot(ctx, S_A7, 4ull * m * F, &rc);
    float* xn = (float*)scratch_slot(ctx, S_QN, 4ull * m, &rc);
    uint32_t* ord_d = (uint32_t*)scratch_slot(ctx, S_QLIST, 4ull * m, &rc);
    if (rc) return rc;
    PK_CHECK(pk_launch_row_norms(xp, m, dim, 1, xn, s));
    PK_CHECK(pk_launch_gemm_exact(xp, cd, ipc, m, C, dim, C, PK_EPI_RAW, nullptr, s));
    PK_CHECK(pk_launch_gemm_exact(xp, fd, ipf, m, F, dim, F, PK_EPI_RAW, nullptr, s));
    ctx->assign_points += m;
    if (ab && ctx->assign_mode != 1) {
        uint32_t* cnt = (uint32_t*)scratch_slot(ctx, S_AB_CNT, 4, &rc);
        if (rc) return rc;
        PK_CHECK(pk_launch_assign_bounded(ipc, ipf, xn, cn, fb, ab->fbpm, st, ab->stmin, ab->stminrow, ab->stabs, m, C,
                                          F, fb_sorted, ctx->assign_mode == 2, cell_out, dist_out, ord_d, cnt, s));
        uint32_t k = 0;
        PK_CHECK(hipMemcpyAsync(&k, cnt, 4, hipMemcpyDeviceToHost, s));
        PK_CHECK(hipStreamSynchronize(s));
        if (k > m) return set_err(ctx, PUCKGPU_EHIP, "assignment listed %u of %u points", k, m);
        ctx->assign_sequential += k;
        if (k) PK_CHECK(pk_launch_assign_cells(ipc, ipf, xn, cn, fb, st, k, C, F, fb_sorted, ord_d, cell_out, dist_out, s));
        PK_CHECK(hipStreamSynchronize(s));
        return PUCKGPU_OK;
    }
    std::vector<uint32_t> ord_h(m), perm_h(m);
    // visit points grouped by nearest coarse cluster (L2 reuse of the
    // stationary rows); results are written by point index
    PK_CHECK(pk_launch_nearest_coarse(ipc, xn, cn, m, C, ord_d, s));
    PK_CHECK(hipMemcpyAsync(ord_h.data(), ord_d, 4ull * m, hipMemcpyDeviceToHost, s));
    PK_CHECK(hipStreamSynchronize(s));
    {
        std::vector<uint32_t> cnt(C + 1, 0);
        for (uint32_t i = 0; i < m; ++i) cnt[ord_h[i] + 1]++;
        for (uint32_t c = 0; c < C; ++c) cnt[c + 1] += cnt[c];
        for (uint32_t i = 0; i < m; ++i) perm_h[cnt[ord_h[i]]++] = i;
    }
    PK_CHECK(hipMemcpyAsync(ord_d, perm_h.data(), 4ull * m, hipMemcpyHostToDevice, s));
    ctx->assign_sequential += m;
    PK_CHECK(pk_launch_assign_cells(ipc, ipf, xn, cn, fb, st, m, C, F, fb_sorted, ord_d, cell_out, dist_out, s));
    // the host vectors above must outlive the copies
    PK_CHECK(hipStreamSynchronize(s));
    return PUCKGPU_OK;
}

}  // namespace

extern "C" {

int puckgpu_server_create(puckgpu_ctx* ctx, const puckgpu_index* ix, const puckgpu_probe* probe,
                          const puckgpu_search_params* sp, const puckgpu_budget_params* bp, uint32_t max_batch,
                          puckgpu_server** out) {
    if (!ctx || !ix || !sp || !out) return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    *out = nullptr;
    if (max_batch == 0 || max_batch > (1u << 16)) return set_err(ctx, PUCKGPU_EINVAL, "max_batch must be in 1..65536");
    if (sp->topk_max == 0) return set_err(ctx, PUCKGPU_EINVAL, "topk_max must be positive");
    if (probe && !bp) return set_err(ctx, PUCKGPU_EINVAL, "a probe needs budget parameters");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    puckgpu_server* sv = new puckgpu_server();
    sv->ctx = ctx;
    sv->ix = ix;
    sv->sp = *sp;
    sv->sp.replay_all = 0;
    sv->probe = probe;
    if (bp) sv->bp = *bp;
    sv->max_batch = max_batch;
    PassCfg cfg;
    ProbeUse pu;
    int rc = server_cfg(sv, cfg, pu);
    for (uint32_t cap = 1; !rc; cap <<= 1) {
        PkBucket* bk = nullptr;
        rc = bucket_create(sv, cap < max_batch ? cap : max_batch, cfg, &bk);
        if (!rc) sv->buckets.push_back(bk);
        if (cap >= max_batch) break;
    }
    if (rc) {
        puckgpu_server_free(sv);
        return rc;
    }
    *out = sv;
    return PUCKGPU_OK;
}

void puckgpu_server_free(puckgpu_server* sv) {
    if (!sv) return;
    (void)hipSetDevice(sv->ctx->device);
    (void)hipStreamSynchronize(sv->ctx->stream);
    for (PkBucket* bk : sv->buckets) bucket_release(bk);
    delete sv;
}

uint32_t puckgpu_server_buckets(const puckgpu_server* sv) { return sv ? (uint32_t)sv->buckets.size() : 0u; }

int puckgpu_server_search(puckgpu_server* sv, uint32_t n, const float* queries, const uint32_t* topk,
                          const uint32_t* scc, float* distance, uint32_t* labels, uint32_t* result_num,
                          int device_ptrs) {
    if (!sv) return PUCKGPU_EINVAL;
    puckgpu_ctx* ctx = sv->ctx;
    if (n == 0) return PUCKGPU_OK;
    if (!queries || !distance || !labels) return set_err(ctx, PUCKGPU_EINVAL, "null array");
    if (n > sv->max_batch) return set_err(ctx, PUCKGPU_EINVAL, "batch of %u above max_batch %u", n, sv->max_batch);
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    PassCfg cfg;
    ProbeUse pu;
    int rc = server_cfg(sv, cfg, pu);
    if (rc) return rc;
    PkBucket* bk = nullptr;
    for (PkBucket* c : sv->buckets)
        if (c->cap >= n) { bk = c; break; }
    cfg.gemm_shape = bk->shape;
    cfg.rerank_rows = bk->rerank_rows;
    hipStream_t s = ctx->stream;
    const uint32_t d = sv->ix->p.dim, os = cfg.ostride;
    const hipMemcpyKind in = device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind out = device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    PK_CHECK(hipMemcpyAsync(bk->qin, queries, sizeof(float) * (size_t)n * d, in, s));
    std::vector<uint32_t> fill;
    if (topk) {
        PK_CHECK(hipMemcpyAsync(bk->qtopk, topk, sizeof(uint32_t) * n, in, s));
    } else {
        fill.assign(n, sv->sp.topk_max);
        PK_CHECK(hipMemcpyAsync(bk->qtopk, fill.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
    }
    std::vector<uint32_t> fill2;
    if (scc) {
        PK_CHECK(hipMemcpyAsync(bk->qscc, scc, sizeof(uint32_t) * n, in, s));
    } else {
        fill2.assign(n, cfg.scc);
        PK_CHECK(hipMemcpyAsync(bk->qscc, fill2.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
    }
    uint32_t nv = n;
    PK_CHECK(hipMemcpyAsync(bk->nvalid, &nv, sizeof(uint32_t), hipMemcpyHostToDevice, s));
    if (!bk->exec || bk->version != sv->ix->version) {
        // one eager pass first (sets kernel attributes outside any capture);
        // an insert since the capture moved the index arrays: capture again
        if ((rc = bucket_enqueue(sv, bk, cfg, s))) return rc;
        PK_CHECK(hipStreamSynchronize(s));
        if ((rc = bucket_capture(sv, bk, cfg))) return rc;
    }
    PK_CHECK(hipGraphLaunch(bk->exec, s));
    uint8_t* hflags = host_pinned(ctx, n);
    if (!hflags) return set_err(ctx, PUCKGPU_ENOMEM, "pinned host allocation failed");
    PK_CHECK(hipMemcpyAsync(hflags, bk->b.flags, n, hipMemcpyDeviceToHost, s));
    PK_CHECK(hipStreamSynchronize(s));
    std::vector<uint32_t> hlist(2 * (size_t)n);
    if (cfg.precision) std::fill(hflags, hflags + n, (uint8_t)0);
    const int nrep = replay_pass(ctx, sv->ix, cfg, bk->b, n, hflags, 0, bk->wa, bk->ra, hlist.data(),
                                 hlist.data() + n, s);
    if (nrep < 0) return nrep;
    PK_CHECK(hipMemcpyAsync(distance, bk->od, sizeof(float) * (size_t)n * os, out, s));
    PK_CHECK(hipMemcpyAsync(labels, bk->oi, sizeof(uint32_t) * (size_t)n * os, out, s));
    if (result_num) PK_CHECK(hipMemcpyAsync(result_num, bk->on, sizeof(uint32_t) * n, out, s));
    PK_CHECK(hipStreamSynchronize(s));
    ctx->counters[4] = (uint64_t)nrep;
    return PUCKGPU_OK;
}

int puckgpu_server_tuned(const puckgpu_server* sv, uint32_t bucket, int32_t* gemm_shape, uint32_t* rerank_rows) {
    if (!sv || bucket >= sv->buckets.size()) return PUCKGPU_EINVAL;
    if (gemm_shape) *gemm_shape = sv->buckets[bucket]->shape;
    if (rerank_rows) *rerank_rows = sv->buckets[bucket]->rerank_rows;
    return PUCKGPU_OK;
}

int puckgpu_set_rerank_rows(puckgpu_ctx* ctx, uint32_t rows) {
    if (!ctx) return PUCKGPU_EINVAL;
    if (rows != 1 && rows != 2 && rows != 4) return set_err(ctx, PUCKGPU_EINVAL, "rerank rows must be 1, 2 or 4");
    ctx->rerank_rows = rows;
    return PUCKGPU_OK;
}

int puckgpu_server_tune(puckgpu_server* sv, const float* sample, uint32_t reps, int32_t* shapes_out,
                        float* times_out) {
    if (!sv || !sample) return sv ? set_err(sv->ctx, PUCKGPU_EINVAL, "null argument") : PUCKGPU_EINVAL;
    puckgpu_ctx* ctx = sv->ctx;
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    PassCfg cfg;
    ProbeUse pu;
    int rc = server_cfg(sv, cfg, pu);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    const uint32_t d = sv->ix->p.dim, ns = pk_gemm_shapes();
    if (reps == 0) reps = 1;
    for (size_t j = 0; j < sv->buckets.size(); ++j) {
        PkBucket* bk = sv->buckets[j];
        PK_CHECK(hipMemcpyAsync(bk->qin, sample, sizeof(float) * (size_t)bk->cap * d, hipMemcpyHostToDevice, s));
        std::vector<uint32_t> t(bk->cap, sv->sp.topk_max), c(bk->cap, cfg.scc);
        PK_CHECK(hipMemcpyAsync(bk->qtopk, t.data(), sizeof(uint32_t) * bk->cap, hipMemcpyHostToDevice, s));
        PK_CHECK(hipMemcpyAsync(bk->qscc, c.data(), sizeof(uint32_t) * bk->cap, hipMemcpyHostToDevice, s));
        PK_CHECK(hipMemcpyAsync(bk->nvalid, &bk->cap, sizeof(uint32_t), hipMemcpyHostToDevice, s));
        float best = 0.f;
        int bshape = 0;
        const uint32_t C = sv->ix->p.coarse_cluster_count, F = sv->ix->p.fine_cluster_count;
        PK_CHECK(pk_launch_prep_queries_ex(bk->qin, bk->cap, d, sv->ix->p.whether_norm, bk->b.q, bk->b.qn,
                                           bk->b.qst, nullptr, nullptr, s));
        for (uint32_t sh = 0; sh < ns; ++sh) {
            // the tuned stage: both center GEMMs of the bucket
            auto gemms = [&]() -> hipError_t {
                hipError_t e = pk_launch_gemm_exact_shape(bk->b.q, sv->ix->coarse, bk->b.cdist, bk->cap, C, d, C,
                                                          PK_EPI_SUB_FROM_BIAS, sv->ix->coarse_norms, (int)sh, s);
                if (e != hipSuccess) return e;
                return pk_launch_gemm_exact_shape(bk->b.q, sv->ix->fine, bk->b.fdist, bk->cap, F, d, F, PK_EPI_NEG,
                                                  nullptr, (int)sh, s);
            };
            PK_CHECK(gemms());  // warm-up
            PK_CHECK(hipEventRecord(ctx->ev[0], s));
            for (uint32_t r = 0; r < reps; ++r) PK_CHECK(gemms());
            PK_CHECK(hipEventRecord(ctx->ev[1], s));
            PK_CHECK(hipEventSynchronize(ctx->ev[1]));
            float ms = 0.f;
            PK_CHECK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
            ms /= (float)reps;
            if (times_out) times_out[j * ns + sh] = ms;
            if (sh == 0 || ms < best) {
                best = ms;
                bshape = (int)sh;
            }
        }
        bk->shape = bshape;
        if (shapes_out) shapes_out[j] = bshape;
        cfg.gemm_shape = bshape;
        // the rerank's rows in flight per lane group: the whole pass timed
        float bestp = 0.f;
        uint32_t brows = 2;
        for (uint32_t rows : {1u, 2u, 4u}) {
            cfg.rerank_rows = rows;
            if ((rc = bucket_enqueue(sv, bk, cfg, s))) return rc;  // warm-up
            PK_CHECK(hipEventRecord(ctx->ev[0], s));
            for (uint32_t r = 0; r < reps; ++r)
                if ((rc = bucket_enqueue(sv, bk, cfg, s))) return rc;
            PK_CHECK(hipEventRecord(ctx->ev[1], s));
            PK_CHECK(hipEventSynchronize(ctx->ev[1]));
            float ms = 0.f;
            PK_CHECK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
            if (rows == 1 || ms < bestp) {
                bestp = ms;
                brows = rows;
            }
        }
        bk->rerank_rows = brows;
        cfg.rerank_rows = brows;
        if ((rc = bucket_capture(sv, bk, cfg))) return rc;
    }
    return PUCKGPU_OK;
}

int puckgpu_last_filter(puckgpu_ctx* ctx, uint32_t n, uint32_t ftk, float* filt_dist, uint32_t* filt_mem,
                        uint32_t* filt_n) {
    if (!ctx) return PUCKGPU_EINVAL;
    if (n != ctx->last_n || ftk != ctx->last_ftk) return set_err(ctx, PUCKGPU_EINVAL, "no matching last search");
    if (ctx->bufs.size() <= S_FN) return set_err(ctx, PUCKGPU_EINVAL, "no last search");
    if (filt_dist) PK_CHECK(hipMemcpy(filt_dist, ctx->bufs[S_FT].p, 4ull * n * ftk, hipMemcpyDeviceToHost));
    if (filt_mem) PK_CHECK(hipMemcpy(filt_mem, ctx->bufs[S_FI].p, 4ull * n * ftk, hipMemcpyDeviceToHost));
    if (filt_n) PK_CHECK(hipMemcpy(filt_n, ctx->bufs[S_FN].p, 4ull * n, hipMemcpyDeviceToHost));
    return PUCKGPU_OK;
}

int puckgpu_last_stage_ms(puckgpu_ctx* ctx, float* ms5, uint64_t* counters13) {
    if (!ctx) return PUCKGPU_EINVAL;
    for (int k = 0; k < 5; ++k)
        if (ms5) ms5[k] = ctx->stage_ms[k];
    for (int k = 0; k < 13; ++k)
        if (counters13) counters13[k] = ctx->counters[k];
    return PUCKGPU_OK;
}

int puckgpu_center_distances(puckgpu_ctx* ctx, const puckgpu_index* ix, uint32_t precision, uint32_t n,
                             const float* queries, float* coarse_out, float* fine_out, int device_ptrs) {
    if (!ctx || !ix || (n && (!queries || !coarse_out || !fine_out)))
        return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    if (precision > 2) return set_err(ctx, PUCKGPU_EINVAL, "precision must be 0, 1 or 2");
    if (n == 0) return PUCKGPU_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    const uint32_t d = ix->p.dim, C = ix->p.coarse_cluster_count, F = ix->p.fine_cluster_count, ld = pkc_ld(d);
    if (precision && ld > 256) return set_err(ctx, PUCKGPU_EINVAL, "reduced precision needs dim <= 256");
    hipStream_t s = ctx->stream;
    int rc = 0;
    if (precision && (rc = ensure_lowp(ctx, ix, (int)precision))) return rc;
    float* q = (float*)scratch_slot(ctx, S_Q, sizeof(float) * (size_t)n * d, &rc);
    float* qn = (float*)scratch_slot(ctx, S_QN, sizeof(float) * n, &rc);
    int32_t* qst = (int32_t*)scratch_slot(ctx, S_QST, sizeof(int32_t) * n, &rc);
    uint16_t* q16 = (uint16_t*)scratch_slot(ctx, S_LP16, sizeof(uint16_t) * (size_t)n * ld, &rc);
    uint8_t* q8 = (uint8_t*)scratch_slot(ctx, S_LP8, (size_t)n * ld, &rc);
    float* qs = (float*)scratch_slot(ctx, S_LPS, sizeof(float) * n, &rc);
    float* co = device_ptrs ? coarse_out : (float*)scratch_slot(ctx, S_CD, sizeof(float) * (size_t)n * C, &rc);
    float* fo = device_ptrs ? fine_out : (float*)scratch_slot(ctx, S_FD, sizeof(float) * (size_t)n * F, &rc);
    if (rc) return rc;
    const float* qin = (const float*)dev_in(ctx, S_IN_Q, queries, sizeof(float) * (size_t)n * d, device_ptrs, &rc);
    if (!qin) return rc;
    PK_CHECK(pk_launch_prep_queries(qin, n, d, ix->p.whether_norm, q, qn, qst, s));
    if (precision == 0) {
        PK_CHECK(pk_launch_gemm_exact(q, ix->coarse, co, n, C, d, C, PK_EPI_SUB_FROM_BIAS, ix->coarse_norms, s));
        PK_CHECK(pk_launch_gemm_exact(q, ix->fine, fo, n, F, d, F, PK_EPI_NEG, nullptr, s));
    } else {
        const bool bf = precision == 1;
        PK_CHECK(pkc_launch_convert(q, n, d, (int)precision, q16, q8, qs, s));
        const void* qa = bf ? (const void*)q16 : (const void*)q8;
        PK_CHECK(pkc_launch_center(qa, qs, bf ? (const void*)ix->coarse16 : (const void*)ix->coarse8, ix->coarse8s, n,
                                   C, d, (int)precision, PK_EPI_SUB_FROM_BIAS, ix->coarse_norms, co, C, s));
        PK_CHECK(pkc_launch_center(qa, qs, bf ? (const void*)ix->fine16 : (const void*)ix->fine8, ix->fine8s, n, F, d,
                                   (int)precision, PK_EPI_NEG, nullptr, fo, F, s));
    }
    if (!device_ptrs) {
        PK_CHECK(hipMemcpyAsync(coarse_out, co, sizeof(float) * (size_t)n * C, hipMemcpyDeviceToHost, s));
        PK_CHECK(hipMemcpyAsync(fine_out, fo, sizeof(float) * (size_t)n * F, hipMemcpyDeviceToHost, s));
    }
    PK_CHECK(hipStreamSynchronize(s));
    return PUCKGPU_OK;
}

int puckgpu_last_scan_ms(puckgpu_ctx* ctx, float* scan_ms) {
    if (!ctx || !scan_ms) return PUCKGPU_EINVAL;
    *scan_ms = ctx->scan_ms;
    return PUCKGPU_OK;
}

int puckgpu_lut(puckgpu_ctx* ctx, uint32_t n, uint32_t dim, const float* queries, const float* codebook,
                uint32_t nsq, float* out, int device_ptrs) {
    if (!ctx || !queries || !codebook || !out) return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    if (nsq == 0 || dim % nsq || !(dim / nsq == 1 || dim / nsq == 2 || dim / nsq == 4))
        return set_err(ctx, PUCKGPU_EINVAL, "dim / nsq must be 1, 2 or 4");
    if (n == 0) return PUCKGPU_OK;
    int rc = 0;
    const float* qd = (const float*)dev_in(ctx, S_A0, queries, 4ull * n * dim, device_ptrs, &rc);
    const float* cb = (const float*)dev_in(ctx, S_A1, codebook, 4ull * nsq * PK_KS * (dim / nsq), device_ptrs, &rc);
    float* o = device_ptrs ? out : (float*)scratch_slot(ctx, S_A2, 4ull * n * nsq * PK_KS, &rc);
    if (rc) return rc;
    PK_CHECK(pk_launch_lut(qd, n, dim, cb, nsq, dim / nsq, o, ctx->stream));
    if (!device_ptrs) PK_CHECK(hipMemcpyAsync(out, o, 4ull * n * nsq * PK_KS, hipMemcpyDeviceToHost, ctx->stream));
    PK_CHECK(hipStreamSynchronize(ctx->stream));
    return PUCKGPU_OK;
}

// ---------------------------------------------------------------------------
int puckgpu_exact_search(puckgpu_ctx* ctx, const void* x, int x_u8, uint64_t n, uint32_t dim, const float* queries,
                         uint32_t nq, uint32_t k, uint32_t extra, uint32_t* ids_out, double* dist_out,
                         int device_ptrs) {
    if (!ctx || (n && !x) || (nq && (!queries || !ids_out || !dist_out))) return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    if (dim == 0 || k == 0) return set_err(ctx, PUCKGPU_EINVAL, "dim and k must be positive");
    if (n >= 0xffffffffull) return set_err(ctx, PUCKGPU_EINVAL, "at most 2^32 - 2 points");
    if (n < k) return set_err(ctx, PUCKGPU_EINVAL, "k = %u above the %llu points", k, (unsigned long long)n);
    const uint32_t kk = k + extra, M = 2 * kk;
    if (M > 1024) return set_err(ctx, PUCKGPU_EINVAL, "2 (k + extra) must be <= 1024");
    if (nq == 0) return PUCKGPU_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    hipStream_t s = ctx->stream;
    int rc = 0;
    const size_t esz = x_u8 ? 1 : 4;
    const void* xd = dev_in(ctx, S_EX_X, x, esz * n * dim, device_ptrs, &rc);
    const float* qd = (const float*)dev_in(ctx, S_EX_Q, queries, 4ull * nq * dim, device_ptrs, &rc);
    if (rc) return rc;
    const uint32_t QC = nq < 2048 ? nq : 2048;
    const uint32_t BC = n < 65536 ? (uint32_t)n : 65536u;
    // candidates per query: min(kk, chunk size) of every base chunk
    uint32_t total = 0;
    for (uint64_t b0 = 0; b0 < n; b0 += BC) total += (uint32_t)std::min<uint64_t>(kk, std::min<uint64_t>(BC, n - b0));
    const uint32_t Mq = M < total ? M : total;
    float* D = (float*)scratch_slot(ctx, S_EX_D, 4ull * QC * BC, &rc);
    float* xf = x_u8 ? (float*)scratch_slot(ctx, S_EX_XF, 4ull * BC * dim, &rc) : nullptr;
    float* xnh = (float*)scratch_slot(ctx, S_EX_XN, 4ull * BC, &rc);
    float* tv = (float*)scratch_slot(ctx, S_EX_TV, 4ull * QC * kk, &rc);
    uint32_t* ti = (uint32_t*)scratch_slot(ctx, S_EX_TI, 4ull * QC * kk, &rc);
    float* cv = (float*)scratch_slot(ctx, S_EX_CV, 4ull * QC * total, &rc);
    uint32_t* ci = (uint32_t*)scratch_slot(ctx, S_EX_CI, 4ull * QC * total, &rc);
    float* mv = (float*)scratch_slot(ctx, S_EX_MV, 4ull * QC * Mq, &rc);
    uint32_t* mp = (uint32_t*)scratch_slot(ctx, S_EX_MP, 4ull * QC * Mq, &rc);
    uint32_t* oi = device_ptrs ? ids_out : (uint32_t*)scratch_slot(ctx, S_EX_OI, 4ull * nq * k, &rc);
    double* od = device_ptrs ? dist_out : (double*)scratch_slot(ctx, S_EX_OD, 8ull * nq * k, &rc);
    if (rc) return rc;
    for (uint32_t q0 = 0; q0 < nq; q0 += QC) {
        const uint32_t m = nq - q0 < QC ? nq - q0 : QC;
        const float* qc = qd + (uint64_t)q0 * dim;
        uint32_t col = 0;
        for (uint64_t b0 = 0; b0 < n; b0 += BC) {
            const uint32_t bn = (uint32_t)std::min<uint64_t>(BC, n - b0);
            const float* xb;
            if (x_u8) {
                PK_CHECK(pk_launch_u8_to_f32((const uint8_t*)xd + b0 * dim, (uint64_t)bn * dim, xf, s));
                xb = xf;
            } else {
                xb = (const float*)xd + b0 * dim;
            }
            PK_CHECK(pk_launch_row_norms(xb, bn, dim, 1, xnh, s));
            PK_CHECK(pk_launch_gemm_exact(qc, xb, D, m, bn, dim, bn, PK_EPI_SUB_FROM_BIAS, xnh, s));
            const uint32_t kc = kk < bn ? kk : bn;
            PK_CHECK(pk_launch_topk_rows(D, m, bn, bn, kc, tv, ti, nullptr, 0, s));
            PK_CHECK(pk_launch_exact_append(tv, ti, m, kc, (uint32_t)b0, cv, ci, total, col, s));
            col += kc;
        }
        PK_CHECK(pk_launch_topk_rows(cv, m, total, total, Mq, mv, mp, nullptr, 0, s));
        PK_CHECK(pk_launch_exact_recheck(xd, x_u8, dim, qc, m, mp, ci, total, Mq, k, oi + (uint64_t)q0 * k,
                                         od + (uint64_t)q0 * k, s));
        report(ctx, "exact_search: %u of %u queries", q0 + m, nq);
    }
    if (!device_ptrs) {
        PK_CHECK(hipMemcpyAsync(ids_out, oi, 4ull * nq * k, hipMemcpyDeviceToHost, s));
        PK_CHECK(hipMemcpyAsync(dist_out, od, 8ull * nq * k, hipMemcpyDeviceToHost, s));
    }
    PK_CHECK(hipStreamSynchronize(s));
    return PUCKGPU_OK;
}

int puckgpu_set_progress(puckgpu_ctx* ctx, puckgpu_progress_fn fn, void* user) {
    if (!ctx) return PUCKGPU_EINVAL;
    ctx->progress = fn;
    ctx->progress_user = user;
    return PUCKGPU_OK;
}

int puckgpu_set_assign_mode(puckgpu_ctx* ctx, int mode) {
    if (!ctx) return PUCKGPU_EINVAL;
    if (mode < 0 || mode > 2) return set_err(ctx, PUCKGPU_EINVAL, "assign mode must be 0, 1 or 2");
    ctx->assign_mode = mode;
    return PUCKGPU_OK;
}

int puckgpu_assign_stats(puckgpu_ctx* ctx, uint64_t* points, uint64_t* sequential) {
    if (!ctx || !points || !sequential) return PUCKGPU_EINVAL;
    *points = ctx->assign_points;
    *sequential = ctx->assign_sequential;
    return PUCKGPU_OK;
}

int puckgpu_assign_cells(puckgpu_ctx* ctx, const float* coarse, const float* fine, uint32_t C, uint32_t F,
                         uint32_t dim, uint32_t n, const float* x, uint32_t* cell_out, float* dist_out,
                         int device_ptrs) {
    if (!ctx || !coarse || !fine || !cell_out || (n && !x)) return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    if (C == 0 || F == 0 || dim == 0) return set_err(ctx, PUCKGPU_EINVAL, "empty geometry");
    if ((uint64_t)C * F >= 0xffffffffull) return set_err(ctx, PUCKGPU_EINVAL, "C*F too large");
    if (n == 0) return PUCKGPU_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    hipStream_t s = ctx->stream;
    int rc = 0;
    const float* cd = (const float*)dev_in(ctx, S_A0, coarse, 4ull * C * dim, device_ptrs, &rc);
    const float* fd = (const float*)dev_in(ctx, S_A1, fine, 4ull * F * dim, device_ptrs, &rc);
    float* cn = (float*)scratch_slot(ctx, S_A2, 4ull * C, &rc);
    float* fnorm = (float*)scratch_slot(ctx, S_A3, 4ull * F, &rc);
    float* fb = (float*)scratch_slot(ctx, S_A4, 4ull * F, &rc);
    float* st = (float*)scratch_slot(ctx, S_A5, 4ull * C * F, &rc);
    if (rc) return rc;
    PK_CHECK(pk_launch_row_norms(cd, C, dim, 1, cn, s));
    PK_CHECK(pk_launch_row_norms(fd, F, dim, 1, fnorm, s));
    PK_CHECK(pk_launch_fine_bound(fd, F, dim, fb, s));
    PK_CHECK(pk_launch_gemm_exact(cd, fd, st, C, F, dim, F, PK_EPI_ADD_BIAS, fnorm, s));
    int fb_sorted = 1;
    std::vector<float> fbpm_h;
    if ((rc = fine_bound_sorted(ctx, fb, F, &fb_sorted, &fbpm_h))) return rc;
    // block lower bounds of the stationary table (pk_assign_bounded)
    AssignBounds ab;
    const AssignBounds* abp = nullptr;
    const uint32_t nb = (F + 31) / 32;
    if (nb <= 256) {
        float* stmin = (float*)scratch_slot(ctx, S_AB_STMIN, 4ull * C * nb, &rc);
        float* srow = (float*)scratch_slot(ctx, S_AB_ROW, 4ull * C, &rc);
        uint32_t* sabs = (uint32_t*)scratch_slot(ctx, S_AB_ABS, 4, &rc);
        float* fbpm = (float*)scratch_slot(ctx, S_AB_FBPM, 4ull * F, &rc);
        if (rc) return rc;
        PK_CHECK(pk_launch_st_blocks(st, C, F, stmin, srow, sabs, s));
        PK_CHECK(hipMemcpyAsync(fbpm, fbpm_h.data(), 4ull * F, hipMemcpyHostToDevice, s));
        ab.stmin = stmin;
        ab.stminrow = srow;
        ab.stabs = sabs;
        ab.fbpm = fbpm;
        abp = &ab;
    }
    const uint32_t CHK = PK_ASSIGN_CHUNK;
    const uint32_t cb = n < CHK ? n : CHK;
    float* xs = (float*)scratch_slot(ctx, S_Q, 4ull * cb * dim, &rc);
    uint32_t* co = (uint32_t*)scratch_slot(ctx, S_OI, 4ull * cb, &rc);
    float* cdo = (float*)scratch_slot(ctx, S_OD, 4ull * cb, &rc);
    if (rc) return rc;
    for (uint32_t p0 = 0; p0 < n; p0 += CHK) {
        const uint32_t m = (n - p0) < CHK ? (n - p0) : CHK;
        const float* xp;
        if (device_ptrs) xp = x + (size_t)p0 * dim;
        else {
            PK_CHECK(hipMemcpyAsync(xs, x + (size_t)p0 * dim, 4ull * m * dim, hipMemcpyHostToDevice, s));
            xp = xs;
        }
        uint32_t* cout = device_ptrs ? cell_out + p0 : co;
        float* dout = device_ptrs ? (dist_out ? dist_out + p0 : nullptr) : cdo;
        if ((rc = assign_chunk(ctx, cd, fd, cn, st, fb, fb_sorted, C, F, dim, m, xp, cout, dout, s, abp))) return rc;
        if (!device_ptrs) {
            PK_CHECK(hipMemcpyAsync(cell_out + p0, co, 4ull * m, hipMemcpyDeviceToHost, s));
            if (dist_out) PK_CHECK(hipMemcpyAsync(dist_out + p0, cdo, 4ull * m, hipMemcpyDeviceToHost, s));
            PK_CHECK(hipStreamSynchronize(s));
        }
    }
    PK_CHECK(hipStreamSynchronize(s));
    return PUCKGPU_OK;
}

int puckgpu_encode_pq(puckgpu_ctx* ctx, const float* coarse, const float* fine, uint32_t C, uint32_t F, uint32_t dim,
                      const float* filter_codebook, uint32_t filter_nsq, const uint8_t* filter_records,
                      const float* pq_codebook, uint32_t pq_nsq, uint32_t n, const float* x, const uint32_t* cell,
                      uint8_t* records_out, int device_ptrs) {
    if (!ctx || !coarse || !fine || !filter_codebook || !pq_codebook || !records_out ||
        (n && (!x || !cell || !filter_records)))
        return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    if (filter_nsq == 0 || pq_nsq == 0 || dim % filter_nsq || dim % pq_nsq || dim / filter_nsq > 4 || dim / pq_nsq > 4)
        return set_err(ctx, PUCKGPU_EINVAL, "dim / filter_nsq and dim / pq_nsq must be <= 4");
    if (n == 0) return PUCKGPU_OK;
    if (C == 0 || F == 0 || (uint64_t)C * F >= 0xffffffffull) return set_err(ctx, PUCKGPU_EINVAL, "bad C/F");
    if (!device_ptrs)
        for (uint32_t i = 0; i < n; ++i)
            if (cell[i] >= C * F) return set_err(ctx, PUCKGPU_EINVAL, "cell[%u] = %u out of range", i, cell[i]);
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    hipStream_t s = ctx->stream;
    int rc = 0;
    const float* cd = (const float*)dev_in(ctx, S_A0, coarse, 4ull * C * dim, device_ptrs, &rc);
    const float* fd = (const float*)dev_in(ctx, S_A1, fine, 4ull * F * dim, device_ptrs, &rc);
    const float* fcb = (const float*)dev_in(ctx, S_A2, filter_codebook, 4ull * filter_nsq * PK_KS * (dim / filter_nsq),
                                            device_ptrs, &rc);
    const float* pcb = (const float*)dev_in(ctx, S_A6, pq_codebook, 4ull * pq_nsq * PK_KS * (dim / pq_nsq),
                                            device_ptrs, &rc);
    if (rc) return rc;
    const uint64_t frec = 4 + filter_nsq, prec = 4 + pq_nsq;
    const uint32_t CHK = 1u << 20;
    const uint32_t cbn = n < CHK ? n : CHK;
    float* xs = device_ptrs ? nullptr : (float*)scratch_slot(ctx, S_A3, 4ull * cbn * dim, &rc);
    uint32_t* cs = device_ptrs ? nullptr : (uint32_t*)scratch_slot(ctx, S_A4, 4ull * cbn, &rc);
    uint8_t* fr = device_ptrs ? nullptr : (uint8_t*)scratch_slot(ctx, S_A7, frec * cbn, &rc);
    uint8_t* ro = device_ptrs ? nullptr : (uint8_t*)scratch_slot(ctx, S_A5, prec * cbn, &rc);
    if (rc) return rc;
    for (uint32_t p0 = 0; p0 < n; p0 += CHK) {
        const uint32_t m = (n - p0) < CHK ? (n - p0) : CHK;
        if (device_ptrs) {
            PK_CHECK(pk_launch_encode_pq(x + (size_t)p0 * dim, cell + p0, m, cd, fd, F, dim, fcb, filter_nsq,
                                         filter_records + (size_t)p0 * frec + 4, frec, pcb, pq_nsq,
                                         records_out + (size_t)p0 * prec, s));
        } else {
            PK_CHECK(hipMemcpyAsync(xs, x + (size_t)p0 * dim, 4ull * m * dim, hipMemcpyHostToDevice, s));
            PK_CHECK(hipMemcpyAsync(cs, cell + p0, 4ull * m, hipMemcpyHostToDevice, s));
            PK_CHECK(hipMemcpyAsync(fr, filter_records + (size_t)p0 * frec, frec * m, hipMemcpyHostToDevice, s));
            PK_CHECK(pk_launch_encode_pq(xs, cs, m, cd, fd, F, dim, fcb, filter_nsq, fr + 4, frec, pcb, pq_nsq, ro, s));
            PK_CHECK(hipMemcpyAsync(records_out + (size_t)p0 * prec, ro, prec * m, hipMemcpyDeviceToHost, s));
            PK_CHECK(hipStreamSynchronize(s));
        }
    }
    PK_CHECK(hipStreamSynchronize(s));
    return PUCKGPU_OK;
}

int puckgpu_encode_filter(puckgpu_ctx* ctx, const float* coarse, const float* fine, uint32_t C, uint32_t F,
                          uint32_t dim, const float* codebook, uint32_t nsq, uint32_t n, const float* x,
                          const uint32_t* cell, uint8_t* records_out, int device_ptrs) {
    if (!ctx || !coarse || !fine || !codebook || !records_out || (n && (!x || !cell)))
        return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    if (nsq == 0 || dim % nsq || dim / nsq > 4) return set_err(ctx, PUCKGPU_EINVAL, "dim / nsq must be <= 4");
    if (n == 0) return PUCKGPU_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    hipStream_t s = ctx->stream;
    int rc = 0;
    if (C == 0 || F == 0 || (uint64_t)C * F >= 0xffffffffull) return set_err(ctx, PUCKGPU_EINVAL, "bad C/F");
    if (!device_ptrs)
        for (uint32_t i = 0; i < n; ++i)
            if (cell[i] >= C * F) return set_err(ctx, PUCKGPU_EINVAL, "cell[%u] = %u out of range", i, cell[i]);
    const float* cd = (const float*)dev_in(ctx, S_A0, coarse, 4ull * C * dim, device_ptrs, &rc);
    const float* fd = (const float*)dev_in(ctx, S_A1, fine, 4ull * F * dim, device_ptrs, &rc);
    const float* cbd = (const float*)dev_in(ctx, S_A2, codebook, 4ull * nsq * PK_KS * (dim / nsq), device_ptrs, &rc);
    if (rc) return rc;
    const uint64_t rec = 4 + nsq;
    const uint32_t CHK = 1u << 20;
    const uint32_t cbn = n < CHK ? n : CHK;
    float* xs = device_ptrs ? nullptr : (float*)scratch_slot(ctx, S_A3, 4ull * cbn * dim, &rc);
    uint32_t* cs = device_ptrs ? nullptr : (uint32_t*)scratch_slot(ctx, S_A4, 4ull * cbn, &rc);
    uint8_t* ro = device_ptrs ? nullptr : (uint8_t*)scratch_slot(ctx, S_A5, rec * cbn, &rc);
    if (rc) return rc;
    for (uint32_t p0 = 0; p0 < n; p0 += CHK) {
        const uint32_t m = (n - p0) < CHK ? (n - p0) : CHK;
        if (device_ptrs) {
            PK_CHECK(pk_launch_encode_filter(x + (size_t)p0 * dim, cell + p0, m, cd, fd, F, dim, cbd, nsq, nullptr,
                                             records_out + (size_t)p0 * rec, rec, 4, s));
        } else {
            PK_CHECK(hipMemcpyAsync(xs, x + (size_t)p0 * dim, 4ull * m * dim, hipMemcpyHostToDevice, s));
            PK_CHECK(hipMemcpyAsync(cs, cell + p0, 4ull * m, hipMemcpyHostToDevice, s));
            PK_CHECK(pk_launch_encode_filter(xs, cs, m, cd, fd, F, dim, cbd, nsq, nullptr, ro, rec, 4, s));
            PK_CHECK(hipMemcpyAsync(records_out + (size_t)p0 * rec, ro, rec * m, hipMemcpyDeviceToHost, s));
            PK_CHECK(hipStreamSynchronize(s));
        }
    }
    PK_CHECK(hipStreamSynchronize(s));
    return PUCKGPU_OK;
}

int puckgpu_kmeans_assign(puckgpu_ctx* ctx, uint32_t dim, const float* centroids, uint32_t k, uint32_t n,
                          const float* x, int32_t* assign_out, float* dist_out, int device_ptrs) {
    if (!ctx || !centroids || !assign_out || (n && !x)) return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    if (k == 0 || dim == 0) return set_err(ctx, PUCKGPU_EINVAL, "empty geometry");
    if (n == 0) return PUCKGPU_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    hipStream_t s = ctx->stream;
    int rc = 0;
    const float* cd = (const float*)dev_in(ctx, S_A0, centroids, 4ull * k * dim, device_ptrs, &rc);
    if (rc) return rc;
    if (dim == 1 || dim == 2 || dim == 4) {
        // sub-vector k-means (filter-PQ codebooks): one sdot per pair, in
        // MKL's order for such lengths, instead of a GEMM with K <= 4
        const uint32_t CHS = 1u << 20;
        float* xs = device_ptrs ? nullptr : (float*)scratch_slot(ctx, S_A4, 4ull * (n < CHS ? n : CHS) * dim, &rc);
        int32_t* ao = device_ptrs ? nullptr : (int32_t*)scratch_slot(ctx, S_A5, 4ull * (n < CHS ? n : CHS), &rc);
        float* dd = device_ptrs ? nullptr : (float*)scratch_slot(ctx, S_A6, 4ull * (n < CHS ? n : CHS), &rc);
        if (rc) return rc;
        for (uint32_t p0 = 0; p0 < n; p0 += CHS) {
            const uint32_t m = (n - p0) < CHS ? (n - p0) : CHS;
            const float* xp = x + (size_t)p0 * dim;
            if (!device_ptrs) {
                PK_CHECK(hipMemcpyAsync(xs, xp, 4ull * m * dim, hipMemcpyHostToDevice, s));
                xp = xs;
            }
            int32_t* a = device_ptrs ? assign_out + p0 : ao;
            float* dv = device_ptrs ? (dist_out ? dist_out + p0 : nullptr) : dd;
            PK_CHECK(pk_launch_kmeans_short(xp, m, dim, cd, k, a, dv, s));
            if (!device_ptrs) {
                PK_CHECK(hipMemcpyAsync(assign_out + p0, ao, 4ull * m, hipMemcpyDeviceToHost, s));
                if (dist_out) PK_CHECK(hipMemcpyAsync(dist_out + p0, dd, 4ull * m, hipMemcpyDeviceToHost, s));
                PK_CHECK(hipStreamSynchronize(s));
            }
        }
        PK_CHECK(hipStreamSynchronize(s));
        return PUCKGPU_OK;
    }
    float* cn = (float*)scratch_slot(ctx, S_A1, 4ull * k, &rc);
    if (rc) return rc;
    PK_CHECK(pk_launch_row_norms(cd, k, dim, 0, cn, s));
    uint32_t CHK = (uint32_t)std::max<uint64_t>(1024, std::min<uint64_t>(1u << 20, (256ull << 20) / (4ull * k)));
    const uint32_t cb = n < CHK ? n : CHK;
    float* ip = (float*)scratch_slot(ctx, S_A2, 4ull * cb * k, &rc);
    float* pn = (float*)scratch_slot(ctx, S_A3, 4ull * cb, &rc);
    float* xs = device_ptrs ? nullptr : (float*)scratch_slot(ctx, S_A4, 4ull * cb * dim, &rc);
    int32_t* ao = device_ptrs ? nullptr : (int32_t*)scratch_slot(ctx, S_A5, 4ull * cb, &rc);
    float* dd = device_ptrs ? nullptr : (float*)scratch_slot(ctx, S_A6, 4ull * cb, &rc);
    if (rc) return rc;
    for (uint32_t p0 = 0; p0 < n; p0 += CHK) {
        const uint32_t m = (n - p0) < CHK ? (n - p0) : CHK;
        const float* xp = x + (size_t)p0 * dim;
        if (!device_ptrs) {
            PK_CHECK(hipMemcpyAsync(xs, xp, 4ull * m * dim, hipMemcpyHostToDevice, s));
            xp = xs;
        }
        PK_CHECK(pk_launch_row_norms(xp, m, dim, 0, pn, s));
        PK_CHECK(pk_launch_gemm_exact(xp, cd, ip, m, k, dim, k, PK_EPI_RAW, nullptr, s));
        int32_t* a = device_ptrs ? assign_out + p0 : ao;
        float* dv = device_ptrs ? (dist_out ? dist_out + p0 : nullptr) : dd;
        PK_CHECK(pk_launch_kmeans_argmin(ip, pn, cn, m, k, a, dv, s));
        if (!device_ptrs) {
            PK_CHECK(hipMemcpyAsync(assign_out + p0, ao, 4ull * m, hipMemcpyDeviceToHost, s));
            if (dist_out) PK_CHECK(hipMemcpyAsync(dist_out + p0, dd, 4ull * m, hipMemcpyDeviceToHost, s));
            PK_CHECK(hipStreamSynchronize(s));
        }
    }
    PK_CHECK(hipStreamSynchronize(s));
    return PUCKGPU_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Realtime insert: RealtimeInsertPuckIndex::insert
// (realtime_insert_puck_index.cpp:273-345) for a batch, merged into the
// memory layout (each cell: loaded points, then inserted points newest first).
namespace {

// first insert: loaded points per cell, host counts, fine bound
int insert_setup(puckgpu_ctx* ctx, puckgpu_index* ix) {
    if (ix->nbase) return PUCKGPU_OK;
    hipStream_t s = ctx->stream;
    const uint32_t F = ix->p.fine_cluster_count, cells = ix->p.coarse_cluster_count * F;
    int rc = 0;
    if ((rc = dalloc(ctx, &ix->nbase, cells)) || (rc = dalloc(ctx, &ix->fbound, F))) return rc;
    PK_CHECK(pk_launch_cell_sizes(ix->cell_start, cells, ix->nbase, s));
    PK_CHECK(pk_launch_fine_bound(ix->fine, F, ix->p.dim, ix->fbound, s));
    ix->h_cnt.assign(cells, 0u);
    PK_CHECK(hipMemcpyAsync(ix->h_cnt.data(), ix->nbase, 4ull * cells, hipMemcpyDeviceToHost, s));
    PK_CHECK(hipStreamSynchronize(s));
    return fine_bound_sorted(ctx, ix->fbound, F, &ix->fb_sorted);
}

// out (host or device) <- device src
int copy_out(puckgpu_ctx* ctx, void* out, const void* src, size_t bytes, int device_ptrs) {
    if (!out || bytes == 0) return PUCKGPU_OK;
    PK_CHECK(hipMemcpyAsync(out, src, bytes, device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                            ctx->stream));
    return PUCKGPU_OK;
}

}  // namespace

extern "C" {

int puckgpu_index_insert(puckgpu_ctx* ctx, puckgpu_index* ix, uint32_t n, const float* features, uint32_t* ids_out,
                         uint32_t* cells_out, uint8_t* filter_records_out, uint8_t* pq_records_out,
                         float* features_out, int device_ptrs) {
    if (!ctx || !ix || (n && !features)) return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    if (ix->raw8) return set_err(ctx, PUCKGPU_EINVAL, "inserts into an index with uint8 raw vectors are not supported");
    if (!ix->raw && !ix->pq_codes)
        return set_err(ctx, PUCKGPU_EINVAL, "the index holds neither raw vectors nor a second-stage PQ");
    if (n == 0) return PUCKGPU_OK;
    const uint32_t d = ix->p.dim, C = ix->p.coarse_cluster_count, F = ix->p.fine_cluster_count, nsq = ix->p.filter_nsq,
                   pnsq = ix->pnsq, N0 = ix->p.total_point_count, cells = C * F;
    if ((uint64_t)N0 + n >= 0xffffffffull) return set_err(ctx, PUCKGPU_EINVAL, "index would exceed 2^32 - 1 points");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    hipStream_t s = ctx->stream;
    int rc = 0;
    if ((rc = insert_setup(ctx, ix))) return rc;
    const size_t rec = 4 + nsq, prec = 4 + pnsq;

    // single_build (hierarchical_cluster_index.cpp:1455-1494): normalisation;
    // a zero-norm point is refused and takes no id
    const float* xd = (const float*)dev_in(ctx, S_INS_X, features, 4ull * n * d, device_ptrs, &rc);
    float* y = (float*)scratch_slot(ctx, S_INS_Y, 4ull * n * d, &rc);
    float* yn = (float*)scratch_slot(ctx, S_QN, 4ull * n, &rc);
    int32_t* st = (int32_t*)scratch_slot(ctx, S_QST, 4ull * n, &rc);
    if (rc) return rc;
    PK_CHECK(pk_launch_prep_queries(xd, n, d, ix->p.whether_norm, y, yn, st, s));
    std::vector<int32_t> st_h(n);
    PK_CHECK(hipMemcpyAsync(st_h.data(), st, 4ull * n, hipMemcpyDeviceToHost, s));
    PK_CHECK(hipStreamSynchronize(s));
    std::vector<uint32_t> acc;
    for (uint32_t i = 0; i < n; ++i)
        if (st_h[i] == 0) acc.push_back(i);
    const uint32_t m = (uint32_t)acc.size();
    std::vector<uint32_t> ids(n, 0xffffffffu);
    for (uint32_t a = 0; a < m; ++a) ids[acc[a]] = N0 + a;
    float* ya = y;
    if (m < n && m > 0) {
        uint32_t* accd = (uint32_t*)scratch_slot(ctx, S_INS_ACC, 4ull * m, &rc);
        ya = (float*)scratch_slot(ctx, S_INS_X, 4ull * m * d, &rc);  // the input copy (if any) is no longer read
        if (rc) return rc;
        PK_CHECK(hipMemcpyAsync(accd, acc.data(), 4ull * m, hipMemcpyHostToDevice, s));
        PK_CHECK(pk_launch_gather_f32(y, d, ya, accd, m, s));
        PK_CHECK(hipStreamSynchronize(s));
    }
    if (m > 0) {
        // nearest_cell_assign over the live stationary values, one point after
        // another as far as it matters: a point that lands in an empty cell
        // gives that cell its stationary value (:326-338) before the later
        // points are assigned again
        uint32_t* cell = (uint32_t*)scratch_slot(ctx, S_INS_CELL, 4ull * m, &rc);
        if (rc) return rc;
        std::vector<uint32_t> hcell(m);
        uint32_t done = 0;
        while (done < m) {
            for (uint32_t p0 = done; p0 < m; p0 += PK_ASSIGN_CHUNK) {
                const uint32_t k = (m - p0) < PK_ASSIGN_CHUNK ? (m - p0) : PK_ASSIGN_CHUNK;
                if ((rc = assign_chunk(ctx, ix->coarse, ix->fine, ix->coarse_norms, ix->stationary, ix->fbound,
                                       ix->fb_sorted, C, F, d, k, ya + (size_t)p0 * d, cell + p0, nullptr, s)))
                    return rc;
            }
            PK_CHECK(hipMemcpyAsync(hcell.data() + done, cell + done, 4ull * (m - done), hipMemcpyDeviceToHost, s));
            PK_CHECK(hipStreamSynchronize(s));
            uint32_t next = m;
            for (uint32_t i = done; i < m; ++i) {
                if (hcell[i] >= cells) return set_err(ctx, PUCKGPU_EHIP, "assignment returned cell %u", hcell[i]);
                if (ix->h_cnt[hcell[i]]++ == 0) {
                    PK_CHECK(pk_launch_insert_stationary(ix->coarse, ix->fine, F, d, cell + i, 1, ix->stationary, s));
                    next = i + 1;
                    break;
                }
            }
            done = next;
        }
        // puck_single_assign (puck_index.cpp:938-988): filter codes, then the
        // second-stage PQ codes on what the filter codewords leave
        uint8_t* recs = (uint8_t*)scratch_slot(ctx, S_INS_REC, rec * m, &rc);
        uint8_t* precs = pnsq ? (uint8_t*)scratch_slot(ctx, S_INS_PQ, prec * m, &rc) : nullptr;
        if (rc) return rc;
        PK_CHECK(pk_launch_encode_filter(ya, cell, m, ix->coarse, ix->fine, F, d, ix->codebook, nsq, nullptr, recs, rec,
                                         4, s));
        if (pnsq)
            PK_CHECK(pk_launch_encode_pq(ya, cell, m, ix->coarse, ix->fine, F, d, ix->codebook, nsq, recs + 4, rec,
                                         ix->pq_codebook, pnsq, precs, s));
        // rank of each point among the batch's points of its cell, newest first
        std::vector<uint32_t> rank(m);
        {
            std::unordered_map<uint32_t, uint32_t> seen;
            for (uint32_t i = m; i-- > 0;) rank[i] = seen[hcell[i]]++;
        }
        uint32_t* rank_d = (uint32_t*)scratch_slot(ctx, S_INS_RANK, 4ull * m, &rc);
        uint32_t* add = (uint32_t*)scratch_slot(ctx, S_INS_ADD, 4ull * (cells + 1), &rc);
        uint32_t* shift = (uint32_t*)scratch_slot(ctx, S_INS_SHIFT, 4ull * (cells + 1), &rc);
        uint32_t* work = (uint32_t*)scratch_slot(ctx, S_INS_WORK, 4ull * pk_exscan_work(cells), &rc);
        uint32_t* dold = (uint32_t*)scratch_slot(ctx, S_INS_DOLD, 4ull * (N0 ? N0 : 1), &rc);
        uint32_t* dnew = (uint32_t*)scratch_slot(ctx, S_INS_DNEW, 4ull * m, &rc);
        if (rc) return rc;
        PK_CHECK(hipMemcpyAsync(rank_d, rank.data(), 4ull * m, hipMemcpyHostToDevice, s));
        PK_CHECK(pk_launch_insert_prepare(cell, m, cells, add, shift, work, s));
        // the merged arrays
        const uint32_t N1 = N0 + m;
        uint32_t *cs1 = nullptr, *m2l1 = nullptr;
        float *off1 = nullptr, *raw1 = nullptr, *poff1 = nullptr;
        uint8_t *codes1 = nullptr, *pcodes1 = nullptr;
        auto release = [&]() {
            for (void* q : {(void*)cs1, (void*)m2l1, (void*)off1, (void*)raw1, (void*)poff1, (void*)codes1,
                            (void*)pcodes1})
                if (q) (void)hipFree(q);
        };
        if ((rc = dalloc(ctx, &cs1, (size_t)cells + 1)) || (rc = dalloc(ctx, &m2l1, N1)) ||
            (rc = dalloc(ctx, &off1, N1)) || (rc = dalloc(ctx, &codes1, (size_t)N1 * nsq)) ||
            (ix->raw && (rc = dalloc(ctx, &raw1, (size_t)N1 * d))) ||
            (pnsq && ((rc = dalloc(ctx, &poff1, N1)) || (rc = dalloc(ctx, &pcodes1, (size_t)N1 * pnsq))))) {
            release();
            return rc;
        }
        hipError_t e = pk_launch_insert_dest(ix->cell_start, cells, ix->nbase, add, shift, N0, dold, cell, rank_d, m,
                                             dnew, m2l1, cs1, s);
        auto mv = [&](const void* src, uint64_t sstr, uint64_t soff, void* dst, uint64_t dstr, uint32_t bytes,
                      const uint32_t* dest, uint32_t cnt) {
            if (e == hipSuccess) e = pk_launch_scatter_rows(src, sstr, soff, dst, dstr, bytes, dest, cnt, s);
        };
        mv(ix->mem2local, 4, 0, m2l1, 4, 4, dold, N0);
        mv(ix->rec_off, 4, 0, off1, 4, 4, dold, N0);
        mv(ix->rec_codes, nsq, 0, codes1, nsq, nsq, dold, N0);
        mv(recs, rec, 0, off1, 4, 4, dnew, m);
        mv(recs, rec, 4, codes1, nsq, nsq, dnew, m);
        if (ix->raw) {
            mv(ix->raw, 4ull * d, 0, raw1, 4ull * d, 4 * d, dold, N0);
            mv(ya, 4ull * d, 0, raw1, 4ull * d, 4 * d, dnew, m);
        }
        if (pnsq) {
            mv(ix->pq_off, 4, 0, poff1, 4, 4, dold, N0);
            mv(ix->pq_codes, pnsq, 0, pcodes1, pnsq, pnsq, dold, N0);
            mv(precs, prec, 0, poff1, 4, 4, dnew, m);
            mv(precs, prec, 4, pcodes1, pnsq, pnsq, dnew, m);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            release();
            return set_err(ctx, PUCKGPU_EHIP, "merging inserted points failed: %s", hipGetErrorString(e));
        }
        for (void* q : {(void*)ix->cell_start, (void*)ix->mem2local, (void*)ix->rec_off, (void*)ix->rec_codes,
                        (void*)ix->raw, (void*)ix->pq_off, (void*)ix->pq_codes})
            if (q) (void)hipFree(q);
        ix->cell_start = cs1;
        ix->mem2local = m2l1;
        ix->rec_off = off1;
        ix->rec_codes = codes1;
        if (ix->raw) ix->raw = raw1;
        if (pnsq) {
            ix->pq_off = poff1;
            ix->pq_codes = pcodes1;
        }
        ix->p.total_point_count = N1;
        ix->version++;
        // outputs of the accepted points, in order
        if ((rc = copy_out(ctx, cells_out, cell, 4ull * m, device_ptrs)) ||
            (rc = copy_out(ctx, filter_records_out, recs, rec * m, device_ptrs)) ||
            (pnsq && (rc = copy_out(ctx, pq_records_out, precs, prec * m, device_ptrs))) ||
            (rc = copy_out(ctx, features_out, ya, 4ull * m * d, device_ptrs)))
            return rc;
    }
    if (ids_out) {
        if (device_ptrs) PK_CHECK(hipMemcpyAsync(ids_out, ids.data(), 4ull * n, hipMemcpyHostToDevice, s));
        else std::memcpy(ids_out, ids.data(), 4ull * n);
    }
    PK_CHECK(hipStreamSynchronize(s));
    return PUCKGPU_OK;
}

uint64_t puckgpu_index_version(const puckgpu_index* ix) { return ix ? ix->version : 0; }

}  // extern "C"

// ---------------------------------------------------------------------------
// k-means on the device: Kmeans::kmeans / kmeans_core (kmeans.cpp:149-329) and
// HierarchicalClusterIndex::train (hierarchical_cluster_index.cpp:882-985).
// The assignments, centroid updates and residuals run on the GPU; the
// generator draws (seeding, the empty-cluster split) and the error sum run
// on the host, in the reference's order.
namespace {

// The reference's Kmeans object: one std::mt19937 (seeded by the caller
// instead of time(0), kmeans.h:47-48), the Lloyd iteration limit, the
// seeding and the centroid-update mode.
struct KmState {
    std::mt19937 rnd;
    int niter = 30;
    bool pp = true;
    int mean_mode = 0;
    explicit KmState(uint32_t seed) : rnd(seed) {}
};

// roulette_selection (kmeans.cpp:48-67): wheel scaled by (float)(1.0 / sum)
// in place, then the first entry whose running subtraction from a uniform
// double goes below zero.
int km_roulette(KmState& st, std::vector<float>& wheel) {
    float total = 0;
    for (float v : wheel) total += v;
    const float a = (float)(1.0 / total);
    for (float& v : wheel) v = a * v;
    std::uniform_real_distribution<double> dis(0, 1.0);
    double rd = dis(st.rnd);
    for (int id = 0; id < (int)wheel.size(); ++id) {
        rd -= wheel[id];
        if (rd < 0) return id;
    }
    return (int)wheel.size() - 1;
}

// Sequential FMA dot product on the host (the engine's order).
float km_dot(const float* a, const float* b, uint32_t n) {
    float acc = 0.0f;
    for (uint32_t i = 0; i < n; ++i) acc = std::fmaf(a[i], b[i], acc);
    return acc;
}

// cblas_sdot's order on the host: MKL's for 1, 2 or 4 floats (rounded
// products, pairwise sums), km_dot otherwise (pk_sdot_short's order).
float km_sdot(const float* a, const float* b, uint32_t n) {
    if (n == 1) return a[0] * b[0];
    if (n == 2) {
        const float p0 = a[0] * b[0], p1 = a[1] * b[1];
        return p0 + p1;
    }
    if (n == 4) {
        const float p0 = a[0] * b[0], p1 = a[1] * b[1], p2 = a[2] * b[2], p3 = a[3] * b[3];
        const float s01 = p0 + p1, s23 = p2 + p3;
        return s01 + s23;
    }
    return km_dot(a, b, n);
}

// nearest_center (kmeans.cpp:221-258) of n device points against k device
// centroids: assign / dis on the device.
int km_assign(puckgpu_ctx* ctx, uint32_t dim, const float* cent, uint32_t k, uint32_t n, const float* x,
              int32_t* assign, float* dis, hipStream_t s) {
    if (n == 0) return PUCKGPU_OK;
    if (dim == 1 || dim == 2 || dim == 4) {
        PK_CHECK(pk_launch_kmeans_short(x, n, dim, cent, k, assign, dis, s));
        return PUCKGPU_OK;
    }
    int rc = 0;
    const uint32_t CHK = (uint32_t)std::max<uint64_t>(1024, std::min<uint64_t>(1u << 20, (256ull << 20) / (4ull * k)));
    const uint32_t cb = n < CHK ? n : CHK;
    float* ip = (float*)scratch_slot(ctx, S_KM_IP, 4ull * cb * k, &rc);
    float* pn = (float*)scratch_slot(ctx, S_KM_PN, 4ull * cb, &rc);
    float* cn = (float*)scratch_slot(ctx, S_KM_CN, 4ull * k, &rc);
    if (rc) return rc;
    PK_CHECK(pk_launch_row_norms(cent, k, dim, 0, cn, s));
    for (uint32_t p0 = 0; p0 < n; p0 += CHK) {
        const uint32_t m = (n - p0) < CHK ? (n - p0) : CHK;
        const float* xp = x + (size_t)p0 * dim;
        PK_CHECK(pk_launch_row_norms(xp, m, dim, 0, pn, s));
        PK_CHECK(pk_launch_gemm_exact(xp, cent, ip, m, k, dim, k, PK_EPI_RAW, nullptr, s));
        PK_CHECK(pk_launch_kmeans_argmin(ip, pn, cn, m, k, assign + p0, dis + p0, s));
    }
    return PUCKGPU_OK;
}

// kmeans_reassign_empty (kmeans.cpp:112-147) on host centroids: every call
// draws the noise scale; each empty cluster copies a donor picked by
// roulette over count^2 - 1 and the pair moves apart by a small Gaussian step.
int km_split(KmState& st, uint32_t dim, uint32_t k, float* cent, const std::vector<uint32_t>& cnt) {
    std::vector<float> proba(k), eps(dim);
    std::normal_distribution<> dn(0, st.rnd() / ((double)RAND_MAX + 1.0));
    for (uint32_t c = 0; c < k; ++c)  // int arithmetic, wrapping as the reference's does
        proba[c] = cnt[c] < 2 ? 0.f : (float)(int32_t)(uint32_t)((int64_t)cnt[c] * cnt[c] - 1);
    int nre = 0;
    for (uint32_t c = 0; c < k; ++c) {
        if (cnt[c] != 0) continue;
        nre++;
        const int j = km_roulette(st, proba);
        std::memcpy(cent + (size_t)c * dim, cent + (size_t)j * dim, sizeof(float) * dim);
        float* cj = cent + (size_t)j * dim;
        const double sc = std::sqrt(km_dot(cj, cj, dim)) * 0.0000001;
        for (float& v : eps) v = (float)dn(st.rnd);
        const float a = (float)sc;
        for (float& v : eps) v = a * v;
        for (uint32_t t = 0; t < dim; ++t) cj[t] = cj[t] + eps[t];
        for (uint32_t t = 0; t < dim; ++t) cent[(size_t)c * dim + t] = cent[(size_t)c * dim + t] - eps[t];
        proba[j] = 0;
    }
    return nre;
}

// kmeans_core (kmeans.cpp:260-329) on device centroids `cent` [k x dim]
// (updated in place) and points x [n x dim]; assign / dis [n] on the device.
int km_core(puckgpu_ctx* ctx, KmState& st, uint32_t d, uint32_t n, uint32_t k, int nit, float* cent,
            const float* x, int32_t* assign, float* dis, double* qerr_out, int* iters, int* split, hipStream_t s) {
    int rc = 0;
    const uint64_t hw = pk_ks_hist_words(n, k);
    uint32_t* perm = (uint32_t*)scratch_slot(ctx, S_KM_PERM, 4ull * (n ? n : 1), &rc);
    uint32_t* starts = (uint32_t*)scratch_slot(ctx, S_KM_STARTS, 4ull * (k + 1), &rc);
    uint32_t* hist = (uint32_t*)scratch_slot(ctx, S_KM_HIST, 4ull * (hw ? hw : 1), &rc);
    uint32_t* off = (uint32_t*)scratch_slot(ctx, S_KM_OFF, 4ull * (hw + 1), &rc);
    uint32_t* work = (uint32_t*)scratch_slot(ctx, S_KM_WORK, 4ull * pk_exscan_work(hw), &rc);
    if (rc) return rc;
    std::vector<uint32_t> st_h(k + 1), cnt(k);
    std::vector<float> dis_h(n), cent_h;
    double qerr = std::numeric_limits<double>::max(), qerr_old;
    int tot = 0, it = 0;
    for (int iter = 1; iter <= nit; ++iter) {
        ++it;
        if ((rc = km_assign(ctx, d, cent, k, n, x, assign, dis, s))) return rc;
        PK_CHECK(pk_launch_sort_by_cluster(assign, n, k, hist, off, work, perm, starts, s));
        PK_CHECK(pk_launch_kmeans_update(x, d, perm, starts, k, cent, st.mean_mode, s));
        PK_CHECK(hipMemcpyAsync(st_h.data(), starts, 4ull * (k + 1), hipMemcpyDeviceToHost, s));
        PK_CHECK(hipMemcpyAsync(dis_h.data(), dis, 4ull * n, hipMemcpyDeviceToHost, s));
        PK_CHECK(hipStreamSynchronize(s));
        bool empty = false;
        for (uint32_t c = 0; c < k; ++c) {
            cnt[c] = st_h[c + 1] - st_h[c];
            empty = empty || cnt[c] == 0;
        }
        int nre;
        if (empty) {
            cent_h.resize((size_t)k * d);
            PK_CHECK(hipMemcpy(cent_h.data(), cent, 4ull * k * d, hipMemcpyDeviceToHost));
            nre = km_split(st, d, k, cent_h.data(), cnt);
            PK_CHECK(hipMemcpy(cent, cent_h.data(), 4ull * k * d, hipMemcpyHostToDevice));
        } else {
            nre = km_split(st, d, k, nullptr, cnt);  // draws the noise scale only
        }
        tot += nre;
        if (tot > (int)(n / 100) && tot > 1000)
            return set_err(ctx, PUCKGPU_EINVAL, "kmeans: reassigned %d times, abandoning", tot);
        qerr_old = qerr;
        qerr = 0;
        for (uint32_t i = 0; i < n; ++i) qerr += dis_h[i];
        if (std::fabs(qerr_old - qerr) < 1e-6 && nre == 0) break;
    }
    *qerr_out = qerr;
    if (iters) *iters = it;
    if (split) *split = tot;
    return PUCKGPU_OK;
}

// Kmeans::kmeans (kmeans.cpp:149-219), redo = 1: seeding (k-means++ over the
// first min(n, 8k) points when n > 8k and n > 8192, kmeans.cpp:173-181; else
// k distinct uniform draws), then kmeans_core.  cent_out [k x dim] and
// assign [n] on the device; err = qerr / n.
int km_run(puckgpu_ctx* ctx, KmState& st, uint32_t d, uint32_t n, uint32_t k, const float* x, float* cent_out,
           int32_t* assign, float* err, int* iters, int* split, hipStream_t s) {
    if (k == 0 || k >= n) return set_err(ctx, PUCKGPU_EINVAL, "k-means needs fewer clusters (%u) than points (%u)", k, n);
    int rc = 0;
    std::vector<uint32_t> sel(k);
    if (st.pp) {
        const uint32_t nsub = ((uint64_t)n > 8ull * k && n > 8u * 1024u) ? 8u * k : n;
        float* pn = (float*)scratch_slot(ctx, S_KM_PN, 4ull * nsub, &rc);
        float* best = (float*)scratch_slot(ctx, S_KM_BEST, 4ull * nsub, &rc);
        if (rc) return rc;
        std::vector<float> best_h(nsub, std::numeric_limits<float>::max()), tmp;
        PK_CHECK(hipMemcpyAsync(best, best_h.data(), 4ull * nsub, hipMemcpyHostToDevice, s));
        PK_CHECK(pk_launch_norms_sdot(x, nsub, d, pn, s));
        sel[0] = st.rnd() % nsub;
        for (uint32_t i = 1; i < k; ++i) {
            PK_CHECK(pk_launch_kmeanspp_step(x, pn, nsub, d, sel[i - 1], best, s));
            PK_CHECK(hipMemcpyAsync(best_h.data(), best, 4ull * nsub, hipMemcpyDeviceToHost, s));
            PK_CHECK(hipStreamSynchronize(s));
            tmp = best_h;
            sel[i] = (uint32_t)km_roulette(st, tmp);
        }
    } else {
        // random_init_center (kmeans.cpp:30-46)
        std::uniform_int_distribution<> dis(0, (int)n - 1);
        std::vector<bool> taken(n, false);
        for (uint32_t got = 0; got < k;) {
            const uint32_t v = (uint32_t)dis(st.rnd);
            if (taken[v]) continue;
            taken[v] = true;
            sel[got++] = v;
        }
    }
    uint32_t* ids = (uint32_t*)scratch_slot(ctx, S_KM_IDS, 4ull * k, &rc);
    float* dis = (float*)scratch_slot(ctx, S_KM_DIS, 4ull * n, &rc);
    if (rc) return rc;
    PK_CHECK(hipMemcpyAsync(ids, sel.data(), 4ull * k, hipMemcpyHostToDevice, s));
    PK_CHECK(pk_launch_gather_f32(x, d, cent_out, ids, k, s));
    double qerr = 0;
    if ((rc = km_core(ctx, st, d, n, k, st.niter, cent_out, x, assign, dis, &qerr, iters, split, s))) return rc;
    *err = (float)(qerr / n);
    return PUCKGPU_OK;
}

// MaxHeap (max_heap.cpp:28-102) over host arrays: sentinel-filled, inserts
// into a non-full heap start at its last free slot, reorder() sorts ascending.
struct KmHeap {
    std::vector<float> v;
    std::vector<uint32_t> t;
    uint32_t size, free_slots;
    explicit KmHeap(uint32_t n) : v(n + 1), t(n + 1, 0), size(n), free_slots(n) {
        float sent;
        const uint32_t b = 0x7f7f7f7fu;
        std::memcpy(&sent, &b, 4);
        std::fill(v.begin(), v.end(), sent);
    }
    void sift(uint32_t hs, uint32_t f, float val, uint32_t tag) {
        for (;;) {
            const uint32_t l = f << 1, r = l + 1;
            if (l > hs) break;
            const uint32_t c = (r > hs || v[l] > v[r]) ? l : r;
            if (v[c] < val) break;
            v[f] = v[c];
            t[f] = t[c];
            f = c;
        }
        v[f] = val;
        t[f] = tag;
    }
    void update(float val, uint32_t tag) {
        if (val >= v[1]) return;
        const uint32_t f = free_slots > 0 ? free_slots-- : 1u;
        sift(size, f, val, tag);
    }
    void reorder() {
        for (uint32_t i = 0; i < size; ++i) {
            const float vb = v[1];
            const uint32_t tb = t[1];
            const uint32_t hs = size - i;
            sift(hs - 1, 1, v[hs], t[hs]);
            v[hs] = vb;
            t[hs] = tb;
        }
    }
};

}  // namespace

extern "C" {

int puckgpu_kmeans(puckgpu_ctx* ctx, uint32_t dim, uint32_t n, uint32_t k, const float* x, int niter, int init_pp,
                   uint32_t seed, int mean_mode, float* centroids_out, int32_t* assign_out, float* err_out,
                   int* iters_out, int* split_out, int device_ptrs) {
    if (!ctx || !x || !centroids_out) return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    if (dim == 0 || niter < 0) return set_err(ctx, PUCKGPU_EINVAL, "bad dim or niter");
    if (k > 16384) return set_err(ctx, PUCKGPU_EINVAL, "k must be <= 16384");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    hipStream_t s = ctx->stream;
    int rc = 0;
    const float* xd = (const float*)dev_in(ctx, S_KM_X, x, 4ull * n * dim, device_ptrs, &rc);
    float* cd = device_ptrs ? centroids_out : (float*)scratch_slot(ctx, S_KM_CENT, 4ull * (k ? k : 1) * dim, &rc);
    int32_t* ad = (device_ptrs && assign_out) ? assign_out
                                              : (int32_t*)scratch_slot(ctx, S_KM_ASSIGN, 4ull * (n ? n : 1), &rc);
    if (rc) return rc;
    KmState st(seed);
    st.pp = init_pp != 0;
    st.mean_mode = mean_mode;
    st.niter = niter;
    float err = 0;
    if ((rc = km_run(ctx, st, dim, n, k, xd, cd, ad, &err, iters_out, split_out, s))) return rc;
    if (!device_ptrs) {
        PK_CHECK(hipMemcpyAsync(centroids_out, cd, 4ull * k * dim, hipMemcpyDeviceToHost, s));
        if (assign_out) PK_CHECK(hipMemcpyAsync(assign_out, ad, 4ull * n, hipMemcpyDeviceToHost, s));
    }
    PK_CHECK(hipStreamSynchronize(s));
    if (err_out) *err_out = err;
    return PUCKGPU_OK;
}

int puckgpu_train_centers(puckgpu_ctx* ctx, uint32_t dim, uint32_t n, const float* x, uint32_t C, uint32_t F,
                          int iterations, int niter, int init_pp, uint32_t seed, int mean_mode, float* coarse_out,
                          float* fine_out, float* err_trace, int* rounds_out, int* niter_out, int device_ptrs) {
    if (!ctx || !x || !coarse_out || !fine_out) return set_err(ctx, PUCKGPU_EINVAL, "null argument");
    if (dim == 0 || iterations < 0 || niter < 0) return set_err(ctx, PUCKGPU_EINVAL, "bad dim, iterations or niter");
    if (C > 16384 || F > 16384) return set_err(ctx, PUCKGPU_EINVAL, "C and F must be <= 16384");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, PUCKGPU_EHIP, "hipSetDevice failed");
    hipStream_t s = ctx->stream;
    int rc = 0;
    const size_t nd = (size_t)n * dim;
    const float* xd = (const float*)dev_in(ctx, S_KM_X, x, 4ull * nd, device_ptrs, &rc);
    float* tv = (float*)scratch_slot(ctx, S_KM_TV, 4ull * (nd ? nd : 1), &rc);
    float* c0 = (float*)scratch_slot(ctx, S_KM_C0, 4ull * (C ? C : 1) * dim, &rc);
    float* f0 = (float*)scratch_slot(ctx, S_KM_F0, 4ull * (F ? F : 1) * dim, &rc);
    float* cb = (float*)scratch_slot(ctx, S_KM_CB, 4ull * (C ? C : 1) * dim, &rc);
    float* fb = (float*)scratch_slot(ctx, S_KM_FB, 4ull * (F ? F : 1) * dim, &rc);
    int32_t* a = (int32_t*)scratch_slot(ctx, S_KM_ASSIGN, 4ull * (n ? n : 1), &rc);
    int32_t* fa = (int32_t*)scratch_slot(ctx, S_KM_FA, 4ull * (n ? n : 1), &rc);
    if (rc) return rc;
    PK_CHECK(hipMemsetAsync(cb, 0, 4ull * C * dim, s));
    PK_CHECK(hipMemsetAsync(fb, 0, 4ull * F * dim, s));
    PK_CHECK(hipMemsetAsync(fa, 0, 4ull * n, s));
    PK_CHECK(hipMemcpyAsync(tv, xd, 4ull * nd, hipMemcpyDeviceToDevice, s));
    KmState st(seed);
    st.pp = init_pp != 0;
    st.mean_mode = mean_mode;
    if (niter > 0) st.niter = niter;
    float min_err = std::numeric_limits<float>::max();
    int rounds = 0;
    for (int ite = 0; ite < iterations; ++ite) {
        ++rounds;
        float err = 0;
        // coarse centers S on the data less the best fine centers so far
        if ((rc = km_run(ctx, st, dim, n, C, tv, c0, a, &err, nullptr, nullptr, s))) return rc;
        if (err_trace) err_trace[2 * ite] = err;
        report(ctx, "train_centers round %d: coarse k-means (%u x %u, k=%u, Lloyd limit %d) error %g", ite + 1, n,
               dim, C, st.niter, (double)err);
        PK_CHECK(pk_launch_sub_assigned(xd, n, dim, c0, a, tv, s));
        // fine centers T on the data less the new coarse centers
        if ((rc = km_run(ctx, st, dim, n, F, tv, f0, a, &err, nullptr, nullptr, s))) return rc;
        if (err_trace) err_trace[2 * ite + 1] = err;
        report(ctx, "train_centers round %d: fine k-means (k=%u) error %g", ite + 1, F, (double)err);
        if ((min_err - err) >= 1e-4) {
            PK_CHECK(hipMemcpyAsync(cb, c0, 4ull * C * dim, hipMemcpyDeviceToDevice, s));
            PK_CHECK(hipMemcpyAsync(fb, f0, 4ull * F * dim, hipMemcpyDeviceToDevice, s));
            PK_CHECK(hipMemcpyAsync(fa, a, 4ull * n, hipMemcpyDeviceToDevice, s));
            min_err = err;
        } else {
            if (st.niter > 80) break;
            st.niter += 10;
        }
        PK_CHECK(pk_launch_sub_assigned(xd, n, dim, fb, fa, tv, s));
    }
    // fine centers by ascending -sqrt(|f|^2 / 2) (a MaxHeap sort, :958-980)
    std::vector<float> fh((size_t)F * dim), ch((size_t)C * dim);
    PK_CHECK(hipMemcpyAsync(fh.data(), fb, 4ull * F * dim, hipMemcpyDeviceToHost, s));
    PK_CHECK(hipMemcpyAsync(ch.data(), cb, 4ull * C * dim, hipMemcpyDeviceToHost, s));
    PK_CHECK(hipStreamSynchronize(s));
    KmHeap heap(F);
    for (uint32_t j = 0; j < F; ++j) {
        const float* f = fh.data() + (size_t)j * dim;
        const float h = km_sdot(f, f, dim) / 2;
        heap.update(0 - std::sqrt(h), j);
    }
    heap.reorder();
    std::vector<float> fo((size_t)F * dim);
    for (uint32_t j = 0; j < F; ++j)
        std::memcpy(fo.data() + (size_t)j * dim, fh.data() + (size_t)heap.t[j + 1] * dim, 4ull * dim);
    const hipMemcpyKind out = device_ptrs ? hipMemcpyHostToDevice : hipMemcpyHostToHost;
    PK_CHECK(hipMemcpy(coarse_out, ch.data(), 4ull * C * dim, out));
    PK_CHECK(hipMemcpy(fine_out, fo.data(), 4ull * F * dim, out));
    if (rounds_out) *rounds_out = rounds;
    if (niter_out) *niter_out = st.niter;
    return PUCKGPU_OK;
}

}  // extern "C"
