// sort.hip -- per-row ascending sort of center distances.
//
// One workgroup per row.  Keys are (order-preserving float bits << 32 | column)
// so equal distances keep column order; a bitonic network in LDS sorts the
// padded row and the first `keep` entries are written out.
//   * coarse selection: keep = search_coarse_count of C (the top-k max-heap of
//     hierarchical_cluster_index.cpp:709-726 keeps the smallest, earlier
//     column on ties at the boundary);
//   * fine order: keep = F (the full heap sort of puck_index.cpp:497-504).
#include "common.h"

#define SORT_THREADS 256

__device__ __forceinline__ void pk_bitonic_u64(uint64_t* keys, uint32_t n2) {
    for (uint32_t k = 2; k <= n2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < (n2 >> 1); i += blockDim.x) {
                uint32_t lo = 2 * i - (i & (j - 1));
                uint32_t hi = lo + j;
                bool up = (lo & k) == 0;
                uint64_t a = keys[lo], b = keys[hi];
                if ((a > b) == up) {
                    keys[lo] = b;
                    keys[hi] = a;
                }
            }
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(SORT_THREADS) void pk_sort_rows(const float* __restrict__ vals, uint32_t L,
                                                             uint64_t ld_in, uint32_t keep,
                                                             float* __restrict__ out_val,
                                                             uint32_t* __restrict__ out_idx, uint32_t n2,
                                                             uint8_t* __restrict__ flags, uint32_t flag_lim,
                                                             const uint32_t* __restrict__ skip_nf) {
    extern __shared__ uint64_t sk[];
    const uint64_t row = blockIdx.x;
    // rows of queries with a fine budget below L are not walked in fine order
    if (skip_nf && skip_nf[row] != 0u && skip_nf[row] < L) return;
    const float* v = vals + row * ld_in;
    for (uint32_t i = threadIdx.x; i < n2; i += blockDim.x)
        sk[i] = i < L ? (((uint64_t)pk_f2ord(v[i]) << 32) | i) : ~0ull;
    __syncthreads();
    pk_bitonic_u64(sk, n2);
    bool tie = false;
    for (uint32_t i = threadIdx.x; i < keep; i += blockDim.x) {
        uint64_t key = sk[i];
        float v = pk_ord2f((uint32_t)(key >> 32));
        out_val[row * keep + i] = v;
        out_idx[row * keep + i] = (uint32_t)(key & 0xffffffffu);
        if (i < flag_lim && i + 1 < L && pk_ord2f((uint32_t)(sk[i + 1] >> 32)) == v) tie = true;
    }
    // equal distances among the first flag_lim entries (or at that boundary):
    // the reference's heap order decides, flag the row for the exact replay
    if (tie && flags) flags[row] = 1u | flags[row];
}

extern "C" hipError_t pk_launch_sort_rows(const float* vals, uint32_t rows, uint32_t L, uint64_t ld_in,
                                          uint32_t keep, float* out_val, uint32_t* out_idx, uint8_t* flags,
                                          uint32_t flag_lim, const uint32_t* skip_nf, hipStream_t s) {
    if (rows == 0 || keep == 0) return hipSuccess;
    if (L > 8192 || keep > L) return hipErrorInvalidValue;
    static size_t allowed = 0;
    if (!allowed) {
        const hipError_t e = pk_allow_lds((const void*)pk_sort_rows, 64 * 1024, &allowed);
        if (e != hipSuccess) return e;
    }
    uint32_t n2 = 2;
    while (n2 < L) n2 <<= 1;
    if (n2 * sizeof(uint64_t) > allowed) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pk_sort_rows, dim3(rows), dim3(SORT_THREADS), n2 * sizeof(uint64_t), s, vals, L, ld_in,
                       keep, out_val, out_idx, n2, flags, flag_lim, skip_nf);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// k smallest entries of a row by (value, column), ascending, without sorting
// the row: a radix select finds the k-th smallest value kc; the entries below
// it and then the first ties at it (column order) are compacted into LDS and
// only those k are sorted.
#define TOPK_CAP 1024

template <typename ValF>
__device__ void pk_topk_block(ValF val, uint32_t L, uint32_t k, float* out_val, uint32_t* out_idx, bool* tie,
                              uint32_t flag_lim) {
    __shared__ uint32_t hist[258];
    __shared__ uint32_t tmp[8];
    __shared__ uint64_t buf[TOPK_CAP];
    auto key = [&](uint32_t i) { return pk_f2ord(val(i)); };
    const uint32_t kc = pk_select_rank(key, L, k, hist, tmp);
    uint32_t nlt = 0, neq = 0;
    for (uint32_t r0 = 0; r0 < L; r0 += blockDim.x) {  // entries below kc (fewer than k)
        const uint32_t i = r0 + threadIdx.x;
        const uint32_t kk = i < L ? key(i) : 0xffffffffu;
        const uint32_t f = (i < L && kk < kc) ? 1u : 0u;
        uint32_t tot;
        const uint32_t pre = pk_block_exscan(f, tmp, &tot);
        if (f) buf[nlt + pre] = ((uint64_t)kk << 32) | i;
        nlt += tot;
    }
    for (uint32_t r0 = 0; r0 < L; r0 += blockDim.x) {  // ties at kc, lowest columns first
        const uint32_t i = r0 + threadIdx.x;
        const uint32_t f = (i < L && key(i) == kc) ? 1u : 0u;
        uint32_t tot;
        const uint32_t pre = pk_block_exscan(f, tmp, &tot);
        if (f && nlt + neq + pre < k) buf[nlt + neq + pre] = ((uint64_t)kc << 32) | i;
        neq += tot;
    }
    uint32_t n2 = 2;
    while (n2 < k) n2 <<= 1;
    for (uint32_t i = k + threadIdx.x; i < n2; i += blockDim.x) buf[i] = ~0ull;
    __syncthreads();
    pk_bitonic_u64(buf, n2);
    bool t = false;
    for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) {
        if (out_val) out_val[i] = pk_ord2f((uint32_t)(buf[i] >> 32));
        if (out_idx) out_idx[i] = (uint32_t)(buf[i] & 0xffffffffu);
        if (i < flag_lim) {
            // equal to the next entry of the full order (inside the k, or a tie cut at the k-th)
            if (i + 1 < k ? (uint32_t)(buf[i + 1] >> 32) == (uint32_t)(buf[i] >> 32) : nlt + neq > k) t = true;
        }
    }
    *tie = t;
}

// The same selection for rows of at most SORT_THREADS * PER entries, with
// every key read once into registers: thread t holds columns t*PER ..
// t*PER+PER-1 (so a scan over threads is column order).  The radix select
// starts at the highest bit in which the row's keys differ (distances of one
// row usually share their sign and top exponent bits, which would send every
// key of the first pass to one histogram bin); the entries below kc and the
// first ties at it are compacted with one packed scan and placed by rank.
template <int PER, typename KeyF>
__device__ void pk_topk_block_reg(KeyF key, uint32_t L, uint32_t k, float* out_val, uint32_t* out_idx, bool* tie,
                                  uint32_t flag_lim) {
    __shared__ uint32_t hist[258];
    __shared__ uint32_t tmp[8];
    __shared__ uint32_t red[8];
    __shared__ uint64_t buf[TOPK_CAP];
    const uint32_t t = threadIdx.x, base = t * PER;
    const uint32_t nt = base < L ? min((uint32_t)PER, L - base) : 0u;
    uint32_t kv[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) kv[j] = (uint32_t)j < nt ? key(base + j) : 0u;
    const uint32_t kc = pk_select_rank_each(
        [&](auto sink) {
#pragma unroll
            for (int j = 0; j < PER; ++j)
                if ((uint32_t)j < nt) sink(kv[j]);
        },
        k, hist, tmp, red);
    uint32_t cl = 0, ce = 0;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        if ((uint32_t)j < nt) {
            cl += kv[j] < kc ? 1u : 0u;
            ce += kv[j] == kc ? 1u : 0u;
        }
    }
    uint32_t tot;
    const uint32_t pre = pk_block_exscan(cl | (ce << 16), tmp, &tot);
    const uint32_t nlt = tot & 0xffffu, neq = tot >> 16;
    uint32_t pl = pre & 0xffffu, pe = nlt + (pre >> 16);
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        if ((uint32_t)j < nt) {
            const uint64_t e = ((uint64_t)kv[j] << 32) | (base + j);
            if (kv[j] < kc) buf[pl++] = e;
            else if (kv[j] == kc) { if (pe < k) buf[pe] = e; ++pe; }
        }
    }
    __syncthreads();
    bool tf = false;
    for (uint32_t i = t; i < k; i += blockDim.x) {
        const uint64_t e = buf[i];
        const uint32_t ek = (uint32_t)(e >> 32);
        uint32_t r = 0, eqa = 0;  // rank; equal keys after it
        for (uint32_t j = 0; j < k; ++j) {
            const uint64_t f = buf[j];
            r += f < e ? 1u : 0u;
            eqa += ((uint32_t)(f >> 32) == ek && f > e) ? 1u : 0u;
        }
        if (out_val) out_val[r] = pk_ord2f(ek);
        if (out_idx) out_idx[r] = (uint32_t)(e & 0xffffffffu);
        // equal to the next entry of the full order (inside the k, or a tie cut at the k-th)
        if (r < flag_lim && (r + 1 < k ? eqa > 0 : nlt + neq > k)) tf = true;
    }
    *tie = tf;
}

#define TOPK_REG_PER 8
#define TOPK_REG_RANK 256  // largest k placed by rank (larger k: the LDS sort)

__global__ __launch_bounds__(SORT_THREADS) void pk_topk_rows(const float* __restrict__ vals, uint32_t L,
                                                             uint64_t ld_in, uint32_t k, float* __restrict__ out_val,
                                                             uint32_t* __restrict__ out_idx, uint8_t* __restrict__ flags,
                                                             uint32_t flag_lim) {
    const uint64_t row = blockIdx.x;
    const float* v = vals + row * ld_in;
    bool tie = false;
    pk_topk_block([&](uint32_t i) { return v[i]; }, L, k, out_val + row * k, out_idx + row * k, &tie, flag_lim);
    __shared__ uint32_t any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    if (tie) any = 1;
    __syncthreads();
    if (threadIdx.x == 0 && any && flags) flags[row] = 1u | flags[row];
}

// Cell distances of the fine clusters of each query's nearest coarse list,
// row[f] = (cd0 + stationary[c0][f]) + fdist[f]; the k smallest, ascending
// (values only).
__global__ __launch_bounds__(SORT_THREADS) void pk_cell_topk(const float* __restrict__ csd,
                                                             const uint32_t* __restrict__ csi, uint32_t cstride,
                                                             const float* __restrict__ fdist,
                                                             const float* __restrict__ stationary, uint32_t F,
                                                             uint32_t k, float* __restrict__ out_val) {
    const uint64_t row = blockIdx.x;
    const float cd = csd[row * cstride];
    const float* st = stationary + (uint64_t)csi[row * cstride] * F;
    const float* fd = fdist + row * F;
    bool tie = false;
    pk_topk_block([&](uint32_t f) { return (cd + st[f]) + fd[f]; }, F, k, out_val + row * k, nullptr, &tie, 0);
}

__global__ __launch_bounds__(SORT_THREADS) void pk_topk_rows_reg(const float* __restrict__ vals, uint32_t L,
                                                                 uint64_t ld_in, uint32_t k,
                                                                 float* __restrict__ out_val,
                                                                 uint32_t* __restrict__ out_idx,
                                                                 uint8_t* __restrict__ flags, uint32_t flag_lim) {
    const uint64_t row = blockIdx.x;
    const float* v = vals + row * ld_in;
    bool tie = false;
    pk_topk_block_reg<TOPK_REG_PER>([&](uint32_t i) { return pk_f2ord(v[i]); }, L, k, out_val + row * k,
                                    out_idx + row * k, &tie, flag_lim);
    __shared__ uint32_t any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    if (tie) any = 1;
    __syncthreads();
    if (threadIdx.x == 0 && any && flags) flags[row] = 1u | flags[row];
}

__global__ __launch_bounds__(SORT_THREADS) void pk_cell_topk_reg(const float* __restrict__ csd,
                                                                 const uint32_t* __restrict__ csi, uint32_t cstride,
                                                                 const float* __restrict__ fdist,
                                                                 const float* __restrict__ stationary, uint32_t F,
                                                                 uint32_t k, float* __restrict__ out_val) {
    const uint64_t row = blockIdx.x;
    const float cd = csd[row * cstride];
    const float* st = stationary + (uint64_t)csi[row * cstride] * F;
    const float* fd = fdist + row * F;
    bool tie = false;
    pk_topk_block_reg<TOPK_REG_PER>([&](uint32_t f) { return pk_f2ord((cd + st[f]) + fd[f]); }, F, k,
                                    out_val + row * k, nullptr, &tie, 0);
}

extern "C" hipError_t pk_launch_topk_rows(const float* vals, uint32_t rows, uint32_t L, uint64_t ld_in, uint32_t k,
                                          float* out_val, uint32_t* out_idx, uint8_t* flags, uint32_t flag_lim,
                                          hipStream_t s) {
    if (rows == 0 || k == 0) return hipSuccess;
    if (k > TOPK_CAP || k > L) return hipErrorInvalidValue;
    if (L <= SORT_THREADS * TOPK_REG_PER && k <= TOPK_REG_RANK)
        hipLaunchKernelGGL(pk_topk_rows_reg, dim3(rows), dim3(SORT_THREADS), 0, s, vals, L, ld_in, k, out_val,
                           out_idx, flags, flag_lim);
    else
        hipLaunchKernelGGL(pk_topk_rows, dim3(rows), dim3(SORT_THREADS), 0, s, vals, L, ld_in, k, out_val, out_idx,
                           flags, flag_lim);
    return hipGetLastError();
}

extern "C" hipError_t pk_launch_cell_topk(const float* csd, const uint32_t* csi, uint32_t cstride, const float* fdist,
                                          const float* stationary, uint32_t F, uint32_t k, uint32_t rows,
                                          float* out_val, hipStream_t s) {
    if (rows == 0 || k == 0) return hipSuccess;
    if (k > TOPK_CAP || k > F) return hipErrorInvalidValue;
    if (F <= SORT_THREADS * TOPK_REG_PER && k <= TOPK_REG_RANK)
        hipLaunchKernelGGL(pk_cell_topk_reg, dim3(rows), dim3(SORT_THREADS), 0, s, csd, csi, cstride, fdist,
                           stationary, F, k, out_val);
    else
        hipLaunchKernelGGL(pk_cell_topk, dim3(rows), dim3(SORT_THREADS), 0, s, csd, csi, cstride, fdist, stationary,
                           F, k, out_val);
    return hipGetLastError();
}
