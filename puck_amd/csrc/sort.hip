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

__device__ void pk_sort_row(const float* __restrict__ vals, uint64_t row, uint32_t L, uint64_t ld_in, uint32_t keep,
                            float* __restrict__ out_val, uint32_t* __restrict__ out_idx, uint32_t n2,
                            uint8_t* __restrict__ flags, uint32_t flag_lim, uint64_t* sk) {
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
    __syncthreads();  // sk is reused by the next row of a list
}

// The same sort with the keys in registers for n2 = 256 * E (E = 4, 8):
// thread t holds elements t * E .. t * E + E - 1.  Exchanges at distances
// below E stay in the thread, those below 64 * E go to lane t ^ (j / E) of
// the same wave (pk_lane_xor64), only the larger ones (three stages at
// n2 = 2048) go through LDS with barriers.  The sorted row is then in sk as
// pk_sort_row leaves it.
template <int E>
__device__ void pk_sort_row_reg(const float* __restrict__ vals, uint64_t row, uint32_t L, uint64_t ld_in,
                                uint32_t keep, float* __restrict__ out_val, uint32_t* __restrict__ out_idx,
                                uint8_t* __restrict__ flags, uint32_t flag_lim, uint64_t* sk) {
    constexpr uint32_t N2 = SORT_THREADS * E;
    const uint32_t t = threadIdx.x, lane = t & 63u;
    const float* v = vals + row * ld_in;
    uint64_t x[E];
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const uint32_t e = t * E + r;
        x[r] = e < L ? (((uint64_t)pk_f2ord(v[e]) << 32) | e) : ~0ull;
    }
#pragma unroll
    for (uint32_t k = 2; k <= N2; k <<= 1) {
#pragma unroll
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            if (j < (uint32_t)E) {
#pragma unroll
                for (int r = 0; r < E; ++r) {
                    if ((r & j) != 0u) continue;
                    const bool up = ((t * E + r) & k) == 0u;
                    const bool lt = x[r] < x[r + j];
                    const uint64_t lo = lt ? x[r] : x[r + j], hi = lt ? x[r + j] : x[r];
                    x[r] = up ? lo : hi;
                    x[r + j] = up ? hi : lo;
                }
            } else if (j < 64u * E) {
                const uint32_t jl = j / E;
                const bool lower = (lane & jl) == 0u;
#pragma unroll
                for (int r = 0; r < E; ++r)
                    x[r] = pk_cx(x[r], pk_lane_xor64(x[r], jl, lane), lower, ((t * E + r) & k) == 0u);
            } else {
#pragma unroll
                for (int r = 0; r < E; ++r) sk[t * E + r] = x[r];
                __syncthreads();
#pragma unroll
                for (int r = 0; r < E; ++r) {
                    const uint32_t e = t * E + r;
                    x[r] = pk_cx(x[r], sk[e ^ j], (e & j) == 0u, (e & k) == 0u);
                }
                __syncthreads();
            }
        }
    }
#pragma unroll
    for (int r = 0; r < E; ++r) sk[t * E + r] = x[r];
    __syncthreads();
    bool tie = false;
    for (uint32_t i = threadIdx.x; i < keep; i += blockDim.x) {
        uint64_t key = sk[i];
        float fv = pk_ord2f((uint32_t)(key >> 32));
        out_val[row * keep + i] = fv;
        out_idx[row * keep + i] = (uint32_t)(key & 0xffffffffu);
        if (i < flag_lim && i + 1 < L && pk_ord2f((uint32_t)(sk[i + 1] >> 32)) == fv) tie = true;
    }
    if (tie && flags) flags[row] = 1u | flags[row];
    __syncthreads();  // sk is reused by the next row of a list
}

__device__ __forceinline__ void pk_sort_row_any(const float* __restrict__ vals, uint64_t row, uint32_t L,
                                                uint64_t ld_in, uint32_t keep, float* __restrict__ out_val,
                                                uint32_t* __restrict__ out_idx, uint32_t n2,
                                                uint8_t* __restrict__ flags, uint32_t flag_lim, uint64_t* sk) {
    if (n2 == SORT_THREADS * 8u) pk_sort_row_reg<8>(vals, row, L, ld_in, keep, out_val, out_idx, flags, flag_lim, sk);
    else if (n2 == SORT_THREADS * 4u)
        pk_sort_row_reg<4>(vals, row, L, ld_in, keep, out_val, out_idx, flags, flag_lim, sk);
    else pk_sort_row(vals, row, L, ld_in, keep, out_val, out_idx, n2, flags, flag_lim, sk);
}

__global__ __launch_bounds__(SORT_THREADS) void pk_sort_rows(const float* __restrict__ vals, uint32_t L,
                                                             uint64_t ld_in, uint32_t keep,
                                                             float* __restrict__ out_val,
                                                             uint32_t* __restrict__ out_idx, uint32_t n2,
                                                             uint8_t* __restrict__ flags, uint32_t flag_lim) {
    extern __shared__ uint64_t sk[];
    const uint64_t row = blockIdx.x;
    pk_sort_row_any(vals, row, L, ld_in, keep, out_val, out_idx, n2, flags, flag_lim, sk);
}

// The rows listed in list[0 .. *count), grid-stride (the rows a walk will use).
__global__ __launch_bounds__(SORT_THREADS) void pk_sort_rows_list(const float* __restrict__ vals, uint32_t L,
                                                                  uint64_t ld_in, uint32_t keep,
                                                                  float* __restrict__ out_val,
                                                                  uint32_t* __restrict__ out_idx, uint32_t n2,
                                                                  const uint32_t* __restrict__ list,
                                                                  const uint32_t* __restrict__ count) {
    extern __shared__ uint64_t sk[];
    const uint32_t n = *count;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x)
        pk_sort_row_any(vals, list[i], L, ld_in, keep, out_val, out_idx, n2, nullptr, 0, sk);
}

extern "C" hipError_t pk_launch_sort_rows_list(const float* vals, uint32_t max_rows, uint32_t L, uint64_t ld_in,
                                               uint32_t keep, float* out_val, uint32_t* out_idx,
                                               const uint32_t* list, const uint32_t* count, hipStream_t s) {
    if (max_rows == 0 || keep == 0) return hipSuccess;
    if (L > 8192 || keep > L) return hipErrorInvalidValue;
    static PkLdsCache lds_cache;
    size_t* slot = pk_lds_slot(lds_cache);
    if (!slot) return hipErrorInvalidDevice;
    if (!*slot) {
        const hipError_t e = pk_allow_lds((const void*)pk_sort_rows_list, 64 * 1024, slot);
        if (e != hipSuccess) return e;
    }
    const size_t allowed = *slot;
    uint32_t n2 = 2;
    while (n2 < L) n2 <<= 1;
    if (n2 * sizeof(uint64_t) > allowed) return hipErrorInvalidValue;
    const uint32_t grid = max_rows < 1024 ? max_rows : 1024;
    hipLaunchKernelGGL(pk_sort_rows_list, dim3(grid), dim3(SORT_THREADS), n2 * sizeof(uint64_t), s, vals, L, ld_in,
                       keep, out_val, out_idx, n2, list, count);
    return hipGetLastError();
}

extern "C" hipError_t pk_launch_sort_rows(const float* vals, uint32_t rows, uint32_t L, uint64_t ld_in,
                                          uint32_t keep, float* out_val, uint32_t* out_idx, uint8_t* flags,
                                          uint32_t flag_lim, hipStream_t s) {
    if (rows == 0 || keep == 0) return hipSuccess;
    if (L > 8192 || keep > L) return hipErrorInvalidValue;
    static PkLdsCache lds_cache;
    size_t* slot = pk_lds_slot(lds_cache);
    if (!slot) return hipErrorInvalidDevice;
    if (!*slot) {
        const hipError_t e = pk_allow_lds((const void*)pk_sort_rows, 64 * 1024, slot);
        if (e != hipSuccess) return e;
    }
    const size_t allowed = *slot;
    uint32_t n2 = 2;
    while (n2 < L) n2 <<= 1;
    if (n2 * sizeof(uint64_t) > allowed) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pk_sort_rows, dim3(rows), dim3(SORT_THREADS), n2 * sizeof(uint64_t), s, vals, L, ld_in,
                       keep, out_val, out_idx, n2, flags, flag_lim);
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

// The same selection by one wave per row (rows of at most 64 * TOPK_WAVE_PER
// entries, k <= TOPK_WAVE_K): lane l holds columns j * 64 + l in registers.
// A radix select of the k-th smallest key kc over a per-wave LDS histogram
// (of the keys that can be among the k smallest, see below),
// starting at the highest bit in which the row's keys differ (distances of
// one row usually share their sign and top exponent bits, which would send
// every key of a first pass to one bin); the entries below kc and then the
// first ties at it in column order are compacted with ballots, and the k
// kept entries are placed by rank.  No block barrier: the wave's LDS
// operations complete in order.
#define TOPK_WAVE_PER 32
#define TOPK_WAVE_K 256
__device__ __forceinline__ void pk_wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// load(kv) fills kv[j] with the key of column min(j * 64 + lane, L - 1) for
// every j, issuing all its loads before combining any (a load under a lane
// condition, or one whose value is used at once, is waited for before the
// next is issued).
template <typename LoadF>
__device__ void pk_topk_wave(LoadF load, uint32_t L, uint32_t k, float* out_val, uint32_t* out_idx, bool* tie,
                             uint32_t flag_lim, uint32_t* hist, uint64_t* buf) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t kv[TOPK_WAVE_PER];
    uint32_t ka = 0xffffffffu, ko = 0u;
    load(kv);
    // Candidates: with k <= 64 and a key in every lane, the largest lane minimum
    // T bounds the k-th smallest key from above (each lane has a key <= T), so
    // only keys <= T enter the histograms (on random columns ~1/8 of the row:
    // LDS atomics are the select's cost).
    uint32_t T = 0xffffffffu;
    if (k <= 64u && L >= 64u) {
        uint32_t mn = 0xffffffffu;
#pragma unroll
        for (int j = 0; j < TOPK_WAVE_PER; ++j)
            if (j * 64u + lane < L) mn = min(mn, kv[j]);
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) mn = max(mn, pk_lane_xor(mn, (uint32_t)o, lane));
        T = mn;
    }
#pragma unroll
    for (int j = 0; j < TOPK_WAVE_PER; ++j) {
        if (j * 64u + lane < L && kv[j] <= T) {
            ka &= kv[j];
            ko |= kv[j];
        }
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        ka &= pk_lane_xor(ka, (uint32_t)o, lane);
        ko |= pk_lane_xor(ko, (uint32_t)o, lane);
    }
    const uint32_t diff = ka ^ ko;
    const int top = diff ? 32 - __clz(diff) : 0;  // bits [top-1, 0] vary
    uint32_t mask = top >= 32 ? 0u : ~((1u << top) - 1u);
    uint32_t prefix = ka & mask, want = k;
    for (int hi = top; hi > 0; hi -= 8) {
        const int shift = hi > 8 ? hi - 8 : 0;
        const uint32_t dm = (1u << (hi - shift)) - 1u;
#pragma unroll
        for (int q = 0; q < 4; ++q) hist[4u * lane + q] = 0u;
        pk_wave_lds_order();
#pragma unroll
        for (int j = 0; j < TOPK_WAVE_PER; ++j)
            if (j * 64u + lane < L && kv[j] <= T && (kv[j] & mask) == prefix)
                atomicAdd(&hist[(kv[j] >> shift) & dm], 1u);
        pk_wave_lds_order();
        // lane l owns bins 4l .. 4l+3: exclusive prefix over lanes, then the bin of rank `want`
        uint32_t h[4], sum = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            h[q] = hist[4u * lane + q];
            sum += h[q];
        }
        const uint32_t inc = pk_wave_incscan(sum);
        uint32_t before = inc - sum, bin = 0, rest = 0;
        const bool mine = sum > 0 && before < want && want <= inc;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (mine && rest == 0 && want <= before + h[q]) {
                bin = 4u * lane + q;
                rest = want - before;
            }
            before += h[q];
        }
        const uint32_t src = (uint32_t)__builtin_ctzll(__ballot(mine));
        bin = __shfl(bin, src, 64);
        want = __shfl(rest, src, 64);
        prefix |= bin << shift;
        mask |= dm << shift;
        pk_wave_lds_order();  // every lane has read the histogram before the next pass clears it
    }
    const uint32_t kc = prefix;
    // entries below kc (fewer than k), then ties at kc by column, into buf
    uint32_t nlt = 0;
#pragma unroll
    for (int j = 0; j < TOPK_WAVE_PER; ++j) {
        const bool lt = j * 64u + lane < L && kv[j] < kc;
        const uint64_t m = __ballot(lt);
        if (lt) buf[nlt + (uint32_t)__popcll(m & below)] = ((uint64_t)kv[j] << 32) | (j * 64u + lane);
        nlt += (uint32_t)__popcll(m);
    }
    uint32_t neq = 0;
#pragma unroll
    for (int j = 0; j < TOPK_WAVE_PER; ++j) {
        const bool eq = j * 64u + lane < L && kv[j] == kc;
        const uint64_t m = __ballot(eq);
        const uint32_t r = nlt + neq + (uint32_t)__popcll(m & below);
        if (eq && r < k) buf[r] = ((uint64_t)kc << 32) | (j * 64u + lane);
        neq += (uint32_t)__popcll(m);
    }
    pk_wave_lds_order();
    bool tf = false;
    if (k <= 64u) {
        // one entry per lane, bitonic-sorted across the wave by (key, column):
        // lane r then holds rank r
        uint64_t e = lane < k ? buf[lane] : ~0ull;
#pragma unroll
        for (uint32_t kk = 2; kk <= 64u; kk <<= 1)
#pragma unroll
            for (uint32_t j = kk >> 1; j > 0; j >>= 1)
                e = pk_cx(e, pk_lane_xor64(e, j, lane), (lane & j) == 0u, (lane & kk) == 0u);
        const uint32_t ek = (uint32_t)(e >> 32);
        const uint32_t nk = (uint32_t)__shfl_down((int)ek, 1, 64);  // key of rank lane + 1
        if (lane < k) {
            if (out_val) out_val[lane] = pk_ord2f(ek);
            if (out_idx) out_idx[lane] = (uint32_t)(e & 0xffffffffu);
            if (lane < flag_lim && (lane + 1u < k ? nk == ek : nlt + neq > k)) tf = true;
        }
        *tie = __ballot(tf) != 0ull;
        return;
    }
    for (uint32_t i = lane; i < k; i += 64u) {
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
    *tie = __ballot(tf) != 0ull;
}

// Per-wave LDS of the wave kernels: histogram (256 words) and k entries.
#define TOPK_WAVE_LDS (256 * 4 + TOPK_WAVE_K * 8)

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
                                                             uint32_t k, float* __restrict__ out_val,
                                                             uint32_t* __restrict__ out_idx) {
    const uint64_t row = blockIdx.x;
    const float cd = csd[row * cstride];
    const float* st = stationary + (uint64_t)csi[row * cstride] * F;
    const float* fd = fdist + row * F;
    bool tie = false;
    pk_topk_block([&](uint32_t f) { return (cd + st[f]) + fd[f]; }, F, k, out_val + row * k,
                  out_idx ? out_idx + row * k : nullptr, &tie, 0);
}

__global__ __launch_bounds__(SORT_THREADS) void pk_topk_rows_wave(const float* __restrict__ vals, uint32_t rows,
                                                                  uint32_t L, uint64_t ld_in, uint32_t k,
                                                                  float* __restrict__ out_val,
                                                                  uint32_t* __restrict__ out_idx,
                                                                  uint8_t* __restrict__ flags, uint32_t flag_lim) {
    __shared__ __align__(16) unsigned char lds[(SORT_THREADS / 64) * TOPK_WAVE_LDS];
    const uint32_t w = threadIdx.x >> 6;
    const uint64_t row = (uint64_t)blockIdx.x * (SORT_THREADS / 64) + w;
    if (row >= rows) return;
    uint64_t* buf = (uint64_t*)(lds + w * TOPK_WAVE_LDS);
    uint32_t* hist = (uint32_t*)(lds + w * TOPK_WAVE_LDS + TOPK_WAVE_K * 8);
    const float* v = vals + row * ld_in;
    bool tie = false;
    const uint32_t lane = threadIdx.x & 63u;
    pk_topk_wave(
        [&](uint32_t(&kv)[TOPK_WAVE_PER]) {
            float x[TOPK_WAVE_PER];
#pragma unroll
            for (int j = 0; j < TOPK_WAVE_PER; ++j) x[j] = v[min(j * 64u + lane, L - 1u)];
#pragma unroll
            for (int j = 0; j < TOPK_WAVE_PER; ++j) kv[j] = pk_f2ord(x[j]);
        },
        L, k, out_val + row * k, out_idx + row * k, &tie, flag_lim, hist, buf);
    if ((threadIdx.x & 63u) == 0 && tie && flags) flags[row] = 1u | flags[row];
}

__global__ __launch_bounds__(SORT_THREADS) void pk_cell_topk_wave(const float* __restrict__ csd,
                                                                  const uint32_t* __restrict__ csi, uint32_t cstride,
                                                                  const float* __restrict__ fdist,
                                                                  const float* __restrict__ stationary, uint32_t F,
                                                                  uint32_t k, uint32_t rows,
                                                                  float* __restrict__ out_val,
                                                                  uint32_t* __restrict__ out_idx) {
    __shared__ __align__(16) unsigned char lds[(SORT_THREADS / 64) * TOPK_WAVE_LDS];
    const uint32_t w = threadIdx.x >> 6;
    const uint64_t row = (uint64_t)blockIdx.x * (SORT_THREADS / 64) + w;
    if (row >= rows) return;
    uint64_t* buf = (uint64_t*)(lds + w * TOPK_WAVE_LDS);
    uint32_t* hist = (uint32_t*)(lds + w * TOPK_WAVE_LDS + TOPK_WAVE_K * 8);
    const float cd = csd[row * cstride];
    const float* st = stationary + (uint64_t)csi[row * cstride] * F;
    const float* fd = fdist + row * F;
    bool tie = false;
    const uint32_t lane = threadIdx.x & 63u;
    pk_topk_wave(
        [&](uint32_t(&kv)[TOPK_WAVE_PER]) {
            float x[TOPK_WAVE_PER], y[TOPK_WAVE_PER];
#pragma unroll
            for (int j = 0; j < TOPK_WAVE_PER; ++j) x[j] = st[min(j * 64u + lane, F - 1u)];
#pragma unroll
            for (int j = 0; j < TOPK_WAVE_PER; ++j) y[j] = fd[min(j * 64u + lane, F - 1u)];
#pragma unroll
            for (int j = 0; j < TOPK_WAVE_PER; ++j) kv[j] = pk_f2ord((cd + x[j]) + y[j]);
        },
        F, k, out_val + row * k, out_idx ? out_idx + row * k : nullptr, &tie, 0, hist, buf);
}

extern "C" hipError_t pk_launch_topk_rows(const float* vals, uint32_t rows, uint32_t L, uint64_t ld_in, uint32_t k,
                                          float* out_val, uint32_t* out_idx, uint8_t* flags, uint32_t flag_lim,
                                          hipStream_t s) {
    if (rows == 0 || k == 0) return hipSuccess;
    if (k > TOPK_CAP || k > L) return hipErrorInvalidValue;
    if (L <= 64 * TOPK_WAVE_PER && k <= TOPK_WAVE_K)
        hipLaunchKernelGGL(pk_topk_rows_wave, dim3((rows + 3) / 4), dim3(SORT_THREADS), 0, s, vals, rows, L, ld_in, k,
                           out_val, out_idx, flags, flag_lim);
    else
        hipLaunchKernelGGL(pk_topk_rows, dim3(rows), dim3(SORT_THREADS), 0, s, vals, L, ld_in, k, out_val, out_idx,
                           flags, flag_lim);
    return hipGetLastError();
}

extern "C" hipError_t pk_launch_cell_topk(const float* csd, const uint32_t* csi, uint32_t cstride, const float* fdist,
                                          const float* stationary, uint32_t F, uint32_t k, uint32_t rows,
                                          float* out_val, uint32_t* out_idx, hipStream_t s) {
    if (rows == 0 || k == 0) return hipSuccess;
    if (k > TOPK_CAP || k > F) return hipErrorInvalidValue;
    if (F <= 64 * TOPK_WAVE_PER && k <= TOPK_WAVE_K)
        hipLaunchKernelGGL(pk_cell_topk_wave, dim3((rows + 3) / 4), dim3(SORT_THREADS), 0, s, csd, csi, cstride, fdist,
                           stationary, F, k, rows, out_val, out_idx);
    else
        hipLaunchKernelGGL(pk_cell_topk, dim3(rows), dim3(SORT_THREADS), 0, s, csd, csi, cstride, fdist, stationary,
                           F, k, out_val, out_idx);
    return hipGetLastError();
}
