// common.h -- device helpers shared by the gfx950 kernels of libpuckgpu.
//
// Arithmetic convention: every source is compiled with -ffp-contract=off, so
// a*b+c is two rounded operations unless fmaf() is written.  Dot products are
// sequential fmaf chains from +0.0f (the engine's exact-arithmetic order); the
// lookup-table and rerank orders follow the reference's SIMD kernels.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

// A launcher's dynamic-LDS allowance per device (kernel attributes are set
// per device): a zero-initialised static of the launcher, slot = the current
// device.
#define PK_MAX_DEVICES 64
struct PkLdsCache {
    size_t v[PK_MAX_DEVICES];
};
static inline size_t* pk_lds_slot(PkLdsCache& c) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= PK_MAX_DEVICES) return nullptr;
    return &c.v[dev];
}

// Lets kernel f take up to `bytes` of dynamic LDS, less its own static LDS
// (a CU holds 160 KB); *allowed receives the dynamic bytes allowed.
static inline hipError_t pk_allow_lds(const void* f, size_t bytes, size_t* allowed) {
    hipFuncAttributes at;
    hipError_t e = hipFuncGetAttributes(&at, f);
    if (e != hipSuccess) return e;
    const size_t cap = 160 * 1024 - at.sharedSizeBytes;
    if (bytes > cap) bytes = cap;
    e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess && allowed) *allowed = bytes;
    return e;
}

#define PK_KS 256

// 0x7f7f7f7f: the value MaxHeap's memset(0x7f) leaves in an empty slot.
__host__ __device__ inline float pk_sentinel() {
    union { uint32_t u; float f; } v;
    v.u = 0x7f7f7f7fu;
    return v.f;
}

// Order-preserving float <-> uint32 map (total order of non-NaN floats).
__device__ __forceinline__ uint32_t pk_f2ord(float f) {
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float pk_ord2f(uint32_t o) {
    uint32_t b = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    return __uint_as_float(b);
}

// Sequential FMA dot product (the engine's exact order).
__device__ __forceinline__ float pk_dot(const float* a, const float* b, uint32_t n) {
    float acc = 0.0f;
    for (uint32_t i = 0; i < n; ++i) acc = fmaf(a[i], b[i], acc);
    return acc;
}

// cblas_sdot of 1, 2 or 4 floats in the order Intel MKL takes for those
// lengths (found by calling it and comparing result bits): rounded products,
// then p0 + p1 or (p0 + p1) + (p2 + p3), no fused multiply-add.  These are the
// dot products with which nearest_center (kmeans.cpp:221-258) picks the
// filter-PQ codes (puck_index.cpp:667-668).  Other n: pk_dot.
__device__ __forceinline__ float pk_sdot_short(const float* a, const float* b, uint32_t n) {
    if (n == 1) return __fmul_rn(a[0], b[0]);
    if (n == 2) return __fadd_rn(__fmul_rn(a[0], b[0]), __fmul_rn(a[1], b[1]));
    if (n == 4)
        return __fadd_rn(__fadd_rn(__fmul_rn(a[0], b[0]), __fmul_rn(a[1], b[1])),
                         __fadd_rn(__fmul_rn(a[2], b[2]), __fmul_rn(a[3], b[3])));
    return pk_dot(a, b, n);
}

// Lookup-table entry ||x - y||^2 for a sub-vector of length 1, 2 or 4, with
// the summation order of faiss fvec_L2sqr_ny_D1/D2/D4
// (puck/puck/distances_simd.cpp:72-133): rounded squares, then pairwise sums.
__device__ __forceinline__ float pk_lut_entry(const float* x, const float* y, uint32_t lsq) {
    if (lsq == 1) {
        float d0 = x[0] - y[0];
        return d0 * d0;
    }
    if (lsq == 2) {
        float d0 = x[0] - y[0], d1 = x[1] - y[1];
        return d0 * d0 + d1 * d1;
    }
    float d0 = x[0] - y[0], d1 = x[1] - y[1], d2 = x[2] - y[2], d3 = x[3] - y[3];
    return (d0 * d0 + d1 * d1) + (d2 * d2 + d3 * d3);
}

// The filter lookup table of one query (quantization.cpp:260-270): entry e
// = pk_lut_entry(q sub-vector e / 256, codebook row e), for the nsq * 256
// entries, written to LDS by the whole block.  Each thread issues the
// codebook loads of PK_LUT_UNROLL entries before computing any of them (one
// load latency per group instead of one per entry; 8 keeps the budget scan
// within two waves per SIMD).
#ifndef PK_LUT_UNROLL
#define PK_LUT_UNROLL 8
#endif
template <int LSQ, typename Put>
__device__ void pk_build_lut_t(const float* __restrict__ qv, const float* __restrict__ codebook, uint32_t n,
                               Put put) {
    constexpr int U = PK_LUT_UNROLL;
    for (uint32_t e0 = threadIdx.x; e0 < n; e0 += U * blockDim.x) {
        float c[U][LSQ];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            // at an in-bounds entry for every lane: a load under a lane
            // condition is waited for before the next one is issued
            const uint32_t e = min(e0 + u * blockDim.x, n - 1u);
            if constexpr (LSQ == 4) {
                const float4 w = *(const float4*)(codebook + (uint64_t)e * 4);
                c[u][0] = w.x;
                c[u][1] = w.y;
                c[u][2] = w.z;
                c[u][3] = w.w;
            } else if constexpr (LSQ == 2) {
                const float2 w = *(const float2*)(codebook + (uint64_t)e * 2);
                c[u][0] = w.x;
                c[u][1] = w.y;
            } else {
                c[u][0] = codebook[e];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t e = e0 + u * blockDim.x;
            if (e < n) put(e, pk_lut_entry(qv + (e / 256u) * LSQ, c[u], LSQ));
        }
    }
}

template <typename Put>
__device__ __forceinline__ void pk_build_lut_put(const float* qv, const float* codebook, uint32_t nsq, uint32_t lsq,
                                                 Put put) {
    if (lsq == 1) pk_build_lut_t<1>(qv, codebook, nsq * 256u, put);
    else if (lsq == 2) pk_build_lut_t<2>(qv, codebook, nsq * 256u, put);
    else pk_build_lut_t<4>(qv, codebook, nsq * 256u, put);
}

__device__ __forceinline__ void pk_build_lut(float* lut, const float* qv, const float* codebook, uint32_t nsq,
                                             uint32_t lsq) {
    pk_build_lut_put(qv, codebook, nsq, lsq, [&](uint32_t e, float v) { lut[e] = v; });
}

// FP16 filter lookup table (a precision form; the reference's table is FP32).
// Entries are scaled by 2^e, rounded to FP16 (nearest even) and summed in
// FP32 in the reference's 4-lane order; the sum is scaled back by 2^-e.  e is
// chosen per query so that the largest possible entry, (|q| + cbn)^2 with cbn
// the largest codeword norm of the codebook, stays at or below 65504 (FP16's
// largest finite value): e = floor(log2(65504 / (sqrt(|q|^2) + cbn)^2)),
// clamped to [-30, 30].  Scaling by a power of two is exact, so a table kept
// in FP32 holding the rounded values times 2^-e (the walk and the per-query
// scan) gives the same sums, bit for bit, as the FP16 table summed and then
// scaled (the split scan).
__device__ __forceinline__ int pk_lut16_exp(float qn, float cbn) {
    const float r = __fadd_rn(__fsqrt_rn(qn), cbn);
    const float B = __fmul_rn(r, r);
    const float x = __fdiv_rn(65504.0f, B);
    int e = 30;
    if (x < INFINITY) {
        int ex = 0;
        (void)frexpf(x, &ex);  // x = m * 2^ex, m in [0.5, 1)
        e = ex - 1;
    }
    return e < -30 ? -30 : (e > 30 ? 30 : e);
}

// One entry of the FP16 table in FP32: round(v * 2^e) * 2^-e.
__device__ __forceinline__ float pk_lut16_q(float v, float s, float inv) {
    return __fmul_rn(__half2float(__float2half_rn(__fmul_rn(v, s))), inv);
}

// The filter table of a search in its precision: FP32, or the FP16 table's
// values in FP32 (lut16 non-zero; qn = |q|^2, cbn as above).
__device__ __forceinline__ void pk_build_filter_lut(float* lut, const float* qv, const float* codebook, uint32_t nsq,
                                                    uint32_t lsq, int lut16, float qn, float cbn) {
    if (!lut16) {
        pk_build_lut(lut, qv, codebook, nsq, lsq);
        return;
    }
    const int e = pk_lut16_exp(qn, cbn);
    const float s = ldexpf(1.0f, e), inv = ldexpf(1.0f, -e);
    pk_build_lut_put(qv, codebook, nsq, lsq, [&](uint32_t i, float v) { lut[i] = pk_lut16_q(v, s, inv); });
}

// Pivot (puck_index.cpp:517, :544): (top - |q|^2) / radius_rate / 2.0 with the
// subtraction in float and the divisions in double, stored as float.
__device__ __forceinline__ float pk_pivot(float top, float qn, double rr) {
    return (float)((double)(top - qn) / rr / 2.0);
}

// Exact squared L2 in the order GCC -Ofast -march=native gives the reference's
// rerank loop (L2Sqr16Ext/L2SqrExt, hnsw_distfunc_opt_impl_inline.h:41-122):
//   per 16-block  l_j = fma(d_j, d_j, d_{8+j}^2);
//                 t = (((l0+l1)+(l2+l3)) + ((l4+l5)+l6)) + (l7 + t)
//   tail (<16)    4-lane fma blocks, (b0+b1)+(b2+b3), then a scalar fma chain.
// Sum over the last dim % 16 elements (left > 0) in the compiled tail order.
// T: float rows, or uint8_t rows of integer features in [0, 255] (each value
// converted exactly to float, so the arithmetic is that of the float rows).
template <typename T>
__device__ __forceinline__ float pk_rerank_tail(const T* __restrict__ x, const float* q, uint32_t left) {
    float b[4] = {0.f, 0.f, 0.f, 0.f};
    uint32_t i = 0;
    for (; i + 4 <= left; i += 4)
        for (int j = 0; j < 4; ++j) {
            float d = (float)x[i + j] - q[i + j];
            b[j] = fmaf(d, d, b[j]);
        }
    float s = (b[0] + b[1]) + (b[2] + b[3]);
    for (; i < left; ++i) {
        float d = (float)x[i] - q[i];
        s = fmaf(d, d, s);
    }
    return s;
}

template <typename T>
__device__ __forceinline__ float pk_rerank_dist(const T* __restrict__ x, const float* q, uint32_t dim) {
    float t = 0.0f;
    uint32_t m = 0;
    for (; m + 16 <= dim; m += 16) {
        float l[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            float d0 = (float)x[m + j] - q[m + j];
            float d1 = (float)x[m + 8 + j] - q[m + 8 + j];
            l[j] = fmaf(d0, d0, d1 * d1);
        }
        t = (((l[0] + l[1]) + (l[2] + l[3])) + ((l[4] + l[5]) + l[6])) + (l[7] + t);
    }
    if (m < dim) t = t + pk_rerank_tail(x + m, q + m, dim - m);
    return t;
}

// Lane exchanges of a wave (64 lanes): the value of lane (lane ^ j) for
// j = 1 .. 32 (DPP quad_perm / row shifts for 1 to 8, ds_swizzle for 16,
// ds_bpermute for 32), and one side of a bitonic compare-exchange.
__device__ __forceinline__ uint32_t pk_lane_xor(uint32_t x, uint32_t j, uint32_t lane) {
    if (j == 1u) return (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0xB1, 0xF, 0xF, false);  // quad_perm 1,0,3,2
    if (j == 2u) return (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x4E, 0xF, 0xF, false);  // quad_perm 2,3,0,1
    if (j == 4u) {
        const uint32_t up = (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x104, 0xF, 0xF, false);  // row_shl:4
        const uint32_t dn = (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x114, 0xF, 0xF, false);  // row_shr:4
        return (lane & 4u) ? dn : up;
    }
    if (j == 8u) {
        const uint32_t up = (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x108, 0xF, 0xF, false);  // row_shl:8
        const uint32_t dn = (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x118, 0xF, 0xF, false);  // row_shr:8
        return (lane & 8u) ? dn : up;
    }
    if (j == 16u) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x401F);  // bitmask mode, xor 16
    return (uint32_t)__shfl_xor((int)x, 32);
}
__device__ __forceinline__ uint64_t pk_lane_xor64(uint64_t x, uint32_t j, uint32_t lane) {
    const uint32_t lo = pk_lane_xor((uint32_t)x, j, lane);
    const uint32_t hi = pk_lane_xor((uint32_t)(x >> 32), j, lane);
    return ((uint64_t)hi << 32) | lo;
}
// compare-exchange seen from one side: the lower element of an ascending pair
// (or the upper of a descending one) keeps the minimum
__device__ __forceinline__ uint64_t pk_cx(uint64_t x, uint64_t y, bool lower, bool up) {
    const bool keep_min = lower == up;
    const bool lt = x < y;
    return (keep_min == lt) ? x : y;
}

// Inclusive prefix sum over the 64 lanes of a wave by DPP moves: row shifts
// 1, 2, 4, 8 within each 16-lane row (lanes before the row's start add 0),
// then row broadcasts of lane 15 into rows 1 and 3 and of lane 31 into rows 2
// and 3 (the rows not enabled add 0).
__device__ __forceinline__ uint32_t pk_wave_incscan(uint32_t x) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, true);   // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, true);   // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, true);   // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, true);   // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);  // row_bcast:15
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return x;
}

// Block-wide exclusive scan of one uint32 per thread (blockDim.x <= 1024,
// multiple of 64).  `tmp` needs blockDim.x/64 + 1 words of LDS.  Returns the
// exclusive prefix; *total receives the block sum.  Must be called by every
// thread of the block.
__device__ __forceinline__ uint32_t pk_block_exscan(uint32_t v, uint32_t* tmp, uint32_t* total) {
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    const uint32_t x = pk_wave_incscan(v);
    if (lane == 63) tmp[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int w = 0; w < nw; ++w) {
            uint32_t t = tmp[w];
            tmp[w] = acc;
            acc += t;
        }
        tmp[nw] = acc;
    }
    __syncthreads();
    uint32_t r = tmp[wid] + x - v;
    *total = tmp[nw];
    __syncthreads();
    return r;
}

// AND and OR of ka / ko over the block (blockDim.x a multiple of 64, at most
// 1024); red needs 2 * blockDim.x / 64 words.  Every thread gets both.
__device__ __forceinline__ void pk_block_andor(uint32_t& ka, uint32_t& ko, uint32_t* red) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        ka &= pk_lane_xor(ka, (uint32_t)o, threadIdx.x & 63u);
        ko |= pk_lane_xor(ko, (uint32_t)o, threadIdx.x & 63u);
    }
    if ((threadIdx.x & 63) == 0) {
        red[2 * (threadIdx.x >> 6)] = ka;
        red[2 * (threadIdx.x >> 6) + 1] = ko;
    }
    __syncthreads();
    ka = 0xffffffffu;
    ko = 0u;
    for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) {
        ka &= red[2 * w];
        ko |= red[2 * w + 1];
    }
    __syncthreads();
}

// Key of rank `want` (1-based, at most the number of keys) among the keys the
// threads hand over: each(sink) calls sink(key) for each of the calling
// thread's keys, and is called once per pass.  MSB-first radix select with
// 8-bit digits that starts at the highest bit in which the keys differ (the
// distances of one row usually share their sign and top exponent bits,
// which would send every key of a first pass to one histogram bin).
// Two block barriers per pass: the histogram is double-buffered (the other
// waves clear the next pass's buffer while wave 0 scans this one) and wave 0
// finds the digit with one wave-wide prefix sum over 4 bins per lane.
// blockDim.x >= 256, a multiple of 64; hist has PK_SEL_HIST words, tmp
// blockDim.x / 64 + 1, red 2 * blockDim.x / 64.  Every thread returns it.
#define PK_SEL_HIST 520
template <typename EachF>
__device__ uint32_t pk_select_rank_each(EachF each, uint32_t want, uint32_t* hist, uint32_t* tmp, uint32_t* red) {
    (void)tmp;
    const uint32_t t = threadIdx.x;
    uint32_t ka = 0xffffffffu, ko = 0u;
    each([&](uint32_t k) {
        ka &= k;
        ko |= k;
    });
    if (t < 256) hist[t] = 0;  // buffer 0, ordered before its use by the barriers of pk_block_andor
    pk_block_andor(ka, ko, red);
    const uint32_t diff = ka ^ ko;
    const int top = diff ? 32 - __clz(diff) : 0;  // bits [top-1, 0] vary
    uint32_t mask = top >= 32 ? 0u : ~((1u << top) - 1u);
    uint32_t prefix = ka & mask;
    int buf = 0;
    for (int hi = top; hi > 0; hi -= 8) {
        const int shift = hi > 8 ? hi - 8 : 0;
        const uint32_t dm = (1u << (hi - shift)) - 1u;
        uint32_t* h = hist + 260 * buf;
        uint32_t* hn = hist + 260 * (buf ^ 1);
        each([&](uint32_t k) {
            if ((k & mask) == prefix) atomicAdd(&h[(k >> shift) & dm], 1u);
        });
        __syncthreads();
        if (t < 64) {
            const uint32_t c0 = h[4 * t], c1 = h[4 * t + 1], c2 = h[4 * t + 2], c3 = h[4 * t + 3];
            const uint32_t sum = (c0 + c1) + (c2 + c3);
            const uint32_t incl = pk_wave_incscan(sum);
            const uint32_t before = incl - sum;
            if (before < want && want <= incl) {  // exactly one lane
                const uint32_t cs[4] = {c0, c1, c2, c3};
                uint32_t bb = before, e = 0;
#pragma unroll
                for (uint32_t j = 0; j < 3; ++j)
                    if (e == j && want > bb + cs[j]) {
                        bb += cs[j];
                        e = j + 1;
                    }
                h[256] = 4 * t + e;
                h[257] = want - bb;
            }
        } else {
            for (uint32_t i = t - 64; i < 256; i += blockDim.x - 64) hn[i] = 0;
        }
        __syncthreads();
        prefix |= h[256] << shift;
        mask |= dm << shift;
        want = h[257];
        buf ^= 1;
    }
    return prefix;
}

// Key of rank `want` (1-based) among key(0..n-1): MSB-first radix select
// with 8-bit digits, keys recomputed on every pass.  blockDim.x == 256 (one
// histogram bin per thread); hist has 258 words.  Every thread returns it.
template <typename KeyF>
__device__ uint32_t pk_select_rank(KeyF key, uint32_t n, uint32_t want, uint32_t* hist, uint32_t* tmp) {
    uint32_t prefix = 0, mask = 0;
    for (int shift = 24; shift >= 0; shift -= 8) {
        hist[threadIdx.x] = 0;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
            const uint32_t k = key(i);
            if ((k & mask) == prefix) atomicAdd(&hist[(k >> shift) & 255u], 1u);
        }
        __syncthreads();
        const uint32_t c = hist[threadIdx.x];
        uint32_t tot;
        const uint32_t before = pk_block_exscan(c, tmp, &tot);
        if (c > 0 && before < want && want <= before + c) {
            hist[256] = threadIdx.x;
            hist[257] = want - before;
        }
        __syncthreads();
        prefix |= hist[256] << shift;
        mask |= 255u << shift;
        want = hist[257];
        __syncthreads();
    }
    return prefix;
}
