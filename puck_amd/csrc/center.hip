// center.hip -- reduced-precision query x center distances on the matrix
// cores (the precision form of the search; the exact form is pk_gemm_exact).
//
//   BF16: q and c rounded to bfloat16, v_mfma_f32_32x32x8_bf16, FP32 sums.
//   FP8 : per-row scaled e4m3 (q / s_q, c / s_c with s = max|x| / 448),
//         v_mfma_f32_32x32x16_fp8_fp8, FP32 sums, times s_q * s_c.
// Epilogue as pk_gemm_exact: SUB_FROM_BIAS (bias[n] - dot) or NEG (-dot).
//
// A workgroup of 4 waves computes a 64 x 64 tile; the whole K extent (dim
// <= 256) of both operands is staged in LDS once, so each output needs a
// single pass of MFMAs.  For the center shapes (K = dim ~ 100) the work is
// bound by writing the m x (C + F) FP32 outputs, not by the matrix cores.
//
// The coarse rows that decide the kept coarse lists are rescored in FP32
// afterwards (pk_rescore_coarse): the (keep + margin) best approximate
// candidates are recomputed with the exact dot order and the keep smallest
// exact values are kept, ascending by (distance, id).
#include "common.h"

typedef short pkc_v4s __attribute__((ext_vector_type(4)));
typedef float pkc_v16f __attribute__((ext_vector_type(16)));

enum { PKC_BF16 = 1, PKC_FP8 = 2 };
#define PKC_SUB_FROM_BIAS 1
#define PKC_NEG 2
#define PKC_FP8_MAX 448.0f

__device__ __forceinline__ uint16_t pkc_f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Row conversion of n x dim FP32 rows: BF16 (out16) or scaled FP8 (out8 +
// scale[row]).  One wave per row.
__global__ void pkc_convert_rows(const float* __restrict__ x, uint32_t n, uint32_t dim, int prec,
                                 uint16_t* __restrict__ out16, uint8_t* __restrict__ out8,
                                 float* __restrict__ scale, uint32_t ld) {
    const uint32_t row = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const float* r = x + (uint64_t)row * dim;
    if (prec == PKC_BF16) {
        for (uint32_t k = lane; k < ld; k += 64) out16[(uint64_t)row * ld + k] = k < dim ? pkc_f2bf(r[k]) : 0;
        return;
    }
    float m = 0.f;
    for (uint32_t k = lane; k < dim; k += 64) m = fmaxf(m, fabsf(r[k]));
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    const float s = m > 0.f ? m / PKC_FP8_MAX : 1.f;
    const float inv = 1.f / s;
    for (uint32_t k = 2 * lane; k < ld; k += 128) {
        const float a = k < dim ? r[k] * inv : 0.f;
        const float b = k + 1 < dim ? r[k + 1] * inv : 0.f;
        const int p = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
        out8[(uint64_t)row * ld + k] = (uint8_t)(p & 0xff);
        if (k + 1 < ld) out8[(uint64_t)row * ld + k + 1] = (uint8_t)((p >> 8) & 0xff);
    }
    if (lane == 0) scale[row] = s;
}

// C[m][n] = epi(dot(A[m], B[n])); A, B converted rows with leading dimension
// ld (a multiple of 16 elements, zero padded past dim).
template <int PREC>
__global__ __launch_bounds__(256) void pkc_center_mfma(const void* __restrict__ A, const float* __restrict__ sa,
                                                       const void* __restrict__ B, const float* __restrict__ sb,
                                                       uint32_t M, uint32_t N, uint32_t ld, int epi,
                                                       const float* __restrict__ bias, float* __restrict__ Cout,
                                                       uint64_t ldc) {
    extern __shared__ __align__(16) unsigned char lds[];
    constexpr uint32_t ES = PREC == PKC_BF16 ? 2 : 1;  // bytes per element
    const uint32_t rowb = ld * ES + 16;                 // LDS row pitch (bytes), padded against bank conflicts
    unsigned char* As = lds;
    unsigned char* Bs = lds + 64 * rowb;
    const uint32_t m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    // stage 64 rows of A and B (16-byte chunks)
    const uint32_t cpr = ld * ES / 16;
    for (uint32_t e = tid; e < 64 * cpr; e += 256) {
        const uint32_t r = e / cpr, c = e % cpr;
        uint4 va = make_uint4(0, 0, 0, 0), vb = make_uint4(0, 0, 0, 0);
        if (m0 + r < M) va = *(const uint4*)((const unsigned char*)A + ((uint64_t)(m0 + r) * ld) * ES + c * 16);
        if (n0 + r < N) vb = *(const uint4*)((const unsigned char*)B + ((uint64_t)(n0 + r) * ld) * ES + c * 16);
        *(uint4*)(As + r * rowb + c * 16) = va;
        *(uint4*)(Bs + r * rowb + c * 16) = vb;
    }
    __syncthreads();
    pkc_v16f acc;
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    const unsigned char* arow = As + (wm * 32 + (lane & 31)) * rowb;
    const unsigned char* brow = Bs + (wn * 32 + (lane & 31)) * rowb;
    if (PREC == PKC_BF16) {
        // K = 8 per instruction; lane holds k = 8s + 4(lane/32) .. +3
        for (uint32_t k = 0; k < ld; k += 8) {
            const uint32_t kk = k + 4 * (lane >> 5);
            const pkc_v4s av = *(const pkc_v4s*)(arow + kk * 2);
            const pkc_v4s bv = *(const pkc_v4s*)(brow + kk * 2);
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
        }
    } else {
        // K = 16 per instruction; lane holds k = 16s + 8(lane/32) .. +7
        for (uint32_t k = 0; k < ld; k += 16) {
            const uint32_t kk = k + 8 * (lane >> 5);
            const long av = *(const long*)(arow + kk);
            const long bv = *(const long*)(brow + kk);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(av, bv, acc, 0, 0, 0);
        }
    }
    const uint32_t gn = n0 + wn * 32 + (lane & 31);
    if (gn >= N) return;
    const float bn = epi == PKC_SUB_FROM_BIAS ? bias[gn] : 0.f;
    const float sbn = PREC == PKC_FP8 ? sb[gn] : 1.f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const uint32_t gm = m0 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (gm >= M) continue;
        float v = acc[r];
        if (PREC == PKC_FP8) v *= sa[gm] * sbn;
        Cout[(uint64_t)gm * ldc + gn] = epi == PKC_SUB_FROM_BIAS ? bn - v : -v;
    }
}

// Exact rescoring of the coarse candidates: cand_id [n x kc] (approximate
// order) -> the keep smallest exact |c|^2/2 - <c,q> (sequential fmaf dot, the
// engine's exact order), ascending by (distance, id), into csd / csi (stride
// keep).  Rows with equal exact values among the first flag_lim (or at that
// boundary) set flags bit 0.
__global__ __launch_bounds__(256) void pkc_rescore_coarse(const float* __restrict__ q, uint32_t dim,
                                                          const float* __restrict__ coarse,
                                                          const float* __restrict__ cnorm,
                                                          const uint32_t* __restrict__ cand_id, uint32_t kc,
                                                          uint32_t keep, float* __restrict__ csd,
                                                          uint32_t* __restrict__ csi, uint8_t* __restrict__ flags,
                                                          uint32_t flag_lim) {
    __shared__ uint64_t keys[512];
    const uint64_t row = blockIdx.x;
    const float* qr = q + row * dim;
    uint32_t n2 = 2;
    while (n2 < kc) n2 <<= 1;
    for (uint32_t i = threadIdx.x; i < n2; i += blockDim.x) {
        uint64_t key = ~0ull;
        if (i < kc) {
            const uint32_t c = cand_id[row * kc + i];
            const float v = cnorm[c] - pk_dot(coarse + (uint64_t)c * dim, qr, dim);
            key = ((uint64_t)pk_f2ord(v) << 32) | c;
        }
        keys[i] = key;
    }
    __syncthreads();
    for (uint32_t k = 2; k <= n2; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < (n2 >> 1); i += blockDim.x) {
                const uint32_t lo = 2 * i - (i & (j - 1)), hi = lo + j;
                const bool up = (lo & k) == 0;
                const uint64_t a = keys[lo], b = keys[hi];
                if ((a > b) == up) {
                    keys[lo] = b;
                    keys[hi] = a;
                }
            }
            __syncthreads();
        }
    bool tie = false;
    for (uint32_t i = threadIdx.x; i < keep; i += blockDim.x) {
        csd[row * keep + i] = pk_ord2f((uint32_t)(keys[i] >> 32));
        csi[row * keep + i] = (uint32_t)keys[i];
        if (i < flag_lim && i + 1 < kc && (uint32_t)(keys[i + 1] >> 32) == (uint32_t)(keys[i] >> 32)) tie = true;
    }
    if (tie && flags) flags[row] = 1u | flags[row];
}

extern "C" uint32_t pkc_ld(uint32_t dim) { return (dim + 15) & ~15u; }

extern "C" hipError_t pkc_launch_convert(const float* x, uint32_t n, uint32_t dim, int prec, uint16_t* out16,
                                         uint8_t* out8, float* scale, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pkc_convert_rows, dim3((n + 3) / 4), dim3(256), 0, s, x, n, dim, prec, out16, out8, scale,
                       pkc_ld(dim));
    return hipGetLastError();
}

extern "C" hipError_t pkc_launch_center(const void* A, const float* sa, const void* B, const float* sb, uint32_t M,
                                        uint32_t N, uint32_t dim, int prec, int epi, const float* bias, float* C,
                                        uint64_t ldc, hipStream_t s) {
    if (M == 0 || N == 0) return hipSuccess;
    const uint32_t ld = pkc_ld(dim);
    if (ld > 256) return hipErrorInvalidValue;
    const size_t smem = 2 * 64 * (size_t)(ld * (prec == PKC_BF16 ? 2 : 1) + 16);
    dim3 grid((N + 63) / 64, (M + 63) / 64);
    if (prec == PKC_BF16)
        hipLaunchKernelGGL(pkc_center_mfma<PKC_BF16>, grid, dim3(256), smem, s, A, sa, B, sb, M, N, ld, epi, bias, C,
                           ldc);
    else
        hipLaunchKernelGGL(pkc_center_mfma<PKC_FP8>, grid, dim3(256), smem, s, A, sa, B, sb, M, N, ld, epi, bias, C,
                           ldc);
    return hipGetLastError();
}

extern "C" hipError_t pkc_launch_rescore_coarse(const float* q, uint32_t dim, const float* coarse, const float* cnorm,
                                                const uint32_t* cand_id, uint32_t kc, uint32_t keep, uint32_t n,
                                                float* csd, uint32_t* csi, uint8_t* flags, uint32_t flag_lim,
                                                hipStream_t s) {
    if (n == 0 || keep == 0) return hipSuccess;
    if (kc > 512 || keep > kc) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pkc_rescore_coarse, dim3(n), dim3(256), 0, s, q, dim, coarse, cnorm, cand_id, kc, keep, csd,
                       csi, flags, flag_lim);
    return hipGetLastError();
}
