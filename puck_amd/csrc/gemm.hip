// gemm.hip -- center-distance contraction C[m][n] = epi( sum_k A[m][k]*B[n][k] ).
//
// pk_gemm_exact: every output is a sequential fmaf chain over k = 0..K-1
// (the engine's exact order), tiled TM x TN x 16 through LDS.  Used for query x center distances in the exact search mode, for
// the C x F stationary table and for build-time assignment, where results
// must match the CPU restatement bit for bit.
//
// Epilogues: RAW = acc; SUB_FROM_BIAS = bias[n] - acc (coarse: |c|^2/2 - <q,c>,
// hierarchical_cluster_index.cpp:713-716); NEG = -acc (fine: puck_index.cpp:500);
// ADD_BIAS = bias[n] + acc (stationary: |f|^2/2 + <f,c>, hierarchical_cluster_index.cpp:318).
#include "common.h"

enum { PK_EPI_RAW = 0, PK_EPI_SUB_FROM_BIAS = 1, PK_EPI_NEG = 2, PK_EPI_ADD_BIAS = 3 };

#define GB_K 16

// Tile TM x TN per workgroup of 16 x 16 threads, (TM/16) x (TN/16) outputs per
// thread.  The tile shape changes only which thread owns an output, never the
// order of its fmaf chain, so every shape gives the same bits.
template <int TM, int TN>
__global__ __launch_bounds__(256) void pk_gemm_exact(const float* __restrict__ A, const float* __restrict__ B,
                                                     float* __restrict__ Cout, uint32_t M, uint32_t N,
                                                     uint32_t K, uint64_t ldc, int epi,
                                                     const float* __restrict__ bias) {
    constexpr int RM = TM / 16, RN = TN / 16;
    __shared__ float As[GB_K][TM + 4];
    __shared__ float Bs[GB_K][TN + 4];
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const uint32_t m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    float acc[RM][RN];
#pragma unroll
    for (int i = 0; i < RM; ++i)
#pragma unroll
        for (int j = 0; j < RN; ++j) acc[i][j] = 0.0f;

    constexpr int PA = TM * GB_K / 256, PB = TN * GB_K / 256;  // tile elements per thread
    for (uint32_t k0 = 0; k0 < K; k0 += GB_K) {
        // every load at a clamped in-bounds index, issued before any is used
        // (loads under a lane condition are waited for one at a time)
        float va[PA], vb[PB];
#pragma unroll
        for (int u = 0; u < PA; ++u) {
            const int idx = tid + u * 256, r = idx >> 4, kk = idx & 15;
            va[u] = A[(uint64_t)min(m0 + r, M - 1u) * K + min(k0 + kk, K - 1u)];
        }
#pragma unroll
        for (int u = 0; u < PB; ++u) {
            const int idx = tid + u * 256, r = idx >> 4, kk = idx & 15;
            vb[u] = B[(uint64_t)min(n0 + r, N - 1u) * K + min(k0 + kk, K - 1u)];
        }
#pragma unroll
        for (int u = 0; u < PA; ++u) {
            const int idx = tid + u * 256, r = idx >> 4, kk = idx & 15;
            As[kk][r] = (m0 + r < M && k0 + kk < K) ? va[u] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < PB; ++u) {
            const int idx = tid + u * 256, r = idx >> 4, kk = idx & 15;
            Bs[kk][r] = (n0 + r < N && k0 + kk < K) ? vb[u] : 0.0f;
        }
        __syncthreads();
        const uint32_t kend = (K - k0) < GB_K ? (K - k0) : GB_K;
        for (uint32_t kk = 0; kk < kend; ++kk) {
            float a[RM], b[RN];
#pragma unroll
            for (int i = 0; i < RM; ++i) a[i] = As[kk][ty * RM + i];
#pragma unroll
            for (int j = 0; j < RN; ++j) b[j] = Bs[kk][tx * RN + j];
#pragma unroll
            for (int i = 0; i < RM; ++i)
#pragma unroll
                for (int j = 0; j < RN; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < RM; ++i) {
        uint32_t gm = m0 + ty * RM + i;
        if (gm >= M) continue;
#pragma unroll
        for (int j = 0; j < RN; ++j) {
            uint32_t gn = n0 + tx * RN + j;
            if (gn >= N) continue;
            float v = acc[i][j];
            if (epi == PK_EPI_SUB_FROM_BIAS) v = bias[gn] - v;
            else if (epi == PK_EPI_NEG) v = -v;
            else if (epi == PK_EPI_ADD_BIAS) v = bias[gn] + v;
            Cout[(uint64_t)gm * ldc + gn] = v;
        }
    }
}

// Tile shapes: 0 = 64 x 64 (default), 1 = 16 x 128, 2 = 32 x 128, 3 = 128 x 64.
extern "C" uint32_t pk_gemm_shapes(void) { return 4; }

extern "C" hipError_t pk_launch_gemm_exact_shape(const float* A, const float* B, float* C, uint32_t M, uint32_t N,
                                                 uint32_t K, uint64_t ldc, int epi, const float* bias, int shape,
                                                 hipStream_t s) {
    if (M == 0 || N == 0) return hipSuccess;
    switch (shape) {
        case 1: {
            dim3 g((N + 127) / 128, (M + 15) / 16);
            hipLaunchKernelGGL((pk_gemm_exact<16, 128>), g, dim3(256), 0, s, A, B, C, M, N, K, ldc, epi, bias);
            break;
        }
        case 2: {
            dim3 g((N + 127) / 128, (M + 31) / 32);
            hipLaunchKernelGGL((pk_gemm_exact<32, 128>), g, dim3(256), 0, s, A, B, C, M, N, K, ldc, epi, bias);
            break;
        }
        case 3: {
            dim3 g((N + 63) / 64, (M + 127) / 128);
            hipLaunchKernelGGL((pk_gemm_exact<128, 64>), g, dim3(256), 0, s, A, B, C, M, N, K, ldc, epi, bias);
            break;
        }
        default: {
            dim3 g((N + 63) / 64, (M + 63) / 64);
            hipLaunchKernelGGL((pk_gemm_exact<64, 64>), g, dim3(256), 0, s, A, B, C, M, N, K, ldc, epi, bias);
        }
    }
    return hipGetLastError();
}

extern "C" hipError_t pk_launch_gemm_exact(const float* A, const float* B, float* C, uint32_t M, uint32_t N,
                                           uint32_t K, uint64_t ldc, int epi, const float* bias,
                                           hipStream_t s) {
    return pk_launch_gemm_exact_shape(A, B, C, M, N, K, ldc, epi, bias, 0, s);
}

// Half squared norms v[i] = dot(x_i, x_i) / 2 (or the plain dot when half == 0).
__global__ void pk_row_norms(const float* __restrict__ X, uint32_t n, uint32_t dim, int half,
                             float* __restrict__ out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* x = X + (uint64_t)i * dim;
    float v = pk_dot(x, x, dim);
    out[i] = half ? v / 2 : v;
}

extern "C" hipError_t pk_launch_row_norms(const float* X, uint32_t n, uint32_t dim, int half, float* out,
                                          hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pk_row_norms, dim3((n + 255) / 256), dim3(256), 0, s, X, n, dim, half, out);
    return hipGetLastError();
}
