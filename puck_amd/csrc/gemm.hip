// gemm.hip -- center-distance contraction C[m][n] = epi( sum_k A[m][k]*B[n][k] ).
//
// pk_gemm_exact: every output is a sequential fmaf chain over k = 0..K-1
// (the engine's exact order), tiled 64x64x16 through LDS, 4x4 outputs per
// thread.  Used for query x center distances in the exact search mode, for
// the C x F stationary table and for build-time assignment, where results
// must match the CPU restatement bit for bit.
//
// Epilogues: RAW = acc; SUB_FROM_BIAS = bias[n] - acc (coarse: |c|^2/2 - <q,c>,
// hierarchical_cluster_index.cpp:713-716); NEG = -acc (fine: puck_index.cpp:500);
// ADD_BIAS = bias[n] + acc (stationary: |f|^2/2 + <f,c>, hierarchical_cluster_index.cpp:318).
#include "common.h"

enum { PK_EPI_RAW = 0, PK_EPI_SUB_FROM_BIAS = 1, PK_EPI_NEG = 2, PK_EPI_ADD_BIAS = 3 };

#define GB_M 64
#define GB_N 64
#define GB_K 16

__global__ __launch_bounds__(256) void pk_gemm_exact(const float* __restrict__ A, const float* __restrict__ B,
                                                     float* __restrict__ Cout, uint32_t M, uint32_t N,
                                                     uint32_t K, uint64_t ldc, int epi,
                                                     const float* __restrict__ bias) {
    __shared__ float As[GB_K][GB_M + 4];
    __shared__ float Bs[GB_K][GB_N + 4];
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const uint32_t m0 = blockIdx.y * GB_M, n0 = blockIdx.x * GB_N;
    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;

    for (uint32_t k0 = 0; k0 < K; k0 += GB_K) {
        // 64 rows x 16 k of A and B: 1024 elements each, 4 per thread.
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int idx = tid + e * 256;
            int r = idx >> 4, kk = idx & 15;
            uint32_t gm = m0 + r, gn = n0 + r, gk = k0 + kk;
            As[kk][r] = (gm < M && gk < K) ? A[(uint64_t)gm * K + gk] : 0.0f;
            Bs[kk][r] = (gn < N && gk < K) ? B[(uint64_t)gn * K + gk] : 0.0f;
        }
        __syncthreads();
        const uint32_t kend = (K - k0) < GB_K ? (K - k0) : GB_K;
        for (uint32_t kk = 0; kk < kend; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        uint32_t gm = m0 + ty * 4 + i;
        if (gm >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            uint32_t gn = n0 + tx * 4 + j;
            if (gn >= N) continue;
            float v = acc[i][j];
            if (epi == PK_EPI_SUB_FROM_BIAS) v = bias[gn] - v;
            else if (epi == PK_EPI_NEG) v = -v;
            else if (epi == PK_EPI_ADD_BIAS) v = bias[gn] + v;
            Cout[(uint64_t)gm * ldc + gn] = v;
        }
    }
}

extern "C" hipError_t pk_launch_gemm_exact(const float* A, const float* B, float* C, uint32_t M, uint32_t N,
                                           uint32_t K, uint64_t ldc, int epi, const float* bias,
                                           hipStream_t s) {
    if (M == 0 || N == 0) return hipSuccess;
    dim3 grid((N + GB_N - 1) / GB_N, (M + GB_M - 1) / GB_M);
    hipLaunchKernelGGL(pk_gemm_exact, grid, dim3(256), 0, s, A, B, C, M, N, K, ldc, epi, bias);
    return hipGetLastError();
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
