// exact.hip -- exact brute-force top-k (the ground truth of recall and the
// labels of the learned probe; no counterpart in the reference's search path,
// ann-benchmarks computes it on the CPU): per query chunk and base chunk, the
// engine's exact GEMM gives |x|^2/2 - <q,x>, the row top-k kernel keeps
// k + extra candidates per chunk, the candidates of all chunks are merged to
// the 2 (k + extra) smallest, and this file's pk_exact_recheck orders those by
// their float64 squared distance (ties by id) and writes the top k.
//
//   pk_u8_to_f32       uint8 rows to float rows (integer features)
//   pk_exact_append    a chunk's candidates into the query's candidate row
//                      (global ids)
//   pk_exact_recheck   float64 distances of the merged candidates, ordered by
//                      (distance, id), top k written
#include "common.h"

__global__ void pk_u8_to_f32(const uint8_t* __restrict__ src, uint64_t n, float* __restrict__ dst) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = (float)src[i];
}

extern "C" hipError_t pk_launch_u8_to_f32(const uint8_t* src, uint64_t n, float* dst, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(pk_u8_to_f32, dim3(blocks < 65536 ? (uint32_t)blocks : 65536u), dim3(256), 0, s, src, n, dst);
    return hipGetLastError();
}

// cand_val / cand_id [rows x stride]: columns [col0, col0 + kk) receive the
// chunk's kk candidates (values and base + column).
__global__ void pk_exact_append(const float* __restrict__ val, const uint32_t* __restrict__ idx, uint32_t rows,
                                uint32_t kk, uint32_t base, float* __restrict__ cand_val,
                                uint32_t* __restrict__ cand_id, uint32_t stride, uint32_t col0) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * kk) return;
    const uint32_t r = i / kk, j = i - r * kk;
    cand_val[(uint64_t)r * stride + col0 + j] = val[i];
    cand_id[(uint64_t)r * stride + col0 + j] = base + idx[i];
}

extern "C" hipError_t pk_launch_exact_append(const float* val, const uint32_t* idx, uint32_t rows, uint32_t kk,
                                             uint32_t base, float* cand_val, uint32_t* cand_id, uint32_t stride,
                                             uint32_t col0, hipStream_t s) {
    if (rows == 0 || kk == 0) return hipSuccess;
    const uint32_t n = rows * kk;
    hipLaunchKernelGGL(pk_exact_append, dim3((n + 255) / 256), dim3(256), 0, s, val, idx, rows, kk, base, cand_val,
                       cand_id, stride, col0);
    return hipGetLastError();
}

// One workgroup per query: M <= 1024 candidates (positions pos[] into the
// query's candidate row), their float64 squared distances to the query, the
// rank of each by (distance, id) counted against all others, the first k
// written.  Rows of T (float or uint8_t).
#define PK_EXACT_M 1024
template <typename T>
__global__ __launch_bounds__(256) void pk_exact_recheck(const T* __restrict__ x, uint32_t dim,
                                                        const float* __restrict__ q, const uint32_t* __restrict__ pos,
                                                        const uint32_t* __restrict__ cand_id, uint32_t stride,
                                                        uint32_t M, uint32_t k, uint32_t* __restrict__ out_id,
                                                        double* __restrict__ out_d) {
    __shared__ double dv[PK_EXACT_M];
    __shared__ uint32_t iv[PK_EXACT_M];
    const uint32_t r = blockIdx.x;
    const float* qr = q + (uint64_t)r * dim;
    for (uint32_t j = threadIdx.x; j < M; j += blockDim.x) {
        const uint32_t id = cand_id[(uint64_t)r * stride + pos[(uint64_t)r * M + j]];
        const T* xr = x + (uint64_t)id * dim;
        double s = 0.0;
        for (uint32_t e = 0; e < dim; ++e) {
            const double d = (double)xr[e] - (double)qr[e];
            s = fma(d, d, s);
        }
        dv[j] = s;
        iv[j] = id;
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < M; j += blockDim.x) {
        const double d = dv[j];
        const uint32_t id = iv[j];
        uint32_t rank = 0;
        for (uint32_t o = 0; o < M; ++o) rank += (dv[o] < d || (dv[o] == d && iv[o] < id)) ? 1u : 0u;
        if (rank < k) {
            out_id[(uint64_t)r * k + rank] = id;
            out_d[(uint64_t)r * k + rank] = d;
        }
    }
}

extern "C" hipError_t pk_launch_exact_recheck(const void* x, int u8, uint32_t dim, const float* q, uint32_t rows,
                                              const uint32_t* pos, const uint32_t* cand_id, uint32_t stride,
                                              uint32_t M, uint32_t k, uint32_t* out_id, double* out_d,
                                              hipStream_t s) {
    if (rows == 0) return hipSuccess;
    if (M > PK_EXACT_M || k > M) return hipErrorInvalidValue;
    if (u8)
        hipLaunchKernelGGL(pk_exact_recheck<uint8_t>, dim3(rows), dim3(256), 0, s, (const uint8_t*)x, dim, q, pos,
                           cand_id, stride, M, k, out_id, out_d);
    else
        hipLaunchKernelGGL(pk_exact_recheck<float>, dim3(rows), dim3(256), 0, s, (const float*)x, dim, q, pos,
                           cand_id, stride, M, k, out_id, out_d);
    return hipGetLastError();
}
