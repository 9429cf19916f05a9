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
                                                             uint8_t* __restrict__ flags, uint32_t flag_lim) {
    extern __shared__ uint64_t sk[];
    const uint64_t row = blockIdx.x;
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
                                          uint32_t flag_lim, hipStream_t s) {
    if (rows == 0 || keep == 0) return hipSuccess;
    if (L > 8192 || keep > L) return hipErrorInvalidValue;
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void*)pk_sort_rows, hipFuncAttributeMaxDynamicSharedMemorySize, 64 * 1024);
        attr = true;
    }
    uint32_t n2 = 2;
    while (n2 < L) n2 <<= 1;
    hipLaunchKernelGGL(pk_sort_rows, dim3(rows), dim3(SORT_THREADS), n2 * sizeof(uint64_t), s, vals, L, ld_in,
                       keep, out_val, out_idx, n2, flags, flag_lim);
    return hipGetLastError();
}
