// ref_driver.cpp -- C entry points over the two reference sources that compile
// on their own: puck/puck/distances_simd.cpp (faiss fvec_L2sqr_ny, the LUT
// kernels) and puck/tinker/method/hnsw_distfunc_opt_impl_inline.h (nmslib
// L2Sqr16Ext / L2SqrExt, the rerank distance).  Built by oracle/Makefile from
// the sources where they lie in the reference checkout, with the reference's
// own flags (-Ofast -march=native, CMakeLists.txt:16-22), into oracle/_ref/.
// TEST INFRASTRUCTURE ONLY: used by tests to pin the oracle's arithmetic.
#include <cstddef>
#include "puck/tinker/method/hnsw_distfunc_opt_impl_inline.h"

namespace faiss {
void fvec_L2sqr_ny(float* dis, const float* x, const float* y, size_t d, size_t ny);
}

extern "C" {
// Quantization::get_dist_table body (quantization.cpp:260-270) for one query.
void ref_dist_table(const float* q, const float* codebook, unsigned nsq, unsigned lsq, float* lut) {
    for (unsigned m = 0; m < nsq; ++m)
        faiss::fvec_L2sqr_ny(lut + m * 256u, q + m * lsq, codebook + (size_t)m * 256u * lsq, lsq, 256u);
}

// rank_topN_points exact distance loop (puck_index.cpp:428-453) for one row.
float ref_rerank_dist(const float* x, const float* q, unsigned dim) {
    alignas(32) float tmp[8];
    size_t qty16 = 16;
    float t = 0.0f;
    unsigned m = 0;
    for (; m + 16 <= dim;) {
        t += similarity::L2Sqr16Ext(x + m, q + m, qty16, tmp);
        m += 16;
    }
    size_t left = dim - m;
    if (left > 0) t += similarity::L2SqrExt(x + m, q + m, left, tmp);
    return t;
}
}
