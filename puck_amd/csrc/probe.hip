// probe.hip -- the learned probe: a small MLP that predicts, per query, how
// many fine clusters (cells of a coarse list, nearest first by cell
// distance) the cell walk needs to reach its top-k.  No counterpart in the
// reference; it feeds the fine budget of the batched search.
//
// Network: x[D] -> Z1 = W1 x + b1 -> H1 = GELU(Z1) -> Z2 = W2 H1 + b2 ->
// H2 = GELU(Z2) -> y = w3 . H2 + b3, GELU(z) = z/2 (1 + erf(z/sqrt 2)).
// Weights are used in BF16 (round-to-nearest-even copies of FP32 master
// weights), activations enter the GEMMs in BF16, every product is
// accumulated in FP32 on the matrix cores (v_mfma_f32_32x32x8_bf16).
//
// Features (D = dim + nc + nf): the normalised query, its nc smallest coarse
// distances |c|^2/2 - <c,q> and the nf smallest cell distances
// (|q-c-f|^2 - |q|^2)/2 = (cd + stationary) + fine distance of its nearest
// coarse list (the fine clusters of that list the walk scores first), each
// standardised with the probe's (mu, 1/sigma).
#include "common.h"

typedef short pk_v4s __attribute__((ext_vector_type(4)));
typedef float pk_v16f __attribute__((ext_vector_type(16)));

// FP32 -> BF16, round to nearest even (NaN kept quiet).
__device__ __forceinline__ uint16_t pk_f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float pk_bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float pk_gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float pk_gelu_grad(float z) {
    // d/dz [z Phi(z)] = Phi(z) + z phi(z)
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

// ---------------------------------------------------------------------------
// C[m, n] = sum_k A(m, k) B(n, k), BF16 operands, FP32 accumulation.
//   A(m, k) = AT ? A[k * lda + m] : A[m * lda + k]
//   B(n, k) = BT ? B[k * ldb + n] : B[n * ldb + k]
// 64 x 64 tile per workgroup of 4 waves (2 x 2 of 32 x 32), K staged through
// LDS 32 at a time, zero padded at the edges.
enum { MM_F32 = 0, MM_BIAS_GELU = 1, MM_GELU_BWD = 2 };

struct PkMmArgs {
    const uint16_t* A;
    const uint16_t* B;
    uint32_t lda, ldb, M, N, K;
    int epi;
    const float* bias;     // [N] (MM_BIAS_GELU)
    float* C;              // fp32 out [M x ldc] (MM_F32: C; MM_BIAS_GELU: Z, may be null)
    uint32_t ldc;
    uint16_t* Hout;        // bf16 out [M x ldc] (MM_BIAS_GELU: GELU(Z); MM_GELU_BWD: dZ)
    const float* Zin;      // [M x ldc] (MM_GELU_BWD: the layer's pre-activation)
};

#define MM_T 64
#define MM_KT 32
#define MM_LDS_STRIDE (MM_KT + 4)  // halves per LDS row (8-byte aligned rows)

template <bool AT, bool BT>
__global__ __launch_bounds__(256) void pk_mm_bf16(PkMmArgs a) {
    __shared__ __align__(16) uint16_t As[MM_T * MM_LDS_STRIDE];
    __shared__ __align__(16) uint16_t Bs[MM_T * MM_LDS_STRIDE];
    const uint32_t m0 = blockIdx.y * MM_T, n0 = blockIdx.x * MM_T;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    pk_v16f acc;
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    for (uint32_t k0 = 0; k0 < a.K; k0 += MM_KT) {
        // stage A and B tiles as [row][k]; every load at an in-bounds index
        // (clamped), then zeroed outside the matrix: loads under a lane
        // condition are waited for one at a time
        constexpr int PT = MM_T * MM_KT / 256;  // elements per thread and tile
        uint16_t va[PT], vb[PT];
#pragma unroll
        for (int u = 0; u < PT; ++u) {
            const int e = tid + u * 256;
            int r, k;
            if (AT) { r = e % MM_T; k = e / MM_T; } else { r = e / MM_KT; k = e % MM_KT; }
            const uint32_t gm = min(m0 + r, a.M - 1u), gk = min(k0 + k, a.K - 1u);
            va[u] = AT ? a.A[(uint64_t)gk * a.lda + gm] : a.A[(uint64_t)gm * a.lda + gk];
        }
#pragma unroll
        for (int u = 0; u < PT; ++u) {
            const int e = tid + u * 256;
            int r, k;
            if (BT) { r = e % MM_T; k = e / MM_T; } else { r = e / MM_KT; k = e % MM_KT; }
            const uint32_t gn = min(n0 + r, a.N - 1u), gk = min(k0 + k, a.K - 1u);
            vb[u] = BT ? a.B[(uint64_t)gk * a.ldb + gn] : a.B[(uint64_t)gn * a.ldb + gk];
        }
#pragma unroll
        for (int u = 0; u < PT; ++u) {
            const int e = tid + u * 256;
            int r, k;
            if (AT) { r = e % MM_T; k = e / MM_T; } else { r = e / MM_KT; k = e % MM_KT; }
            As[r * MM_LDS_STRIDE + k] = (m0 + r < a.M && k0 + k < a.K) ? va[u] : (uint16_t)0;
        }
#pragma unroll
        for (int u = 0; u < PT; ++u) {
            const int e = tid + u * 256;
            int r, k;
            if (BT) { r = e % MM_T; k = e / MM_T; } else { r = e / MM_KT; k = e % MM_KT; }
            Bs[r * MM_LDS_STRIDE + k] = (n0 + r < a.N && k0 + k < a.K) ? vb[u] : (uint16_t)0;
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < MM_KT / 8; ++s) {
            // lane l: row l%32 of the wave's A (B) sub-tile, k = 8s + 4(l/32) .. +3
            const int kk = s * 8 + 4 * (lane >> 5);
            const pk_v4s av = *(const pk_v4s*)(As + (wm * 32 + (lane & 31)) * MM_LDS_STRIDE + kk);
            const pk_v4s bv = *(const pk_v4s*)(Bs + (wn * 32 + (lane & 31)) * MM_LDS_STRIDE + kk);
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // acc[r] holds C[8 (r/4) + 4 (lane/32) + r%4][lane%32] of the wave tile
    const uint32_t gn = n0 + wn * 32 + (lane & 31);
    if (gn >= a.N) return;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const uint32_t gm = m0 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (gm >= a.M) continue;
        const uint64_t o = (uint64_t)gm * a.ldc + gn;
        float v = acc[r];
        if (a.epi == MM_F32) {
            a.C[o] = v;
        } else if (a.epi == MM_BIAS_GELU) {
            v += a.bias[gn];
            if (a.C) a.C[o] = v;
            a.Hout[o] = pk_f2bf(pk_gelu(v));
        } else {
            a.Hout[o] = pk_f2bf(v * pk_gelu_grad(a.Zin[o]));
        }
    }
}

static hipError_t pk_mm(bool at, bool bt, const PkMmArgs& a, hipStream_t s) {
    if (a.M == 0 || a.N == 0) return hipSuccess;
    dim3 grid((a.N + MM_T - 1) / MM_T, (a.M + MM_T - 1) / MM_T);
    if (!at && !bt) hipLaunchKernelGGL((pk_mm_bf16<false, false>), grid, dim3(256), 0, s, a);
    else if (!at && bt) hipLaunchKernelGGL((pk_mm_bf16<false, true>), grid, dim3(256), 0, s, a);
    else if (at && bt) hipLaunchKernelGGL((pk_mm_bf16<true, true>), grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((pk_mm_bf16<true, false>), grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Raw features of n queries from the search pipeline's sorted rows: the
// query, the nc smallest coarse distances (csd, stride cstride) and the nfine
// smallest cell distances of the nearest list (cell, stride nfine,
// ascending).  An empty cell's distance carries the stationary sentinel
// sqrt(FLT_MAX); such an entry (at or above half the sentinel) takes the
// row's largest populated-cell distance instead (0 for a list with no
// point), so that features stay in the range of the data.
#define PK_EMPTY_CELL_FEATURE 9.2233715e18f
__global__ void pk_probe_raw_features(const float* __restrict__ q, uint32_t dim, const float* __restrict__ csd,
                                      uint32_t cstride, uint32_t nc, const float* __restrict__ cell,
                                      uint32_t nfine, uint32_t n, float* __restrict__ out) {
    const uint32_t D = dim + nc + nfine;
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (uint64_t)n * D) return;
    const uint32_t i = (uint32_t)(e / D), j = (uint32_t)(e % D);
    float v;
    if (j < dim) v = q[(uint64_t)i * dim + j];
    else if (j < dim + nc) v = csd[(uint64_t)i * cstride + (j - dim)];
    else {
        const float* row = cell + (uint64_t)i * nfine;
        const uint32_t jj = j - dim - nc;
        v = row[jj];
        if (v >= PK_EMPTY_CELL_FEATURE) {
            v = 0.f;
            for (uint32_t t = jj; t-- > 0;)
                if (row[t] < PK_EMPTY_CELL_FEATURE) {
                    v = row[t];
                    break;
                }
        }
    }
    out[e] = v;
}

// Standardise raw features to the BF16 network input.
__global__ void pk_probe_standardize(const float* __restrict__ raw, uint32_t n, uint32_t D,
                                     const float* __restrict__ mu, const float* __restrict__ inv_sigma,
                                     uint16_t* __restrict__ x) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (uint64_t)n * D) return;
    const uint32_t j = (uint32_t)(e % D);
    x[e] = pk_f2bf((raw[e] - mu[j]) * inv_sigma[j]);
}

// y[b] = sum_j w3[j] H2[b, j] + b3 (one wave per row); with targets also the
// loss gradient dy = 2 (y - t) / n and the squared error.
__global__ void pk_probe_head(const uint16_t* __restrict__ H2, uint32_t H, const uint16_t* __restrict__ w3,
                              const float* __restrict__ b3, uint32_t n, const float* __restrict__ target,
                              float* __restrict__ y, float* __restrict__ dy, float* __restrict__ sqerr) {
    const uint32_t row = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    float s = 0.f;
    for (uint32_t j = lane; j < H; j += 64) s += pk_bf2f(w3[j]) * pk_bf2f(H2[(uint64_t)row * H + j]);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) {
        const float v = s + b3[0];
        if (y) y[row] = v;
        if (target) {
            const float d = v - target[row];
            dy[row] = 2.0f * d / (float)n;
            sqerr[row] = d * d;
        }
    }
}

// dZ2[b, j] = bf16(dy[b] w3[j] GELU'(Z2[b, j]))
__global__ void pk_probe_head_bwd(const float* __restrict__ dy, const uint16_t* __restrict__ w3,
                                  const float* __restrict__ Z2, uint32_t n, uint32_t H, uint16_t* __restrict__ dZ2) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (uint64_t)n * H) return;
    const uint32_t b = (uint32_t)(e / H), j = (uint32_t)(e % H);
    dZ2[e] = pk_f2bf(dy[b] * pk_bf2f(w3[j]) * pk_gelu_grad(Z2[e]));
}

// Column sums: out[j] = sum_b v(b, j), v BF16 (bias grads) or the w3 / b3
// gradient sum_b dy[b] H2[b, j] (mode 1) / sum_b dy[b] (mode 2).  One
// workgroup per column, fixed-order tree reduction.
__global__ __launch_bounds__(256) void pk_probe_colsum(const uint16_t* __restrict__ v, uint32_t n, uint32_t H,
                                                       const float* __restrict__ dy, int mode,
                                                       float* __restrict__ out) {
    __shared__ float red[256];
    const uint32_t j = blockIdx.x;
    float s = 0.f;
    for (uint32_t b = threadIdx.x; b < n; b += 256) {
        if (mode == 0) s += pk_bf2f(v[(uint64_t)b * H + j]);
        else if (mode == 1) s += dy[b] * pk_bf2f(v[(uint64_t)b * H + j]);
        else s += dy[b];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = red[0];
}

// Sum of n floats into out[0] (fixed-order tree, one workgroup).
__global__ __launch_bounds__(256) void pk_probe_sum(const float* __restrict__ v, uint32_t n, float* __restrict__ out) {
    __shared__ float red[256];
    float s = 0.f;
    for (uint32_t b = threadIdx.x; b < n; b += 256) s += v[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// Adam on the packed FP32 parameters (first `count` entries), then refresh
// the BF16 copies of the three weight blocks.
__global__ void pk_probe_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                              const float* __restrict__ g, uint32_t count, float lr, float b1, float b2, float eps,
                              float c1, float c2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / c1, vh = vi / c2;
    p[i] -= lr * mh / (sqrtf(vh) + eps);
}

__global__ void pk_to_bf16(const float* __restrict__ x, uint32_t n, uint16_t* __restrict__ y) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = pk_f2bf(x[i]);
}

// Fine budget from the prediction: ceil(exp(y + margin)) clamped to [min, F];
// F (the reference's walk) when the prediction is not finite or the budget
// would exceed fallback_frac * F.
__global__ void pk_probe_budget(const float* __restrict__ y, uint32_t n, float margin, uint32_t nf_min,
                                float fallback_frac, uint32_t F, uint32_t* __restrict__ nf) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float t = y[i] + margin;
    uint32_t b = F;
    if (isfinite(t)) {
        const float e = expf(fminf(t, 30.f));
        if (e < fallback_frac * (float)F) {
            const float c = ceilf(e);
            b = c < (float)nf_min ? nf_min : (c > (float)F ? F : (uint32_t)c);
        }
    }
    nf[i] = b;
}

// ---------------------------------------------------------------------------
// Host side.  Packed FP32 parameter layout (D inputs, H hidden):
//   W1 [H x D], b1 [H], W2 [H x H], b2 [H], w3 [H], b3 [1]   (trained)
//   mu [D], inv_sigma [D]                                    (feature scaling)
extern "C" {

uint32_t pk_probe_trained_count(uint32_t D, uint32_t H) { return H * D + H + H * H + H + H + 1; }
uint32_t pk_probe_param_count(uint32_t D, uint32_t H) { return pk_probe_trained_count(D, H) + 2 * D; }

struct PkProbeOffsets {
    uint32_t W1, b1, W2, b2, w3, b3, mu, isg, total, trained;
};

static PkProbeOffsets pk_probe_offsets(uint32_t D, uint32_t H) {
    PkProbeOffsets o;
    o.W1 = 0;
    o.b1 = o.W1 + H * D;
    o.W2 = o.b1 + H;
    o.b2 = o.W2 + H * H;
    o.w3 = o.b2 + H;
    o.b3 = o.w3 + H;
    o.trained = o.b3 + 1;
    o.mu = o.trained;
    o.isg = o.mu + D;
    o.total = o.isg + D;
    return o;
}

// Workspace of a forward (+ backward) pass over n rows.
struct PkProbeWork {
    uint16_t* X;    // n x D
    float* Z1;      // n x H
    uint16_t* H1;   // n x H
    float* Z2;      // n x H
    uint16_t* H2;   // n x H
    float* y;       // n
    float* dy;      // n
    float* sq;      // n
    uint16_t* dZ2;  // n x H
    uint16_t* dZ1;  // n x H
};

size_t pk_probe_work_bytes(uint32_t n, uint32_t D, uint32_t H) {
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    return al(2ull * n * D) + 2 * al(4ull * n * H) + 4 * al(2ull * n * H) + 3 * al(4ull * n);
}

static PkProbeWork pk_probe_carve(void* base, uint32_t n, uint32_t D, uint32_t H) {
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    char* p = (char*)base;
    PkProbeWork w;
    w.X = (uint16_t*)p; p += al(2ull * n * D);
    w.Z1 = (float*)p; p += al(4ull * n * H);
    w.Z2 = (float*)p; p += al(4ull * n * H);
    w.H1 = (uint16_t*)p; p += al(2ull * n * H);
    w.H2 = (uint16_t*)p; p += al(2ull * n * H);
    w.dZ2 = (uint16_t*)p; p += al(2ull * n * H);
    w.dZ1 = (uint16_t*)p; p += al(2ull * n * H);
    w.y = (float*)p; p += al(4ull * n);
    w.dy = (float*)p; p += al(4ull * n);
    w.sq = (float*)p; p += al(4ull * n);
    return w;
}

hipError_t pk_launch_probe_raw_features(const float* q, uint32_t dim, const float* csd, uint32_t cstride, uint32_t nc,
                                        const float* cell, uint32_t nfine, uint32_t n, float* out, hipStream_t s) {
    const uint64_t tot = (uint64_t)n * (dim + nc + nfine);
    if (tot == 0) return hipSuccess;
    hipLaunchKernelGGL(pk_probe_raw_features, dim3((uint32_t)((tot + 255) / 256)), dim3(256), 0, s, q, dim, csd,
                       cstride, nc, cell, nfine, n, out);
    return hipGetLastError();
}

hipError_t pk_launch_to_bf16(const float* x, uint32_t n, uint16_t* y, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pk_to_bf16, dim3((n + 255) / 256), dim3(256), 0, s, x, n, y);
    return hipGetLastError();
}

// Forward pass over n rows of raw features.  params: packed FP32; wbf: BF16
// copy of the packed trained block.  With target != null also the backward
// pass: grads (packed trained layout) and loss_sum[0] = sum of squared errors.
hipError_t pk_probe_pass(const float* params, const uint16_t* wbf, uint32_t D, uint32_t H, uint32_t n,
                         const float* raw, const float* target, void* work, float* grads, float* loss_sum,
                         float* y_out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const PkProbeOffsets o = pk_probe_offsets(D, H);
    PkProbeWork w = pk_probe_carve(work, n, D, H);
    const uint64_t nd = (uint64_t)n * D, nh = (uint64_t)n * H;
    hipLaunchKernelGGL(pk_probe_standardize, dim3((uint32_t)((nd + 255) / 256)), dim3(256), 0, s, raw, n, D,
                       params + o.mu, params + o.isg, w.X);
    hipError_t e;
    PkMmArgs a{};
    // Z1 = X W1^T + b1, H1 = GELU(Z1)
    a.A = w.X; a.lda = D; a.B = wbf + o.W1; a.ldb = D; a.M = n; a.N = H; a.K = D;
    a.epi = MM_BIAS_GELU; a.bias = params + o.b1; a.C = target ? w.Z1 : nullptr; a.ldc = H; a.Hout = w.H1;
    if ((e = pk_mm(false, false, a, s)) != hipSuccess) return e;
    // Z2 = H1 W2^T + b2, H2 = GELU(Z2)
    a.A = w.H1; a.lda = H; a.B = wbf + o.W2; a.ldb = H; a.M = n; a.N = H; a.K = H;
    a.bias = params + o.b2; a.C = target ? w.Z2 : nullptr; a.Hout = w.H2;
    if ((e = pk_mm(false, false, a, s)) != hipSuccess) return e;
    float* y = y_out ? y_out : w.y;
    hipLaunchKernelGGL(pk_probe_head, dim3((n + 3) / 4), dim3(256), 0, s, w.H2, H, wbf + o.w3, params + o.b3, n,
                       target, y, w.dy, w.sq);
    if (!target) return hipGetLastError();
    // backward
    hipLaunchKernelGGL(pk_probe_sum, dim3(1), dim3(256), 0, s, w.sq, n, loss_sum);
    hipLaunchKernelGGL(pk_probe_colsum, dim3(H), dim3(256), 0, s, w.H2, n, H, w.dy, 1, grads + o.w3);
    hipLaunchKernelGGL(pk_probe_colsum, dim3(1), dim3(256), 0, s, w.H2, n, H, w.dy, 2, grads + o.b3);
    hipLaunchKernelGGL(pk_probe_head_bwd, dim3((uint32_t)((nh + 255) / 256)), dim3(256), 0, s, w.dy, wbf + o.w3,
                       w.Z2, n, H, w.dZ2);
    // gW2 = dZ2^T H1, gb2 = colsum(dZ2)
    a = PkMmArgs{};
    a.A = w.dZ2; a.lda = H; a.B = w.H1; a.ldb = H; a.M = H; a.N = H; a.K = n; a.epi = MM_F32;
    a.C = grads + o.W2; a.ldc = H;
    if ((e = pk_mm(true, true, a, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(pk_probe_colsum, dim3(H), dim3(256), 0, s, w.dZ2, n, H, nullptr, 0, grads + o.b2);
    // dZ1 = bf16((dZ2 W2) * GELU'(Z1))
    a = PkMmArgs{};
    a.A = w.dZ2; a.lda = H; a.B = wbf + o.W2; a.ldb = H; a.M = n; a.N = H; a.K = H; a.epi = MM_GELU_BWD;
    a.ldc = H; a.Hout = w.dZ1; a.Zin = w.Z1;
    if ((e = pk_mm(false, true, a, s)) != hipSuccess) return e;
    // gW1 = dZ1^T X, gb1 = colsum(dZ1)
    a = PkMmArgs{};
    a.A = w.dZ1; a.lda = H; a.B = w.X; a.ldb = D; a.M = H; a.N = D; a.K = n; a.epi = MM_F32;
    a.C = grads + o.W1; a.ldc = D;
    if ((e = pk_mm(true, true, a, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(pk_probe_colsum, dim3(H), dim3(256), 0, s, w.dZ1, n, H, nullptr, 0, grads + o.b1);
    return hipGetLastError();
}

hipError_t pk_probe_adam_step(float* params, float* m, float* v, const float* grads, uint16_t* wbf, uint32_t D,
                              uint32_t H, float lr, float b1, float b2, float eps, uint32_t step, hipStream_t s) {
    const uint32_t cnt = pk_probe_trained_count(D, H);
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(pk_probe_adam, dim3((cnt + 255) / 256), dim3(256), 0, s, params, m, v, grads, cnt, lr, b1,
                       b2, eps, c1, c2);
    hipLaunchKernelGGL(pk_to_bf16, dim3((cnt + 255) / 256), dim3(256), 0, s, params, cnt, wbf);
    return hipGetLastError();
}

hipError_t pk_launch_probe_budget(const float* y, uint32_t n, float margin, uint32_t nf_min, float fallback_frac,
                                  uint32_t F, uint32_t* nf, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pk_probe_budget, dim3((n + 255) / 256), dim3(256), 0, s, y, n, margin, nf_min,
                       fallback_frac, F, nf);
    return hipGetLastError();
}

}  // extern "C"
