// build.hip -- index construction and load-time derivations on gfx950.
//
//   pk_assign_cells     nearest_cell_assign (hierarchical_cluster_index.cpp:1042-1085)
//                       on the inner products of pk_gemm_exact; one wave per point
//   pk_encode_filter    filter-PQ codes and offsets (puck_index.cpp:602-709 with
//                       nearest_center, kmeans.cpp:221-258, and
//                       set_static_value_of_formula, quantization.cpp:244-258)
//   pk_kmeans_argmin    nearest_center's argmin on a GEMM chunk (kmeans.cpp:240-254)
//   pk_kmeans_short     nearest_center for points of 1, 2 or 4 floats with MKL's
//                       order for those dot products (the filter-PQ codebook k-means)
//   pk_finalize_cells   empty-cell sentinel and per-coarse min_dist_offset
//                       (puck_index.cpp:287-306)
//   pk_gather_*         permutation of records / features into memory order
#include "common.h"
#include <float.h>

// ---------------------------------------------------------------------------
// nearest_cell_assign on one point per wave.
//   ct_c  = (|c|^2/2 - <x,c>) + |x|^2/2,  coarse visited in (ct, c) order
//   score = (ct - <x,f>) + stationary[c][f], fine in stored order, strict '<'
//   break when sqrt(best) - fbound[f] < sqrt(ct)   (fbound = -sqrt(|f|^2/2))
// Coarse clusters whose f = 0 check already breaks under the best found in
// the nearest coarse cluster are never visited by the reference either; the
// remaining ones are replayed exactly in (ct, c) order.
__device__ __forceinline__ void pk_wave_argmin(float& v, uint32_t& idx) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        float ov = __shfl_xor(v, o, 64);
        uint32_t oi = __shfl_xor(idx, o, 64);
        if (ov < v || (ov == v && oi < idx)) { v = ov; idx = oi; }
    }
}

// Process one coarse cluster: returns updated (best, cell).  A wave covers
// 4 * 64 consecutive fine centers per step: each lane scores 4 of them, an
// exclusive wave scan of the lane minima gives the best-so-far seen by every
// position, and the first position whose triangle check fails ends the
// cluster.  The check is skipped for a whole step when even the smallest
// best-so-far of the step cannot fail it at the step's last position.
__device__ void pk_assign_one_coarse(uint32_t c, float ct, const float* __restrict__ ipf, const float* __restrict__ st,
                                     const float* __restrict__ fbound, uint32_t F, int fb_sorted, float& best,
                                     int32_t& cell) {
    const int lane = threadIdx.x & 63;
    const float sct = __fsqrt_rn(ct);
    const float* strow = st + (uint64_t)c * F;
    for (uint32_t base = 0; base < F; base += 256) {
        const uint32_t f0 = base + 4 * lane;
        float sc[4];
        bool val[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t f = f0 + e;
            val[e] = f < F;
            sc[e] = val[e] ? (ct - ipf[f]) + strow[f] : FLT_MAX;
        }
        float lmin = sc[0];
#pragma unroll
        for (int e = 1; e < 4; ++e) lmin = sc[e] < lmin ? sc[e] : lmin;
        float incl = lmin;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            float y = __shfl_up(incl, o, 64);
            if (lane >= o) incl = y < incl ? y : incl;
        }
        const float stepmin = __shfl(incl, 63, 64);
        const float lo = stepmin < best ? stepmin : best;
        const uint32_t flast = (base + 256 <= F ? base + 256 : F) - 1;
        // sqrt(run) - fbound[f] >= sqrt(lo) - fbound[flast] for every position
        // of the step (run >= lo, fbound non-decreasing along the stored order)
        // (needs fbound non-decreasing: fine centers stored by descending norm)
        const bool nobreak = fb_sorted && !(__fsqrt_rn(lo) - fbound[flast] < sct);
        uint32_t bpos = 256u;
        if (!nobreak) {
            float excl = __shfl_up(incl, 1, 64);
            float run = lane == 0 ? best : (excl < best ? excl : best);
            int bl = 4;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if (bl == 4 && val[e] && (__fsqrt_rn(run) - fbound[f0 + e] < sct)) bl = e;
                run = sc[e] < run ? sc[e] : run;
            }
            uint64_t bm = __ballot(bl < 4);
            uint32_t blane = bm ? (uint32_t)__ffsll((unsigned long long)bm) - 1 : 64u;
            bpos = blane < 64 ? 4 * blane + (uint32_t)__shfl(bl, (int)blane, 64) : 256u;
        }
        if (stepmin < best || bpos < 256) {
            float v = FLT_MAX;
            uint32_t vi = 0xffffffffu;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                uint32_t pos = 4 * lane + e;
                if (val[e] && pos < bpos && (vi == 0xffffffffu || sc[e] < v)) { v = sc[e]; vi = pos; }
            }
            pk_wave_argmin(v, vi);
            if (vi != 0xffffffffu && v < best) {
                best = v;
                cell = (int32_t)(c * F + base + vi);
            }
        }
        if (bpos < 256) return;
    }
}

__device__ __forceinline__ void pk_wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One point per wave.  The coarse terms are sorted once per point (wave-level
// bitonic sort of (ct, c) keys in LDS); coarse clusters are visited in that
// order until one fails the triangle check at f = 0 -- every later one fails
// it too, because sqrt(ct) does not decrease and best does not increase.
__global__ __launch_bounds__(256) void pk_assign_cells(const float* __restrict__ ipc, const float* __restrict__ ipf,
                                                       const float* __restrict__ xn, const float* __restrict__ cn,
                                                       const float* __restrict__ fbound,
                                                       const float* __restrict__ st, uint32_t n, uint32_t C,
                                                       uint32_t F, uint32_t C2, int fb_sorted,
                                                       const uint32_t* __restrict__ order,
                                                       uint32_t* __restrict__ cell_out,
                                                       float* __restrict__ dist_out) {
    extern __shared__ __align__(16) unsigned char asmem[];
    uint64_t* keys = (uint64_t*)asmem + (threadIdx.x >> 6) * C2;
    const uint32_t wpt = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (wpt >= n) return;
    const uint32_t pt = order ? order[wpt] : wpt;
    const float* pc = ipc + (uint64_t)pt * C;
    const float* pf = ipf + (uint64_t)pt * F;
    const float x2 = xn[pt];
    for (uint32_t c = lane; c < C2; c += 64) {
        uint64_t k = ~0ull;
        if (c < C) k = ((uint64_t)pk_f2ord((cn[c] - pc[c]) + x2) << 32) | c;
        keys[c] = k;
    }
    pk_wave_lds_sync();
    for (uint32_t k = 2; k <= C2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = lane; i < (C2 >> 1); i += 64) {
                uint32_t lo = 2 * i - (i & (j - 1));
                uint32_t hi = lo + j;
                bool up = (lo & k) == 0;
                uint64_t x = keys[lo], y = keys[hi];
                if ((x > y) == up) {
                    keys[lo] = y;
                    keys[hi] = x;
                }
            }
            pk_wave_lds_sync();
        }
    }
    float best = FLT_MAX;
    int32_t cell = -1;
    const float fb0 = fbound[0];
    for (uint32_t r = 0; r < C; ++r) {
        const uint64_t key = keys[r];
        const float ct = pk_ord2f((uint32_t)(key >> 32));
        if (__fsqrt_rn(best) - fb0 < __fsqrt_rn(ct)) break;
        pk_assign_one_coarse((uint32_t)(key & 0xffffffffu), ct, pf, st, fbound, F, fb_sorted, best, cell);
    }
    if (lane == 0) {
        cell_out[pt] = (uint32_t)cell;
        if (dist_out) dist_out[pt] = best;
    }
}

// ---------------------------------------------------------------------------
// nearest_cell_assign without the sequential walk over all coarse clusters.
//
// The reference visits coarse clusters in (ct, c) order and keeps the first
// strict minimum of score = (ct - <x,f>) + stationary over the cells it visits;
// a cell is skipped when the triangle check fails (break) under the running
// best.  Its answer is reproduced exactly in three steps per point:
//   1. the nearest coarse cluster c0 (lexicographic (ct, c) minimum) is
//      scanned exactly as the reference scans its first list (from best =
//      FLT_MAX) -> best B and cell;
//   2. every other cell whose float score could be <= B is evaluated: cells
//      are skipped in blocks of PK_AB fine centers (stored order) when the
//      block's lower bound ct + min st[c][block] - max <x,f in block> exceeds
//      B by more than the rounding of either expression (2^-18 * (|ct| +
//      max|st| + max|<x,f>|), far above the 4u bound of two float additions),
//      and whole lists on the row minimum first;
//   3. the lexicographic (score, ct, c, f) minimum over the evaluated cells is
//      the reference's answer when the reference visits that cell, which the
//      certificate sqrt(score) - max(fbound[0..f]) >= sqrt(ct) guarantees: the
//      running best never drops below that score (no visited cell scores lower),
//      so no check up to f can break.  A point failing the certificate (or
//      with a negative score) is listed and assigned by pk_assign_cells
//      (list_all lists every point: the sequential path's test).
// Every cell with a smaller score, or an equal score earlier in the visiting
// order, is evaluated, so the minimum is the first one the reference meets.
#define PK_AB 32u
#define PK_AB_NBMAX 256u  // blocks per row: F <= 8192

// Per coarse row of the stationary table: the minimum of each block of PK_AB
// fine centers, the row minimum, and the row's largest magnitude (folded into
// *stabs_bits with an unsigned max on the float bits of |st|).
__global__ void pk_st_blocks(const float* __restrict__ st, uint32_t C, uint32_t F, uint32_t nb,
                             float* __restrict__ stmin, float* __restrict__ stminrow, uint32_t* __restrict__ stabs_bits) {
    const uint32_t c = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (c >= C) return;
    const float* row = st + (uint64_t)c * F;
    float rmin = FLT_MAX, amax = 0.0f;
    for (uint32_t j = lane; j < nb; j += 64) {
        float m = FLT_MAX;
        const uint32_t f1 = (j + 1) * PK_AB < F ? (j + 1) * PK_AB : F;
        for (uint32_t f = j * PK_AB; f < f1; ++f) {
            const float v = row[f];
            m = v < m ? v : m;
            const float a = fabsf(v);
            amax = a > amax ? a : amax;
        }
        stmin[(uint64_t)c * nb + j] = m;
        rmin = m < rmin ? m : rmin;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float r2 = __shfl_xor(rmin, o, 64), a2 = __shfl_xor(amax, o, 64);
        rmin = r2 < rmin ? r2 : rmin;
        amax = a2 > amax ? a2 : amax;
    }
    if (lane == 0) {
        stminrow[c] = rmin;
        atomicMax(stabs_bits, __float_as_uint(amax));
    }
}

extern "C" hipError_t pk_launch_st_blocks(const float* st, uint32_t C, uint32_t F, float* stmin, float* stminrow,
                                          uint32_t* stabs_bits, hipStream_t s) {
    const uint32_t nb = (F + PK_AB - 1) / PK_AB;
    if (nb > PK_AB_NBMAX) return hipErrorInvalidValue;
    hipError_t e = hipMemsetAsync(stabs_bits, 0, sizeof(uint32_t), s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(pk_st_blocks, dim3((C + 3) / 4), dim3(256), 0, s, st, C, F, nb, stmin, stminrow, stabs_bits);
    return hipGetLastError();
}

struct PkCellKey {
    float v, ct;
    uint32_t c, f;  // c = 0xffffffff: none
};

__device__ __forceinline__ bool pk_key_less(const PkCellKey& a, const PkCellKey& b) {
    if (b.c == 0xffffffffu) return a.c != 0xffffffffu;
    if (a.c == 0xffffffffu) return false;
    if (a.v != b.v) return a.v < b.v;
    if (a.ct != b.ct) return a.ct < b.ct;
    if (a.c != b.c) return a.c < b.c;
    return a.f < b.f;
}

__global__ __launch_bounds__(256) void pk_assign_bounded(
    const float* __restrict__ ipc, const float* __restrict__ ipf, const float* __restrict__ xn,
    const float* __restrict__ cn, const float* __restrict__ fbound, const float* __restrict__ fbpm,
    const float* __restrict__ st, const float* __restrict__ stmin, const float* __restrict__ stminrow,
    const uint32_t* __restrict__ stabs_bits, uint32_t n, uint32_t C, uint32_t F, uint32_t nb, int fb_sorted,
    int list_all, uint32_t* __restrict__ cell_out, float* __restrict__ dist_out, uint32_t* __restrict__ fb_list,
    uint32_t* __restrict__ fb_count) {
    __shared__ float bmax_s[4][PK_AB_NBMAX];
    const int lane = threadIdx.x & 63;
    const uint32_t pt = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pt >= n) return;  // wave-uniform; the kernel has no block barrier
    float* bmax = bmax_s[threadIdx.x >> 6];
    const float* pc = ipc + (uint64_t)pt * C;
    const float* pf = ipf + (uint64_t)pt * F;
    const float x2 = xn[pt];
    // 1. c0 and the reference's scan of it
    float ct0 = FLT_MAX;
    uint32_t c0 = 0xffffffffu;
    for (uint32_t c = lane; c < C; c += 64) {
        const float ct = (cn[c] - pc[c]) + x2;
        if (c0 == 0xffffffffu || ct < ct0) {  // c ascends per lane: '<' keeps the smaller id on ties
            ct0 = ct;
            c0 = c;
        }
    }
    pk_wave_argmin(ct0, c0);
    float B = FLT_MAX;
    int32_t cell0 = -1;
    pk_assign_one_coarse(c0, ct0, pf, st, fbound, F, fb_sorted, B, cell0);
    // block maxima of <x,f>, their maximum, and max |<x,f>|
    float gmax = -FLT_MAX, amax = 0.0f;
    for (uint32_t j = lane; j < nb; j += 64) {
        float m = -FLT_MAX;
        const uint32_t f1 = (j + 1) * PK_AB < F ? (j + 1) * PK_AB : F;
        for (uint32_t f = j * PK_AB; f < f1; ++f) {
            const float v = pf[f];
            m = v > m ? v : m;
            const float a = fabsf(v);
            amax = a > amax ? a : amax;
        }
        bmax[j] = m;
        gmax = m > gmax ? m : gmax;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float g2 = __shfl_xor(gmax, o, 64), a2 = __shfl_xor(amax, o, 64);
        gmax = g2 > gmax ? g2 : gmax;
        amax = a2 > amax ? a2 : amax;
    }
    pk_wave_lds_sync();
    const float stabs = __uint_as_float(*stabs_bits);
    // 2. cells that could score <= B
    PkCellKey best{FLT_MAX, FLT_MAX, 0xffffffffu, 0u};
    for (uint32_t cb0 = 0; cb0 < C; cb0 += 64) {
        const uint32_t c = cb0 + lane;
        const uint32_t cl = c < C ? c : C - 1u;
        const float ct = (cn[cl] - pc[cl]) + x2;
        const float srow = stminrow[cl];
        const float marg = 0x1p-18f * ((fabsf(ct) + stabs) + amax);
        const bool cand = c < C && c != c0 && !(((ct + srow) - gmax) > B + marg);
        uint64_t cm = __ballot(cand);
        while (cm) {
            const int src = __ffsll((unsigned long long)cm) - 1;
            cm &= cm - 1;
            const uint32_t cc = cb0 + (uint32_t)src;
            const float cct = __shfl(ct, src, 64), cmg = __shfl(marg, src, 64);
            const float* srowp = stmin + (uint64_t)cc * nb;
            const float* strow = st + (uint64_t)cc * F;
            for (uint32_t j0 = 0; j0 < nb; j0 += 64) {
                const uint32_t j = j0 + lane;
                const uint32_t jl = j < nb ? j : nb - 1u;
                const float lb = (cct + srowp[jl]) - bmax[jl];
                uint64_t bm = __ballot(j < nb && !(lb > B + cmg));
                while (bm) {
                    // two blocks per pass: lanes 0..31 score the first, 32..63 the second
                    const int b0 = __ffsll((unsigned long long)bm) - 1;
                    bm &= bm - 1;
                    int b1 = -1;
                    if (bm) {
                        b1 = __ffsll((unsigned long long)bm) - 1;
                        bm &= bm - 1;
                    }
                    const int mine = lane < 32 ? b0 : b1;
                    const uint32_t f = (j0 + (uint32_t)(mine < 0 ? b0 : mine)) * PK_AB + (lane & 31);
                    const uint32_t fl = f < F ? f : F - 1u;
                    const float s = (cct - pf[fl]) + strow[fl];
                    if (mine >= 0 && f < F) {
                        const PkCellKey k{s, cct, cc, f};
                        if (pk_key_less(k, best)) best = k;
                    }
                }
            }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        PkCellKey k{__shfl_xor(best.v, o, 64), __shfl_xor(best.ct, o, 64), (uint32_t)__shfl_xor((int)best.c, o, 64),
                    (uint32_t)__shfl_xor((int)best.f, o, 64)};
        if (pk_key_less(k, best)) best = k;
    }
    // 3. the first coarse list's answer, or the certified minimum
    const PkCellKey k0{B, ct0, c0, (uint32_t)cell0 - c0 * F};
    if (lane == 0) {
        bool ok = true;
        uint32_t cell = (uint32_t)cell0;
        float dist = B;
        if (pk_key_less(best, k0)) {
            ok = best.v >= 0.0f && !(__fsqrt_rn(best.v) - fbpm[best.f] < __fsqrt_rn(best.ct));
            cell = best.c * F + best.f;
            dist = best.v;
        }
        if (ok && !list_all) {
            cell_out[pt] = cell;
            if (dist_out) dist_out[pt] = dist;
        } else {
            fb_list[atomicAdd(fb_count, 1u)] = pt;
        }
    }
}

extern "C" hipError_t pk_launch_assign_bounded(const float* ipc, const float* ipf, const float* xn, const float* cn,
                                               const float* fbound, const float* fbpm, const float* st,
                                               const float* stmin, const float* stminrow, const uint32_t* stabs_bits,
                                               uint32_t n, uint32_t C, uint32_t F, int fb_sorted, int list_all,
                                               uint32_t* cell_out, float* dist_out, uint32_t* fb_list,
                                               uint32_t* fb_count, hipStream_t s) {
    const uint32_t nb = (F + PK_AB - 1) / PK_AB;
    if (nb > PK_AB_NBMAX) return hipErrorInvalidValue;
    hipError_t e = hipMemsetAsync(fb_count, 0, sizeof(uint32_t), s);
    if (e != hipSuccess) return e;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pk_assign_bounded, dim3((n + 3) / 4), dim3(256), 0, s, ipc, ipf, xn, cn, fbound, fbpm, st,
                       stmin, stminrow, stabs_bits, n, C, F, nb, fb_sorted, list_all, cell_out, dist_out, fb_list,
                       fb_count);
    return hipGetLastError();
}

// Nearest coarse term per point (used to order points so that waves running
// together share stationary-table rows in L2).
__global__ void pk_nearest_coarse(const float* __restrict__ ipc, const float* __restrict__ xn,
                                  const float* __restrict__ cn, uint32_t n, uint32_t C, uint32_t* __restrict__ out) {
    const uint32_t pt = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (pt >= n) return;
    const float* pc = ipc + (uint64_t)pt * C;
    const float x2 = xn[pt];
    float bv = FLT_MAX;
    uint32_t bi = 0xffffffffu;
    for (uint32_t c = lane; c < C; c += 64) {
        float ct = (cn[c] - pc[c]) + x2;
        if (bi == 0xffffffffu || ct < bv) { bv = ct; bi = c; }
    }
    pk_wave_argmin(bv, bi);
    if (lane == 0) out[pt] = bi;
}

extern "C" hipError_t pk_launch_nearest_coarse(const float* ipc, const float* xn, const float* cn, uint32_t n,
                                               uint32_t C, uint32_t* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pk_nearest_coarse, dim3((n + 3) / 4), dim3(256), 0, s, ipc, xn, cn, n, C, out);
    return hipGetLastError();
}

extern "C" hipError_t pk_launch_assign_cells(const float* ipc, const float* ipf, const float* xn, const float* cn,
                                             const float* fbound, const float* st, uint32_t n, uint32_t C,
                                             uint32_t F, int fb_sorted, const uint32_t* order, uint32_t* cell_out,
                                             float* dist_out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    uint32_t C2 = 64;
    while (C2 < C) C2 <<= 1;
    size_t smem = sizeof(uint64_t) * 4 * C2;
    static PkLdsCache lds_cache;
    size_t* slot = pk_lds_slot(lds_cache);
    if (!slot) return hipErrorInvalidDevice;
    if (!*slot) {
        const hipError_t e = pk_allow_lds((const void*)pk_assign_cells, 160 * 1024, slot);
        if (e != hipSuccess) return e;
    }
    const size_t allowed = *slot;
    if (smem > allowed) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pk_assign_cells, dim3((n + 3) / 4), dim3(256), smem, s, ipc, ipf, xn, cn, fbound, st, n, C, F,
                       C2, fb_sorted, order, cell_out, dist_out);
    return hipGetLastError();
}

// fbound[f] = 0 - sqrt(dot(f,f)/2)   (hierarchical_cluster_index.cpp:1186-1195)
__global__ void pk_fine_bound(const float* __restrict__ fine, uint32_t F, uint32_t dim, float* __restrict__ out) {
    uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const float* v = fine + (uint64_t)f * dim;
    float h = pk_dot(v, v, dim) / 2;
    out[f] = 0.0f - __fsqrt_rn(h);
}

extern "C" hipError_t pk_launch_fine_bound(const float* fine, uint32_t F, uint32_t dim, float* out, hipStream_t s) {
    hipLaunchKernelGGL(pk_fine_bound, dim3((F + 255) / 256), dim3(256), 0, s, fine, F, dim, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Filter-PQ encode: one thread per point, one subspace's codebook in LDS at a
// time.  residual r = (x - c) - f; code_m = argmin_k (float)((double)(pn + cn_k)
// - 2.0 * (double)dot(r_m, cb_k)) with strict '<' from FLT_MAX, the three dot
// products of lsq floats in MKL's order for such lengths (pk_sdot_short);
// offset = 2 * dot(decoded, c + f), accumulated in subspace order.
__global__ __launch_bounds__(256) void pk_encode_filter(const float* __restrict__ X, const uint32_t* __restrict__ cell,
                                                        uint32_t n, const float* __restrict__ coarse,
                                                        const float* __restrict__ fine, uint32_t F, uint32_t dim,
                                                        const float* __restrict__ codebook, uint32_t nsq,
                                                        float* __restrict__ off_out, uint8_t* __restrict__ codes_out,
                                                        uint64_t rec_stride, uint64_t codes_offset) {
    __shared__ float cb[PK_KS * 4];
    __shared__ float cbn[PK_KS];
    const uint32_t lsq = dim / nsq;
    const uint32_t pt = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = pt < n;
    uint32_t cc = 0, ff = 0;
    if (valid) {
        cc = cell[pt] / F;
        ff = cell[pt] % F;
    }
    const float* x = X + (uint64_t)(valid ? pt : 0) * dim;
    const float* cv = coarse + (uint64_t)cc * dim;
    const float* fv = fine + (uint64_t)ff * dim;
    float offacc = 0.0f;
    for (uint32_t m = 0; m < nsq; ++m) {
        __syncthreads();
        for (uint32_t e = threadIdx.x; e < PK_KS * lsq; e += blockDim.x)
            cb[e] = codebook[(uint64_t)m * PK_KS * lsq + e];
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < PK_KS; k += blockDim.x) cbn[k] = pk_sdot_short(cb + k * lsq, cb + k * lsq, lsq);
        __syncthreads();
        if (!valid) continue;
        float r[4];
        for (uint32_t j = 0; j < lsq; ++j) {
            uint32_t i = m * lsq + j;
            r[j] = (x[i] - cv[i]) - fv[i];
        }
        float pn = pk_sdot_short(r, r, lsq);
        float best = FLT_MAX;
        uint32_t arg = 0;
        for (uint32_t k = 0; k < PK_KS; ++k) {
            float dt = pk_sdot_short(r, cb + k * lsq, lsq);
            float d = (float)((double)(pn + cbn[k]) - 2.0 * (double)dt);
            if (d < best) { best = d; arg = k; }
        }
        codes_out[(uint64_t)pt * rec_stride + codes_offset + m] = (uint8_t)arg;
        for (uint32_t j = 0; j < lsq; ++j) {
            uint32_t i = m * lsq + j;
            float v = cv[i] + fv[i];
            offacc = fmaf(cb[arg * lsq + j], v, offacc);
        }
    }
    if (valid) {
        float off = 2 * offacc;
        if (off_out) off_out[pt] = off;
        else {
            uint8_t* dst = codes_out + (uint64_t)pt * rec_stride;
            uint32_t b = __float_as_uint(off);
            dst[0] = b & 0xff; dst[1] = (b >> 8) & 0xff; dst[2] = (b >> 16) & 0xff; dst[3] = (b >> 24) & 0xff;
        }
    }
}

extern "C" hipError_t pk_launch_encode_filter(const float* X, const uint32_t* cell, uint32_t n, const float* coarse,
                                              const float* fine, uint32_t F, uint32_t dim, const float* codebook,
                                              uint32_t nsq, float* off_out, uint8_t* codes_out,
                                              uint64_t rec_stride, uint64_t codes_offset, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (nsq == 0 || dim % nsq != 0 || dim / nsq > 4) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pk_encode_filter, dim3((n + 255) / 256), dim3(256), 0, s, X, cell, n, coarse, fine, F, dim,
                       codebook, nsq, off_out, codes_out, rec_stride, codes_offset);
    return hipGetLastError();
}

// Second-stage PQ encode (whether_pq = true; puck_index.cpp:602-709 with
// quantizations = {filter, pq}): one thread per point, one PQ subspace's
// codebook in LDS at a time.  The residual the PQ encodes is
// r1 = ((x - c) - f) - dec_f, dec_f the point's filter codewords (the filter
// quantizer's saxpy, :672-675); code_m = argmin_k (float)((double)(pn + cn_k)
// - 2.0 * (double)dot(r1_m, cb_k)) with strict '<' from FLT_MAX and the short
// dot products in MKL's order; offset = 2 * dot(dec_pq, (c + f) + dec_f)
// accumulated in subspace order (set_static_value_of_formula after the
// filter's, quantization.cpp:244-258).  fcodes: the point's filter codes
// (frec_stride bytes apart); records_out rows of 4 + pnsq bytes.
__global__ __launch_bounds__(256) void pk_encode_pq(const float* __restrict__ X, const uint32_t* __restrict__ cell,
                                                    uint32_t n, const float* __restrict__ coarse,
                                                    const float* __restrict__ fine, uint32_t F, uint32_t dim,
                                                    const float* __restrict__ fcb, uint32_t fnsq,
                                                    const uint8_t* __restrict__ fcodes, uint64_t frec_stride,
                                                    const float* __restrict__ pcb, uint32_t pnsq,
                                                    uint8_t* __restrict__ records_out) {
    __shared__ float cb[PK_KS * 4];
    __shared__ float cbn[PK_KS];
    const uint32_t flsq = dim / fnsq, plsq = dim / pnsq;
    const uint32_t pt = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = pt < n;
    uint32_t cc = 0, ff = 0;
    if (valid) {
        cc = cell[pt] / F;
        ff = cell[pt] % F;
    }
    const float* x = X + (uint64_t)(valid ? pt : 0) * dim;
    const float* cv = coarse + (uint64_t)cc * dim;
    const float* fv = fine + (uint64_t)ff * dim;
    const uint8_t* fc = fcodes + (uint64_t)(valid ? pt : 0) * frec_stride;
    const uint64_t rec = 4 + pnsq;
    float offacc = 0.0f;
    for (uint32_t m = 0; m < pnsq; ++m) {
        __syncthreads();
        for (uint32_t e = threadIdx.x; e < PK_KS * plsq; e += blockDim.x)
            cb[e] = pcb[(uint64_t)m * PK_KS * plsq + e];
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < PK_KS; k += blockDim.x)
            cbn[k] = pk_sdot_short(cb + k * plsq, cb + k * plsq, plsq);
        __syncthreads();
        if (!valid) continue;
        float r[4], df[4];
        for (uint32_t j = 0; j < plsq; ++j) {
            const uint32_t i = m * plsq + j, mf = i / flsq;
            df[j] = fcb[((uint64_t)mf * PK_KS + fc[mf]) * flsq + (i - mf * flsq)];
            r[j] = ((x[i] - cv[i]) - fv[i]) - df[j];
        }
        float pn = pk_sdot_short(r, r, plsq);
        float best = FLT_MAX;
        uint32_t arg = 0;
        for (uint32_t k = 0; k < PK_KS; ++k) {
            float dt = pk_sdot_short(r, cb + k * plsq, plsq);
            float d = (float)((double)(pn + cbn[k]) - 2.0 * (double)dt);
            if (d < best) { best = d; arg = k; }
        }
        records_out[(uint64_t)pt * rec + 4 + m] = (uint8_t)arg;
        for (uint32_t j = 0; j < plsq; ++j) {
            const uint32_t i = m * plsq + j;
            const float v = (cv[i] + fv[i]) + df[j];
            offacc = fmaf(cb[arg * plsq + j], v, offacc);
        }
    }
    if (valid) {
        const uint32_t b = __float_as_uint(2 * offacc);
        uint8_t* dst = records_out + (uint64_t)pt * rec;
        dst[0] = b & 0xff; dst[1] = (b >> 8) & 0xff; dst[2] = (b >> 16) & 0xff; dst[3] = (b >> 24) & 0xff;
    }
}

extern "C" hipError_t pk_launch_encode_pq(const float* X, const uint32_t* cell, uint32_t n, const float* coarse,
                                          const float* fine, uint32_t F, uint32_t dim, const float* fcb, uint32_t fnsq,
                                          const uint8_t* fcodes, uint64_t frec_stride, const float* pcb,
                                          uint32_t pnsq, uint8_t* records_out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (fnsq == 0 || pnsq == 0 || dim % fnsq || dim % pnsq || dim / fnsq > 4 || dim / pnsq > 4)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(pk_encode_pq, dim3((n + 255) / 256), dim3(256), 0, s, X, cell, n, coarse, fine, F, dim, fcb,
                       fnsq, fcodes, frec_stride, pcb, pnsq, records_out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// nearest_center argmin over a chunk of GEMM inner products ip[n x k]:
//   d = (float)((double)(pn + cn[c]) - 2.0 * (double)ip), strict '<' from FLT_MAX.
__global__ __launch_bounds__(256) void pk_kmeans_argmin(const float* __restrict__ ip, const float* __restrict__ pn,
                                                        const float* __restrict__ cn, uint32_t n, uint32_t k,
                                                        int32_t* __restrict__ assign, float* __restrict__ dis) {
    const uint32_t pt = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (pt >= n) return;
    const float p = pn[pt];
    const float* row = ip + (uint64_t)pt * k;
    float bv = FLT_MAX;
    uint32_t bi = 0xffffffffu;
    for (uint32_t c = lane; c < k; c += 64) {
        float d = (float)((double)(p + cn[c]) - 2.0 * (double)row[c]);
        if (d < bv) { bv = d; bi = c; }
    }
    pk_wave_argmin(bv, bi);
    if (lane == 0) {
        assign[pt] = bi == 0xffffffffu ? 0 : (int32_t)bi;
        if (dis) dis[pt] = bv;
    }
}

extern "C" hipError_t pk_launch_kmeans_argmin(const float* ip, const float* pn, const float* cn, uint32_t n,
                                              uint32_t k, int32_t* assign, float* dis, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pk_kmeans_argmin, dim3((n + 3) / 4), dim3(256), 0, s, ip, pn, cn, n, k, assign, dis);
    return hipGetLastError();
}

// nearest_center for points and centroids of 1, 2 or 4 floats (the
// filter-PQ codebook k-means on sub-vectors): the reference takes one
// cblas_sdot per (point, centroid) pair and per norm, so every dot product
// here is pk_sdot_short; d = (float)((double)(pn + cn) - 2.0 * (double)dot),
// strict '<' from FLT_MAX in centroid order.  One thread per point, the
// centroids and their norms staged through LDS in tiles of KM_TILE.
#define KM_TILE 1024u

__global__ __launch_bounds__(256) void pk_kmeans_short(const float* __restrict__ x, uint32_t n, uint32_t dim,
                                                       const float* __restrict__ cent, uint32_t k,
                                                       int32_t* __restrict__ assign, float* __restrict__ dis) {
    __shared__ float cs[KM_TILE * 4];
    __shared__ float csn[KM_TILE];
    const uint32_t pt = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = pt < n;
    float xv[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    const float* xp = x + (uint64_t)(valid ? pt : 0) * dim;
    for (uint32_t j = 0; j < dim; ++j) xv[j] = xp[j];
    const float pn = pk_sdot_short(xv, xv, dim);
    float best = FLT_MAX;
    uint32_t arg = 0;
    for (uint32_t c0 = 0; c0 < k; c0 += KM_TILE) {
        const uint32_t ct = (k - c0) < KM_TILE ? (k - c0) : KM_TILE;
        __syncthreads();
        for (uint32_t e = threadIdx.x; e < ct * dim; e += blockDim.x) cs[e] = cent[(uint64_t)c0 * dim + e];
        __syncthreads();
        for (uint32_t c = threadIdx.x; c < ct; c += blockDim.x) csn[c] = pk_sdot_short(cs + c * dim, cs + c * dim, dim);
        __syncthreads();
        for (uint32_t c = 0; c < ct; ++c) {
            const float d = (float)((double)(pn + csn[c]) - 2.0 * (double)pk_sdot_short(xv, cs + c * dim, dim));
            if (d < best) { best = d; arg = c0 + c; }
        }
    }
    if (valid) {
        assign[pt] = (int32_t)arg;
        if (dis) dis[pt] = best;
    }
}

extern "C" hipError_t pk_launch_kmeans_short(const float* x, uint32_t n, uint32_t dim, const float* cent, uint32_t k,
                                             int32_t* assign, float* dis, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (!(dim == 1 || dim == 2 || dim == 4) || k == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pk_kmeans_short, dim3((n + 255) / 256), dim3(256), 0, s, x, n, dim, cent, k, assign, dis);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Empty cells get stationary = sqrt(FLT_MAX); min_dist_offset[c] = min over the
// coarse cluster's cells (starting from FLT_MAX).  One block per coarse cluster.
__global__ __launch_bounds__(256) void pk_finalize_cells(float* __restrict__ st, const uint32_t* __restrict__ cs,
                                                         uint32_t F, float* __restrict__ mdo) {
    __shared__ float red[256];
    const uint32_t c = blockIdx.x;
    float mn = FLT_MAX;
    const float empty = __fsqrt_rn(FLT_MAX);
    for (uint32_t f = threadIdx.x; f < F; f += blockDim.x) {
        uint64_t i = (uint64_t)c * F + f;
        float v = st[i];
        if (cs[i] == cs[i + 1]) {
            v = empty;
            st[i] = v;
        }
        mn = v < mn ? v : mn;
    }
    red[threadIdx.x] = mn;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            float b = red[threadIdx.x + o];
            if (b < red[threadIdx.x]) red[threadIdx.x] = b;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) mdo[c] = red[0];
}

extern "C" hipError_t pk_launch_finalize_cells(float* st, const uint32_t* cs, uint32_t C, uint32_t F, float* mdo,
                                               hipStream_t s) {
    hipLaunchKernelGGL(pk_finalize_cells, dim3(C), dim3(256), 0, s, st, cs, F, mdo);
    return hipGetLastError();
}

// Gather rows into memory order: dst[m] = src[mem2local[m]] (row of `bytes`).
// One element per work-item in a grid-stride loop: an index's arrays hold
// more elements (100M x 128 features) than one dispatch has work-items (2^32).
#define GATHER_BLOCKS 4096u

__global__ void pk_gather_rows(const uint8_t* __restrict__ src, uint64_t src_stride, uint64_t src_off,
                               uint8_t* __restrict__ dst, uint64_t dst_stride, uint32_t bytes,
                               const uint32_t* __restrict__ mem2local, uint64_t total) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < total; gid += step) {
        const uint64_t m = gid / bytes;
        const uint32_t b = (uint32_t)(gid % bytes);
        dst[m * dst_stride + b] = src[(uint64_t)mem2local[m] * src_stride + src_off + b];
    }
}

__global__ void pk_gather_f32(const float* __restrict__ src, uint32_t width, float* __restrict__ dst,
                              const uint32_t* __restrict__ mem2local, uint64_t total) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < total; gid += step) {
        const uint64_t m = gid / width;
        const uint32_t b = (uint32_t)(gid % width);
        dst[gid] = src[(uint64_t)mem2local[m] * width + b];
    }
}

static unsigned gather_blocks(uint64_t total) {
    const uint64_t b = (total + 255) / 256;
    return (unsigned)(b < GATHER_BLOCKS ? b : GATHER_BLOCKS);
}

extern "C" hipError_t pk_launch_gather_rows(const uint8_t* src, uint64_t src_stride, uint64_t src_off, uint8_t* dst,
                                            uint64_t dst_stride, uint32_t bytes, const uint32_t* mem2local,
                                            uint32_t n, hipStream_t s) {
    if (n == 0 || bytes == 0) return hipSuccess;
    const uint64_t total = (uint64_t)n * bytes;
    hipLaunchKernelGGL(pk_gather_rows, dim3(gather_blocks(total)), dim3(256), 0, s, src, src_stride, src_off, dst,
                       dst_stride, bytes, mem2local, total);
    return hipGetLastError();
}

extern "C" hipError_t pk_launch_gather_f32(const float* src, uint32_t width, float* dst, const uint32_t* mem2local,
                                           uint32_t n, hipStream_t s) {
    if (n == 0 || width == 0) return hipSuccess;
    const uint64_t total = (uint64_t)n * width;
    hipLaunchKernelGGL(pk_gather_f32, dim3(gather_blocks(total)), dim3(256), 0, s, src, width, dst, mem2local,
                       total);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Realtime insert (RealtimeInsertPuckIndex, realtime_insert_puck_index.cpp):
// a batch of inserted points is merged into the memory layout, where each
// cell holds its loaded points (ascending offset), then its inserted points
// newest first -- the order in which the reference offers them to the filter
// heap (:398-453).  Per cell: nbase loaded points, add[k] points of this
// batch; shift[k] = exclusive prefix sum of add.

// add[cell[i]] += 1
__global__ void pk_cell_histogram(const uint32_t* __restrict__ cell, uint32_t n, uint32_t* __restrict__ add) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&add[cell[i]], 1u);
}

// out[k] = sum of in[0..k) for k in [0, n]: block sums, a one-block scan of
// them, then each block's prefix.  1024 elements per block, at most 1024 * 64
// blocks.
#define PK_SCAN_B 1024u
__device__ __forceinline__ uint32_t pk_block_exscan_1024(uint32_t v, uint32_t* sh, uint32_t* total) {
    // Hillis-Steele over 1024 threads in LDS
    const uint32_t t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (uint32_t o = 1; o < 1024u; o <<= 1) {
        const uint32_t a = t >= o ? sh[t - o] : 0u;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    const uint32_t incl = sh[t];
    if (total) *total = sh[1023];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(1024) void pk_scan_blocks(const uint32_t* __restrict__ in, uint64_t n,
                                                       uint32_t* __restrict__ out, uint32_t* __restrict__ bsum) {
    __shared__ uint32_t sh[1024];
    const uint64_t i = (uint64_t)blockIdx.x * PK_SCAN_B + threadIdx.x;
    const uint32_t v = i < n ? in[i] : 0u;
    uint32_t tot;
    const uint32_t ex = pk_block_exscan_1024(v, sh, &tot);
    if (i <= n) out[i] = ex;
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void pk_scan_sums(uint32_t* __restrict__ bsum, uint32_t nb) {
    __shared__ uint32_t sh[1024];
    // 64 sums per thread, in order
    constexpr uint32_t PER = 64;
    uint32_t loc[PER];
    uint32_t s = 0;
    for (uint32_t j = 0; j < PER; ++j) {
        const uint32_t i = threadIdx.x * PER + j;
        loc[j] = i < nb ? bsum[i] : 0u;
        s += loc[j];
    }
    uint32_t run = pk_block_exscan_1024(s, sh, nullptr);
    for (uint32_t j = 0; j < PER; ++j) {
        const uint32_t i = threadIdx.x * PER + j;
        if (i < nb) bsum[i] = run;
        run += loc[j];
    }
}

__global__ __launch_bounds__(1024) void pk_scan_add(uint32_t* __restrict__ out, uint64_t n,
                                                    const uint32_t* __restrict__ bsum) {
    const uint64_t i = (uint64_t)blockIdx.x * PK_SCAN_B + threadIdx.x;
    if (i <= n) out[i] += bsum[blockIdx.x];
}

// work: at least pk_exscan_work(n) words
extern "C" uint64_t pk_exscan_work(uint64_t n) { return (n + 1 + PK_SCAN_B - 1) / PK_SCAN_B; }

extern "C" hipError_t pk_launch_exscan(const uint32_t* in, uint64_t n, uint32_t* out, uint32_t* work, hipStream_t s) {
    const uint64_t nb = pk_exscan_work(n);
    if (nb > 1024ull * 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pk_scan_blocks, dim3((unsigned)nb), dim3(1024), 0, s, in, n, out, work);
    hipLaunchKernelGGL(pk_scan_sums, dim3(1), dim3(1024), 0, s, work, (uint32_t)nb);
    hipLaunchKernelGGL(pk_scan_add, dim3((unsigned)nb), dim3(1024), 0, s, out, n, work);
    return hipGetLastError();
}

// Destination of every current memory position m in the merged layout: the
// cell k holding m (last k with cs[k] <= m), its offset o in the cell; loaded
// points move by shift[k], inserted ones also by add[k] (this batch goes in
// front of them).
__global__ void pk_insert_dest_old(const uint32_t* __restrict__ cs, uint32_t cells, const uint32_t* __restrict__ nbase,
                                   const uint32_t* __restrict__ add, const uint32_t* __restrict__ shift, uint32_t n,
                                   uint32_t* __restrict__ dest) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    uint32_t lo = 0, hi = cells;  // cs[lo] <= m < cs[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cs[mid] <= m) lo = mid; else hi = mid;
    }
    const uint32_t o = m - cs[lo];
    dest[m] = m + shift[lo] + (o >= nbase[lo] ? add[lo] : 0u);
}

// Destination of batch point i (cell[i], rank[i] among the batch's points of
// its cell, newest first): right after the cell's loaded points.
__global__ void pk_insert_dest_new(const uint32_t* __restrict__ cell, const uint32_t* __restrict__ rank, uint32_t n,
                                   const uint32_t* __restrict__ cs, const uint32_t* __restrict__ shift,
                                   const uint32_t* __restrict__ nbase, uint32_t n_old, uint32_t* __restrict__ dest,
                                   uint32_t* __restrict__ m2l_new) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = cell[i];
    const uint32_t d = cs[k] + shift[k] + nbase[k] + rank[i];
    dest[i] = d;
    m2l_new[d] = n_old + i;  // local id: the count of points before it
}

// cs_new[k] = cs[k] + shift[k], k in [0, cells]
__global__ void pk_insert_cell_start(const uint32_t* __restrict__ cs, const uint32_t* __restrict__ shift,
                                     uint32_t cells, uint32_t* __restrict__ out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k <= cells) out[k] = cs[k] + shift[k];
}

// nbase[k] = cs[k + 1] - cs[k]
__global__ void pk_cell_sizes(const uint32_t* __restrict__ cs, uint32_t cells, uint32_t* __restrict__ out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < cells) out[k] = cs[k + 1] - cs[k];
}

// Row scatter: row i of src (stride src_stride, from byte src_off, `bytes`
// long) to row dest[i] of dst (stride dst_stride).  4-byte words when every
// offset and length is a multiple of 4, else bytes; grid-stride.
__global__ void pk_scatter_rows_w(const uint32_t* __restrict__ src, uint64_t src_stride, uint64_t src_off,
                                  uint32_t* __restrict__ dst, uint64_t dst_stride, uint32_t words,
                                  const uint32_t* __restrict__ dest, uint64_t total) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < total; gid += step) {
        const uint64_t i = gid / words;
        const uint32_t w = (uint32_t)(gid % words);
        dst[(uint64_t)dest[i] * dst_stride + w] = src[i * src_stride + src_off + w];
    }
}

__global__ void pk_scatter_rows_b(const uint8_t* __restrict__ src, uint64_t src_stride, uint64_t src_off,
                                  uint8_t* __restrict__ dst, uint64_t dst_stride, uint32_t bytes,
                                  const uint32_t* __restrict__ dest, uint64_t total) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < total; gid += step) {
        const uint64_t i = gid / bytes;
        const uint32_t b = (uint32_t)(gid % bytes);
        dst[(uint64_t)dest[i] * dst_stride + b] = src[i * src_stride + src_off + b];
    }
}

extern "C" hipError_t pk_launch_scatter_rows(const void* src, uint64_t src_stride, uint64_t src_off, void* dst,
                                             uint64_t dst_stride, uint32_t bytes, const uint32_t* dest, uint32_t n,
                                             hipStream_t s) {
    if (n == 0 || bytes == 0) return hipSuccess;
    if (src_stride % 4 == 0 && src_off % 4 == 0 && dst_stride % 4 == 0 && bytes % 4 == 0 &&
        ((uintptr_t)src & 3u) == 0 && ((uintptr_t)dst & 3u) == 0) {
        const uint64_t total = (uint64_t)n * (bytes / 4);
        hipLaunchKernelGGL(pk_scatter_rows_w, dim3(gather_blocks(total)), dim3(256), 0, s, (const uint32_t*)src,
                           src_stride / 4, src_off / 4, (uint32_t*)dst, dst_stride / 4, bytes / 4, dest, total);
    } else {
        const uint64_t total = (uint64_t)n * bytes;
        hipLaunchKernelGGL(pk_scatter_rows_b, dim3(gather_blocks(total)), dim3(256), 0, s, (const uint8_t*)src,
                           src_stride, src_off, (uint8_t*)dst, dst_stride, bytes, dest, total);
    }
    return hipGetLastError();
}

// Stationary value of a cell that receives its first point by insert:
// sdot(c, f) - sdot(f, f) (realtime_insert_puck_index.cpp:326-338).
__global__ void pk_insert_stationary(const float* __restrict__ coarse, const float* __restrict__ fine, uint32_t F,
                                     uint32_t dim, const uint32_t* __restrict__ cells, uint32_t n,
                                     float* __restrict__ st) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = cells[i];
    const float* c = coarse + (uint64_t)(k / F) * dim;
    const float* f = fine + (uint64_t)(k % F) * dim;
    float v = pk_dot(c, f, dim);
    v -= pk_dot(f, f, dim);
    st[k] = v;
}

extern "C" hipError_t pk_launch_insert_prepare(const uint32_t* cell, uint32_t n, uint32_t cells, uint32_t* add,
                                               uint32_t* shift, uint32_t* work, hipStream_t s) {
    hipError_t e = hipMemsetAsync(add, 0, sizeof(uint32_t) * ((uint64_t)cells + 1), s);
    if (e != hipSuccess) return e;
    if (n) hipLaunchKernelGGL(pk_cell_histogram, dim3((n + 255) / 256), dim3(256), 0, s, cell, n, add);
    return pk_launch_exscan(add, cells, shift, work, s);
}

extern "C" hipError_t pk_launch_insert_dest(const uint32_t* cs, uint32_t cells, const uint32_t* nbase,
                                            const uint32_t* add, const uint32_t* shift, uint32_t n_old,
                                            uint32_t* dest_old, const uint32_t* cell, const uint32_t* rank,
                                            uint32_t n_new, uint32_t* dest_new, uint32_t* m2l_new,
                                            uint32_t* cs_new, hipStream_t s) {
    if (n_old)
        hipLaunchKernelGGL(pk_insert_dest_old, dim3((n_old + 255) / 256), dim3(256), 0, s, cs, cells, nbase, add,
                           shift, n_old, dest_old);
    if (n_new)
        hipLaunchKernelGGL(pk_insert_dest_new, dim3((n_new + 255) / 256), dim3(256), 0, s, cell, rank, n_new, cs,
                           shift, nbase, n_old, dest_new, m2l_new);
    hipLaunchKernelGGL(pk_insert_cell_start, dim3((cells + 1 + 255) / 256), dim3(256), 0, s, cs, shift, cells,
                       cs_new);
    return hipGetLastError();
}

extern "C" hipError_t pk_launch_cell_sizes(const uint32_t* cs, uint32_t cells, uint32_t* out, hipStream_t s) {
    hipLaunchKernelGGL(pk_cell_sizes, dim3((cells + 255) / 256), dim3(256), 0, s, cs, cells, out);
    return hipGetLastError();
}

extern "C" hipError_t pk_launch_insert_stationary(const float* coarse, const float* fine, uint32_t F, uint32_t dim,
                                                  const uint32_t* cells, uint32_t n, float* st, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pk_insert_stationary, dim3((n + 255) / 256), dim3(256), 0, s, coarse, fine, F, dim, cells, n,
                       st);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// k-means training (Kmeans::kmeans_core, kmeans.cpp:260-329, and
// HierarchicalClusterIndex::train, hierarchical_cluster_index.cpp:882-985).
//
// Centroid update in the reference's arithmetic: each centroid's points are
// added in point order with float additions (cblas_saxpy with alpha 1), then
// scaled by (float)(1.0 / count) (cblas_sscal).  The points are first put in
// cluster order by a stable counting sort, so one workgroup per centroid can
// add them in that order.

#define KS_THREADS 256u
#define KS_ITEMS 8u
#define KS_BLOCK (KS_THREADS * KS_ITEMS)

// hist[c * nb + b] = points of block b (KS_BLOCK consecutive points) in cluster c
__global__ __launch_bounds__(KS_THREADS) void pk_ks_hist(const int32_t* __restrict__ key, uint32_t n, uint32_t k,
                                                         uint32_t nb, uint32_t* __restrict__ hist) {
    extern __shared__ uint32_t cnt[];
    for (uint32_t c = threadIdx.x; c < k; c += blockDim.x) cnt[c] = 0;
    __syncthreads();
    const uint32_t b0 = blockIdx.x * KS_BLOCK;
    for (uint32_t r = 0; r < KS_ITEMS; ++r) {
        const uint32_t i = b0 + r * KS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(uint32_t)key[i]], 1u);
    }
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < k; c += blockDim.x) hist[(uint64_t)c * nb + blockIdx.x] = cnt[c];
}

// Stable scatter: point i goes to off[key * nb + block] + (points of its
// block with the same key before it).  Rounds of 256 points in point order;
// inside a round the four waves take their turn, and inside a wave the lanes
// with equal keys are found with one ballot per key bit.
__global__ __launch_bounds__(KS_THREADS) void pk_ks_scatter(const int32_t* __restrict__ key, uint32_t n, uint32_t k,
                                                            uint32_t nb, uint32_t kbits,
                                                            const uint32_t* __restrict__ off,
                                                            uint32_t* __restrict__ perm) {
    extern __shared__ uint32_t run[];
    for (uint32_t c = threadIdx.x; c < k; c += blockDim.x) run[c] = off[(uint64_t)c * nb + blockIdx.x];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint32_t b0 = blockIdx.x * KS_BLOCK;
    for (uint32_t r = 0; r < KS_ITEMS; ++r) {
        const uint32_t i = b0 + r * KS_THREADS + threadIdx.x;
        const bool valid = i < n;
        const uint32_t kv = valid ? (uint32_t)key[i] : 0u;
        uint64_t eq = __ballot(valid);
        for (uint32_t bit = 0; bit < kbits; ++bit) {
            const uint64_t ones = __ballot(valid && ((kv >> bit) & 1u));
            eq &= ((kv >> bit) & 1u) ? ones : ~ones;
        }
        for (uint32_t w = 0; w < KS_THREADS / 64u; ++w) {
            if (wave == w && valid) {
                const uint32_t base = run[kv];
                perm[base + (uint32_t)__popcll(eq & lt)] = i;
                // the group's lowest lane advances the counter (written from
                // the value read, so the read comes first)
                if ((eq & lt) == 0ull) run[kv] = base + (uint32_t)__popcll(eq);
            }
            __syncthreads();
        }
    }
}

extern "C" uint64_t pk_ks_hist_words(uint32_t n, uint32_t k) {
    return (uint64_t)k * ((n + KS_BLOCK - 1) / KS_BLOCK);
}

// perm [n]: point indices by cluster, point order inside a cluster; starts
// [k + 1]: each cluster's first position.  hist / off: pk_ks_hist_words
// words each; work: pk_exscan_work of that many.
extern "C" hipError_t pk_launch_sort_by_cluster(const int32_t* key, uint32_t n, uint32_t k, uint32_t* hist,
                                                uint32_t* off, uint32_t* work, uint32_t* perm, uint32_t* starts,
                                                hipStream_t s);

__global__ void pk_ks_starts(const uint32_t* __restrict__ off, uint32_t k, uint32_t nb, uint32_t n,
                             uint32_t* __restrict__ starts) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < k) starts[c] = off[(uint64_t)c * nb];
    if (c == k) starts[k] = n;
}

extern "C" hipError_t pk_launch_sort_by_cluster(const int32_t* key, uint32_t n, uint32_t k, uint32_t* hist,
                                                uint32_t* off, uint32_t* work, uint32_t* perm, uint32_t* starts,
                                                hipStream_t s) {
    if (k == 0 || k > 16384) return hipErrorInvalidValue;
    const uint32_t nb = (n + KS_BLOCK - 1) / KS_BLOCK;
    uint32_t kbits = 0;
    while ((1u << kbits) < k) ++kbits;
    const size_t lds = sizeof(uint32_t) * k;
    if (nb) {
        hipLaunchKernelGGL(pk_ks_hist, dim3(nb), dim3(KS_THREADS), lds, s, key, n, k, nb, hist);
        hipError_t e = pk_launch_exscan(hist, (uint64_t)k * nb, off, work, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(pk_ks_scatter, dim3(nb), dim3(KS_THREADS), lds, s, key, n, k, nb, kbits, off, perm);
        hipLaunchKernelGGL(pk_ks_starts, dim3((k + 1 + 255) / 256), dim3(256), 0, s, off, k, nb, n, starts);
    } else {
        hipError_t e = hipMemsetAsync(starts, 0, sizeof(uint32_t) * (k + 1ull), s);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

// One workgroup per centroid, one thread per dimension: acc starts at 0
// (mean_mode) or at the previous centroid with element [0][0] cleared (the
// reference clears only the first 4 bytes, kmeans.cpp:275); the centroid's
// points are added in point order, then the sum is scaled by
// (float)(1.0 / count) (inf for an empty cluster, which the split replaces).
__global__ __launch_bounds__(256) void pk_kmeans_update(const float* __restrict__ x, uint32_t dim,
                                                        const uint32_t* __restrict__ perm,
                                                        const uint32_t* __restrict__ starts,
                                                        float* __restrict__ cent, int mean_mode) {
    const uint32_t c = blockIdx.x;
    const uint32_t b = starts[c], e = starts[c + 1];
    const float alpha = (float)(1.0 / (double)(e - b));
    for (uint32_t t = threadIdx.x; t < dim; t += blockDim.x) {
        float acc = (mean_mode || (c == 0 && t == 0)) ? 0.0f : cent[(uint64_t)c * dim + t];
        uint32_t i = b;
        for (; i + 8 <= e; i += 8) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = x[(uint64_t)perm[i + u] * dim + t];
#pragma unroll
            for (int u = 0; u < 8; ++u) acc = acc + v[u];
        }
        for (; i < e; ++i) acc = acc + x[(uint64_t)perm[i] * dim + t];
        cent[(uint64_t)c * dim + t] = alpha * acc;
    }
}

extern "C" hipError_t pk_launch_kmeans_update(const float* x, uint32_t dim, const uint32_t* perm,
                                              const uint32_t* starts, uint32_t k, float* cent, int mean_mode,
                                              hipStream_t s) {
    if (k == 0) return hipSuccess;
    const uint32_t nt = dim >= 256 ? 256u : (dim >= 64 ? ((dim + 63) / 64) * 64 : 64u);
    hipLaunchKernelGGL(pk_kmeans_update, dim3(k), dim3(nt), 0, s, x, dim, perm, starts, cent, mean_mode);
    return hipGetLastError();
}

// out[i] = x[i] - c[assign[i]] (cblas_saxpy with alpha -1)
__global__ void pk_sub_assigned(const float* __restrict__ x, uint64_t total, uint32_t dim,
                                const float* __restrict__ cent, const int32_t* __restrict__ assign,
                                float* __restrict__ out) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += step) {
        const uint64_t i = g / dim;
        const uint32_t t = (uint32_t)(g % dim);
        out[g] = x[g] - cent[(uint64_t)assign[i] * dim + t];
    }
}

extern "C" hipError_t pk_launch_sub_assigned(const float* x, uint32_t n, uint32_t dim, const float* cent,
                                             const int32_t* assign, float* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint64_t total = (uint64_t)n * dim;
    hipLaunchKernelGGL(pk_sub_assigned, dim3(gather_blocks(total)), dim3(256), 0, s, x, total, dim, cent, assign,
                       out);
    return hipGetLastError();
}

// Squared norms in cblas_sdot's order (MKL's for 1, 2 or 4 floats, the
// sequential chain otherwise).
__global__ void pk_norms_sdot(const float* __restrict__ x, uint32_t n, uint32_t dim, float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = pk_sdot_short(x + (uint64_t)i * dim, x + (uint64_t)i * dim, dim);
}

extern "C" hipError_t pk_launch_norms_sdot(const float* x, uint32_t n, uint32_t dim, float* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pk_norms_sdot, dim3((n + 255) / 256), dim3(256), 0, s, x, n, dim, out);
    return hipGetLastError();
}

// One k-means++ step (kmeanspp_init_center, kmeans.cpp:86-102): the squared
// distance to the last chosen point, float(double(pn_j + pn_sel) - 2 sdot),
// lowers each point's best.
__global__ void pk_kmeanspp_step(const float* __restrict__ x, const float* __restrict__ pn, uint32_t n, uint32_t dim,
                                 uint32_t sel, float* __restrict__ best) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const float t = (float)((double)(pn[j] + pn[sel]) -
                            2.0 * (double)pk_sdot_short(x + (uint64_t)j * dim, x + (uint64_t)sel * dim, dim));
    if (t < best[j]) best[j] = t;
}

extern "C" hipError_t pk_launch_kmeanspp_step(const float* x, const float* pn, uint32_t n, uint32_t dim,
                                              uint32_t sel, float* best, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pk_kmeanspp_step, dim3((n + 255) / 256), dim3(256), 0, s, x, pn, n, dim, sel, best);
    return hipGetLastError();
}
