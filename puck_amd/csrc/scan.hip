// scan.hip -- the budget scan (a query's fine budget b < F; no counterpart in
// the reference) as two data-parallel kernels:
//
//   pk_scan_cells      one workgroup per query: in each walked coarse list, the
//                      cut at the b-th smallest cell distance (cd + stationary)
//                      + fine distance (the reference's cell distance
//                      expression, puck_index.cpp:535), and the query's list of
//                      selected non-empty cells {first memory id, points before
//                      it, cell distance}
//   pk_scan_score_cut  one workgroup per query: the query's filter lookup table
//                      in LDS (quantization.cpp:260-270), then every point of
//                      every selected cell gets the reference's filter distance
//                      (compute_quantized_distance, puck_index.cpp:346-390:
//                      t = fl(2 * cell_dist + offset) + lookup sum in the 4-lane
//                      order of puck_index.cpp:75-113) as a (distance, memory
//                      id) key kept in registers (the cell list in LDS, a
//                      point's cell by binary search; consecutive threads take
//                      consecutive points: the record loads coalesce); then the
//                      filter_topk smallest keys (radix select on the distance
//                      bits, then on the ids of the keys tied at the cut) into
//                      the filter rows
//
// The result is the one of the per-query scan (pk_scan_budget in search.hip):
// the filter set is the filter_topk smallest (distance, memory id) keys over
// every point of the selected cells, which does not depend on the order in
// which points are scored.  A query with more than sc_capp points or
// PK_SC_MAXC cells, or whose index has F > PK_SC_CPT * PK_SC_CELL_THREADS, is
// marked and left to the per-query kernel (pk_scan_overflow_list builds its
// list).
#include "common.h"
#include "launch_args.h"

#define PK_SC_OVERFLOW 0xffffffffu
#define PK_SC_MAXC 512u  // cells per query in the scoring kernel's LDS

__device__ __forceinline__ bool pk_sc_budget_query(const PkWalkArgs& a, uint32_t qi) {
    return a.q_nf && a.q_nf[qi] != 0u && a.q_nf[qi] < a.F;
}

// ---------------------------------------------------------------------------
#ifndef PK_SC_CELL_THREADS
#define PK_SC_CELL_THREADS 256
#endif
#define PK_SC_CPT (2048 / PK_SC_CELL_THREADS)  // columns (fine clusters) per thread: F <= 2048
__global__ __launch_bounds__(PK_SC_CELL_THREADS) void pk_scan_cells(PkWalkArgs a) {
    __shared__ uint32_t hist[PK_SEL_HIST];
    __shared__ uint32_t tmp[PK_SC_CELL_THREADS / 64 + 1];
    __shared__ uint32_t red[2 * PK_SC_CELL_THREADS / 64];
    const uint32_t tid = threadIdx.x, NT = PK_SC_CELL_THREADS;
    const uint32_t qi = a.qlist ? a.qlist[blockIdx.x] : blockIdx.x;
    if (!pk_sc_budget_query(a, qi)) return;
    uint32_t* meta = a.sc_meta + 2ull * qi;
    if (a.qstatus[qi] != 0) {
        if (tid == 0) {
            meta[0] = 0;
            meta[1] = 0;
            if (a.stats) {
                for (int k = 0; k < 4; ++k) a.stats[qi * 4 + k] = 0;
            }
        }
        return;
    }
    const uint32_t F = a.F, b = a.q_nf[qi];
    if (F > PK_SC_CPT * NT) {
        if (tid == 0) meta[0] = PK_SC_OVERFLOW;
        return;
    }
    uint32_t scc = a.scc;
    if (a.q_scc) scc = a.q_scc[qi] < scc ? a.q_scc[qi] : scc;
    const float* fdq = a.fdist + (uint64_t)qi * F;
    uint4* cells = a.sc_cells + (uint64_t)qi * PK_SC_MAXC;
    uint32_t npt = 0, ncell = 0;
    bool ovf = false;
    for (uint32_t l = 0; l < scc; ++l) {
        const uint32_t c = a.csel_id[(uint64_t)qi * a.cstride + l];
        const float cd = a.csel_d[(uint64_t)qi * a.cstride + l];
        const float* stc = a.stationary + (uint64_t)c * F;
        const uint32_t* csc = a.cell_start + (uint64_t)c * F;
        // per column slot: the cell's point count when selected (else 0), its
        // first memory id and its cell distance key
        uint32_t kv[PK_SC_CPT], c0[PK_SC_CPT], c1[PK_SC_CPT];
        // the nearest list with the probe's sorted cell distances: the b-th
        // smallest is entry b - 1 when it is below the last entry (no tie can
        // reach past the sorted prefix), and the selected cells are the entries
        // up to its last tie (block-uniform test)
        bool fast = false;
        if (l == 0 && a.top_f && b < a.top_k) {
            const float* td = a.top_d + (uint64_t)qi * a.top_k;
            fast = td[b - 1] < td[a.top_k - 1];
        }
        if (fast) {
            const float* td = a.top_d + (uint64_t)qi * a.top_k;
            const uint32_t* tf = a.top_f + (uint64_t)qi * a.top_k;
            const float vb = td[b - 1];
            const uint32_t j = tid < a.top_k ? tid : a.top_k - 1u;  // in-bounds for every lane
            const float v = td[j];
            const uint32_t f = tf[j];
            const uint32_t s0 = csc[f], s1 = csc[f + 1];
            const bool sel = tid < a.top_k && v <= vb;
#pragma unroll
            for (int jj = 0; jj < PK_SC_CPT; ++jj) {
                kv[jj] = pk_f2ord(v);
                c0[jj] = s0;
                c1[jj] = (jj == 0 && sel) ? s1 - s0 : 0u;
            }
        } else {
            // every load of the list at once, at in-bounds columns (a load under
            // a lane condition is waited for before the next one is issued)
            float sv[PK_SC_CPT], fv[PK_SC_CPT];
#pragma unroll
            for (int j = 0; j < PK_SC_CPT; ++j) {
                const uint32_t k = min(j * NT + tid, F - 1u);
                sv[j] = stc[k];
                fv[j] = fdq[k];
                c0[j] = csc[k];
                c1[j] = csc[k + 1];
            }
#pragma unroll
            for (int j = 0; j < PK_SC_CPT; ++j) kv[j] = pk_f2ord((cd + sv[j]) + fv[j]);
            // the list's cut: b-th smallest cell key
            const uint32_t kcut = pk_select_rank_each(
                [&](auto sink) {
#pragma unroll
                    for (int j = 0; j < PK_SC_CPT; ++j)
                        if (j * NT + tid < F) sink(kv[j]);
                },
                b, hist, tmp, red);
#pragma unroll
            for (int j = 0; j < PK_SC_CPT; ++j) {
                const bool sel = j * NT + tid < F && kv[j] <= kcut;
                c1[j] = sel ? c1[j] - c0[j] : 0u;  // the cell's point count when selected
            }
        }
        uint32_t lp = 0, lc = 0;
#pragma unroll
        for (int j = 0; j < PK_SC_CPT; ++j) {
            lp += c1[j];
            lc += c1[j] > 0 ? 1u : 0u;
        }
        uint32_t tp, tc;
        uint32_t pb = pk_block_exscan(lp, tmp, &tp);
        uint32_t cb = pk_block_exscan(lc, tmp, &tc);
        if (npt + tp > a.sc_capp || ncell + tc > PK_SC_MAXC) {  // block-uniform
            ovf = true;
            break;
        }
        // the selected non-empty cells, in column order within each thread
#pragma unroll
        for (int j = 0; j < PK_SC_CPT; ++j) {
            if (c1[j] > 0) {
                cells[ncell + cb] = make_uint4(c0[j], npt + pb, __float_as_uint(pk_ord2f(kv[j])), 0u);
                ++cb;
                pb += c1[j];
            }
        }
        npt += tp;
        ncell += tc;
    }
    if (tid == 0) {
        meta[0] = ovf ? PK_SC_OVERFLOW : npt;
        meta[1] = ncell;
        if (a.stats && !ovf) {
            a.stats[qi * 4 + 0] = 0;
            a.stats[qi * 4 + 1] = 0;
            a.stats[qi * 4 + 2] = ncell;
            a.stats[qi * 4 + 3] = npt;
        }
    }
}

// ---------------------------------------------------------------------------
// Filter distances of PP points per thread: item i = j * NT + tid of the
// round (consecutive threads, consecutive records); NV: nsq / 16 code vectors
// per record (0: byte loads for any nsq).  T: the table's element type, float
// (the reference's) or __half (the FP16 form: entries scaled by 2^e, the sum
// scaled back by inv = 2^-e).
#define PK_SC_PP 2
__device__ __forceinline__ float pk_sc_lv(float v) { return v; }
__device__ __forceinline__ float pk_sc_lv(__half v) { return __half2float(v); }

// The query's selected cells in LDS: first memory id, points before the
// cell, cell distance bits; pre[nc] = n.
struct PkScCells {
    const uint32_t* c0;
    const uint32_t* pre;
    const uint32_t* cd;
    uint32_t nc;
};

// Point i's {memory id, cell distance bits}: the cell with pre <= i < next pre.
__device__ __forceinline__ uint2 pk_sc_point(const PkScCells& sc, uint32_t i) {
    uint32_t lo = 0, hi = sc.nc;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sc.pre[mid] <= i) lo = mid;
        else hi = mid;
    }
    return make_uint2(sc.c0[lo] + (i - sc.pre[lo]), sc.cd[lo]);
}

// Filter distances of PP points per thread, returned as keys in registers:
// kout[h] is the key (distance bits, memory id) of point i0 + h * NT +
// threadIdx.x (consecutive threads, consecutive records), or ~0 past n.
template <int NT, int NV, typename T>
__device__ __forceinline__ void pk_sc_score_keys(const PkWalkArgs& a, const T* __restrict__ lut, float inv,
                                                 const PkScCells& sc, uint32_t i0, uint32_t n,
                                                 uint64_t (&kout)[PK_SC_PP]) {
    const uint32_t nsq = a.nsq;
    uint2 it[PK_SC_PP];
    float off[PK_SC_PP];
    uint4 cw[PK_SC_PP][NV > 0 ? NV : 1];
#pragma unroll
    for (int h = 0; h < PK_SC_PP; ++h) {
        const uint32_t i = i0 + h * NT + threadIdx.x;
        it[h] = pk_sc_point(sc, i < n ? i : n - 1u);
    }
#pragma unroll
    for (int h = 0; h < PK_SC_PP; ++h) {
        off[h] = a.rec_off[it[h].x];
        if constexpr (NV > 0) {
            const uint4* cp = (const uint4*)(a.rec_codes + (uint64_t)it[h].x * nsq);
#pragma unroll
            for (int g = 0; g < NV; ++g) cw[h][g] = cp[g];
        }
    }
#pragma unroll
    for (int h = 0; h < PK_SC_PP; ++h) {
        const uint32_t i = i0 + h * NT + threadIdx.x;
        const uint32_t P = it[h].x;
        float tv = 2.0f * __uint_as_float(it[h].y) + off[h];  // 2x is exact: one rounding, as in double
        float acc[4] = {0.f, 0.f, 0.f, 0.f};
        if constexpr (NV > 0) {
#pragma unroll
            for (int g = 0; g < NV; ++g) {
                const uint32_t ws[4] = {cw[h][g].x, cw[h][g].y, cw[h][g].z, cw[h][g].w};
#pragma unroll
                for (int hh = 0; hh < 4; ++hh) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const uint32_t code = (ws[hh] >> (8 * j)) & 0xffu;
                        acc[j] = acc[j] + pk_sc_lv(lut[(16 * g + 4 * hh + j) * PK_KS + code]);
                    }
                }
            }
        } else {
            const uint8_t* codes = a.rec_codes + (uint64_t)P * nsq;
            for (uint32_t q = 0; q < nsq; ++q) acc[q & 3] = acc[q & 3] + pk_sc_lv(lut[q * PK_KS + codes[q]]);
        }
        const float sum = (acc[0] + acc[1]) + (acc[2] + acc[3]);
        if constexpr (sizeof(T) == sizeof(float)) tv = tv + sum;
        else tv = tv + __fmul_rn(sum, inv);
        kout[h] = i < n ? (((uint64_t)pk_f2ord(tv) << 32) | P) : ~0ull;
    }
}

// LDS of pk_scan_score_cut: the lookup table (FP32 or FP16), then the query
// (the cell list is static).
__host__ __device__ inline size_t pk_scan_lut_bytes(uint32_t nsq, int h16) {
    return (((size_t)nsq * PK_KS * (h16 ? 2u : 4u)) + 15u) & ~(size_t)15u;
}

// ---------------------------------------------------------------------------
// Scoring and filter cut of one query in one workgroup: the keys of the
// query's points stay in registers (PK_SC_KR rounds of PK_SC_PP per thread,
// sc_capp <= PK_SC_PP * PK_SC_KR * NT), then the filter_topk smallest
// (distance bits, memory id) are selected (radix select on the distance
// bits, then on the ids of the keys tied at the cut) and written to the
// filter rows.
template <int NT>
struct PkScCut {
    static constexpr int KR = 8192 / (PK_SC_PP * NT);  // rounds
    static constexpr int KPT = KR * PK_SC_PP;         // keys per thread
};

template <int NT, int NV, bool H16>
__global__ __launch_bounds__(NT) void pk_scan_score_cut(PkWalkArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint32_t hist[PK_SEL_HIST];
    __shared__ uint32_t tmp[NT / 64 + 1];
    __shared__ uint32_t red[2 * NT / 64];
    __shared__ uint32_t cnt[2];
    __shared__ uint32_t s_c0[PK_SC_MAXC], s_pre[PK_SC_MAXC + 1], s_cd[PK_SC_MAXC];
    constexpr int KR = PkScCut<NT>::KR, KPT = PkScCut<NT>::KPT;
    const uint32_t qi = a.qlist ? a.qlist[blockIdx.x] : blockIdx.x;
    const uint32_t tid = threadIdx.x;
    if (!pk_sc_budget_query(a, qi)) return;
    const uint32_t nit = a.sc_meta[2ull * qi];  // points, or the overflow mark
    if (nit == PK_SC_OVERFLOW) return;  // the per-query kernel's query
    if (tid == 0) {
        if (a.flags) a.flags[qi] = 0;  // deterministic: nothing to replay
        if (a.tie_n) a.tie_n[qi] = 0;
    }
    if (a.qstatus[qi] != 0 || nit == 0) {
        if (tid == 0) a.out_n[qi] = 0;
        return;
    }
    const uint32_t n = nit, nc = a.sc_meta[2ull * qi + 1];
    float* qv = (float*)(smem + pk_scan_lut_bytes(a.nsq, H16));
    const float* qg = a.q + (uint64_t)qi * a.dim;
    for (uint32_t i = tid; i < a.dim; i += NT) qv[i] = qg[i];
    const uint4* cg = a.sc_cells + (uint64_t)qi * PK_SC_MAXC;
    for (uint32_t i = tid; i < nc; i += NT) {
        const uint4 e = cg[i];
        s_c0[i] = e.x;
        s_pre[i] = e.y;
        s_cd[i] = e.z;
    }
    if (tid == 0) s_pre[nc] = n;
    __syncthreads();
    const PkScCells cl{s_c0, s_pre, s_cd, nc};
    uint64_t x[KPT];
    if constexpr (H16) {
        __half* lut = (__half*)smem;
        const int e = pk_lut16_exp(a.qn[qi], a.cb_nmax);
        const float sc = ldexpf(1.0f, e), inv = ldexpf(1.0f, -e);
        pk_build_lut_put(qv, a.codebook, a.nsq, a.lsq,
                         [&](uint32_t i, float v) { lut[i] = __float2half_rn(__fmul_rn(v, sc)); });
        __syncthreads();
#pragma unroll
        for (int r = 0; r < KR; ++r) {
            uint64_t k2[PK_SC_PP];
            if (r * PK_SC_PP * NT < n) pk_sc_score_keys<NT, NV>(a, lut, inv, cl, r * PK_SC_PP * NT, n, k2);
            else
#pragma unroll
                for (int h = 0; h < PK_SC_PP; ++h) k2[h] = ~0ull;
#pragma unroll
            for (int h = 0; h < PK_SC_PP; ++h) x[r * PK_SC_PP + h] = k2[h];
        }
    } else {
        float* lut = (float*)smem;
        pk_build_lut(lut, qv, a.codebook, a.nsq, a.lsq);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < KR; ++r) {
            uint64_t k2[PK_SC_PP];
            if (r * PK_SC_PP * NT < n) pk_sc_score_keys<NT, NV>(a, lut, 1.0f, cl, r * PK_SC_PP * NT, n, k2);
            else
#pragma unroll
                for (int h = 0; h < PK_SC_PP; ++h) k2[h] = ~0ull;
#pragma unroll
            for (int h = 0; h < PK_SC_PP; ++h) x[r * PK_SC_PP + h] = k2[h];
        }
    }
    // key j of this thread is item (j / PP) * PP * NT + (j % PP) * NT + tid;
    // keys past n are ~0 and never selected (ftk < n below)
    auto valid = [&](int j) { return (uint32_t)((j / PK_SC_PP) * PK_SC_PP * NT + (j % PK_SC_PP) * NT) + tid < n; };
    const uint32_t ftk = a.ftk;
    uint64_t K = ~0ull;
    if (n > ftk) {
        const uint32_t kc = pk_select_rank_each(
            [&](auto sink) {
#pragma unroll
                for (int j = 0; j < KPT; ++j)
                    if (valid(j)) sink((uint32_t)(x[j] >> 32));
            },
            ftk, hist, tmp, red);
        uint32_t lt = 0, eq = 0;
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
            const uint32_t hv = (uint32_t)(x[j] >> 32);
            const bool v = valid(j);
            lt += (v && hv < kc) ? 1u : 0u;
            eq += (v && hv == kc) ? 1u : 0u;
        }
        if (tid == 0) cnt[0] = cnt[1] = 0;
        __syncthreads();
        atomicAdd(&cnt[0], lt);
        atomicAdd(&cnt[1], eq);
        __syncthreads();
        const uint32_t nlt = cnt[0], neq = cnt[1];
        if (nlt + neq == ftk) {
            K = ((uint64_t)kc << 32) | 0xffffffffull;
        } else {
            const uint32_t idc = pk_select_rank_each(
                [&](auto sink) {
#pragma unroll
                    for (int j = 0; j < KPT; ++j)
                        if (valid(j) && (uint32_t)(x[j] >> 32) == kc) sink((uint32_t)x[j]);
                },
                ftk - nlt, hist, tmp, red);
            K = ((uint64_t)kc << 32) | idc;
        }
    }
    uint32_t keep = 0;
#pragma unroll
    for (int j = 0; j < KPT; ++j) keep += (valid(j) && x[j] <= K) ? 1u : 0u;
    uint32_t tot;
    uint32_t p = pk_block_exscan(keep, tmp, &tot);
    float* ot = a.out_t + (uint64_t)qi * ftk;
    uint32_t* oid = a.out_id + (uint64_t)qi * ftk;
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
        if (valid(j) && x[j] <= K) {
            ot[p] = pk_ord2f((uint32_t)(x[j] >> 32));
            oid[p] = (uint32_t)x[j];
            ++p;
        }
    }
    if (tid == 0) a.out_n[qi] = tot;
}

// Ordered list of the batch's budget queries marked for the per-query kernel
// (one workgroup of 1024 threads, block scans in index order).
__global__ __launch_bounds__(1024) void pk_scan_overflow_list(PkWalkArgs a, uint32_t n) {
    __shared__ uint32_t tmp[1024 / 64 + 1];
    uint32_t base = 0;
    for (uint32_t i0 = 0; i0 < n; i0 += 1024) {
        const uint32_t i = i0 + threadIdx.x;
        const bool o = i < n && pk_sc_budget_query(a, i) && a.qstatus[i] == 0 &&
                       a.sc_meta[2ull * i] == PK_SC_OVERFLOW;
        uint32_t tot;
        const uint32_t p = pk_block_exscan(o ? 1u : 0u, tmp, &tot);
        if (o) a.sc_olist[base + p] = i;
        base += tot;
    }
    if (threadIdx.x == 0) *a.sc_ocount = base;
}

// LDS of pk_scan_score_cut: the lookup table and the query.
extern "C" size_t pk_scan_score_smem(uint32_t dim, uint32_t nsq, int h16) {
    return pk_scan_lut_bytes(nsq, h16) + sizeof(float) * ((dim + 3) & ~3u);
}

// Whether the three kernels serve an index of these shapes (else every budget
// query takes the per-query kernel): a lookup table within 64 KB of LDS (its
// FP32 size, so that both precisions take the same kernels).
extern "C" int pk_scan_split_ok(uint32_t dim, uint32_t nsq, uint32_t F) {
    (void)F;
    return pk_scan_score_smem(dim, nsq, 0) <= 64 * 1024;
}

// Enqueues the split scan over nq queries in two parts: the cell selection
// and the list of queries it leaves to the per-query kernel (pre), then the
// scoring and the filter cut (post).  The per-query kernel needs only the
// list, so a caller can run it on another stream between the two parts.
static hipError_t pk_scan_split_check(const PkWalkArgs& a) {
    if (!a.q_nf || !a.fdist || !a.sc_cells || !a.sc_meta || !a.sc_olist || !a.sc_ocount)
        return hipErrorInvalidValue;
    if (a.sc_capp > 8192u || a.sc_capp < a.ftk) return hipErrorInvalidValue;
    if (pk_scan_score_smem(a.dim, a.nsq, 0) > 64 * 1024) return hipErrorInvalidValue;
    return hipSuccess;
}

extern "C" hipError_t pk_launch_scan_split_pre(const PkWalkArgs* args, uint32_t nq, hipStream_t s) {
    if (nq == 0) return hipSuccess;
    const hipError_t e = pk_scan_split_check(*args);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(pk_scan_cells, dim3(nq), dim3(PK_SC_CELL_THREADS), 0, s, *args);
    hipLaunchKernelGGL(pk_scan_overflow_list, dim3(1), dim3(1024), 0, s, *args, nq);
    return hipGetLastError();
}

extern "C" hipError_t pk_launch_scan_split_post(const PkWalkArgs* args, uint32_t nq, hipStream_t s) {
    if (nq == 0) return hipSuccess;
    const PkWalkArgs& a = *args;
    const hipError_t e = pk_scan_split_check(a);
    if (e != hipSuccess) return e;
    constexpr int NT = 512;
    if (a.sc_capp > (uint32_t)(PkScCut<NT>::KPT * NT)) return hipErrorInvalidValue;
    const size_t smem = pk_scan_score_smem(a.dim, a.nsq, a.lut16);
    const uint32_t nv = (a.nsq & 15u) == 0 && a.nsq <= 64 ? a.nsq / 16u : 0u;
    if (a.lut16) {
        switch (nv) {
        case 1: hipLaunchKernelGGL((pk_scan_score_cut<NT, 1, true>), dim3(nq), dim3(NT), smem, s, a); break;
        case 2: hipLaunchKernelGGL((pk_scan_score_cut<NT, 2, true>), dim3(nq), dim3(NT), smem, s, a); break;
        case 3: hipLaunchKernelGGL((pk_scan_score_cut<NT, 3, true>), dim3(nq), dim3(NT), smem, s, a); break;
        case 4: hipLaunchKernelGGL((pk_scan_score_cut<NT, 4, true>), dim3(nq), dim3(NT), smem, s, a); break;
        default: hipLaunchKernelGGL((pk_scan_score_cut<NT, 0, true>), dim3(nq), dim3(NT), smem, s, a); break;
        }
    } else {
        switch (nv) {
        case 1: hipLaunchKernelGGL((pk_scan_score_cut<NT, 1, false>), dim3(nq), dim3(NT), smem, s, a); break;
        case 2: hipLaunchKernelGGL((pk_scan_score_cut<NT, 2, false>), dim3(nq), dim3(NT), smem, s, a); break;
        case 3: hipLaunchKernelGGL((pk_scan_score_cut<NT, 3, false>), dim3(nq), dim3(NT), smem, s, a); break;
        case 4: hipLaunchKernelGGL((pk_scan_score_cut<NT, 4, false>), dim3(nq), dim3(NT), smem, s, a); break;
        default: hipLaunchKernelGGL((pk_scan_score_cut<NT, 0, false>), dim3(nq), dim3(NT), smem, s, a); break;
        }
    }
    return hipGetLastError();
}
