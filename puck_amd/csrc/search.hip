// search.hip -- the per-query stages of PuckIndex::search on gfx950.
//
//   pk_prep_queries   normalisation (hierarchical_cluster_index.cpp:1513-1536)
//                     and |q|^2 (puck_index.cpp:515)
//   pk_lut            filter lookup table (quantization.cpp:260-270), test entry
//   pk_walk           lookup table in LDS + pivot-pruned cell walk + filter-PQ
//                     scan into the top-filter_topk set (puck_index.cpp:485-557,
//                     :346-390), one workgroup per query
//   pk_rerank         exact FP32 rerank of the filter survivors into top-k
//                     (puck_index.cpp:392-457, whether_pq=false)
//
// pk_walk reproduces the reference's sequential walk exactly while doing the
// per-point work in parallel.  The reference visits cells in (coarse rank l,
// fine rank idx) order; a cell is skipped by `break` (fd + min_dist >= pivot)
// or `continue` (stationary >= max_stationary), and the pivot and
// max_stationary only decrease.  For a chunk of 256 cells the kernel
//   1. computes every point value t = fl(2*cell_dist + offset) + LUT sum of the
//      cells that can still be scanned, keeping those below the current top;
//   2. bounds the pivot from below with P_low, the pivot the chunk would
//      reach if all of those values were accepted;
//   3. classifies each cell as sure-scanned (no break under P_low and
//      stationary < min(M, P_low - cd - fd_prev)), sure-skipped (no break
//      under P_low and stationary >= M) or uncertain;
//   4. merges the values of the longest sure prefix into the sorted top set in
//      one step, then replays the first uncertain cell exactly with the now
//      exact state, and repeats from the next cell.
// Because the final filter heap of the reference is the multiset of the
// filter_topk smallest values offered (arrival order only breaks exact ties),
// merging a sure segment at once gives the same set as the reference's
// one-by-one heap updates.  Exact ties are broken by arrival order.
#include "common.h"
#include "launch_args.h"

#define WALK_THREADS 256
#define WALK_CH 256      // cells per chunk (one per thread)
#define WALK_PMAX 512    // points buffered per chunk
#define WALK_FTK_MAX 2048   // filter_topk limit (instances: <= 1024, <= 2048)


#define PK_TIE_MAX 16
#define PK_TIE_OVERFLOW 255

// ---------------------------------------------------------------------------
// Rows at or past *nvalid (when given) are padding of a fixed-size batch:
// status 2, skipped by every later stage.  flags (when given) are cleared,
// with a kernel rather than a memset so that the pass can be graph-captured.
// One query: y = x normalised (whether_norm) or copied, |y|^2, status (0, or
// -1 for a zero-norm query under whether_norm); y may alias x.
__device__ __forceinline__ void pk_prep_one(const float* x, float* y, uint32_t dim, int whether_norm, float* qn,
                                            int32_t* status) {
    int st = 0;
    if (whether_norm) {
        float norm = __fsqrt_rn(pk_dot(x, x, dim));
        if (norm < 1e-6f) {
            st = -1;
            for (uint32_t k = 0; k < dim; ++k) y[k] = x[k];
        } else {
            float alpha = (float)(1.0 / (double)norm);
            for (uint32_t k = 0; k < dim; ++k) y[k] = alpha * x[k];
        }
    } else {
        for (uint32_t k = 0; k < dim; ++k) y[k] = x[k];
    }
    *qn = pk_dot(y, y, dim);
    *status = st;
}

__global__ void pk_prep_queries(const float* __restrict__ qin, uint32_t nq, uint32_t dim, int whether_norm,
                                float* __restrict__ qout, float* __restrict__ qn,
                                int32_t* __restrict__ status, const uint32_t* __restrict__ nvalid,
                                uint8_t* __restrict__ flags) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    if (flags) flags[i] = 0;
    float* y = qout + (uint64_t)i * dim;
    if (nvalid && i >= *nvalid) {
        for (uint32_t k = 0; k < dim; ++k) y[k] = 0.f;
        qn[i] = 0.f;
        status[i] = 2;
        return;
    }
    pk_prep_one(qin + (uint64_t)i * dim, y, dim, whether_norm, qn + i, status + i);
}

// The same for dim <= PREP_LDS_DIM: 64 queries per block of 256 threads,
// staged through LDS (rows of dim + 1 words, conflict-free): every thread
// issues its share of the loads (16-byte when vec4: dim % 4 == 0 and the
// input 16-byte aligned) before storing
// any, threads 0..63 run the per-query arithmetic, and the rows go back
// coalesced.
#define PREP_LDS_DIM 255
#define PREP_ROWS 64
__global__ __launch_bounds__(256) void pk_prep_queries_lds(const float* __restrict__ qin, uint32_t nq, uint32_t dim,
                                                           int whether_norm, float* __restrict__ qout,
                                                           float* __restrict__ qn, int32_t* __restrict__ status,
                                                           const uint32_t* __restrict__ nvalid,
                                                           uint8_t* __restrict__ flags, int vec4) {
    extern __shared__ float tile[];
    const uint32_t q0 = blockIdx.x * PREP_ROWS, nrow = nq - q0 < PREP_ROWS ? nq - q0 : PREP_ROWS;
    const uint32_t ld = dim + 1u, t = threadIdx.x, nt = blockDim.x;
    const uint32_t total = nrow * dim;
    const float* src = qin + (uint64_t)q0 * dim;
    float* dst = qout + (uint64_t)q0 * dim;
    if (vec4) {
        const uint32_t t4 = total >> 2;  // float4s of the block's rows
        for (uint32_t e0 = t; e0 < t4; e0 += 8u * nt) {
            float4 v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint32_t e = e0 + u * nt;
                v[u] = ((const float4*)src)[e < t4 ? e : t4 - 1u];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint32_t e = (e0 + u * nt) * 4u;
                if (e < total) {
                    const uint32_t r = e / dim, k = e - r * dim;  // 4 | dim: the float4 stays in one row
                    float* row = tile + r * ld + k;
                    row[0] = v[u].x;
                    row[1] = v[u].y;
                    row[2] = v[u].z;
                    row[3] = v[u].w;
                }
            }
        }
    } else {
        for (uint32_t e0 = t; e0 < total; e0 += 8u * nt) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint32_t e = e0 + u * nt;
                v[u] = src[e < total ? e : total - 1u];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint32_t e = e0 + u * nt;
                if (e < total) {
                    const uint32_t r = e / dim;
                    tile[r * ld + (e - r * dim)] = v[u];
                }
            }
        }
    }
    __syncthreads();
    if (t < nrow) {
        const uint32_t i = q0 + t;
        float* row = tile + t * ld;
        if (flags) flags[i] = 0;
        if (nvalid && i >= *nvalid) {
            for (uint32_t k = 0; k < dim; ++k) row[k] = 0.f;
            qn[i] = 0.f;
            status[i] = 2;
        } else {
            pk_prep_one(row, row, dim, whether_norm, qn + i, status + i);
        }
    }
    __syncthreads();
    for (uint32_t e = t; e < total; e += nt) {
        const uint32_t r = e / dim;
        dst[e] = tile[r * ld + (e - r * dim)];
    }
}

extern "C" hipError_t pk_launch_prep_queries_ex(const float* qin, uint32_t nq, uint32_t dim, int whether_norm,
                                                float* qout, float* qn, int32_t* status, const uint32_t* nvalid,
                                                uint8_t* flags, hipStream_t s) {
    if (nq == 0) return hipSuccess;
    if (dim <= PREP_LDS_DIM)
        hipLaunchKernelGGL(pk_prep_queries_lds, dim3((nq + PREP_ROWS - 1) / PREP_ROWS), dim3(256),
                           sizeof(float) * PREP_ROWS * (dim + 1), s, qin, nq, dim, whether_norm, qout, qn, status,
                           nvalid, flags, (dim % 4 == 0 && ((uintptr_t)qin & 15) == 0) ? 1 : 0);
    else
        hipLaunchKernelGGL(pk_prep_queries, dim3((nq + 127) / 128), dim3(128), 0, s, qin, nq, dim, whether_norm,
                           qout, qn, status, nvalid, flags);
    return hipGetLastError();
}

extern "C" hipError_t pk_launch_prep_queries(const float* qin, uint32_t nq, uint32_t dim, int whether_norm,
                                             float* qout, float* qn, int32_t* status, hipStream_t s) {
    return pk_launch_prep_queries_ex(qin, nq, dim, whether_norm, qout, qn, status, nullptr, nullptr, s);
}

// ---------------------------------------------------------------------------
__global__ void pk_lut(const float* __restrict__ q, uint32_t dim, const float* __restrict__ codebook,
                       uint32_t nsq, uint32_t lsq, float* __restrict__ lut) {
    const uint64_t qi = blockIdx.y;
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nsq * PK_KS) return;
    uint32_t m = e / PK_KS;
    lut[qi * nsq * PK_KS + e] = pk_lut_entry(q + qi * dim + m * lsq, codebook + (uint64_t)e * lsq, lsq);
}

extern "C" hipError_t pk_launch_lut(const float* q, uint32_t nq, uint32_t dim, const float* codebook,
                                    uint32_t nsq, uint32_t lsq, float* lut, hipStream_t s) {
    if (nq == 0) return hipSuccess;
    hipLaunchKernelGGL(pk_lut, dim3((nsq * PK_KS + 255) / 256, nq), dim3(256), 0, s, q, dim, codebook, nsq, lsq,
                       lut);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Walk kernel helpers.  All are called by every thread of the block with
// block-uniform arguments.

struct PkWalkSmem {
    uint64_t* S_key;   // ftk      sorted top set keys (t bits << 32 | seq)
    uint32_t* S_id;    // ftk      memory ids
    uint64_t* W_key;   // PMAX     chunk candidates (sorted after build)
    uint32_t* P_id;    // PMAX     memory id by local point index
    uint16_t* P_cell;  // PMAX     chunk cell by local point index
    uint16_t* X_pos;   // PMAX     selected W positions for a merge
    uint8_t* W_done;   // PMAX
    float* c_fd;       // CH
    float* c_st;       // CH
    uint32_t* c_b;     // CH
    uint32_t* c_pre;   // CH + 1
    uint8_t* c_cls;    // CH
    float* lut;        // nsq*256
    float* qv;         // dim
    uint32_t* red;     // 16 words of reduction scratch
    uint32_t* rb;      // 2 x 4 words: per-wave partials of pk_wmin / pk_wmax (two sets, alternating)
    uint64_t* rbm;     // 2 x 8 words: per-wave ballots of the cell classification (two sets, alternating)
    float* P_t;        // PMAX     value by local point index
    uint8_t* P_sel;    // PMAX     selected for the current merge (exact mode)
    uint2* H;          // ftk + 1  reference heap (exact mode), slot i = heap index i
    uint32_t* td_val;  // PK_TIE_MAX  dropped-at-tie entries: distance bits, memory id
    uint32_t* td_id;
    uint32_t rbs = 0;  // set of rb the next reduction writes (block-uniform, in registers)
    uint32_t rbms = 0; // set of rbm the next classification writes
};

// The reference MaxHeap (max_heap.cpp:52-100) on an LDS array.  Entries are
// (value bits, tag) pairs; slot i holds heap index i (slot 0 unused).
__device__ __forceinline__ float pk_hv(uint2 e) { return __uint_as_float(e.x); }

// The heap is run by one full wave (64 active lanes, uniform arguments).  A
// sift-down (MaxHeap::insert) takes the five levels below the current node
// in one step: lane i holds node i of that 62-node subtree (lanes
// [2^d - 2, 2^(d+1) - 2) are depth d).  Each lane decides whether its parent
// takes it (the reference's rule: the right child when present and not
// left > right; the sibling's entry comes by a lane exchange), is on the
// path when every ancestor in the subtree is taken too (bit tests of one
// ballot), and the path's entries up to the first one below the sifted value
// move up one level in one masked store.
__device__ __forceinline__ bool pk_bit(uint64_t m, uint32_t i) { return (m >> i) & 1ull; }

__device__ void pk_heap_insert_wave(uint2* H, uint32_t heap_size, uint32_t father, float v, uint32_t t) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t lv = 31u - __clz(lane + 2u);
    const uint32_t off = lane + 2u - (1u << lv);
    const bool inl = lane < 62u;
    heap_size = __builtin_amdgcn_readfirstlane(heap_size);
    const uint32_t vb = __builtin_amdgcn_readfirstlane(__float_as_uint(v));
    t = __builtin_amdgcn_readfirstlane(t);
    v = __uint_as_float(vb);
    for (;;) {
        father = __builtin_amdgcn_readfirstlane(father);
        const uint32_t idx = inl ? (father << lv) + off : 0xffffffffu;
        const bool ex = inl && idx <= heap_size;
        uint2 e = make_uint2(0u, 0u);
        if (ex) e = H[idx];
        const float ev = __uint_as_float(e.x);
        // the sibling is lane ^ 1 (node idx ^ 1): a DPP swap within the quad
        const float sv = __uint_as_float(pk_lane_xor(e.x, 1u, lane));
        const bool sex = inl && (idx ^ 1u) <= heap_size;
        const bool chosen = (off & 1u) ? (ex && !(sv > ev)) : (ex && !(sex && !(ev > sv)));
        const uint64_t C = __ballot(chosen);
        bool onp = chosen;
#pragma unroll
        for (uint32_t k = 1; k < 5; ++k)
            if (k < lv) onp = onp && pk_bit(C, (1u << k) - 2u + (off >> (lv - k)));
        const uint64_t P = __ballot(onp);
        const uint64_t S = __ballot(onp && ev < v);           // path entries below v
        const uint64_t mv = S ? (P & ((S & (~S + 1ull)) - 1ull)) : P;  // the path before the first
        if (mv == 0ull) break;
        if (pk_bit(mv, lane)) H[idx >> 1] = e;
        const uint32_t last = 63u - (uint32_t)__builtin_clzll(mv);
        const uint32_t ld = 31u - __clz(last + 2u);
        father = (father << ld) + (last + 2u - (1u << ld));
        // stopped inside the subtree, or at a depth-5 node without children
        if (S != 0ull || ld < 5u || 2u * father > heap_size) break;
    }
    if (lane == 0) H[father] = make_uint2(vb, t);
}

// An LDS word read by every lane of the wave, as a wave-uniform value (LDS
// loads are otherwise taken as per-lane, which turns the heap's scalar
// control flow into exec-mask loops).
__device__ __forceinline__ uint32_t pk_uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float pk_uni_top(const uint2* H) { return __uint_as_float(pk_uni(H[1].x)); }

__device__ void pk_heap_update_wave(uint2* H, uint32_t size, uint32_t* default_cnt, float v, uint32_t t) {
    if (v >= pk_uni_top(H)) return;
    uint32_t father = 1;
    if (*default_cnt > 0) father = (*default_cnt)--;
    pk_heap_insert_wave(H, size, father, v, t);
}

__device__ void pk_heap_reorder_wave(uint2* H, uint32_t size) {
    for (uint32_t i = 0; i < size; i++) {
        const uint2 top = H[1];
        const uint32_t hs = size - i;
        const uint2 last = H[hs];
        pk_heap_insert_wave(H, hs - 1, 1, __uint_as_float(pk_uni(last.x)), pk_uni(last.y));
        if ((threadIdx.x & 63u) == 0) H[hs] = top;
    }
}

// Offers get(0..n-1) (value, tag) to the heap in order, as a loop of
// max_heap_update calls would; run by one full wave.  Each group of 64 is
// first tested against the heap top at the start of the group (the top only
// decreases, so a value that fails then fails at its turn too): LT offers a
// value only when it is below the top (callers that test `<` first),
// otherwise unless it is at or above it (max_heap_update's own test).
template <bool LT, typename GetF>
__device__ void pk_heap_offer_wave(uint2* H, uint32_t size, uint32_t* default_cnt, uint32_t n, GetF get) {
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t b = 0; b < n; b += 64) {
        const uint32_t i = b + lane;
        float v = 0.f;
        uint32_t tg = 0;
        if (i < n) get(i, v, tg);
        const float top = pk_hv(H[1]);
        uint64_t m = __ballot(i < n && (LT ? v < top : !(v >= top)));
        while (m) {
            const uint32_t j = (uint32_t)__builtin_ctzll(m);
            m &= m - 1;
            const float vj = __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(v), j));
            const uint32_t tj = __builtin_amdgcn_readlane(tg, j);
            if (LT && !(vj < pk_uni_top(H))) continue;
            pk_heap_update_wave(H, size, default_cnt, vj, tj);
        }
    }
}

__device__ void pk_heap_init(uint2* H, uint32_t size) {
    const uint2 sent = make_uint2(0x7f7f7f7fu, 0u);
    for (uint32_t i = threadIdx.x; i <= size; i += blockDim.x) H[i] = sent;
}

// Block min / max of the walk's 256 threads with one barrier: a wave
// reduction by lane exchanges, one partial per wave in rb, alternating between
// two sets of rb (the next reduction's barrier separates this one's reads from
// the writes of the one after it).  Every thread of the block calls them in
// the same order.
__device__ __forceinline__ uint32_t pk_wmin(uint32_t v, PkWalkSmem& s) {
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (uint32_t j = 1; j < 64u; j <<= 1) v = min(v, pk_lane_xor(v, j, lane));
    uint32_t* slot = s.rb + 4u * s.rbs;
    if (lane == 0) slot[threadIdx.x >> 6] = v;
    __syncthreads();
    s.rbs ^= 1u;
    return min(min(slot[0], slot[1]), min(slot[2], slot[3]));
}
__device__ __forceinline__ uint32_t pk_wmax(uint32_t v, PkWalkSmem& s) {
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (uint32_t j = 1; j < 64u; j <<= 1) v = max(v, pk_lane_xor(v, j, lane));
    uint32_t* slot = s.rb + 4u * s.rbs;
    if (lane == 0) slot[threadIdx.x >> 6] = v;
    __syncthreads();
    s.rbs ^= 1u;
    return max(max(slot[0], slot[1]), max(slot[2], slot[3]));
}

// Bitonic sort of the n <= WALK_PMAX keys v[0..n) by the walk's 256 threads
// with the keys in registers: element e sits in wave e / 128, lane e % 64, slot
// (e / 64) % 2; padded with ~0 to n2 = 2^ceil(log2 n).  Exchanges at distance
// 64 stay in the lane, those below 64 move between lanes (pk_lane_xor64); only
// the distances >= 128 (three stages at n2 = 512) go through LDS with barriers.
static_assert(WALK_THREADS == 256 && WALK_PMAX == 512, "pk_walk_sort holds 2 keys per thread in 4 waves");
__device__ __forceinline__ void pk_walk_sort(uint64_t* v, uint32_t n) {
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t n2 = 2;
    while (n2 < n) n2 <<= 1;
    const uint32_t e0 = w * 128u + lane, e1 = e0 + 64u;
    const bool on = w * 128u < n2;
    uint64_t x0 = ~0ull, x1 = ~0ull;
    if (on) {  // a wave reads and (at distances >= 128) first writes only its own entries
        if (e0 < n) x0 = v[e0];
        if (e1 < n) x1 = v[e1];
    }
#pragma unroll
    for (uint32_t k = 2; k <= WALK_PMAX; k <<= 1) {
        if (k > n2) continue;  // block-uniform
#pragma unroll
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            if (j >= 128u) {
                if (on) { v[e0] = x0; v[e1] = x1; }
                __syncthreads();
                if (on) {
                    const uint64_t y0 = v[e0 ^ j], y1 = v[e1 ^ j];
                    x0 = pk_cx(x0, y0, (e0 & j) == 0u, (e0 & k) == 0u);
                    x1 = pk_cx(x1, y1, (e1 & j) == 0u, (e1 & k) == 0u);
                }
                __syncthreads();
            } else if (j == 64u) {
                // e0 and e1 = e0 + 64 are the pair; they agree on bit k (k >= 128)
                const bool up = (e0 & k) == 0u;
                const bool lt = x0 < x1;
                const uint64_t lo = lt ? x0 : x1, hi = lt ? x1 : x0;
                x0 = up ? lo : hi;
                x1 = up ? hi : lo;
            } else {
                const bool lower = (lane & j) == 0u;
                const uint64_t y0 = pk_lane_xor64(x0, j, lane);
                x0 = pk_cx(x0, y0, lower, (e0 & k) == 0u);
                if (n2 > 64u) {  // at n2 <= 64 slot 1 holds only padding
                    const uint64_t y1 = pk_lane_xor64(x1, j, lane);
                    x1 = pk_cx(x1, y1, lower, (e1 & k) == 0u);
                }
            }
        }
    }
    if (on) {
        if (e0 < n) v[e0] = x0;
        if (e1 < n) v[e1] = x1;
    }
    __syncthreads();
}

// Values of local points [p_lo, p_hi) of the chunk's active cells; keeps those
// below `thr`.  Returns the number kept (W_key[0..n)).
__device__ __forceinline__ uint32_t pk_walk_values(const PkWalkArgs& a, PkWalkSmem& s, uint32_t cend, uint32_t p_lo,
                                   uint32_t p_hi, float cd, float thr, uint32_t seqbase) {
    if (threadIdx.x == 0) s.red[2] = 0;
    __syncthreads();
    const uint32_t nsq = a.nsq;
    const bool vec = (nsq & 15) == 0 && nsq <= 64;
    const uint32_t ng = nsq >> 4;
    // two points per thread and pass, every record load issued before the
    // first is used (a lane past p_hi loads the last point again: loads under
    // a lane condition are waited for one at a time)
    for (uint32_t pb = p_lo; pb < p_hi; pb += 2u * blockDim.x) {
        uint32_t mids[2], ts[2];
        float offs[2];
        uint4 cw[2][4];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t p = pb + h * blockDim.x + threadIdx.x;
            const uint32_t pc = p < p_hi ? p : p_hi - 1u;
            // cell containing pc: last t with c_pre[t] <= pc (c_pre[cend] = total)
            uint32_t lo = 0, hi = cend;  // invariant c_pre[lo] <= pc < c_pre[hi]
            while (hi - lo > 1) {
                uint32_t mid = (lo + hi) >> 1;
                if (s.c_pre[mid] <= pc) lo = mid; else hi = mid;
            }
            ts[h] = lo;
            mids[h] = s.c_b[lo] + (pc - s.c_pre[lo]);
            offs[h] = a.rec_off[mids[h]];
            if (vec) {
                const uint4* cp = (const uint4*)(a.rec_codes + (uint64_t)mids[h] * nsq);
#pragma unroll
                for (int g = 0; g < 4; ++g)
                    if ((uint32_t)g < ng) cw[h][g] = cp[g];
            }
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t p = pb + h * blockDim.x + threadIdx.x;
            if (p >= p_hi) continue;
            const uint32_t t = ts[h], mid = mids[h];
            const float cell_dist = (cd + s.c_st[t]) + s.c_fd[t];
            float tv = 2.0f * cell_dist + offs[h];   // 2x is exact: one rounding, as in double
            if (!(tv < thr)) continue;
            float acc[4] = {0.f, 0.f, 0.f, 0.f};
            if (vec) {
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    if ((uint32_t)g >= ng) break;
                    const uint32_t ws[4] = {cw[h][g].x, cw[h][g].y, cw[h][g].z, cw[h][g].w};
#pragma unroll
                    for (int hh = 0; hh < 4; ++hh) {
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const uint32_t code = (ws[hh] >> (8 * j)) & 0xffu;
                            acc[j] = acc[j] + s.lut[(16 * g + 4 * hh + j) * PK_KS + code];
                        }
                    }
                }
            } else {
                const uint8_t* codes = a.rec_codes + (uint64_t)mid * nsq;
                for (uint32_t m = 0; m < nsq; ++m) acc[m & 3] = acc[m & 3] + s.lut[m * PK_KS + codes[m]];
            }
            tv = tv + ((acc[0] + acc[1]) + (acc[2] + acc[3]));
            if (!(tv < thr)) continue;
            uint32_t slot = atomicAdd(&s.red[2], 1u);
            s.W_key[slot] = ((uint64_t)pk_f2ord(tv) << 32) | (uint64_t)(seqbase + p);
            s.P_id[p - p_lo] = mid;
            s.P_cell[p - p_lo] = (uint16_t)t;
            s.P_t[p - p_lo] = tv;
        }
    }
    __syncthreads();
    const uint32_t n = s.red[2];  // rewritten only after later barriers
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s.W_done[i] = 0;
    // sort W_key[0..n) ascending (the keys are distinct: their low word is
    // the point's sequence number); the sort ends with a barrier
    if (n > 1) pk_walk_sort(s.W_key, n);
    else __syncthreads();
    return n;
}

// Number of keys in sorted arr[0..n) strictly below key.
__device__ __forceinline__ uint32_t pk_lower_bound(const uint64_t* arr, uint32_t n, uint64_t key) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (arr[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Selects W positions r < nW with !W_done[r] and cell (by P_cell) in [clo, chi)
// and c_cls[cell] == want (want == 0xff: any cell, used for big cells and T_low).
// Compacts them into X_pos; returns the count.
__device__ __forceinline__ uint32_t pk_walk_select(PkWalkSmem& s, uint32_t nW, uint32_t clo, uint32_t chi, uint32_t want,
                                   uint32_t seqbase, uint32_t p_lo) {
    // per 256 entries: the wave's prefix by ballot, the waves' counts through
    // the alternating per-wave slots of pk_wmin (one barrier)
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t total_sel = 0;
    for (uint32_t base = 0; base < nW; base += blockDim.x) {
        uint32_t r = base + threadIdx.x;
        bool sel = false;
        if (r < nW && !s.W_done[r]) {
            uint32_t p = (uint32_t)(s.W_key[r] & 0xffffffffu) - seqbase - p_lo;
            uint32_t cell = s.P_cell[p];
            sel = cell >= clo && cell < chi && (want == 0xffu || s.c_cls[cell] == want);
        }
        const uint64_t m = __ballot(sel);
        uint32_t* slot = s.rb + 4u * s.rbs;
        if (lane == 0) slot[w] = (uint32_t)__popcll(m);
        __syncthreads();
        s.rbs ^= 1u;
        uint32_t before = 0, tot = 0;
#pragma unroll
        for (uint32_t q = 0; q < 4u; ++q) {
            const uint32_t c = slot[q];
            before += q < w ? c : 0u;
            tot += c;
        }
        if (sel) s.X_pos[total_sel + before + (uint32_t)__popcll(m & below)] = (uint16_t)r;
        total_sel += tot;
    }
    __syncthreads();
    return total_sel;
}

// Merge the m selected W entries (X_pos) into the sorted top set S (ns
// entries, capacity ftk).  Marks them consumed.  Returns the new size.
// When dry != 0 nothing is written; instead the key of rank ftk-1 of the
// union is stored in red[8..9] (or ~0 when the union has fewer than ftk).
template <int FTKM, bool EX>
__device__ __forceinline__ uint32_t pk_walk_merge(const PkWalkArgs& a, PkWalkSmem& s, uint32_t ns, uint32_t m,
                                  uint32_t seqbase, uint32_t p_lo, int dry, uint32_t plen = 0) {
    const uint32_t ftk = a.ftk;
    if (dry && threadIdx.x == 0) {
        s.red[8] = 0xffffffffu;
        s.red[9] = 0xffffffffu;
    }
    if (m == 0) {
        if (dry) {
            __syncthreads();
            if (threadIdx.x == 0 && ns >= ftk) {
                uint64_t k = s.S_key[ftk - 1];
                s.red[8] = (uint32_t)(k >> 32);
                s.red[9] = (uint32_t)k;
            }
            __syncthreads();
        }
        return ns;
    }
    if (!dry && threadIdx.x == 0) {
        s.red[10] = 0u;
        s.red[11] = 0u;
    }
    uint64_t sk[FTKM / WALK_THREADS];
    uint32_t sid[FTKM / WALK_THREADS], spos[FTKM / WALK_THREADS];
    uint64_t xk[WALK_PMAX / WALK_THREADS];
    uint32_t xid[WALK_PMAX / WALK_THREADS], xpos[WALK_PMAX / WALK_THREADS];
    // ranks of S entries among X (X keys via W_key[X_pos[.]], sorted since W is)
#pragma unroll
    for (int e = 0; e < FTKM / WALK_THREADS; ++e) {
        uint32_t i = threadIdx.x + e * WALK_THREADS;
        spos[e] = 0xffffffffu;
        if (i < ns) {
            uint64_t key = s.S_key[i];
            uint32_t lo = 0, hi = m;
            while (lo < hi) {
                uint32_t mid = (lo + hi) >> 1;
                if (s.W_key[s.X_pos[mid]] < key) lo = mid + 1; else hi = mid;
            }
            sk[e] = key;
            sid[e] = s.S_id[i];
            spos[e] = i + lo;
        }
    }
#pragma unroll
    for (int e = 0; e < WALK_PMAX / WALK_THREADS; ++e) {
        uint32_t r = threadIdx.x + e * WALK_THREADS;
        xpos[e] = 0xffffffffu;
        if (r < m) {
            uint32_t wpos = s.X_pos[r];
            uint64_t key = s.W_key[wpos];
            uint32_t p = (uint32_t)(key & 0xffffffffu) - seqbase - p_lo;
            xk[e] = key;
            xid[e] = s.P_id[p];
            xpos[e] = r + pk_lower_bound(s.S_key, ns, key);
        }
    }
    __syncthreads();
    const uint32_t target = ftk - 1;
    if (dry) {
#pragma unroll
        for (int e = 0; e < FTKM / WALK_THREADS; ++e)
            if (spos[e] == target) { s.red[8] = (uint32_t)(sk[e] >> 32); s.red[9] = (uint32_t)sk[e]; }
#pragma unroll
        for (int e = 0; e < WALK_PMAX / WALK_THREADS; ++e)
            if (xpos[e] == target) { s.red[8] = (uint32_t)(xk[e] >> 32); s.red[9] = (uint32_t)xk[e]; }
        __syncthreads();
        return ns;
    }
    // boundary tie: the first dropped key has the same distance as the last
    // kept (red[10], red[11]: cleared before the searches' barrier)
#pragma unroll
    for (int e = 0; e < FTKM / WALK_THREADS; ++e) {
        if (spos[e] < ftk) { s.S_key[spos[e]] = sk[e]; s.S_id[spos[e]] = sid[e]; }
        if (spos[e] == ftk) s.red[10] = (uint32_t)(sk[e] >> 32);
        if (spos[e] == ftk - 1) s.red[11] = (uint32_t)(sk[e] >> 32);
    }
#pragma unroll
    for (int e = 0; e < WALK_PMAX / WALK_THREADS; ++e) {
        if (xpos[e] < ftk) { s.S_key[xpos[e]] = xk[e]; s.S_id[xpos[e]] = xid[e]; }
        if (xpos[e] == ftk) s.red[10] = (uint32_t)(xk[e] >> 32);
        if (xpos[e] == ftk - 1) s.red[11] = (uint32_t)(xk[e] >> 32);
    }
    for (uint32_t r = threadIdx.x; r < m; r += blockDim.x) {
        uint32_t wpos = s.X_pos[r];
        s.W_done[wpos] = 1;
        if (EX) s.P_sel[(uint32_t)(s.W_key[wpos] & 0xffffffffu) - seqbase - p_lo] = 1;
    }
    __syncthreads();
    // Entries dropped with the same distance as the new top: the reference may
    // have kept any of them instead.  Remember them (the list only keeps
    // entries at the current top; older ones can no longer be in either set).
    if (!EX && s.red[10] != 0u && s.red[10] == s.red[11]) {
        const uint32_t K = s.red[11];
        if (threadIdx.x == 0) {
            uint32_t n0 = s.red[14], w = 0;
            if (n0 != PK_TIE_OVERFLOW) {
                for (uint32_t i = 0; i < n0; ++i)
                    if (s.td_val[i] == K) { s.td_val[w] = K; s.td_id[w] = s.td_id[i]; ++w; }
                s.red[14] = w;
            }
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < FTKM / WALK_THREADS; ++e)
            if (spos[e] != 0xffffffffu && spos[e] >= ftk && (uint32_t)(sk[e] >> 32) == K) {
                uint32_t slot = atomicAdd(&s.red[14], 1u);
                if (slot < PK_TIE_MAX) { s.td_val[slot] = K; s.td_id[slot] = sid[e]; }
            }
#pragma unroll
        for (int e = 0; e < WALK_PMAX / WALK_THREADS; ++e)
            if (xpos[e] != 0xffffffffu && xpos[e] >= ftk && (uint32_t)(xk[e] >> 32) == K) {
                uint32_t slot = atomicAdd(&s.red[14], 1u);
                if (slot < PK_TIE_MAX) { s.td_val[slot] = K; s.td_id[slot] = xid[e]; }
            }
        __syncthreads();
        if (threadIdx.x == 0 && s.red[14] > PK_TIE_MAX) s.red[14] = PK_TIE_OVERFLOW;
        __syncthreads();
    }
    if (EX) {
        // replay the accepted values in arrival order through the reference heap
        uint32_t cnt = 0;
        for (uint32_t base = 0; base < plen; base += blockDim.x) {
            uint32_t p = base + threadIdx.x;
            uint32_t sel = (p < plen && s.P_sel[p]) ? 1u : 0u;
            uint32_t tot;
            uint32_t pre = pk_block_exscan(sel, s.red + 4, &tot);
            if (sel) { s.X_pos[cnt + pre] = (uint16_t)p; s.P_sel[p] = 0; }
            cnt += tot;
        }
        __syncthreads();
        if (threadIdx.x < 64) {
            uint32_t dc = pk_uni(s.red[13]);
            pk_heap_offer_wave<false>(s.H, ftk, &dc, cnt, [&](uint32_t r, float& v, uint32_t& tg) {
                const uint32_t p = s.X_pos[r];
                v = s.P_t[p];
                tg = s.P_id[p];
            });
            if (threadIdx.x == 0) s.red[13] = dc;
        }
        __syncthreads();
    }
    uint32_t nn = ns + m;
    return nn < ftk ? nn : ftk;
}

// k-th smallest (1-based) of the 64-bit keys v[0..n) in LDS: MSB-first radix
// select, 8-bit digits.  blockDim.x >= 256; hist has 258 words.
__device__ uint64_t pk_select_rank64(const uint64_t* v, uint32_t n, uint32_t want, uint32_t* hist, uint32_t* tmp) {
    uint64_t prefix = 0, mask = 0;
    for (int shift = 56; shift >= 0; shift -= 8) {
        if (threadIdx.x < 256) hist[threadIdx.x] = 0;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
            const uint64_t k = v[i];
            if ((k & mask) == prefix) atomicAdd(&hist[(uint32_t)(k >> shift) & 255u], 1u);
        }
        __syncthreads();
        const uint32_t c = threadIdx.x < 256 ? hist[threadIdx.x] : 0u;
        uint32_t tot;
        const uint32_t before = pk_block_exscan(c, tmp, &tot);
        if (c > 0 && before < want && want <= before + c) {
            hist[256] = threadIdx.x;
            hist[257] = want - before;
        }
        __syncthreads();
        prefix |= (uint64_t)hist[256] << shift;
        mask |= 255ull << shift;
        want = hist[257];
        __syncthreads();
    }
    return prefix;
}

// Keeps the keys <= K of v[0..n) (in place, order kept); returns the count.
__device__ uint32_t pk_keep_le(uint64_t* v, uint32_t n, uint64_t K, uint32_t* tmp) {
    uint32_t out = 0;
    for (uint32_t r0 = 0; r0 < n; r0 += blockDim.x) {
        const uint32_t i = r0 + threadIdx.x;
        const uint64_t x = i < n ? v[i] : ~0ull;
        const uint32_t keep = (i < n && x <= K) ? 1u : 0u;
        uint32_t tot;
        const uint32_t pre = pk_block_exscan(keep, tmp, &tot);  // all reads of this round precede the writes
        if (keep) v[out + pre] = x;
        out += tot;
        __syncthreads();
    }
    return out;
}

// The same with one scan: thread t takes the contiguous entries
// [t*c, t*c + c) (c = ceil(n / blockDim.x) <= KEEP_REG) into registers, so
// the kept entries keep their order.  out == v: in place; otherwise out gets
// the kept keys split into (distance, id) rows out_t / out_id (out null).
#ifndef KEEP_REG
#define KEEP_REG 16
#endif
__device__ uint32_t pk_keep_le_reg(uint64_t* v, uint32_t n, uint64_t K, uint32_t* tmp, float* out_t,
                                   uint32_t* out_id) {
    const uint32_t c = (n + blockDim.x - 1) / blockDim.x, b0 = threadIdx.x * c;
    uint64_t x[KEEP_REG];
    uint32_t keep = 0;
#pragma unroll
    for (int j = 0; j < KEEP_REG; ++j) {
        const uint32_t i = b0 + j;
        x[j] = ((uint32_t)j < c && i < n) ? v[i] : ~0ull;
        keep += ((uint32_t)j < c && i < n && x[j] <= K) ? 1u : 0u;
    }
    uint32_t tot;
    uint32_t p = pk_block_exscan(keep, tmp, &tot);  // every read precedes every write
#pragma unroll
    for (int j = 0; j < KEEP_REG; ++j) {
        const uint32_t i = b0 + j;
        if ((uint32_t)j < c && i < n && x[j] <= K) {
            if (out_t) {
                out_t[p] = pk_ord2f((uint32_t)(x[j] >> 32));
                out_id[p] = (uint32_t)x[j];
            } else {
                v[p] = x[j];
            }
            ++p;
        }
    }
    __syncthreads();
    return tot;
}

// Budget scan of one query (fine budget b < F; no counterpart in the
// reference).  In each walked coarse list the cells whose cell distance
// (cd + stationary) + fine distance is at most the b-th smallest of the list
// are scored in full: every point's filter distance is computed as the
// reference's compute_quantized_distance computes it, and the filter set is
// the ftk smallest (distance, memory id) keys, kept in an unsorted LDS buffer
// (cap entries, R points scored per round).  When the buffer could overflow it is
// cut back to the keys whose distance is at most the ftk-th smallest distance
// (a select on the distance bits); the exact (distance, id) cut to ftk
// entries is made once, at the end.  The result does not depend on any
// visiting order, so it needs no replay.
//
// Cells are taken in segments of SCAN_THREADS * SCAN_CPT (thread t holds
// cells seg + j * SCAN_THREADS + t in registers; one segment when F fits),
// and the selected cells of a segment are scored in batches of up to SCAN_CB
// cells, so a list costs one chain of record loads per R points rather than
// one per window of cells.
#ifndef SCAN_THREADS
#define SCAN_THREADS 256
#endif
#ifndef SCAN_CPT
#define SCAN_CPT 8
#endif
#ifndef SCAN_CB
#define SCAN_CB 512
#endif
#ifndef SCAN_PP
#define SCAN_PP 4  // points per thread and round
#endif
#define SCAN_RED 32  // reduction words: 0-3 counters, 4-13 scan, 16-31 and/or (up to 512 threads)

template <int FTKM>
__device__ void pk_scan_budget(const PkWalkArgs& a, PkWalkSmem& s, uint32_t qi, uint32_t b, uint64_t* B,
                               uint32_t cap, uint32_t R, uint32_t* hist) {
    const uint32_t F = a.F, ftk = a.ftk, nsq = a.nsq;
    const uint32_t tid = threadIdx.x, NT = SCAN_THREADS, SEG = SCAN_THREADS * SCAN_CPT;
    uint32_t* tmp = s.red + 4;
    uint32_t* rao = s.red + 16;
    uint32_t n = 0;
    uint64_t K = ~0ull;  // keys above K cannot enter the filter set
    uint32_t scc = a.scc;
    if (a.q_scc) scc = a.q_scc[qi] < scc ? a.q_scc[qi] : scc;
    const float* fdq = a.fdist + (uint64_t)qi * F;
    uint32_t st_cells = 0, st_pts = 0;
    auto hi32_each = [&](auto sink) {
        for (uint32_t i = tid; i < n; i += NT) sink((uint32_t)(B[i] >> 32));
    };
    for (uint32_t l = 0; l < scc; ++l) {
        const uint32_t c = a.csel_id[(uint64_t)qi * a.cstride + l];
        const float cd = a.csel_d[(uint64_t)qi * a.cstride + l];
        const float* stc = a.stationary + (uint64_t)c * F;
        const uint32_t* csc = a.cell_start + (uint64_t)c * F;
        const bool one = F <= SEG;
        float cv[SCAN_CPT];
        if (one) {
            // unconditional loads at in-bounds cells (conditional ones are issued one at a time)
#pragma unroll
            for (int j = 0; j < SCAN_CPT; ++j) {
                const uint32_t k = min(j * NT + tid, F - 1u);
                cv[j] = (cd + stc[k]) + fdq[k];
            }
        }
        if (l == 0) {
            // the lookup table, built while the first list's cell keys load
            pk_build_filter_lut(s.lut, s.qv, a.codebook, nsq, a.lsq, a.lut16, a.qn[qi], a.cb_nmax);
            __syncthreads();
        }
        // the list's cut: b-th smallest cell key
        uint32_t kcut;
        if (one)
            kcut = pk_select_rank_each(
                [&](auto sink) {
#pragma unroll
                    for (int j = 0; j < SCAN_CPT; ++j)
                        if (j * NT + tid < F) sink(pk_f2ord(cv[j]));
                },
                b, hist, tmp, rao);
        else
            kcut = pk_select_rank_each(
                [&](auto sink) {
                    for (uint32_t k = tid; k < F; k += NT) sink(pk_f2ord((cd + stc[k]) + fdq[k]));
                },
                b, hist, tmp, rao);
        for (uint32_t seg = 0; seg < F; seg += SEG) {
            // selected non-empty cells of the segment held by this thread
            uint32_t cb[SCAN_CPT], sz[SCAN_CPT];
            uint32_t ns = 0, np = 0;
            if (!one) {
#pragma unroll
                for (int j = 0; j < SCAN_CPT; ++j) {
                    const uint32_t k = min(seg + j * NT + tid, F - 1u);
                    cv[j] = (cd + stc[k]) + fdq[k];
                }
            }
#pragma unroll
            for (int j = 0; j < SCAN_CPT; ++j) {  // cell starts of every held cell, in bounds
                const uint32_t k = min(seg + j * NT + tid, F - 1u);
                cb[j] = csc[k];
                sz[j] = csc[k + 1];
            }
#pragma unroll
            for (int j = 0; j < SCAN_CPT; ++j) {
                const uint32_t k = seg + j * NT + tid;
                sz[j] = (k < F && pk_f2ord(cv[j]) <= kcut) ? sz[j] - cb[j] : 0u;
                ns += sz[j] > 0 ? 1u : 0u;
                np += sz[j];
            }
            uint32_t nsel, npts;
            const uint32_t pos0 = pk_block_exscan(ns, tmp, &nsel);
            const uint32_t pp0 = pk_block_exscan(np, tmp, &npts);
            st_cells += nsel;
            st_pts += npts;
            for (uint32_t b0 = 0; b0 < nsel; b0 += SCAN_CB) {
                const uint32_t ne = nsel - b0 < SCAN_CB ? nsel - b0 : SCAN_CB;
                // the batch's cells with their first point (segment order)
                uint32_t pos = pos0, pp = pp0;
#pragma unroll
                for (int j = 0; j < SCAN_CPT; ++j) {
                    if (sz[j] > 0) {
                        if (pos >= b0 && pos < b0 + ne) {
                            s.c_b[pos - b0] = cb[j];
                            s.c_fd[pos - b0] = cv[j];
                            s.c_pre[pos - b0] = pp;
                        }
                        if (pos == b0 + ne) s.red[1] = pp;  // first point of the next batch
                        ++pos;
                        pp += sz[j];
                    }
                }
                __syncthreads();
                const uint32_t pbase = s.c_pre[0];
                const uint32_t pend = b0 + ne < nsel ? s.red[1] : npts;
                if (tid == 0) s.c_pre[ne] = pend;
                const uint32_t tot = pend - pbase;
                __syncthreads();
                for (uint32_t p0 = 0; p0 < tot; p0 += R) {
                    if (n + R > cap) {
                        const uint32_t kc = pk_select_rank_each(hi32_each, ftk, hist, tmp, rao);
                        K = ((uint64_t)kc << 32) | 0xffffffffull;
                        n = n <= KEEP_REG * NT ? pk_keep_le_reg(B, n, K, tmp, nullptr, nullptr) : pk_keep_le(B, n, K, tmp);
                        if (n + R > cap) {  // many equal distances at the cut: cut exactly
                            K = pk_select_rank64(B, n, ftk, hist, tmp);
                            n = n <= KEEP_REG * NT ? pk_keep_le_reg(B, n, K, tmp, nullptr, nullptr)
                                                   : pk_keep_le(B, n, K, tmp);
                        }
                    }
                    if (tid == 0) s.red[2] = n;
                    __syncthreads();
                    // all records of the round first, then their filter distances
                    constexpr int PP = SCAN_PP;
                    const bool vec = (nsq & 15) == 0 && nsq <= 64;
                    const uint32_t ng = nsq >> 4;
                    uint4 cw[PP][4];
                    float off[PP], cdl[PP];
                    uint32_t mids[PP];
#pragma unroll
                    for (int h = 0; h < PP; ++h) {
                        mids[h] = 0xffffffffu;
                        if (h * NT >= R) continue;  // uniform
                        // a lane past the end loads the last point again: loads under a
                        // lane condition would be issued one at a time
                        const uint32_t p = p0 + h * NT + tid;
                        const uint32_t P = pbase + (p < tot ? p : tot - 1u);
                        uint32_t lo = 0, hi = ne;  // c_pre[lo] <= P < c_pre[hi]
                        while (hi - lo > 1) {
                            const uint32_t mid = (lo + hi) >> 1;
                            if (s.c_pre[mid] <= P) lo = mid; else hi = mid;
                        }
                        const uint32_t mid = s.c_b[lo] + (P - s.c_pre[lo]);
                        mids[h] = p < tot ? mid : 0xffffffffu;
                        cdl[h] = s.c_fd[lo];
                        off[h] = a.rec_off[mid];
                        if (vec) {
                            const uint4* cp = (const uint4*)(a.rec_codes + (uint64_t)mid * nsq);
#pragma unroll
                            for (int g = 0; g < 4; ++g)
                                if ((uint32_t)g < ng) cw[h][g] = cp[g];
                        }
                    }
#pragma unroll
                    for (int h = 0; h < PP; ++h) {
                        const bool valid = mids[h] != 0xffffffffu;
                        uint64_t kk = ~0ull;
                        if (valid) {
                            float tv = 2.0f * cdl[h] + off[h];  // 2x is exact: one rounding, as in double
                            float acc[4] = {0.f, 0.f, 0.f, 0.f};
                            if (vec) {
#pragma unroll
                                for (int g = 0; g < 4; ++g) {
                                    if ((uint32_t)g >= ng) break;
                                    const uint32_t ws[4] = {cw[h][g].x, cw[h][g].y, cw[h][g].z, cw[h][g].w};
#pragma unroll
                                    for (int hh = 0; hh < 4; ++hh) {
#pragma unroll
                                        for (int j = 0; j < 4; ++j) {
                                            const uint32_t code = (ws[hh] >> (8 * j)) & 0xffu;
                                            acc[j] = acc[j] + s.lut[(16 * g + 4 * hh + j) * PK_KS + code];
                                        }
                                    }
                                }
                            } else {
                                const uint8_t* codes = a.rec_codes + (uint64_t)mids[h] * nsq;
                                for (uint32_t m = 0; m < nsq; ++m)
                                    acc[m & 3] = acc[m & 3] + s.lut[m * PK_KS + codes[m]];
                            }
                            tv = tv + ((acc[0] + acc[1]) + (acc[2] + acc[3]));
                            kk = ((uint64_t)pk_f2ord(tv) << 32) | mids[h];
                        }
                        // one LDS atomic per wave for its kept keys
                        const bool keep = valid && kk <= K;
                        const uint64_t bal = __ballot(keep);
                        if (bal) {
                            const uint32_t lane = tid & 63u;
                            uint32_t base = 0;
                            if (lane == 0) base = atomicAdd(&s.red[2], (uint32_t)__popcll(bal));
                            base = __shfl(base, 0, 64);
                            if (keep) B[base + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull))] = kk;
                        }
                    }
                    __syncthreads();
                    n = s.red[2];
                    __syncthreads();
                }
            }
        }
    }
    if (n > ftk) {
        // exact cut: the ftk smallest (distance, id) keys
        const uint32_t kc = pk_select_rank_each(hi32_each, ftk, hist, tmp, rao);
        uint32_t nle = 0;
        {
            if (tid == 0) s.red[3] = 0;
            __syncthreads();
            uint32_t c = 0;
            for (uint32_t i = tid; i < n; i += NT) c += (uint32_t)(B[i] >> 32) <= kc ? 1u : 0u;
            atomicAdd(&s.red[3], c);
            __syncthreads();
            nle = s.red[3];
            __syncthreads();
        }
        K = nle == ftk ? (((uint64_t)kc << 32) | 0xffffffffull) : pk_select_rank64(B, n, ftk, hist, tmp);
    }
    // the kept entries (all when n <= ftk) straight to the output rows
    if (n <= KEEP_REG * NT) {
        n = pk_keep_le_reg(B, n, K, tmp, a.out_t + (uint64_t)qi * ftk, a.out_id + (uint64_t)qi * ftk);
    } else {
        n = pk_keep_le(B, n, K, tmp);
        for (uint32_t r = tid; r < n; r += blockDim.x) {
            a.out_t[(uint64_t)qi * ftk + r] = pk_ord2f((uint32_t)(B[r] >> 32));
            a.out_id[(uint64_t)qi * ftk + r] = (uint32_t)B[r];
        }
    }
    if (tid == 0) {
        a.out_n[qi] = n;
        if (a.flags) a.flags[qi] = 0;  // deterministic: nothing to replay
        if (a.tie_n) a.tie_n[qi] = 0;
        if (a.stats) {
            a.stats[qi * 4 + 0] = 0;
            a.stats[qi * 4 + 1] = 0;
            a.stats[qi * 4 + 2] = st_cells;
            a.stats[qi * 4 + 3] = st_pts;
        }
    }
}

__device__ __forceinline__ float pk_walk_top(const PkWalkSmem& s, uint32_t ns, uint32_t ftk) {
    return ns >= ftk ? pk_ord2f((uint32_t)(s.S_key[ftk - 1] >> 32)) : pk_sentinel();
}

template <int FTKM, bool EX>
__device__ __forceinline__ void pk_walk_query(const PkWalkArgs& a, uint32_t qi) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x;
    const uint32_t ftk = a.ftk, F = a.F, nsq = a.nsq;
    // carve LDS (8-byte arrays first)
    PkWalkSmem s;
    unsigned char* ptr = smem;
    s.S_key = (uint64_t*)ptr; ptr += sizeof(uint64_t) * ((ftk + 1) & ~1u);
    s.W_key = (uint64_t*)ptr; ptr += sizeof(uint64_t) * WALK_PMAX;
    s.rbm = (uint64_t*)ptr; ptr += sizeof(uint64_t) * 16;
    s.S_id = (uint32_t*)ptr; ptr += sizeof(uint32_t) * ((ftk + 3) & ~3u);
    s.P_id = (uint32_t*)ptr; ptr += sizeof(uint32_t) * WALK_PMAX;
    s.c_b = (uint32_t*)ptr; ptr += sizeof(uint32_t) * WALK_CH;
    s.c_pre = (uint32_t*)ptr; ptr += sizeof(uint32_t) * (WALK_CH + 4);
    s.red = (uint32_t*)ptr; ptr += sizeof(uint32_t) * 16;
    s.rb = (uint32_t*)ptr; ptr += sizeof(uint32_t) * 8;
    s.c_fd = (float*)ptr; ptr += sizeof(float) * WALK_CH;
    s.c_st = (float*)ptr; ptr += sizeof(float) * WALK_CH;
    s.qv = (float*)ptr; ptr += sizeof(float) * ((a.dim + 3) & ~3u);
    s.lut = (float*)ptr; ptr += sizeof(float) * nsq * PK_KS;
    s.P_cell = (uint16_t*)ptr; ptr += sizeof(uint16_t) * WALK_PMAX;
    s.X_pos = (uint16_t*)ptr; ptr += sizeof(uint16_t) * WALK_PMAX;
    s.W_done = (uint8_t*)ptr; ptr += WALK_PMAX;
    s.c_cls = (uint8_t*)ptr; ptr += WALK_CH;
    s.P_t = (float*)ptr; ptr += sizeof(float) * WALK_PMAX;
    s.P_sel = (uint8_t*)ptr; ptr += WALK_PMAX;
    ptr = (unsigned char*)(((uintptr_t)ptr + 15) & ~(uintptr_t)15);
    s.H = (uint2*)ptr; ptr += EX ? sizeof(uint2) * ((ftk + 2) & ~1u) : 0;
    s.td_val = (uint32_t*)ptr; ptr += sizeof(uint32_t) * PK_TIE_MAX;
    s.td_id = (uint32_t*)ptr; ptr += sizeof(uint32_t) * PK_TIE_MAX;

    if (a.q_nf && a.q_nf[qi] != 0u && a.q_nf[qi] < F) return;  // pk_scan_budget_kernel's query
    const bool fine_exact = a.fine_given && a.flags && (a.flags[qi] & 8u);
    uint32_t st_seg = 0, st_unc = 0, st_cells = 0, st_pts = 0;
    bool tie_dec = false;
    uint32_t tie_sub = 0;
    for (uint32_t i = tid; i < WALK_PMAX; i += blockDim.x) s.P_sel[i] = 0;
    if (EX) pk_heap_init(s.H, ftk);
    if (tid == 0) { s.red[13] = ftk; s.red[14] = 0u; }
    __syncthreads();
    if (a.qstatus[qi] != 0) {
        if (tid == 0) {
            a.out_n[qi] = 0;
            if (a.tie_n) a.tie_n[qi] = 0;
            if (a.stats)
                for (int k = 0; k < 4; ++k) a.stats[qi * 4 + k] = 0;
        }
        return;
    }
    const float* qg = a.q + (uint64_t)qi * a.dim;
    for (uint32_t i = tid; i < a.dim; i += blockDim.x) s.qv[i] = qg[i];
    __syncthreads();
    pk_build_filter_lut(s.lut, s.qv, a.codebook, nsq, a.lsq, a.lut16, a.qn[qi], a.cb_nmax);
    __syncthreads();

    const float qn = a.qn[qi];
    const double rr = a.rr;
    const float* fsd = a.fsort_d + (uint64_t)qi * F;
    const uint32_t* fsi = a.fsort_id + (uint64_t)qi * F;
    uint32_t ns = 0;
    float T = pk_sentinel();
    float P = pk_pivot(T, qn, rr);
    uint32_t seqbase = 0;
    const uint32_t scs = a.cstride;
    uint32_t scc = a.scc;
    if (a.q_scc) scc = a.q_scc[qi] < scc ? a.q_scc[qi] : scc;

    for (uint32_t l = 0; l < scc; ++l) {
        const uint32_t c = a.csel_id[(uint64_t)qi * scs + l];
        const float cd = a.csel_d[(uint64_t)qi * scs + l];
        const float mind = a.mdo[c] + cd;
        const float fd0 = fsd[0];
        // the reference breaks at the list's first fine cluster: skip the list
        // without staging a chunk (puck_index.cpp:528-530)
        if (fd0 + mind >= P) continue;
        float M = (P - cd) - fd0;
        const float* stc = a.stationary + (uint64_t)c * F;
        const uint32_t* csc = a.cell_start + (uint64_t)c * F;
        uint32_t j0 = 0;
        bool done = false;
        // the next chunk's cell data, loaded while this chunk is processed
        // (sorted fine entry once the chunk's end is known, the cell's
        // stationary distance and start after the chunk's values pass)
        uint32_t pf_at = 0xffffffffu, pf_k = 0, pf_b = 0, pf_n = 0;
        float pf_fd = 0.f, pf_st = 0.f;
        while (!done && j0 < F) {
            const uint32_t j = j0 + tid;
            const bool valid = j < F;
            float fdj = 0.f, stj = 0.f;
            uint32_t bj = 0, nj = 0;
            if (valid && pf_at == j0) {
                fdj = pf_fd;
                stj = pf_st;
                bj = pf_b;
                nj = pf_n;
            } else if (valid) {
                fdj = fsd[j];
                uint32_t k = fsi[j];
                stj = stc[k];
                bj = csc[k];
                nj = csc[k + 1] - bj;
            }
            const float Phi = P, Mhi = M, Thi = T;
            const bool brk = !valid || (fdj + mind >= Phi);
            const uint32_t Bhi = pk_wmin(brk ? (uint32_t)tid : 0xffffffffu, s);
            const uint32_t bhi = Bhi == 0xffffffffu ? WALK_CH : Bhi;
            const bool act = ((uint32_t)tid < bhi) && !(stj >= Mhi) && nj > 0;
            const uint32_t cnt = act ? nj : 0u;
            uint32_t tot;
            const uint32_t pre = pk_block_exscan(cnt, s.red + 4, &tot);
            uint32_t cut = 0xffffffffu;
            if (act && pre + cnt > WALK_PMAX) cut = (pre == 0) ? (uint32_t)tid + 1 : (uint32_t)tid;
            const uint32_t cutm = pk_wmin(cut, s);
            uint32_t cend = cutm < bhi ? cutm : bhi;
            // do not split a fine-distance tie group across chunks: end the
            // chunk at the group's first cell when the group starts inside it
            if (j0 + cend < F && cend > 1 && fsd[j0 + cend] == fsd[j0 + cend - 1]) {
                uint32_t g = cend - 1;
                while (g > 0 && fsd[j0 + g - 1] == fsd[j0 + g]) --g;
                if (g > 0) cend = g;
            }
            const uint32_t jn = j0 + cend + tid;
            const bool pfv = jn < F;
            if (pfv) {
                pf_fd = fsd[jn];
                pf_k = fsi[jn];
            }
            pf_at = j0 + cend;
            if ((uint32_t)tid < cend) {
                s.c_fd[tid] = fdj;
                s.c_st[tid] = stj;
                s.c_b[tid] = bj;
                s.c_pre[tid] = pre;
            }
            if ((uint32_t)tid == cend) s.c_pre[cend] = pre;
            if (cend == WALK_CH && tid == 0) s.c_pre[WALK_CH] = tot;
            const float fdprev0 = j0 > 0 ? fsd[j0 - 1] : fd0;
            // equal fine distances across a chunk boundary: the reference's heap
            // order of the tie group decides the visiting order; replay exactly
            if (j0 > 0 && fdprev0 == fsd[j0]) tie_dec = true;
            __syncthreads();
            const uint32_t ptot = s.c_pre[cend];
            const bool big = ptot > WALK_PMAX;
            uint32_t nW = 0;
            float Plow = Phi;
            if (!big) {
                nW = pk_walk_values(a, s, cend, 0, ptot, cd, Thi, seqbase);
                uint32_t m = pk_walk_select(s, nW, 0, WALK_CH, 0xffu, seqbase, 0);
                pk_walk_merge<FTKM, EX>(a, s, ns, m, seqbase, 0, 1);
                uint32_t hi = s.red[8], lo = s.red[9];
                __syncthreads();
                float Tlow = (hi == 0xffffffffu && lo == 0xffffffffu) ? pk_sentinel() : pk_ord2f(hi);
                Plow = pk_pivot(Tlow, qn, rr);
            }
            if (pfv) {
                pf_st = stc[pf_k];
                pf_b = csc[pf_k];
                pf_n = csc[pf_k + 1] - pf_b;
            }
            // While the top set is not full the pivot is exactly the sentinel
            // pivot: cells up to the one whose points complete the fill are
            // classified with it; later cells with P_low.
            uint32_t tfill = 0;
            if (ns < ftk) {
                uint32_t tf = pk_wmin(((uint32_t)tid < cend && ns + s.c_pre[tid + 1] >= ftk) ? (uint32_t)tid
                                                                                             : 0xffffffffu, s);
                tfill = tf == 0xffffffffu ? cend : tf + 1;
            }
            uint32_t t0 = 0;
            while (true) {
                st_seg++;
                const bool inr = (uint32_t)tid >= t0 && (uint32_t)tid < cend;
                uint32_t cls = 0;  // 0 uncertain, 1 scan, 2 skip
                if (inr) {
                    const float f = s.c_fd[tid], sv = s.c_st[tid];
                    // a cell whose fine distance ties with a neighbour may be
                    // visited after it in the reference's order: bound its
                    // max_stationary with its own fine distance
                    const bool tied = (tid > 0 ? s.c_fd[tid - 1] == f : (j0 > 0 && fdprev0 == f)) ||
                                      ((uint32_t)tid + 1 < cend && s.c_fd[tid + 1] == f);
                    const float fprev = tied ? f : (tid > 0 ? s.c_fd[tid - 1] : fdprev0);
                    const float Pc = (uint32_t)tid < tfill ? Phi : Plow;
                    const bool nobrk = f + mind < Pc;
                    const bool forced = big && (uint32_t)tid == cend - 1;
                    if (!forced && nobrk) {
                        float bound = (Pc - cd) - fprev;
                        bound = M < bound ? M : bound;
                        if (sv >= M) cls = 2;
                        else if (sv < bound) cls = 1;
                    }
                    s.c_cls[tid] = (uint8_t)cls;
                }
                // a member of a fine-distance tie group is sure only when its
                // neighbours in the group are: the group is then decided as a unit
                __syncthreads();
                if (inr && cls != 0) {
                    const float f = s.c_fd[tid];
                    const bool tp = (uint32_t)tid > t0 && s.c_fd[tid - 1] == f && s.c_cls[tid - 1] == 0;
                    const bool tn = (uint32_t)tid + 1 < cend && s.c_fd[tid + 1] == f && s.c_cls[tid + 1] == 0;
                    if (tp || tn) cls = 0;
                }
                __syncthreads();
                if (inr) s.c_cls[tid] = (uint8_t)cls;
                // u: the first uncertain cell (cend if none); last1: one past the
                // last sure-scan cell before u (0 if none); from the waves'
                // ballots after one barrier (alternating sets of rbm)
                uint32_t u = cend, last1 = 0;
                {
                    const uint64_t b0 = __ballot(inr && cls == 0), b1 = __ballot(inr && cls == 1);
                    uint64_t* slot = s.rbm + 8u * s.rbms;
                    if ((tid & 63) == 0) {
                        slot[2 * (tid >> 6)] = b0;
                        slot[2 * (tid >> 6) + 1] = b1;
                    }
                    __syncthreads();
                    s.rbms ^= 1u;
#pragma unroll
                    for (uint32_t q = 4; q-- > 0;) {
                        const uint64_t m0 = slot[2 * q];
                        if (m0) u = q * 64u + (uint32_t)__builtin_ctzll(m0);
                    }
#pragma unroll
                    for (uint32_t q = 0; q < 4u; ++q) {
                        const uint32_t lim = u > q * 64u ? u - q * 64u : 0u;
                        uint64_t m1 = slot[2 * q + 1];
                        if (lim < 64u) m1 &= (1ull << lim) - 1ull;
                        if (m1) last1 = q * 64u + 64u - (uint32_t)__builtin_clzll(m1);
                    }
                }
                if (nW > 0) {
                    uint32_t m = pk_walk_select(s, nW, t0, u, 1u, seqbase, 0);
                    ns = pk_walk_merge<FTKM, EX>(a, s, ns, m, seqbase, 0, 0, ptot);
                }
                T = pk_walk_top(s, ns, ftk);
                P = pk_pivot(T, qn, rr);
                if (last1 > 0) {
                    const float cand = (P - cd) - s.c_fd[last1 - 1];
                    M = cand < M ? cand : M;
                }
                if (u >= cend) break;
                st_unc++;
                const float fu = s.c_fd[u], su = s.c_st[u];
                {
                    // fine-distance tie group starting at u: the reference may
                    // visit its members in another order (its heap sort of the
                    // fine list breaks ties by heap position)
                    const bool back_tie = u > 0 ? s.c_fd[u - 1] == fu : (j0 > 0 && fdprev0 == fu);
                    uint32_t gend = u;
                    while (gend + 1 < cend && s.c_fd[gend + 1] == fu) ++gend;
                    const bool fwd_tie = gend + 1 == cend && j0 + cend < F && fsd[j0 + cend] == fu;
                    const uint32_t gs = gend - u + 1;
                    if (back_tie) tie_sub |= 0x10u;
                    if (fwd_tie) tie_sub |= 0x20u;
                    if (gs > 2) tie_sub |= 0x40u;
                    if (gs == 2 && !back_tie && !fwd_tie && !(big && gend == cend - 1)) {
                        // decide the pair under both visiting orders (dry merges)
                        if (tid == 0) { s.c_cls[u] = 3; s.c_cls[u + 1] = 4; }
                        __syncthreads();
                        uint32_t outc[2];
                        for (int ord = 0; ord < 2; ++ord) {
                            float Pw = P, Mw = M;
                            uint32_t res = 0;  // bit0 first scanned, bit1 second scanned, bit2 list ended
                            for (int k = 0; k < 2; ++k) {
                                const uint32_t cell = (ord == 0) ? u + k : u + 1 - k;
                                if (fu + mind >= Pw) { res |= 4u; break; }
                                if (s.c_st[cell] >= Mw) continue;
                                res |= (cell == u) ? 1u : 2u;
                                if (k == 0) {
                                    uint32_t mm = pk_walk_select(s, nW, cell, cell + 1, cell == u ? 3u : 4u, seqbase, 0);
                                    pk_walk_merge<FTKM, EX>(a, s, ns, mm, seqbase, 0, 1);
                                    const uint32_t hi = s.red[8], lo = s.red[9];
                                    __syncthreads();
                                    const float T1 = (hi == 0xffffffffu && lo == 0xffffffffu) ? T : pk_ord2f(hi);
                                    Pw = pk_pivot(T1, qn, rr);
                                    const float cand = (Pw - cd) - fu;
                                    Mw = cand < Mw ? cand : Mw;
                                }
                            }
                            outc[ord] = res;
                        }
                        if (outc[0] == outc[1]) {
                            const uint32_t res = outc[0];
                            if (res & 3u) {
                                if (tid == 0) {
                                    s.c_cls[u] = (res & 1u) ? 5 : 6;
                                    s.c_cls[u + 1] = (res & 2u) ? 5 : 6;
                                }
                                __syncthreads();
                                if (nW > 0) {
                                    uint32_t mm = pk_walk_select(s, nW, u, u + 2, 5u, seqbase, 0);
                                    ns = pk_walk_merge<FTKM, EX>(a, s, ns, mm, seqbase, 0, 0, ptot);
                                }
                                T = pk_walk_top(s, ns, ftk);
                                P = pk_pivot(T, qn, rr);
                                const float cand = (P - cd) - fu;
                                M = cand < M ? cand : M;
                            }
                            if (res & 4u) { done = true; break; }
                            t0 = u + 2;
                            continue;
                        }
                        tie_dec = true;
                        tie_sub |= 0x80u;
                    } else if (gs > 1 || back_tie || fwd_tie) {
                        tie_dec = true;
                    }
                }
                if (fu + mind >= P) { done = true; break; }
                if (!(su >= M)) {
                    if (big && u == cend - 1) {
                        for (uint32_t p_lo = 0; p_lo < ptot; p_lo += WALK_PMAX) {
                            uint32_t p_hi = p_lo + WALK_PMAX < ptot ? p_lo + WALK_PMAX : ptot;
                            uint32_t nb = pk_walk_values(a, s, cend, p_lo, p_hi, cd, T, seqbase);
                            uint32_t m = pk_walk_select(s, nb, 0, WALK_CH, 0xffu, seqbase, p_lo);
                            ns = pk_walk_merge<FTKM, EX>(a, s, ns, m, seqbase, p_lo, 0, p_hi - p_lo);
                            T = pk_walk_top(s, ns, ftk);
                        }
                    } else if (nW > 0) {
                        if (tid == 0) s.c_cls[u] = 3;
                        __syncthreads();
                        uint32_t m = pk_walk_select(s, nW, u, u + 1, 3u, seqbase, 0);
                        ns = pk_walk_merge<FTKM, EX>(a, s, ns, m, seqbase, 0, 0, ptot);
                    }
                    T = pk_walk_top(s, ns, ftk);
                    P = pk_pivot(T, qn, rr);
                    const float cand = (P - cd) - fu;
                    M = cand < M ? cand : M;
                }
                t0 = u + 1;
            }
            st_pts += ptot;
            st_cells += cend;
            seqbase += ptot;
            if (done) break;
            if (cend == bhi && bhi < WALK_CH) break;  // sure break (or end of the fine list)
            j0 += cend;
        }
    }
    if (EX) {
        // the filter heap as rank_topN_points reads it: heap order when full,
        // heap-sorted ascending when not full (puck_index.cpp:552-554)
        if (tid < 64 && ns < ftk) pk_heap_reorder_wave(s.H, ftk);
        __syncthreads();
        for (uint32_t r = tid; r < ns; r += blockDim.x) {
            a.out_t[(uint64_t)qi * ftk + r] = pk_hv(s.H[r + 1]);
            a.out_id[(uint64_t)qi * ftk + r] = s.H[r + 1].y;
        }
    } else {
        for (uint32_t r = tid; r < ns; r += blockDim.x) {
            a.out_t[(uint64_t)qi * ftk + r] = pk_ord2f((uint32_t)(s.S_key[r] >> 32));
            a.out_id[(uint64_t)qi * ftk + r] = s.S_id[r];
        }
    }
    if (tid == 0) {
        a.out_n[qi] = ns;
        if (a.flags) {
            if (a.fine_given) a.flags[qi] = (tie_dec && !fine_exact) ? (uint8_t)(8u | tie_sub) : (uint8_t)0;
            else if (tie_dec) a.flags[qi] |= (uint8_t)(8u | tie_sub);
        }
        if (a.tie_n) {
            uint32_t tn = 0;
            if (ns >= ftk) {
                const uint32_t K = (uint32_t)(s.S_key[ftk - 1] >> 32);
                if (s.red[14] == PK_TIE_OVERFLOW) tn = PK_TIE_OVERFLOW;
                else
                    for (uint32_t i = 0; i < s.red[14]; ++i)
                        if (s.td_val[i] == K) a.tie_ids[(uint64_t)qi * PK_TIE_MAX + tn++] = s.td_id[i];
            }
            a.tie_n[qi] = (uint8_t)tn;
        }
        if (a.stats) {
            a.stats[qi * 4 + 0] = st_seg;
            a.stats[qi * 4 + 1] = st_unc;
            a.stats[qi * 4 + 2] = st_cells;
            a.stats[qi * 4 + 3] = st_pts;
        }
    }
}

// pk_walk: one query per workgroup (qlist or identity).  pk_walk_list:
// grid-stride
// over the device list wlist[0..*wcount) of the queries that take the walk
// (the others of the batch are budget-scan queries); every workgroup leaves
// the loop when the list is exhausted.
// EX: the reference-heap form (replays), compiled apart so that the fast
// form does not carry its registers
template <int FTKM, bool EX>
__global__ __launch_bounds__(WALK_THREADS) void pk_walk(PkWalkArgs a) {
    pk_walk_query<FTKM, EX>(a, a.qlist ? a.qlist[blockIdx.x] : blockIdx.x);
}

template <int FTKM>
__global__ __launch_bounds__(WALK_THREADS) void pk_walk_list(PkWalkArgs a) {
    const uint32_t n = *a.wcount;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        pk_walk_query<FTKM, false>(a, a.wlist[i]);
        __syncthreads();
    }
}

// Ordered list of the queries of a batch that take the walk (no fine budget
// below F): one workgroup, block scans in index order.
__global__ __launch_bounds__(1024) void pk_build_walk_list(const uint32_t* __restrict__ q_nf, uint32_t n, uint32_t F,
                                                          uint32_t* __restrict__ list, uint32_t* __restrict__ count) {
    __shared__ uint32_t tmp[20];
    uint32_t base = 0;
    for (uint32_t r0 = 0; r0 < n; r0 += blockDim.x) {
        const uint32_t i = r0 + threadIdx.x;
        const uint32_t w = (i < n && !(q_nf[i] != 0u && q_nf[i] < F)) ? 1u : 0u;
        uint32_t tot;
        const uint32_t pre = pk_block_exscan(w, tmp, &tot);
        if (w) list[base + pre] = i;
        base += tot;
    }
    if (threadIdx.x == 0) *count = base;
}

// Workgroup -> query order for the per-query kernels (walk, budget scan,
// rerank): the queries grouped by their nearest coarse list (counting sort
// over csi[q * cstride]) and the sorted order dealt to the 8 XCDs in
// contiguous slices (workgroup b is dispatched to XCD b % 8), so that the
// queries reading one list's records and rows run on one L2, close in time.
// Each query is still processed once, by the same code: results do not
// depend on the order.  One workgroup of 1024 threads; C words of LDS.
#define PK_XCDS 8u
__global__ __launch_bounds__(1024) void pk_build_query_order(const uint32_t* __restrict__ csi, uint32_t cstride,
                                                            uint32_t m, uint32_t C, uint32_t* __restrict__ order) {
    extern __shared__ uint32_t bin[];  // C counters, then bin starts
    __shared__ uint32_t tmp[20];
    const uint32_t t = threadIdx.x, nt = blockDim.x;
    for (uint32_t i = t; i < C; i += nt) bin[i] = 0;
    __syncthreads();
    for (uint32_t q = t; q < m; q += nt) {
        const uint32_t c = csi[(uint64_t)q * cstride];
        atomicAdd(&bin[c < C ? c : C - 1u], 1u);
    }
    __syncthreads();
    // exclusive prefix over the bins: thread t owns bins [t*per, t*per + per)
    const uint32_t per = (C + nt - 1u) / nt, b0 = t * per;
    uint32_t sum = 0;
    for (uint32_t i = b0; i < b0 + per && i < C; ++i) sum += bin[i];
    uint32_t tot;
    uint32_t run = pk_block_exscan(sum, tmp, &tot);
    for (uint32_t i = b0; i < b0 + per && i < C; ++i) {
        const uint32_t c = bin[i];
        bin[i] = run;
        run += c;
    }
    __syncthreads();
    // sorted position -> workgroup: XCD x takes positions [start_x, start_x + n_x),
    // n_x = ceil((m - x) / 8) workgroups x, x + 8, x + 16, ...
    for (uint32_t q = t; q < m; q += nt) {
        const uint32_t c = csi[(uint64_t)q * cstride];
        const uint32_t pos = atomicAdd(&bin[c < C ? c : C - 1u], 1u);
        uint32_t x = 0, start = 0;
        for (; x + 1 < PK_XCDS; ++x) {
            const uint32_t nx = m > x ? (m - x + PK_XCDS - 1u) / PK_XCDS : 0u;
            if (pos < start + nx) break;
            start += nx;
        }
        order[x + PK_XCDS * (pos - start)] = q;
    }
}

extern "C" hipError_t pk_launch_build_query_order(const uint32_t* csi, uint32_t cstride, uint32_t m, uint32_t C,
                                                  uint32_t* order, hipStream_t s) {
    if (m == 0) return hipSuccess;
    if (C == 0 || C > 8192) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pk_build_query_order, dim3(1), dim3(1024), sizeof(uint32_t) * C, s, csi, cstride, m, C, order);
    return hipGetLastError();
}

extern "C" hipError_t pk_launch_build_walk_list(const uint32_t* q_nf, uint32_t n, uint32_t F, uint32_t* list,
                                                uint32_t* count, hipStream_t s) {
    hipLaunchKernelGGL(pk_build_walk_list, dim3(1), dim3(1024), 0, s, q_nf, n, F, list, count);
    return hipGetLastError();
}

// The budget scan as a kernel of its own (its register needs differ from the
// walk's): one block per query, returning at once for queries without a
// budget below F.  Same LDS layout as pk_walk.
template <int FTKM>
__device__ void pk_scan_budget_one(const PkWalkArgs& a, uint32_t qi, unsigned char* smem) {
    const uint32_t ftk = a.ftk, F = a.F, nsq = a.nsq;
    if (!(a.q_nf && a.q_nf[qi] != 0u && a.q_nf[qi] < F)) return;
    // own LDS plan: LUT, small arrays, then the candidate buffer (a.scan_cap keys)
    PkWalkSmem s;
    unsigned char* ptr = smem;
    s.lut = (float*)ptr; ptr += sizeof(float) * nsq * PK_KS;
    s.qv = (float*)ptr; ptr += sizeof(float) * ((a.dim + 3) & ~3u);
    s.red = (uint32_t*)ptr; ptr += sizeof(uint32_t) * SCAN_RED;
    s.c_b = (uint32_t*)ptr; ptr += sizeof(uint32_t) * SCAN_CB;
    s.c_pre = (uint32_t*)ptr; ptr += sizeof(uint32_t) * (SCAN_CB + 4);
    s.c_fd = (float*)ptr; ptr += sizeof(float) * SCAN_CB;
    uint32_t* hist = (uint32_t*)ptr; ptr += sizeof(uint32_t) * PK_SEL_HIST;
    ptr = (unsigned char*)(((uintptr_t)ptr + 15) & ~(uintptr_t)15);
    uint64_t* B = (uint64_t*)ptr;
    (void)ftk;
    if (a.qstatus[qi] != 0) {
        if (threadIdx.x == 0) {
            a.out_n[qi] = 0;
            if (a.flags) a.flags[qi] = 0;
            if (a.tie_n) a.tie_n[qi] = 0;
            if (a.stats)
                for (int k = 0; k < 4; ++k) a.stats[qi * 4 + k] = 0;
        }
        return;
    }
    const float* qg = a.q + (uint64_t)qi * a.dim;
    for (uint32_t i = threadIdx.x; i < a.dim; i += blockDim.x) s.qv[i] = qg[i];
    __syncthreads();
    (void)nsq;
    pk_scan_budget<FTKM>(a, s, qi, a.q_nf[qi], B, a.scan_cap, a.scan_round, hist);
}

template <int FTKM>
__global__ __launch_bounds__(SCAN_THREADS) void pk_scan_budget_kernel(PkWalkArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    pk_scan_budget_one<FTKM>(a, a.qlist ? a.qlist[blockIdx.x] : blockIdx.x, smem);
}

// The same over the device list sc_olist[0..*sc_ocount) (the budget queries
// the split scan of scan.hip leaves to this kernel), grid-stride.
template <int FTKM>
__global__ __launch_bounds__(SCAN_THREADS) void pk_scan_budget_list(PkWalkArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t cnt = *a.sc_ocount;
    for (uint32_t i = blockIdx.x; i < cnt; i += gridDim.x) {
        pk_scan_budget_one<FTKM>(a, a.sc_olist[i], smem);
        __syncthreads();  // every thread is done with the LDS before the next query
    }
}

extern "C" size_t pk_scan_smem_fixed(uint32_t dim, uint32_t nsq) {
    size_t b = sizeof(float) * nsq * PK_KS + sizeof(float) * ((dim + 3) & ~3u) + sizeof(uint32_t) * SCAN_RED +
               sizeof(uint32_t) * SCAN_CB + sizeof(uint32_t) * (SCAN_CB + 4) + sizeof(float) * SCAN_CB +
               sizeof(uint32_t) * PK_SEL_HIST;
    return (b + 15) & ~(size_t)15;
}

extern "C" size_t pk_walk_smem_bytes(uint32_t dim, uint32_t nsq, uint32_t ftk, int exact) {
    size_t b = 0;
    b += sizeof(float) * WALK_PMAX + WALK_PMAX + 2 * sizeof(uint32_t) * PK_TIE_MAX + 16;
    if (exact) b += sizeof(uint2) * ((ftk + 2) & ~1u);
    b += sizeof(uint64_t) * ((ftk + 1) & ~1u);
    b += sizeof(uint64_t) * WALK_PMAX + sizeof(uint64_t) * 16;
    b += sizeof(uint32_t) * ((ftk + 3) & ~3u);
    b += sizeof(uint32_t) * WALK_PMAX;
    b += sizeof(uint32_t) * WALK_CH;
    b += sizeof(uint32_t) * (WALK_CH + 4);
    b += sizeof(uint32_t) * 16 + sizeof(uint32_t) * 8;
    b += sizeof(float) * WALK_CH * 2;
    b += sizeof(float) * ((dim + 3) & ~3u);
    b += sizeof(float) * nsq * PK_KS;
    b += sizeof(uint16_t) * WALK_PMAX * 2;
    b += WALK_PMAX + WALK_CH;
    return b;
}

extern "C" uint32_t pk_walk_ftk_max(void) { return WALK_FTK_MAX; }

// With fine budgets (q_nf) and a second stream s2, the walk over the queries
// without a budget runs on s2 concurrently with the budget scan on s (event
// fork / join, capturable in a graph); ev_mid is recorded on s right after
// the scan.
extern "C" int pk_scan_split_ok(uint32_t dim, uint32_t nsq, uint32_t F);
extern "C" hipError_t pk_launch_scan_split_pre(const PkWalkArgs* args, uint32_t nq, hipStream_t s);
extern "C" hipError_t pk_launch_scan_split_post(const PkWalkArgs* args, uint32_t nq, hipStream_t s);

extern "C" hipError_t pk_launch_walk(const PkWalkArgs* args, uint32_t nq, hipStream_t s, hipEvent_t ev_mid,
                                     hipStream_t s2, hipEvent_t ev_fork, hipEvent_t ev_join) {
    if (nq == 0) return hipSuccess;
    if (args->ftk == 0 || args->ftk > WALK_FTK_MAX || args->cstride < args->scc) return hipErrorInvalidValue;
    if (args->q_nf && !args->fdist) return hipErrorInvalidValue;
    if (args->wcount && args->exact) return hipErrorInvalidValue;  // the list form is the fast walk only
    size_t smem = pk_walk_smem_bytes(args->dim, args->nsq, args->ftk, args->exact);
    // budget scan: fill the same footprint as the walk, at least 80 KB (two
    // blocks per CU) and at least ftk + SCAN_THREADS candidates, up to
    // SCAN_PP points per thread and round
    PkWalkArgs sa = *args;
    const size_t fixed = pk_scan_smem_fixed(args->dim, args->nsq);
    size_t want = smem > 80 * 1024 ? smem : 80 * 1024;
    size_t cap = want > fixed ? (want - fixed) / 8 : 0;
    if (cap < (size_t)args->ftk + SCAN_THREADS) cap = (size_t)args->ftk + SCAN_THREADS;
    size_t rnd = (cap - args->ftk) & ~(size_t)(SCAN_THREADS - 1);
    if (rnd > SCAN_PP * SCAN_THREADS) rnd = SCAN_PP * SCAN_THREADS;
    sa.scan_cap = (uint32_t)cap;
    sa.scan_round = (uint32_t)rnd;
    const size_t smem_scan = fixed + 8 * cap;
    const uint32_t wgrid = args->wcount ? (nq < 1024 ? nq : 1024) : nq;
    static PkLdsCache lds_cache;
    size_t* slot = pk_lds_slot(lds_cache);
    if (!slot) return hipErrorInvalidDevice;
    if (!*slot) {
        const void* fns[] = {(const void*)pk_walk<1024, false>, (const void*)pk_walk<2048, false>,
                             (const void*)pk_walk<1024, true>, (const void*)pk_walk<2048, true>,
                             (const void*)pk_walk_list<1024>, (const void*)pk_walk_list<2048>,
                             (const void*)pk_scan_budget_kernel<1024>, (const void*)pk_scan_budget_kernel<2048>,
                             (const void*)pk_scan_budget_list<1024>, (const void*)pk_scan_budget_list<2048>};
        size_t lo = 160 * 1024;
        for (const void* f : fns) {
            size_t a = 0;
            const hipError_t e = pk_allow_lds(f, 160 * 1024, &a);
            if (e != hipSuccess) return e;
            lo = a < lo ? a : lo;
        }
        *slot = lo;
    }
    const size_t allowed = *slot;
    if (args->q_nf && smem_scan > allowed) return hipErrorInvalidValue;
    if (smem > allowed) return hipErrorInvalidValue;
    // budget scan first (its own launch, timed between ev_mid and the end),
    // then the walk for the queries without a budget
    auto launch_walk = [&](hipStream_t st) {
        const bool lst = args->wcount != nullptr;
        const bool ex = args->exact != 0;
        if (args->ftk <= 1024) {
            if (lst) hipLaunchKernelGGL(pk_walk_list<1024>, dim3(wgrid), dim3(WALK_THREADS), smem, st, *args);
            else if (ex) hipLaunchKernelGGL((pk_walk<1024, true>), dim3(wgrid), dim3(WALK_THREADS), smem, st, *args);
            else hipLaunchKernelGGL((pk_walk<1024, false>), dim3(wgrid), dim3(WALK_THREADS), smem, st, *args);
        } else {
            if (lst) hipLaunchKernelGGL(pk_walk_list<2048>, dim3(wgrid), dim3(WALK_THREADS), smem, st, *args);
            else if (ex) hipLaunchKernelGGL((pk_walk<2048, true>), dim3(wgrid), dim3(WALK_THREADS), smem, st, *args);
            else hipLaunchKernelGGL((pk_walk<2048, false>), dim3(wgrid), dim3(WALK_THREADS), smem, st, *args);
        }
    };
    const bool fork = args->q_nf && s2 && ev_fork && ev_join;
    hipStream_t sw = fork ? s2 : s;
    hipError_t e;
    if (fork) {
        if ((e = hipEventRecord(ev_fork, s)) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(s2, ev_fork, 0)) != hipSuccess) return e;
        launch_walk(sw);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (args->q_nf) {
        if (args->sc_cells && pk_scan_split_ok(args->dim, args->nsq, args->F)) {
            // the split scan; the per-query kernel takes the queries its cell
            // selection leaves, on the second stream beside the scoring and cut
            if ((e = pk_launch_scan_split_pre(&sa, nq, s)) != hipSuccess) return e;
            if (fork) {
                if ((e = hipEventRecord(ev_fork, s)) != hipSuccess) return e;
                if ((e = hipStreamWaitEvent(s2, ev_fork, 0)) != hipSuccess) return e;
            }
            const uint32_t lg = nq < 1024 ? nq : 1024;
            if (args->ftk <= 1024)
                hipLaunchKernelGGL(pk_scan_budget_list<1024>, dim3(lg), dim3(SCAN_THREADS), smem_scan, sw, sa);
            else
                hipLaunchKernelGGL(pk_scan_budget_list<2048>, dim3(lg), dim3(SCAN_THREADS), smem_scan, sw, sa);
            if ((e = hipGetLastError()) != hipSuccess) return e;
            if ((e = pk_launch_scan_split_post(&sa, nq, s)) != hipSuccess) return e;
        } else if (args->ftk <= 1024) {
            hipLaunchKernelGGL(pk_scan_budget_kernel<1024>, dim3(nq), dim3(SCAN_THREADS), smem_scan, s, sa);
        } else {
            hipLaunchKernelGGL(pk_scan_budget_kernel<2048>, dim3(nq), dim3(SCAN_THREADS), smem_scan, s, sa);
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (ev_mid) (void)hipEventRecord(ev_mid, s);
    if (fork) {
        if ((e = hipEventRecord(ev_join, s2)) != hipSuccess) return e;
        return hipStreamWaitEvent(s, ev_join, 0);
    }
    launch_walk(s);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------

#define RERANK_THREADS 256

// Fast mode: distances sorted by (distance, filter rank).  Exact mode: the
// result MaxHeap of rank_topN_points fed in filter-heap order, then reorder().
// keys[0 .. K1-1] = the K1 smallest of the distinct keys[0 .. n-1], ascending
// (the rest of keys is left as it was).  A radix select of the K1-th smallest
// high word, starting at the highest bit in which the high words differ, then
// the entries at or below it placed by rank.  Returns false (keys untouched)
// when more than PK_PREFIX_CAP entries share or undercut that high word.
#define PK_PREFIX_CAP 256
__device__ bool pk_sorted_prefix(uint64_t* keys, uint32_t n, uint32_t K1) {
    __shared__ uint32_t hist[PK_SEL_HIST];
    __shared__ uint32_t tmp[8];
    __shared__ uint32_t red[9];
    __shared__ uint64_t sb[PK_PREFIX_CAP];
    const uint32_t t = threadIdx.x, nb = blockDim.x;
    if (t == 0) red[8] = 0;
    const uint32_t kc = pk_select_rank_each(
        [&](auto sink) {
            for (uint32_t i = t; i < n; i += nb) sink((uint32_t)(keys[i] >> 32));
        },
        K1, hist, tmp, red);
    for (uint32_t i = t; i < n; i += nb) {
        const uint64_t e = keys[i];
        if ((uint32_t)(e >> 32) <= kc) {
            const uint32_t p = atomicAdd(&red[8], 1u);
            if (p < PK_PREFIX_CAP) sb[p] = e;
        }
    }
    __syncthreads();
    const uint32_t c = red[8];
    if (c > PK_PREFIX_CAP) return false;  // uniform across the block
    for (uint32_t i = t; i < c; i += nb) {
        const uint64_t e = sb[i];
        uint32_t r = 0;
        for (uint32_t j = 0; j < c; ++j) r += sb[j] < e ? 1u : 0u;
        if (r < K1) keys[r] = e;
    }
    __syncthreads();
    return true;
}

// Second-stage PQ distance of a filter survivor (rank_topN_points,
// puck_index.cpp:402-427): (filter distance - |q|^2) + PQ offset, then + the
// LUT sum of its PQ codes in the 4-lane order of lookup_dist_table (:75-113).
__device__ __forceinline__ float pk_pq_dist(const PkRerankArgs& a, const float* __restrict__ lut, uint32_t mid,
                                            float fd, float qn) {
    const uint32_t nsq = a.pnsq;
    float t = (fd - qn) + a.pq_off[mid];
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    const uint8_t* codes = a.pq_codes + (uint64_t)mid * nsq;
    uint32_t m = 0;
    if ((nsq & 15u) == 0) {
        // 16 codes per load, issued before the lookups of a group
        for (; m < nsq; m += 16) {
            const uint4 w = *(const uint4*)(codes + m);
            const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
            for (int hh = 0; hh < 4; ++hh)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[j] = acc[j] + lut[(m + 4 * hh + j) * PK_KS + ((ws[hh] >> (8 * j)) & 0xffu)];
        }
    } else {
        for (; m < nsq; ++m) acc[m & 3] = acc[m & 3] + lut[m * PK_KS + codes[m]];
    }
    return t + ((acc[0] + acc[1]) + (acc[2] + acc[3]));
}

// T: the element type of the raw rows (float, or uint8_t for the uint8
// storage form: integer features, converted exactly).  RP: rows per 8-lane
// group and iteration of the distance loop (the gathers in flight; a tuning
// choice, every value gives the same bits).
template <typename T, int RP>
__global__ __launch_bounds__(RERANK_THREADS) void pk_rerank(PkRerankArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const T* raw = (const T*)(sizeof(T) == 1 ? (const void*)a.raw8 : (const void*)a.raw);
    const uint32_t qi = a.qlist ? a.qlist[blockIdx.x] : blockIdx.x;
    uint32_t topk = a.topk;
    if (a.q_topk) topk = (a.q_topk[qi] >= 1u && a.q_topk[qi] <= a.topk) ? a.q_topk[qi] : 0u;
    const uint32_t n = (a.qstatus[qi] == 0 && topk > 0) ? a.filt_n[qi] : 0u;
    const uint32_t ftk = a.ftk, ostr = a.out_stride;
    const bool bud = a.q_nf && a.q_nf[qi] != 0u && a.q_nf[qi] < a.F;
    const bool exact = a.exact && !bud;
    uint32_t n2 = 2;
    while (n2 < ftk) n2 <<= 1;
    uint64_t* keys = (uint64_t*)smem;
    float* qv = (float*)(smem + sizeof(uint64_t) * n2);
    float* dist = qv + ((a.dim + 3) & ~3u);
    uint2* RH = (uint2*)(((uintptr_t)(dist + ((ftk + 3) & ~3u)) + 15) & ~(uintptr_t)15);
    float* plut = (float*)(RH + ((a.topk + 2) & ~1u));  // PQ lookup table (whether_pq)
    const bool pq = a.pq_codes != nullptr;
    for (uint32_t i = threadIdx.x; i < a.dim; i += blockDim.x) qv[i] = a.q[(uint64_t)qi * a.dim + i];
    __syncthreads();
    const float qn = pq ? a.qn[qi] : 0.f;
    if (pq && n > 0) {
        pk_build_lut(plut, qv, a.pq_codebook, a.pnsq, a.plsq);
        __syncthreads();
    }
    // filter-boundary tie group (entries whose filter distance equals the final
    // top, kept or dropped): smallest exact distance among its members
    const uint32_t tn = (!exact && !bud && a.tie_n && n >= ftk) ? a.tie_n[qi] : 0u;
    const float tfin = n >= ftk ? a.filt_t[(uint64_t)qi * ftk + ftk - 1] : 0.f;
    __shared__ uint32_t tie_min;
    if (threadIdx.x == 0) tie_min = tn == PK_TIE_OVERFLOW ? 0u : 0xffffffffu;
    // 8 lanes per row: lane j holds element j of each 8-lane half block (the
    // compiled order of pk_rerank_dist); shuffles combine the lanes in that
    // order, so every lane ends with the same bits as pk_rerank_dist
    // Two rows per group and iteration, every load of a 128-element block of
    // both rows issued before the first is used (the gathers are latency
    // bound: one row is a few cache lines).
    if (pq) {
        // one survivor per thread
        for (uint32_t r = threadIdx.x; r < n2; r += blockDim.x) {
            const uint32_t rc = r < ftk ? r : ftk - 1u;
            const uint32_t fid = a.filt_id[(uint64_t)qi * ftk + rc];
            const float fd = a.filt_t[(uint64_t)qi * ftk + rc];
            const uint32_t mid = r < n ? fid : 0u;
            const float t = pk_pq_dist(a, plut, mid, fd, qn);
            uint64_t key = ~0ull;
            if (r < n) {
                key = ((uint64_t)pk_f2ord(t) << 32) | r;
                if (exact || bud) dist[r] = t;
            }
            if (!exact) keys[r] = key;
        }
    } else if (sizeof(T) == 1 && (a.dim & 15u) == 0) {
        // uint8 rows of a multiple of 16 bytes: lane j of a row's 8-lane group
        // loads whole 16-element blocks (one 16-byte load) and forms the
        // block's ((l0+l1)+(l2+l3))+((l4+l5)+l6) and l7 with
        // l_i = fma(d_i, d_i, d_{8+i}^2); the blocks are then chained in order,
        // t = S + (l7 + t) -- the arithmetic of the float rows' loop below
        const uint32_t g = threadIdx.x >> 3, j = threadIdx.x & 7, ng = blockDim.x >> 3;
        const int src = (int)(threadIdx.x & 63u & ~7u);
        const uint32_t dim = a.dim, nfull = dim >> 4;
        for (uint32_t r0 = 0; r0 < n2; r0 += RP * ng) {
            uint32_t rr[RP];
            const uint8_t* x[RP];
            float t[RP];
#pragma unroll
            for (int h = 0; h < RP; ++h) {
                rr[h] = r0 + h * ng + g;
                const uint32_t fid = a.filt_id[(uint64_t)qi * ftk + (rr[h] < ftk ? rr[h] : ftk - 1u)];
                x[h] = (const uint8_t*)raw + (uint64_t)(rr[h] < n ? fid : 0u) * dim;
                t[h] = 0.0f;
            }
            for (uint32_t k0 = 0; k0 < nfull; k0 += 8) {
                const uint32_t sb = k0 + j < nfull ? k0 + j : k0;  // in-bounds block for every lane
                uint4 w[RP];
#pragma unroll
                for (int h = 0; h < RP; ++h) w[h] = *(const uint4*)(x[h] + 16 * sb);
                float S[RP], L7[RP];
#pragma unroll
                for (int h = 0; h < RP; ++h) {
                    const uint32_t ws[4] = {w[h].x, w[h].y, w[h].z, w[h].w};
                    float d[16];
#pragma unroll
                    for (int i = 0; i < 16; ++i)
                        d[i] = (float)((ws[i >> 2] >> (8 * (i & 3))) & 0xffu) - qv[16 * sb + i];
                    float l[8];
#pragma unroll
                    for (int i = 0; i < 8; ++i) l[i] = fmaf(d[i], d[i], d[8 + i] * d[8 + i]);
                    S[h] = ((l[0] + l[1]) + (l[2] + l[3])) + ((l[4] + l[5]) + l[6]);
                    L7[h] = l[7];
                }
                const uint32_t nb = nfull - k0 < 8 ? nfull - k0 : 8;
                for (uint32_t jj = 0; jj < nb; ++jj) {
#pragma unroll
                    for (int h = 0; h < RP; ++h) {
                        const float sj = __shfl(S[h], src + (int)jj, 64), lj = __shfl(L7[h], src + (int)jj, 64);
                        t[h] = sj + (lj + t[h]);
                    }
                }
            }
#pragma unroll
            for (int h = 0; h < RP; ++h) {
                const uint32_t r = rr[h];
                if (j == 0 && r < n2) {
                    uint64_t key = ~0ull;
                    if (r < n) {
                        key = ((uint64_t)pk_f2ord(t[h]) << 32) | r;
                        if (exact || bud) dist[r] = t[h];
                    }
                    if (!exact) keys[r] = key;
                }
            }
        }
    } else {
        constexpr int SB = 8;  // 16-element steps per block
        const uint32_t g = threadIdx.x >> 3, j = threadIdx.x & 7, ng = blockDim.x >> 3;
        const int src = (int)(threadIdx.x & 63u & ~7u);
        const uint32_t dim = a.dim, nfull = dim >> 4;
        for (uint32_t r0 = 0; r0 < n2; r0 += RP * ng) {
            uint32_t rr[RP], mid[RP];
            const T* x[RP];
            float t[RP];
#pragma unroll
            for (int h = 0; h < RP; ++h) {
                rr[h] = r0 + h * ng + g;
                // in-row load for every lane (a lane-conditional load is waited for at once)
                const uint32_t fid = a.filt_id[(uint64_t)qi * ftk + (rr[h] < ftk ? rr[h] : ftk - 1u)];
                mid[h] = rr[h] < n ? fid : 0u;
                x[h] = raw + (uint64_t)mid[h] * dim;
                t[h] = 0.0f;
            }
            for (uint32_t s0 = 0; s0 < nfull; s0 += SB) {
                float v0[RP][SB], v1[RP][SB];
#pragma unroll
                for (int h = 0; h < RP; ++h)
#pragma unroll
                    for (int st = 0; st < SB; ++st)
                        if (s0 + st < nfull) {
                            v0[h][st] = (float)x[h][16 * (s0 + st) + j];
                            v1[h][st] = (float)x[h][16 * (s0 + st) + 8 + j];
                        }
#pragma unroll
                for (int st = 0; st < SB; ++st) {
                    if (s0 + st >= nfull) break;
                    const uint32_t m = 16 * (s0 + st);
                    const float q0 = qv[m + j], q1 = qv[m + 8 + j];
#pragma unroll
                    for (int h = 0; h < RP; ++h) {
                        const float d0 = v0[h][st] - q0;
                        const float d1 = v1[h][st] - q1;
                        const float l = fmaf(d0, d0, d1 * d1);
                        float lj[8];
#pragma unroll
                        for (int k = 0; k < 8; ++k) lj[k] = __shfl(l, src + k, 64);
                        t[h] = (((lj[0] + lj[1]) + (lj[2] + lj[3])) + ((lj[4] + lj[5]) + lj[6])) + (lj[7] + t[h]);
                    }
                }
            }
#pragma unroll
            for (int h = 0; h < RP; ++h) {
                if (16 * nfull < dim) t[h] = t[h] + pk_rerank_tail(x[h] + 16 * nfull, qv + 16 * nfull, dim - 16 * nfull);
                const uint32_t r = rr[h];
                if (j == 0 && r < n2) {
                    uint64_t key = ~0ull;
                    if (r < n) {
                        key = ((uint64_t)pk_f2ord(t[h]) << 32) | r;
                        if (exact || bud) dist[r] = t[h];
                    }
                    if (!exact) keys[r] = key;
                }
            }
        }
    }
    __syncthreads();
    if (tn > 0 && tn != PK_TIE_OVERFLOW) {
        for (uint32_t r = threadIdx.x; r < n + tn; r += blockDim.x) {
            float d;
            if (r < n) {
                if (!(a.filt_t[(uint64_t)qi * ftk + r] == tfin)) continue;
                d = pk_ord2f((uint32_t)(keys[r] >> 32));
            } else {
                uint32_t mid = a.tie_ids[(uint64_t)qi * PK_TIE_MAX + (r - n)];
                d = pq ? pk_pq_dist(a, plut, mid, tfin, qn) : pk_rerank_dist(raw + (uint64_t)mid * a.dim, qv, a.dim);
            }
            atomicMin(&tie_min, pk_f2ord(d));
        }
    }
    __syncthreads();
    if (exact) {
        if (threadIdx.x < 64) {
            for (uint32_t r = threadIdx.x; r <= topk; r += 64) RH[r] = make_uint2(0x7f7f7f7fu, 0xffffffffu);
            uint32_t dc = topk;
            pk_heap_offer_wave<true>(RH, topk, &dc, n, [&](uint32_t r, float& v, uint32_t& tg) {
                v = dist[r];
                tg = a.mem2local[a.filt_id[(uint64_t)qi * ftk + r]];
            });
            pk_heap_reorder_wave(RH, topk);
            if (threadIdx.x == 0) a.out_n[qi] = topk - dc;
        }
        __syncthreads();
        for (uint32_t r = threadIdx.x; r < ostr; r += blockDim.x) {
            const bool in = r < topk;
            a.out_d[(uint64_t)qi * ostr + r] = in ? pk_hv(RH[r + 1]) : pk_sentinel();
            a.out_id[(uint64_t)qi * ostr + r] = in ? RH[r + 1].y : 0xffffffffu;
        }
        return;
    }
    // only the first min(n, topk + 1) entries of the sorted order are read
    const uint32_t kp = n < topk + 1 ? n : topk + 1;
    auto sort_prefix = [&]() {
        if (n > 1 && !pk_sorted_prefix(keys, n, kp)) {
            uint32_t m2 = 2;
            while (m2 < n) m2 <<= 1;
            for (uint32_t k = 2; k <= m2; k <<= 1)
                for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                    for (uint32_t i = threadIdx.x; i < (m2 >> 1); i += blockDim.x) {
                        uint32_t lo = 2 * i - (i & (j - 1));
                        uint32_t hi = lo + j;
                        bool up = (lo & k) == 0;
                        uint64_t x = keys[lo], y = keys[hi];
                        if ((x > y) == up) {
                            keys[lo] = y;
                            keys[hi] = x;
                        }
                    }
                    __syncthreads();
                }
        }
    };
    sort_prefix();
    // budget scan: equal distances rank by local id.  The keys carry the
    // filter rank, so only when two of the first kp sorted entries share a
    // distance are the keys rebuilt with local ids (one id lookup per
    // survivor, in the rare case only) and the prefix sorted again.
    bool local_keys = false;
    if (bud) {
        __shared__ uint32_t btie;
        if (threadIdx.x == 0) btie = 0u;
        __syncthreads();
        for (uint32_t r = threadIdx.x; r + 1 < kp; r += blockDim.x)
            if ((uint32_t)(keys[r] >> 32) == (uint32_t)(keys[r + 1] >> 32)) btie = 1u;
        __syncthreads();
        if (btie) {
            for (uint32_t r = threadIdx.x; r < n2; r += blockDim.x)
                keys[r] = r < n ? ((uint64_t)pk_f2ord(dist[r]) << 32) | a.mem2local[a.filt_id[(uint64_t)qi * ftk + r]]
                                : ~0ull;
            __syncthreads();
            sort_prefix();
            local_keys = true;
        }
    }
    const uint32_t kout = n < topk ? n : topk;
    bool tie = false;
    for (uint32_t r = threadIdx.x; r < ostr; r += blockDim.x) {
        float d = pk_sentinel();
        uint32_t id = 0xffffffffu;
        if (r < kout) {
            uint64_t key = keys[r];
            d = pk_ord2f((uint32_t)(key >> 32));
            uint32_t rr = (uint32_t)(key & 0xffffffffu);
            id = local_keys ? rr : a.mem2local[a.filt_id[(uint64_t)qi * ftk + rr]];
            if (!bud && r + 1 < n && pk_ord2f((uint32_t)(keys[r + 1] >> 32)) == d) tie = true;
        }
        a.out_d[(uint64_t)qi * ostr + r] = d;
        a.out_id[(uint64_t)qi * ostr + r] = id;
    }
    if (tie && a.flags) a.flags[qi] = 4u | a.flags[qi];
    // a tie-group member at or inside the k-th distance: which members the
    // reference kept decides the output
    if (threadIdx.x == 0 && tn > 0 && a.flags && topk > 0) {
        const uint32_t kth = kout == topk && kout > 0 ? (uint32_t)(keys[kout - 1] >> 32) : 0xffffffffu;
        if (tie_min <= kth) a.flags[qi] = 2u | a.flags[qi];
    }
    if (threadIdx.x == 0) a.out_n[qi] = kout;
}

extern "C" hipError_t pk_launch_rerank(const PkRerankArgs* args, uint32_t nq, hipStream_t s) {
    if (nq == 0) return hipSuccess;
    if (args->out_stride < args->topk) return hipErrorInvalidValue;
    uint32_t n2 = 2;
    while (n2 < args->ftk) n2 <<= 1;
    size_t smem = sizeof(uint64_t) * n2 + sizeof(float) * ((args->dim + 3) & ~3u) +
                  sizeof(float) * ((args->ftk + 3) & ~3u) + 16 + 8 * ((args->topk + 2) & ~1u);
    if (args->pq_codes) {
        if (!args->qn || !args->pq_off || !args->pq_codebook || args->pnsq == 0) return hipErrorInvalidValue;
        smem += sizeof(float) * (size_t)args->pnsq * PK_KS;
    }
    static PkLdsCache lds_cache;
    size_t* slot = pk_lds_slot(lds_cache);
    if (!slot) return hipErrorInvalidDevice;
    if (!*slot) {
        const void* fns[] = {(const void*)pk_rerank<float, 1>, (const void*)pk_rerank<float, 2>,
                             (const void*)pk_rerank<float, 4>, (const void*)pk_rerank<uint8_t, 1>,
                             (const void*)pk_rerank<uint8_t, 2>, (const void*)pk_rerank<uint8_t, 4>};
        size_t lo = 160 * 1024;
        for (const void* f : fns) {
            size_t al = 0;
            const hipError_t e = pk_allow_lds(f, 160 * 1024, &al);
            if (e != hipSuccess) return e;
            lo = al < lo ? al : lo;
        }
        *slot = lo;
    }
    if (smem > *slot) return hipErrorInvalidValue;
    const uint32_t rows = args->rows == 1 || args->rows == 4 ? args->rows : 2u;
    const bool u8 = args->raw8 && !args->pq_codes;
#define PK_RR_LAUNCH(T, R) hipLaunchKernelGGL((pk_rerank<T, R>), dim3(nq), dim3(RERANK_THREADS), smem, s, *args)
    if (u8) {
        if (rows == 1) PK_RR_LAUNCH(uint8_t, 1);
        else if (rows == 4) PK_RR_LAUNCH(uint8_t, 4);
        else PK_RR_LAUNCH(uint8_t, 2);
    } else {
        if (rows == 1) PK_RR_LAUNCH(float, 1);
        else if (rows == 4) PK_RR_LAUNCH(float, 4);
        else PK_RR_LAUNCH(float, 2);
    }
#undef PK_RR_LAUNCH
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Exact center order for flagged queries: the coarse top-scc MaxHeap
// (hierarchical_cluster_index.cpp:709-726) and the fine full heap sort
// (puck_index.cpp:497-504), replayed by one wave on LDS copies of the rows.
__device__ void pk_exact_center_order_one(uint32_t qi, const float* __restrict__ cdist, uint32_t C,
                                                             const float* __restrict__ fdist, uint32_t F,
                                                             uint32_t scs, uint32_t scap,
                                                             const uint32_t* __restrict__ q_scc,
                                                             const uint8_t* __restrict__ flags, int force,
                                                             float* __restrict__ csel_d, uint32_t* __restrict__ csel_id,
                                                             float* __restrict__ fsort_d, uint32_t* __restrict__ fsort_id) {
    extern __shared__ __align__(16) unsigned char smem[];
    // coarse heap order for a coarse-distance tie (flag 1), fine heap order for
    // a decision on a fine-distance tie (flag 8); both when forced
    const bool do_coarse = force || (flags[qi] & 1u);
    const bool do_fine = force || (flags[qi] & 8u);
    // heap capacity = this query's search_coarse_count (its tie order depends
    // on it), at most scap; rows keep the stride scs
    const uint32_t scc = q_scc ? (q_scc[qi] < scap ? q_scc[qi] : scap) : scap;
    const uint32_t L = C > F ? C : F;
    uint2* H = (uint2*)smem;                           // L + 2 entries
    float* row = (float*)(H + ((L + 2) & ~1u));        // max(C, F)
    // coarse top-scc (hierarchical_cluster_index.cpp:709-726)
    if (do_coarse) {
        for (uint32_t i = threadIdx.x; i < C; i += blockDim.x) row[i] = cdist[(uint64_t)qi * C + i];
        if (scc > 0) pk_heap_init(H, scc);
        __syncthreads();
        if (threadIdx.x < 64 && scc > 0) {
            uint32_t dc = scc;
            pk_heap_offer_wave<true>(H, scc, &dc, C, [&](uint32_t c, float& v, uint32_t& tg) {
                v = row[c];
                tg = c;
            });
            pk_heap_reorder_wave(H, scc);
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < scc; i += blockDim.x) {
            csel_d[(uint64_t)qi * scs + i] = pk_hv(H[i + 1]);
            csel_id[(uint64_t)qi * scs + i] = H[i + 1].y;
        }
        __syncthreads();
    }
    if (!do_fine) return;
    // fine full heap sort (puck_index.cpp:497-504)
    int below = 1;
    for (uint32_t i = threadIdx.x; i < F; i += blockDim.x) {
        row[i] = fdist[(uint64_t)qi * F + i];
        if (row[i] >= __uint_as_float(0x7f7f7f7fu)) below = 0;
    }
    pk_heap_init(H, F);
    // Update k places row[k] at position F - k and sifts it down (Floyd's
    // construction over positions F .. 1), since the top stays the empty
    // sentinel until the last update and no value reaches it.  Positions of
    // one depth sift down disjoint subtrees, so each depth runs in parallel,
    // deepest first; leaves need no sift.
    if (__syncthreads_and(below)) {
        for (uint32_t p = 1 + threadIdx.x; p <= F; p += blockDim.x) H[p] = make_uint2(__float_as_uint(row[F - p]), F - p);
        __syncthreads();
        int D = 0;
        while ((2u << D) <= F / 2u) ++D;  // depth of the last internal node F / 2
        for (; D >= 0; --D) {
            const uint32_t p_end = min((2u << D) - 1u, F / 2u);
            for (uint32_t p = (1u << D) + threadIdx.x; p <= p_end; p += blockDim.x) {
                const uint2 e = H[p];
                const float v = pk_hv(e);
                uint32_t f = p;
                for (;;) {
                    const uint32_t l = 2u * f;
                    if (l > F) break;
                    uint2 c = H[l];
                    uint32_t ci = l;
                    if (l + 1u <= F && !(pk_hv(c) > pk_hv(H[l + 1u]))) {
                        c = H[l + 1u];
                        ci = l + 1u;
                    }
                    if (pk_hv(c) < v) break;
                    H[f] = c;
                    f = ci;
                }
                H[f] = e;
            }
            __syncthreads();
        }
        if (threadIdx.x < 64) pk_heap_reorder_wave(H, F);
    } else if (threadIdx.x < 64) {
        uint32_t dc = F;
        pk_heap_offer_wave<false>(H, F, &dc, F, [&](uint32_t k, float& v, uint32_t& tg) {
            v = row[k];
            tg = k;
        });
        pk_heap_reorder_wave(H, F);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < F; i += blockDim.x) {
        fsort_d[(uint64_t)qi * F + i] = pk_hv(H[i + 1]);
        fsort_id[(uint64_t)qi * F + i] = H[i + 1].y;
    }
}

__global__ __launch_bounds__(256) void pk_exact_center_order(const float* __restrict__ cdist, uint32_t C,
                                                             const float* __restrict__ fdist, uint32_t F,
                                                             uint32_t scs, uint32_t scap,
                                                             const uint32_t* __restrict__ q_scc,
                                                             const uint32_t* __restrict__ qlist,
                                                             const uint8_t* __restrict__ flags, int force,
                                                             float* __restrict__ csel_d, uint32_t* __restrict__ csel_id,
                                                             float* __restrict__ fsort_d, uint32_t* __restrict__ fsort_id) {
    pk_exact_center_order_one(qlist[blockIdx.x], cdist, C, fdist, F, scs, scap, q_scc, flags, force, csel_d, csel_id,
                              fsort_d, fsort_id);
}

// The same over a list built on the device (list[0 .. *count)), grid-stride:
// the batch's queries with a coarse or fine tie, ordered while the batch's
// rerank runs on the other stream.
__global__ __launch_bounds__(256) void pk_exact_center_order_dev(const float* __restrict__ cdist, uint32_t C,
                                                                 const float* __restrict__ fdist, uint32_t F,
                                                                 uint32_t scs, uint32_t scap,
                                                                 const uint32_t* __restrict__ q_scc,
                                                                 const uint32_t* __restrict__ list,
                                                                 const uint32_t* __restrict__ count,
                                                                 const uint8_t* __restrict__ flags,
                                                                 float* __restrict__ csel_d,
                                                                 uint32_t* __restrict__ csel_id,
                                                                 float* __restrict__ fsort_d,
                                                                 uint32_t* __restrict__ fsort_id) {
    const uint32_t n = *count;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        pk_exact_center_order_one(list[i], cdist, C, fdist, F, scs, scap, q_scc, flags, 0, csel_d, csel_id, fsort_d,
                                  fsort_id);
        __syncthreads();  // LDS reused by the next query
    }
}

// list[0 .. *count) = the i < n with flags[i] & mask, ascending (one workgroup).
__global__ __launch_bounds__(1024) void pk_list_flagged(const uint8_t* __restrict__ flags, uint32_t n, uint32_t mask,
                                                       uint32_t* __restrict__ list, uint32_t* __restrict__ count) {
    __shared__ uint32_t tmp[1024 / 64 + 1];
    uint32_t base = 0;
    for (uint32_t i0 = 0; i0 < n; i0 += blockDim.x) {
        const uint32_t i = i0 + threadIdx.x;
        const uint32_t sel = (i < n && (flags[i] & mask)) ? 1u : 0u;
        uint32_t tot;
        const uint32_t pre = pk_block_exscan(sel, tmp, &tot);
        if (sel) list[base + pre] = i;
        base += tot;
    }
    if (threadIdx.x == 0) *count = base;
}

extern "C" hipError_t pk_launch_exact_center_order_dev(const float* cdist, uint32_t C, const float* fdist, uint32_t F,
                                                       uint32_t scs, uint32_t scap, const uint32_t* q_scc,
                                                       const uint8_t* flags, uint32_t mask, uint32_t n,
                                                       uint32_t* list, uint32_t* count, float* csel_d,
                                                       uint32_t* csel_id, float* fsort_d, uint32_t* fsort_id,
                                                       hipStream_t s) {
    if (scap > scs) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    const uint32_t L = C > F ? C : F;
    size_t smem = sizeof(uint2) * ((L + 2) & ~1u) + sizeof(float) * L;
    static PkLdsCache lds_cache;
    size_t* slot = pk_lds_slot(lds_cache);
    if (!slot) return hipErrorInvalidDevice;
    if (!*slot) {
        const hipError_t e = pk_allow_lds((const void*)pk_exact_center_order_dev, 160 * 1024, slot);
        if (e != hipSuccess) return e;
    }
    if (smem > *slot) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pk_list_flagged, dim3(1), dim3(1024), 0, s, flags, n, mask, list, count);
    hipLaunchKernelGGL(pk_exact_center_order_dev, dim3(n < 64 ? n : 64), dim3(256), smem, s, cdist, C, fdist, F, scs,
                       scap, q_scc, list, count, flags, csel_d, csel_id, fsort_d, fsort_id);
    return hipGetLastError();
}

extern "C" hipError_t pk_launch_exact_center_order(const float* cdist, uint32_t C, const float* fdist, uint32_t F,
                                                   uint32_t scs, uint32_t scap, const uint32_t* q_scc,
                                                   const uint32_t* qlist, uint32_t nlist, const uint8_t* flags,
                                                   int force, float* csel_d, uint32_t* csel_id, float* fsort_d,
                                                   uint32_t* fsort_id, hipStream_t s) {
    if (scap > scs) return hipErrorInvalidValue;
    if (nlist == 0) return hipSuccess;
    const uint32_t L = C > F ? C : F;
    size_t smem = sizeof(uint2) * ((L + 2) & ~1u) + sizeof(float) * L;
    static PkLdsCache lds_cache;
    size_t* slot = pk_lds_slot(lds_cache);
    if (!slot) return hipErrorInvalidDevice;
    if (!*slot) {
        const hipError_t e = pk_allow_lds((const void*)pk_exact_center_order, 160 * 1024, slot);
        if (e != hipSuccess) return e;
    }
    const size_t allowed = *slot;
    if (smem > allowed) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pk_exact_center_order, dim3(nlist), dim3(256), smem, s, cdist, C, fdist, F, scs, scap,
                       q_scc, qlist, flags, force, csel_d, csel_id, fsort_d, fsort_id);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fine-budget rank of given cells (labels of the learned probe).  For query
// qi and entry e, cell = cells[qi * k + e] (coarse * F + fine, or ~0): when
// the coarse center is among the query's first scc sorted coarse centers, the
// smallest fine budget under which the walk of that list scores the cell,
// 1 + #cells of the list with a smaller cell-distance key; otherwise ~0.
__global__ __launch_bounds__(256) void pk_budget_rank(const float* __restrict__ csd,
                                                      const uint32_t* __restrict__ csi, uint32_t cstride,
                                                      uint32_t scc, const float* __restrict__ fdist,
                                                      const float* __restrict__ stationary, uint32_t C,
                                                      uint32_t F, const uint32_t* __restrict__ cells, uint32_t k,
                                                      uint32_t* __restrict__ ranks) {
    __shared__ uint32_t cnt;
    const uint32_t qi = blockIdx.x;
    const float* fd = fdist + (uint64_t)qi * F;
    for (uint32_t e = 0; e < k; ++e) {
        const uint32_t cell = cells[(uint64_t)qi * k + e];
        uint32_t r = 0xffffffffu;
        if (cell != 0xffffffffu && cell / F < C) {
            const uint32_t c = cell / F, f = cell % F;
            uint32_t l = 0xffffffffu;
            for (uint32_t i = 0; i < scc; ++i)
                if (csi[(uint64_t)qi * cstride + i] == c) { l = i; break; }
            if (l != 0xffffffffu) {
                const float cd = csd[(uint64_t)qi * cstride + l];
                const float* st = stationary + (uint64_t)c * F;
                const uint32_t key = pk_f2ord((cd + st[f]) + fd[f]);
                if (threadIdx.x == 0) cnt = 0;
                __syncthreads();
                uint32_t m = 0;
                for (uint32_t j = threadIdx.x; j < F; j += blockDim.x)
                    if (pk_f2ord((cd + st[j]) + fd[j]) < key) ++m;
                atomicAdd(&cnt, m);
                __syncthreads();
                r = cnt + 1;
                __syncthreads();
            }
        }
        if (threadIdx.x == 0) ranks[(uint64_t)qi * k + e] = r;
    }
}

extern "C" hipError_t pk_launch_budget_rank(const float* csd, const uint32_t* csi, uint32_t cstride, uint32_t scc,
                                            const float* fdist, const float* stationary, uint32_t C, uint32_t F,
                                            const uint32_t* cells, uint32_t k, uint32_t nq, uint32_t* ranks,
                                            hipStream_t s) {
    if (nq == 0 || k == 0) return hipSuccess;
    hipLaunchKernelGGL(pk_budget_rank, dim3(nq), dim3(256), 0, s, csd, csi, cstride, scc, fdist, stationary, C, F,
                       cells, k, ranks);
    return hipGetLastError();
}
