// launch_args.h -- argument blocks of the walk / budget-scan and rerank
// launches (search.hip), filled by the pipeline in capi.hip.
#pragma once
#include <cstdint>

struct PkWalkArgs {
    const float* q;          // nq x dim (normalised when whether_norm)
    const float* qn;         // nq      |q|^2
    const int32_t* qstatus;  // nq      0 = searchable
    const float* csel_d;     // nq x scc
    const uint32_t* csel_id; // nq x scc
    const float* fsort_d;    // nq x F
    const uint32_t* fsort_id;
    const float* codebook;   // nsq x 256 x lsq
    const float* stationary; // C x F (empty cell: sqrt(FLT_MAX))
    const uint32_t* cell_start; // C*F + 1
    const float* mdo;        // C
    const float* rec_off;    // N
    const uint8_t* rec_codes;// N x nsq
    uint32_t dim, C, F, nsq, lsq, scc, ftk;
    double rr;
    float* out_t;            // nq x ftk
    uint32_t* out_id;        // nq x ftk
    uint32_t* out_n;         // nq
    uint32_t* stats;         // nq x 4 or null
    const uint32_t* qlist;   // block -> query index (null: identity)
    uint8_t* flags;          // nq: bit 3 set when a decision touched a fine-distance tie group
    int exact;               // replay the reference MaxHeap (insertion order and ties)
    uint32_t* tie_ids;       // nq x PK_TIE_MAX: memory ids dropped at a tie with the final top
    uint8_t* tie_n;          // nq: count, PK_TIE_OVERFLOW when the list overflowed
    const uint32_t* q_scc;   // nq or null: per-query coarse lists walked (clamped to scc)
    const uint32_t* q_nf;    // nq or null: per-query fine budget (0 or >= F: all cells)
    uint32_t cstride;        // csel row stride (>= scc)
    const float* fdist;      // nq x F unsorted fine distances -<f,q> (budget scan)
    uint32_t scan_cap;       // budget scan: candidate buffer entries (set by the launcher)
    uint32_t scan_round;     // budget scan: points scored per round (multiple of its block size)
    const uint32_t* wlist;   // with wcount: the queries that take the walk (pk_build_walk_list)
    const uint32_t* wcount;
    int fine_given;          // a query whose flags hold bit 3 on entry has its fine row in the
                             // reference's heap order: its fine-distance ties need no flag, and
                             // its flags are rewritten (not or-ed) with what this walk finds
    // split budget scan (scan.hip); sc_cells null: every budget query takes the per-query kernel
    uint4* sc_cells;         // nq x 512: selected cells {first memory id, points before, cell distance bits, 0}
    uint32_t* sc_meta;       // nq x 2: points (0xffffffff: left to the per-query kernel), cells
    uint32_t* sc_olist;      // nq: the queries left to the per-query kernel
    uint32_t* sc_ocount;     // 1
    uint32_t sc_capp;
    // the probe's sorted smallest cell distances of the nearest coarse list
    // (values and fine ids, nq x top_k; null without a probe): the cell cut
    // of that list is read from them when the budget is below top_k
    const float* top_d;
    const uint32_t* top_f;
    uint32_t top_k;
    // filter table precision: 0 FP32 (the reference's), 1 FP16 (scaled per
    // query by a power of two; cb_nmax = the largest codeword norm)
    int lut16;
    float cb_nmax;
};

struct PkRerankArgs {
    const float* q;          // nq x dim
    const int32_t* qstatus;
    const float* filt_t;     // nq x ftk
    const uint32_t* filt_id; // memory ids
    const uint32_t* filt_n;
    const float* raw;        // N x dim, memory order
    const uint8_t* raw8;     // N x dim uint8 rows (the uint8 storage form), or null
    const uint32_t* mem2local;
    uint32_t dim, ftk, topk;
    float* out_d;            // nq x topk
    uint32_t* out_id;
    uint32_t* out_n;
    const uint32_t* qlist;   // block -> query index (null: identity)
    uint8_t* flags;          // bit 2: exact distances tie inside the top-k (+1); bit 1: a
                             // filter-boundary tie group reaches the top-k
    int exact;               // replay the reference result heap over the filter heap order
    const uint32_t* tie_ids; // from the walk (may be null)
    const uint8_t* tie_n;
    const uint32_t* q_topk;  // nq or null: per-query top-k (1..topk; other values give no result)
    uint32_t out_stride;     // row stride of out_d / out_id (>= topk)
    const uint32_t* q_nf;    // nq or null: fine budgets; a query with 0 < budget < F came from the
    uint32_t F;              // budget scan and ranks by (distance, local id), without replay
    // second-stage PQ (whether_pq = true; pq_codes null: exact L2 over raw)
    const float* qn;         // nq      |q|^2
    const float* pq_off;     // N       PQ offsets, memory order
    const uint8_t* pq_codes; // N x pnsq
    const float* pq_codebook;// pnsq x 256 x plsq
    uint32_t pnsq, plsq;
    uint32_t rows;           // rows per 8-lane group in flight: 1, 2 or 4 (another value: 2)
};
