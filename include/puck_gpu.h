/*
 * puck_gpu.h -- C ABI of libpuckgpu.so, the MI355X (gfx950) engine for Puck's
 * batched search path (PuckIndex / HierarchicalClusterIndex, whether_pq false
 * or true, with RealtimeInsertPuckIndex's inserts).
 *
 * Plain C types only; no C++ exceptions cross this boundary.  Every call
 * returns 0 on success or a negative code (the reference's convention: 0 = OK,
 * < 0 = error); puckgpu_last_error() gives the message of the last failure on
 * a context.  Pointers are host pointers unless `device_ptrs` is non-zero, in
 * which case every array argument of that call is a device (HBM) pointer on
 * the context's device and the call runs on the context's stream and returns
 * after it has finished.  A context may be used from one host thread at a time.
 *
 * Reference interfaces replaced (paths relative to the heaoxiang/puck checkout):
 *   puckgpu_index_load     HierarchicalClusterIndex::init / read_coodbooks /
 *                          convert_local_to_memory_idx / read_feature_index
 *                          (puck/hierarchical_cluster/hierarchical_cluster_index.cpp:220-255,
 *                          :257-324; puck/puck/puck_index.cpp:173-309)
 *   puckgpu_search         PySearcher::search + ParallelFor + PuckIndex::search
 *                          (pyapi_wrapper/py_api_wrapper.cpp:111-171;
 *                          puck/puck/puck_index.cpp:559-600; puck/index.h:30-81)
 *   puckgpu_search_batch   the same, one Request (own topk) per query, with
 *                          per-query search_coarse_count and fine budget
 *   puckgpu_assign_cells   HierarchicalClusterIndex::assign / nearest_cell_assign
 *                          (hierarchical_cluster_index.cpp:1042-1138)
 *   puckgpu_encode_filter  PuckIndex::puck_assign, filter quantization only
 *                          (puck/puck/puck_index.cpp:602-709)
 *   puckgpu_kmeans_assign  nearest_center, the k-means assignment step
 *                          (puck/hierarchical_cluster/kmeans.cpp:221-258)
 *   puckgpu_kmeans         Kmeans::kmeans / kmeans_core / kmeans_reassign_empty
 *                          (puck/hierarchical_cluster/kmeans.cpp:30-329)
 *   puckgpu_train_centers  HierarchicalClusterIndex::train
 *                          (puck/hierarchical_cluster/hierarchical_cluster_index.cpp:882-985)
 *   puckgpu_index_attach_pq  PuckIndex::read_coodbooks / read_feature_index for
 *                          whether_pq = true (the PQ codebook and pq_data.dat,
 *                          puck/puck/puck_index.cpp:173-227)
 *   puckgpu_encode_pq      PuckIndex::puck_assign with quantizations = {filter,
 *                          pq} (puck/puck/puck_index.cpp:602-709)
 *   puckgpu_index_insert   RealtimeInsertPuckIndex::insert and its search
 *                          overrides (puck/puck/realtime_insert_puck_index.cpp:
 *                          273-345, :398-541), the file appends left to the
 *                          caller
 * Entry points without a counterpart in the reference (each says so below):
 * the learned probe (puckgpu_probe_*, puckgpu_search_learned,
 * puckgpu_budget_ranks), the batched server (puckgpu_server_*), the
 * reduced-precision center distances (puckgpu_search_params.precision,
 * puckgpu_center_distances) and the measurement hooks.
 */
#ifndef PUCK_GPU_H
#define PUCK_GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct puckgpu_ctx puckgpu_ctx;
typedef struct puckgpu_index puckgpu_index;

/* Error codes. */
#define PUCKGPU_OK 0
#define PUCKGPU_EINVAL (-1)   /* bad argument (the reference's -1 of search / init) */
#define PUCKGPU_EHIP (-2)     /* HIP runtime failure */
#define PUCKGPU_ENOMEM (-3)   /* device allocation failed */

/* Index geometry and build flags (index.dat fields, hierarchical_cluster_index.cpp:522-567). */
typedef struct {
    uint32_t dim;                  /* feature_dim */
    uint32_t coarse_cluster_count; /* C */
    uint32_t fine_cluster_count;   /* F */
    uint32_t filter_nsq;           /* filter PQ subspaces; dim / filter_nsq in {1, 2, 4} */
    uint32_t total_point_count;    /* N */
    int32_t whether_norm;          /* queries are L2-normalised before search */
} puckgpu_index_params;

/* Search knobs (IndexConf, puck/index_conf.cpp:28-95, :152-198). */
typedef struct {
    uint32_t topk_max;            /* FLAGS_topk: a request with topk > topk_max fails with -1 */
    uint32_t search_coarse_count; /* clamped to C */
    uint32_t filter_topk;         /* >= topk, <= puckgpu_filter_topk_max() */
    double radius_rate;           /* >= 1.0 */
    uint32_t replay_all;          /* non-zero: replay every query with the reference heaps (tests) */
    uint32_t precision;           /* center distances: 0 = FP32 in the engine's exact order (the default);
                                     1 = BF16, 2 = per-row scaled FP8 e4m3, on the matrix cores with FP32
                                     accumulation (dim <= 256).  With 1 or 2 the coarse candidates near the
                                     search_coarse_count boundary are rescored in the exact FP32 order and
                                     the rerank stays exact FP32, so returned distances are exact; the
                                     fine distances, cell distances and filter distances carry the reduced
                                     precision and results are not bit-identical to the reference.  No
                                     tie replay is made in these forms.
                                     Or-ed with PUCKGPU_LUT_FP16: the filter lookup table in FP16 (see
                                     below), with either center precision. */
} puckgpu_search_params;

/* Precision flag (no counterpart in the reference, whose filter table is FP32,
 * quantization.cpp:260-270): the filter lookup table of each query is built in
 * FP16 -- every entry ||q_m - c||^2 is computed in FP32, scaled by 2^e and
 * rounded to FP16 (nearest even), where e = floor(log2(65504 / (|q| + cbn)^2))
 * clamped to [-30, 30] and cbn is the largest codeword norm, so that no entry
 * overflows; lookups are summed in FP32 in the reference's 4-lane order and the
 * sum is scaled by 2^-e.  The budget scan keeps the table in FP16 in LDS (half
 * the space); the walk and the per-query scan hold the same rounded values in
 * FP32, which gives the same sums bit for bit.  Filter distances carry the
 * rounding (relative error below 2^-11 per entry); the rerank stays exact. */
#define PUCKGPU_LUT_FP16 16u

/* Context: one device, one stream, reusable scratch.  device = HIP ordinal. */
int puckgpu_create(int device, puckgpu_ctx** out);
void puckgpu_destroy(puckgpu_ctx* ctx);
const char* puckgpu_last_error(const puckgpu_ctx* ctx);
uint32_t puckgpu_filter_topk_max(void);

/* Load an index from arrays in the reference file layouts, local (file) order:
 *   coarse [C x dim] (coarse.dat rows), fine [F x dim] (fine.dat rows, stored
 *   order), filter_codebook [nsq x 256 x dim/nsq] (filter_codebook.dat rows),
 *   cell_assign [N] (cell_assign.dat), filter_records [N x (4 + nsq)]
 *   (filter_data.dat: f32 offset then nsq code bytes), features [N x dim]
 *   (all_data.feat.bin rows without the int32 length prefix) or NULL for an
 *   index that ranks by a second-stage PQ (whether_pq = true, which does not
 *   load the raw vectors: puck_index.cpp:198-227; attach the PQ with
 *   puckgpu_index_attach_pq before searching).
 * The engine derives |c|^2/2, the C x F stationary table, the memory order
 * (cells ascending, then offset, then local id), empty-cell sentinels and the
 * per-coarse minimum offsets, as PuckIndex::init does. */
int puckgpu_index_load(puckgpu_ctx* ctx, const puckgpu_index_params* params, const float* coarse,
                       const float* fine, const float* filter_codebook, const uint32_t* cell_assign,
                       const uint8_t* filter_records, const float* features, int device_ptrs,
                       puckgpu_index** out);
void puckgpu_index_free(puckgpu_index* index);
/* Raw vectors stored as uint8 (the precision form of integer-valued data, no
 * counterpart in the reference, which keeps float rows in all_data.feat.bin,
 * puck_index.cpp:198-227): features [N x dim] uint8 in local-id order,
 * attached to an index loaded with features = NULL and whether_norm = 0.  The
 * exact rerank converts each value to float, so for data whose values are the
 * integers 0..255 its results are those of the float rows, at a quarter of
 * the bytes.  Inserts into such an index are refused. */
int puckgpu_index_attach_raw_u8(puckgpu_ctx* ctx, puckgpu_index* index, const uint8_t* features, int device_ptrs);
/* Second-stage PQ of a whether_pq = true index: pq_nsq subspaces of
 * dim / pq_nsq floats (1, 2 or 4; pq_nsq <= 128), pq_codebook [pq_nsq x 256 x
 * dim / pq_nsq] (pq_codebook.dat rows), pq_records [N x (4 + pq_nsq)] in local
 * order (pq_data.dat: f32 offset then pq_nsq code bytes).  The rerank then
 * ranks the filter survivors by rank_topN_points' PQ distance
 * (puck_index.cpp:402-427): (filter distance - |q|^2) + offset + the sum of
 * the query's PQ lookup table over the codes, instead of exact L2 over the
 * raw vectors. */
int puckgpu_index_attach_pq(puckgpu_ctx* ctx, puckgpu_index* index, uint32_t pq_nsq, const float* pq_codebook,
                            const uint8_t* pq_records, int device_ptrs);
/* Realtime insert of n points [n x dim], in order (RealtimeInsertPuckIndex::
 * insert, realtime_insert_puck_index.cpp:273-345): each is normalised
 * (whether_norm; a norm below 1e-6 refuses the point, which takes no id),
 * assigned to its nearest cell over the index's live stationary values
 * (nearest_cell_assign: empty cells hold the sentinel), and encoded with the
 * filter PQ (and the second-stage PQ of a whether_pq index) as the build
 * does.  A cell that had no point takes the stationary value
 * sdot(c, f) - sdot(f, f) (:326-338); min_dist_offset keeps its loaded value.
 * The points are searchable when the call returns: every cell holds its
 * loaded points, then its inserted points newest first, the order in which
 * the reference offers them to the filter heap (:398-453); the rerank reads
 * their stored features or PQ codes (:455-541).  Outputs (any may be NULL):
 * ids_out [n] the local id given (the count of points before it) or
 * 0xffffffff for a refused point; for the m accepted points, in order,
 * cells_out [m], filter_records_out [m x (4 + filter_nsq)] (filter_data.dat
 * rows), pq_records_out [m x (4 + pq_nsq)] (pq_data.dat rows) and
 * features_out [m x dim] (the stored, normalised rows) -- what the reference
 * appends to its index files.  The call rewrites the index's point arrays
 * (each cell's range moves), so a batch costs one pass over them: insert in
 * batches.  Servers created on the index capture their graphs again on their
 * next call. */
int puckgpu_index_insert(puckgpu_ctx* ctx, puckgpu_index* index, uint32_t n, const float* features,
                         uint32_t* ids_out, uint32_t* cells_out, uint8_t* filter_records_out,
                         uint8_t* pq_records_out, float* features_out, int device_ptrs);
/* Number of layout changes (insert calls that added points) of the index. */
uint64_t puckgpu_index_version(const puckgpu_index* index);
/* Copy derived arrays to host buffers (any may be NULL): coarse_norms [C],
 * stationary [C*F], cell_start [C*F+1], min_dist_offset [C], mem2local [N]. */
int puckgpu_index_export(puckgpu_ctx* ctx, const puckgpu_index* index, float* coarse_norms, float* stationary,
                         uint32_t* cell_start, float* min_dist_offset, uint32_t* mem2local);

/* Batched search of n queries [n x dim].  Outputs have row stride topk:
 * distance [n x topk] ascending squared L2 (on normalised vectors when
 * whether_norm), labels [n x topk] local ids, result_num [n] (may be NULL).
 * Unfilled slots hold distance 0x7f7f7f7f and label 0xffffffff.  A query whose
 * norm is below 1e-6 under whether_norm gets result_num 0.  Returns -1 without
 * touching the outputs when topk > topk_max or topk == 0. */
int puckgpu_search(puckgpu_ctx* ctx, const puckgpu_index* index, const puckgpu_search_params* params,
                   uint32_t n, const float* queries, uint32_t topk, float* distance, uint32_t* labels,
                   uint32_t* result_num, int device_ptrs);
/* puckgpu_search with device pointers that allocates nothing: no hipMalloc,
 * hipFree or other allocating call.  No counterpart in the reference; for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid it).  Every pass buffer lives in one array of
 * 64 MiB that each device holds for the life of the process (so do the host
 * staging arrays of the flags and replay lists), so calls that overlap on one
 * device race; the array keeps the last call's contents, and nothing in it
 * needs to start at zero.  queries [n x dim], distance / labels [n x topk] and
 * result_num [n] (may be NULL) are device pointers on the context's device;
 * results are those of puckgpu_search.  Limits: n <= 4096, precision 0, and
 * the call's buffers (per query 4 (dim + 2 C + 3 F + 2 search_coarse_count +
 * 2 filter_topk) bytes and about 100 more) within 64 MiB: a call above them
 * returns PUCKGPU_EINVAL and touches nothing.  The filter rows of such a call
 * are not kept for puckgpu_last_filter. */
int puckgpu_search_noalloc(puckgpu_ctx* ctx, const puckgpu_index* index, const puckgpu_search_params* params,
                           uint32_t n, const float* queries, uint32_t topk, float* distance, uint32_t* labels,
                           uint32_t* result_num);

/* Per-query parameters of puckgpu_search_batch.  The reference's Request
 * carries only topk (puck/index.h:60-69); search_coarse_count is a process
 * flag there (index_conf.cpp:64) and the fine budget has no counterpart, so
 * they live in this side struct and Request stays as it is.  Each pointer is
 * NULL (the batch-wide value applies) or an array of n entries, host or device
 * memory as the call's device_ptrs says. */
typedef struct {
    const uint32_t* topk;                /* 1..topk_max; another value gives that query result_num 0 */
    const uint32_t* search_coarse_count; /* coarse lists walked, clamped to params->search_coarse_count
                                            (the batch maximum, itself clamped to C) */
    const uint32_t* fine_budget;         /* b: 0 or >= F = the reference's cell walk.  0 < b < F = the
                                            budget scan: the coarse lists are the nearest by (distance,
                                            id); in each, every point of the cells whose cell distance
                                            (|q-c-f|^2 - |q|^2)/2 is at most the b-th smallest of the
                                            list (ties included) gets the reference's filter distance;
                                            the filter set is the filter_topk smallest (distance, memory
                                            id) and the result the topk smallest (exact distance, local
                                            id) of it */
} puckgpu_query_params;

/* Batched search with per-query parameters (PySearcher::search's batch, one
 * Request per query with its own topk; py_api_wrapper.cpp:154-171).  Output
 * rows have stride params->topk_max; entries past a query's topk hold the
 * unfilled-slot values.  filter_topk must be >= topk_max. */
int puckgpu_search_batch(puckgpu_ctx* ctx, const puckgpu_index* index, const puckgpu_search_params* params,
                         const puckgpu_query_params* query_params, uint32_t n, const float* queries,
                         float* distance, uint32_t* labels, uint32_t* result_num, int device_ptrs);
/* ---- Learned probe ------------------------------------------------------
 * No counterpart in the reference.  A small MLP predicts, per query, the log
 * of the fine budget (puckgpu_query_params.fine_budget) the cell walk needs
 * to reach its top-k; the walk with every cell (the reference's) is its
 * fallback.  Network: in = dim + coarse_features +
 * fine_features inputs (the normalised query, its smallest coarse distances
 * |c|^2/2 - <c,q>, and the smallest cell distances (|q-c-f|^2 - |q|^2)/2 of
 * the fine clusters of its nearest coarse list, both ascending), each
 * standardised as (x - mu) / sigma; two hidden GELU layers of `hidden` units;
 * one output.  Weights are used in BF16 with FP32 accumulation.
 * Packed FP32 parameters (puckgpu_probe_param_count floats):
 *   W1 [hidden x in], b1 [hidden], W2 [hidden x hidden], b2 [hidden],
 *   w3 [hidden], b3 [1], mu [in], 1/sigma [in]. */
typedef struct puckgpu_probe puckgpu_probe;
typedef struct {
    uint32_t coarse_features; /* smallest coarse distances used (<= C) */
    uint32_t fine_features;   /* smallest cell distances of the nearest list used (<= F) */
    uint32_t hidden;          /* units per hidden layer */
} puckgpu_probe_shape;
typedef struct {
    float lr, beta1, beta2, eps;
} puckgpu_adam_params;
/* Budget from the prediction y: ceil(exp(y + margin)) clamped to
 * [min_budget, F]; all F (the reference's walk) when y is not finite or
 * exp(y + margin) >= fallback_frac * F. */
typedef struct {
    float margin;
    uint32_t min_budget;
    float fallback_frac;
} puckgpu_budget_params;

int puckgpu_probe_create(puckgpu_ctx* ctx, uint32_t dim, const puckgpu_probe_shape* shape, puckgpu_probe** out);
void puckgpu_probe_free(puckgpu_probe* probe);
uint32_t puckgpu_probe_param_count(const puckgpu_probe* probe);
/* Load all parameters (resets the Adam state) / read them back. */
int puckgpu_probe_set_params(puckgpu_ctx* ctx, puckgpu_probe* probe, const float* params, int device_ptrs);
int puckgpu_probe_get_params(puckgpu_ctx* ctx, const puckgpu_probe* probe, float* params, int device_ptrs);
/* Raw (unstandardised) features of n queries [n x in], computed by the search
 * pipeline's own center GEMM and sorts; fine_order [n x F] (may be NULL)
 * receives each query's ascending fine-distance (-<f,q>) order (fine ids). */
int puckgpu_probe_features(puckgpu_ctx* ctx, const puckgpu_index* index, const puckgpu_probe* probe, uint32_t n,
                           const float* queries, float* features, uint32_t* fine_order, int device_ptrs);
/* One training step on a batch: forward, mean-squared-error loss against
 * target [n], backward.  With adam != NULL the parameters take one Adam step;
 * loss (host, may be NULL) receives the batch's mean squared error, grads
 * (may be NULL) the gradient of that loss for the trained block (the first
 * param_count - 2*in floats of the packed layout). */
int puckgpu_probe_train_step(puckgpu_ctx* ctx, puckgpu_probe* probe, uint32_t n, const float* features,
                             const float* target, const puckgpu_adam_params* adam, float* loss, float* grads,
                             int device_ptrs);
/* Network output y [n] for raw features [n x in]. */
int puckgpu_probe_predict(puckgpu_ctx* ctx, const puckgpu_probe* probe, uint32_t n, const float* features,
                          float* y, int device_ptrs);
/* puckgpu_search_batch with the fine budget of every query predicted by the
 * probe inside the same pipeline (features, network and budget run on the
 * search's stream between the center sorts and the walk).  query_params'
 * fine_budget is ignored; budget_out [n] (may be NULL) receives the budgets. */
int puckgpu_search_learned(puckgpu_ctx* ctx, const puckgpu_index* index, const puckgpu_probe* probe,
                           const puckgpu_search_params* params, const puckgpu_budget_params* budget,
                           const puckgpu_query_params* query_params, uint32_t n, const float* queries,
                           float* distance, uint32_t* labels, uint32_t* result_num, uint32_t* budget_out,
                           int device_ptrs);

/* ---- Batched serving ------------------------------------------------------
 * No counterpart in the reference (PySearcher::search runs one Request per
 * thread, py_api_wrapper.cpp:111-171).  A server owns, for every power-of-two
 * batch size up to max_batch, device buffers and a HIP graph of the whole
 * device pipeline (captured on first use); a batch is padded to its bucket and
 * served by one graph launch.  Queries whose result depends on the order of
 * exactly equal distances are replayed outside the graph.  With a probe the
 * fine budgets are predicted in the graph (puckgpu_search_learned).  The
 * server uses the context's stream; one call at a time per server. */
typedef struct puckgpu_server puckgpu_server;
int puckgpu_server_create(puckgpu_ctx* ctx, const puckgpu_index* index, const puckgpu_probe* probe,
                          const puckgpu_search_params* params, const puckgpu_budget_params* budget,
                          uint32_t max_batch, puckgpu_server** out);
void puckgpu_server_free(puckgpu_server* server);
uint32_t puckgpu_server_buckets(const puckgpu_server* server);
/* One batch of n <= max_batch queries; topk / search_coarse_count [n] may be
 * NULL (params->topk_max / params->search_coarse_count).  Outputs as
 * puckgpu_search_batch (row stride topk_max). */
int puckgpu_server_search(puckgpu_server* server, uint32_t n, const float* queries, const uint32_t* topk,
                          const uint32_t* search_coarse_count, float* distance, uint32_t* labels,
                          uint32_t* result_num, int device_ptrs);
/* Autotuner: for every bucket, times the two center GEMMs in each tile shape
 * on the first cap rows of sample [max_batch x dim] (host) over reps
 * repetitions and keeps the fastest shape, then times the bucket's whole pass
 * with the rerank's rows in flight per lane group set to 1, 2 and 4 and keeps
 * the fastest (every choice gives the same bits); graphs are re-captured.
 * shapes_out [buckets] and times_out [buckets x shapes] (ms per repetition of
 * the GEMMs) may be NULL; puckgpu_server_tuned reports a bucket's choices. */
int puckgpu_server_tune(puckgpu_server* server, const float* sample, uint32_t reps, int32_t* shapes_out,
                        float* times_out);
int puckgpu_server_tuned(const puckgpu_server* server, uint32_t bucket, int32_t* gemm_shape, uint32_t* rerank_rows);
/* Rows in flight per 8-lane group of the exact rerank in this context's
 * non-server searches: 1, 2 (default) or 4.  A tuning choice with no
 * counterpart in the reference; results do not depend on it. */
int puckgpu_set_rerank_rows(puckgpu_ctx* ctx, uint32_t rows);

/* Fine-budget ranks (labels of the learned probe): for query i and entry e of
 * cells [n x k] (cell id coarse * F + fine, or 0xffffffff), the smallest
 * fine_budget under which the walk of that coarse list scores the cell -- 1 +
 * the number of cells of the list with a smaller cell distance -- or
 * 0xffffffff when the entry is empty or its coarse center is not among the
 * query's search_coarse_count nearest.  ranks [n x k]. */
int puckgpu_budget_ranks(puckgpu_ctx* ctx, const puckgpu_index* index, uint32_t search_coarse_count, uint32_t n,
                         const float* queries, uint32_t k, const uint32_t* cells, uint32_t* ranks,
                         int device_ptrs);
/* Stage output of the last search on this context (host buffers):
 * filter-stage survivors per query (filter_topk entries per row, ascending
 * filter distance -- in no particular order for a query searched with a fine
 * budget -- memory ids) and their count.  For tests and diagnostics. */
int puckgpu_last_filter(puckgpu_ctx* ctx, uint32_t n, uint32_t filter_topk, float* filt_dist, uint32_t* filt_mem,
                        uint32_t* filt_n);
/* Per-stage device times (ms) of the last search, summed over its passes:
 * [0] query prep + center GEMM, [1] center sort, [2] walk + filter scan,
 * [3] rerank, [4] exact-tie replay.  counters13 (may be NULL) receives: the
 * walk's segments, uncertain cells, examined cells and scanned points summed
 * over queries; the number of queries replayed with the reference heaps
 * because their result depends on the order of exactly equal distances; and
 * how many queries raised each cause (coarse-distance tie, filter-boundary
 * tie, exact-distance tie in the top-k, decision on a fine-distance tie) and,
 * for the last cause, its detail (tie with an earlier cell, tie across a
 * chunk boundary, group of three or more, visiting orders disagree). */
int puckgpu_last_stage_ms(puckgpu_ctx* ctx, float* ms5, uint64_t* counters13);

/* Query x center distances of n queries as the search computes them in the
 * given precision (0 exact, 1 BF16, 2 FP8; see puckgpu_search_params):
 * coarse_out [n x C] = |c|^2/2 - <c,q>, fine_out [n x F] = -<f,q>, on the
 * normalised query under whether_norm.  Test and measurement hook. */
int puckgpu_center_distances(puckgpu_ctx* ctx, const puckgpu_index* index, uint32_t precision, uint32_t n,
                             const float* queries, float* coarse_out, float* fine_out, int device_ptrs);
/* Device time (ms) of the budget-scan launch of the last search (part of
 * stage [2]); -1 when the search launched no budget scan. */
int puckgpu_last_scan_ms(puckgpu_ctx* ctx, float* scan_ms);

/* Filter lookup table for n queries (test hook): out [n x nsq x 256]. */
int puckgpu_lut(puckgpu_ctx* ctx, uint32_t n, uint32_t dim, const float* queries, const float* filter_codebook,
                uint32_t nsq, float* out, int device_ptrs);

/* Build-time nearest cell of n points [n x dim]: cell_out [n] = coarse*F + fine,
 * dist_out [n] (may be NULL) = the winning score.  The answer is the
 * reference's (the first strict minimum over the cells it visits in (coarse
 * term, coarse id, fine) order); the default mode finds it without the
 * sequential walk: the nearest coarse list is scanned as the reference scans
 * it, cells are skipped by block lower bounds, and the minimum of the rest is
 * accepted when a triangle-check certificate shows the reference visits it;
 * points without the certificate are walked sequentially. */
int puckgpu_assign_cells(puckgpu_ctx* ctx, const float* coarse, const float* fine, uint32_t C, uint32_t F,
                         uint32_t dim, uint32_t n, const float* x, uint32_t* cell_out, float* dist_out,
                         int device_ptrs);
/* Assignment strategy of puckgpu_assign_cells (no counterpart in the
 * reference, which always walks): 0 = bounded with certificate (default),
 * 1 = the sequential walk for every point, 2 = bounded, then every point
 * walked again (test of the fallback path).  Counters: points assigned and
 * points walked sequentially since the context was created. */
int puckgpu_set_assign_mode(puckgpu_ctx* ctx, int mode);
/* Exact brute-force top-k of nq queries [nq x dim] float32 in n points
 * [n x dim] (float32, or uint8 integer features with x_u8 = 1): ids_out
 * [nq x k] and dist_out [nq x k] (float64 squared L2), ordered by (distance,
 * id).  Candidates come from the engine's exact FP32 GEMM (|x|^2/2 - <q,x>)
 * and row top-k, k + extra per 65,536-point chunk, merged to the 2 (k + extra)
 * smallest, then ordered in float64; 2 (k + extra) <= 1024.  The ground
 * truth of recall@k as ann-benchmarks defines it (metrics.py:15-40) and the
 * labels of the learned probe; no counterpart in the reference's search. */
int puckgpu_exact_search(puckgpu_ctx* ctx, const void* x, int x_u8, uint64_t n, uint32_t dim, const float* queries,
                         uint32_t nq, uint32_t k, uint32_t extra, uint32_t* ids_out, double* dist_out,
                         int device_ptrs);
/* Progress lines of long calls (puckgpu_train_centers: one per k-means run),
 * passed to fn(message, user) on the calling thread; fn = NULL stops them.
 * No counterpart in the reference, which logs through glog. */
typedef void (*puckgpu_progress_fn)(const char* message, void* user);
int puckgpu_set_progress(puckgpu_ctx* ctx, puckgpu_progress_fn fn, void* user);
int puckgpu_assign_stats(puckgpu_ctx* ctx, uint64_t* points, uint64_t* sequential);
/* Filter-PQ records of n points: records_out [n x (4 + nsq)] in the
 * filter_data.dat layout. */
int puckgpu_encode_filter(puckgpu_ctx* ctx, const float* coarse, const float* fine, uint32_t C, uint32_t F,
                          uint32_t dim, const float* filter_codebook, uint32_t nsq, uint32_t n, const float* x,
                          const uint32_t* cell, uint8_t* records_out, int device_ptrs);
/* Second-stage PQ records of n points: records_out [n x (4 + pq_nsq)] in the
 * pq_data.dat layout.  filter_records [n x (4 + filter_nsq)] are the points'
 * filter records (puckgpu_encode_filter): their codewords are subtracted
 * from the residual x - c - f before the PQ codes are chosen, and the offset
 * is 2 <PQ decode, c + f + filter decode>. */
int puckgpu_encode_pq(puckgpu_ctx* ctx, const float* coarse, const float* fine, uint32_t C, uint32_t F, uint32_t dim,
                      const float* filter_codebook, uint32_t filter_nsq, const uint8_t* filter_records,
                      const float* pq_codebook, uint32_t pq_nsq, uint32_t n, const float* x, const uint32_t* cell,
                      uint8_t* records_out, int device_ptrs);
/* k-means assignment: assign_out [n] nearest of k centroids [k x dim]
 * (nearest_center's formula and strict '<'), dist_out [n] may be NULL. */
int puckgpu_kmeans_assign(puckgpu_ctx* ctx, uint32_t dim, const float* centroids, uint32_t k, uint32_t n,
                          const float* x, int32_t* assign_out, float* dist_out, int device_ptrs);
/* Kmeans::kmeans (kmeans.cpp:149-219, redo 1): seeding, then niter Lloyd
 * iterations of kmeans_core (:260-329) on n points x [n x dim].  The
 * reference's generator is a std::mt19937 seeded with time(0); here it is
 * seeded with `seed` and drawn in the reference's sequence (k-means++ over
 * the first min(n, 8k) points when n > 8k and n > 8192, init_pp = 1, the
 * reference's default; otherwise k distinct uniform picks; each iteration
 * then draws the empty-cluster split's noise scale, and one roulette pick
 * and dim Gaussian steps per empty cluster).  Assignments are
 * nearest_center's.  Centroid update: mean_mode = 0 is the reference's
 * arithmetic, which clears only the first 4 bytes of the centroids before
 * adding the points in point order (kmeans.cpp:275), so a centroid becomes
 * (previous + sum) / count except element [0][0]; mean_mode = 1 is the plain
 * mean, the same float additions from zero.  Stops early when the error sum
 * changes by less than 1e-6 with no split.  centroids_out [k x dim],
 * assign_out [n] (may be NULL: the last assignment, made before the last
 * update), err_out: error sum / n, iters_out / split_out (may be NULL). */
int puckgpu_kmeans(puckgpu_ctx* ctx, uint32_t dim, uint32_t n, uint32_t k, const float* x, int niter, int init_pp,
                   uint32_t seed, int mean_mode, float* centroids_out, int32_t* assign_out, float* err_out,
                   int* iters_out, int* split_out, int device_ptrs);
/* HierarchicalClusterIndex::train (hierarchical_cluster_index.cpp:882-985):
 * `iterations` rounds (kmeans_iterations_count, 10 by default) of coarse
 * k-means on the points less their best fine centers so far, then fine
 * k-means on the points less their new coarse centers; a round whose fine
 * error is at least 1e-4 below the best keeps both, otherwise the Lloyd
 * limit (niter, or the reference's 30 when niter is 0) grows by 10 and the
 * rounds end once it exceeds 80.  One generator (seed) serves every k-means
 * run.  Fine centers come out in
 * ascending -sqrt(|f|^2/2) order (a MaxHeap sort, :958-980), i.e. by
 * descending norm.  coarse_out [C x dim], fine_out [F x dim], err_trace
 * [2 x iterations] (coarse and fine error of each round run, may be NULL),
 * rounds_out / niter_out (may be NULL). */
int puckgpu_train_centers(puckgpu_ctx* ctx, uint32_t dim, uint32_t n, const float* x, uint32_t C, uint32_t F,
                          int iterations, int niter, int init_pp, uint32_t seed, int mean_mode, float* coarse_out,
                          float* fine_out, float* err_trace, int* rounds_out, int* niter_out, int device_ptrs);

#ifdef __cplusplus
}
#endif
#endif
