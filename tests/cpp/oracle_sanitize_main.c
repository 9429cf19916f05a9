/* Drives the CPU oracle (oracle/puck_oracle.c) through index creation, the
 * walk and the budget scan (single queries and a per-query batch), the heap,
 * cell assignment, filter-PQ encoding and nearest_center on a small random
 * index, for a build with AddressSanitizer / UndefinedBehaviorSanitizer
 * (tests/test_oracle.py::test_oracle_clean_under_sanitizers).  Test
 * infrastructure only. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
typedef struct { uint64_t a, b, c, d; } po_stats;
void* po_index_create(uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, int, const float*, const float*, const float*,
                      const uint32_t*, const uint8_t*, const float*);
void po_index_destroy(void*);
int po_search_one(const void*, const float*, uint32_t, uint32_t, uint32_t, uint32_t, double, uint32_t, float*,
                  uint32_t*, uint32_t*, float*, uint32_t*, uint32_t*, po_stats*, int*, int*);
int po_search_batch_pq(const void*, uint32_t, const float*, const uint32_t*, uint32_t, const uint32_t*, uint32_t,
                       const uint32_t*, uint32_t, double, float*, uint32_t*, uint32_t*, int);
int po_assign_cells(const float*, const float*, uint32_t, uint32_t, uint32_t, const float*, uint32_t, uint32_t*,
                    float*, int);
int po_encode_filter(const float*, const float*, uint32_t, uint32_t, const float*, uint32_t, const float*,
                     const uint32_t*, uint32_t, uint8_t*, int);
int po_nearest_center(uint32_t, const float*, uint32_t, const float*, uint32_t, int32_t*, float*, int);
uint32_t po_heap_run(uint32_t, const float*, const uint32_t*, uint32_t, int, float*, uint32_t*);
static float frand(void) { return (float)rand() / RAND_MAX * 20.0f - 10.0f; }
int main(void) {
    srand(7);
    const uint32_t dim = 16, C = 5, F = 6, nsq = 8, N = 700, lsq = dim / nsq;
    float* coarse = malloc(sizeof(float) * C * dim); float* fine = malloc(sizeof(float) * F * dim);
    float* cb = malloc(sizeof(float) * nsq * 256 * lsq); float* x = malloc(sizeof(float) * N * dim);
    for (uint32_t i = 0; i < C * dim; ++i) coarse[i] = frand() * 3;
    for (uint32_t i = 0; i < F * dim; ++i) fine[i] = frand();
    for (uint32_t i = 0; i < nsq * 256 * lsq; ++i) cb[i] = frand() * 0.3f;
    for (uint32_t i = 0; i < N * dim; ++i) x[i] = frand() * 3;
    uint32_t* cell = malloc(sizeof(uint32_t) * N);
    po_assign_cells(coarse, fine, C, F, dim, x, N, cell, NULL, 2);
    uint8_t* rec = malloc((size_t)N * (4 + nsq));
    po_encode_filter(coarse, fine, F, dim, cb, nsq, x, cell, N, rec, 2);
    int32_t* a = malloc(sizeof(int32_t) * 2000);
    po_nearest_center(lsq, cb, 256, x, N * dim / lsq > 2000 ? 2000 : N * dim / lsq, a, NULL, 2);
    void* ix = po_index_create(dim, C, F, nsq, N, 0, coarse, fine, cb, cell, rec, x);
    if (!ix) { printf("index create failed\n"); return 1; }
    float q[16]; float d[100]; uint32_t ids[100], n, fn; float fv[2048]; uint32_t ft[2048]; po_stats st;
    int co[8]; int fo[8];
    unsigned runs = 0;
    for (int t = 0; t < 40; ++t) {
        for (uint32_t j = 0; j < dim; ++j) q[j] = frand() * 3;
        uint32_t topk = 1 + t % 50, ftk = topk + (t * 37) % 900, scc = 1 + t % 7, nf = (uint32_t)(t % 8);
        if (ftk > 2048) ftk = 2048;
        po_search_one(ix, q, topk, 100, scc, ftk, 1.0 + 0.1 * (t % 3), nf, d, ids, &n, fv, ft, &fn, &st,
                      scc <= 8 ? co : NULL, fo);
        ++runs;
    }
    uint32_t nq = 30; float* Q = malloc(sizeof(float) * nq * dim);
    uint32_t tk[30], sc[30], nb[30];
    for (uint32_t i = 0; i < nq; ++i) { tk[i] = i % 5 == 0 ? 0 : 1 + i % 20; sc[i] = 1 + i % 6; nb[i] = i % 7; }
    for (uint32_t i = 0; i < nq * dim; ++i) Q[i] = frand() * 3;
    float* D = malloc(sizeof(float) * nq * 20); uint32_t* I = malloc(sizeof(uint32_t) * nq * 20); uint32_t R[30];
    po_search_batch_pq(ix, nq, Q, tk, 20, sc, 6, nb, 300, 1.0, D, I, R, 2);
    float hv[64]; uint32_t hid[64]; float vals[200]; uint32_t tags[200];
    for (int i = 0; i < 200; ++i) { vals[i] = (float)(rand() % 50); tags[i] = i; }
    po_heap_run(64, vals, tags, 200, 1, hv, hid);
    po_index_destroy(ix);
    free(coarse); free(fine); free(cb); free(x); free(cell); free(rec); free(a); free(Q); free(D); free(I);
    printf("oracle driver done: %u single searches, batch of %u\n", runs, nq);
    return 0;
}
