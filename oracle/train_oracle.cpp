/*
 * train_oracle.cpp -- CPU restatement of Puck's k-means (puck/hierarchical_cluster/
 * kmeans.cpp) and of HierarchicalClusterIndex::train's coarse / fine loop
 * (puck/hierarchical_cluster/hierarchical_cluster_index.cpp:882-985).
 *
 * TEST INFRASTRUCTURE ONLY.  Compiled with puck_oracle.c into
 * oracle/liboracle_train.so; used only by tests/.  The product library never
 * links, loads or calls it.
 *
 * The reference draws its randomness from one std::mt19937 per Kmeans object
 * seeded with time(0) (kmeans.h:47-48) through libstdc++'s
 * uniform_int_distribution, uniform_real_distribution<double> and
 * normal_distribution<double>; here the engine is seeded with an argument and
 * the same library distributions are used, so a run is reproducible and
 * follows the reference's draw sequence.  Dot products follow po_sdot_short
 * (MKL's order for 1, 2 or 4 floats, the sequential FMA chain otherwise), as
 * the rest of the oracle.  cblas_snrm2 (the empty-cluster split's scale) is
 * taken as sqrtf(po_dot(c, c)): MKL's own order is unpinned.
 *
 * Parity choice: kmeans_core clears only the first 4 bytes of the centroids
 * before accumulating (kmeans.cpp:275, memset(centroids, 0,
 * sizeof(centroids[0] * k * d)) is sizeof(float)), so each new centroid is
 * (previous centroid + sum of its points) / count, except element [0][0].
 * mean_mode = 0 reproduces that; mean_mode = 1 clears every centroid (the
 * plain Lloyd mean).
 */
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <random>
#include <vector>

extern "C" {
float po_dot(const float* a, const float* b, uint32_t n);
float po_sdot_short(const float* a, const float* b, uint32_t n);
int po_nearest_center(uint32_t dim, const float* centroids, uint32_t k, const float* x, uint32_t n, int32_t* assign,
                      float* dis, int nthreads);
uint32_t po_heap_run(uint32_t cap, const float* vals, const uint32_t* tags, uint32_t n, int reorder, float* out_val,
                     uint32_t* out_tag);
}

namespace {

struct Kmeans {
    std::mt19937 rnd;
    int niter = 30;  // KmeansParams (kmeans.h:31-42): redo 1, niter 30
    bool pp;
    int mean_mode;
    Kmeans(uint32_t seed, bool kmeans_pp, int mean) : rnd(seed), pp(kmeans_pp), mean_mode(mean) {}

    // random_init_center (kmeans.cpp:30-46)
    void random_init(size_t total, size_t k, std::vector<size_t>& ids) {
        ids.clear();
        std::uniform_int_distribution<> dis(0, (int)total - 1);
        std::vector<bool> taken(total, false);
        while (ids.size() < k) {
            size_t s = dis(rnd);
            if (taken[s]) continue;
            ids.push_back(s);
            taken[s] = true;
        }
    }

    // roulette_selection (kmeans.cpp:48-67)
    int roulette(std::vector<float>& wheel) {
        float total = 0;
        for (float v : wheel) total += v;
        const float alpha = (float)(1.0 / total);  // cblas_sscal(n, 1.0 / total, ...)
        for (float& v : wheel) v = alpha * v;
        std::uniform_real_distribution<double> dis(0, 1.0);
        double rd = dis(rnd);
        for (int id = 0; id < (int)wheel.size(); ++id) {
            rd -= wheel[id];
            if (rd < 0) return id;
        }
        return (int)wheel.size() - 1;
    }

    // kmeanspp_init_center (kmeans.cpp:69-110) over the first `total` rows
    void pp_init(size_t total, size_t k, uint32_t dim, const float* x, std::vector<size_t>& ids) {
        std::vector<float> best(total, std::numeric_limits<float>::max()), tmp(total);
        ids.assign(k, 0);
        ids[0] = rnd() % total;
        std::vector<float> pn(total);
        for (size_t j = 0; j < total; ++j) pn[j] = po_sdot_short(x + j * dim, x + j * dim, dim);
        for (size_t i = 1; i < k; ++i) {
            const size_t sel = ids[i - 1];
            const float* last = x + sel * dim;
            for (size_t j = 0; j < total; ++j) {
                float t = (float)((double)(pn[j] + pn[sel]) - 2.0 * (double)po_sdot_short(x + j * dim, last, dim));
                if (t < best[j]) best[j] = t;
            }
            tmp = best;
            ids[i] = (size_t)roulette(tmp);
        }
    }

    // kmeans_reassign_empty (kmeans.cpp:112-147)
    int reassign_empty(uint32_t dim, size_t k, float* cent, const int* nassign) {
        std::vector<float> proba(k), eps(dim);
        std::normal_distribution<> dn(0, rnd() / ((double)RAND_MAX + 1.0));
        for (size_t c = 0; c < k; ++c)  // int products, wrapping as the reference's int arithmetic does
            proba[c] = nassign[c] < 2 ? 0.f : (float)(int32_t)(uint32_t)((int64_t)nassign[c] * nassign[c] - 1);
        int nre = 0;
        for (size_t c = 0; c < k; ++c) {
            if (nassign[c] != 0) continue;
            nre++;
            const int j = roulette(proba);
            std::memcpy(cent + c * dim, cent + (size_t)j * dim, sizeof(float) * dim);
            const float* cj = cent + (size_t)j * dim;
            const double s = sqrtf(po_dot(cj, cj, dim)) * 0.0000001;
            for (float& v : eps) v = (float)dn(rnd);
            const float a = (float)s;
            for (float& v : eps) v = a * v;
            for (uint32_t t = 0; t < dim; ++t) cent[(size_t)j * dim + t] = cent[(size_t)j * dim + t] + eps[t];
            for (uint32_t t = 0; t < dim; ++t) cent[c * dim + t] = cent[c * dim + t] - eps[t];
            proba[j] = 0;
        }
        return nre;
    }

    // kmeans_core (kmeans.cpp:260-329).  Returns 0 / -1; qerr, iterations run
    // and clusters split through the pointers.
    int core(uint32_t d, size_t n, size_t k, int nit, float* cent, const float* v, int* assign, int* nassign,
             float* dis, double* qerr_out, int* iters, int* split, double* trace = nullptr) {
        double qerr = std::numeric_limits<double>::max(), qerr_old;
        int tot = 0, it = 0;
        for (int iter = 1; iter <= nit; ++iter) {
            ++it;
            po_nearest_center(d, cent, (uint32_t)k, v, (uint32_t)n, assign, dis, 0);
            std::memset(nassign, 0, k * sizeof(int));
            if (mean_mode) std::memset(cent, 0, sizeof(float) * k * d);
            else cent[0] = 0.f;
            for (size_t i = 0; i < n; ++i) {
                if (assign[i] < 0 || (size_t)assign[i] >= k) return -1;
                nassign[assign[i]]++;
                float* c = cent + (size_t)assign[i] * d;
                for (uint32_t t = 0; t < d; ++t) c[t] = c[t] + v[i * d + t];
            }
            for (size_t i = 0; i < k; ++i) {
                const float a = (float)(1.0 / nassign[i]);
                for (uint32_t t = 0; t < d; ++t) cent[i * d + t] = a * cent[i * d + t];
            }
            const int nre = reassign_empty(d, k, cent, nassign);
            tot += nre;
            if (tot > (int)(n / 100) && tot > 1000) return -1;
            qerr_old = qerr;
            qerr = 0;
            for (size_t i = 0; i < n; ++i) qerr += dis[i];
            if (trace) trace[it - 1] = qerr;
            if (std::fabs(qerr_old - qerr) < 1e-6 && nre == 0) break;
        }
        *qerr_out = qerr;
        if (iters) *iters = it;
        if (split) *split = tot;
        return 0;
    }

    // Kmeans::kmeans (kmeans.cpp:149-219), redo = 1.  Returns qerr / n, or -1.
    float run(uint32_t d, size_t n, size_t k, const float* x, float* cent_out, int* assign_out, int* iters,
              int* split) {
        if (k >= n) return -1;
        std::vector<float> cent(k * d), dis(n);
        std::vector<int> assign(n), nassign(k);
        std::vector<size_t> sel;
        if (pp) {
            const uint32_t nsub = (n > k * 8 && n > 8 * 1024) ? (uint32_t)(k * 8) : (uint32_t)n;
            pp_init(nsub, k, d, x, sel);
        } else {
            random_init(n, k, sel);
        }
        for (size_t i = 0; i < k; ++i) std::memcpy(cent.data() + i * d, x + sel[i] * d, sizeof(float) * d);
        double qerr = 0;
        if (core(d, n, k, niter, cent.data(), x, assign.data(), nassign.data(), dis.data(), &qerr, iters, split))
            return -1;
        std::memcpy(cent_out, cent.data(), sizeof(float) * k * d);
        if (assign_out) std::memcpy(assign_out, assign.data(), sizeof(int) * n);
        return (float)(qerr / n);
    }
};

}  // namespace

extern "C" {

/* Kmeans::kmeans with a fresh generator seeded `seed`: centroids_out [k x dim],
 * assign_out [n] (may be NULL), iters_out / split_out (may be NULL): Lloyd
 * iterations run and empty clusters split.  Returns the mean squared error
 * (qerr / n), or -1. */
float pot_kmeans(uint32_t dim, uint32_t n, uint32_t k, const float* x, int niter, int init_pp, uint32_t seed,
                 int mean_mode, float* centroids_out, int32_t* assign_out, int* iters_out, int* split_out) {
    Kmeans km(seed, init_pp != 0, mean_mode);
    km.niter = niter;
    return km.run(dim, n, k, x, centroids_out, assign_out, iters_out, split_out);
}

/* kmeans_core from given centroids (centroids [k x dim], updated in place),
 * generator seeded `seed` (used by the empty-cluster split).  Returns 0 / -1. */
int pot_kmeans_core(uint32_t dim, uint32_t n, uint32_t k, const float* x, int niter, uint32_t seed, int mean_mode,
                    float* centroids, int32_t* assign_out, float* dis_out, double* qerr_out, int* iters_out,
                    int* split_out, double* qerr_trace) {
    Kmeans km(seed, false, mean_mode);
    std::vector<int> nassign(k);
    return km.core(dim, n, k, niter, centroids, x, assign_out, nassign.data(), dis_out, qerr_out, iters_out,
                   split_out, qerr_trace);
}

/* HierarchicalClusterIndex::train (hierarchical_cluster_index.cpp:882-985):
 * `iterations` rounds (kmeans_iterations_count) of coarse k-means on the data
 * less the best fine centers so far, fine k-means on the data less the new
 * coarse centers; a round whose fine error is at least 1e-4 below the best
 * keeps both, otherwise niter grows by 10 (the loop ends once it is above
 * 80).  Fine centers are then stored in ascending -sqrt(|f|^2/2) order (a
 * MaxHeap sort).  niter: the first Lloyd limit (0: the reference's 30).
 * err_trace [2 * iterations]: coarse and fine errors per
 * round run; rounds_out: rounds run; niter_out: the Lloyd iteration limit
 * reached.  Returns 0 / -1. */
int pot_train(uint32_t dim, uint32_t n, const float* x, uint32_t C, uint32_t F, int iterations, int niter,
              int init_pp, uint32_t seed, int mean_mode, float* coarse_out, float* fine_out, float* err_trace,
              int* rounds_out, int* niter_out) {
    Kmeans km(seed, init_pp != 0, mean_mode);
    if (niter > 0) km.niter = niter;
    const size_t nd = (size_t)n * dim;
    std::vector<float> tv(x, x + nd), cinit((size_t)C * dim), finit((size_t)F * dim);
    std::vector<float> cbest((size_t)C * dim, 0.f), fbest((size_t)F * dim, 0.f);
    std::vector<int> assign(n), fassign(n, 0);
    float min_err = std::numeric_limits<float>::max();
    int rounds = 0;
    for (int ite = 0; ite < iterations; ++ite) {
        ++rounds;
        float err = km.run(dim, n, C, tv.data(), cinit.data(), assign.data(), nullptr, nullptr);
        if (err < 0) return -1;
        if (err_trace) err_trace[2 * ite] = err;
        for (size_t i = 0; i < n; ++i)
            for (uint32_t t = 0; t < dim; ++t) tv[i * dim + t] = x[i * dim + t] - cinit[(size_t)assign[i] * dim + t];
        err = km.run(dim, n, F, tv.data(), finit.data(), assign.data(), nullptr, nullptr);
        if (err < 0) return -1;
        if (err_trace) err_trace[2 * ite + 1] = err;
        if ((min_err - err) >= 1e-4) {
            cbest = cinit;
            fbest = finit;
            fassign = assign;
            min_err = err;
        } else {
            if (km.niter > 80) break;
            km.niter += 10;
        }
        for (size_t i = 0; i < n; ++i)
            for (uint32_t t = 0; t < dim; ++t) tv[i * dim + t] = x[i * dim + t] - fbest[(size_t)fassign[i] * dim + t];
    }
    std::vector<float> key(F), sv(F);
    std::vector<uint32_t> tag(F), st(F);
    for (uint32_t j = 0; j < F; ++j) {
        const float* f = fbest.data() + (size_t)j * dim;
        const float h = po_sdot_short(f, f, dim) / 2;
        key[j] = 0 - std::sqrt(h);
        tag[j] = j;
    }
    po_heap_run(F, key.data(), tag.data(), F, 1, sv.data(), st.data());
    std::memcpy(coarse_out, cbest.data(), sizeof(float) * C * dim);
    for (uint32_t j = 0; j < F; ++j)
        std::memcpy(fine_out + (size_t)j * dim, fbest.data() + (size_t)st[j] * dim, sizeof(float) * dim);
    if (rounds_out) *rounds_out = rounds;
    if (niter_out) *niter_out = km.niter;
    return 0;
}

}  // extern "C"
