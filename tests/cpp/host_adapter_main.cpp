// Drives puck::GpuPuckIndex the way demo/search_client.cpp drives PuckIndex:
// init, then Request/Response searches.  Usage:
//   host_adapter_test <index_path> <queries.f32> <nq> <dim> <topk> <FLAGS_topk> <scc> <filter_topk> <out.bin> [build|threads=T]
// Writes nq*topk floats, nq*topk uint32 labels and nq uint32 counts to out.bin.
// threads=T: T threads search concurrently (thread t takes queries t, t + T,
// ...), as the reference's ParallelFor does; the coalescing counters are
// printed as "batches B queries Q largest L".
//   host_adapter_test <index_path> type
// prints load_index_type() as an integer.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <thread>
#include <vector>

#include "../../puck_amd/host/gpu_puck_index.h"

int main(int argc, char** argv) {
    if (argc == 3 && std::string(argv[2]) == "type") {
        puck::GpuIndexFlags tf;
        tf.index_path = argv[1];
        printf("%d\n", (int)puck::load_index_type(tf));
        return 0;
    }
    if (argc < 10) return 2;
    puck::GpuIndexFlags f;
    f.index_path = argv[1];
    const uint32_t nq = atoi(argv[3]), dim = atoi(argv[4]), topk = atoi(argv[5]);
    f.topk = atoi(argv[6]);
    f.search_coarse_count = atoi(argv[7]);
    f.filter_topk = atoi(argv[8]);
    puck::GpuPuckIndex index(f);
    int threads = 0;
    if (argc > 10 && std::string(argv[10]).rfind("threads=", 0) == 0) threads = atoi(argv[10] + 8);
    if (argc > 10 && std::string(argv[10]) == "build" && index.build() != 0) {
        fprintf(stderr, "build failed: %s\n", index.last_error());
        return 3;
    }
    if (index.init() != 0) {
        fprintf(stderr, "init failed: %s\n", index.last_error());
        return 4;
    }
    std::vector<float> q((size_t)nq * dim);
    std::ifstream in(argv[2], std::ios::binary);
    in.read((char*)q.data(), (std::streamsize)(q.size() * 4));
    std::vector<float> D((size_t)nq * topk, 0.f);
    std::vector<uint32_t> I((size_t)nq * topk, 0u), R(nq, 0u);
    auto one = [&](uint32_t i) {
        puck::Request req;
        req.topk = topk;
        req.feature = q.data() + (size_t)i * dim;
        puck::Response res;
        res.distance = D.data() + (size_t)i * topk;
        res.local_idx = I.data() + (size_t)i * topk;
        if (index.search(&req, &res) != 0) return false;
        R[i] = res.result_num;
        return true;
    };
    if (threads > 0) {
        std::vector<std::thread> th;
        std::vector<int> ok(threads, 1);
        for (int t = 0; t < threads; ++t)
            th.emplace_back([&, t] {
                for (uint32_t i = t; i < nq; i += threads)
                    if (!one(i)) ok[t] = 0;
            });
        for (auto& x : th) x.join();
        for (int t = 0; t < threads; ++t)
            if (!ok[t]) {
                fprintf(stderr, "search failed: %s\n", index.last_error());
                return 5;
            }
        printf("batches %llu queries %llu largest %llu\n", (unsigned long long)index.coalesced_batches(),
               (unsigned long long)index.coalesced_queries(), (unsigned long long)index.largest_coalesced_batch());
    } else {
        for (uint32_t i = 0; i < nq; ++i)
            if (!one(i)) {
                fprintf(stderr, "search failed: %s\n", index.last_error());
                return 5;
            }
    }
    // a request above FLAGS_topk must fail with -1
    puck::Request bad;
    bad.topk = f.topk + 1;
    bad.feature = q.data();
    puck::Response br;
    if (index.search(&bad, &br) != -1) return 6;
    std::ofstream out(argv[9], std::ios::binary);
    out.write((const char*)D.data(), (std::streamsize)(D.size() * 4));
    out.write((const char*)I.data(), (std::streamsize)(I.size() * 4));
    out.write((const char*)R.data(), (std::streamsize)(R.size() * 4));
    return 0;
}
