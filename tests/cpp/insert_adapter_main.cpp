// Drives puck::GpuRealtimeInsertPuckIndex the way demo/insert_demo.cpp drives
// RealtimeInsertPuckIndex: init, inserts with labels, then Request/Response
// searches and get_label on every result.  Usage:
//   insert_adapter_test <index_path> <inserts.f32> <n_ins> <dim> <queries.f32> <nq> <topk> <FLAGS_topk> <scc>
//                       <filter_topk> <out.bin>
// Inserted point i gets the label "ins_<i>".  Writes nq*topk floats, nq*topk
// uint32 labels, nq uint32 counts to out.bin and the results' labels, one per
// line, to out.bin.labels.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "../../puck_amd/host/gpu_puck_index.h"

static std::vector<float> read_f32(const char* path, size_t count) {
    std::vector<float> v(count);
    std::ifstream in(path, std::ios::binary);
    in.read((char*)v.data(), (std::streamsize)(count * 4));
    return v;
}

int main(int argc, char** argv) {
    if (argc < 12) return 2;
    puck::GpuIndexFlags f;
    f.index_path = argv[1];
    const uint32_t n_ins = atoi(argv[3]), dim = atoi(argv[4]), nq = atoi(argv[6]), topk = atoi(argv[7]);
    f.topk = atoi(argv[8]);
    f.search_coarse_count = atoi(argv[9]);
    f.filter_topk = atoi(argv[10]);
    puck::GpuRealtimeInsertPuckIndex index(f);
    if (index.init() != 0) {
        fprintf(stderr, "init failed: %s\n", index.last_error());
        return 4;
    }
    std::vector<float> x = read_f32(argv[2], (size_t)n_ins * dim);
    for (uint32_t i = 0; i < n_ins; ++i) {
        puck::InsertRequest r;
        r.label = "ins_" + std::to_string(i);
        r.feature = x.data() + (size_t)i * dim;
        r.dim = dim;
        if (index.insert(&r) != 0) {
            fprintf(stderr, "insert %u failed: %s\n", i, index.last_error());
            return 5;
        }
    }
    puck::InsertRequest bad;
    if (index.insert(&bad) != -1) return 6;  // null feature
    std::vector<float> q = read_f32(argv[5], (size_t)nq * dim);
    std::vector<float> D((size_t)nq * topk, 0.f);
    std::vector<uint32_t> I((size_t)nq * topk, 0u), R(nq, 0u);
    std::ofstream lab(std::string(argv[11]) + ".labels");
    for (uint32_t i = 0; i < nq; ++i) {
        puck::Request req;
        req.topk = topk;
        req.feature = q.data() + (size_t)i * dim;
        puck::Response res;
        res.distance = D.data() + (size_t)i * topk;
        res.local_idx = I.data() + (size_t)i * topk;
        if (index.search(&req, &res) != 0) {
            fprintf(stderr, "search failed: %s\n", index.last_error());
            return 7;
        }
        R[i] = res.result_num;
        for (uint32_t j = 0; j < res.result_num; ++j) {
            std::string l;
            if (index.get_label(res.local_idx[j], l) != 0) {
                fprintf(stderr, "get_label(%u) failed\n", res.local_idx[j]);
                return 8;
            }
            lab << l << "\n";
        }
    }
    std::ofstream out(argv[11], std::ios::binary);
    out.write((const char*)D.data(), (std::streamsize)(D.size() * 4));
    out.write((const char*)I.data(), (std::streamsize)(I.size() * 4));
    out.write((const char*)R.data(), (std::streamsize)(R.size() * 4));
    return 0;
}
