// gpu_puck_index.h -- the reference's puck::Index API served by libpuckgpu.so.
//
// Request / Response / Index have the shape of puck/index.h:30-81 so that a
// caller written against the reference compiles unchanged; GpuPuckIndex
// stands in for PuckIndex (puck/puck/puck_index.h), whether_pq false or true.
// The host code reaches the GPU only through include/puck_gpu.h.
#pragma once
#include <cstdint>
#include <fstream>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/puck_gpu.h"
#include "group_commit.h"

namespace puck {

enum class IndexType { HIERARCHICAL_CLUSTER = 0, PUCK = 1, TINKER = 2 };

struct Request {
    uint32_t topk = 100;
    const float* feature = nullptr;
    virtual ~Request() { feature = nullptr; }
};

struct Response {
    float* distance = nullptr;
    uint32_t* local_idx = nullptr;
    uint32_t result_num = 0;
    virtual ~Response() {
        distance = nullptr;
        local_idx = nullptr;
        result_num = 0;
    }
};

class Index {
public:
    virtual ~Index() {}
    virtual int init() = 0;
    virtual int search(const Request* request, Response* response) = 0;
    virtual int train() = 0;
    virtual int build() = 0;
};

// The gflags this path reads (puck/gflags/puck_gflags.cpp:26-81), as plain
// fields with the reference's defaults.
struct GpuIndexFlags {
    std::string index_path = "puck_index";
    std::string index_file_name = "index.dat";
    std::string feature_file_name = "all_data.feat.bin";
    std::string coarse_codebook_file_name = "coarse.dat";
    std::string fine_codebook_file_name = "fine.dat";
    std::string cell_assign_file_name = "cell_assign.dat";
    std::string filter_codebook_file_name = "filter_codebook.dat";
    std::string filter_data_file_name = "filter_data.dat";
    std::string pq_codebook_file_name = "learn_codebooks.dat";
    std::string pq_data_file_name = "pq_data.dat";
    std::string label_file_name = "all_data.url";
    uint32_t topk = 100;
    uint32_t search_coarse_count = 200;
    int32_t filter_topk = -1;  // unset: 11 * topk (index_conf.cpp:178-180)
    double radius_rate = 1.0;
    int device = 0;
};

// The index type written in index.dat (load_index_type,
// hierarchical_cluster_index.cpp:1544-1549); IndexType(-1) when the file
// cannot be read.
IndexType load_index_type(const GpuIndexFlags& flags = GpuIndexFlags());

class GpuPuckIndex : public Index {
public:
    explicit GpuPuckIndex(const GpuIndexFlags& flags = GpuIndexFlags());
    ~GpuPuckIndex() override;
    // Reads index.dat and the index files, loads them into HBM (0 / -1).  A
    // whether_pq index loads learn_codebooks.dat and pq_data.dat in place of
    // the raw vectors (puck_index.cpp:198-227).
    int init() override;
    // One query (puck_index.cpp:559-600): -1 for topk > FLAGS_topk or a null
    // feature, leaving the response untouched.  Re-entrant: calls from several
    // threads (the reference's ParallelFor callers, py_api_wrapper.cpp:111-152)
    // are coalesced into one batched device search per group of callers that
    // arrive while the previous group runs (group_commit.h); every query's
    // result is the one it gets alone.
    int search(const Request* request, Response* response) override;
    // n queries, row stride topk (PySearcher::search semantics).
    int search_batch(uint32_t n, const float* queries, uint32_t topk, float* distance, uint32_t* labels,
                     uint32_t* result_num = nullptr);
    // Assigns all_data.feat.bin to cells and writes cell_assign.dat,
    // filter_data.dat (and pq_data.dat when index.dat says whether_pq) and
    // index.dat with the codebooks already on disk.
    int build() override;
    // Training runs in puck_amd.train (Python); returns -1 here.
    int train() override;
    const char* last_error() const { return _err.c_str(); }
    uint32_t dim() const { return _dim; }
    // Coalescing counters of search(): device batches run and queries served.
    uint64_t coalesced_batches() const { return _gc.batches(); }
    uint64_t coalesced_queries() const { return _gc.items(); }
    uint64_t largest_coalesced_batch() const { return _gc.largest_batch(); }

protected:
    struct PendingSearch {
        const float* feature;
        uint32_t topk;
        float* distance;
        uint32_t* local_idx;
        uint32_t result_num;
        int rc;
    };
    int init_locked();
    void run_coalesced(PendingSearch* const* items, size_t n);
    GpuIndexFlags _flags;
    puckgpu_ctx* _ctx = nullptr;
    puckgpu_index* _index = nullptr;
    puckgpu_search_params _sp{};
    uint32_t _dim = 0, _filter_nsq = 0, _pq_nsq = 0;
    bool _whether_pq = false;
    std::string _err;
    std::mutex _mu;  // one context: device calls (batches, inserts) from several threads are serialised
    GroupCommit<PendingSearch> _gc{4096};
    // staging of a coalesced batch (used under _mu)
    std::vector<float> _cq, _cd;
    std::vector<uint32_t> _ctopk, _ci, _cn;
};

// InsertRequest of realtime_insert_puck_index.h:142-149.
struct InsertRequest {
    std::string label;
    const float* feature = nullptr;
    uint32_t dim = 0;
};

// RealtimeInsertPuckIndex (puck/puck/realtime_insert_puck_index.h): init()
// takes the point count from the label file and cuts the index files to it
// (a crash between the appends of one insert leaves longer files,
// :52-129); insert() puts the point into HBM, searchable when it returns,
// then appends it to filter_data.dat, pq_data.dat or all_data.feat.bin,
// cell_assign.dat and the label file (:346-396).  Deviation: the reference
// appends to the files before the point becomes searchable; here a failed
// append returns -1 with the point already in the index.
class GpuRealtimeInsertPuckIndex : public GpuPuckIndex {
public:
    explicit GpuRealtimeInsertPuckIndex(const GpuIndexFlags& flags = GpuIndexFlags());
    int init() override;
    // 0, or -1 (null request or feature, a zero-norm feature under
    // whether_norm, a device or file error).
    int insert(const InsertRequest* request);
    // The label of local id label_id (get_label, :543-561): 0 / -1.
    int get_label(uint32_t label_id, std::string& label);

private:
    std::vector<std::string> _labels;
    std::ofstream _f_cell, _f_filter, _f_data, _f_label;
};

}  // namespace puck
