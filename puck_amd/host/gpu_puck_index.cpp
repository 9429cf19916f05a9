// gpu_puck_index.cpp -- see gpu_puck_index.h.
#include "gpu_puck_index.h"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <vector>

namespace puck {
namespace {

struct Conf {
    int32_t index_type = -1;
    uint32_t dim = 0, N = 0, filter_nsq = 0, ks = 0, nsq = 0, ks2 = 0, ip2cos = 0, C = 0, F = 0;
    bool whether_norm = false, whether_filter = false, whether_pq = false;
};

// index.dat: size_t length, then the packed fields of
// HierarchicalClusterIndex::save_model_config (hierarchical_cluster_index.cpp:522-543).
bool read_conf(const std::string& path, Conf* c) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    uint64_t len = 0;
    in.read((char*)&len, 8);
    std::vector<char> b(43);
    in.read(b.data(), 43);
    if (!in || len < 43) return false;
    const char* p = b.data();
    auto rd32 = [&](void* dst) { std::memcpy(dst, p, 4); p += 4; };
    auto rdb = [&](bool* dst) { *dst = *p != 0; p += 1; };
    rd32(&c->index_type); rd32(&c->dim); rdb(&c->whether_norm); rd32(&c->N); rdb(&c->whether_filter);
    rd32(&c->filter_nsq); rd32(&c->ks); rdb(&c->whether_pq); rd32(&c->nsq); rd32(&c->ks2); rd32(&c->ip2cos);
    rd32(&c->C); rd32(&c->F);
    return true;
}

bool write_conf(const std::string& path, const Conf& c) {
    std::ofstream out(path, std::ios::binary);
    if (!out) return false;
    std::vector<char> b;
    auto w32 = [&](const void* v) { b.insert(b.end(), (const char*)v, (const char*)v + 4); };
    auto wb = [&](bool v) { b.push_back(v ? 1 : 0); };
    w32(&c.index_type); w32(&c.dim); wb(c.whether_norm); w32(&c.N); wb(c.whether_filter); w32(&c.filter_nsq);
    w32(&c.ks); wb(c.whether_pq); w32(&c.nsq); w32(&c.ks2); w32(&c.ip2cos); w32(&c.C); w32(&c.F);
    uint64_t len = b.size();
    out.write((const char*)&len, 8);
    out.write(b.data(), (std::streamsize)b.size());
    return (bool)out;
}

// fvecs rows (int32 dim + floats), checking every prefix (read_fvec_format).
bool read_fvecs(const std::string& path, uint32_t dim, uint64_t n, std::vector<float>* out) {
    std::ifstream in(path, std::ios::binary | std::ios::ate);
    if (!in) return false;
    uint64_t bytes = (uint64_t)in.tellg();
    if (bytes != n * (4 + 4ull * dim)) return false;
    in.seekg(0);
    out->resize(n * dim);
    for (uint64_t i = 0; i < n; ++i) {
        int32_t d = 0;
        in.read((char*)&d, 4);
        if (d != (int32_t)dim) return false;
        in.read((char*)(out->data() + i * dim), 4ll * dim);
    }
    return (bool)in;
}

template <typename T>
bool read_raw(const std::string& path, uint64_t count, std::vector<T>* out) {
    std::ifstream in(path, std::ios::binary | std::ios::ate);
    if (!in || (uint64_t)in.tellg() != count * sizeof(T)) return false;
    in.seekg(0);
    out->resize(count);
    in.read((char*)out->data(), (std::streamsize)(count * sizeof(T)));
    return (bool)in;
}

std::string join(const std::string& a, const std::string& b) { return a + "/" + b; }

}  // namespace

IndexType load_index_type(const GpuIndexFlags& flags) {
    Conf c;
    if (!read_conf(join(flags.index_path, flags.index_file_name), &c)) return (IndexType)-1;
    return (IndexType)c.index_type;
}

GpuPuckIndex::GpuPuckIndex(const GpuIndexFlags& flags) : _flags(flags) {}

GpuPuckIndex::~GpuPuckIndex() {
    if (_index) puckgpu_index_free(_index);
    if (_ctx) puckgpu_destroy(_ctx);
}

int GpuPuckIndex::init() {
    std::lock_guard<std::mutex> g(_mu);
    return init_locked();
}

int GpuPuckIndex::init_locked() {
    const std::string& P = _flags.index_path;
    Conf c;
    if (!read_conf(join(P, _flags.index_file_name), &c)) { _err = "model file read error"; return -1; }
    if (c.index_type != (int32_t)IndexType::PUCK) { _err = "index_type is not PUCK"; return -1; }
    if (c.ip2cos) { _err = "ip2cos indexes are not served by the GPU engine"; return -1; }
    if (c.filter_nsq == 0 || c.dim % c.filter_nsq) { _err = "filter_nsq does not divide dim"; return -1; }
    const uint32_t lsq = c.dim / c.filter_nsq;
    if (c.whether_pq && (c.nsq == 0 || c.dim % c.nsq)) { _err = "nsq does not divide dim"; return -1; }
    std::vector<float> coarse, fine, cb, feat, pcb;
    std::vector<uint32_t> cell;
    std::vector<uint8_t> rec, prec;
    bool ok = read_fvecs(join(P, _flags.coarse_codebook_file_name), c.dim, c.C, &coarse) &&
              read_fvecs(join(P, _flags.fine_codebook_file_name), c.dim, c.F, &fine) &&
              read_fvecs(join(P, _flags.filter_codebook_file_name), lsq, 256ull * c.filter_nsq, &cb) &&
              read_raw(join(P, _flags.cell_assign_file_name), c.N, &cell) &&
              read_raw(join(P, _flags.filter_data_file_name), (uint64_t)c.N * (4 + c.filter_nsq), &rec);
    // whether_pq: the PQ codes replace the raw vectors, which are not read
    if (ok && c.whether_pq)
        ok = read_fvecs(join(P, _flags.pq_codebook_file_name), c.dim / c.nsq, 256ull * c.nsq, &pcb) &&
             read_raw(join(P, _flags.pq_data_file_name), (uint64_t)c.N * (4 + c.nsq), &prec);
    else if (ok)
        ok = read_fvecs(join(P, _flags.feature_file_name), c.dim, c.N, &feat);
    if (!ok) {
        _err = "check file length has error";
        return -1;
    }
    // adaptive_search_param (index_conf.cpp:152-198)
    uint32_t ftk = _flags.filter_topk < 0 ? 11 * _flags.topk : (uint32_t)_flags.filter_topk;
    if (_flags.radius_rate < 1.0) { _err = "radius_rate should >= 1.0"; return -1; }
    if (ftk < _flags.topk) { _err = "filter_topk should >= topk"; return -1; }
    if (ftk > puckgpu_filter_topk_max()) { _err = "filter_topk above the engine's maximum"; return -1; }
    if (!_ctx && puckgpu_create(_flags.device, &_ctx) != 0) { _err = "no GPU context"; return -1; }
    if (_index) { puckgpu_index_free(_index); _index = nullptr; }
    puckgpu_index_params ip{c.dim, c.C, c.F, c.filter_nsq, c.N, c.whether_norm ? 1 : 0};
    if (puckgpu_index_load(_ctx, &ip, coarse.data(), fine.data(), cb.data(), cell.data(), rec.data(),
                           c.whether_pq ? nullptr : feat.data(), 0, &_index) != 0 ||
        (c.whether_pq &&
         puckgpu_index_attach_pq(_ctx, _index, c.nsq, pcb.data(), prec.data(), 0) != 0)) {
        _err = puckgpu_last_error(_ctx);
        if (_index) { puckgpu_index_free(_index); _index = nullptr; }
        return -1;
    }
    _sp.topk_max = _flags.topk;
    _sp.search_coarse_count = _flags.search_coarse_count < c.C ? _flags.search_coarse_count : c.C;
    _sp.filter_topk = ftk;
    _sp.radius_rate = _flags.radius_rate;
    _sp.replay_all = 0;
    _dim = c.dim;
    _filter_nsq = c.filter_nsq;
    _whether_pq = c.whether_pq;
    _pq_nsq = c.whether_pq ? c.nsq : 0;
    return 0;
}

int GpuPuckIndex::search(const Request* request, Response* response) {
    if (!request || !response || request->feature == nullptr || request->topk > _flags.topk) {
        std::lock_guard<std::mutex> g(_mu);
        _err = "topk should <= topk, or feature is nullptr";
        return -1;
    }
    if (request->topk > 0 && (!response->distance || !response->local_idx)) {
        std::lock_guard<std::mutex> g(_mu);
        _err = "response buffers are nullptr";
        return -1;
    }
    PendingSearch p{request->feature, request->topk, response->distance, response->local_idx, 0, -1};
    _gc.submit(&p, [this](PendingSearch* const* items, size_t n) { run_coalesced(items, n); });
    if (p.rc == 0) response->result_num = p.result_num;
    return p.rc;
}

// One device batch for the queued single-query calls: per-query topk
// (puckgpu_search_batch), rows of stride FLAGS_topk, each caller's first topk
// entries copied to its response buffers.
void GpuPuckIndex::run_coalesced(PendingSearch* const* items, size_t n) {
    std::lock_guard<std::mutex> g(_mu);
    if (!_index) {
        _err = "search before init";
        for (size_t i = 0; i < n; ++i) items[i]->rc = -1;
        return;
    }
    const uint32_t K = _sp.topk_max;
    _cq.resize(n * _dim);
    _ctopk.resize(n);
    _cd.resize(n * K);
    _ci.resize(n * K);
    _cn.resize(n);
    for (size_t i = 0; i < n; ++i) {
        std::memcpy(_cq.data() + i * _dim, items[i]->feature, sizeof(float) * _dim);
        _ctopk[i] = items[i]->topk;
    }
    puckgpu_query_params qp{};
    qp.topk = _ctopk.data();
    const int rc = puckgpu_search_batch(_ctx, _index, &_sp, &qp, (uint32_t)n, _cq.data(), _cd.data(), _ci.data(),
                                        _cn.data(), 0);
    if (rc != 0) _err = puckgpu_last_error(_ctx);
    for (size_t i = 0; i < n; ++i) {
        PendingSearch* p = items[i];
        p->rc = rc == 0 ? 0 : -1;
        if (rc != 0) continue;
        if (p->topk > 0) {
            std::memcpy(p->distance, _cd.data() + i * K, sizeof(float) * p->topk);
            std::memcpy(p->local_idx, _ci.data() + i * K, sizeof(uint32_t) * p->topk);
        }
        p->result_num = _cn[i];
    }
}

int GpuPuckIndex::search_batch(uint32_t n, const float* queries, uint32_t topk, float* distance, uint32_t* labels,
                               uint32_t* result_num) {
    std::lock_guard<std::mutex> g(_mu);
    if (!_index) { _err = "search before init"; return -1; }
    int rc = puckgpu_search(_ctx, _index, &_sp, n, queries, topk, distance, labels, result_num, 0);
    if (rc != 0) _err = puckgpu_last_error(_ctx);
    return rc == 0 ? 0 : -1;
}

int GpuPuckIndex::build() {
    std::lock_guard<std::mutex> g(_mu);
    const std::string& P = _flags.index_path;
    Conf c;
    if (!read_conf(join(P, _flags.index_file_name), &c)) { _err = "model file read error"; return -1; }
    if (c.filter_nsq == 0 || c.dim % c.filter_nsq) { _err = "filter_nsq does not divide dim"; return -1; }
    const uint32_t lsq = c.dim / c.filter_nsq;
    if (c.whether_pq && (c.nsq == 0 || c.dim % c.nsq)) { _err = "nsq does not divide dim"; return -1; }
    std::vector<float> coarse, fine, cb, feat, pcb;
    if (!read_fvecs(join(P, _flags.coarse_codebook_file_name), c.dim, c.C, &coarse) ||
        !read_fvecs(join(P, _flags.fine_codebook_file_name), c.dim, c.F, &fine) ||
        !read_fvecs(join(P, _flags.filter_codebook_file_name), lsq, 256ull * c.filter_nsq, &cb) ||
        !read_fvecs(join(P, _flags.feature_file_name), c.dim, c.N, &feat) ||
        (c.whether_pq && !read_fvecs(join(P, _flags.pq_codebook_file_name), c.dim / c.nsq, 256ull * c.nsq, &pcb))) {
        _err = "check file length has error";
        return -1;
    }
    if (!_ctx && puckgpu_create(_flags.device, &_ctx) != 0) { _err = "no GPU context"; return -1; }
    std::vector<uint32_t> cell(c.N);
    std::vector<uint8_t> rec((uint64_t)c.N * (4 + c.filter_nsq));
    if (puckgpu_assign_cells(_ctx, coarse.data(), fine.data(), c.C, c.F, c.dim, c.N, feat.data(), cell.data(), nullptr,
                             0) != 0 ||
        puckgpu_encode_filter(_ctx, coarse.data(), fine.data(), c.C, c.F, c.dim, cb.data(), c.filter_nsq, c.N,
                              feat.data(), cell.data(), rec.data(), 0) != 0) {
        _err = puckgpu_last_error(_ctx);
        return -1;
    }
    // whether_pq: the PQ encodes what the filter codewords leave (puck_index.cpp:602-709)
    std::vector<uint8_t> prec;
    if (c.whether_pq) {
        prec.resize((uint64_t)c.N * (4 + c.nsq));
        if (puckgpu_encode_pq(_ctx, coarse.data(), fine.data(), c.C, c.F, c.dim, cb.data(), c.filter_nsq, rec.data(),
                              pcb.data(), c.nsq, c.N, feat.data(), cell.data(), prec.data(), 0) != 0) {
            _err = puckgpu_last_error(_ctx);
            return -1;
        }
    }
    std::ofstream oc(join(P, _flags.cell_assign_file_name), std::ios::binary);
    oc.write((const char*)cell.data(), (std::streamsize)(4ull * c.N));
    std::ofstream orc(join(P, _flags.filter_data_file_name), std::ios::binary);
    orc.write((const char*)rec.data(), (std::streamsize)rec.size());
    bool pq_ok = true;
    if (c.whether_pq) {
        std::ofstream opq(join(P, _flags.pq_data_file_name), std::ios::binary);
        opq.write((const char*)prec.data(), (std::streamsize)prec.size());
        pq_ok = (bool)opq;
    }
    c.index_type = (int32_t)IndexType::PUCK;
    c.whether_filter = true;
    if (!oc || !orc || !pq_ok || !write_conf(join(P, _flags.index_file_name), c)) {
        _err = "write index files failed";
        return -1;
    }
    return 0;
}

int GpuPuckIndex::train() {
    _err = "training runs in puck_amd.train";
    return -1;
}

// ---------------------------------------------------------------------------
namespace {

// Cut a file to `bytes` when it is longer (RealtimeInsertPuckIndex::reorganize,
// realtime_insert_puck_index.cpp:110-129); false when it is shorter.
bool cut_to(const std::string& path, uint64_t bytes) {
    std::ifstream in(path, std::ios::binary | std::ios::ate);
    if (!in) return false;
    const uint64_t have = (uint64_t)in.tellg();
    if (have == bytes) return true;
    if (have < bytes) return false;
    std::vector<char> b(bytes);
    in.seekg(0);
    in.read(b.data(), (std::streamsize)bytes);
    in.close();
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    out.write(b.data(), (std::streamsize)bytes);
    return (bool)out;
}

}  // namespace

GpuRealtimeInsertPuckIndex::GpuRealtimeInsertPuckIndex(const GpuIndexFlags& flags) : GpuPuckIndex(flags) {}

int GpuRealtimeInsertPuckIndex::init() {
    std::lock_guard<std::mutex> g(_mu);
    const std::string& P = _flags.index_path;
    // read_labels (:131-152): the point count is the label file's line count
    _labels.clear();
    {
        std::ifstream in(join(P, _flags.label_file_name));
        if (!in) { _err = "open label file error"; return -1; }
        std::string line;
        while (std::getline(in, line)) _labels.push_back(line);
    }
    Conf c;
    if (!read_conf(join(P, _flags.index_file_name), &c)) { _err = "model file read error"; return -1; }
    const uint64_t N = _labels.size();
    // reorganize_inserted_index (:52-108)
    if (N > 0) {
        bool ok = cut_to(join(P, _flags.filter_data_file_name), N * (4ull + c.filter_nsq)) &&
                  cut_to(join(P, _flags.cell_assign_file_name), N * 4ull);
        if (ok && c.whether_pq) ok = cut_to(join(P, _flags.pq_data_file_name), N * (4ull + c.nsq));
        else if (ok) ok = cut_to(join(P, _flags.feature_file_name), N * (4ull + 4ull * c.dim));
        if (!ok) { _err = "index files shorter than the label file"; return -1; }
    }
    c.N = (uint32_t)N;
    if (!write_conf(join(P, _flags.index_file_name), c)) { _err = "model file write error"; return -1; }
    if (init_locked() != 0) return -1;
    // IndexFileHandle::init (:585-611): append handles
    _f_cell.open(join(P, _flags.cell_assign_file_name), std::ios::binary | std::ios::app);
    _f_filter.open(join(P, _flags.filter_data_file_name), std::ios::binary | std::ios::app);
    _f_data.open(join(P, _whether_pq ? _flags.pq_data_file_name : _flags.feature_file_name),
                 std::ios::binary | std::ios::app);
    _f_label.open(join(P, _flags.label_file_name), std::ios::app);
    if (!_f_cell || !_f_filter || !_f_data || !_f_label) { _err = "open index files for append failed"; return -1; }
    return 0;
}

int GpuRealtimeInsertPuckIndex::insert(const InsertRequest* request) {
    std::lock_guard<std::mutex> g(_mu);
    if (!request || !request->feature) { _err = "null request or feature"; return -1; }
    if (!_index) { _err = "insert before init"; return -1; }
    uint32_t id = 0, cell = 0;
    std::vector<uint8_t> rec(4 + _filter_nsq), prec(4 + _pq_nsq);
    std::vector<float> feat(_dim);
    if (puckgpu_index_insert(_ctx, _index, 1, request->feature, &id, &cell, rec.data(),
                             _whether_pq ? prec.data() : nullptr, feat.data(), 0) != 0) {
        _err = puckgpu_last_error(_ctx);
        return -1;
    }
    if (id == 0xffffffffu) { _err = "feature norm below 1e-6, could not be normalised"; return -1; }
    // append_index (:346-396): the quantized rows, the feature row (fvecs),
    // the cell, then the label
    _f_filter.write((const char*)rec.data(), (std::streamsize)rec.size());
    if (_whether_pq) {
        _f_data.write((const char*)prec.data(), (std::streamsize)prec.size());
    } else {
        const int32_t d = (int32_t)_dim;
        _f_data.write((const char*)&d, 4);
        _f_data.write((const char*)feat.data(), 4ll * _dim);
    }
    _f_cell.write((const char*)&cell, 4);
    _f_label << request->label << "\n";
    _f_filter.flush();
    _f_data.flush();
    _f_cell.flush();
    _f_label.flush();
    if (!_f_filter || !_f_data || !_f_cell || !_f_label) { _err = "append to index files failed"; return -1; }
    _labels.push_back(request->label);
    return 0;
}

int GpuRealtimeInsertPuckIndex::get_label(uint32_t label_id, std::string& label) {
    std::lock_guard<std::mutex> g(_mu);
    if (label_id >= _labels.size()) return -1;
    label = _labels[label_id];
    return 0;
}

}  // namespace puck
