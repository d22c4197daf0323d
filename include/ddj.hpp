// ddj.hpp -- C++ classes with the reference's names over the C ABI (ddj.h).
//
// Header-only: Encoding (compression/encoding.hpp:18-113), CompressionTree
// (tree/compression_tree.hpp), CompressionOptimizer
// (optimizer/compression_optimizer.hpp:28-96) and ParallelTSCompressor
// (parallel_ts_compressor.hpp) keep the reference's method names and
// argument meaning; device memory is a DeviceBuffer owned by a Context
// instead of SharedCudaPtr<char>, and errors throw ddj::Error (a
// std::runtime_error carrying the ddj_status) instead of exit(1).
#ifndef DDJ_HPP_
#define DDJ_HPP_

#include <algorithm>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "ddj.h"

namespace ddj {

class Error : public std::runtime_error {
public:
    Error(ddj_status s, const std::string& msg) : std::runtime_error(msg), status(s) {}
    ddj_status status;
};

class Context;

// Device memory owned by a Context (freed with the object).
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    DeviceBuffer(ddj_ctx* ctx, ddj_buf b) : ctx_(ctx), b_(b) {}
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : ctx_(o.ctx_), b_(o.b_) { o.b_ = ddj_buf{nullptr, 0}; }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) {
            reset();
            ctx_ = o.ctx_;
            b_ = o.b_;
            o.b_ = ddj_buf{nullptr, 0};
        }
        return *this;
    }
    ~DeviceBuffer() { reset(); }
    void* get() const { return b_.dev; }
    size_t size() const { return b_.bytes; }
    ddj_buf raw() const { return b_; }
    std::vector<char> copyToHost() const {  // CudaPtr::copyToHost (core/cuda_ptr.hpp:115-121)
        std::vector<char> h(b_.bytes);
        if (b_.bytes) check(ddj_memcpy_d2h(ctx_, h.data(), b_.dev, b_.bytes));
        return h;
    }
    void reset() {
        if (ctx_ && b_.dev) ddj_buf_free(ctx_, &b_);
        b_ = ddj_buf{nullptr, 0};
    }
    static void check(ddj_status s, ddj_ctx* ctx = nullptr) {
        if (s != DDJ_OK) throw Error(s, ctx ? ddj_last_error(ctx) : "ddj call failed");
    }

private:
    ddj_ctx* ctx_ = nullptr;
    ddj_buf b_{nullptr, 0};
};

// A GPU, a HIP stream and a caching allocator.
class Context {
public:
    explicit Context(int device = 0) {
        ddj_ctx* c = nullptr;
        DeviceBuffer::check(ddj_ctx_create(&c, device));
        ctx_.reset(c, ddj_ctx_destroy);
    }
    ddj_ctx* get() const { return ctx_.get(); }
    void check(ddj_status s) const { DeviceBuffer::check(s, ctx_.get()); }
    DeviceBuffer alloc(size_t bytes) const {
        ddj_buf b;
        check(ddj_buf_alloc(get(), bytes, &b));
        return DeviceBuffer(get(), b);
    }
    DeviceBuffer upload(const void* host, size_t bytes) const {  // CudaPtr::fillFromHost
        DeviceBuffer d = alloc(bytes);
        if (bytes) check(ddj_memcpy_h2d(get(), d.get(), host, bytes));
        return d;
    }
    void sync() const { check(ddj_ctx_sync(get())); }

private:
    std::shared_ptr<ddj_ctx> ctx_;
};

// Encoding (compression/encoding.hpp:18-113) for one EncodingType.
class Encoding {
public:
    Encoding(const Context& ctx, int encoding_type) : ctx_(ctx), enc_(encoding_type) {}
    unsigned GetNumberOfResults(int dtype = DDJ_D_INT) const { return info(dtype).first; }
    std::vector<int> GetReturnTypes(int dtype) const { return info(dtype).second; }
    std::vector<DeviceBuffer> Encode(const DeviceBuffer& data, int dtype) const {
        const unsigned k = 1 + std::max(GetNumberOfResults(dtype), 1u);
        std::vector<ddj_buf> outs(k);
        ctx_.check(ddj_encode(ctx_.get(), enc_, dtype, data.get(), data.size(), outs.data()));
        std::vector<DeviceBuffer> r;
        for (auto& b : outs) r.emplace_back(ctx_.get(), b);
        return r;
    }
    DeviceBuffer Decode(const std::vector<DeviceBuffer>& in, int dtype) const {
        std::vector<ddj_buf> ins;
        for (auto& b : in) ins.push_back(b.raw());
        ddj_buf out;
        ctx_.check(ddj_decode(ctx_.get(), enc_, dtype, ins.data(), (uint32_t)ins.size(), &out));
        return DeviceBuffer(ctx_.get(), out);
    }
    size_t GetMetadataSize(const DeviceBuffer& data, int dtype) const { return sizes(data, dtype).first; }
    size_t GetCompressedSize(const DeviceBuffer& data, int dtype) const { return sizes(data, dtype).second; }
    static double GetCompressionRatio(size_t in, size_t out) {  // encoding.hpp:78-84
        return (out == 0 || out > in) ? 1.0 : (double)in / (double)out;
    }

private:
    std::pair<unsigned, std::vector<int>> info(int dtype) const {
        uint32_t n = 0;
        int rt[2] = {0, 0};
        ctx_.check(ddj_encoding_info(enc_, dtype, &n, rt));
        return {n, std::vector<int>(rt, rt + n)};
    }
    std::pair<size_t, size_t> sizes(const DeviceBuffer& data, int dtype) const {
        size_t m = 0, d = 0;
        ctx_.check(ddj_predict_sizes(ctx_.get(), enc_, dtype, data.get(), data.size(), &m, &d));
        return {m, d};
    }
    Context ctx_;
    int enc_;
};

// CompressionTree built from a pre-order path (PathGenerator::GenerateTree).
class CompressionTree {
public:
    CompressionTree(const Context& ctx, std::vector<int> path = {}, int dtype = DDJ_D_INT)
        : ctx_(ctx), path_(std::move(path)), dtype_(dtype) {}
    DeviceBuffer Compress(const DeviceBuffer& data) {
        ddj_buf blob;
        ratios_.assign(path_.size(), 1.0);
        ctx_.check(ddj_tree_compress(ctx_.get(), path_.data(), path_.size(), dtype_, data.get(), data.size(), &blob,
                                     ratios_.data()));
        return DeviceBuffer(ctx_.get(), blob);
    }
    DeviceBuffer Decompress(const DeviceBuffer& blob) {
        ddj_buf out;
        int dt = 0;
        ctx_.check(ddj_tree_decompress(ctx_.get(), blob.get(), blob.size(), &out, &dt));
        dtype_ = dt;
        return DeviceBuffer(ctx_.get(), out);
    }
    size_t GetPredictedSizeAfterCompression(const DeviceBuffer& data) const {
        size_t s = 0;
        ctx_.check(ddj_tree_predict_size(ctx_.get(), path_.data(), path_.size(), dtype_, data.get(), data.size(), &s));
        return s;
    }
    double GetCompressionRatio() const { return ratios_.empty() ? 1.0 : ratios_[0]; }
    const std::vector<int>& Path() const { return path_; }

private:
    Context ctx_;
    std::vector<int> path_;
    int dtype_;
    std::vector<double> ratios_;
};

// CompressionOptimizer (optimizer/compression_optimizer.hpp:28-96).
class CompressionOptimizer {
public:
    explicit CompressionOptimizer(const Context& ctx) : ctx_(ctx) {
        ddj_optimizer* o = nullptr;
        ctx_.check(ddj_optimizer_create(ctx_.get(), &o));
        opt_.reset(o, ddj_optimizer_destroy);
    }
    CompressionTree OptimizeTree(const DeviceBuffer& data, int dtype) const {
        return CompressionTree(ctx_, plan(DDJ_PLAN_OPTIMIZE30, data, dtype), dtype);
    }
    DeviceBuffer CompressData(const DeviceBuffer& part, int dtype) {
        ddj_buf blob;
        ctx_.check(ddj_optimizer_compress(opt_.get(), dtype, part.get(), part.size(), &blob));
        return DeviceBuffer(ctx_.get(), blob);
    }
    std::vector<int> GetOptimalTree() const {
        std::vector<int> p(4096);
        size_t n = 0;
        ctx_.check(ddj_optimizer_tree(opt_.get(), p.data(), p.size(), &n));
        p.resize(n);
        return p;
    }

private:
    std::vector<int> plan(int mode, const DeviceBuffer& data, int dtype) const {
        std::vector<int> p(4096);
        size_t n = 0;
        double ratio = 0;
        uint64_t cand = 0;
        ctx_.check(ddj_plan(ctx_.get(), mode, dtype, data.get(), data.size(), p.data(), p.size(), &n, &ratio, &cand));
        p.resize(n);
        return p;
    }
    Context ctx_;
    std::shared_ptr<ddj_optimizer> opt_;
};

// ParallelTSCompressor (parallel_ts_compressor.cpp:17-102) over host columns.
class ParallelTSCompressor {
public:
    ParallelTSCompressor(const Context& ctx, std::vector<int> dtypes, size_t batch_size = 1000000)
        : ctx_(ctx), dtypes_(std::move(dtypes)), batch_(batch_size) {
        make();
    }
    void SetBatchSize(size_t size) {
        batch_ = size;
        make();
    }
    void Compress(const std::vector<const void*>& columns, size_t rows, const std::string& path, bool append = false) {
        ctx_.check(ddj_ts_compress(ts_.get(), columns.data(), rows, path.c_str(), append ? 1 : 0));
    }
    // Calls sink(columns, rows) for every decompressed batch.
    template <typename Sink> void Decompress(const std::string& path, Sink&& sink) const {
        ddj_ts_reader* r = nullptr;
        ctx_.check(ddj_ts_reader_open(ctx_.get(), path.c_str(), (uint32_t)dtypes_.size(), dtypes_.data(), &r));
        std::unique_ptr<ddj_ts_reader, void (*)(ddj_ts_reader*)> guard(r, ddj_ts_reader_close);
        std::vector<std::vector<char>> cols;
        std::vector<void*> ptrs;
        for (int dt : dtypes_) {
            cols.emplace_back(batch_ * width(dt));
            ptrs.push_back(cols.back().data());
        }
        for (;;) {
            size_t rows = 0;
            ctx_.check(ddj_ts_reader_next(r, ptrs.data(), batch_, &rows));
            if (!rows) return;
            sink(ptrs, rows);
        }
    }

private:
    static size_t width(int dt) {
        switch (dt) {
        case DDJ_D_TIME:
        case DDJ_D_DOUBLE: return 8;
        case DDJ_D_SHORT: return 2;
        case DDJ_D_CHAR: return 1;
        default: return 4;
        }
    }
    void make() {
        ddj_ts_compressor* t = nullptr;
        ctx_.check(ddj_ts_create(ctx_.get(), (uint32_t)dtypes_.size(), dtypes_.data(), batch_, &t));
        ts_.reset(t, ddj_ts_destroy);
    }
    Context ctx_;
    std::vector<int> dtypes_;
    size_t batch_;
    std::shared_ptr<ddj_ts_compressor> ts_;
};

}  // namespace ddj

#endif
