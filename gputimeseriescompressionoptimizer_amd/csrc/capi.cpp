// The C ABI (include/ddj.h): status codes at the boundary, C++ inside.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <sys/stat.h>
#include <cstring>
#include <mutex>
#include <thread>

#include "codecs.h"
#include "learned.h"
#include "planner.h"
#include "tree.h"

using namespace ddj;

// Sets up a stream, control block, pinned staging ring and allocator on
// `device`; on failure everything made so far is released.
static void fini_ctx(Ctx& c);
static hipError_t init_ctx(Ctx& c, int device) {
    c.device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking);
    if (e == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
            c.cu_count = prop.multiProcessorCount;
        e = hipMalloc(reinterpret_cast<void**>(&c.control), sizeof(Control));
    }
    if (e == hipSuccess) e = hipMemset(c.control, 0, sizeof(Control));
    if (e == hipSuccess) {
        c.staging_size = 4u << 20;
        e = hipHostMalloc(reinterpret_cast<void**>(&c.staging), c.staging_size, hipHostMallocDefault);
        if (e == hipSuccess && hipHostGetDevicePointer(reinterpret_cast<void**>(&c.staging_dev), c.staging, 0) != hipSuccess) {
            (void)hipGetLastError();  // tables are then uploaded
            c.staging_dev = nullptr;
        }
    }
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&c.readback), Ctx::kReadbackSize + 64, hipHostMallocDefault);
    if (e == hipSuccess) memset(c.readback, 0, Ctx::kReadbackSize + 64);
    if (e == hipSuccess && hipHostGetDevicePointer(reinterpret_cast<void**>(&c.readback_dev), c.readback, 0) != hipSuccess) {
        (void)hipGetLastError();  // read-backs then go through the runtime's copies
        c.readback_dev = nullptr;
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        fini_ctx(c);
        return e;
    }
    c.alloc.reset(new Allocator(device));
    // the device address of the no-allocation decoder's scratch, taken once
    // per device; a failure is reported by that entry point, not here
    (void)noalloc_init_device(device);
    return hipSuccess;
}

static void fini_ctx(Ctx& c) {
    (void)hipSetDevice(c.device);
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    c.alloc.reset();
    if (c.control) (void)hipFree(c.control);
    if (c.staging) (void)hipHostFree(c.staging);
    if (c.readback) (void)hipHostFree(c.readback);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    c.control = nullptr;
    c.staging = nullptr;
    c.staging_dev = nullptr;
    c.readback = nullptr;
    c.readback_dev = nullptr;
    c.stream = nullptr;
}

// A Ctx that releases its device resources when destroyed.
struct OwnedCtx {
    Ctx c;
    ~OwnedCtx() { fini_ctx(c); }
};

// The captured graph of the last batch of FL-leaf columns, replayed while
// the batch layout (pointers, sizes, types, trees) stays the same.
struct BatchGraph {
    std::vector<std::tuple<const void*, size_t, int, int>> key;
    uint32_t range_rows_log2 = 0, wg_per_cu = 0;
    std::unique_ptr<DevBuf> cols, ranges, tiles, arena;
    std::vector<BatchColumn> host_cols;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~BatchGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

struct ddj_ctx {
    Ctx c;
    std::vector<std::unique_ptr<OwnedCtx>> pool;  // worker contexts of ddj_batch_compress
    std::unique_ptr<BatchGraph> bgraph;
};

struct ddj_optimizer {
    ddj_ctx* ctx;
    std::unique_ptr<Optimizer> opt;
};

namespace {

// The device error flag (2 = encoded input failed a device-side check, 3 = a
// requested row lies past the end of its column) is
// zero between calls.  A call whose kernels may raise it arms it; the flag
// is then read and cleared when the call ends, also when it fails.
void check_device_error(Ctx* c, bool report = true) {
    if (!c->error_flag_armed) return;
    c->error_flag_armed = false;
    uint32_t err = 0;
    d2h(c, &err, &c->control->error_flag, 4);
    if (!err) return;
    const uint32_t zero = 0;
    h2d(c, &c->control->error_flag, &zero, 4);
    if (report && err == 3) fail(DDJ_EINVAL, "row index beyond the end of the column");
    if (report) fail(DDJ_ECORRUPT, "encoded input failed a device-side check");
}

template <typename F> ddj_status guarded(ddj_ctx* ctx, F&& f) {
    if (!ctx) return DDJ_EINVAL;
    try {
        HIPC(hipSetDevice(ctx->c.device));
        f();
        check_device_error(&ctx->c);
        return DDJ_OK;
    } catch (const Error& e) {
        ctx->c.last_error = e.msg;
        try {
            check_device_error(&ctx->c, false);
        } catch (const Error&) {
        }
        return e.status;
    } catch (const std::bad_alloc&) {
        ctx->c.last_error = "host allocation failed";
        return DDJ_ENOMEM;
    } catch (const std::exception& e) {
        ctx->c.last_error = e.what();
        return DDJ_EINVAL;
    }
}

ddj_buf to_buf(Ctx* c, const Part& p) {
    // hand out a buffer the caller owns: reuse the block when the part owns
    // all of it, otherwise copy (host-resident metadata is uploaded)
    if (p.n == 0) return ddj_buf{nullptr, 0};
    if (p.on_host()) return make_buf(c, p.host->data(), p.n).release();
    if (p.owner && p.owner.use_count() == 1 && p.owner->bytes_ptr() == p.p) {
        DevBuf& b = *p.owner;
        b.set_size(p.n);
        return b.release();
    }
    DevBuf b(c, p.n);
    d2d(c, b.get(), p.p, p.n);
    return b.release();
}

std::vector<DView> views(const ddj_buf* ins, uint32_t n) {
    std::vector<DView> v;
    for (uint32_t i = 0; i < n; i++) v.push_back(DView{static_cast<const uint8_t*>(ins[i].dev), ins[i].bytes});
    return v;
}

}  // namespace

extern "C" {

ddj_status ddj_ctx_create(ddj_ctx** out, int device) {
    if (!out) return DDJ_EINVAL;
    *out = nullptr;
    auto* ctx = new (std::nothrow) ddj_ctx;
    if (!ctx) return DDJ_ENOMEM;
    const hipError_t e = init_ctx(ctx->c, device);
    if (e != hipSuccess) {
        fprintf(stderr, "ddj_ctx_create: %s\n", hipGetErrorString(e));
        delete ctx;
        return e == hipErrorOutOfMemory ? DDJ_ENOMEM : DDJ_EHIP;
    }
    *out = ctx;
    return DDJ_OK;
}

void ddj_ctx_destroy(ddj_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->c.device);
    (void)hipStreamSynchronize(ctx->c.stream);
    ctx->bgraph.reset();
    ctx->pool.clear();
    fini_ctx(ctx->c);
    delete ctx;
}

ddj_status ddj_ctx_sync(ddj_ctx* ctx) {
    return guarded(ctx, [&] { HIPC(hipStreamSynchronize(ctx->c.stream)); });
}

const char* ddj_last_error(ddj_ctx* ctx) { return ctx ? ctx->c.last_error.c_str() : "no context"; }

void ddj_ctx_trim(ddj_ctx* ctx) {
    if (ctx && ctx->c.alloc) ctx->c.alloc->trim();
}

void* ddj_ctx_stream(ddj_ctx* ctx) { return ctx ? (void*)ctx->c.stream : nullptr; }

static void drain_timers(Ctx* c) {
    HIPC(hipStreamSynchronize(c->stream));
    for (auto& kv : c->ktimes) {
        for (auto& ev : kv.second.pending) {
            float ms = 0;
            HIPC(hipEventElapsedTime(&ms, ev.first, ev.second));
            kv.second.total_ms += ms;
            kv.second.launches++;
            (void)hipEventDestroy(ev.first);
            (void)hipEventDestroy(ev.second);
        }
        kv.second.pending.clear();
    }
}

ddj_status ddj_profile_enable(ddj_ctx* ctx, int on) {
    return guarded(ctx, [&] {
        drain_timers(&ctx->c);
        ctx->c.profiling = on != 0;
    });
}

ddj_status ddj_profile_kernel(ddj_ctx* ctx, const char* kernel, double* total_ms, uint64_t* launches,
                              double* algo_bytes) {
    return guarded(ctx, [&] {
        if (!kernel || !total_ms || !launches) fail(DDJ_EINVAL, "null argument");
        drain_timers(&ctx->c);
        auto it = ctx->c.ktimes.find(kernel);
        *total_ms = it == ctx->c.ktimes.end() ? 0.0 : it->second.total_ms;
        *launches = it == ctx->c.ktimes.end() ? 0 : it->second.launches;
        if (algo_bytes) *algo_bytes = it == ctx->c.ktimes.end() ? 0.0 : it->second.bytes;
    });
}

ddj_status ddj_profile_reset(ddj_ctx* ctx) {
    return guarded(ctx, [&] {
        drain_timers(&ctx->c);
        ctx->c.ktimes.clear();
    });
}

ddj_status ddj_buf_alloc(ddj_ctx* ctx, size_t bytes, ddj_buf* out) {
    return guarded(ctx, [&] {
        if (!out) fail(DDJ_EINVAL, "null output");
        DevBuf b(&ctx->c, bytes);
        *out = b.release();
    });
}

void ddj_buf_free(ddj_ctx* ctx, ddj_buf* buf) {
    if (!ctx || !buf) return;
    if (buf->dev) ctx->c.alloc->free(buf->dev);
    buf->dev = nullptr;
    buf->bytes = 0;
}

ddj_status ddj_buf_detach(ddj_ctx* ctx, ddj_buf* buf) {
    return guarded(ctx, [&] {
        if (!buf) fail(DDJ_EINVAL, "null buffer");
        if (buf->dev && !ctx->c.alloc->detach(buf->dev)) fail(DDJ_EINVAL, "buffer not owned by this context");
    });
}

ddj_status ddj_memcpy_h2d(ddj_ctx* ctx, void* dst, const void* src, size_t bytes) {
    return guarded(ctx, [&] { h2d(&ctx->c, dst, src, bytes); });
}

ddj_status ddj_memcpy_d2h(ddj_ctx* ctx, void* dst, const void* src, size_t bytes) {
    return guarded(ctx, [&] { d2h_any(&ctx->c, dst, src, bytes); });
}

ddj_status ddj_memcpy_d2d(ddj_ctx* ctx, void* dst, const void* src, size_t bytes) {
    return guarded(ctx, [&] { d2d(&ctx->c, dst, src, bytes); });
}

ddj_status ddj_encoding_info(int enc, int dtype, uint32_t* n, int* rt) {
    try {
        if (!enc_ok(enc) || !dtype_ok(dtype) || !n || !rt) return DDJ_EINVAL;
        *n = (uint32_t)n_results(enc);
        return_types(enc, dtype, rt);
        return DDJ_OK;
    } catch (const Error& e) {
        return e.status;
    }
}

ddj_status ddj_encode(ddj_ctx* ctx, int enc, int dtype, const void* d_in, size_t in_bytes, ddj_buf* outs) {
    return guarded(ctx, [&] {
        if (!outs || (in_bytes && !d_in)) fail(DDJ_EINVAL, "null argument");
        auto parts = encode(&ctx->c, enc, dtype, DView{static_cast<const uint8_t*>(d_in), in_bytes});
        for (size_t i = 0; i < parts.size(); i++) outs[i] = to_buf(&ctx->c, parts[i]);
    });
}

ddj_status ddj_decode(ddj_ctx* ctx, int enc, int dtype, const ddj_buf* ins, uint32_t n_ins, ddj_buf* out) {
    return guarded(ctx, [&] {
        if (!ins || !out) fail(DDJ_EINVAL, "null argument");
        Part p = decode(&ctx->c, enc, dtype, views(ins, n_ins));
        *out = to_buf(&ctx->c, p);
    });
}

ddj_status ddj_predict_sizes(ddj_ctx* ctx, int enc, int dtype, const void* d_in, size_t in_bytes, size_t* meta,
                             size_t* data) {
    return guarded(ctx, [&] {
        if (!meta || !data) fail(DDJ_EINVAL, "null argument");
        predict_sizes(&ctx->c, enc, dtype, DView{static_cast<const uint8_t*>(d_in), in_bytes}, meta, data);
    });
}

ddj_status ddj_stats(ddj_ctx* ctx, int dtype, const void* d_in, size_t in_bytes, ddj_stats_t* out) {
    return guarded(ctx, [&] {
        if (!out || !dtype_ok(dtype)) fail(DDJ_EINVAL, "bad argument");
        *out = stats(&ctx->c, dtype, DView{static_cast<const uint8_t*>(d_in), in_bytes});
    });
}

ddj_status ddj_fl_pack(ddj_ctx* ctx, int dtype, int bits, const void* d_in, size_t n, void* d_out) {
    return guarded(ctx, [&] {
        if (!dtype_ok(dtype) || dtype == DDJ_D_FLOAT || dtype == DDJ_D_DOUBLE) fail(DDJ_EINVAL, "integer types only");
        if (bits <= 0 || bits > 8 * (int)dtype_size(dtype)) fail(DDJ_EINVAL, "bit width out of range");
        launch_fl_pack(&ctx->c, dtype, d_in, n, nullptr, bits, d_out);
    });
}

ddj_status ddj_fl_unpack(ddj_ctx* ctx, int dtype, int bits, const void* d_in, size_t in_bytes, size_t n, void* d_out) {
    return guarded(ctx, [&] {
        if (!dtype_ok(dtype) || dtype == DDJ_D_FLOAT || dtype == DDJ_D_DOUBLE) fail(DDJ_EINVAL, "integer types only");
        launch_fl_unpack(&ctx->c, dtype, static_cast<const uint8_t*>(d_in), in_bytes, n, bits, d_out);
    });
}

static void check_int_type(int dtype) {
    if (!dtype_ok(dtype) || dtype == DDJ_D_FLOAT || dtype == DDJ_D_DOUBLE) fail(DDJ_EINVAL, "integer types only");
}

ddj_status ddj_fl_fetch(ddj_ctx* ctx, int dtype, int bits, int lanes, const void* d_stream, size_t stream_bytes,
                        size_t n, const uint64_t* d_rows, size_t count, void* d_out) {
    return guarded(ctx, [&] {
        check_int_type(dtype);
        if (count && (!d_stream || !d_rows || !d_out)) fail(DDJ_EINVAL, "null argument");
        launch_fl_fetch(&ctx->c, dtype, bits, lanes, static_cast<const uint8_t*>(d_stream), stream_bytes, n, d_rows,
                        count, d_out);
    });
}

size_t ddj_fl_lanes_bytes(int dtype, int bits, size_t n) {
    if (!dtype_ok(dtype) || dtype == DDJ_D_FLOAT || dtype == DDJ_D_DOUBLE) return 0;
    if (bits <= 0 || bits > 8 * (int)dtype_size(dtype)) return 0;
    return fl_lanes_bytes(dtype, bits, n);
}

ddj_status ddj_fl_pack_lanes(ddj_ctx* ctx, int dtype, int bits, const void* d_in, size_t n, void* d_out) {
    return guarded(ctx, [&] {
        check_int_type(dtype);
        const size_t es = dtype_size(dtype);
        if (n && (!d_in || !d_out)) fail(DDJ_EINVAL, "null argument");
        if (reinterpret_cast<uintptr_t>(d_in) % es || reinterpret_cast<uintptr_t>(d_out) % es)
            fail(DDJ_EINVAL, "column and stream must be aligned to the row type");
        launch_fl_pack_lanes(&ctx->c, dtype, bits, d_in, n, d_out);
    });
}

ddj_status ddj_fl_unpack_lanes(ddj_ctx* ctx, int dtype, int bits, const void* d_in, size_t in_bytes, size_t n,
                               void* d_out) {
    return guarded(ctx, [&] {
        check_int_type(dtype);
        const size_t es = dtype_size(dtype);
        if (n && (!d_in || !d_out)) fail(DDJ_EINVAL, "null argument");
        if (reinterpret_cast<uintptr_t>(d_in) % es || reinterpret_cast<uintptr_t>(d_out) % es)
            fail(DDJ_EINVAL, "column and stream must be aligned to the row type");
        if (bits <= 0 || bits > 8 * (int)es) fail(DDJ_EINVAL, "bit width out of range");
        if (in_bytes < fl_lanes_bytes(dtype, bits, n)) fail(DDJ_EINVAL, "stream shorter than n values of b bits");
        launch_fl_unpack_lanes(&ctx->c, dtype, bits, d_in, n, d_out);
    });
}

ddj_status ddj_tree_compress(ddj_ctx* ctx, const int* preorder_path, size_t path_len, int dtype, const void* d_in,
                             size_t in_bytes, ddj_buf* blob, double* node_ratios) {
    return guarded(ctx, [&] {
        if (!preorder_path || !blob || !dtype_ok(dtype)) fail(DDJ_EINVAL, "bad argument");
        NodeP root = tree_from_path(preorder_path, path_len, dtype);
        fix_tree(root);
        Part p = compress_tree(&ctx->c, root, DView{static_cast<const uint8_t*>(d_in), in_bytes});
        *blob = to_buf(&ctx->c, p);
        if (node_ratios) {
            size_t i = 0;
            std::function<void(const NodeP&)> rec = [&](const NodeP& n) {
                if (i < path_len) node_ratios[i] = n->ratio;
                i++;
                for (auto& ch : n->ch) rec(ch);
            };
            rec(root);
        }
    });
}

ddj_status ddj_tree_decompress(ddj_ctx* ctx, const void* d_blob, size_t blob_bytes, ddj_buf* out, int* dtype) {
    return guarded(ctx, [&] {
        if (!out) fail(DDJ_EINVAL, "null output");
        int dt = 0;
        Part p = decompress_tree(&ctx->c, DView{static_cast<const uint8_t*>(d_blob), blob_bytes}, &dt);
        *out = to_buf(&ctx->c, p);
        if (dtype) *dtype = dt;
    });
}

ddj_status ddj_tree_decompress_rows(ddj_ctx* ctx, const void* d_blob, size_t blob_bytes, size_t row_begin,
                                   size_t row_count, ddj_buf* out, int* dtype) {
    return guarded(ctx, [&] {
        if (!out) fail(DDJ_EINVAL, "null output");
        int dt = 0;
        Part p = decompress_tree_rows(&ctx->c, DView{static_cast<const uint8_t*>(d_blob), blob_bytes}, row_begin,
                                      row_count, &dt);
        *out = to_buf(&ctx->c, p);
        if (dtype) *dtype = dt;
    });
}

ddj_status ddj_tree_decompress_noalloc(ddj_ctx* ctx, const void* d_blob, size_t blob_bytes, void* d_out,
                                      size_t out_cap, size_t* out_bytes, int* dtype) {
    return guarded(ctx, [&] {
        tree_decompress_noalloc(&ctx->c, static_cast<const uint8_t*>(d_blob), blob_bytes, d_out, out_cap, out_bytes,
                                dtype);
    });
}

ddj_status ddj_tree_predict_size(ddj_ctx* ctx, const int* preorder_path, size_t path_len, int dtype, const void* d_in,
                                 size_t in_bytes, size_t* size) {
    return guarded(ctx, [&] {
        if (!preorder_path || !size || !dtype_ok(dtype)) fail(DDJ_EINVAL, "bad argument");
        NodeP root = tree_from_path(preorder_path, path_len, dtype);
        *size = predict_tree(&ctx->c, root, DView{static_cast<const uint8_t*>(d_in), in_bytes});
    });
}

static void write_path(const NodeP& t, int* path_out, size_t cap, size_t* len) {
    auto pre = preorder(t);
    if (pre.size() > cap) fail(DDJ_EINVAL, "path buffer too small");
    memcpy(path_out, pre.data(), pre.size() * sizeof(int));
    *len = pre.size();
}

ddj_status ddj_plan(ddj_ctx* ctx, int mode, int dtype, const void* d_in, size_t in_bytes, int* path_out,
                    size_t path_cap, size_t* path_len, double* ratio, uint64_t* n_candidates) {
    return guarded(ctx, [&] {
        if (!path_out || !path_len || !dtype_ok(dtype)) fail(DDJ_EINVAL, "bad argument");
        DView in{static_cast<const uint8_t*>(d_in), in_bytes};
        NodeP t;
        if (mode == DDJ_PLAN_EXHAUSTIVE) {
            size_t sz = 0;
            t = plan_exhaustive(&ctx->c, in, dtype, &sz, n_candidates);
        } else if (mode == DDJ_PLAN_OPTIMIZE30) {
            t = optimize_tree(&ctx->c, in, dtype);
            if (n_candidates) *n_candidates = 30;
        } else {
            fail(DDJ_EUNSUPPORTED, "planner mode not available");
        }
        if (ratio) *ratio = t->ratio;
        write_path(t, path_out, path_cap, path_len);
    });
}

ddj_status ddj_optimizer_create(ddj_ctx* ctx, ddj_optimizer** out) {
    return guarded(ctx, [&] {
        if (!out) fail(DDJ_EINVAL, "null output");
        auto* o = new ddj_optimizer;
        o->ctx = ctx;
        o->opt.reset(new Optimizer(&ctx->c));
        *out = o;
    });
}

void ddj_optimizer_destroy(ddj_optimizer* opt) { delete opt; }

ddj_status ddj_optimizer_compress(ddj_optimizer* opt, int dtype, const void* d_in, size_t in_bytes, ddj_buf* blob) {
    if (!opt) return DDJ_EINVAL;
    return guarded(opt->ctx, [&] {
        if (!blob || !dtype_ok(dtype)) fail(DDJ_EINVAL, "bad argument");
        Part p = opt->opt->compress_data(DView{static_cast<const uint8_t*>(d_in), in_bytes}, dtype);
        *blob = to_buf(&opt->ctx->c, p);
    });
}

ddj_status ddj_optimizer_tree(ddj_optimizer* opt, int* path_out, size_t path_cap, size_t* path_len) {
    if (!opt) return DDJ_EINVAL;
    return guarded(opt->ctx, [&] {
        if (!opt->opt->tree()) fail(DDJ_EINVAL, "no tree yet");
        write_path(opt->opt->tree(), path_out, path_cap, path_len);
    });
}

// ------------------------------------------------------------- column API
// ParallelTSCompressor (parallel_ts_compressor.cpp:17-102) with its
// CompressionTask / DecompressionTask (optimizer/compression_task.cpp:35-77,
// optimizer/decompression_task.cpp:19-46): batches of rows, one stateful
// optimizer per column, each column's blob written as [size_t bytes][blob]
// in column order, batch after batch.

}  // extern "C"

struct ddj_planner {
    ddj_ctx* ctx;
    std::unique_ptr<PlannerNet> net;
    int cls = kClassInt;  // the column class its candidates and weights are for
};

struct ddj_ts_compressor {
    ddj_ctx* ctx;
    std::vector<int> dtypes;
    size_t batch_rows;
    std::vector<std::unique_ptr<Optimizer>> opts;
};

struct ddj_ts_reader {
    ddj_ctx* ctx;
    std::vector<int> dtypes;
    FILE* f;
};

extern "C" {

ddj_status ddj_ts_create(ddj_ctx* ctx, uint32_t n_columns, const int* dtypes, size_t batch_rows,
                         ddj_ts_compressor** out) {
    return guarded(ctx, [&] {
        if (!out || !dtypes || n_columns == 0 || batch_rows == 0) fail(DDJ_EINVAL, "bad argument");
        auto* t = new ddj_ts_compressor;
        t->ctx = ctx;
        t->batch_rows = batch_rows;
        for (uint32_t i = 0; i < n_columns; i++) {
            if (!dtype_ok(dtypes[i])) {
                delete t;
                fail(DDJ_EINVAL, "bad column type");
            }
            t->dtypes.push_back(dtypes[i]);
            t->opts.emplace_back(new Optimizer(&ctx->c));
        }
        *out = t;
    });
}

void ddj_ts_destroy(ddj_ts_compressor* t) { delete t; }

ddj_status ddj_ts_compress(ddj_ts_compressor* t, const void* const* host_columns, size_t rows, const char* out_path,
                           int append) {
    if (!t) return DDJ_EINVAL;
    return guarded(t->ctx, [&] {
        if (!host_columns || !out_path) fail(DDJ_EINVAL, "bad argument");
        FILE* f = fopen(out_path, append ? "ab" : "wb");
        if (!f) fail(DDJ_EINVAL, std::string("cannot open ") + out_path);
        std::unique_ptr<FILE, int (*)(FILE*)> guard(f, fclose);
        Ctx* c = &t->ctx->c;
        std::vector<uint8_t> host;
        for (size_t r0 = 0; r0 < rows; r0 += t->batch_rows) {
            const size_t nr = std::min(t->batch_rows, rows - r0);
            for (size_t k = 0; k < t->dtypes.size(); k++) {
                const size_t es = dtype_size(t->dtypes[k]);
                DevBuf d(c, nr * es);
                h2d(c, d.get(), static_cast<const uint8_t*>(host_columns[k]) + r0 * es, nr * es);
                Part blob = t->opts[k]->compress_data(DView{d.bytes_ptr(), nr * es}, t->dtypes[k]);
                host.resize(blob.n);
                d2h(c, host.data(), blob.p, blob.n);
                const size_t sz = blob.n;
                if (fwrite(&sz, sizeof sz, 1, f) != 1 || (sz && fwrite(host.data(), 1, sz, f) != sz))
                    fail(DDJ_EINVAL, "write failed");
            }
        }
    });
}

ddj_status ddj_ts_reader_open(ddj_ctx* ctx, const char* in_path, uint32_t n_columns, const int* dtypes,
                              ddj_ts_reader** out) {
    return guarded(ctx, [&] {
        if (!out || !in_path || !dtypes || n_columns == 0) fail(DDJ_EINVAL, "bad argument");
        for (uint32_t i = 0; i < n_columns; i++)
            if (!dtype_ok(dtypes[i])) fail(DDJ_EINVAL, "bad column type");
        FILE* f = fopen(in_path, "rb");
        if (!f) fail(DDJ_EINVAL, std::string("cannot open ") + in_path);
        auto* r = new ddj_ts_reader;
        r->ctx = ctx;
        r->dtypes.assign(dtypes, dtypes + n_columns);
        r->f = f;
        *out = r;
    });
}

void ddj_ts_reader_close(ddj_ts_reader* r) {
    if (!r) return;
    if (r->f) fclose(r->f);
    delete r;
}

ddj_status ddj_ts_reader_next(ddj_ts_reader* r, void* const* host_columns, size_t cap_rows, size_t* rows) {
    if (!r) return DDJ_EINVAL;
    return guarded(r->ctx, [&] {
        if (!host_columns || !rows) fail(DDJ_EINVAL, "bad argument");
        *rows = 0;
        Ctx* c = &r->ctx->c;
        // Every column of the batch is read and decoded before any is copied
        // out; on an error the file goes back to where the batch began and
        // the caller's buffers are left as they were.
        const off_t start = ftello(r->f);
        struct stat st;
        if (start < 0 || fstat(fileno(r->f), &st) != 0) fail(DDJ_EINVAL, "cannot read the column file");
        const uint64_t file_bytes = (uint64_t)st.st_size;
        struct Rewind {
            FILE* f;
            off_t at;
            bool armed = true;
            ~Rewind() {
                if (armed) fseeko(f, at, SEEK_SET);
            }
        } rewind_on_error{r->f, start};
        std::vector<uint8_t> host;
        std::vector<Part> cols;
        size_t batch = 0;
        for (size_t k = 0; k < r->dtypes.size(); k++) {
            size_t sz = 0;
            if (fread(&sz, sizeof sz, 1, r->f) != 1) {
                if (k == 0 && feof(r->f)) {  // end of file
                    rewind_on_error.armed = false;
                    return;
                }
                fail(DDJ_ECORRUPT, "truncated column file");
            }
            const off_t at = ftello(r->f);
            if (at < 0 || sz > file_bytes - (uint64_t)at) fail(DDJ_ECORRUPT, "blob length beyond the end of the file");
            host.resize(sz);
            if (sz && fread(host.data(), 1, sz, r->f) != sz) fail(DDJ_ECORRUPT, "truncated blob");
            DevBuf d(c, sz);
            h2d(c, d.get(), host.data(), sz);
            int dt = -1;
            Part out = decompress_tree(c, DView{d.bytes_ptr(), sz}, &dt);
            if (dt != r->dtypes[k]) fail(DDJ_ECORRUPT, "column type differs from the blob's");
            const size_t es = dtype_size(dt), nr = out.n / es;
            if (k == 0) batch = nr;
            if (nr != batch) fail(DDJ_ECORRUPT, "columns of one batch differ in length");
            if (nr > cap_rows) fail(DDJ_EINVAL, "batch larger than the capacity given");
            cols.push_back(std::move(out));
        }
        for (size_t k = 0; k < cols.size(); k++) d2h(c, host_columns[k], cols[k].p, cols[k].n);
        rewind_on_error.armed = false;
        *rows = batch;
    });
}

}  // extern "C"

// ---------------------------------------------------------- learned planner

extern "C" {

ddj_status ddj_planner_candidates(int* paths, size_t cap, size_t* offsets) {
    return ddj_planner_candidates_class(kClassInt, paths, cap, offsets);
}

ddj_status ddj_planner_candidates_class(int cls, int* paths, size_t cap, size_t* offsets) {
    if (cls != kClassInt && cls != kClassFloat) return DDJ_EINVAL;
    const auto& c = candidate_paths(cls);
    size_t total = 0;
    for (auto& p : c) total += p.size();
    if (!offsets || (paths && cap < total)) return DDJ_EINVAL;
    size_t off = 0;
    for (size_t i = 0; i < c.size(); i++) {
        offsets[i] = off;
        if (paths) memcpy(paths + off, c[i].data(), sizeof(int) * c[i].size());
        off += c[i].size();
    }
    offsets[c.size()] = off;
    return DDJ_OK;
}

ddj_status ddj_planner_features(ddj_ctx* ctx, int dtype, const void* d_in, size_t in_bytes, float* out) {
    return guarded(ctx, [&] {
        if (!out || !dtype_ok(dtype)) fail(DDJ_EINVAL, "bad argument");
        auto f = block_features(&ctx->c, dtype, DView{static_cast<const uint8_t*>(d_in), in_bytes});
        memcpy(out, f.data(), sizeof(float) * f.size());
    });
}

ddj_status ddj_planner_labels(ddj_ctx* ctx, int dtype, const void* d_in, size_t in_bytes, float* out) {
    return guarded(ctx, [&] {
        if (!out || !dtype_ok(dtype)) fail(DDJ_EINVAL, "bad argument");
        auto l = candidate_labels(&ctx->c, dtype, DView{static_cast<const uint8_t*>(d_in), in_bytes});
        memcpy(out, l.data(), sizeof(float) * l.size());
    });
}

ddj_status ddj_planner_create(ddj_ctx* ctx, uint64_t seed, ddj_planner** out) {
    return ddj_planner_create_class(ctx, kClassInt, seed, out);
}

ddj_status ddj_planner_create_class(ddj_ctx* ctx, int cls, uint64_t seed, ddj_planner** out) {
    return guarded(ctx, [&] {
        if (!out) fail(DDJ_EINVAL, "null output");
        if (cls != kClassInt && cls != kClassFloat) fail(DDJ_EINVAL, "unknown column class");
        auto* p = new ddj_planner;
        p->ctx = ctx;
        p->cls = cls;
        p->net.reset(new PlannerNet(&ctx->c, seed));
        *out = p;
    });
}

void ddj_planner_destroy(ddj_planner* p) { delete p; }

ddj_status ddj_planner_params(ddj_planner* p, float* get_out, const float* set_in) {
    if (!p) return DDJ_EINVAL;
    return guarded(p->ctx, [&] {
        if (set_in) p->net->set(set_in);
        if (get_out) p->net->get(get_out);
    });
}

ddj_status ddj_planner_gradient(ddj_planner* p, const float* X, const float* T, size_t n, float* grad_out,
                                float* loss) {
    if (!p) return DDJ_EINVAL;
    return guarded(p->ctx, [&] {
        if (!X || !T) fail(DDJ_EINVAL, "bad argument");
        const float l = p->net->gradient(X, T, n, grad_out);
        if (loss) *loss = l;
    });
}

ddj_status ddj_planner_train(ddj_planner* p, const float* X, const float* T, size_t n, int epochs, int batch,
                             float lr, uint64_t seed, float* loss) {
    if (!p) return DDJ_EINVAL;
    return guarded(p->ctx, [&] {
        if (!X || !T) fail(DDJ_EINVAL, "bad argument");
        const float l = p->net->train(X, T, n, epochs, batch, lr, seed);
        if (loss) *loss = l;
    });
}

ddj_status ddj_planner_predict(ddj_planner* p, const float* X, size_t n, float* Y) {
    if (!p) return DDJ_EINVAL;
    return guarded(p->ctx, [&] {
        if (!X || !Y) fail(DDJ_EINVAL, "bad argument");
        p->net->predict(X, n, Y);
    });
}

ddj_status ddj_planner_plan(ddj_planner* p, int dtype, const void* d_in, size_t in_bytes, float min_margin,
                            int* path_out, size_t path_cap, size_t* path_len, int* used_fallback, float* margin) {
    if (!p) return DDJ_EINVAL;
    return guarded(p->ctx, [&] {
        if (!dtype_ok(dtype) || !path_len) fail(DDJ_EINVAL, "bad argument");
        if (planner_class(dtype) != p->cls) fail(DDJ_EINVAL, "the planner's weights are for another column class");
        Ctx* c = &p->ctx->c;
        const DView data{static_cast<const uint8_t*>(d_in), in_bytes};
        auto f = block_features(c, dtype, data);
        float y[kMlpOutputs];
        p->net->predict(f.data(), 1, y);
        int best = -1, second = -1;  // among the candidates this type can take
        for (int i = 0; i < kMlpOutputs; i++) {
            if (!candidate_valid(i, dtype)) continue;
            if (best < 0 || y[i] < y[best]) {
                second = best;
                best = i;
            } else if (second < 0 || y[i] < y[second]) {
                second = i;
            }
        }
        if (best < 0 || second < 0) fail(DDJ_EHIP, "fewer than two candidate trees for this type");
        const float m = y[second] - y[best];
        if (margin) *margin = m;
        NodeP tree;
        if (m < min_margin) {
            tree = plan_exhaustive(c, data, dtype, nullptr, nullptr);
            if (used_fallback) *used_fallback = 1;
        } else {
            const auto& cp = candidate_paths(p->cls)[best];
            tree = tree_from_path(cp.data(), cp.size(), dtype);
            fix_tree(tree);
            if (used_fallback) *used_fallback = 0;
        }
        write_path(tree, path_out, path_cap, path_len);
    });
}

ddj_status ddj_planner_plan_checked(ddj_planner* p, int dtype, const void* d_in, size_t in_bytes, float within,
                                    int max_checked, int* path_out, size_t path_cap, size_t* path_len, int* checked) {
    if (!p) return DDJ_EINVAL;
    return guarded(p->ctx, [&] {
        if (!dtype_ok(dtype) || !path_len || !(within >= 0.0f)) fail(DDJ_EINVAL, "bad argument");
        if (planner_class(dtype) != p->cls) fail(DDJ_EINVAL, "the planner's weights are for another column class");
        Ctx* c = &p->ctx->c;
        const DView data{static_cast<const uint8_t*>(d_in), in_bytes};
        auto f = block_features(c, dtype, data);
        float y[kMlpOutputs];
        p->net->predict(f.data(), 1, y);
        std::vector<int> order;
        for (int i = 0; i < kMlpOutputs; i++)
            if (candidate_valid(i, dtype)) order.push_back(i);
        if (order.empty()) fail(DDJ_EHIP, "no candidate tree for this type");
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return y[a] < y[b]; });
        // the candidates predicted within `within` of the best, best first
        std::vector<int> near;
        for (int i : order) {
            if ((int)near.size() >= std::max(1, max_checked) || y[i] - y[order[0]] > within) break;
            near.push_back(i);
        }
        size_t bk = 0;  // the smallest exact size; ties go to the better prediction
        if (near.size() > 1) {
            const std::vector<size_t> sz = candidate_sizes(c, dtype, data, near);
            for (size_t k = 1; k < near.size(); k++)
                if (sz[k] < sz[bk]) bk = k;
        }
        const int pick = near[bk];
        if (checked) *checked = near.size() > 1 ? (int)near.size() : 0;
        const auto& cp = candidate_paths(p->cls)[(size_t)pick];
        NodeP tree = tree_from_path(cp.data(), cp.size(), dtype);
        fix_tree(tree);
        write_path(tree, path_out, path_cap, path_len);
    });
}

}  // extern "C"

// ------------------------------------------------------------ batched trees

namespace {

// 1: afl -> none, 2: delta -> afl -> none on an aligned int32 / int64 column
// (the graph-captured batch path); 0: anything else.
int fl_leaf_kind(const ddj_column& col) {
    if (col.dtype != DDJ_D_INT && col.dtype != DDJ_D_TIME) return 0;
    if ((reinterpret_cast<uintptr_t>(col.d_in) & 15) != 0 || !col.preorder) return 0;
    NodeP t = tree_from_path(col.preorder, col.path_len, col.dtype);
    fix_tree(t);
    const std::vector<int> p = preorder(t);
    const size_t n = col.in_bytes / dtype_size(col.dtype);
    if (p == std::vector<int>{DDJ_ENC_AFL, DDJ_ENC_NONE} && n >= 1) return 1;
    if (p == std::vector<int>{DDJ_ENC_DELTA, DDJ_ENC_AFL, DDJ_ENC_NONE} && n >= 2) return 2;
    return 0;
}

// Compresses the FL-leaf columns `idx` of `cols` through the captured graph
// (captured on first use of a layout) into blobs owned by ctx.
void batch_fl_graph(ddj_ctx* ctx, const ddj_column* cols, const std::vector<uint32_t>& idx,
                    const ddj_batch_config& cfg, ddj_buf* blobs) {
    Ctx* c = &ctx->c;
    std::vector<std::tuple<const void*, size_t, int, int>> key;
    for (uint32_t i : idx) key.emplace_back(cols[i].d_in, cols[i].in_bytes, cols[i].dtype, fl_leaf_kind(cols[i]));
    if (!ctx->bgraph || ctx->bgraph->key != key || ctx->bgraph->range_rows_log2 != cfg.range_rows_log2 ||
        ctx->bgraph->wg_per_cu != cfg.wg_per_cu) {
        ctx->bgraph.reset();
        std::unique_ptr<BatchGraph> g(new BatchGraph);
        g->key = key;
        g->range_rows_log2 = cfg.range_rows_log2;
        g->wg_per_cu = cfg.wg_per_cu;
        std::vector<BatchItem> ranges, tiles;
        uint64_t arena = 16;
        const uint64_t kRange = 1ull << cfg.range_rows_log2;  // rows per min/max item
        for (size_t k = 0; k < idx.size(); k++) {
            const ddj_column& col = cols[idx[k]];
            BatchColumn d{};
            d.x = static_cast<const uint8_t*>(col.d_in);
            d.es = (uint32_t)dtype_size(col.dtype);
            d.n = col.in_bytes / d.es;
            d.delta = std::get<3>(key[k]) == 2 ? 1 : 0;
            d.dtype = (uint32_t)col.dtype;
            g->host_cols.push_back(d);
            const uint64_t m = d.delta ? d.n - 1 : d.n;
            for (uint64_t s0 = 0; s0 < m; s0 += kRange)
                ranges.push_back(BatchItem{(uint32_t)k, 0, s0, std::min(m, s0 + kRange)});
            for (uint64_t t = 0; t * kFlTileRows < m; t++) tiles.push_back(BatchItem{(uint32_t)k, 0, t, 0});
            arena += ((d.delta ? 8 + d.es : 0) + 28 + m * d.es + 15) & ~15ull;
        }
        g->cols.reset(new DevBuf(c, sizeof(BatchColumn) * g->host_cols.size()));
        g->ranges.reset(new DevBuf(c, sizeof(BatchItem) * std::max<size_t>(ranges.size(), 1)));
        g->tiles.reset(new DevBuf(c, sizeof(BatchItem) * std::max<size_t>(tiles.size(), 1)));
        g->arena.reset(new DevBuf(c, arena));
        h2d(c, g->cols->get(), g->host_cols.data(), sizeof(BatchColumn) * g->host_cols.size());
        if (!ranges.empty()) h2d(c, g->ranges->get(), ranges.data(), sizeof(BatchItem) * ranges.size());
        if (!tiles.empty()) h2d(c, g->tiles->get(), tiles.data(), sizeof(BatchItem) * tiles.size());
        HIPC(hipStreamSynchronize(c->stream));
        HIPC(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        launch_batch_fl(c, g->cols->get<BatchColumn>(), (int)g->host_cols.size(), g->ranges->get<BatchItem>(),
                        (int)ranges.size(), g->tiles->get<BatchItem>(), (int)tiles.size(), g->arena->bytes_ptr(),
                        cfg.wg_per_cu);
        HIPC(hipStreamEndCapture(c->stream, &g->graph));
        HIPC(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
        ctx->bgraph = std::move(g);
    }
    BatchGraph* g = ctx->bgraph.get();
    HIPC(hipGraphLaunch(g->exec, c->stream));
    std::vector<BatchColumn> out(g->host_cols.size());
    d2h(c, out.data(), g->cols->get(), sizeof(BatchColumn) * out.size());
    // every blob into its own allocation, one copy launch for all of them
    std::vector<CopyPair> pairs;
    uint64_t max_bytes = 8;
    for (size_t k = 0; k < idx.size(); k++) {
        DevBuf b(c, out[k].blob_bytes);
        pairs.push_back(CopyPair{g->arena->bytes_ptr() + out[k].out_off, b.bytes_ptr(), out[k].blob_bytes});
        max_bytes = std::max<uint64_t>(max_bytes, out[k].blob_bytes);
        blobs[idx[k]] = b.release();
    }
    DevBuf dp(c, sizeof(CopyPair) * pairs.size());
    h2d(c, dp.get(), pairs.data(), sizeof(CopyPair) * pairs.size());
    launch_copy_pairs(c, dp.get<CopyPair>(), (int)pairs.size(), max_bytes);
    HIPC(hipStreamSynchronize(c->stream));
}

ddj_batch_config resolve(const ddj_batch_config* in) {
    ddj_batch_config c = in ? *in : ddj_batch_config{0, 0, 0};
    if (c.workers == 0) c.workers = 4;  // the hardware queues one process gets
    if (c.range_rows_log2 == 0) c.range_rows_log2 = 18;
    if (c.wg_per_cu == 0) c.wg_per_cu = 4;
    if (c.workers > 16 || c.range_rows_log2 < 12 || c.range_rows_log2 > 24 || c.wg_per_cu > 16)
        fail(DDJ_EINVAL, "batch configuration out of range");
    return c;
}

void batch_run(ddj_ctx* ctx, const ddj_column* cols, uint32_t n_cols, const ddj_batch_config& cfg, ddj_buf* blobs) {
    if ((!cols || !blobs) && n_cols) fail(DDJ_EINVAL, "bad argument");
    const uint32_t workers = cfg.workers;
    while (ctx->pool.size() < workers) {
        std::unique_ptr<OwnedCtx> w(new OwnedCtx);
        const hipError_t e = init_ctx(w->c, ctx->c.device);
        if (e != hipSuccess) fail(e == hipErrorOutOfMemory ? DDJ_ENOMEM : DDJ_EHIP, hipGetErrorString(e));
        ctx->pool.push_back(std::move(w));
    }
    for (uint32_t i = 0; i < n_cols; i++) blobs[i] = ddj_buf{nullptr, 0};
    HIPC(hipStreamSynchronize(ctx->c.stream));  // inputs written on the caller's stream are in place
    // FL-leaf columns go through the captured graph on the caller's
    // stream; the others through the worker contexts, concurrently
    std::vector<uint32_t> graph_cols, other_cols;
    for (uint32_t i = 0; i < n_cols; i++) (fl_leaf_kind(cols[i]) ? graph_cols : other_cols).push_back(i);
    std::atomic<uint32_t> next{0};
    std::mutex mu;  // the caller's allocator and the first error
    ddj_status first_st = DDJ_OK;
    std::string first_msg;
    auto work = [&](Ctx* w) {
        (void)hipSetDevice(w->device);
        for (;;) {
            const uint32_t q = next.fetch_add(1);
            if (q >= other_cols.size()) return;
            const uint32_t i = other_cols[q];
            try {
                NodeP t = tree_from_path(cols[i].preorder, cols[i].path_len, cols[i].dtype);
                fix_tree(t);
                Part blob = compress_tree(w, t, DView{static_cast<const uint8_t*>(cols[i].d_in), cols[i].in_bytes});
                HIPC(hipStreamSynchronize(w->stream));
                check_device_error(w);
                // the blob's block moves from the worker's allocator to the caller's
                const ddj_buf b = blob.owner->release();
                size_t cls = 0;
                if (!w->alloc->detach(b.dev, &cls)) fail(DDJ_EHIP, "blob not owned by the worker");
                {
                    std::lock_guard<std::mutex> lk(mu);
                    ctx->c.alloc->adopt(b.dev, cls);
                }
                blobs[i] = ddj_buf{b.dev, blob.n};
            } catch (const Error& e) {
                std::lock_guard<std::mutex> lk(mu);
                if (first_st == DDJ_OK) {
                    first_st = e.status;
                    first_msg = "column " + std::to_string(i) + ": " + e.msg;
                }
            } catch (const std::exception& e) {
                std::lock_guard<std::mutex> lk(mu);
                if (first_st == DDJ_OK) {
                    first_st = DDJ_EINVAL;
                    first_msg = "column " + std::to_string(i) + ": " + e.what();
                }
            }
        }
    };
    std::vector<std::thread> th;
    for (uint32_t k = 0; k < workers && k < other_cols.size(); k++) th.emplace_back(work, &ctx->pool[k]->c);
    if (!graph_cols.empty()) {
        try {
            batch_fl_graph(ctx, cols, graph_cols, cfg, blobs);
        } catch (const Error& e) {
            std::lock_guard<std::mutex> lk(mu);
            if (first_st == DDJ_OK) {
                first_st = e.status;
                first_msg = "graph batch: " + e.msg;
            }
        }
    }
    for (auto& t : th) t.join();
    if (first_st != DDJ_OK) {
        for (uint32_t i = 0; i < n_cols; i++)
            if (blobs[i].dev) ctx->c.alloc->free(blobs[i].dev), blobs[i] = ddj_buf{nullptr, 0};
        fail(first_st, first_msg);
    }
}

}  // namespace

extern "C" {

ddj_status ddj_batch_compress(ddj_ctx* ctx, const ddj_column* cols, uint32_t n_cols, uint32_t workers,
                              ddj_buf* blobs) {
    return guarded(ctx, [&] {
        ddj_batch_config c{workers > 16 ? 16 : workers, 0, 0};
        batch_run(ctx, cols, n_cols, resolve(&c), blobs);
    });
}

ddj_status ddj_batch_compress_config(ddj_ctx* ctx, const ddj_column* cols, uint32_t n_cols,
                                     const ddj_batch_config* config, ddj_buf* blobs) {
    return guarded(ctx, [&] { batch_run(ctx, cols, n_cols, resolve(config), blobs); });
}

ddj_status ddj_batch_autotune(ddj_ctx* ctx, const ddj_column* cols, uint32_t n_cols, ddj_batch_config* best,
                              double* best_ms) {
    return guarded(ctx, [&] {
        if (!best) fail(DDJ_EINVAL, "null output");
        std::vector<ddj_buf> blobs(std::max<uint32_t>(n_cols, 1));
        auto once = [&](const ddj_batch_config& cfg) {
            batch_run(ctx, cols, n_cols, cfg, blobs.data());
            HIPC(hipStreamSynchronize(ctx->c.stream));
            for (uint32_t i = 0; i < n_cols; i++) ddj_buf_free(ctx, &blobs[i]);
        };
        auto timed = [&](const ddj_batch_config& cfg) {
            once(cfg);  // graph capture, worker contexts, allocator warm-up
            const auto t0 = std::chrono::steady_clock::now();
            once(cfg);
            return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        };
        bool has_graph = false;
        for (uint32_t i = 0; i < n_cols; i++) has_graph = has_graph || fl_leaf_kind(cols[i]) != 0;
        ddj_batch_config cur = resolve(nullptr);
        double cur_ms = 1e300;
        for (uint32_t w : {2u, 4u, 8u, 16u}) {
            const ddj_batch_config cfg{w, cur.range_rows_log2, cur.wg_per_cu};
            const double ms = timed(cfg);
            if (ms < cur_ms) cur_ms = ms, cur = cfg;
        }
        if (has_graph) {
            const ddj_batch_config base = cur;
            for (uint32_t r : {16u, 18u, 20u})
                for (uint32_t g : {2u, 4u, 8u}) {
                    const ddj_batch_config cfg{base.workers, r, g};
                    if (cfg.range_rows_log2 == base.range_rows_log2 && cfg.wg_per_cu == base.wg_per_cu) continue;
                    const double ms = timed(cfg);
                    if (ms < cur_ms) cur_ms = ms, cur = cfg;
                }
        }
        *best = cur;
        if (best_ms) *best_ms = cur_ms;
    });
}

}  // extern "C"
