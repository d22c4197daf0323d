// Tree decompression into caller memory with no device allocation: the
// variant of ddj_tree_decompress for callers that cannot allocate (code under
// HIP graph capture, serving loops, harnesses that forbid hipMalloc).
//
// Trees: none, afl -> none and delta -> afl -> none over integer columns, the
// FL-leaf shapes whose decode needs no intermediate column.  The only device
// scratch is the fused FL-unpack + delta decode's per-tile sums and prefixes
// (two 64-bit words per 4096-row tile, plus one), kept in `scratch` below:
// 1 MiB, so at most 65,535 tiles, i.e. 268,431,361 rows per call.
//
// Each GPU holds its own copy of `scratch` for the life of the process.  In
// host code the array's name is not a device address: set-up takes each
// device's address once with hipGetSymbolAddress (with that device current)
// and every launch uses that address.  Calls share the array, so calls that
// overlap on one GPU (two contexts of the same device, two host threads)
// race on it; every word a call reads it first writes, so nothing needs
// clearing between calls.
#include <mutex>

#include "kernels.h"
#include "tree.h"

namespace ddj {

namespace {

constexpr size_t kScratchBytes = 1u << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint64_t scratch[kScratchBytes / sizeof(uint64_t)];

std::mutex g_mu;
uint64_t* g_scratch[kMaxDevices] = {};
hipError_t g_status[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};

}  // namespace

hipError_t noalloc_init_device(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_done[device]) return g_status[device];
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess && cur != device) e = hipSetDevice(device);
    void* p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur >= 0 && cur != device) (void)hipSetDevice(cur);
    g_status[device] = e;
    g_scratch[device] = e == hipSuccess ? static_cast<uint64_t*>(p) : nullptr;
    g_done[device] = true;
    return e;
}

namespace {

uint64_t* scratch_of(int device) {
    const hipError_t e = noalloc_init_device(device);
    if (e != hipSuccess || !g_scratch[device])
        fail(DDJ_EHIP, std::string("no-allocation scratch unavailable: ") + hipGetErrorString(e));
    return g_scratch[device];
}

bool is_int(int dt) { return dt == DDJ_D_INT || dt == DDJ_D_TIME || dt == DDJ_D_SHORT || dt == DDJ_D_CHAR; }

struct Hdr {
    int16_t enc, dtype;
    int32_t meta_len;
};

// Launches with per-kernel timing switched off: a timed launch creates HIP
// events, which this path does not.
struct NoTiming {
    Ctx* c;
    bool saved;
    explicit NoTiming(Ctx* cc) : c(cc), saved(cc->profiling) { c->profiling = false; }
    ~NoTiming() { c->profiling = saved; }
};

}  // namespace

void tree_decompress_noalloc(Ctx* c, const uint8_t* blob, size_t blob_bytes, void* out, size_t out_cap,
                             size_t* out_bytes, int* dtype_out) {
    if (!blob || blob_bytes < 8) fail(DDJ_ECORRUPT, "blob shorter than one node header");
    if (!out_bytes) fail(DDJ_EINVAL, "null output size");
    uint64_t* sc = scratch_of(c->device);
    NoTiming nt(c);
    // the node records of the supported shapes fit in the first 64 bytes:
    // one read-back into the context's pinned landing area
    uint8_t head[64] = {};
    const size_t hb = blob_bytes < sizeof head ? blob_bytes : sizeof head;
    d2h_any(c, head, blob, hb);
    size_t at = 0;
    auto header = [&]() -> Hdr {
        Hdr h;
        if (at + 8 > hb) fail(DDJ_EUNSUPPORTED, "tree shape not supported without allocation");
        memcpy(&h, head + at, 8);
        at += 8;
        if (h.meta_len < 0) fail(DDJ_ECORRUPT, "negative metadata length");
        return h;
    };
    const Hdr root = header();
    const int dt = root.dtype;
    if (!dtype_ok(dt)) fail(DDJ_ECORRUPT, "bad data type in the root header");
    const size_t es = dtype_size(dt);
    // the none leaf: [size_t length][data], which must end the blob
    auto leaf = [&](size_t& data_off, size_t& data_len) {
        if (at + 8 > hb) fail(DDJ_ECORRUPT, "truncated leaf");
        uint64_t len;
        memcpy(&len, head + at, 8);
        at += 8;
        if (len != blob_bytes - at) fail(DDJ_ECORRUPT, "leaf length does not end the blob");
        data_off = at;
        data_len = (size_t)len;
    };
    if (reinterpret_cast<uintptr_t>(out) % es != 0) fail(DDJ_EINVAL, "output not aligned to its row type");
    size_t data_off = 0, data_len = 0;
    if (root.enc == DDJ_ENC_NONE) {
        if (root.meta_len != 8) fail(DDJ_ECORRUPT, "none metadata is one size_t");
        leaf(data_off, data_len);
        if (data_len > out_cap) fail(DDJ_EINVAL, "output capacity smaller than the column");
        if (data_len) HIPC(hipMemcpyAsync(out, blob + data_off, data_len, hipMemcpyDeviceToDevice, c->stream));
        *out_bytes = data_len;
        if (dtype_out) *dtype_out = dt;
        return;
    }
    if (!is_int(dt)) fail(DDJ_EUNSUPPORTED, "tree shape not supported without allocation");
    // optional delta root: its metadata is the first row
    bool delta = false;
    size_t first_off = 0;
    Hdr afl = root;
    if (root.enc == DDJ_ENC_DELTA) {
        if ((size_t)root.meta_len != es) fail(DDJ_ECORRUPT, "delta metadata is one row");
        delta = true;
        first_off = at;
        at += es;
        afl = header();
        if (afl.dtype != dt) fail(DDJ_EUNSUPPORTED, "tree shape not supported without allocation");
    }
    // AFL of an empty input has empty metadata (afl_encoding.cu:21-24)
    if (afl.enc != DDJ_ENC_AFL || (afl.meta_len != 4 && afl.meta_len != 0))
        fail(DDJ_EUNSUPPORTED, "tree shape not supported without allocation");
    int bits = 0, rest = 0;
    if (afl.meta_len == 4) {
        if (at + 4 > hb) fail(DDJ_ECORRUPT, "truncated FL metadata");
        bits = (int8_t)head[at];
        rest = (int8_t)head[at + 1];
        at += 4;
    }
    const Hdr lh = header();
    if (lh.enc != DDJ_ENC_NONE || lh.dtype != dt || lh.meta_len != 8)
        fail(DDJ_EUNSUPPORTED, "tree shape not supported without allocation");
    leaf(data_off, data_len);
    // DecodeAfl (compression/afl/afl_encoding.cu:74-87): rows = (8 * bytes - rest) / b
    size_t m = 0;
    if (data_len) {
        if (bits <= 0 || bits > (int)(8 * es)) fail(DDJ_ECORRUPT, "FL bit width out of range");
        const long long bits_total = (long long)data_len * 8 - rest;
        if (bits_total < 0) fail(DDJ_ECORRUPT, "FL rest beyond the stream");
        m = (size_t)(bits_total / bits);
    }
    const size_t n = delta ? m + 1 : m;
    if (n * es > out_cap) fail(DDJ_EINVAL, "output capacity smaller than the column");
    if (delta && m == 0) {
        HIPC(hipMemcpyAsync(out, blob + first_off, es, hipMemcpyDeviceToDevice, c->stream));
    } else if (delta) {
        const size_t words = fl_unpack_delta_scratch_words(m);
        if (words * sizeof(uint64_t) > kScratchBytes)
            fail(DDJ_EINVAL, "column longer than the no-allocation scratch allows (268,431,361 rows)");
        launch_fl_unpack_delta(c, (int)es, blob + first_off, blob + data_off, data_len, m, bits, out, sc);
    } else {
        launch_fl_unpack(c, dt, blob + data_off, data_len, m, bits, out);
    }
    *out_bytes = n * es;
    if (dtype_out) *dtype_out = dt;
}

}  // namespace ddj
