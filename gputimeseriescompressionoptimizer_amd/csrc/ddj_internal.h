// Internal declarations of libddj_hip.so: context, allocator, error
// plumbing and the host-side launchers of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ddj.h"

namespace ddj {

// Thrown inside the library, turned into a ddj_status at the C boundary.
struct Error {
    ddj_status status;
    std::string msg;
};

[[noreturn]] void fail(ddj_status st, const std::string& msg);
void hip_check(hipError_t e, const char* what, const char* file, int line);
#define HIPC(x) ::ddj::hip_check((x), #x, __FILE__, __LINE__)

size_t dtype_size(int dtype);
inline bool dtype_ok(int dtype) { return dtype >= DDJ_D_INT && dtype <= DDJ_D_CHAR; }

// Stream-ordered caching allocator: every kernel of a context runs on one
// stream, so a block freed by the host can be handed to the next request
// without synchronising.  Requests round up to a size class; blocks go back
// to a per-class free list.  Thread-safe: a batched call's worker threads
// hand blocks over while the calling thread allocates.
class Allocator {
public:
    explicit Allocator(int device) : device_(device) {}
    ~Allocator();
    void* alloc(size_t bytes);
    void free(void* p);
    // Hands a live block to the caller (released later with hipFree);
    // *cls receives its size class.
    bool detach(void* p, size_t* cls = nullptr);
    // Takes over a block detached from another allocator of the same device
    // (its stream work on the block must be complete).
    void adopt(void* p, size_t cls);
    void trim();
    size_t in_use() const { return in_use_; }

private:
    static size_t round_size(size_t bytes);
    void trim_locked();
    std::mutex mu_;
    int device_;
    std::map<size_t, std::vector<void*>> free_;
    std::map<void*, size_t> live_;
    size_t in_use_ = 0;
};

// Device-side control block: counters that must start at zero for the
// "last workgroup finishes the reduction" pattern, and the error flag.
struct Control {
    uint32_t counters[64];
    uint32_t error_flag;
    uint32_t pad[63];
};

// Per-kernel HIP-event timings, recorded on the context stream around each
// launch while profiling is enabled.
struct KernelTimes {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    double total_ms = 0;
    uint64_t launches = 0;
    double bytes = 0;  // algorithmic bytes (one read of the inputs, one write of the outputs)
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::unique_ptr<Allocator> alloc;
    Control* control = nullptr;  // device
    int cu_count = 256;
    std::string last_error;
    std::mutex mu;
    bool profiling = false;
    std::map<std::string, KernelTimes> ktimes;
    // Pinned staging ring for small uploads: h2d copies the source here and
    // returns without waiting; the ring restarts after the stream drains.
    uint8_t* staging = nullptr;
    uint8_t* staging_dev = nullptr;  // its device address, when the runtime gives one
    size_t staging_size = 0, staging_off = 0;
    // Pinned landing area for small read-backs (a copy into pageable memory
    // is staged by the runtime itself, which costs a further round trip).
    // The area holds kReadbackSize bytes and, after them, the completion
    // flag that the read-back launch sets to readback_seq.
    static constexpr size_t kReadbackSize = 64 << 10;
    uint8_t* readback = nullptr;
    uint8_t* readback_dev = nullptr;  // its device address, when the runtime gives one
    uint32_t readback_seq = 0;
    // set when a launch may raise the device error flag (device-side input
    // checks); the C ABI reads the flag back only then
    bool error_flag_armed = false;
    // Host copies of small device ranges that do not change while they are
    // registered (the metadata heads of a blob being decoded): read_small
    // serves reads inside them without a device round trip.
    struct Shadow {
        const uint8_t* dev;
        size_t n;
        const uint8_t* host;
    };
    std::vector<Shadow> shadows;
};

// Adds the algorithmic bytes of a launch to its kernel's tally.
inline void kernel_bytes(Ctx* c, const char* name, double b) {
    if (c->profiling) c->ktimes[name].bytes += b;
}

// Brackets one kernel launch with events when profiling is on.
class KernelTimer {
public:
    KernelTimer(Ctx* c, const char* name) : c_(c), name_(name) {
        if (!c->profiling) return;
        if (hipEventCreate(&a_) != hipSuccess || hipEventCreate(&b_) != hipSuccess) {
            a_ = b_ = nullptr;
            return;
        }
        (void)hipEventRecord(a_, c->stream);
    }
    ~KernelTimer() {
        if (!a_) return;
        (void)hipEventRecord(b_, c_->stream);
        c_->ktimes[name_].pending.push_back({a_, b_});
    }

private:
    Ctx* c_;
    const char* name_;
    hipEvent_t a_ = nullptr, b_ = nullptr;
};

// RAII device buffer from the context allocator.
class DevBuf {
public:
    DevBuf() = default;
    DevBuf(Ctx* c, size_t bytes) : ctx_(c), bytes_(bytes) { p_ = bytes ? c->alloc->alloc(bytes) : nullptr; }
    ~DevBuf() { reset(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept { *this = std::move(o); }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            reset();
            ctx_ = o.ctx_;
            p_ = o.p_;
            bytes_ = o.bytes_;
            o.p_ = nullptr;
            o.bytes_ = 0;
        }
        return *this;
    }
    void reset() {
        if (p_ && ctx_) ctx_->alloc->free(p_);
        p_ = nullptr;
        bytes_ = 0;
    }
    template <typename T = void> T* get() const { return static_cast<T*>(p_); }
    uint8_t* bytes_ptr() const { return static_cast<uint8_t*>(p_); }
    size_t size() const { return bytes_; }
    void set_size(size_t b) { bytes_ = b; }  // logical size <= allocated size
    // Hand the block to a ddj_buf (the caller becomes the owner).
    ddj_buf release() {
        ddj_buf b{p_, bytes_};
        p_ = nullptr;
        bytes_ = 0;
        return b;
    }

private:
    Ctx* ctx_ = nullptr;
    void* p_ = nullptr;
    size_t bytes_ = 0;
};

// A read-only view of device bytes (input of an encoder or decoder).
struct DView {
    const uint8_t* p = nullptr;
    size_t n = 0;
    template <typename T> const T* as() const { return reinterpret_cast<const T*>(p); }
};

inline DView view(const DevBuf& b) { return DView{b.bytes_ptr(), b.size()}; }

void d2h(Ctx* c, void* dst, const void* src, size_t bytes);  // synchronises the stream
// Several small read-backs with one stream synchronisation.
struct ReadBack {
    void* dst;
    const void* src;
    size_t bytes;
};
// Read-backs of library-owned device memory (small ones take one launch
// that writes the landing area directly).  d2h_any is for pointers a
// caller passed in, which the runtime's copy checks.
void d2h_many(Ctx* c, const ReadBack* rb, size_t count);
void d2h_any(Ctx* c, void* dst, const void* src, size_t bytes);
void h2d(Ctx* c, void* dst, const void* src, size_t bytes);  // stream-ordered
// A slot of `bytes` in the staging ring for a small table that kernels read
// directly through the returned device address (no upload launch); the
// caller fills staging + (address - staging_dev) before its launch.  Null
// when the ring has no device address or the table is too big.  Like a
// staged upload, the slot stays valid until the stream next drains.
const uint8_t* staging_slot(Ctx* c, size_t bytes);
void d2d(Ctx* c, void* dst, const void* src, size_t bytes);  // stream-ordered
DevBuf make_buf(Ctx* c, const void* host, size_t bytes);     // stream-ordered upload
// A few bytes of device memory to the host: from a registered shadow when
// one covers them, else a synchronous read-back.
void read_small(Ctx* c, void* dst, const void* src, size_t bytes);

}  // namespace ddj
