// Context, stream, caching allocator and copies.
#include "ddj_internal.h"
#include "kernels.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <thread>

namespace ddj {

[[noreturn]] void fail(ddj_status st, const std::string& msg) { throw Error{st, msg}; }

void hip_check(hipError_t e, const char* what, const char* file, int line) {
    if (e == hipSuccess) return;
    char buf[512];
    snprintf(buf, sizeof buf, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) fail(DDJ_ENOMEM, buf);
    fail(DDJ_EHIP, buf);
}

size_t dtype_size(int dtype) {
    switch (dtype) {
    case DDJ_D_INT: return 4;
    case DDJ_D_TIME: return 8;
    case DDJ_D_FLOAT: return 4;
    case DDJ_D_DOUBLE: return 8;
    case DDJ_D_SHORT: return 2;
    case DDJ_D_CHAR: return 1;
    default: return 0;
    }
}

// Size classes: powers of two up to 1 MiB, then multiples of 2 MiB.  Every
// block also has 64 spare bytes so kernels may touch whole 16-byte vectors
// past a logical end.
size_t Allocator::round_size(size_t bytes) {
    bytes += 64;
    if (bytes <= (1u << 20)) {
        size_t s = 256;
        while (s < bytes) s <<= 1;
        return s;
    }
    const size_t g = 2u << 20;
    return (bytes + g - 1) / g * g;
}

void* Allocator::alloc(size_t bytes) {
    std::lock_guard<std::mutex> lk(mu_);
    size_t s = round_size(bytes);
    auto it = free_.find(s);
    if (it != free_.end() && !it->second.empty()) {
        void* p = it->second.back();
        it->second.pop_back();
        live_[p] = s;
        in_use_ += s;
        return p;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, s);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        trim_locked();  // give cached blocks back and retry once
        e = hipMalloc(&p, s);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            fail(DDJ_ENOMEM, "device allocation of " + std::to_string(s) + " bytes failed");
        }
    }
    live_[p] = s;
    in_use_ += s;
    return p;
}

void Allocator::free(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(mu_);
    auto it = live_.find(p);
    if (it == live_.end()) return;
    free_[it->second].push_back(p);
    in_use_ -= it->second;
    live_.erase(it);
}

bool Allocator::detach(void* p, size_t* cls) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = live_.find(p);
    if (it == live_.end()) return false;
    if (cls) *cls = it->second;
    in_use_ -= it->second;
    live_.erase(it);
    return true;
}

void Allocator::adopt(void* p, size_t cls) {
    std::lock_guard<std::mutex> lk(mu_);
    live_[p] = cls;
    in_use_ += cls;
}

void Allocator::trim() {
    std::lock_guard<std::mutex> lk(mu_);
    trim_locked();
}

void Allocator::trim_locked() {
    (void)hipDeviceSynchronize();
    for (auto& kv : free_)
        for (void* p : kv.second) (void)hipFree(p);
    free_.clear();
}

Allocator::~Allocator() {
    trim_locked();
    for (auto& kv : live_) (void)hipFree(kv.first);
}

void d2h_any(Ctx* c, void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    const bool small = c->readback && bytes <= Ctx::kReadbackSize;
    HIPC(hipMemcpyAsync(small ? c->readback : dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    if (small) memcpy(dst, c->readback, bytes);
    c->staging_off = 0;
}

void d2h(Ctx* c, void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    const ReadBack rb{dst, src, bytes};
    d2h_many(c, &rb, 1);
}

void d2h_many(Ctx* c, const ReadBack* rb, size_t count) {
    size_t total = 0;
    for (size_t i = 0; i < count; i++) total += (rb[i].bytes + 63) & ~(size_t)63;
    if (!c->readback || total > Ctx::kReadbackSize) {
        for (size_t i = 0; i < count; i++) d2h_any(c, rb[i].dst, rb[i].src, rb[i].bytes);
        return;
    }
    size_t off = 0;
    if (c->readback_dev && count <= (size_t)kToHostPieces) {
        ToHostPieces t{};
        for (size_t i = 0; i < count; i++) {
            if (rb[i].bytes) {
                t.src[t.count] = static_cast<const uint8_t*>(rb[i].src);
                t.off[t.count] = (uint32_t)off;
                t.bytes[t.count] = (uint32_t)rb[i].bytes;
                t.count++;
            }
            off += (rb[i].bytes + 63) & ~(size_t)63;
        }
        const uint32_t seq = ++c->readback_seq;
        uint32_t* flag = reinterpret_cast<uint32_t*>(c->readback + Ctx::kReadbackSize);
        launch_to_host(c, t, c->readback_dev, reinterpret_cast<uint32_t*>(c->readback_dev + Ctx::kReadbackSize), seq);
        // Poll the flag (a blocking stream wait sleeps once the work outlasts
        // its short spin, and wakes late).  After ~200 us the loop yields
        // between polls, leaving the cores to threads with host work; after
        // ~1 ms it also asks the runtime whether the stream has gone idle,
        // so that a failed launch is reported instead of waited on.
        const auto t0 = std::chrono::steady_clock::now();
        int phase = 0;  // 0 spin, 1 yield, 2 yield and query the stream
        for (uint32_t spins = 1;; spins++) {
            if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) == seq) break;
            if (phase) std::this_thread::yield();
            if ((spins & 255) != 0) continue;
            if (phase < 2) {
                const auto us = std::chrono::duration_cast<std::chrono::microseconds>(
                                    std::chrono::steady_clock::now() - t0).count();
                phase = us > 1000 ? 2 : us > 200 ? 1 : 0;
                continue;
            }
            const hipError_t q = hipStreamQuery(c->stream);
            if (q == hipErrorNotReady) continue;
            HIPC(q);  // idle: the flag is set by now
            if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) != seq) fail(DDJ_EHIP, "read-back flag not set");
            break;
        }
    } else {
        for (size_t i = 0; i < count; i++) {
            if (rb[i].bytes)
                HIPC(hipMemcpyAsync(c->readback + off, rb[i].src, rb[i].bytes, hipMemcpyDeviceToHost, c->stream));
            off += (rb[i].bytes + 63) & ~(size_t)63;
        }
        HIPC(hipStreamSynchronize(c->stream));
    }
    off = 0;
    for (size_t i = 0; i < count; i++) {
        if (rb[i].bytes) memcpy(rb[i].dst, c->readback + off, rb[i].bytes);
        off += (rb[i].bytes + 63) & ~(size_t)63;
    }
    c->staging_off = 0;
}

void read_small(Ctx* c, void* dst, const void* src, size_t bytes) {
    const uint8_t* s = static_cast<const uint8_t*>(src);
    for (const auto& sh : c->shadows) {
        if (s >= sh.dev && s + bytes <= sh.dev + sh.n) {
            memcpy(dst, sh.host + (s - sh.dev), bytes);
            return;
        }
    }
    d2h_any(c, dst, src, bytes);
}

void h2d(Ctx* c, void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    if (c->staging && bytes <= c->staging_size / 4) {
        const size_t need = (bytes + 63) & ~(size_t)63;
        if (c->staging_off + need > c->staging_size) {
            HIPC(hipStreamSynchronize(c->stream));  // every staged copy has landed
            c->staging_off = 0;
        }
        uint8_t* s = c->staging + c->staging_off;
        memcpy(s, src, bytes);
        HIPC(hipMemcpyAsync(dst, s, bytes, hipMemcpyHostToDevice, c->stream));
        c->staging_off += need;
        return;
    }
    HIPC(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPC(hipStreamSynchronize(c->stream));  // the source may be a temporary
    c->staging_off = 0;
}

const uint8_t* staging_slot(Ctx* c, size_t bytes) {
    if (!c->staging_dev || !bytes || bytes > c->staging_size / 4) return nullptr;
    const size_t need = (bytes + 63) & ~(size_t)63;
    if (c->staging_off + need > c->staging_size) {
        HIPC(hipStreamSynchronize(c->stream));  // every staged copy and table has been read
        c->staging_off = 0;
    }
    const uint8_t* d = c->staging_dev + c->staging_off;
    c->staging_off += need;
    return d;
}

void d2d(Ctx* c, void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    HIPC(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
}

DevBuf make_buf(Ctx* c, const void* host, size_t bytes) {
    DevBuf b(c, bytes);
    h2d(c, b.get(), host, bytes);
    return b;
}

}  // namespace ddj
