/*
 * fl_cpu.cpp -- CPU baseline: the reference's FL host loops timed on the host.
 *
 * TEST / MEASUREMENT INFRASTRUCTURE ONLY (bench.py's cpu_baseline leg and the
 * oracle tests load it; the product library never does).
 *
 * run_afl_compress_cpu<T,1> / run_afl_decompress_cpu<T,1>
 * (compression/afl/afl_gpu.cu:6-65) are the reference's only CPU path: one
 * host loop over "threads" t, each packing values [t*W, t*W+W) into words
 * [t*b, t*b+b).  The groups are
 * independent, so the all-cores variant hands each host thread a contiguous
 * range of groups and the stream is identical.  Built twice by the Makefile
 * (-O3 -march=x86-64-v3 and -march=x86-64-v4); the caller loads the widest
 * one the host supports.
 */
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <thread>
#include <type_traits>
#include <vector>

namespace {

/* The loop bodies with the reference's host bit macros (GETNBITS /
 * GETNPBITS = shift and mask, core/cuda_macros.cuh:6-7, :101-160) in T's own
 * width, so the compiler sees what the reference's host build sees.  b = W is
 * the raw copy the device's bfe gives (the host macro's shift by W is
 * undefined there). */
template <typename T> struct Loops {
    typedef typename std::make_unsigned<T>::type U;
    static constexpr unsigned W = 8 * sizeof(T);
    static inline U mask(unsigned n) { return n >= W ? (U)~(U)0 : (U)(((U)1 << n) - 1); }

    /* afl_compress_base_gpu<T,1> (afl_gpu.cu:124-163) */
    static inline void pack_group(unsigned bits, unsigned long data_id, unsigned long comp_id, const T* data, T* out,
                                  unsigned long length, unsigned long comprLength) {
        U value = 0;
        unsigned v1_pos = 0, v1_len;
        unsigned long pos = comp_id, pos_data = data_id;
        for (unsigned i = 0; i < W && pos_data < length; ++i) {
            const U v1 = (U)data[pos_data];
            pos_data += 1;
            if (v1_pos >= W - bits && pos < comprLength) {
                v1_len = W - v1_pos;
                value = (U)(value | (U)((v1 & mask(v1_len)) << v1_pos));
                out[pos] = (T)value;
                v1_pos = bits - v1_len;
                value = (U)((v1 >> v1_len) & mask(v1_pos));
                pos += 1;
            } else {
                value = (U)(value | (U)((v1 & mask(bits)) << v1_pos));
                v1_pos += bits;
            }
        }
        if (pos_data >= length && pos_data < length + 1 && pos < comprLength) out[pos] = (T)value;
    }

    /* afl_decompress_base_gpu<T,1> (afl_gpu.cu:165-195); the one word past
     * the stream's end that the reference may read is taken as 0. */
    static inline void unpack_group(unsigned bits, unsigned long comp_id, unsigned long data_id, const T* in, T* data,
                                    unsigned long length, unsigned long words) {
        unsigned long pos = comp_id, pos_decomp = data_id;
        unsigned v1_pos = 0, v1_len;
        if (pos_decomp >= length) return;
        U v1 = pos < words ? (U)in[pos] : 0, ret;
        for (unsigned i = 0; i < W && pos_decomp < length; ++i) {
            if (v1_pos > W - bits) {
                v1_len = W - v1_pos;
                ret = (U)((v1 >> (v1_pos & (W - 1))) & mask(v1_len));
                pos += 1;
                v1 = pos < words ? (U)in[pos] : 0;
                v1_pos = bits - v1_len;
                ret = (U)(ret | (U)((v1 & mask(v1_pos)) << v1_len));
            } else {
                ret = (U)((v1 >> v1_pos) & mask(bits));
                v1_pos += bits;
            }
            data[pos_decomp] = (T)ret;
            pos_decomp += 1;
        }
    }
};

template <typename T> void pack_range(unsigned bits, const T* in, T* out, unsigned long n, unsigned long words,
                                      unsigned long g0, unsigned long g1) {
    const unsigned long W = 8 * sizeof(T);
    if (bits == W) {
        const unsigned long a = g0 * W, b = std::min(g1 * W, n);
        if (b > a) memcpy(out + a, in + a, (b - a) * sizeof(T));
        return;
    }
    for (unsigned long t = g0; t < g1; t++) Loops<T>::pack_group(bits, t * W, t * bits, in, out, n, words);
}

template <typename T> void unpack_range(unsigned bits, const T* in, T* out, unsigned long n, unsigned long words,
                                        unsigned long g0, unsigned long g1) {
    const unsigned long W = 8 * sizeof(T);
    if (bits == W) {
        const unsigned long a = g0 * W, b = std::min(g1 * W, n);
        if (b > a) memcpy(out + a, in + a, (b - a) * sizeof(T));
        return;
    }
    for (unsigned long t = g0; t < g1; t++) Loops<T>::unpack_group(bits, t * bits, t * W, in, out, n, words);
}

/* A sense-reversing spin barrier: the pool's threads live for the whole
 * measurement, so a timed repetition costs no thread start-up. */
struct Barrier {
    std::atomic<int> waiting{0};
    std::atomic<int> phase{0};
    int n;
    explicit Barrier(int k) : n(k) {}
    void wait() {
        const int ph = phase.load(std::memory_order_acquire);
        if (waiting.fetch_add(1, std::memory_order_acq_rel) == n - 1) {
            waiting.store(0, std::memory_order_relaxed);
            phase.store(ph + 1, std::memory_order_release);
        } else {
            while (phase.load(std::memory_order_acquire) == ph) std::this_thread::yield();
        }
    }
};

template <typename T>
int bench_typed(int bits, const void* in_, size_t n, void* packed_, size_t words, void* out_, int threads, int reps,
                double* pack_s, double* unpack_s) {
    const T* in = (const T*)in_;
    T* packed = (T*)packed_;
    T* out = (T*)out_;
    const unsigned long W = 8 * sizeof(T);
    const unsigned long groups = (n + W - 1) / W;
    const int total = reps + 2; /* two untimed warm-ups */
    std::vector<double> tp(total), tu(total);
    Barrier bar(threads);
    auto worker = [&](int id) {
        const unsigned long g0 = groups * id / threads, g1 = groups * (id + 1) / threads;
        for (int r = 0; r < total; r++) {
            bar.wait();
            auto t0 = std::chrono::steady_clock::now();
            pack_range<T>(bits, in, packed, n, words, g0, g1);
            bar.wait();
            auto t1 = std::chrono::steady_clock::now();
            unpack_range<T>(bits, packed, out, n, words, g0, g1);
            bar.wait();
            auto t2 = std::chrono::steady_clock::now();
            if (id == 0) {
                tp[r] = std::chrono::duration<double>(t1 - t0).count();
                tu[r] = std::chrono::duration<double>(t2 - t1).count();
            }
        }
    };
    std::vector<std::thread> pool;
    for (int i = 1; i < threads; i++) pool.emplace_back(worker, i);
    worker(0);
    for (auto& t : pool) t.join();
    auto median = [&](std::vector<double> v) {
        v.erase(v.begin(), v.begin() + 2);
        std::sort(v.begin(), v.end());
        return v[v.size() / 2];
    };
    *pack_s = median(tp);
    *unpack_s = median(tu);
    return 0;
}

} // namespace

extern "C" {

unsigned flcpu_hardware_concurrency() { return std::thread::hardware_concurrency(); }

/* Packs `in` (n values of width_bytes) to `bits`-bit fields in `packed`
 * (words = ceil(n*bits/W) words) and unpacks them into `out`, `reps` timed
 * times after two warm-ups, on `threads` host threads; returns the median
 * seconds of each direction. */
int flcpu_bench(int width_bytes, int bits, const void* in, size_t n, void* packed, size_t words, void* out,
                int threads, int reps, double* pack_s, double* unpack_s) {
    if (threads < 1 || reps < 1 || bits < 1 || bits > 8 * width_bytes) return -1;
    switch (width_bytes) {
    case 4: return bench_typed<int32_t>(bits, in, n, packed, words, out, threads, reps, pack_s, unpack_s);
    case 8: return bench_typed<int64_t>(bits, in, n, packed, words, out, threads, reps, pack_s, unpack_s);
    case 2: return bench_typed<int16_t>(bits, in, n, packed, words, out, threads, reps, pack_s, unpack_s);
    case 1: return bench_typed<int8_t>(bits, in, n, packed, words, out, threads, reps, pack_s, unpack_s);
    default: return -1;
    }
}
}
