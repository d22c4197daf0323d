// Random access into FL streams, and the interleaved FL layout.
//
// Random access (afl_decompress_value_gpu / afl_decompress_base_value_gpu,
// reference compression/afl/afl_gpu.cu:113-121, 197-223): any value of a
// b-bit stream is found from its row alone, so a set of rows is fetched with
// one or two aligned 32-bit loads per row and no decode of the rest.
//
// Interleaved layout (run_afl_compress_gpu<T,32> / run_afl_decompress_gpu<T,32>,
// afl_gpu.cu:67-111, the CWARP_SIZE = 32 instantiation the reference
// benchmarks in afl_encoding_benchmark.cpp:44): the column is cut into blocks
// of 32 lanes x W values (W = bits of T); lane l of block k owns the values
// k*32*W + l + 32*i (i < W) and packs them LSB first into b words of T, its
// word j stored at k*32*b + l + 32*j.  A thread is a lane: its W loads and b
// stores are each 32 consecutive words across the lanes, so both sides
// coalesce without staging.  The last block is stored whole, its missing
// values packed as zero.
#include "dispatch.h"
#include "k_common.h"
#include "kernels.h"

namespace ddj {

namespace {

// Bits [pos, pos + len) (len <= 64) of the little-endian LSB-first bit
// stream at `in` (any byte address), from the aligned 32-bit words that hold
// them: only words holding a wanted bit are read.
__device__ __forceinline__ uint64_t stream_bits(const uint8_t* in, uint64_t pos, int len) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(in + (pos >> 3));
    const uint32_t* w = reinterpret_cast<const uint32_t*>(a & ~uintptr_t(3));
    const int sh = 8 * (int)(a & 3) + (int)(pos & 7);  // < 32
    const int nw = (sh + len + 31) >> 5;                // 1..3 words
    uint64_t lo = w[0];
    if (nw > 1) lo |= (uint64_t)w[1] << 32;
    uint64_t v = lo >> sh;
    if (nw > 2) v |= (uint64_t)w[2] << (64 - sh);  // nw == 3 needs sh > 0
    return v & lowmask64((unsigned)len);
}

template <typename T, int LANES>
__global__ __launch_bounds__(kBlock) void k_fl_fetch(const uint8_t* __restrict__ in, size_t n, int b,
                                                     const uint64_t* __restrict__ rows, size_t count,
                                                     T* __restrict__ out, uint32_t* err) {
    constexpr int W = 8 * sizeof(T);
    bool bad = false;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += (size_t)gridDim.x * kBlock) {
        const uint64_t r = rows[i];
        uint64_t v = 0;
        if (r >= n) {
            bad = true;
        } else if (LANES == 1) {
            v = stream_bits(in, r * (uint64_t)b, b);
        } else {
            const uint64_t blk = r / (32 * W), rin = r % (32 * W);
            const uint64_t lane = rin % 32, k = rin / 32;
            const uint64_t bit = k * (uint64_t)b, j = bit / W;
            const int o = (int)(bit % W);
            const uint64_t q = blk * 32 * (uint64_t)b + lane + 32 * j;  // word holding the value's first bit
            if (o + b <= W) {
                v = stream_bits(in, q * W + o, b);
            } else {  // the rest is in the lane's next word, 32 words on
                const int lo = W - o;
                v = stream_bits(in, q * W + o, lo) | (stream_bits(in, (q + 32) * W, b - lo) << lo);
            }
        }
        out[i] = (T)v;
    }
    if (bad) __hip_atomic_store(err, 3u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fl_pack_lanes(const T* __restrict__ x, size_t n, int b,
                                                          T* __restrict__ out, size_t lanes_total) {
    using U = typename UnsignedOf<T>::type;
    constexpr int W = 8 * sizeof(T);
    const size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= lanes_total) return;
    const size_t blk = t >> 5;
    const size_t r0 = blk * 32 * W + (t & 31);
    U* o = reinterpret_cast<U*>(out) + blk * 32 * (size_t)b + (t & 31);
    const uint64_t mask = lowmask64((unsigned)b);
    uint64_t acc = 0;  // pending bits, LSB first
    int fill = 0;      // < W
#pragma unroll
    for (int i = 0; i < W; i++) {
        const size_t r = r0 + 32 * (size_t)i;
        const uint64_t v = r < n ? ((uint64_t)(U)x[r] & mask) : 0;
        acc |= v << fill;
        fill += b;
        if (fill >= W) {
            *o = (U)acc;
            o += 32;
            fill -= W;
            if constexpr (W == 64)
                acc = fill ? v >> (b - fill) : 0;
            else
                acc >>= W;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fl_unpack_lanes(const T* __restrict__ in, size_t n, int b,
                                                            T* __restrict__ out, size_t lanes_total) {
    using U = typename UnsignedOf<T>::type;
    constexpr int W = 8 * sizeof(T);
    const size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= lanes_total) return;
    const size_t blk = t >> 5;
    const size_t r0 = blk * 32 * W + (t & 31);
    const U* p = reinterpret_cast<const U*>(in) + blk * 32 * (size_t)b + (t & 31);
    const uint64_t mask = lowmask64((unsigned)b);
    uint64_t cur = (uint64_t)*p;
    int pos = 0;  // bits of cur consumed
#pragma unroll
    for (int i = 0; i < W; i++) {
        uint64_t v;
        if (pos + b <= W) {
            v = (cur >> pos) & mask;
            pos += b;
        } else {  // 0 < pos: the value continues in the lane's next word
            const int lo = W - pos;
            v = cur >> pos;
            p += 32;
            cur = (uint64_t)*p;
            v = (v | (cur << lo)) & mask;
            pos = b - lo;
        }
        if (pos == W && i + 1 < W) {  // the lane's b words are used up exactly at i = W - 1
            p += 32;
            cur = (uint64_t)*p;
            pos = 0;
        }
        const size_t r = r0 + 32 * (size_t)i;
        if (r < n) out[r] = (T)(U)v;
    }
}

unsigned grid_for(size_t threads) { return (unsigned)((threads + kBlock - 1) / kBlock); }

}  // namespace

size_t fl_lanes_bytes(int dtype, int bits, size_t n) {
    const size_t es = dtype_size(dtype), W = 8 * es;
    const size_t blocks = (n + 32 * W - 1) / (32 * W);
    return blocks * 32 * (size_t)bits * es;
}

void launch_fl_fetch(Ctx* c, int dtype, int bits, int lanes, const uint8_t* in, size_t in_bytes, size_t n,
                     const uint64_t* rows, size_t count, void* out) {
    if (count == 0) return;
    const size_t es = dtype_size(dtype);
    if (bits <= 0 || bits > (int)(8 * es)) fail(DDJ_EINVAL, "FL bit width out of range");
    if (lanes != 1 && lanes != 32) fail(DDJ_EINVAL, "FL layout has 1 or 32 lanes");
    const size_t need = lanes == 1 ? (size_t)(((unsigned __int128)n * bits + 7) / 8) : fl_lanes_bytes(dtype, bits, n);
    if (in_bytes < need) fail(DDJ_EINVAL, "stream shorter than n values of b bits");
    arm_error_flag(c);
    uint32_t* err = &c->control->error_flag;
    const unsigned grid = std::min<unsigned>(grid_for(count), (unsigned)c->cu_count * 16);
    kernel_bytes(c, "k_fl_fetch", (double)count * (8.0 + 2.0 * es));
    dispatch_int(es, [&](auto tag) {
        using T = decltype(tag);
        KernelTimer kt_(c, "k_fl_fetch");
        if (lanes == 1)
            hipLaunchKernelGGL((k_fl_fetch<T, 1>), dim3(grid), dim3(kBlock), 0, c->stream, in, n, bits, rows, count,
                               static_cast<T*>(out), err);
        else
            hipLaunchKernelGGL((k_fl_fetch<T, 32>), dim3(grid), dim3(kBlock), 0, c->stream, in, n, bits, rows, count,
                               static_cast<T*>(out), err);
    });
    HIPC(hipGetLastError());
}

void launch_fl_pack_lanes(Ctx* c, int dtype, int bits, const void* x, size_t n, void* out) {
    if (n == 0) return;
    const size_t es = dtype_size(dtype), W = 8 * es;
    if (bits <= 0 || bits > (int)W) fail(DDJ_EINVAL, "FL bit width out of range");
    const size_t lanes_total = (n + 32 * W - 1) / (32 * W) * 32;
    kernel_bytes(c, "k_fl_pack_lanes", (double)n * es + (double)fl_lanes_bytes(dtype, bits, n));
    dispatch_int(es, [&](auto tag) {
        using T = decltype(tag);
        KernelTimer kt_(c, "k_fl_pack_lanes");
        hipLaunchKernelGGL(k_fl_pack_lanes<T>, dim3(grid_for(lanes_total)), dim3(kBlock), 0, c->stream,
                           static_cast<const T*>(x), n, bits, static_cast<T*>(out), lanes_total);
    });
    HIPC(hipGetLastError());
}

void launch_fl_unpack_lanes(Ctx* c, int dtype, int bits, const void* in, size_t n, void* out) {
    if (n == 0) return;
    const size_t es = dtype_size(dtype), W = 8 * es;
    if (bits <= 0 || bits > (int)W) fail(DDJ_EINVAL, "FL bit width out of range");
    const size_t lanes_total = (n + 32 * W - 1) / (32 * W) * 32;
    kernel_bytes(c, "k_fl_unpack_lanes", (double)n * es + (double)fl_lanes_bytes(dtype, bits, n));
    dispatch_int(es, [&](auto tag) {
        using T = decltype(tag);
        KernelTimer kt_(c, "k_fl_unpack_lanes");
        hipLaunchKernelGGL(k_fl_unpack_lanes<T>, dim3(grid_for(lanes_total)), dim3(kBlock), 0, c->stream,
                           static_cast<const T*>(in), n, bits, static_cast<T*>(out), lanes_total);
    });
    HIPC(hipGetLastError());
}

}  // namespace ddj
