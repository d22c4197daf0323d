// FL bit-packing (the reference's AFL encoding as it is actually called,
// run_afl_compress_gpu<T,1> / run_afl_decompress_gpu<T,1>,
// compression/afl/afl_gpu.cu:67-195).
//
// The reference stream is a contiguous LSB-first bit stream in words of T:
// value i occupies stream bits [i*b, (i+1)*b), stream bit p being bit p%W of
// word p/W.  On a little-endian machine that stream is the same byte sequence
// for every word width, so both kernels work on 32-bit words whatever T is;
// only the total length is rounded to whole T words.
//
// Pack: a workgroup owns a tile of 4096 values (a whole number of output
// words for any b), stages the values in LDS with 16-byte loads and emits
// the tile's 128*b output words, one word per thread-iteration, each word
// OR-ing the <= ceil(32/b)+1 values that overlap it.  Loads and stores are
// both fully coalesced.
//
// Unpack: a workgroup stages the tile's input in LDS from the aligned
// 16-byte chunks that hold it (an aligned chunk containing a valid byte never
// crosses a page, so the stream may start at any byte address) and every
// thread extracts values with a 64/96-bit funnel shift.
#include <type_traits>

#include "dispatch.h"
#include "k_common.h"
#include "kernels.h"
#include "k_fl_tile.h"

namespace ddj {

namespace {

template <typename T, bool DELTA = false>
__global__ __launch_bounds__(kBlock) void k_fl_pack(const T* __restrict__ x, size_t n, const MinMaxResult* mm,
                                                    int fixed_bits, uint8_t* __restrict__ out, size_t out_bytes) {
    __shared__ __attribute__((aligned(16))) T vals[kFlLdsVals];
    int b = fixed_bits;
    if (mm) b = bitlen_of<T>(from_order_key<T>(mm->min_key), from_order_key<T>(mm->max_key));
    fl_pack_tile<T, DELTA>(x, n, b, out, out_bytes, blockIdx.x, vals);
}

template <typename T, int MAXB>
__global__ __launch_bounds__(kBlock) void k_fl_unpack(const uint8_t* __restrict__ in, size_t in_bytes, size_t n,
                                                      int b, T* __restrict__ out, size_t vbase) {
    // rows [vbase, vbase + n) of the stream into out[0, n)
    __shared__ __attribute__((aligned(16))) uint32_t w[fl_stage_words<MAXB>()];
    const size_t v0 = (size_t)blockIdx.x * kFlTile;
    const int nv = (int)((n - v0) < (size_t)kFlTile ? (n - v0) : (size_t)kFlTile);
    const int a = stage_fl_words(in, in_bytes, vbase + v0, nv, b, w);
    __syncthreads();
    const uint64_t mask = lowmask64(b);
    constexpr int V = 16 / (int)sizeof(T);
    if (nv == kFlTile && (reinterpret_cast<uintptr_t>(out + v0) & 15) == 0) {
        // whole tile: each lane unpacks V consecutive values into one 16-byte
        // store (a cached store: RLE's values and lengths are read next)
        u32x4* o4 = reinterpret_cast<u32x4*>(out + v0);
        for (int i = threadIdx.x; i < kFlTile / V; i += kBlock) {
            T vals[V];
#pragma unroll
            for (int k = 0; k < V; k++) vals[k] = (T)fl_value(w, V * i + k, b, a, mask);
            u32x4 q;
            __builtin_memcpy(&q, vals, 16);
            o4[i] = q;
        }
    } else {
        // consecutive lanes take consecutive values: their LDS reads spread over the banks
        for (int k = threadIdx.x; k < nv; k += kBlock) out[v0 + k] = (T)fl_value(w, k, b, a, mask);
    }
}

// DeltaEncoding::Decode over AflEncoding::Decode without writing the deltas
// (delta_encoding.cu:74-102 on afl_encoding.cu:74-113): k_fl_delta_sums
// unpacks each tile of 4096 b-bit deltas and writes their sum, k_scan_tiles
// turns the sums into tile prefixes, and k_fl_unpack_delta_scan unpacks the
// tile again and prefix-sums it (wrapping in T) on top of the first row and
// its prefix.  The packed stream is b/8 bytes a row, so reading it twice
// costs far less than writing and re-reading the deltas.  out has m + 1 rows.
template <int MAXB>
__global__ __launch_bounds__(kBlock) void k_fl_delta_sums(const uint8_t* __restrict__ in, size_t in_bytes, size_t m,
                                                          int b, uint64_t* __restrict__ vals, size_t vbase) {
    // sums of the deltas [vbase, vbase + m), one per tile of 4096
    __shared__ __attribute__((aligned(16))) uint32_t w[fl_stage_words<MAXB>()];
    __shared__ uint64_t red[kBlock / kWave];
    const size_t v0 = (size_t)blockIdx.x * kFlTile;
    const int nv = (int)((m - v0) < (size_t)kFlTile ? (m - v0) : (size_t)kFlTile);
    const int a = stage_fl_words(in, in_bytes, vbase + v0, nv, b, w);
    __syncthreads();
    const uint64_t mask = lowmask64(b);
    uint64_t sum = 0;
    for (int k = threadIdx.x; k < nv; k += kBlock) sum += fl_value(w, k, b, a, mask);
    uint64_t total;
    block_exclusive_scan<uint64_t>(sum, &total, red);
    if (threadIdx.x == 0) vals[blockIdx.x] = total;
}

template <typename T, int MAXB>
__global__ __launch_bounds__(kBlock) void k_fl_unpack_delta_scan(const uint8_t* first_bytes, const uint8_t* __restrict__ in,
                                                                 size_t in_bytes, size_t m, int b, T* __restrict__ out,
                                                                 const uint64_t* __restrict__ prefix, size_t vbase) {
    // out[0] = first + prefix[0]; out[i + 1] = out[i] + d[vbase + i] for i < m
    // the staged words, then the tile's results
    __shared__ __attribute__((aligned(16))) uint32_t w[cmax(fl_stage_words<MAXB>(), kFlTile * (int)sizeof(T) / 4)];
    __shared__ uint64_t red[kBlock / kWave];
    constexpr int kPer = kFlTile / kBlock;  // 16 consecutive values per thread
    T first;
    __builtin_memcpy(&first, first_bytes, sizeof(T));
    const uint32_t tile = blockIdx.x;
    const size_t v0 = (size_t)tile * kFlTile;
    const int nv = (int)((m - v0) < (size_t)kFlTile ? (m - v0) : (size_t)kFlTile);
    const int a = stage_fl_words(in, in_bytes, vbase + v0, nv, b, w);
    __syncthreads();
    const uint64_t mask = lowmask64(b);
    uint64_t d[kPer], sum = 0;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        const int i = threadIdx.x * kPer + k;
        d[k] = i < nv ? fl_value(w, i, b, a, mask) : 0;
        sum += d[k];
    }
    uint64_t total;
    const uint64_t excl = block_exclusive_scan<uint64_t>(sum, &total, red);  // a barrier: w is free again
    T* res = reinterpret_cast<T*>(w);
    uint64_t run = ubits(first) + prefix[tile] + excl;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        run += d[k];
        res[threadIdx.x * kPer + k] = (T)run;
    }
    __syncthreads();
    // out[v0 + 1 ..]: single rows up to a 16-byte boundary, then 16-byte
    // stores from consecutive lanes, then single rows
    T* dst = out + v0 + 1;
    constexpr int V = 16 / (int)sizeof(T);
    int head = (int)(((16 - (reinterpret_cast<uintptr_t>(dst) & 15)) & 15) / sizeof(T));
    head = head < nv ? head : nv;
    int nq = (nv - head) / V;
    if ((reinterpret_cast<uintptr_t>(dst) % sizeof(T)) != 0) {  // not on a row boundary: rows only
        head = 0;
        nq = 0;
    }
    if ((int)threadIdx.x < head) dst[threadIdx.x] = res[threadIdx.x];
    u32x4* d4 = reinterpret_cast<u32x4*>(dst + head);
    for (int i = threadIdx.x; i < nq; i += kBlock) {
        T vals[V];
#pragma unroll
        for (int k = 0; k < V; k++) vals[k] = res[head + V * i + k];
        u32x4 q;
        __builtin_memcpy(&q, vals, 16);
        __builtin_nontemporal_store(q, d4 + i);
    }
    for (int i = head + V * nq + threadIdx.x; i < nv; i += kBlock) dst[i] = res[i];
    if (tile == 0 && threadIdx.x == 0) out[0] = (T)(ubits(first) + prefix[0]);
}

}  // namespace

namespace {
// Calls f(std::integral_constant<int, MAXB>) with the narrowest staging class
// (16, 32 or 64 bits) that holds b.
template <typename F> void with_maxb(int b, F&& f) {
    if (b <= 16)
        f(std::integral_constant<int, 16>{});
    else if (b <= 32)
        f(std::integral_constant<int, 32>{});
    else
        f(std::integral_constant<int, 64>{});
}
}  // namespace

void launch_fl_pack(Ctx* c, int dtype, const void* x, size_t n, const MinMaxResult* mm, int bits, void* out,
                    size_t out_bytes) {
    if (n == 0) return;
    const int esize = (int)dtype_size(dtype);
    dispatch_int(esize, [&](auto tag) {
        using T = decltype(tag);
        if (out_bytes == 0) {  // the widest b, rounded up to whole words
            const int maxb = mm ? 8 * (int)sizeof(T) : bits;
            const uint64_t W = 8 * sizeof(T);
            out_bytes = ((size_t)((maxb * (uint64_t)n + W - 1) / W) * sizeof(T) + 3) & ~size_t(3);
        }
        const size_t tiles = (n + kFlTile - 1) / kFlTile;
        { KernelTimer kt_(c, "k_fl_pack"); hipLaunchKernelGGL(k_fl_pack<T>, dim3((unsigned)tiles), dim3(kBlock), 0, c->stream, static_cast<const T*>(x),
                           n, mm, bits, static_cast<uint8_t*>(out), out_bytes); }
        HIPC(hipGetLastError());
    });
}

void launch_fl_pack_delta(Ctx* c, int elem_size, const void* x, size_t m, const MinMaxResult* mm, void* out, int bits,
                          size_t out_bytes) {
    if (m == 0) return;
    dispatch_int(elem_size, [&](auto tag) {
        using T = decltype(tag);
        if (out_bytes == 0) {  // the widest b, rounded up to whole words
            const uint64_t W = 8 * sizeof(T);
            out_bytes = ((size_t)((W * m + W - 1) / W) * sizeof(T) + 3) & ~size_t(3);
        }
        const size_t tiles = (m + kFlTile - 1) / kFlTile;
        { KernelTimer kt_(c, "k_fl_pack"); hipLaunchKernelGGL((k_fl_pack<T, true>), dim3((unsigned)tiles), dim3(kBlock), 0, c->stream,
                           static_cast<const T*>(x), m, mm, bits, static_cast<uint8_t*>(out), out_bytes); }
        HIPC(hipGetLastError());
    });
}

size_t fl_unpack_delta_scratch_words(size_t m) { return 2 * ((m + kFlTile - 1) / kFlTile) + 1; }

void launch_fl_unpack_delta(Ctx* c, int elem_size, const uint8_t* first, const uint8_t* in, size_t in_bytes, size_t m,
                           int bits, void* out, uint64_t* tile_scratch) {
    dispatch_int(elem_size, [&](auto tag) {
        using T = decltype(tag);
        if (bits <= 0 || bits > 8 * (int)sizeof(T)) fail(DDJ_ECORRUPT, "FL bit width out of range");
        if (m == 0) {
            d2d(c, out, first, sizeof(T));
            return;
        }
        const size_t tiles = (m + kFlTile - 1) / kFlTile;
        std::unique_ptr<TileScan> own_scan;
        uint64_t* vals = tile_scratch;
        if (!vals) {
            own_scan = std::make_unique<TileScan>(c, tiles);
            vals = own_scan->vals;
        }
        uint64_t* prefix = vals + tiles;
        kernel_bytes(c, "k_fl_delta_sums", (double)((m * (uint64_t)bits + 7) / 8));
        with_maxb(bits, [&](auto mb) {
            constexpr int MAXB = decltype(mb)::value;
            { KernelTimer kt_(c, "k_fl_delta_sums"); hipLaunchKernelGGL(k_fl_delta_sums<MAXB>, dim3((unsigned)tiles), dim3(kBlock), 0, c->stream,
                               in, in_bytes, m, bits, vals, (size_t)0); }
            HIPC(hipGetLastError());
            launch_scan_tiles(c, vals, tiles, prefix);
            { KernelTimer kt_(c, "k_fl_unpack_delta_scan"); hipLaunchKernelGGL((k_fl_unpack_delta_scan<T, MAXB>), dim3((unsigned)tiles), dim3(kBlock), 0, c->stream,
                               first, in, in_bytes, m, bits, static_cast<T*>(out), (const uint64_t*)prefix, (size_t)0); }
        });
        HIPC(hipGetLastError());
    });
}

void launch_fl_unpack(Ctx* c, int dtype, const uint8_t* in, size_t in_bytes, size_t n, int bits, void* out,
                      size_t first_row) {
    if (n == 0) return;
    const int esize = (int)dtype_size(dtype);
    dispatch_int(esize, [&](auto tag) {
        using T = decltype(tag);
        if (bits <= 0 || bits > 8 * (int)sizeof(T)) fail(DDJ_ECORRUPT, "FL bit width out of range");
        const size_t tiles = (n + kFlTile - 1) / kFlTile;
        with_maxb(bits, [&](auto mb) {
            constexpr int MAXB = decltype(mb)::value;
            KernelTimer kt_(c, "k_fl_unpack");
            hipLaunchKernelGGL((k_fl_unpack<T, MAXB>), dim3((unsigned)tiles), dim3(kBlock), 0, c->stream, in, in_bytes, n,
                               bits, static_cast<T*>(out), first_row);
        });
        HIPC(hipGetLastError());
    });
}

void launch_fl_unpack_delta_rows(Ctx* c, int elem_size, const uint8_t* first, const uint8_t* in, size_t in_bytes,
                                 int bits, size_t begin, size_t count, void* out) {
    if (count == 0) return;
    dispatch_int(elem_size, [&](auto tag) {
        using T = decltype(tag);
        if (bits <= 0 || bits > 8 * (int)sizeof(T)) fail(DDJ_ECORRUPT, "FL bit width out of range");
        // tile sums of the deltas before the range (part A) and inside it
        // (part B) in one array: one scan gives part B's prefixes on top of
        // the sum of everything before the range
        const size_t mb = count - 1;
        const size_t ta = (begin + kFlTile - 1) / kFlTile;
        const size_t tb = mb ? (mb + kFlTile - 1) / kFlTile : 1;
        TileScan ts(c, ta + tb);
        kernel_bytes(c, "k_fl_delta_sums", (double)(((begin + mb) * (uint64_t)bits + 7) / 8));
        with_maxb(bits, [&](auto mbits) {
            constexpr int MAXB = decltype(mbits)::value;
            if (ta) {
                KernelTimer kt_(c, "k_fl_delta_sums");
                hipLaunchKernelGGL(k_fl_delta_sums<MAXB>, dim3((unsigned)ta), dim3(kBlock), 0, c->stream, in, in_bytes,
                                   begin, bits, ts.vals, (size_t)0);
            }
            {
                KernelTimer kt_(c, "k_fl_delta_sums");
                hipLaunchKernelGGL(k_fl_delta_sums<MAXB>, dim3((unsigned)tb), dim3(kBlock), 0, c->stream, in, in_bytes,
                                   mb, bits, ts.vals + ta, begin);
            }
            HIPC(hipGetLastError());
            ts.scan(c);
            KernelTimer kt_(c, "k_fl_unpack_delta_scan");
            hipLaunchKernelGGL((k_fl_unpack_delta_scan<T, MAXB>), dim3((unsigned)tb), dim3(kBlock), 0, c->stream, first, in,
                               in_bytes, mb, bits, static_cast<T*>(out), (const uint64_t*)(ts.prefix + ta), begin);
        });
        HIPC(hipGetLastError());
    });
}

}  // namespace ddj
