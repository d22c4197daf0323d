// Batched compression of many columns whose trees end in FL: afl -> none
// and delta -> afl -> none on int32 / int64 columns.  Four launches cover
// the whole batch and contain no host round trip, so they are captured once
// into a hipGraph and replayed:
//
//   k_batch_init    reset every column's min/max keys
//   k_batch_minmax  persistent workgroups drain a static queue of
//                   (column, row range) items; each folds its range's
//                   min/max (of the values or of their differences) into
//                   the column with 64-bit atomics
//   k_batch_layout  one workgroup: bit widths, blob sizes, an exclusive scan
//                   of the sizes into arena offsets, and every node header
//                   and metadata byte of every blob
//   k_batch_pack    persistent workgroups drain the (column, FL tile) queue,
//                   packing straight into each blob's data region
//
// Blobs come out byte-identical to CompressionTree::Compress on each column.
#include <cstdint>

#include "dispatch.h"
#include "k_common.h"
#include "k_fl_tile.h"
#include "kernels.h"

namespace ddj {

namespace {

__device__ __forceinline__ uint64_t key_of(int64_t v, uint32_t es) {
    return es == 8 ? (uint64_t)v ^ (1ull << 63) : (uint64_t)(v + (1ll << 31));
}
__device__ __forceinline__ int64_t value_of(uint64_t k, uint32_t es) {
    return es == 8 ? (int64_t)(k ^ (1ull << 63)) : (int64_t)k - (1ll << 31);
}

__global__ __launch_bounds__(kBlock) void k_batch_init(BatchColumn* cols, int ncol) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < ncol; i += gridDim.x * kBlock) {
        cols[i].min_key = ~0ull;
        cols[i].max_key = 0;
    }
}

// Min/max (native compares) of the value stream rows [start, end) of one
// column: x itself, or with delta the differences x[i+1] - x[i].  Lanes take
// consecutive 16-byte vectors; with delta, the row after a vector is the
// next lane's first (the wave's last lane, and the range's last vector,
// load it).  start is a multiple of 16 / sizeof(T) (ranges are 2^k rows).
template <typename T>
__device__ __forceinline__ void range_minmax(const T* __restrict__ x, bool delta, uint64_t start, uint64_t end, T& mn,
                                             T& mx) {
    using U = typename UnsignedOf<T>::type;
    constexpr int V = 16 / (int)sizeof(T);
    const uint64_t v0 = start / V, v1 = end / V;
    const int lane = threadIdx.x & (kWave - 1);
    const u32x4* x4 = reinterpret_cast<const u32x4*>(x);
    for (uint64_t b = v0; b < v1; b += kBlock) {  // uniform trip count: every lane takes part in the shuffles
        const uint64_t i = b + threadIdx.x;
        const bool valid = i < v1;
        u32x4 w = valid ? x4[i] : u32x4{0u, 0u, 0u, 0u};
        T nx = T(0);
        if (delta && valid && (lane == kWave - 1 || i + 1 >= v1)) nx = x[(i + 1) * V];  // row (i+1)V <= end exists
        const uint32_t nb0 = (uint32_t)__shfl_down((int)w.x, 1, kWave);
        const uint32_t nb1 = (uint32_t)__shfl_down((int)w.y, 1, kWave);
        if (!valid) continue;
        T v[V + 1];
        __builtin_memcpy(v, &w, 16);
        if (delta) {
            if (lane == kWave - 1 || i + 1 >= v1) {
                v[V] = nx;
            } else {
                const uint64_t nb = ((uint64_t)nb1 << 32) | nb0;
                __builtin_memcpy(&v[V], &nb, sizeof(T));
            }
        }
#pragma unroll
        for (int k = 0; k < V; k++) {
            const T d = delta ? (T)(U)((U)v[k + 1] - (U)v[k]) : v[k];
            mn = d < mn ? d : mn;
            mx = d > mx ? d : mx;
        }
    }
    for (uint64_t r = v1 * V + threadIdx.x; r < end; r += kBlock) {
        const T d = delta ? (T)(U)((U)x[r + 1] - (U)x[r]) : x[r];
        mn = d < mn ? d : mn;
        mx = d > mx ? d : mx;
    }
}

__global__ __launch_bounds__(kBlock) void k_batch_minmax(BatchColumn* cols, const BatchItem* items, int nitems) {
    __shared__ uint64_t smn[kBlock / kWave], smx[kBlock / kWave];
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {  // uniform per workgroup
        const BatchItem item = items[it];
        const BatchColumn c = cols[item.col];
        uint64_t mn = ~0ull, mx = 0;
        if (c.es == 8) {
            int64_t a = INT64_MAX, z = INT64_MIN;
            range_minmax<int64_t>(reinterpret_cast<const int64_t*>(c.x), c.delta, item.start, item.end, a, z);
            if (a <= z) {
                mn = key_of(a, 8);
                mx = key_of(z, 8);
            }
        } else {
            int32_t a = INT32_MAX, z = INT32_MIN;
            range_minmax<int32_t>(reinterpret_cast<const int32_t*>(c.x), c.delta, item.start, item.end, a, z);
            if (a <= z) {
                mn = key_of(a, 4);
                mx = key_of(z, 4);
            }
        }
        mn = wave_min(mn);
        mx = wave_max(mx);
        const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
        if (lane == 0) {
            smn[wid] = mn;
            smx[wid] = mx;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < kBlock / kWave; w++) {
                mn = smn[w] < mn ? smn[w] : mn;
                mx = smx[w] > mx ? smx[w] : mx;
            }
            atomicMin(reinterpret_cast<unsigned long long*>(&cols[item.col].min_key), (unsigned long long)mn);
            atomicMax(reinterpret_cast<unsigned long long*>(&cols[item.col].max_key), (unsigned long long)mx);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void put16(uint8_t* p, int16_t v) { __builtin_memcpy(p, &v, 2); }
__device__ __forceinline__ void put32(uint8_t* p, int32_t v) { __builtin_memcpy(p, &v, 4); }
__device__ __forceinline__ void put64(uint8_t* p, uint64_t v) { __builtin_memcpy(p, &v, 8); }
__device__ __forceinline__ void header(uint8_t* p, int enc, int dtype, int metalen) {  // encoding_metadata_header.hpp:13-17
    put16(p, (int16_t)enc);
    put16(p + 2, (int16_t)dtype);
    put32(p + 4, metalen);
}

// One workgroup.  Blob of afl -> none: [hdr afl][b, rest, 0, 0][hdr none]
// [size_t data bytes][data]; delta -> afl -> none puts [hdr delta][x[0]] first.
__global__ __launch_bounds__(kBlock) void k_batch_layout(BatchColumn* cols, int ncol, uint8_t* arena) {
    __shared__ uint64_t red[kBlock / kWave];
    __shared__ uint64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int c0 = 0; c0 < ncol; c0 += kBlock) {
        const int i = c0 + threadIdx.x;
        uint64_t slot = 0;
        BatchColumn c;
        int b = 0;
        uint64_t m = 0, data_bytes = 0, head = 0;
        if (i < ncol) {
            c = cols[i];
            const int64_t mn = value_of(c.min_key, c.es), mx = value_of(c.max_key, c.es);
            // BITLEN (core/cuda_macros.cuh:168-175)
            if (mn < 0) {
                b = 8 * (int)c.es;
            } else {
                uint64_t v = (uint64_t)mx;
                b = 1;
                while (v >>= 1) b++;
            }
            m = c.delta ? c.n - 1 : c.n;
            const uint64_t W = 8 * c.es;
            data_bytes = (b * m + W - 1) / W * c.es;
            head = (c.delta ? 8 + c.es : 0) + 8 + 4 + 8 + 8;
            slot = (head + data_bytes + 15) & ~15ull;  // blobs start 16-byte aligned
        }
        uint64_t total;
        const uint64_t excl = block_exclusive_scan<uint64_t>(slot, &total, red);
        if (i < ncol) {
            const uint64_t off = carry + excl;
            uint8_t* p = arena + off;
            if (c.delta) {
                header(p, DDJ_ENC_DELTA, (int)c.dtype, (int)c.es);
                for (uint32_t k = 0; k < c.es; k++) p[8 + k] = c.x[k];  // DeltaEncoding metadata: the first value
                p += 8 + c.es;
            }
            header(p, DDJ_ENC_AFL, (int)c.dtype, 4);
            p[8] = (uint8_t)b;
            p[9] = (uint8_t)((data_bytes * 8 - m * (uint64_t)b) & 0xff);
            p[10] = 0;
            p[11] = 0;
            header(p + 12, DDJ_ENC_NONE, (int)c.dtype, 8);
            put64(p + 20, data_bytes);
            cols[i].bits = (uint32_t)b;
            cols[i].out_off = off;
            cols[i].blob_bytes = head + data_bytes;
            cols[i].data_off = off + head;
            cols[i].data_bytes = data_bytes;
        }
        __syncthreads();
        if (threadIdx.x == 0) carry += total;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_batch_pack(const BatchColumn* cols, const BatchItem* tiles, int ntiles,
                                                       uint8_t* arena) {
    __shared__ __attribute__((aligned(16))) int64_t vals[kFlLdsVals];
    for (int it = blockIdx.x; it < ntiles; it += gridDim.x) {  // uniform per workgroup
        const BatchItem item = tiles[it];
        const BatchColumn c = cols[item.col];
        uint8_t* out = arena + c.data_off;
        const size_t out_bytes = c.data_bytes;
        const uint64_t m = c.delta ? c.n - 1 : c.n;
        const int b = (int)c.bits;
        if (c.es == 8) {
            const int64_t* x = reinterpret_cast<const int64_t*>(c.x);
            if (c.delta)
                fl_pack_tile<int64_t, true>(x, m, b, out, out_bytes, item.start, vals);
            else
                fl_pack_tile<int64_t, false>(x, m, b, out, out_bytes, item.start, vals);
        } else {
            const int32_t* x = reinterpret_cast<const int32_t*>(c.x);
            int32_t* v32 = reinterpret_cast<int32_t*>(vals);
            if (c.delta)
                fl_pack_tile<int32_t, true>(x, m, b, out, out_bytes, item.start, v32);
            else
                fl_pack_tile<int32_t, false>(x, m, b, out, out_bytes, item.start, v32);
        }
        __syncthreads();  // vals is reused by the next item
    }
}

}  // namespace

void launch_batch_fl(Ctx* c, BatchColumn* d_cols, int ncol, const BatchItem* d_ranges, int nranges,
                     const BatchItem* d_tiles, int ntiles, uint8_t* arena, unsigned wg_per_cu) {
    const unsigned persistent = (unsigned)c->cu_count * wg_per_cu;
    hipLaunchKernelGGL(k_batch_init, dim3((ncol + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, d_cols, ncol);
    hipLaunchKernelGGL(k_batch_minmax, dim3(std::min<unsigned>(persistent, (unsigned)nranges)), dim3(kBlock), 0,
                       c->stream, d_cols, d_ranges, nranges);
    hipLaunchKernelGGL(k_batch_layout, dim3(1), dim3(kBlock), 0, c->stream, d_cols, ncol, arena);
    hipLaunchKernelGGL(k_batch_pack, dim3(std::min<unsigned>(persistent, (unsigned)ntiles)), dim3(kBlock), 0,
                       c->stream, d_cols, d_tiles, ntiles, arena);
    HIPC(hipGetLastError());
}

}  // namespace ddj
