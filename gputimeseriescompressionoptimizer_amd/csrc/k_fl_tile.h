// The FL tile packer shared by the single-column and batched kernels.
#pragma once

#include "k_common.h"

namespace ddj {

constexpr int kFlTile = 4096;

// Packs tile `tile` (kFlTile values) of an FL stream of n values with b
// bits into out (32-bit words, out_words of them).  vals: LDS for
// kFlTile + 16 / sizeof(T) values.  DELTA: value i is x[i+1] - x[i]
// (wrapping), read straight from x (n + 1 rows).
template <typename T, bool DELTA>
__device__ __forceinline__ void fl_pack_tile(const T* __restrict__ x, size_t n, int b, uint32_t* __restrict__ out,
                                             size_t out_words, size_t tile, T* vals) {
    const size_t v0 = tile * kFlTile;
    const int nv = (int)((n - v0) < (size_t)kFlTile ? (n - v0) : (size_t)kFlTile);
    const T* src = x + v0;
    if (nv == kFlTile && (reinterpret_cast<uintptr_t>(src) & 15) == 0) {
        const u32x4* s4 = reinterpret_cast<const u32x4*>(src);
        u32x4* d4 = reinterpret_cast<u32x4*>(vals);
        constexpr int N4 = kFlTile * (int)sizeof(T) / 16;
#pragma unroll
        for (int i = threadIdx.x; i < N4; i += kBlock) d4[i] = __builtin_nontemporal_load(s4 + i);
        if (DELTA && threadIdx.x == 0) vals[kFlTile] = src[kFlTile];  // x has n + 1 rows
    } else {
        for (int i = threadIdx.x; i < kFlTile + (DELTA ? 1 : 0); i += kBlock) vals[i] = i < nv + (DELTA ? 1 : 0) ? src[i] : T(0);
    }
    __syncthreads();
    if (DELTA) {  // in place: vals[i] = vals[i+1] - vals[i]
        using U = typename UnsignedOf<T>::type;
        T d[kFlTile / kBlock];
#pragma unroll
        for (int k = 0; k < kFlTile / kBlock; k++) {
            const int i = threadIdx.x + k * kBlock;
            d[k] = i < nv ? (T)(U)((U)vals[i + 1] - (U)vals[i]) : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kFlTile / kBlock; k++) vals[threadIdx.x + k * kBlock] = d[k];
        __syncthreads();
    }
    const int tile_words = kFlTile / 32 * b;  // 128 * b
    const size_t w0 = tile * (size_t)tile_words;
    const uint64_t mask = lowmask64(b);
    // q / b as (q * ceil(2^32 / b)) >> 32: exact for every b <= 64 and
    // q < 2^18 + 64 (the error term stays below 1/b)
    const uint64_t rb = ((1ull << 32) + (uint64_t)b - 1) / (uint64_t)b;
    for (int j = threadIdx.x; j < tile_words; j += kBlock) {
        if (w0 + j >= out_words) break;
        const int bit0 = j * 32;
        const int va = (int)(((uint64_t)bit0 * rb) >> 32);
        int vb = (int)(((uint64_t)(bit0 + 31) * rb) >> 32);
        vb = vb < kFlTile - 1 ? vb : kFlTile - 1;
        uint64_t word = 0;
        for (int v = va; v <= vb; v++) {
            const uint64_t val = ubits(vals[v]) & mask;
            const int off = v * b - bit0;
            word |= off >= 0 ? (val << off) : (val >> (-off));
        }
        out[w0 + j] = (uint32_t)word;
    }
}

}  // namespace ddj
