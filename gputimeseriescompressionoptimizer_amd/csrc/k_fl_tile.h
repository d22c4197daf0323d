// The FL tile packer shared by the single-column and batched kernels, and
// the staging of an FL stream's row range in LDS shared by the unpack and
// merge kernels.
#pragma once

#include "k_common.h"

namespace ddj {

constexpr int kFlTile = 4096;
constexpr int kFlGroups = kFlTile / 32;  // 32 values per packing thread
// LDS the packer needs, in values of T: the tile with one pad slot per 32
// values (so that the packing threads' reads fall in distinct banks), the
// row after the tile, and slack.  The packed words reuse the same bytes.
constexpr int kFlLdsVals = kFlTile + kFlGroups + 4;
static_assert(kBlock >= kFlGroups, "one packing thread per 32-value group");

__device__ __forceinline__ int fl_slot(int v) { return v + (v >> 5); }

// Packs tile `tile` (kFlTile values) of an FL stream of n values with b
// bits (b <= 8 * sizeof(T)) into the stream's bytes at out (any byte
// address), writing no byte at or past out_bytes.  vals: LDS for kFlLdsVals
// values, 16-byte aligned.  DELTA: value i is x[i+1] - x[i] (wrapping), read
// straight from x (n + 1 rows).
//
// The tile is staged in LDS with coalesced 16-byte loads; thread g < 128
// then takes values [32g, 32g + 32) into registers and packs them into
// exactly b words (32 * b bits), LSB first, which land at words [g*b,
// g*b + b) of the tile; the tile's 128 * b words go out as 16-byte stores,
// realigned to the destination's byte offset, with single bytes at either
// end (so that a stream can start and end inside a word it shares with its
// neighbours in a blob).
template <typename T, bool DELTA>
__device__ __forceinline__ void fl_pack_tile(const T* __restrict__ x, size_t n, int b, uint8_t* __restrict__ out,
                                             size_t out_bytes, size_t tile, T* vals) {
    using U = typename UnsignedOf<T>::type;
    const size_t v0 = tile * kFlTile;
    const int nv = (int)((n - v0) < (size_t)kFlTile ? (n - v0) : (size_t)kFlTile);
    const T* src = x + v0;
    if (nv == kFlTile && (reinterpret_cast<uintptr_t>(src) & 15) == 0) {
        constexpr int V = 16 / (int)sizeof(T);  // values per 16-byte load, all in one 32-value group
        constexpr int N4 = kFlTile / V;
        static_assert(N4 % kBlock == 0, "whole loads per thread");
        const u32x4* s4 = reinterpret_cast<const u32x4*>(src);
        u32x4 q[N4 / kBlock];
#pragma unroll
        for (int r = 0; r < N4 / kBlock; r++) q[r] = s4[threadIdx.x + r * kBlock];
#pragma unroll
        for (int r = 0; r < N4 / kBlock; r++) {
            T t[V];
            __builtin_memcpy(t, &q[r], 16);
            const int p = fl_slot((threadIdx.x + r * kBlock) * V);
#pragma unroll
            for (int k = 0; k < V; k++) vals[p + k] = t[k];
        }
        if (DELTA && threadIdx.x == 0) vals[fl_slot(kFlTile)] = src[kFlTile];  // x has n + 1 rows
    } else {
        for (int i = threadIdx.x; i < kFlTile + (DELTA ? 1 : 0); i += kBlock)
            vals[fl_slot(i)] = i < nv + (DELTA ? 1 : 0) ? src[i] : T(0);
    }
    __syncthreads();
    const int g = threadIdx.x;
    U v[33];
    if (g < kFlGroups) {
#pragma unroll
        for (int k = 0; k < 32; k++) v[k] = (U)vals[33 * g + k];
        if (DELTA) v[32] = (U)vals[33 * g + 33];  // fl_slot(32 * (g + 1))
    }
    __syncthreads();  // the packed words overwrite vals
    uint32_t* words = reinterpret_cast<uint32_t*>(vals);
    if (g < kFlGroups) {
        const uint64_t mask = lowmask64(b);
        uint32_t* w = words + g * b;
        uint64_t acc = 0;  // pending bits, LSB first
        int nb = 0, k = 0;
        // every thread has the same b, so the branches are uniform
        auto push = [&](uint64_t bits, int len) {  // len <= 32
            acc |= bits << nb;
            nb += len;
            if (nb >= 32) {
                w[k++] = (uint32_t)acc;
                acc >>= 32;
                nb -= 32;
            }
        };
        const int lim = nv - 32 * g;  // values of the group inside the stream
        if constexpr (sizeof(T) <= 4) {
            // b <= 32: a 32-bit accumulator; nb depends on b and i only, so
            // it is uniform and the shifts take scalar amounts
            const uint32_t m32 = (uint32_t)mask;
            uint32_t a32 = 0;
            int nb32 = 0;
#pragma unroll
            for (int i = 0; i < 32; i++) {
                const U d = DELTA ? (i < lim ? (U)(v[i + 1] - v[i]) : U(0)) : v[i];
                const uint32_t val = (uint32_t)d & m32;
                a32 |= val << nb32;
                const int nn = nb32 + b;
                if (nn >= 32) {
                    w[k++] = a32;
                    a32 = nb32 ? (val >> (32 - nb32)) : 0u;  // the bits that did not fit
                    nb32 = nn - 32;
                } else {
                    nb32 = nn;
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < 32; i++) {
                // past the stream's end the deltas, like the values, are zero
                const U d = DELTA ? (i < lim ? (U)(v[i + 1] - v[i]) : U(0)) : v[i];
                const uint64_t val = (uint64_t)d & mask;
                if (b <= 32) {
                    push(val, b);
                } else {
                    push(val & 0xffffffffull, 32);
                    push(val >> 32, b - 32);
                }
            }
        }
    }
    __syncthreads();
    const int tile_bytes = 4 * kFlGroups * b;
    const size_t B0 = tile * (size_t)tile_bytes;
    if (B0 >= out_bytes) return;
    const int L = (int)((out_bytes - B0) < (size_t)tile_bytes ? (out_bytes - B0) : (size_t)tile_bytes);
    uint8_t* d = out + B0;
    const uint8_t* wb = reinterpret_cast<const uint8_t*>(words);
    int head = (int)((16 - (reinterpret_cast<uintptr_t>(d) & 15)) & 15);  // bytes before a 16-byte boundary
    head = head < L ? head : L;
    if ((int)threadIdx.x < head) d[threadIdx.x] = wb[threadIdx.x];
    const int r = head & 3;  // the same for every word of the body
    auto src_word = [&](int s) -> uint32_t {  // tile bytes [s, s + 4), s + 4 <= L
        const int q = s >> 2;
        if (r == 0) return words[q];
        return (uint32_t)((((uint64_t)words[q + 1] << 32) | words[q]) >> (8 * r));
    };
    const int nq = (L - head) / 16;
    u32x4* d4 = reinterpret_cast<u32x4*>(d + head);
    for (int i = threadIdx.x; i < nq; i += kBlock) {
        const int s = head + 16 * i;
        __builtin_nontemporal_store(u32x4{src_word(s), src_word(s + 4), src_word(s + 8), src_word(s + 12)}, d4 + i);
    }
    for (int j = head + 16 * nq + threadIdx.x; j < L; j += kBlock) d[j] = wb[j];
}

// Staged words of one tile at b <= MAXB bits: at most kFlTile * MAXB bits,
// plus the bytes before the stream start in its first 16-byte chunk and one
// zero chunk after it.  The unpack kernels come in MAXB = 16, 32, 64, so that
// narrow streams stage in less LDS and more workgroups fit on a CU.
template <int MAXB> constexpr int fl_stage_words() { return kFlTile * MAXB / 32 + 16; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }

// Bytes [lo, hi) of a 32-bit word kept, the others cleared.
__device__ __forceinline__ uint32_t byte_mask(int lo, int hi) {
    lo = lo < 0 ? 0 : (lo > 4 ? 4 : lo);
    hi = hi < 0 ? 0 : (hi > 4 ? 4 : hi);
    if (hi <= lo) return 0u;
    const uint32_t upto = hi == 4 ? ~0u : ((1u << (8 * hi)) - 1u);
    const uint32_t from = lo == 4 ? 0u : ~((1u << (8 * lo)) - 1u);
    return upto & from;
}

// Stages the b-bit values [v0, v0 + nv) of the stream `in` (any byte
// address; v0 any row, so the first value may start inside a byte) in LDS
// as 32-bit words, from the 16-byte chunks that hold them (an aligned
// 16-byte chunk holding a valid byte never crosses a page), four chunks in
// flight per thread.  Returns the bit offset of the first value inside w[0];
// bytes outside the stream are zero, and one zero chunk follows the last
// staged one.
__device__ __forceinline__ int stage_fl_words(const uint8_t* __restrict__ in, size_t in_bytes, size_t v0, int nv, int b,
                                              uint32_t* w) {
    const size_t B0 = v0 * (size_t)b / 8;
    const int s0 = (int)((v0 * (size_t)b) & 7);  // 0 whenever v0 is a multiple of 8
    const size_t need = ((size_t)s0 + (size_t)nv * b + 7) / 8;
    const size_t avail_total = in_bytes > B0 ? in_bytes - B0 : 0;
    const int avail = (int)(need < avail_total ? need : avail_total);
    const uint8_t* base = in + B0;
    const int a = (int)(reinterpret_cast<uintptr_t>(base) & 15);
    const u32x4* wp4 = reinterpret_cast<const u32x4*>(base - a);
    const int n4_valid = (a + avail + 15) / 16;
    const int n4_fill = n4_valid + 1;  // <= fl_stage_words<MAXB>() / 4 for b <= MAXB
    u32x4* w4 = reinterpret_cast<u32x4*>(w);
    constexpr int U = 4;
    for (int i0 = threadIdx.x; i0 < n4_fill; i0 += U * kBlock) {
        u32x4 v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = i0 + u * kBlock;
            v[u] = i < n4_valid ? wp4[i] : u32x4{0u, 0u, 0u, 0u};
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = i0 + u * kBlock;
            if (i < n4_fill) {
                const int b0 = 16 * i;  // first byte of the chunk, relative to wp4
                u32x4 m = v[u];
                if (b0 < a || b0 + 16 > a + avail) {  // only the chunks at either end hold bytes outside the stream
                    m.x &= byte_mask(a - b0, a + avail - b0);
                    m.y &= byte_mask(a - b0 - 4, a + avail - b0 - 4);
                    m.z &= byte_mask(a - b0 - 8, a + avail - b0 - 8);
                    m.w &= byte_mask(a - b0 - 12, a + avail - b0 - 12);
                }
                w4[i] = m;
            }
        }
    }
    return 8 * a + s0;
}

// Value k (0-based in the tile) from staged words whose first value starts
// at bit `abit` of w[0].
__device__ __forceinline__ uint64_t fl_value(const uint32_t* w, int k, int b, int abit, uint64_t mask) {
    const uint64_t p = (uint64_t)k * b + abit;
    const uint32_t q = (uint32_t)(p >> 5), s = (uint32_t)(p & 31);
    uint64_t val = ((uint64_t)w[q] | ((uint64_t)w[q + 1] << 32)) >> s;
    if (s + b > 64) val |= (uint64_t)w[q + 2] << (64 - s);
    return val & mask;
}

}  // namespace ddj
