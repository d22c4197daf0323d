// Variable-length outputs: stable split by a predicate with the packed
// stencil (patch / const / dict), its inverse merge, RLE encode/decode, dict
// code packing.  Each output pass takes its tile's offset from a counting
// pass and k_scan_tiles (reduce, then scan): the counting passes read the
// input once more (split, RLE encode) or only the stencil / lengths (merge,
// RLE decode), which costs less on MI355X than chaining 16K tiles through a
// look-back.
//
// Tiles hold 4096 rows; thread t of a workgroup owns the 16 consecutive rows
// [16t, 16t+16) of the tile, i.e. exactly two stencil bytes.  Inputs are
// staged through LDS with striped (coalesced) loads, outputs are compacted in
// LDS and written back striped.
#include <vector>
#include <limits>
#include <type_traits>

#include "dispatch.h"
#include "k_common.h"
#include "kernels.h"
#include "k_fl_tile.h"

namespace ddj {

namespace {

constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;
// LDS index of tile row i: one pad slot every 16 rows spreads the blocked
// (16 consecutive rows per thread) accesses over the banks.
constexpr int kTilePad = kTile + kTile / 16;
__device__ __forceinline__ int P(int i) { return i + (i >> 4); }

// The split predicate.  Kernels take its kind as a template argument K (the
// switch folds away, so the per-row test is one or a few compares).
template <typename T> struct Pred {
    int kind;
    T a, b;
    T set[4];
    int set_n;
    uint32_t active;  // bit j: set[j] is in use (j < set_n)
    template <int K> __device__ __forceinline__ bool eval(T v, uint32_t& code) const {
        code = 0;
        switch (K) {
        case PRED_LOWER: return v < a;                // LowerOperator (util/stencil/stencil_operators.hpp)
        case PRED_INSIDE: return !(v > b || v < a);   // OutsideOperator: true inside [low, high]
        case PRED_NOT_EQUAL: return v != a;           // _constEncodeStencilKernel (const_encoding.cu:19-25)
        default: {                                    // getMostFrequentStencilKernel (dict_encoding.cu:15-34)
            // the dictionary holds at most 4 keys: bit j of m is a match of
            // key j, and the code ORs every matching index, as
            // _compressUniqueKernel does (indices 1, 2, 3 contribute 1, 2, 3).
            // Integers, not per-key masks, so that 16 rows x 4 keys of
            // comparisons never sit in scalar registers at once.
            const uint32_t m = ((v == set[0]) ? 1u : 0u) | ((v == set[1]) ? 2u : 0u) |
                               ((v == set[2]) ? 4u : 0u) | ((v == set[3]) ? 8u : 0u);
            const uint32_t mm = m & active;
            code = ((mm >> 1) & 3u) | ((mm & 8u) ? 3u : 0u);
            return mm != 0;
        }
        }
    }
};

template <typename T> T bits_to(uint64_t b) {
    T v;
    using U = typename UnsignedOf<T>::type;
    U u = (U)b;
    memcpy(&v, &u, sizeof(T));
    return v;
}

// Rows of each tile that satisfy the predicate (the split's counting pass).
// MM (integer types): also the order keys of the minimum and maximum of the
// rows going to each side, keys[4 t .. 4 t + 3] = {yes min, yes max, no min,
// no max}, so that an FL encoder of either output needs no pass of its own.
template <typename T, bool MM, int K>
__global__ __launch_bounds__(kBlock) void k_split_count(const T* __restrict__ x, size_t n, Pred<T> p,
                                                        uint64_t* __restrict__ vals, uint64_t* __restrict__ keys) {
    __shared__ uint64_t red[kBlock / kWave];
    __shared__ unsigned long long mm[4];
    const size_t base = (size_t)blockIdx.x * kTile;
    const int cnt = (int)((n - base) < (size_t)kTile ? (n - base) : (size_t)kTile);
    uint32_t c = 0, code;
    // yes side: ylo / yhi, no side: nlo / nhi
    T ylo = std::numeric_limits<T>::max(), yhi = std::numeric_limits<T>::min();
    T nlo = ylo, nhi = yhi;
    if constexpr (MM) {
        if (threadIdx.x < 4) mm[threadIdx.x] = (threadIdx.x & 1) ? 0ull : ~0ull;
        __syncthreads();  // before any thread folds into mm
    }
    auto take = [&](T v) {
        const bool f = p.template eval<K>(v, code);
        c += f ? 1u : 0u;
        if constexpr (MM) {
            const T a = f ? v : ylo, b = f ? v : yhi, d = f ? nlo : v, e = f ? nhi : v;
            ylo = a < ylo ? a : ylo;
            yhi = b > yhi ? b : yhi;
            nlo = d < nlo ? d : nlo;
            nhi = e > nhi ? e : nhi;
        }
    };
    if (cnt == kTile && (reinterpret_cast<uintptr_t>(x + base) & 15) == 0) {
        constexpr int V = 16 / sizeof(T);
        const u32x4* s4 = reinterpret_cast<const u32x4*>(x + base);
#pragma unroll
        for (int q = 0; q < kTile / V / kBlock; q++) {
            const u32x4 w = s4[threadIdx.x + kBlock * q];
            T v[V];
            __builtin_memcpy(v, &w, 16);
#pragma unroll
            for (int k = 0; k < V; k++) take(v[k]);
        }
    } else {
        for (int i = threadIdx.x; i < cnt; i += kBlock) take(x[base + i]);
    }
    if constexpr (MM) {  // wave extremes, then one LDS atomic per wave and key
        uint64_t k4[4];
        if constexpr (sizeof(T) <= 4) {
            k4[0] = order_key((T)wave_extreme_i32<false>((int32_t)ylo));
            k4[1] = order_key((T)wave_extreme_i32<true>((int32_t)yhi));
            k4[2] = order_key((T)wave_extreme_i32<false>((int32_t)nlo));
            k4[3] = order_key((T)wave_extreme_i32<true>((int32_t)nhi));
        } else {
            k4[0] = wave_min(order_key(ylo));
            k4[1] = wave_max(order_key(yhi));
            k4[2] = wave_min(order_key(nlo));
            k4[3] = wave_max(order_key(nhi));
        }
        if ((threadIdx.x & (kWave - 1)) == 0) {
            atomicMin(&mm[0], (unsigned long long)k4[0]);
            atomicMax(&mm[1], (unsigned long long)k4[1]);
            atomicMin(&mm[2], (unsigned long long)k4[2]);
            atomicMax(&mm[3], (unsigned long long)k4[3]);
        }
    }
    uint64_t total;
    block_exclusive_scan<uint64_t>((uint64_t)c, &total, red);  // a barrier: mm is complete
    if (threadIdx.x == 0) vals[blockIdx.x] = total;
    if (MM && threadIdx.x < 4) keys[4 * (size_t)blockIdx.x + threadIdx.x] = mm[threadIdx.x];
}


template <typename T, bool CODES, bool WRITE_NO, int K>
__global__ __launch_bounds__(kBlock) void k_split(const T* __restrict__ x, size_t n, Pred<T> p,
                                                  uint8_t* __restrict__ stencil, T* __restrict__ yes,
                                                  uint8_t* __restrict__ no_bytes, uint8_t* __restrict__ codes,
                                                  uint32_t* __restrict__ edges, const uint64_t* __restrict__ prefix,
                                                  uint64_t* d_count) {
    constexpr int kCodeWords = kTile / 16 + 2;  // 2-bit code words a tile's codes can touch
    __shared__ T buf[kTilePad];
    __shared__ uint32_t cwl[CODES ? kCodeWords : 1];  // those words, aligned to the output's word grid
    __shared__ uint64_t red[kBlock / kWave];
    const uint32_t tile = blockIdx.x;
    const size_t base = (size_t)tile * kTile;
    const int cnt = (int)((n - base) < (size_t)kTile ? (n - base) : (size_t)kTile);
    const T padv = x[n - 1];  // stencil padding repeats the last row (reference stencil.cu:26-27)
    const uint64_t Y0 = prefix[tile];  // loaded with the tile
    if (CODES)
        for (int q = threadIdx.x; q < kCodeWords; q += kBlock) cwl[q] = 0;
    T v[kItems];
    if (cnt == kTile && (reinterpret_cast<uintptr_t>(x + base) & 15) == 0) {  // uniform
        // a whole tile: each thread's 16 rows straight into registers
        constexpr int V = 16 / sizeof(T);
        const u32x4* s4 = reinterpret_cast<const u32x4*>(x + base + (size_t)threadIdx.x * kItems);
#pragma unroll
        for (int q = 0; q < kItems / V; q++) {
            const u32x4 w = s4[q];
            __builtin_memcpy(&v[q * V], &w, 16);
        }
    } else {
        for (int i = threadIdx.x; i < kTile; i += kBlock) buf[P(i)] = i < cnt ? x[base + i] : padv;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kItems; k++) v[k] = buf[P(threadIdx.x * kItems + k)];
    }
    uint32_t mask = 0;
    uint32_t cchunk = 0;  // CODES: the thread's codes, 2 bits each, in row order
    int c = 0;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        uint32_t code;
        bool f = p.template eval<K>(v[k], code);
        mask |= (uint32_t)f << k;
        if (threadIdx.x * kItems + k < cnt && f) {
            if (CODES) cchunk |= (code & 3u) << (2 * c);
            c++;
        }
    }
    const size_t g0 = base + (size_t)threadIdx.x * kItems;
    if (g0 < n) stencil[1 + g0 / 8] = (uint8_t)(mask & 0xff);
    if (g0 + 8 < n) stencil[2 + g0 / 8] = (uint8_t)(mask >> 8);
    if (tile == 0 && threadIdx.x == 0) stencil[0] = (uint8_t)(n % 8);
    uint64_t total;
    const uint64_t excl = block_exclusive_scan<uint64_t>((uint64_t)c, &total, red);
    // codes j in [Y0, Y0 + total) go to field j % 16 of 2-bit word j / 16
    // (_compressUniqueKernel, unique_encoding.cu:48-76): tile-local code i
    // lands in cwl[(lead + i) / 16]
    const int lead = (int)(Y0 % 16);
    if (CODES && c > 0) {  // the chunk covers code bits [2 (lead + excl), +2c): at most two words
        const int bit = 2 * (lead + (int)excl), sh = bit & 31;
        atomicOr(&cwl[bit >> 5], cchunk << sh);
        if (sh && sh + 2 * c > 32) atomicOr(&cwl[(bit >> 5) + 1], cchunk >> (32 - sh));
    }
    // compaction without branches (with codes, only the no rows are kept);
    // the guard for rows past the column's end is needed on its last tile only
    const int nbase = CODES ? 0 : (int)total;
    int yr = (int)excl, nr = nbase + threadIdx.x * kItems - (int)excl;
    auto place = [&](int k, int f) {
        if (CODES) {
            if (!f) buf[P(nr)] = v[k];
        } else {
            buf[P(f ? yr : nr)] = v[k];
        }
    };
    if (cnt == kTile) {  // uniform
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            const int f = (mask >> k) & 1;
            place(k, f);
            yr += f;
            nr += 1 - f;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            const int f = (mask >> k) & 1;
            if (threadIdx.x * kItems + k < cnt) place(k, f);
            yr += f;
            nr += 1 - f;
        }
    }
    __syncthreads();
    if (CODES) {
        // words shared with a neighbouring tile (the first and the last, when
        // partly filled) go to edges[2 tile], edges[2 tile + 1]: k_code_edges
        // assembles them (device-scope atomics on the words cost more)
        uint32_t* cw = reinterpret_cast<uint32_t*>(codes);
        const uint64_t w_first = Y0 / 16, w_end = (Y0 + total + 15) / 16;
        const int nt = (int)total, nw = (int)(w_end - w_first);
        for (int q = threadIdx.x; q < nw; q += kBlock) {
            const int i0 = 16 * q - lead;  // tile-local index of the word's field 0
            if (i0 >= 0 && i0 + 16 <= nt)
                cw[w_first + q] = cwl[q];
            else
                edges[2 * (size_t)tile + (q == 0 ? 0 : 1)] = cwl[q];
        }
    } else if (yes) {
        for (int i = threadIdx.x; i < (int)total; i += kBlock) yes[Y0 + i] = buf[P(i)];
    }
    if (WRITE_NO) {
        const size_t N0 = base - Y0;
        const int nn = cnt - (int)total;
        if ((reinterpret_cast<uintptr_t>(no_bytes) % sizeof(T)) == 0) {
            T* no = reinterpret_cast<T*>(no_bytes);
            for (int i = threadIdx.x; i < nn; i += kBlock) no[N0 + i] = buf[P(nbase + i)];
        } else {
            // rows stored unchanged into a blob at any byte address: the
            // tile's bytes go out as 16-byte stores realigned to that address,
            // with single bytes at either end
            auto row_bits = [&](int e) -> uint64_t { return e < nn ? ubits(buf[P(nbase + e)]) : 0ull; };  // bit pattern
            auto word = [&](int w) -> uint32_t {  // 32-bit word w of the tile's row bytes
                if constexpr (sizeof(T) == 8) {
                    const uint64_t v = row_bits(w >> 1);
                    return (w & 1) ? (uint32_t)(v >> 32) : (uint32_t)v;
                } else if constexpr (sizeof(T) == 4) {
                    return (uint32_t)row_bits(w);
                } else if constexpr (sizeof(T) == 2) {
                    return (uint32_t)row_bits(2 * w) | ((uint32_t)row_bits(2 * w + 1) << 16);
                } else {
                    return (uint32_t)row_bits(4 * w) | ((uint32_t)row_bits(4 * w + 1) << 8) |
                           ((uint32_t)row_bits(4 * w + 2) << 16) | ((uint32_t)row_bits(4 * w + 3) << 24);
                }
            };
            uint8_t* d = no_bytes + N0 * sizeof(T);
            const int L = nn * (int)sizeof(T);
            int head = (int)((16 - (reinterpret_cast<uintptr_t>(d) & 15)) & 15);
            head = head < L ? head : L;
            if ((int)threadIdx.x < head) d[threadIdx.x] = (uint8_t)(word(threadIdx.x >> 2) >> (8 * (threadIdx.x & 3)));
            const uint32_t r = (uint32_t)(head & 3);
            const int nq = (L - head) / 16;
            u32x4* d4 = reinterpret_cast<u32x4*>(d + head);
            for (int i = threadIdx.x; i < nq; i += kBlock) {
                const int q = (head + 16 * i) >> 2;  // first word of the chunk's bytes
                const uint32_t w0 = word(q), w1 = word(q + 1), w2 = word(q + 2), w3 = word(q + 3), w4 = word(q + 4);
                __builtin_nontemporal_store(u32x4{__builtin_amdgcn_alignbyte(w1, w0, r), __builtin_amdgcn_alignbyte(w2, w1, r),
                                                  __builtin_amdgcn_alignbyte(w3, w2, r), __builtin_amdgcn_alignbyte(w4, w3, r)},
                                            d4 + i);
            }
            for (int j = head + 16 * nq + threadIdx.x; j < L; j += kBlock) d[j] = (uint8_t)(word(j >> 2) >> (8 * (j & 3)));
        }
    }
    if (threadIdx.x == 0 && base + cnt == n) d_count[0] = Y0 + total;
}

// The split of a patch node whose outputs go to FL leaves (patch -> (afl ->
// none), (afl -> none)), with the FL encoding fused in: the stencil and the
// compaction of k_split, then each side's rows are packed at their bit
// offset in their FL stream (a side's row r at bits [r b, r b + b)), so that
// neither side is ever written unpacked.  Thread g < 128 packs yes rows
// [32g, 32g + 32) of the tile, thread 128 + g the no rows, into LDS words
// already at the stream's bit offset (a group starts at bit (first row * b)
// mod 32, so the LDS words are the stream's words); the word a group shares
// with the next goes through a carry slot, ORed in after a barrier.  Words
// the tile fills alone are stored at any byte address (16-byte stores
// realigned to the destination, single bytes at the ends); the first and
// last word of a side that the tile shares with its neighbours go to
// edges[4 t + 2 side + 0/1], assembled by k_fl_edges.
template <typename T, int K>
__global__ __launch_bounds__(kBlock) void k_split_fl(const T* __restrict__ x, size_t n, Pred<T> p,
                                                     uint8_t* __restrict__ stencil, FlSide yside, FlSide nside,
                                                     const uint64_t* __restrict__ prefix, uint64_t* d_count) {
    using U = typename UnsignedOf<T>::type;
    static_assert(sizeof(T) * kTilePad >= 4 * (kTile / 32 + 2) * 8 * sizeof(T), "packed words fit the tile buffer");
    __shared__ __attribute__((aligned(16))) T buf[kTilePad];
    __shared__ uint64_t red[kBlock / kWave];
    const uint32_t tile = blockIdx.x;
    const size_t base = (size_t)tile * kTile;
    const int cnt = (int)((n - base) < (size_t)kTile ? (n - base) : (size_t)kTile);
    const T padv = x[n - 1];
    const uint64_t Y0 = prefix[tile];
    T v[kItems];
    if (cnt == kTile && (reinterpret_cast<uintptr_t>(x + base) & 15) == 0) {  // uniform
        // a whole tile: each thread's 16 rows straight into registers
        constexpr int V = 16 / sizeof(T);
        const u32x4* s4 = reinterpret_cast<const u32x4*>(x + base + (size_t)threadIdx.x * kItems);
#pragma unroll
        for (int q = 0; q < kItems / V; q++) {
            const u32x4 w = s4[q];
            __builtin_memcpy(&v[q * V], &w, 16);
        }
    } else {
        for (int i = threadIdx.x; i < kTile; i += kBlock) buf[P(i)] = i < cnt ? x[base + i] : padv;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kItems; k++) v[k] = buf[P(threadIdx.x * kItems + k)];
    }
    uint32_t mask = 0;
    int c = 0;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        uint32_t code;
        const bool f = p.template eval<K>(v[k], code);
        mask |= (uint32_t)f << k;
        if (threadIdx.x * kItems + k < cnt && f) c++;
    }
    const size_t g0 = base + (size_t)threadIdx.x * kItems;
    if (g0 < n) stencil[1 + g0 / 8] = (uint8_t)(mask & 0xff);
    if (g0 + 8 < n) stencil[2 + g0 / 8] = (uint8_t)(mask >> 8);
    if (tile == 0 && threadIdx.x == 0) stencil[0] = (uint8_t)(n % 8);
    uint64_t total;
    const uint64_t excl = block_exclusive_scan<uint64_t>((uint64_t)c, &total, red);
    // compaction without branches: each row's slot is a select, and the
    // guard for rows past the column's end is needed on its last tile only
    int yr = (int)excl, nr = (int)total + threadIdx.x * kItems - (int)excl;
    if (cnt == kTile) {  // uniform
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            const int f = (mask >> k) & 1;
            buf[P(f ? yr : nr)] = v[k];
            yr += f;
            nr += 1 - f;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            const int f = (mask >> k) & 1;
            if (threadIdx.x * kItems + k < cnt) buf[P(f ? yr : nr)] = v[k];
            yr += f;
            nr += 1 - f;
        }
    }
    __syncthreads();
    const int cy = (int)total, cn = cnt - cy;
    // pack: thread g < 128 the yes group g, thread 128 + g the no group g
    static_assert((kBlock / 2) % kWave == 0, "each wave packs one side");
    const bool is_yes = threadIdx.x < kBlock / 2;  // uniform per wave
    const int grp = threadIdx.x & (kBlock / 2 - 1);
    const int cs = is_yes ? cy : cn;
    // width and start bit are the same for every group of a side, so
    // scalar: the pack's word boundaries are then uniform branches
    const int b = __builtin_amdgcn_readfirstlane(is_yes ? yside.bits : nside.bits);
    U w32[32];
    const bool packs = 32 * grp < cs;
    if (packs) {
        const int sb = is_yes ? 0 : cy;
#pragma unroll
        for (int k = 0; k < 32; k++) {
            const int i = 32 * grp + k;
            w32[k] = i < cs ? (U)buf[P(sb + i)] : U(0);
        }
    }
    __syncthreads();  // the packed words overwrite buf
    // Each side's words are laid out on the stream's word grid: local word q
    // is stream word wf + q, the side's first row starting at bit o of word
    // 0.  Group g then starts at bit o of its first word, whose low o bits
    // are the previous group's last bits (carry[side][g - 1]).
    __shared__ uint32_t carry[2][kTile / 32];
    uint32_t* words = reinterpret_cast<uint32_t*>(buf);
    const int wy = (cy + 31) / 32 * yside.bits + 1;  // yes words (and the partial word after them), then the no words
    const int o_me = __builtin_amdgcn_readfirstlane((int)(((is_yes ? Y0 : (uint64_t)(base - Y0)) * (uint64_t)b) & 31));
    if (packs) {
        uint32_t* w = words + (is_yes ? 0 : wy) + grp * b;
        const uint64_t m = lowmask64((unsigned)b);
        uint64_t acc = 0;
        int nb = o_me, k = 0;
        auto push = [&](uint64_t bits, int len) {  // len <= 32
            acc |= bits << nb;
            nb += len;
            if (nb >= 32) {
                w[k++] = (uint32_t)acc;
                acc >>= 32;
                nb -= 32;
            }
        };
        if constexpr (sizeof(T) <= 4) {
            // b <= 32: a 32-bit accumulator whose bit count is uniform
            const uint32_t m32 = (uint32_t)m;
            uint32_t a32 = 0;
            int nb32 = o_me;
#pragma unroll
            for (int i = 0; i < 32; i++) {
                const uint32_t val = (uint32_t)w32[i] & m32;
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
            carry[is_yes ? 0 : 1][grp] = a32;  // the group's last o bits
        } else {
#pragma unroll
            for (int i = 0; i < 32; i++) {
                const uint64_t val = (uint64_t)w32[i] & m;
                if (b <= 32) {
                    push(val, b);
                } else {
                    push(val & 0xffffffffull, 32);
                    push(val >> 32, b - 32);
                }
            }
            carry[is_yes ? 0 : 1][grp] = (uint32_t)acc;  // the group's last o bits
        }
    }
    __syncthreads();
    if (packs) {  // each group's first word takes the previous group's last bits
        uint32_t* w = words + (is_yes ? 0 : wy) + grp * b;
        if (grp > 0) w[0] |= carry[is_yes ? 0 : 1][grp - 1];
        if (32 * (grp + 1) >= cs) w[b] = carry[is_yes ? 0 : 1][grp];  // the last group: the word after it
    }
    __syncthreads();
#pragma unroll
    for (int side = 0; side < 2; side++) {
        const FlSide fs = side == 0 ? yside : nside;
        const int ns = side == 0 ? cy : cn;
        if (ns == 0) continue;  // uniform
        const uint64_t R0 = side == 0 ? Y0 : (uint64_t)(base - Y0);
        const int bb = fs.bits;
        const uint64_t bit0 = R0 * (uint64_t)bb;
        const uint32_t o = (uint32_t)(bit0 & 31);
        const uint64_t wf = bit0 >> 5;
        const uint64_t L = (uint64_t)ns * bb;
        const int nw = (int)((o + L + 31) >> 5);
        const uint32_t* lw = words + (side == 0 ? 0 : wy);
        auto outw = [&](int q) -> uint32_t { return lw[q]; };  // stream word wf + q as this tile fills it
        const int q_lo = o ? 1 : 0;
        const int q_hi = (int)((o + L) >> 5);  // words this tile fills to their end
        if (threadIdx.x == 0 && (o || q_hi == 0)) fs.edges[4 * (size_t)tile + 2 * side] = outw(0);
        if (threadIdx.x == 1 && nw > 1 && q_hi < nw) fs.edges[4 * (size_t)tile + 2 * side + 1] = outw(nw - 1);
        if (q_hi <= q_lo) continue;
        // stream bytes [4 (wf + q_lo), 4 (wf + q_hi)), clipped to the stream
        uint64_t B0 = 4 * (wf + (uint64_t)q_lo), B1 = 4 * (wf + (uint64_t)q_hi);
        if (B1 > fs.bytes) B1 = fs.bytes;
        if (B1 <= B0) continue;
        uint8_t* d = fs.dst + B0;
        const int Lb = (int)(B1 - B0);
        auto byte_at = [&](int j) -> uint8_t { return (uint8_t)(outw(q_lo + (j >> 2)) >> (8 * (j & 3))); };
        int head = (int)((16 - (reinterpret_cast<uintptr_t>(d) & 15)) & 15);
        head = head < Lb ? head : Lb;
        if ((int)threadIdx.x < head) d[threadIdx.x] = byte_at(threadIdx.x);
        const uint32_t r = (uint32_t)(head & 3);
        const int nq = (Lb - head) / 16;
        u32x4* d4 = reinterpret_cast<u32x4*>(d + head);
        for (int i = threadIdx.x; i < nq; i += kBlock) {
            const int q = q_lo + ((head + 16 * i) >> 2);
            const uint32_t a0 = outw(q), a1 = outw(q + 1), a2 = outw(q + 2), a3 = outw(q + 3), a4 = outw(q + 4);
            if (r)
                __builtin_nontemporal_store(u32x4{__builtin_amdgcn_alignbyte(a1, a0, r), __builtin_amdgcn_alignbyte(a2, a1, r),
                                                  __builtin_amdgcn_alignbyte(a3, a2, r), __builtin_amdgcn_alignbyte(a4, a3, r)},
                                            d4 + i);
            else
                __builtin_nontemporal_store(u32x4{a0, a1, a2, a3}, d4 + i);
        }
        for (int j = head + 16 * nq + threadIdx.x; j < Lb; j += kBlock) d[j] = byte_at(j);
    }
    if (base + cnt == n && threadIdx.x < 2) {
        // A 64-bit stream ends on an 8-byte boundary: the word after the
        // side's last bit (if any) holds no row, so no tile writes it.
        const FlSide fs = threadIdx.x == 0 ? yside : nside;
        const uint64_t rows = threadIdx.x == 0 ? Y0 + total : n - (Y0 + total);
        for (uint64_t j = 4 * ((rows * (uint64_t)fs.bits + 31) >> 5); j < fs.bytes; j++) fs.dst[j] = 0;
        if (threadIdx.x == 0) d_count[0] = Y0 + total;
    }
}

// Stream words of a fused patch split that several tiles fill: the tile
// holding a word's first bit assembles it from its own last-word edge and
// the first-word edges of the following tiles whose rows reach into it, and
// writes its bytes (those inside the stream).  Side 0 = yes rows (tile t
// starts at row prefix[t]), side 1 = no rows (t * 4096 - prefix[t]).
__global__ __launch_bounds__(kBlock) void k_fl_edges(const uint64_t* __restrict__ prefix, size_t tiles, size_t n,
                                                     FlSide yside, FlSide nside) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= 2 * tiles) return;
    const size_t t = i >> 1;
    const int side = (int)(i & 1);
    const FlSide fs = side ? nside : yside;
    auto rows0 = [&](size_t u) -> uint64_t { return side ? (uint64_t)(u * kTile) - prefix[u] : prefix[u]; };
    auto rows_of = [&](size_t u) -> uint64_t {
        const uint64_t c = (n - u * kTile) < (uint64_t)kTile ? (n - u * kTile) : (uint64_t)kTile;
        const uint64_t y = prefix[u + 1] - prefix[u];
        return side ? c - y : y;
    };
    const uint64_t ns = rows_of(t);
    if (ns == 0) return;
    const uint64_t b = (uint64_t)fs.bits;
    const uint64_t first = rows0(t) * b, end = (rows0(t) + ns) * b;
    const uint64_t wf = first >> 5, wl = (end - 1) >> 5;
    if (32 * wl < first || end >= 32 * wl + 32) return;  // the word's bit 0 lies earlier, or the tile fills it
    uint32_t w = fs.edges[4 * t + 2 * side + (wl == wf ? 0 : 1)];
    for (size_t u = t + 1; u < tiles; u++) {
        if (rows0(u) * b >= 32 * wl + 32) break;
        if (rows_of(u)) w |= fs.edges[4 * u + 2 * side];  // its first word is this one
    }
    for (int j = 0; j < 4; j++)
        if (4 * wl + j < fs.bytes) fs.dst[4 * wl + j] = (uint8_t)(w >> (8 * j));
}

// Code words that several tiles fill: tile t's codes are [prefix[t],
// prefix[t+1]); a word whose field 0 lies in tile t and which t leaves
// partly empty is assembled by t's thread from t's contribution and the
// first-word contributions of the following tiles that reach into it.
// Every other word was stored whole by k_split.
__global__ __launch_bounds__(kBlock) void k_code_edges(const uint64_t* __restrict__ prefix, size_t tiles,
                                                       const uint32_t* __restrict__ edges, uint32_t* __restrict__ cw) {
    const size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= tiles) return;
    const uint64_t y0 = prefix[t], y1 = prefix[t + 1];
    if (y1 == y0) return;
    const uint64_t wf = y0 / 16, wl = (y1 - 1) / 16, c0 = 16 * wl;
    if (c0 < y0 || y1 >= c0 + 16) return;  // field 0 in an earlier tile, or the word is whole
    uint32_t w = edges[2 * t + (wl == wf ? 0 : 1)];
    for (size_t u = t + 1; u < tiles; u++) {
        const uint64_t u0 = prefix[u];
        if (u0 >= c0 + 16) break;
        if (prefix[u + 1] > u0) w |= edges[2 * u];  // its first word is this one
    }
    cw[wl] = w;
}

// Splitter::Merge (reference util/splitter/splitter.cu:140-184) driven by a
// packed stencil: row i takes the next `yes` value when bit i is set.
// Set stencil bits of each tile (the merge's counting pass): one wave per
// tile, lane l taking the tile's stencil bytes [8l, 8l + 8).  Bits past row
// n - 1 (the stencil's padding) are not counted.
__global__ __launch_bounds__(kBlock) void k_stencil_count(const uint8_t* __restrict__ bits, size_t n, size_t tiles,
                                                          uint64_t* __restrict__ vals) {
    const size_t tile = (size_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    if (tile >= tiles) return;  // uniform per wave
    const int lane = threadIdx.x & (kWave - 1);
    const size_t r0 = tile * kTile + (size_t)lane * 64;
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const size_t r = r0 + 8 * k;
        if (r < n) {
            const size_t valid = n - r;
            uint32_t byte = bits[r / 8];
            if (valid < 8) byte &= (1u << valid) - 1u;
            c += __popc(byte);
        }
    }
    c = wave_sum(c);
    if (lane == 0) vals[tile] = c;
}

// The "yes" rows of a merge given as dictionary codes: row j is
// dict[code j], code j being the `bits`-wide field j % (32 / bits) of the
// little-endian word j / (32 / bits) (_decompressUniqueKernel,
// unique_encoding.cu:183-204).  Codes at or beyond K (corrupt input) take
// the last key.
struct DictCodes {
    const uint8_t* words;  // any byte address
    const uint8_t* dict;
    uint32_t bits, K;
    int pshift;  // log2(32 / bits) when bits is a power of two, else -1
};
constexpr uint32_t kLdsDict = 64;  // dictionaries up to this size are staged in LDS

// Code j: word j / per read as two aligned 32-bit loads joined by a funnel
// shift (an aligned word holding a byte of the stream never crosses a page).
template <typename T> __device__ __forceinline__ T dict_key(const DictCodes& d, const T* sdict, uint32_t word, uint32_t f) {
    uint32_t code = d.bits >= 32 ? word : (word >> (d.bits * f)) & ((1u << d.bits) - 1u);
    code = code < d.K ? code : d.K - 1;
    if (d.K <= kLdsDict) return sdict[code];
    T v;
    __builtin_memcpy(&v, d.dict + sizeof(T) * code, sizeof(T));
    return v;
}
template <typename T>
__device__ __forceinline__ T dict_value(const DictCodes& d, const T* sdict, uint64_t j) {
    const uint64_t per = 32 / d.bits;
    const uint64_t w = d.pshift >= 0 ? (j >> d.pshift) : j / per;
    const uint32_t f = (uint32_t)(j - w * per);
    const uintptr_t a = reinterpret_cast<uintptr_t>(d.words + 4 * w);
    const uint32_t* wa = reinterpret_cast<const uint32_t*>(a & ~(uintptr_t)3);
    const uint32_t sh = (uint32_t)(a & 3) * 8;
    const uint32_t word = sh ? __builtin_amdgcn_alignbit(wa[1], wa[0], sh) : wa[0];
    return dict_key<T>(d, sdict, word, f);
}

// Rows [s0, s0 + len) of the T stream at src (any byte address) into LDS
// rows dst0 + e, with aligned 16-byte loads (an aligned 16-byte chunk
// holding a valid byte never crosses a page); rows at or past `avail` read
// as zero.  Returns whether the range ran past avail (corrupt input).
// A stream whose address is not a multiple of sizeof(T) (an output stored
// unchanged in a blob, after metadata of odd length) is realigned on the
// way: every row then starts the same r = address % sizeof(T) bytes past the
// chunk grid's multiples of sizeof(T), and a row starting in chunk k is cut
// from chunks k and k + 1 by byte funnel shifts.
template <typename T>
__device__ __forceinline__ bool stage_rows(const uint8_t* __restrict__ src, uint64_t s0, int len, uint64_t avail,
                                           T* buf, int dst0) {
    constexpr int S = (int)sizeof(T), V = 16 / S;
    if (len <= 0) return false;
    const int ok = s0 >= avail ? 0 : (int)((avail - s0) < (uint64_t)len ? (avail - s0) : (uint64_t)len);
    for (int e = ok + threadIdx.x; e < len; e += kBlock) buf[P(dst0 + e)] = T(0);
    if (ok > 0) {
        const uintptr_t a = reinterpret_cast<uintptr_t>(src + s0 * S);
        const int sh = (int)(a & 15);  // bytes before the first row in its chunk
        const u32x4* v4 = reinterpret_cast<const u32x4*>(a & ~(uintptr_t)15);
        const int nv = (sh + ok * S + 15) / 16;
        if (sh % S == 0) {  // rows on the chunk grid (uniform)
            for (int k = threadIdx.x; k < nv; k += kBlock) {
                const u32x4 w = v4[k];
                T ev[V];
                __builtin_memcpy(ev, &w, 16);
                const int e0 = V * k - sh / S;
#pragma unroll
                for (int j = 0; j < V; j++) {
                    const int e = e0 + j;
                    if (e >= 0 && e < ok) buf[P(dst0 + e)] = ev[j];
                }
            }
        } else if constexpr (S > 1) {
            const int r = sh % S;
            for (int k = threadIdx.x; k < nv; k += kBlock) {
                const u32x4 c0 = v4[k];
                const u32x4 c1 = k + 1 < nv ? v4[k + 1] : u32x4{0u, 0u, 0u, 0u};
                const uint32_t w[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
                const int e0 = (16 * k + r - sh) / S;  // the first row starting in chunk k
#pragma unroll
                for (int j = 0; j < V; j++) {
                    const int e = e0 + j;
                    // row e's bytes: [r + j S, r + j S + S) of the 32-byte window
                    T v;
                    if constexpr (S == 8) {
                        const uint32_t rr = (uint32_t)(r & 3);  // r >> 2 is 0 or 1 for every thread
                        uint32_t lo, hi;
                        if ((r >> 2) == 0) {
                            lo = __builtin_amdgcn_alignbyte(w[2 * j + 1], w[2 * j], rr);
                            hi = __builtin_amdgcn_alignbyte(w[2 * j + 2], w[2 * j + 1], rr);
                        } else {
                            lo = __builtin_amdgcn_alignbyte(w[2 * j + 2], w[2 * j + 1], rr);
                            hi = __builtin_amdgcn_alignbyte(w[(2 * j + 3) & 7], w[2 * j + 2], rr);
                        }
                        const uint64_t u = ((uint64_t)hi << 32) | lo;
                        __builtin_memcpy(&v, &u, 8);
                    } else {  // S = 2 or 4: the row lies in words q, q + 1
                        const int off = r + j * S;
                        const uint32_t u = __builtin_amdgcn_alignbyte(w[(off >> 2) + 1], w[off >> 2], (uint32_t)(off & 3));
                        __builtin_memcpy(&v, &u, S);
                    }
                    if (e >= 0 && e < ok) buf[P(dst0 + e)] = v;
                }
            }
        }
    }
    return ok < len;
}

// Code words a tile's codes can touch when codes are at most 8 bits wide,
// plus a chunk of realignment slack.
constexpr int kCodeLdsWords = kTile / 4 + 2 + 8;

// Merge: out[i] = bit i ? yes[k++] : no[j++], with `no` the constant
// `no_const` (CONST_NO) and `yes` from dictionary codes (CODES).  Rows past
// n_yes / n_no (a stencil with more set or clear bits than the outputs hold:
// corrupt input) read as zero and raise the error flag.
//
// The tile's yes count is prefix[tile + 1] - prefix[tile], so its yes rows
// [Y0, Y0 + total) and no rows [N0, N0 + cnt - total) are contiguous ranges
// known at the start: they are staged in LDS (yes first, then no) with
// 16-byte loads, the stencil scan gives each row its source slot, and the
// tile is written back striped.  Codes of at most 8 bits are staged the
// same way (as the bytes of their words) and looked up per row.  FL: both
// sides are FL streams (patch -> (afl -> none), (afl -> none)); each side's
// packed words for the tile's rows are staged in LDS and unpacked into the
// same slots, so that the sides are never written unpacked.
template <typename T, bool CONST_NO, bool CODES, bool FL>
__global__ __launch_bounds__(kBlock) void k_merge(const uint8_t* __restrict__ bits, size_t n, const uint8_t* __restrict__ yes,
                                                  size_t n_yes, DictCodes dc, const uint8_t* __restrict__ no, size_t n_no,
                                                  T no_const, T* __restrict__ out, const uint64_t* __restrict__ prefix,
                                                  uint32_t* error, FlSrc fy, FlSrc fn) {
    using U = typename UnsignedOf<T>::type;
    __shared__ T buf[kTilePad];
    __shared__ uint64_t red[kBlock / kWave];
    __shared__ T sdict[CODES ? kLdsDict : 1];
    __shared__ __attribute__((aligned(16))) uint32_t cw[CODES ? kCodeLdsWords : 4];
    __shared__ __attribute__((aligned(16))) uint32_t wfl[FL ? fl_stage_words<8 * (int)sizeof(T)>() : 4];
    const uint32_t tile = blockIdx.x;
    const size_t base = (size_t)tile * kTile;
    const int cnt = (int)((n - base) < (size_t)kTile ? (n - base) : (size_t)kTile);
    const uint64_t Y0 = prefix[tile];
    const int total = (int)(prefix[tile + 1] - Y0);  // 0 <= total <= cnt: counted from the same stencil
    const size_t N0 = base - Y0;
    bool bad = false;
    // CODES: bytes of code words [w_lo, w_hi] staged from chunk boundary `al`
    uint64_t w_lo = 0;
    int off0 = 0;
    bool staged = false;
    if constexpr (CODES) {
        if (dc.K <= kLdsDict && threadIdx.x < dc.K)
            __builtin_memcpy(&sdict[threadIdx.x], dc.dict + sizeof(T) * threadIdx.x, sizeof(T));
        const uint64_t j_end = Y0 + (uint64_t)total < n_yes ? Y0 + (uint64_t)total : n_yes;  // codes that exist
        if (Y0 + (uint64_t)total > n_yes) bad = true;
        if (dc.bits <= 8 && j_end > Y0) {
            staged = true;
            const uint64_t per = 32 / dc.bits;
            w_lo = Y0 / per;
            const uint64_t w_hi = (j_end - 1) / per;
            const uintptr_t a = reinterpret_cast<uintptr_t>(dc.words + 4 * w_lo);
            off0 = (int)(a & 15);
            const u32x4* c4 = reinterpret_cast<const u32x4*>(a & ~(uintptr_t)15);
            const int nchunks = (off0 + 4 * (int)(w_hi - w_lo + 1) + 15) / 16;  // <= kCodeLdsWords / 4
            u32x4* l4 = reinterpret_cast<u32x4*>(cw);
            for (int k = threadIdx.x; k < nchunks; k += kBlock) l4[k] = c4[k];
        }
    } else if constexpr (FL) {
        // one side at a time through wfl; rows past a stream read as zero
        if (Y0 + (uint64_t)total > n_yes || N0 + (uint64_t)(cnt - total) > n_no) bad = true;
#pragma unroll
        for (int side = 0; side < 2; side++) {
            const FlSrc& f = side == 0 ? fy : fn;
            const int rows = side == 0 ? total : cnt - total;
            if (rows == 0) continue;  // uniform
            const int a = stage_fl_words(side == 0 ? yes : no, f.bytes, side == 0 ? Y0 : N0, rows, f.bits, wfl);
            __syncthreads();
            const uint64_t mask = lowmask64((unsigned)f.bits);
            const int dst0 = side == 0 ? 0 : total;
            for (int k = threadIdx.x; k < rows; k += kBlock) buf[P(dst0 + k)] = (T)(U)fl_value(wfl, k, f.bits, a, mask);
            __syncthreads();  // wfl is restaged for the other side
        }
    } else {
        bad |= stage_rows<T>(yes, Y0, total, n_yes, buf, 0);
    }
    if (!CONST_NO && !FL) bad |= stage_rows<T>(no, N0, cnt - total, n_no, buf, total);
    const size_t g0 = base + (size_t)threadIdx.x * kItems;
    uint32_t mask = 0;
    if (g0 < n) mask = bits[g0 / 8];
    if (g0 + 8 < n) mask |= (uint32_t)bits[g0 / 8 + 1] << 8;
    const int mine = cnt - threadIdx.x * kItems;
    if (mine <= 0)
        mask = 0;
    else if (mine < kItems)
        mask &= (1u << mine) - 1u;
    uint64_t tsum;
    const uint64_t excl = block_exclusive_scan<uint64_t>((uint64_t)__popc(mask), &tsum, red);  // a barrier: LDS staged
    if (bad) __hip_atomic_store(error, 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    T v[kItems];
    int yr = (int)excl, nr = total + threadIdx.x * kItems - (int)excl;
    // CODES, staged: the thread's yes rows are the consecutive codes j0 =
    // Y0 + excl, j0 + 1, ...; its first code's word and field are found
    // once, then codes are shifted out of the current word, the next word
    // read from LDS every 32 / bits codes
    auto code_word = [&](int b) -> uint32_t {  // staged code bytes [b, b + 4)
        const int q = b >> 2, r = b & 3;
        return r ? __builtin_amdgcn_alignbyte(cw[q + 1], cw[q], (uint32_t)r) : cw[q];
    };
    uint32_t cur = 0, cper = 1;
    int cq = 0, fleft = 1, ylim = 0;
    if constexpr (CODES) {
        if (staged) {
            cper = 32u / dc.bits;
            const uint64_t j0 = Y0 + excl;
            const uint64_t w0 = dc.pshift >= 0 ? (j0 >> dc.pshift) : j0 / cper;
            const uint32_t f0 = (uint32_t)(j0 - w0 * cper);
            cq = off0 + 4 * (int)(w0 - w_lo);
            cur = code_word(cq) >> (dc.bits * f0);
            fleft = (int)(cper - f0);
            ylim = j0 < n_yes ? (int)((n_yes - j0) < (uint64_t)kItems ? (n_yes - j0) : (uint64_t)kItems) : 0;
        }
    }
    bool merged = false;  // uniform
    if constexpr (CODES) {
        if (staged && dc.bits <= 2) {
            // 1- and 2-bit codes: the thread's (at most 16) codes lie in a
            // 64-bit window of the staged words, and the (at most 4) values
            // sit in registers; each row is then a select, no branch
            const uint64_t win = (uint64_t)cur | ((uint64_t)code_word(cq + 4) << (32 - (int)dc.bits * ((int)cper - fleft)));
            const uint32_t cm = (1u << dc.bits) - 1u, kmax = dc.K - 1;
            const T d0 = sdict[0], d1 = sdict[kmax < 1 ? kmax : 1], d2 = sdict[kmax < 2 ? kmax : 2],
                    d3 = sdict[kmax < 3 ? kmax : 3];
            int i = 0;
#pragma unroll
            for (int k = 0; k < kItems; k++) {
                const int f = (mask >> k) & 1;
                uint32_t code = (uint32_t)(win >> (dc.bits * i)) & cm;
                code = code < kmax ? code : kmax;
                T dv = code == 0 ? d0 : code == 1 ? d1 : code == 2 ? d2 : d3;
                dv = i < ylim ? dv : T(0);  // codes past the stream read as zero (flagged above)
                const T nv = CONST_NO ? no_const : buf[P(nr < kTile ? nr : kTile - 1)];
                v[k] = f ? dv : nv;
                i += f;
                nr += 1 - f;
            }
            merged = true;
        }
    }
#pragma unroll
    for (int k = 0; k < kItems && !merged; k++) {
        if constexpr (!CODES) {  // rows from LDS: a select, no branch
            const int f = (mask >> k) & 1;
            const int idx = f ? yr : nr;
            const T t = buf[P(idx < kTile ? idx : kTile - 1)];
            v[k] = (CONST_NO && !f) ? no_const : t;
            yr += f;
            nr += 1 - f;
            continue;
        }
        if ((mask >> k) & 1) {
            if constexpr (CODES) {
                T val = T(0);
                if (staged) {
                    uint32_t code = cur & ((1u << dc.bits) - 1u);
                    cur >>= dc.bits;
                    if (--fleft == 0) {
                        cq += 4;
                        cur = code_word(cq);
                        fleft = (int)cper;
                    }
                    if (yr - (int)excl < ylim) {  // codes past the stream read as zero (flagged above)
                        code = code < dc.K ? code : dc.K - 1;
                        if (dc.K <= kLdsDict)
                            val = sdict[code];
                        else
                            __builtin_memcpy(&val, dc.dict + sizeof(T) * code, sizeof(T));
                    }
                } else {
                    const uint64_t j = Y0 + (uint64_t)yr;
                    if (j < n_yes) val = dict_value<T>(dc, sdict, j);
                }
                v[k] = val;
            } else {
                v[k] = buf[P(yr < kTile ? yr : kTile - 1)];
            }
            yr++;
        } else {
            v[k] = CONST_NO ? no_const : buf[P(nr < kTile ? nr : kTile - 1)];
            nr++;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kItems; k++)
        if (threadIdx.x * kItems + k < cnt) buf[P(threadIdx.x * kItems + k)] = v[k];
    __syncthreads();
    if (cnt == kTile && (reinterpret_cast<uintptr_t>(out + base) & 15) == 0) {
        // whole tile: consecutive lanes store consecutive 16-byte vectors
        constexpr int V = 16 / (int)sizeof(T);
        u32x4* o4 = reinterpret_cast<u32x4*>(out + base);
        for (int i = threadIdx.x; i < kTile / V; i += kBlock) {
            T vals[V];
#pragma unroll
            for (int k = 0; k < V; k++) vals[k] = buf[P(V * i + k)];
            u32x4 w;
            __builtin_memcpy(&w, vals, 16);
            __builtin_nontemporal_store(w, o4 + i);
        }
    } else {
        for (int i = threadIdx.x; i < cnt; i += kBlock) out[base + i] = buf[P(i)];
    }
}


// RleEncoding (reference compression/rle/rle_encode.cu:18-62): a run starts
// at row 0 and wherever !(x[i] == x[i-1]); the run's value is its first row.
// The value of lane - 1 (any 1-, 2-, 4- or 8-byte type); lane 0 gets its own.
template <typename T> __device__ __forceinline__ T shfl_up_value(T v) {
    using U = typename UnsignedOf<T>::type;
    U u;
    __builtin_memcpy(&u, &v, sizeof(T));
    U r;
    if constexpr (sizeof(T) == 8) {
        r = (U)__shfl_up((unsigned long long)u, 1, kWave);
    } else {
        r = (U)__shfl_up((unsigned)u, 1, kWave);
    }
    T out;
    __builtin_memcpy(&out, &r, sizeof(T));
    return out;
}

// Runs starting in each tile (the RLE encoder's counting pass).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_rle_count(const T* __restrict__ x, size_t n, uint64_t* __restrict__ vals) {
    __shared__ uint64_t red[kBlock / kWave];
    const size_t base = (size_t)blockIdx.x * kTile;
    const int cnt = (int)((n - base) < (size_t)kTile ? (n - base) : (size_t)kTile);
    uint32_t c = 0;
    if (cnt == kTile && (reinterpret_cast<uintptr_t>(x + base) & 15) == 0) {
        constexpr int V = 16 / sizeof(T);
        constexpr int Q = kTile / V / kBlock;
        const u32x4* s4 = reinterpret_cast<const u32x4*>(x + base);
        const bool lane0 = (threadIdx.x & (kWave - 1)) == 0;
        u32x4 w[Q];
        T before[Q];  // lane 0: the row before its chunk, loaded with the chunks
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const size_t g = base + (size_t)V * (threadIdx.x + kBlock * q);
            w[q] = s4[threadIdx.x + kBlock * q];
            if (lane0 && g > 0) before[q] = x[g - 1];
        }
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const size_t g = base + (size_t)V * (threadIdx.x + kBlock * q);
            T v[V];
            __builtin_memcpy(v, &w[q], 16);
            // the row before this chunk ends the previous lane's chunk
            T prev = shfl_up_value(v[V - 1]);
            if (lane0) prev = g > 0 ? before[q] : v[0];
            c += (g == 0 || !(v[0] == prev)) ? 1u : 0u;
#pragma unroll
            for (int k = 1; k < V; k++) c += !(v[k] == v[k - 1]) ? 1u : 0u;
        }
    } else {
        for (int i = threadIdx.x; i < cnt; i += kBlock) {
            const size_t g = base + i;
            c += (g == 0 || !(x[g] == x[g - 1])) ? 1u : 0u;
        }
    }
    uint64_t total;
    block_exclusive_scan<uint64_t>((uint64_t)c, &total, red);
    if (threadIdx.x == 0) vals[blockIdx.x] = total;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_rle_encode(const T* __restrict__ x, size_t n, T* __restrict__ values,
                                                       uint32_t* __restrict__ starts,
                                                       const uint64_t* __restrict__ prefix, uint64_t* d_runs) {
    __shared__ T buf[kTilePad + 1];
    __shared__ uint64_t red[kBlock / kWave];
    const uint32_t tile = blockIdx.x;
    const size_t base = (size_t)tile * kTile;
    const int cnt = (int)((n - base) < (size_t)kTile ? (n - base) : (size_t)kTile);
    const uint64_t s_prefix = prefix[tile];  // loaded with the tile
    if (cnt == kTile && (reinterpret_cast<uintptr_t>(x + base) & 15) == 0) {
        // buf[i + 1] = x[base + i]; buf[0] = the row before the tile
        constexpr int V = 16 / sizeof(T);
        const u32x4* s4 = reinterpret_cast<const u32x4*>(x + base);
        for (int i = threadIdx.x; i < kTile / V; i += kBlock) {
            const u32x4 w = s4[i];
            T vals[V];
            __builtin_memcpy(vals, &w, 16);
#pragma unroll
            for (int k = 0; k < V; k++) buf[P(V * i + k + 1)] = vals[k];
        }
        if (threadIdx.x == 0) buf[P(0)] = base > 0 ? x[base - 1] : T(0);
    } else {
        for (int i = threadIdx.x; i < kTile + 1; i += kBlock) {
            const long long g = (long long)base + i - 1;
            buf[P(i)] = (g >= 0 && i <= cnt) ? x[g] : T(0);
        }
    }
    __syncthreads();
    uint32_t mask = 0;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        const int idx = threadIdx.x * kItems + k;
        if (idx < cnt) {
            const bool f = (base + idx == 0) || !(buf[P(idx + 1)] == buf[P(idx)]);
            mask |= (uint32_t)f << k;
        }
    }
    uint64_t total;
    const uint64_t excl = block_exclusive_scan<uint64_t>((uint64_t)__popc(mask), &total, red);
    uint64_t r = s_prefix + excl;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        if ((mask >> k) & 1) {
            const int idx = threadIdx.x * kItems + k;
            values[r] = buf[P(idx + 1)];
            starts[r] = (uint32_t)(base + idx);
            r++;
        }
    }
    if (threadIdx.x == 0 && base + cnt == n) d_runs[0] = s_prefix + total;
}

__global__ __launch_bounds__(kBlock) void k_rle_lengths(const uint32_t* __restrict__ starts, const uint64_t* d_runs,
                                                        size_t n, int32_t* __restrict__ lengths) {
    const size_t runs = *d_runs;
    const size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t r = (size_t)blockIdx.x * kBlock + threadIdx.x; r < runs; r += stride) {
        const uint64_t e = r + 1 < runs ? starts[r + 1] : n;
        lengths[r] = (int32_t)(e - starts[r]);
    }
}

// Run length r of an RLE stream at any byte address; a negative length is
// corrupt input and counts as an empty run.
__device__ __forceinline__ uint64_t run_length(const uint8_t* lens, size_t r, bool aligned, bool& bad) {
    int32_t l;
    if (aligned)
        l = reinterpret_cast<const int32_t*>(lens)[r];
    else
        __builtin_memcpy(&l, lens + 4 * r, 4);
    if (l < 0) {
        bad = true;
        return 0;
    }
    return (uint64_t)l;
}

// Rows covered by each tile of 4096 runs (the decoder's counting pass); a
// negative length raises the error flag (2: corrupt input).
__global__ __launch_bounds__(kBlock) void k_length_sums(const uint8_t* __restrict__ lens, size_t R,
                                                        uint64_t* __restrict__ vals, uint32_t* error) {
    __shared__ uint64_t red[kBlock / kWave];
    const size_t base = (size_t)blockIdx.x * kTile;
    const int cnt = (int)((R - base) < (size_t)kTile ? (R - base) : (size_t)kTile);
    const bool aligned = (reinterpret_cast<uintptr_t>(lens) & 3) == 0;
    bool bad = false;
    uint64_t s = 0;
    for (int i = threadIdx.x; i < cnt; i += kBlock) s += run_length(lens, base + i, aligned, bad);
    if (bad) __hip_atomic_store(error, 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint64_t total;
    block_exclusive_scan<uint64_t>(s, &total, red);
    if (threadIdx.x == 0) vals[blockIdx.x] = total;
}

// Exclusive scan of int32 run lengths into uint64 offsets[R+1].
__global__ __launch_bounds__(kBlock) void k_scan_lengths(const uint8_t* __restrict__ lens, size_t R,
                                                         uint64_t* __restrict__ offsets,
                                                         const uint64_t* __restrict__ prefix) {
    __shared__ uint64_t buf[kTile + kTile / 32];
    __shared__ uint64_t red[kBlock / kWave];
    auto pad = [](int i) { return i + (i >> 5); };
    const uint32_t tile = blockIdx.x;
    const size_t base = (size_t)tile * kTile;
    const int cnt = (int)((R - base) < (size_t)kTile ? (R - base) : (size_t)kTile);
    const bool aligned = (reinterpret_cast<uintptr_t>(lens) & 3) == 0;
    bool bad = false;
    for (int i = threadIdx.x; i < kTile; i += kBlock) buf[pad(i)] = i < cnt ? run_length(lens, base + i, aligned, bad) : 0;
    __syncthreads();
    uint64_t loc[kItems], s = 0;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        loc[k] = buf[pad(threadIdx.x * kItems + k)];
        s += loc[k];
    }
    uint64_t total;
    const uint64_t excl = block_exclusive_scan<uint64_t>(s, &total, red);  // a barrier: buf is free again
    const uint64_t s_prefix = prefix[tile];
    uint64_t run = s_prefix + excl;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        buf[pad(threadIdx.x * kItems + k)] = run;
        run += loc[k];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += kBlock) offsets[base + i] = buf[pad(i)];
    if (threadIdx.x == 0 && base + cnt == R) offsets[R] = s_prefix + total;
}

// Run expansion: a workgroup owns 4096 output rows, finds the first run
// overlapping them by binary search, stages the overlapping runs in LDS and
// gives every row its run by a binary search in LDS.
// Expansion tiles: TILE output rows per workgroup; LDS holds up to TILE
// runs of a tile (all of them, unless some runs are empty).

// tile_run[t] = the run holding row t * TILE (offsets are non-decreasing:
// lengths were checked non-negative by the scan); tile_run[tiles] = R.
template <int TILE>
__global__ __launch_bounds__(kBlock) void k_rle_tile_runs(const uint64_t* __restrict__ offsets, size_t R, size_t tiles,
                                                          uint32_t* __restrict__ tile_run) {
    const size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t r = (size_t)blockIdx.x * kBlock + threadIdx.x; r < R; r += stride) {
        const uint64_t s = offsets[r], e = offsets[r + 1];
        const uint64_t t0 = (s + TILE - 1) / TILE, t1 = (e + TILE - 1) / TILE;  // tiles starting in [s, e)
        for (uint64_t t = t0; t < t1 && t < tiles; t++) tile_run[t] = (uint32_t)r;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) tile_run[tiles] = (uint32_t)R;
}

// Run expansion: a workgroup owns TILE output rows; its runs are
// tile_run[b] .. tile_run[b + 1] (the run holding the next tile's first row).
// They are staged in LDS and every row finds its run by binary search there;
// a tile overlapping more runs than LDS holds (only possible with empty
// runs) searches the offsets in global memory instead.
template <typename T, int TILE>
__global__ __launch_bounds__(kBlock) void k_rle_expand(const uint64_t* __restrict__ offsets, size_t R,
                                                       const uint8_t* __restrict__ values, size_t total,
                                                       const uint32_t* __restrict__ tile_run, T* __restrict__ out) {
    __shared__ uint64_t soff[TILE + 2];
    __shared__ T sval[TILE + 1];
    const size_t o0 = (size_t)blockIdx.x * TILE;
    const int cnt = (int)((total - o0) < (size_t)TILE ? (total - o0) : (size_t)TILE);
    const size_t r0 = tile_run[blockIdx.x];
    size_t r1 = tile_run[blockIdx.x + 1];  // last run needed (inclusive)
    r1 = r1 < R ? r1 : R - 1;
    r1 = r1 < r0 ? r0 : r1;
    const size_t nr = r1 - r0 + 1;
    if (nr <= (size_t)TILE + 1) {
        for (int i = threadIdx.x; i <= (int)nr; i += kBlock) soff[i] = offsets[r0 + i];
        for (int i = threadIdx.x; i < (int)nr; i += kBlock) __builtin_memcpy(&sval[i], values + sizeof(T) * (r0 + i), sizeof(T));
        __syncthreads();
        // thread t writes rows [R t, R t + R) of the tile: one binary search
        // for the first, then the runs advance monotonically
        constexpr int R = TILE / kBlock;
        const int k0 = threadIdx.x * R;
        if (k0 < cnt) {
            const uint64_t pos0 = o0 + k0;
            int lo = 0, hi = (int)nr;  // soff[lo] <= pos0 < soff[hi]
            while (hi - lo > 1) {
                const int mid = (lo + hi) / 2;
                if (soff[mid] <= pos0)
                    lo = mid;
                else
                    hi = mid;
            }
            T v[R];
#pragma unroll
            for (int k = 0; k < R; k++) {
                while (lo + 1 < (int)nr && soff[lo + 1] <= pos0 + k) lo++;
                v[k] = sval[lo];
            }
            if (k0 + R <= cnt && (R * sizeof(T)) % 16 == 0) {
                u32x4* d4 = reinterpret_cast<u32x4*>(out + pos0);  // 16-byte aligned: R * sizeof(T) % 16 == 0
#pragma unroll
                for (int q = 0; q < (int)(R * sizeof(T) / 16); q++) {
                    u32x4 w;
                    __builtin_memcpy(&w, reinterpret_cast<const uint8_t*>(v) + 16 * q, 16);
                    d4[q] = w;
                }
            } else {
                for (int k = 0; k < R && k0 + k < cnt; k++) out[pos0 + k] = v[k];
            }
        }
    } else {
        for (int k = threadIdx.x; k < cnt; k += kBlock) {
            const uint64_t pos = o0 + k;
            size_t lo = r0, hi = r1 + 1;  // offsets[lo] <= pos < offsets[hi]
            while (hi - lo > 1) {
                const size_t mid = (lo + hi) / 2;
                if (offsets[mid] <= pos)
                    lo = mid;
                else
                    hi = mid;
            }
            T v;
            __builtin_memcpy(&v, values + sizeof(T) * lo, sizeof(T));
            out[pos] = v;
        }
    }
}

inline unsigned grid_for(Ctx* c, size_t items) {
    size_t g = (items + kBlock - 1) / kBlock;
    size_t cap = (size_t)c->cu_count * 32;
    if (g > cap) g = cap;
    return (unsigned)(g ? g : 1);
}

}  // namespace

namespace {
// Calls f(std::integral_constant<int, kind>) for a predicate kind.
template <typename F> void with_kind(int kind, F&& f) {
    switch (kind) {
    case PRED_LOWER: f(std::integral_constant<int, PRED_LOWER>{}); break;
    case PRED_INSIDE: f(std::integral_constant<int, PRED_INSIDE>{}); break;
    case PRED_NOT_EQUAL: f(std::integral_constant<int, PRED_NOT_EQUAL>{}); break;
    case PRED_IN_SET: f(std::integral_constant<int, PRED_IN_SET>{}); break;
    default: fail(DDJ_EINVAL, "unknown split predicate");
    }
}
}  // namespace

namespace {
template <typename T> Pred<T> make_pred(const PredArgs& pa) {
    Pred<T> p;
    p.kind = pa.kind;
    p.a = bits_to<T>(pa.a_bits);
    p.b = bits_to<T>(pa.b_bits);
    for (int j = 0; j < 4; j++) p.set[j] = bits_to<T>(pa.set_bits[j]);
    p.set_n = pa.set_n;
    p.active = pa.set_n >= 4 ? 0xFu : pa.set_n <= 0 ? 0u : (1u << pa.set_n) - 1u;
    return p;
}
}  // namespace

SplitPlan launch_split_count(Ctx* c, int dtype, const void* x, size_t n, const PredArgs& pa, MinMaxResult* mm_yes,
                             MinMaxResult* mm_no) {
    SplitPlan plan;
    plan.dtype = dtype;
    plan.x = x;
    plan.n = n;
    plan.p = pa;
    if (n == 0) return plan;
    dispatch_native(dtype, [&](auto tag) {
        using T = decltype(tag);
        const Pred<T> p = make_pred<T>(pa);
        const size_t tiles = (n + kTile - 1) / kTile;
        plan.ts = std::make_shared<TileScan>(c, tiles);
        TileScan& ts = *plan.ts;
        auto xs = static_cast<const T*>(x);
        kernel_bytes(c, "k_split_count", (double)n * sizeof(T));
        const bool mm = mm_yes && mm_no;
        if ((mm_yes || mm_no) && !(mm && std::is_integral<T>::value))
            fail(DDJ_EINVAL, "split min/max outputs need an integer type and both outputs");
        DevBuf keys;
        if (mm) keys = DevBuf(c, 32 * tiles);
        const dim3 grid((unsigned)tiles), block(kBlock);
        with_kind(p.kind, [&](auto kind) {
            constexpr int K = decltype(kind)::value;
            KernelTimer kt_(c, "k_split_count");
            if constexpr (std::is_integral<T>::value) {
                if (mm) {
                    hipLaunchKernelGGL((k_split_count<T, true, K>), grid, block, 0, c->stream, xs, n, p, ts.vals,
                                       keys.get<uint64_t>());
                    return;
                }
            }
            hipLaunchKernelGGL((k_split_count<T, false, K>), grid, block, 0, c->stream, xs, n, p, ts.vals, nullptr);
        });
        HIPC(hipGetLastError());
        ts.scan(c, mm ? keys.get<uint64_t>() : nullptr, mm_yes, mm_no);
    });
    return plan;
}

void launch_split_write(Ctx* c, const SplitPlan& plan, uint8_t* stencil, void* yes, uint8_t* no, uint8_t* codes,
                        uint64_t* d_counts) {
    if (plan.n == 0) return;
    dispatch_native(plan.dtype, [&](auto tag) {
        using T = decltype(tag);
        const Pred<T> p = make_pred<T>(plan.p);
        if (codes && p.kind != PRED_IN_SET) fail(DDJ_EINVAL, "split codes need a dictionary predicate");
        const size_t n = plan.n, tiles = plan.ts->tiles;
        auto xs = static_cast<const T*>(plan.x);
        auto ys = static_cast<T*>(yes);
        const uint64_t* pre = plan.ts->prefix;
        const dim3 grid((unsigned)tiles), block(kBlock);
        DevBuf edges;
        if (codes) edges = DevBuf(c, 8 * tiles);
        uint32_t* ed = codes ? edges.get<uint32_t>() : nullptr;
        with_kind(p.kind, [&](auto kind) {
            constexpr int K = decltype(kind)::value;
            KernelTimer kt_(c, "k_split");
            if constexpr (K == PRED_IN_SET) {
                if (codes) {
                    hipLaunchKernelGGL((k_split<T, true, true, K>), grid, block, 0, c->stream, xs, n, p, stencil, ys, no,
                                       codes, ed, pre, d_counts);
                    return;
                }
            }
            if (no)
                hipLaunchKernelGGL((k_split<T, false, true, K>), grid, block, 0, c->stream, xs, n, p, stencil, ys, no,
                                   codes, ed, pre, d_counts);
            else
                hipLaunchKernelGGL((k_split<T, false, false, K>), grid, block, 0, c->stream, xs, n, p, stencil, ys, no,
                                   codes, ed, pre, d_counts);
        });
        if (codes) {
            HIPC(hipGetLastError());
            KernelTimer kt_(c, "k_code_edges");
            hipLaunchKernelGGL(k_code_edges, dim3((unsigned)((tiles + kBlock - 1) / kBlock)), block, 0, c->stream, pre,
                               tiles, ed, reinterpret_cast<uint32_t*>(codes));
        }
        HIPC(hipGetLastError());
    });
}

void launch_split(Ctx* c, int dtype, const void* x, size_t n, const PredArgs& pa, uint8_t* stencil, void* yes,
                  void* no, uint8_t* codes, uint64_t* d_counts, MinMaxResult* mm_yes, MinMaxResult* mm_no) {
    if (n == 0) return;
    const SplitPlan plan = launch_split_count(c, dtype, x, n, pa, mm_yes, mm_no);
    launch_split_write(c, plan, stencil, yes, static_cast<uint8_t*>(no), codes, d_counts);
}

void launch_split_fl(Ctx* c, const SplitPlan& plan, uint8_t* stencil, FlSide yes, FlSide no, uint64_t* d_counts) {
    if (plan.n == 0) return;
    if (plan.p.kind != PRED_LOWER && plan.p.kind != PRED_INSIDE) fail(DDJ_EINVAL, "FL split: patch predicates only");
    const size_t tiles = plan.ts->tiles;
    DevBuf edges(c, 16 * tiles);
    yes.edges = edges.get<uint32_t>();
    no.edges = yes.edges;  // one array for both sides: 4 words a tile
    dispatch_int((int)dtype_size(plan.dtype), [&](auto tag) {
        using T = decltype(tag);
        if (plan.dtype == DDJ_D_FLOAT || plan.dtype == DDJ_D_DOUBLE) fail(DDJ_EINVAL, "FL split: integer columns only");
        if (yes.bits <= 0 || yes.bits > 8 * (int)sizeof(T) || no.bits <= 0 || no.bits > 8 * (int)sizeof(T))
            fail(DDJ_EINVAL, "FL split: bit width out of range");
        const Pred<T> p = make_pred<T>(plan.p);
        const uint64_t* pre = plan.ts->prefix;
        kernel_bytes(c, "k_split_fl", (double)plan.n * sizeof(T) + 1 + (double)((plan.n + 7) / 8) + (double)yes.bytes +
                                       (double)no.bytes);
        {
            KernelTimer kt_(c, "k_split_fl");
            with_kind(p.kind, [&](auto kind) {
                constexpr int K = decltype(kind)::value;
                if constexpr (K == PRED_LOWER || K == PRED_INSIDE)
                    hipLaunchKernelGGL((k_split_fl<T, K>), dim3((unsigned)tiles), dim3(kBlock), 0, c->stream,
                                       static_cast<const T*>(plan.x), plan.n, p, stencil, yes, no, pre, d_counts);
            });
        }
        HIPC(hipGetLastError());
        KernelTimer kt_(c, "k_fl_edges");
        hipLaunchKernelGGL(k_fl_edges, dim3((unsigned)((2 * tiles + kBlock - 1) / kBlock)), dim3(kBlock), 0, c->stream,
                           pre, tiles, plan.n, yes, no);
        HIPC(hipGetLastError());
    });
}

namespace {

// Stencil counts, their scan and the merge; `dc.words` set: yes rows from codes.
template <typename T>
void merge_impl(Ctx* c, const uint8_t* bits, size_t n, const uint8_t* yes, size_t n_yes, const DictCodes& dc,
                const uint8_t* no, size_t n_no, bool const_no, T no_const, T* out) {
    const size_t tiles = (n + kTile - 1) / kTile;
    TileScan ts(c, tiles);
    arm_error_flag(c);
    const unsigned cg = (unsigned)((tiles + kBlock / kWave - 1) / (kBlock / kWave));
    kernel_bytes(c, "k_stencil_count", (double)((n + 7) / 8));
    { KernelTimer kt_(c, "k_stencil_count"); hipLaunchKernelGGL(k_stencil_count, dim3(cg), dim3(kBlock), 0, c->stream, bits, n, tiles, ts.vals); }
    HIPC(hipGetLastError());
    ts.scan(c);
    const uint64_t* pre = ts.prefix;
    uint32_t* err = &c->control->error_flag;
    const dim3 grid((unsigned)tiles);
    KernelTimer kt_(c, "k_merge");
    if (dc.words) {
        kernel_bytes(c, "k_merge", (double)((n + 7) / 8) + (double)n_yes * dc.bits / 8 + (double)n_no * sizeof(T) +
                                       (double)n * sizeof(T));
        hipLaunchKernelGGL((k_merge<T, false, true, false>), grid, dim3(kBlock), 0, c->stream, bits, n, yes, n_yes, dc, no, n_no,
                           T(0), out, pre, err, FlSrc{}, FlSrc{});
    } else if (const_no) {
        kernel_bytes(c, "k_merge", (double)((n + 7) / 8) + (double)n_yes * sizeof(T) + (double)n * sizeof(T));
        hipLaunchKernelGGL((k_merge<T, true, false, false>), grid, dim3(kBlock), 0, c->stream, bits, n, yes, n_yes, dc, no, n_no,
                           no_const, out, pre, err, FlSrc{}, FlSrc{});
    } else {
        kernel_bytes(c, "k_merge", (double)((n + 7) / 8) + 2.0 * (double)n * sizeof(T));
        hipLaunchKernelGGL((k_merge<T, false, false, false>), grid, dim3(kBlock), 0, c->stream, bits, n, yes, n_yes, dc, no, n_no,
                           T(0), out, pre, err, FlSrc{}, FlSrc{});
    }
    HIPC(hipGetLastError());
}

// The merge of a patch whose sides are FL streams.
template <typename T>
void merge_fl_impl(Ctx* c, const uint8_t* bits, size_t n, const uint8_t* yes, size_t n_yes, FlSrc fy,
                   const uint8_t* no, size_t n_no, FlSrc fn, T* out) {
    const size_t tiles = (n + kTile - 1) / kTile;
    TileScan ts(c, tiles);
    arm_error_flag(c);
    const unsigned cg = (unsigned)((tiles + kBlock / kWave - 1) / (kBlock / kWave));
    kernel_bytes(c, "k_stencil_count", (double)((n + 7) / 8));
    { KernelTimer kt_(c, "k_stencil_count"); hipLaunchKernelGGL(k_stencil_count, dim3(cg), dim3(kBlock), 0, c->stream, bits, n, tiles, ts.vals); }
    HIPC(hipGetLastError());
    ts.scan(c);
    KernelTimer kt_(c, "k_merge_fl");
    kernel_bytes(c, "k_merge_fl", (double)((n + 7) / 8) + (double)fy.bytes + (double)fn.bytes + (double)n * sizeof(T));
    hipLaunchKernelGGL((k_merge<T, false, false, true>), dim3((unsigned)tiles), dim3(kBlock), 0, c->stream, bits, n, yes,
                       n_yes, DictCodes{}, no, n_no, T(0), out, ts.prefix, &c->control->error_flag, fy, fn);
    HIPC(hipGetLastError());
}

}  // namespace

void launch_merge_fl(Ctx* c, int dtype, const uint8_t* bits, size_t n, const uint8_t* yes, size_t yes_bytes,
                     size_t n_yes, int by, const uint8_t* no, size_t no_bytes, size_t n_no, int bn, void* out) {
    if (n == 0) return;
    dispatch_int((int)dtype_size(dtype), [&](auto tag) {
        using T = decltype(tag);
        if (dtype == DDJ_D_FLOAT || dtype == DDJ_D_DOUBLE) fail(DDJ_EINVAL, "FL merge: integer columns only");
        const int W = 8 * (int)sizeof(T);
        if ((n_yes && (by <= 0 || by > W)) || (n_no && (bn <= 0 || bn > W))) fail(DDJ_ECORRUPT, "FL merge: bit width out of range");
        // a side with no rows is never staged: give it a width that keeps the arithmetic in range
        merge_fl_impl<T>(c, bits, n, yes, n_yes, FlSrc{yes_bytes, n_yes ? by : 1}, no, n_no,
                         FlSrc{no_bytes, n_no ? bn : 1}, static_cast<T*>(out));
    });
}

void launch_merge(Ctx* c, int dtype, const uint8_t* bits, size_t n, const uint8_t* yes, size_t n_yes,
                  const uint8_t* no, size_t n_no, bool const_no, uint64_t no_const, void* out) {
    if (n == 0) return;
    dispatch_native(dtype, [&](auto tag) {
        using T = decltype(tag);
        merge_impl<T>(c, bits, n, yes, n_yes, DictCodes{}, no,
                      n_no, const_no, bits_to<T>(no_const), static_cast<T*>(out));
    });
}

void launch_merge_codes(Ctx* c, int dtype, const uint8_t* bits, size_t n, const uint8_t* words, size_t m, int code_bits,
                        const uint8_t* dict, uint32_t K, const uint8_t* no, size_t n_no, void* out) {
    if (n == 0) return;
    if (code_bits < 1 || code_bits > 32 || K == 0) fail(DDJ_ECORRUPT, "dictionary code width out of range");
    DictCodes dc{words, dict, (uint32_t)code_bits, K, -1};
    if ((code_bits & (code_bits - 1)) == 0) {
        int l = 0;
        while ((1 << l) < code_bits) l++;
        dc.pshift = 5 - l;
    }
    dispatch_native(dtype, [&](auto tag) {
        using T = decltype(tag);
        merge_impl<T>(c, bits, n, nullptr, m, dc, no, n_no, false, T(0),
                      static_cast<T*>(out));
    });
}



uint64_t launch_rle_encode(Ctx* c, int dtype, const void* x, size_t n, int32_t* lengths, void* values, uint64_t* d_runs,
                           MinMaxResult* h_values, MinMaxResult* h_lengths) {
    if (n == 0) return 0;
    if (n >= (1ull << 31)) fail(DDJ_EUNSUPPORTED, "RLE run lengths are int32 (reference rle_encoding.hpp:35)");
    return dispatch_native(dtype, [&](auto tag) -> uint64_t {
        using T = decltype(tag);
        const size_t tiles = (n + kTile - 1) / kTile;
        TileScan ts(c, tiles);
        kernel_bytes(c, "k_rle_count", (double)n * sizeof(T));
        { KernelTimer kt_(c, "k_rle_count"); hipLaunchKernelGGL(k_rle_count<T>, dim3((unsigned)tiles), dim3(kBlock), 0, c->stream,
                           static_cast<const T*>(x), n, ts.vals); }
        HIPC(hipGetLastError());
        ts.scan(c);
        DevBuf starts(c, 4 * n);
        { KernelTimer kt_(c, "k_rle_encode"); hipLaunchKernelGGL(k_rle_encode<T>, dim3((unsigned)tiles), dim3(kBlock), 0, c->stream,
                           static_cast<const T*>(x), n, static_cast<T*>(values), starts.get<uint32_t>(),
                           (const uint64_t*)ts.prefix, d_runs); }
        HIPC(hipGetLastError());
        // lengths (and the outputs' min/max) over the run count the device
        // holds, so that one read-back brings everything the next nodes need
        { KernelTimer kt_(c, "k_rle_lengths"); hipLaunchKernelGGL(k_rle_lengths, dim3(grid_for(c, n)), dim3(kBlock), 0, c->stream,
                           starts.get<uint32_t>(), (const uint64_t*)d_runs, n, lengths); }
        HIPC(hipGetLastError());
        // the two min/max reductions' partials come back with the run count
        // and are folded on the host
        const bool mv = h_values && std::is_integral<T>::value;
        DevBuf pv, pl;
        const int gv = mv ? launch_minmax_partials(c, dtype, values, n, d_runs, pv) : 0;
        const int gl = h_lengths ? launch_minmax_partials(c, DDJ_D_INT, lengths, n, d_runs, pl) : 0;
        std::vector<uint64_t> hv(3 * (size_t)gv), hl(3 * (size_t)gl);
        uint64_t runs = 0;
        ReadBack rb[3] = {{&runs, d_runs, 8}, {nullptr, nullptr, 0}, {nullptr, nullptr, 0}};
        size_t k = 1;
        if (mv) rb[k++] = ReadBack{hv.data(), pv.get(), 8 * hv.size()};
        if (h_lengths) rb[k++] = ReadBack{hl.data(), pl.get(), 8 * hl.size()};
        d2h_many(c, rb, k);
        if (mv) *h_values = fold_minmax_partials(hv.data(), gv, 3);
        if (h_lengths) *h_lengths = fold_minmax_partials(hl.data(), gl, 3);
        return runs;
    });
}

DevBuf rle_decode_dev(Ctx* c, int dtype, const uint8_t* lengths, const uint8_t* values, size_t runs) {
    // rle_decode.cu:46-67: scan of the lengths sizes the output (one read-back), then runs expand.
    if (runs == 0) return DevBuf();
    return dispatch_native(dtype, [&](auto tag) -> DevBuf {
        using T = decltype(tag);
        const size_t tiles = (runs + kTile - 1) / kTile;
        TileScan ts(c, tiles);
        arm_error_flag(c);
        kernel_bytes(c, "k_length_sums", 4.0 * (double)runs);
        kernel_bytes(c, "k_scan_lengths", 4.0 * (double)runs + 8.0 * (double)(runs + 1));
        { KernelTimer kt_(c, "k_length_sums"); hipLaunchKernelGGL(k_length_sums, dim3((unsigned)tiles), dim3(kBlock), 0, c->stream, lengths, runs,
                           ts.vals, &c->control->error_flag); }
        HIPC(hipGetLastError());
        ts.scan(c);
        DevBuf offsets(c, 8 * (runs + 1));
        { KernelTimer kt_(c, "k_scan_lengths"); hipLaunchKernelGGL(k_scan_lengths, dim3((unsigned)tiles), dim3(kBlock), 0, c->stream, lengths, runs,
                           offsets.get<uint64_t>(), (const uint64_t*)ts.prefix); }
        HIPC(hipGetLastError());
        // the output size and the error flag in one read-back: a negative
        // length fails here; otherwise nothing launched so far raised the
        // flag and the expansion cannot, so the call ends without reading it
        uint64_t total = 0;
        uint32_t err = 0;
        const ReadBack rb[2] = {{&total, offsets.get<uint64_t>() + runs, 8}, {&err, &c->control->error_flag, 4}};
        d2h_many(c, rb, 2);
        if (err) {
            const uint32_t zero = 0;
            h2d(c, &c->control->error_flag, &zero, 4);
            c->error_flag_armed = false;
            fail(DDJ_ECORRUPT, "encoded input failed a device-side check");
        }
        c->error_flag_armed = false;
        DevBuf out(c, total * sizeof(T));
        if (total == 0) return out;
        kernel_bytes(c, "k_rle_expand", (double)total * sizeof(T) + (double)runs * (8 + sizeof(T)));
        auto expand = [&](auto tile_c) {
            constexpr int TILE = decltype(tile_c)::value;
            const size_t etiles = (total + TILE - 1) / TILE;
            DevBuf tile_run(c, 4 * (etiles + 1));
            { KernelTimer kt_(c, "k_rle_tile_runs"); hipLaunchKernelGGL(k_rle_tile_runs<TILE>, dim3(grid_for(c, runs)), dim3(kBlock), 0, c->stream,
                               offsets.get<uint64_t>(), runs, etiles, tile_run.get<uint32_t>()); }
            { KernelTimer kt_(c, "k_rle_expand"); hipLaunchKernelGGL((k_rle_expand<T, TILE>), dim3((unsigned)etiles), dim3(kBlock), 0, c->stream,
                               offsets.get<uint64_t>(), runs, values, (size_t)total, tile_run.get<uint32_t>(), out.get<T>()); }
            HIPC(hipGetLastError());
        };
        // 2048-row tiles: 1024-row tiles (more, shorter-lived workgroups) and
        // 8192-row tiles (fewer, each row walking further) measured slower
        expand(std::integral_constant<int, 2048>{});
        return out;
    });
}

}  // namespace ddj
