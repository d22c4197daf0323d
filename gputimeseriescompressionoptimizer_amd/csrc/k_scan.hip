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
#include "dispatch.h"
#include "k_common.h"
#include "kernels.h"

namespace ddj {

namespace {

constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;
// LDS index of tile row i: one pad slot every 16 rows spreads the blocked
// (16 consecutive rows per thread) accesses over the banks.
constexpr int kTilePad = kTile + kTile / 16;
__device__ __forceinline__ int P(int i) { return i + (i >> 4); }

template <typename T> struct Pred {
    int kind;
    T a, b;
    T set[4];
    int set_n;
    __device__ __forceinline__ bool operator()(T v, uint32_t& code) const {
        code = 0;
        switch (kind) {
        case PRED_LOWER: return v < a;                // LowerOperator (util/stencil/stencil_operators.hpp)
        case PRED_INSIDE: return !(v > b || v < a);   // OutsideOperator: true inside [low, high]
        case PRED_NOT_EQUAL: return v != a;           // _constEncodeStencilKernel (const_encoding.cu:19-25)
        default: {                                    // getMostFrequentStencilKernel (dict_encoding.cu:15-34)
            bool r = false;
#pragma unroll
            for (int j = 0; j < 4; j++)               // the dictionary holds at most 4 keys
                if (j < set_n && v == set[j]) {
                    r = true;
                    code |= (uint32_t)j;              // _compressUniqueKernel ORs every matching index
                }
            return r;
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
template <typename T>
__global__ __launch_bounds__(kBlock) void k_split_count(const T* __restrict__ x, size_t n, Pred<T> p,
                                                        uint64_t* __restrict__ vals) {
    __shared__ uint64_t red[kBlock / kWave];
    const size_t base = (size_t)blockIdx.x * kTile;
    const int cnt = (int)((n - base) < (size_t)kTile ? (n - base) : (size_t)kTile);
    uint32_t c = 0, code;
    if (cnt == kTile && (reinterpret_cast<uintptr_t>(x + base) & 15) == 0) {
        constexpr int V = 16 / sizeof(T);
        const u32x4* s4 = reinterpret_cast<const u32x4*>(x + base);
#pragma unroll
        for (int q = 0; q < kTile / V / kBlock; q++) {
            const u32x4 w = __builtin_nontemporal_load(s4 + threadIdx.x + kBlock * q);
            T v[V];
            __builtin_memcpy(v, &w, 16);
#pragma unroll
            for (int k = 0; k < V; k++) c += p(v[k], code) ? 1u : 0u;
        }
    } else {
        for (int i = threadIdx.x; i < cnt; i += kBlock) c += p(x[base + i], code) ? 1u : 0u;
    }
    uint64_t total;
    block_exclusive_scan<uint64_t>((uint64_t)c, &total, red);
    if (threadIdx.x == 0) vals[blockIdx.x] = total;
}

template <typename T, bool CODES, bool WRITE_NO>
__global__ __launch_bounds__(kBlock) void k_split(const T* __restrict__ x, size_t n, Pred<T> p,
                                                  uint8_t* __restrict__ stencil, T* __restrict__ yes,
                                                  T* __restrict__ no, uint8_t* __restrict__ codes,
                                                  const uint64_t* __restrict__ prefix, uint64_t* d_count) {
    __shared__ T buf[kTilePad];
    __shared__ uint8_t cbuf[CODES ? kTile : 1];
    __shared__ uint64_t red[kBlock / kWave];
    const uint32_t tile = blockIdx.x;
    const size_t base = (size_t)tile * kTile;
    const int cnt = (int)((n - base) < (size_t)kTile ? (n - base) : (size_t)kTile);
    const T padv = x[n - 1];  // stencil padding repeats the last row (reference stencil.cu:26-27)
    if (cnt == kTile && (reinterpret_cast<uintptr_t>(x + base) & 15) == 0) {
        constexpr int V = 16 / sizeof(T);
        const u32x4* s4 = reinterpret_cast<const u32x4*>(x + base);
        for (int i = threadIdx.x; i < kTile / V; i += kBlock) {
            const u32x4 w = __builtin_nontemporal_load(s4 + i);
            T vals[V];
            __builtin_memcpy(vals, &w, 16);
#pragma unroll
            for (int k = 0; k < V; k++) buf[P(V * i + k)] = vals[k];
        }
    } else {
        for (int i = threadIdx.x; i < kTile; i += kBlock) buf[P(i)] = i < cnt ? x[base + i] : padv;
    }
    __syncthreads();
    T v[kItems];
    uint32_t mask = 0;
    uint64_t cpack = 0;
    int c = 0;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        v[k] = buf[P(threadIdx.x * kItems + k)];
        uint32_t code;
        bool f = p(v[k], code);
        mask |= (uint32_t)f << k;
        cpack |= (uint64_t)code << (4 * k);
        if (threadIdx.x * kItems + k < cnt) c += f;
    }
    const size_t g0 = base + (size_t)threadIdx.x * kItems;
    if (g0 < n) stencil[1 + g0 / 8] = (uint8_t)(mask & 0xff);
    if (g0 + 8 < n) stencil[2 + g0 / 8] = (uint8_t)(mask >> 8);
    if (tile == 0 && threadIdx.x == 0) stencil[0] = (uint8_t)(n % 8);
    uint64_t total;
    const uint64_t excl = block_exclusive_scan<uint64_t>((uint64_t)c, &total, red);
    const uint64_t Y0 = prefix[tile];
    int yr = (int)excl, nr = threadIdx.x * kItems - (int)excl;
    const int nbase = CODES ? 0 : (int)total;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        if (threadIdx.x * kItems + k >= cnt) break;
        if ((mask >> k) & 1) {
            if (CODES)
                cbuf[yr] = (uint8_t)((cpack >> (4 * k)) & 15);
            else
                buf[P(yr)] = v[k];
            yr++;
        } else {
            buf[P(nbase + nr)] = v[k];
            nr++;
        }
    }
    __syncthreads();
    if (CODES || yes)
        for (int i = threadIdx.x; i < (int)total; i += kBlock) {
            if (CODES)
                codes[Y0 + i] = cbuf[i];
            else
                yes[Y0 + i] = buf[P(i)];
        }
    if (WRITE_NO) {
        const size_t N0 = base - Y0;
        for (int i = threadIdx.x; i < cnt - (int)total; i += kBlock) no[N0 + i] = buf[P(nbase + i)];
    }
    if (threadIdx.x == 0 && base + cnt == n) d_count[0] = Y0 + total;
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

template <typename T, bool CONST_NO>
__global__ __launch_bounds__(kBlock) void k_merge(const uint8_t* __restrict__ bits, size_t n, const T* __restrict__ yes,
                                                  const T* __restrict__ no, T no_const, T* __restrict__ out,
                                                  const uint64_t* __restrict__ prefix) {
    __shared__ T buf[kTilePad];
    __shared__ uint64_t red[kBlock / kWave];
    const uint32_t tile = blockIdx.x;
    const size_t base = (size_t)tile * kTile;
    const int cnt = (int)((n - base) < (size_t)kTile ? (n - base) : (size_t)kTile);
    const size_t g0 = base + (size_t)threadIdx.x * kItems;
    uint32_t mask = 0;
    if (g0 < n) mask = bits[g0 / 8];
    if (g0 + 8 < n) mask |= (uint32_t)bits[g0 / 8 + 1] << 8;
    const int mine = cnt - threadIdx.x * kItems;
    if (mine <= 0)
        mask = 0;
    else if (mine < kItems)
        mask &= (1u << mine) - 1u;
    uint64_t total;
    const uint64_t excl = block_exclusive_scan<uint64_t>((uint64_t)__popc(mask), &total, red);
    const uint64_t Y0 = prefix[tile];
    const size_t N0 = base - Y0;
    for (int i = threadIdx.x; i < cnt; i += kBlock) {
        if (i < (int)total)
            buf[P(i)] = yes[Y0 + i];
        else
            buf[P(i)] = CONST_NO ? no_const : no[N0 + (i - (int)total)];
    }
    __syncthreads();
    T v[kItems];
    int yr = (int)excl, nr = (int)total + threadIdx.x * kItems - (int)excl;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        if (threadIdx.x * kItems + k < cnt) v[k] = ((mask >> k) & 1) ? buf[P(yr++)] : buf[P(nr++)];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kItems; k++)
        if (threadIdx.x * kItems + k < cnt) buf[P(threadIdx.x * kItems + k)] = v[k];
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += kBlock) out[base + i] = buf[P(i)];
}

// _compressUniqueKernel word layout (reference compression/unique/unique_encoding.cu:48-76):
// code j of the m coded rows sits at bits [(j % P) * b, ...) of word j / P.
__global__ __launch_bounds__(kBlock) void k_pack_codes(const uint8_t* __restrict__ codes, size_t m, int b,
                                                       uint32_t* __restrict__ words) {
    const int P = 32 / b;
    const size_t nw = (m + P - 1) / P;
    const size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t w = (size_t)blockIdx.x * kBlock + threadIdx.x; w < nw; w += stride) {
        uint32_t r = 0;
        for (int i = 0; i < P; i++) {
            size_t j = w * P + i;
            if (j < m) r |= (uint32_t)codes[j] << (b * i);
        }
        words[w] = r;
    }
}

// _decompressUniqueKernel (unique_encoding.cu:183-204); words and the
// dictionary may sit at any byte offset.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_unpack_codes(const uint8_t* __restrict__ words, size_t m, int b,
                                                         const uint8_t* __restrict__ dict, T* __restrict__ out) {
    const int P = 32 / b;
    const uint32_t mask = b >= 32 ? ~0u : ((1u << b) - 1u);
    const size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t j = (size_t)blockIdx.x * kBlock + threadIdx.x; j < m; j += stride) {
        uint32_t w;
        __builtin_memcpy(&w, words + 4 * (j / P), 4);
        uint32_t code = (w >> (b * (j % P))) & mask;
        T v;
        __builtin_memcpy(&v, dict + sizeof(T) * code, sizeof(T));
        out[j] = v;
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
        const u32x4* s4 = reinterpret_cast<const u32x4*>(x + base);
#pragma unroll
        for (int q = 0; q < kTile / V / kBlock; q++) {
            const int i = threadIdx.x + kBlock * q;
            const u32x4 w = __builtin_nontemporal_load(s4 + i);
            T v[V];
            __builtin_memcpy(v, &w, 16);
            const size_t g = base + (size_t)V * i;
            // the row before this chunk ends the previous lane's chunk
            T prev = shfl_up_value(v[V - 1]);
            if ((threadIdx.x & (kWave - 1)) == 0) prev = g > 0 ? x[g - 1] : v[0];
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
    if (cnt == kTile && (reinterpret_cast<uintptr_t>(x + base) & 15) == 0) {
        // buf[i + 1] = x[base + i]; buf[0] = the row before the tile
        constexpr int V = 16 / sizeof(T);
        const u32x4* s4 = reinterpret_cast<const u32x4*>(x + base);
        for (int i = threadIdx.x; i < kTile / V; i += kBlock) {
            const u32x4 w = __builtin_nontemporal_load(s4 + i);
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
    const uint64_t s_prefix = prefix[tile];
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

__global__ __launch_bounds__(kBlock) void k_rle_lengths(const uint32_t* __restrict__ starts, size_t runs, size_t n,
                                                        int32_t* __restrict__ lengths) {
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
constexpr int kExpTile = 2048;

// tile_run[t] = the run holding row t * kExpTile (offsets are non-decreasing:
// lengths were checked non-negative by the scan); tile_run[tiles] = R.
__global__ __launch_bounds__(kBlock) void k_rle_tile_runs(const uint64_t* __restrict__ offsets, size_t R, size_t tiles,
                                                          uint32_t* __restrict__ tile_run) {
    const size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t r = (size_t)blockIdx.x * kBlock + threadIdx.x; r < R; r += stride) {
        const uint64_t s = offsets[r], e = offsets[r + 1];
        const uint64_t t0 = (s + kExpTile - 1) / kExpTile, t1 = (e + kExpTile - 1) / kExpTile;  // tiles starting in [s, e)
        for (uint64_t t = t0; t < t1 && t < tiles; t++) tile_run[t] = (uint32_t)r;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) tile_run[tiles] = (uint32_t)R;
}

// Run expansion: a workgroup owns kExpTile output rows; its runs are
// tile_run[b] .. tile_run[b + 1] (the run holding the next tile's first row).
// They are staged in LDS and every row finds its run by binary search there;
// a tile overlapping more runs than LDS holds (only possible with empty runs)
// searches the offsets in global memory instead.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_rle_expand(const uint64_t* __restrict__ offsets, size_t R,
                                                       const uint8_t* __restrict__ values, size_t total,
                                                       const uint32_t* __restrict__ tile_run, T* __restrict__ out) {
    __shared__ uint64_t soff[kExpTile + 2];
    __shared__ T sval[kExpTile + 1];
    const size_t o0 = (size_t)blockIdx.x * kExpTile;
    const int cnt = (int)((total - o0) < (size_t)kExpTile ? (total - o0) : (size_t)kExpTile);
    const size_t r0 = tile_run[blockIdx.x];
    size_t r1 = tile_run[blockIdx.x + 1];  // last run needed (inclusive)
    r1 = r1 < R ? r1 : R - 1;
    r1 = r1 < r0 ? r0 : r1;
    const size_t nr = r1 - r0 + 1;
    if (nr <= (size_t)kExpTile + 1) {
        for (int i = threadIdx.x; i <= (int)nr; i += kBlock) soff[i] = offsets[r0 + i];
        for (int i = threadIdx.x; i < (int)nr; i += kBlock) __builtin_memcpy(&sval[i], values + sizeof(T) * (r0 + i), sizeof(T));
        __syncthreads();
        for (int k = threadIdx.x; k < cnt; k += kBlock) {
            const uint64_t pos = o0 + k;
            int lo = 0, hi = (int)nr;  // soff[lo] <= pos < soff[hi]
            while (hi - lo > 1) {
                const int mid = (lo + hi) / 2;
                if (soff[mid] <= pos)
                    lo = mid;
                else
                    hi = mid;
            }
            out[pos] = sval[lo];
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

void launch_split(Ctx* c, int dtype, const void* x, size_t n, const PredArgs& pa, uint8_t* stencil, void* yes,
                  void* no, uint8_t* codes, uint64_t* d_counts) {
    if (n == 0) return;
    dispatch_native(dtype, [&](auto tag) {
        using T = decltype(tag);
        Pred<T> p;
        p.kind = pa.kind;
        p.a = bits_to<T>(pa.a_bits);
        p.b = bits_to<T>(pa.b_bits);
        for (int j = 0; j < 4; j++) p.set[j] = bits_to<T>(pa.set_bits[j]);
        p.set_n = pa.set_n;
        const size_t tiles = (n + kTile - 1) / kTile;
        TileScan ts(c, tiles);
        auto xs = static_cast<const T*>(x);
        auto ys = static_cast<T*>(yes);
        auto ns = static_cast<T*>(no);
        kernel_bytes(c, "k_split_count", (double)n * sizeof(T));
        { KernelTimer kt_(c, "k_split_count"); hipLaunchKernelGGL(k_split_count<T>, dim3((unsigned)tiles), dim3(kBlock), 0, c->stream, xs, n, p, ts.vals); }
        HIPC(hipGetLastError());
        ts.scan(c);
        const uint64_t* pre = ts.prefix;
        if (codes) {
            { KernelTimer kt_(c, "k_split"); hipLaunchKernelGGL((k_split<T, true, true>), dim3((unsigned)tiles), dim3(kBlock), 0, c->stream, xs, n, p,
                               stencil, ys, ns, codes, pre, d_counts); }
        } else if (no) {
            { KernelTimer kt_(c, "k_split"); hipLaunchKernelGGL((k_split<T, false, true>), dim3((unsigned)tiles), dim3(kBlock), 0, c->stream, xs, n, p,
                               stencil, ys, ns, codes, pre, d_counts); }
        } else {
            { KernelTimer kt_(c, "k_split"); hipLaunchKernelGGL((k_split<T, false, false>), dim3((unsigned)tiles), dim3(kBlock), 0, c->stream, xs, n,
                               p, stencil, ys, ns, codes, pre, d_counts); }
        }
        HIPC(hipGetLastError());
    });
}

void launch_merge(Ctx* c, int dtype, const uint8_t* bits, size_t n, const uint8_t* yes, const uint8_t* no,
                  uint64_t no_const, void* out) {
    if (n == 0) return;
    dispatch_native(dtype, [&](auto tag) {
        using T = decltype(tag);
        const size_t tiles = (n + kTile - 1) / kTile;
        TileScan ts(c, tiles);
        const unsigned cg = (unsigned)((tiles + kBlock / kWave - 1) / (kBlock / kWave));
        kernel_bytes(c, "k_stencil_count", (double)((n + 7) / 8));
        kernel_bytes(c, "k_merge", (double)((n + 7) / 8) + 2.0 * (double)n * sizeof(T));
        { KernelTimer kt_(c, "k_stencil_count"); hipLaunchKernelGGL(k_stencil_count, dim3(cg), dim3(kBlock), 0, c->stream, bits, n, tiles, ts.vals); }
        HIPC(hipGetLastError());
        ts.scan(c);
        const uint64_t* pre = ts.prefix;
        auto ys = reinterpret_cast<const T*>(yes);
        auto ns = reinterpret_cast<const T*>(no);
        if (no)
            { KernelTimer kt_(c, "k_merge"); hipLaunchKernelGGL((k_merge<T, false>), dim3((unsigned)tiles), dim3(kBlock), 0, c->stream, bits, n, ys,
                               ns, T(0), static_cast<T*>(out), pre); }
        else
            { KernelTimer kt_(c, "k_merge"); hipLaunchKernelGGL((k_merge<T, true>), dim3((unsigned)tiles), dim3(kBlock), 0, c->stream, bits, n, ys,
                               ns, bits_to<T>(no_const), static_cast<T*>(out), pre); }
        HIPC(hipGetLastError());
    });
}

void launch_pack_codes(Ctx* c, const uint8_t* codes, size_t m, int bits, uint32_t* words) {
    if (!m) return;
    const size_t nw = (m + (32 / bits) - 1) / (32 / bits);
    { KernelTimer kt_(c, "k_pack_codes"); hipLaunchKernelGGL(k_pack_codes, dim3(grid_for(c, nw)), dim3(kBlock), 0, c->stream, codes, m, bits, words); }
    HIPC(hipGetLastError());
}

void launch_unpack_codes(Ctx* c, int dtype, const uint8_t* words, size_t m, int bits, const uint8_t* dict, void* out) {
    if (!m) return;
    dispatch_native(dtype, [&](auto tag) {
        using T = decltype(tag);
        { KernelTimer kt_(c, "k_unpack_codes"); hipLaunchKernelGGL(k_unpack_codes<T>, dim3(grid_for(c, m)), dim3(kBlock), 0, c->stream, words, m, bits, dict,
                           static_cast<T*>(out)); }
        HIPC(hipGetLastError());
    });
}

void launch_rle_encode(Ctx* c, int dtype, const void* x, size_t n, int32_t* lengths, void* values, uint64_t* d_runs) {
    if (n == 0) return;
    if (n >= (1ull << 31)) fail(DDJ_EUNSUPPORTED, "RLE run lengths are int32 (reference rle_encoding.hpp:35)");
    dispatch_native(dtype, [&](auto tag) {
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
        uint64_t runs = 0;
        d2h(c, &runs, d_runs, 8);
        { KernelTimer kt_(c, "k_rle_lengths"); hipLaunchKernelGGL(k_rle_lengths, dim3(grid_for(c, runs)), dim3(kBlock), 0, c->stream, starts.get<uint32_t>(),
                           runs, n, lengths); }
        HIPC(hipGetLastError());
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
        uint64_t total = 0;
        d2h(c, &total, offsets.get<uint64_t>() + runs, 8);
        DevBuf out(c, total * sizeof(T));
        if (total == 0) return out;
        const size_t etiles = (total + kExpTile - 1) / kExpTile;
        DevBuf tile_run(c, 4 * (etiles + 1));
        { KernelTimer kt_(c, "k_rle_tile_runs"); hipLaunchKernelGGL(k_rle_tile_runs, dim3(grid_for(c, runs)), dim3(kBlock), 0, c->stream,
                           offsets.get<uint64_t>(), runs, etiles, tile_run.get<uint32_t>()); }
        kernel_bytes(c, "k_rle_expand", (double)total * sizeof(T) + (double)runs * (8 + sizeof(T)));
        { KernelTimer kt_(c, "k_rle_expand"); hipLaunchKernelGGL(k_rle_expand<T>, dim3((unsigned)etiles), dim3(kBlock), 0, c->stream,
                           offsets.get<uint64_t>(), runs, values, (size_t)total, tile_run.get<uint32_t>(), out.get<T>()); }
        HIPC(hipGetLastError());
        return out;
    });
}

}  // namespace ddj
