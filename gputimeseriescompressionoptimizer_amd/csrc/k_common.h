// Device helpers shared by the HIP kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "kernels.h"

namespace ddj {

constexpr int kWave = 64;

// Native 16-byte vector (works with the nontemporal load/store builtins).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr int kBlock = 256;  // 4 waves per workgroup

// ------------------------------------------------------------ type helpers
template <typename T> struct UnsignedOf;
template <> struct UnsignedOf<int8_t> { using type = uint8_t; };
template <> struct UnsignedOf<int16_t> { using type = uint16_t; };
template <> struct UnsignedOf<int32_t> { using type = uint32_t; };
template <> struct UnsignedOf<int64_t> { using type = uint64_t; };
template <> struct UnsignedOf<uint32_t> { using type = uint32_t; };
template <> struct UnsignedOf<uint64_t> { using type = uint64_t; };
template <> struct UnsignedOf<float> { using type = uint32_t; };
template <> struct UnsignedOf<double> { using type = uint64_t; };

template <typename T> __device__ __forceinline__ uint64_t ubits(T v) {
    using U = typename UnsignedOf<T>::type;
    U u;
    __builtin_memcpy(&u, &v, sizeof(T));
    return (uint64_t)u;
}

// Order-preserving map of a value to an unsigned key (min/max by atomics or
// integer compares).  Floats: flip the sign bit for positives, all bits for
// negatives.
template <typename T> __device__ __forceinline__ uint64_t order_key(T v) {
    if constexpr (std::is_floating_point<T>::value) {
        uint64_t b = ubits(v);
        constexpr int W = 8 * sizeof(T);
        const uint64_t sign = 1ull << (W - 1);
        const uint64_t mask = (W == 64) ? ~0ull : ((1ull << W) - 1);
        return (b & sign) ? (~b & mask) : (b | sign);
    } else {
        constexpr int W = 8 * sizeof(T);
        if constexpr (W == 64)
            return (uint64_t)(int64_t)v ^ (1ull << 63);
        else
            return (uint64_t)((int64_t)v + (1ll << (W - 1)));
    }
}
template <typename T> __device__ __host__ __forceinline__ T from_order_key(uint64_t k) {
    if constexpr (std::is_floating_point<T>::value) {
        constexpr int W = 8 * sizeof(T);
        const uint64_t sign = 1ull << (W - 1);
        const uint64_t mask = (W == 64) ? ~0ull : ((1ull << W) - 1);
        uint64_t b = (k & sign) ? (k & ~sign) : (~k & mask);
        T v;
        using U = typename UnsignedOf<T>::type;
        U u = (U)b;
        __builtin_memcpy(&v, &u, sizeof(T));
        return v;
    } else {
        constexpr int W = 8 * sizeof(T);
        if constexpr (W == 64)
            return (T)(int64_t)(k ^ (1ull << 63));
        else
            return (T)((int64_t)k - (1ll << (W - 1)));
    }
}

// ---------------------------------------------------------- wave reductions
template <typename T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
template <typename T> __device__ __forceinline__ T wave_min(T v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, kWave);
        v = w < v ? w : v;
    }
    return v;
}
template <typename T> __device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, kWave);
        v = w > v ? w : v;
    }
    return v;
}
__device__ __forceinline__ uint64_t wave_or(uint64_t v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v |= __shfl_xor(v, o, kWave);
    return v;
}

// Minimum (MAX: maximum) of a 32-bit value over the 64 lanes of a wave:
// DPP steps inside every row of 16 lanes (swap neighbours, swap pairs, mirror
// half-rows, mirror rows), then the four rows' results by readlane.  The
// result is the same in every lane.
template <bool MAX> __device__ __forceinline__ int32_t wave_extreme_i32(int32_t v) {
    auto op = [](int32_t a, int32_t b) { return MAX ? (a > b ? a : b) : (a < b ? a : b); };
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, false));  // row_half_mirror
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, false));  // row_mirror
    const int32_t r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    const int32_t r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    return op(op(r0, r1), op(r2, r3));
}

// v shifted across lanes by a DPP control (lanes the control or the row /
// bank masks leave without a source read 0), for 32- or 64-bit integers.
template <int CTRL, int ROWM, int BANKM, typename T> __device__ __forceinline__ T dpp_shift(T v) {
    if constexpr (sizeof(T) == 8) {
        const uint64_t u = (uint64_t)v;
        const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, ROWM, BANKM, false);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, ROWM, BANKM, false);
        return (T)(((uint64_t)hi << 32) | lo);
    } else {
        return (T)(uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, ROWM, BANKM, false);
    }
}

// Inclusive prefix sum over the 64 lanes of a wave (32- or 64-bit
// integers), by DPP: inside each row of 16 lanes, the sum of the last four
// lanes (row_shr 1, 2, 3 of the input), then of the last eight (row_shr 4
// into banks 1-3) and sixteen (row_shr 8 into banks 2-3); across rows,
// lane 15 into row 1 and lane 47 into row 3 (row_bcast 15), then lane 31
// into rows 2 and 3 (row_bcast 31).
template <typename T> __device__ __forceinline__ T wave_inclusive_scan(T v) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
    T x = v + dpp_shift<0x111, 0xF, 0xF>(v);  // row_shr:1
    x += dpp_shift<0x112, 0xF, 0xF>(v);       // row_shr:2
    x += dpp_shift<0x113, 0xF, 0xF>(v);       // row_shr:3
    x += dpp_shift<0x114, 0xF, 0xE>(x);       // row_shr:4, banks 1-3
    x += dpp_shift<0x118, 0xF, 0xC>(x);       // row_shr:8, banks 2-3
    x += dpp_shift<0x142, 0xA, 0xF>(x);       // row_bcast:15, rows 1 and 3
    x += dpp_shift<0x143, 0xC, 0xF>(x);       // row_bcast:31, rows 2 and 3
    return x;
}

// Block-wide exclusive scan of one value per thread (NT threads).
// Returns the exclusive prefix; *total receives the block sum.
template <typename T, int NT = kBlock>
__device__ __forceinline__ T block_exclusive_scan(T v, T* total, T* smem /* NT/kWave */) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    T inc = wave_inclusive_scan(v);
    if (lane == kWave - 1) smem[wid] = inc;
    __syncthreads();
    T wave_off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NT / kWave; w++) {
        T s = smem[w];
        if (w < wid) wave_off += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return wave_off + inc - v;
}

// ------------------------------------------------------ reduce, then scan
// Variable-length outputs (split, merge, RLE, delta decode, GFC) take their
// tile offsets from two cheap launches ahead of the output pass: a counting
// pass writes one value per tile, k_scan_tiles (one workgroup) turns them
// into exclusive prefixes, and the output pass reads prefix[tile].  No
// workgroup waits on another, so nothing spins.
constexpr int kScanThreads = 1024;
constexpr int kScanRows = 16;  // rows of kScanThreads values per step of k_scan_tiles

// Relaxed agent-scope stores and loads for values that one workgroup
// publishes and another reads within the same launch.
__device__ __forceinline__ void st_agent(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// "Last workgroup finishes" without acquire/release fences (which would
// write back and invalidate the XCD's whole L2 once per workgroup): every
// thread has written its share of the partial with st_agent; the waitcnt
// makes those stores complete at the device coherence point before the
// barrier, and only then does thread 0 take a ticket.  The last workgroup
// reads the partials with relaxed agent-scope loads, which go to that point.
__device__ __forceinline__ bool last_block_arrives(uint32_t* counter, bool* sm_flag) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        *sm_flag = (prev == gridDim.x - 1);
    }
    __syncthreads();
    return *sm_flag;
}

// -------------------------------------------------------------- bit helpers
__device__ __host__ __forceinline__ uint64_t lowmask64(unsigned n) { return n >= 64 ? ~0ull : ((1ull << n) - 1ull); }

// BITLEN (reference core/cuda_macros.cuh:168-175) from min and max.
template <typename T> __device__ __host__ __forceinline__ int bitlen_of(T mn, T mx) {
    if constexpr (std::is_floating_point<T>::value) {
        return 8 * sizeof(T);
    } else {
        if (mn < 0) return 8 * sizeof(T);
        uint64_t m = (uint64_t)mx;
        int r = 0;
        while (m >>= 1) r++;
        return r + 1;
    }
}

}  // namespace ddj
