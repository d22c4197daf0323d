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

// Inclusive prefix sum over the 64 lanes of a wave.
template <typename T> __device__ __forceinline__ T wave_inclusive_scan(T v) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        T w = __shfl_up(v, o, kWave);
        if (lane >= o) v += w;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread (kBlock threads).
// Returns the exclusive prefix; *total receives the block sum.
template <typename T> __device__ __forceinline__ T block_exclusive_scan(T v, T* total, T* smem /* kBlock/kWave */) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    T inc = wave_inclusive_scan(v);
    if (lane == kWave - 1) smem[wid] = inc;
    __syncthreads();
    T wave_off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; w++) {
        T s = smem[w];
        if (w < wid) wave_off += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return wave_off + inc - v;
}

// ------------------------------------------------ decoupled look-back scan
// Tiles take tickets in start order, publish their aggregate, and add up the
// predecessors' values until one with an inclusive prefix is found.
//
// A tile's state and value travel together in one 64-bit status word
// (bits 63..62 state, 61..0 value), so a relaxed agent-scope atomic load sees
// a consistent pair and no acquire/release fence is needed: such fences make
// the compiler write back and invalidate the whole L2 of the XCD
// (buffer_wbl2 / buffer_inv), which stalls every tile of a streaming kernel.
// Wide values (full 64-bit wrapping sums) use two status words carrying the
// 32-bit halves; a pair is taken only when both words show the same state.
//
// One wave inspects 64 predecessors per step.  Every wait is bounded: on
// expiry the error flag is raised and the missing prefix counts as zero, so
// a kernel always finishes.
constexpr uint64_t kStAggregate = 1ull << 62, kStInclusive = 2ull << 62;
constexpr uint64_t kStValue = (1ull << 62) - 1;
constexpr uint32_t kSpinLimit = 1u << 22;

__device__ __forceinline__ uint64_t ld_status(uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <bool WIDE> __device__ __forceinline__ void lb_publish(const Lookback& lb, uint32_t tile, uint64_t state,
                                                                uint64_t v) {
    if constexpr (WIDE) {
        st_status(&lb.status[2 * (size_t)tile], state | (v & 0xffffffffull));
        st_status(&lb.status[2 * (size_t)tile + 1], state | (v >> 32));
    } else {
        st_status(&lb.status[tile], state | (v & kStValue));
    }
}

// Reads predecessor `idx`; returns false while it has published nothing.
template <bool WIDE> __device__ __forceinline__ bool lb_read(const Lookback& lb, size_t idx, uint64_t& state,
                                                             uint64_t& v) {
    if constexpr (WIDE) {
        const uint64_t a = ld_status(&lb.status[2 * idx]);
        const uint64_t b = ld_status(&lb.status[2 * idx + 1]);
        state = a & ~kStValue;
        v = (b << 32) | (a & 0xffffffffull);
        return state != 0 && state == (b & ~kStValue);
    } else {
        const uint64_t a = ld_status(&lb.status[idx]);
        state = a & ~kStValue;
        v = a & kStValue;
        return state != 0;
    }
}

// Exclusive prefix of `tile`.  Called by all 64 lanes of one wave with the
// same tile and aggregate.  Narrow values must stay below 2^62 (counts), or
// only their low bits matter (sums that wrap at 32 bits or fewer).
template <bool WIDE> __device__ uint64_t lookback_wave(const Lookback& lb, uint32_t tile, uint64_t aggregate) {
    const int lane = threadIdx.x & (kWave - 1);
    if (tile == 0) {
        if (lane == 0) lb_publish<WIDE>(lb, 0, kStInclusive, aggregate);
        return 0;
    }
    if (lane == 0) lb_publish<WIDE>(lb, tile, kStAggregate, aggregate);
    uint64_t excl = 0;
    int64_t end = tile;  // window [end - 64, end)
    for (;;) {
        const int64_t idx = end - 1 - lane;
        uint64_t state = kStInclusive, v = 0;  // before tile 0: an inclusive zero
        bool ready = idx < 0;
        uint32_t spins = 0;
        for (;;) {
            if (!ready) ready = lb_read<WIDE>(lb, (size_t)idx, state, v);
            if (__all(ready)) break;
            if (++spins > kSpinLimit) {
                if (!ready) {
                    __hip_atomic_store(lb.error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    state = kStInclusive;
                    v = 0;
                }
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        const uint64_t inc = __ballot(state == kStInclusive);
        if (inc) {
            const int first = __builtin_ctzll(inc);
            excl += wave_sum<uint64_t>(lane <= first ? v : 0);
            break;
        }
        excl += wave_sum<uint64_t>(v);
        end -= kWave;
    }
    if (lane == 0) lb_publish<WIDE>(lb, tile, kStInclusive, excl + aggregate);
    return excl;
}

// Ticket of the calling workgroup (all threads get it).
__device__ __forceinline__ uint32_t take_ticket(const Lookback& lb, uint32_t* smem) {
    if (threadIdx.x == 0) *smem = atomicAdd(lb.ticket, 1u);
    __syncthreads();
    uint32_t t = *smem;
    __syncthreads();
    return t;
}

// Block-wide look-back: wave 0 resolves the prefix, every thread gets it.
template <bool WIDE> __device__ __forceinline__ uint64_t block_lookback(const Lookback& lb, uint32_t tile,
                                                                       uint64_t aggregate, uint64_t* smem) {
    if (threadIdx.x < kWave) {
        const uint64_t p = lookback_wave<WIDE>(lb, tile, aggregate);
        if (threadIdx.x == 0) *smem = p;
    }
    __syncthreads();
    return *smem;
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
