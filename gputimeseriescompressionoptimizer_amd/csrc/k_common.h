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
// Tiles publish their aggregate and add up the predecessors' values until
// one with an inclusive prefix is found.
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

// Exclusive prefix of `tile`, resolved by the whole workgroup (kBlock
// threads, all of which call it with the same tile and aggregate).  Each
// step inspects the kLbWindow = kBlock * kLbPerThread predecessors before the
// current window (thread t, entry j: tile end - 1 - t - kBlock * j), all
// loads in flight together.  The window must be wide: a tile only turns
// inclusive after its own look-back, so in steady state the nearest
// inclusive predecessor lies about (tiles started per microsecond) x (one
// device round trip) tiles back, and tiles resolve at most one window per
// round trip.  Narrow values must stay below 2^62 (counts), or only their low
// bits matter (sums that wrap at 32 bits or fewer).
constexpr int kLbPerThread = 8;
constexpr int kLbWindow = kBlock * kLbPerThread;

template <bool WIDE> __device__ uint64_t block_lookback(const Lookback& lb, uint32_t tile, uint64_t aggregate) {
    __shared__ uint64_t red[kBlock / kWave];
    __shared__ uint32_t first_inc[kBlock / kWave];
    const int t = threadIdx.x, lane = t & (kWave - 1), wid = t / kWave;
    if (tile == 0) {
        if (t == 0) lb_publish<WIDE>(lb, 0, kStInclusive, aggregate);
        __syncthreads();  // the call is a barrier on every path
        return 0;
    }
    if (t == 0) lb_publish<WIDE>(lb, tile, kStAggregate, aggregate);
    uint64_t excl = 0;
    int64_t end = tile;  // window [end - kLbWindow, end)
    for (;;) {
        uint64_t state[kLbPerThread], v[kLbPerThread];
        bool ready[kLbPerThread];
#pragma unroll
        for (int j = 0; j < kLbPerThread; j++) {
            state[j] = kStInclusive;  // before tile 0: an inclusive zero
            v[j] = 0;
            ready[j] = end - 1 - t - (int64_t)kBlock * j < 0;
        }
        uint32_t spins = 0;
        for (;;) {
            bool all = true;
#pragma unroll
            for (int j = 0; j < kLbPerThread; j++) {
                if (!ready[j]) ready[j] = lb_read<WIDE>(lb, (size_t)(end - 1 - t - (int64_t)kBlock * j), state[j], v[j]);
                all = all && ready[j];
            }
            if (__syncthreads_and(all)) break;
            if (++spins > kSpinLimit) {
#pragma unroll
                for (int j = 0; j < kLbPerThread; j++)
                    if (!ready[j]) {
                        __hip_atomic_store(lb.error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        state[j] = kStInclusive;
                        v[j] = 0;
                    }
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        // the nearest inclusive predecessor has the smallest distance t + kBlock * j
        uint32_t mine = 0xffffffffu;
#pragma unroll
        for (int j = kLbPerThread - 1; j >= 0; j--)
            if (state[j] == kStInclusive) mine = (uint32_t)(t + kBlock * j);
        uint32_t wmin = mine;
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) {
            const uint32_t w = __shfl_xor(wmin, o, kWave);
            wmin = w < wmin ? w : wmin;
        }
        if (lane == 0) first_inc[wid] = wmin;
        __syncthreads();
        uint32_t first = 0xffffffffu;
#pragma unroll
        for (int w = 0; w < kBlock / kWave; w++) first = first_inc[w] < first ? first_inc[w] : first;
        uint64_t part = 0;
#pragma unroll
        for (int j = 0; j < kLbPerThread; j++)
            if ((uint32_t)(t + kBlock * j) <= first) part += v[j];
        const uint64_t ws = wave_sum<uint64_t>(part);
        if (lane == 0) red[wid] = ws;
        __syncthreads();
        uint64_t sum = 0;
#pragma unroll
        for (int w = 0; w < kBlock / kWave; w++) sum += red[w];
        excl += sum;
        __syncthreads();  // red / first_inc are rewritten by the next step
        if (first != 0xffffffffu) break;
        end -= kLbWindow;
    }
    if (t == 0) lb_publish<WIDE>(lb, tile, kStInclusive, excl + aggregate);
    return excl;
}

// Tile of the calling workgroup.  Workgroups are dispatched in increasing
// blockIdx order on every XCD (round-robin over the XCDs), so the lowest
// unfinished tile is always resident and the look-back always progresses;
// the bounded waits cover anything else.  A ticket counter (tiles numbered
// in start order) would cost one device-scope atomic on a single address per
// tile, and those serialise at roughly 15-20 ns each: 16K tiles of a 64M-row
// column would spend ~300 us there.
__device__ __forceinline__ uint32_t tile_index() { return blockIdx.x; }

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
