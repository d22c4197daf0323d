// Elementwise kernels: delta, scale, float split/compose (AFL on floats),
// floatToInt and its inverse, zero-fill, gathered copies; and the delta
// decode scan (single pass, decoupled look-back).
#include "dispatch.h"
#include "k_common.h"
#include "kernels.h"

namespace ddj {

namespace {

inline unsigned grid_for(Ctx* c, size_t work_items, int per_thread = 1) {
    size_t threads = (work_items + per_thread - 1) / per_thread;
    size_t g = (threads + kBlock - 1) / kBlock;
    size_t cap = (size_t)c->cu_count * 32;
    if (g > cap) g = cap;
    if (g == 0) g = 1;
    return (unsigned)g;
}

// DeltaEncoding deltaEncodeKernel (reference compression/delta/delta_encoding.cu:19-27):
// out[i] = x[i+1] - x[i], wrapping in T.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_delta_encode(const T* __restrict__ x, size_t n, T* __restrict__ out) {
    constexpr int V = 16 / sizeof(T);
    const size_t m = n - 1;
    const size_t stride = (size_t)gridDim.x * kBlock;
    const bool vec = ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
    const size_t nv = vec ? m / V : 0;
    for (size_t g = (size_t)blockIdx.x * kBlock + threadIdx.x; g < nv; g += stride) {
        const size_t i0 = g * V;
        u32x4 a = reinterpret_cast<const u32x4*>(x)[g];
        T v[V + 1];
        __builtin_memcpy(v, &a, 16);
        v[V] = x[i0 + V];  // i0 + V <= m < n
        T d[V];
#pragma unroll
        for (int k = 0; k < V; k++) d[k] = (T)(ubits(v[k + 1]) - ubits(v[k]));
        u32x4 o;
        __builtin_memcpy(&o, d, 16);
        __builtin_nontemporal_store(o, reinterpret_cast<u32x4*>(out) + g);
    }
    for (size_t i = nv * V + (size_t)blockIdx.x * kBlock + threadIdx.x; i < m; i += stride)
        out[i] = (T)(ubits(x[i + 1]) - ubits(x[i]));
}

// x +/- s wrapping in T (scaleEncodeKernel / scaleDecodeKernel,
// reference compression/scale/scale_encoding.cu:21-28, :70-76).
template <typename T, bool ADD>
__global__ __launch_bounds__(kBlock) void k_add_scalar(const T* __restrict__ x, size_t n, uint64_t s,
                                                       T* __restrict__ out) {
    constexpr int V = 16 / sizeof(T);
    const size_t stride = (size_t)gridDim.x * kBlock;
    const bool vec = ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
    const size_t nv = vec ? n / V : 0;
    for (size_t g = (size_t)blockIdx.x * kBlock + threadIdx.x; g < nv; g += stride) {
        u32x4 a = reinterpret_cast<const u32x4*>(x)[g];
        T v[V];
        __builtin_memcpy(v, &a, 16);
#pragma unroll
        for (int k = 0; k < V; k++) v[k] = ADD ? (T)(ubits(v[k]) + s) : (T)(ubits(v[k]) - s);
        __builtin_memcpy(&a, v, 16);
        __builtin_nontemporal_store(a, reinterpret_cast<u32x4*>(out) + g);
    }
    for (size_t i = nv * V + (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
        out[i] = ADD ? (T)(ubits(x[i]) + s) : (T)(ubits(x[i]) - s);
}

// _splitFloatKernel (reference compression/afl/afl_encoding.cu:58-72)
__global__ __launch_bounds__(kBlock) void k_float_split(const float* __restrict__ x, size_t n, int32_t* mant,
                                                        int32_t* expo, int32_t* sign) {
    const size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        uint32_t u = __float_as_uint(x[i]);
        mant[i] = (int32_t)(u & 0x7fffffu);
        expo[i] = (int32_t)((u >> 23) & 0xffu);
        sign[i] = (int32_t)(u >> 31);
    }
}

// _composeFloatKernel (afl_encoding.cu:169-183).  The sign comes from a
// packed stencil, or is constant (sign_const 0/1) when the stencil is null.
__global__ __launch_bounds__(kBlock) void k_float_compose(const int32_t* __restrict__ mant,
                                                          const int32_t* __restrict__ expo, const uint8_t* bits,
                                                          int sign_const, size_t n, float* __restrict__ out) {
    const size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        uint32_t s = bits ? ((bits[i >> 3] >> (i & 7)) & 1u) : (uint32_t)sign_const;
        uint32_t u = (s << 31) | (((uint32_t)expo[i] & 0xffu) << 23) | ((uint32_t)mant[i] & 0x7fffffu);
        out[i] = __uint_as_float(u);
    }
}

__constant__ double kPow10[11] = {1e0, 1e1, 1e2, 1e3, 1e4, 1e5, 1e6, 1e7, 1e8, 1e9, 1e10};

// FloatingPointToIntegerOperator (reference util/transform/transform_operators.hpp:109-119):
// int(lrint(x * 10^p)) with the product in double.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_float_to_int(const T* __restrict__ x, size_t n, int p,
                                                         int32_t* __restrict__ out) {
    const double f = kPow10[p];
    const size_t stride = (size_t)gridDim.x * kBlock;
    auto conv = [&](T v) -> int32_t { return (int32_t)(long)__builtin_rint(__dmul_rn((double)v, f)); };
    // 4 rows per thread and step: 16-byte loads and stores
    // (4- and 8-byte T; narrower T takes the row loop below)
    constexpr int L = sizeof(T) >= 4 ? (int)sizeof(T) / 4 : 1;  // 16-byte loads per 4 rows
    const bool vec = sizeof(T) >= 4 && ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
    const size_t n4 = vec ? n / 4 : 0;
    for (size_t g = (size_t)blockIdx.x * kBlock + threadIdx.x; g < n4; g += stride) {
        u32x4 a[L];
#pragma unroll
        for (int l = 0; l < L; l++) a[l] = reinterpret_cast<const u32x4*>(x + 4 * g)[l];
        T v[4] = {};
        if constexpr (sizeof(T) >= 4) __builtin_memcpy(v, a, 16 * L);
        int32_t r[4];
#pragma unroll
        for (int k = 0; k < 4; k++) r[k] = conv(v[k]);
        u32x4 o;
        __builtin_memcpy(&o, r, 16);
        __builtin_nontemporal_store(o, reinterpret_cast<u32x4*>(out + 4 * g));
    }
    for (size_t i = 4 * n4 + (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) out[i] = conv(x[i]);
}

// IntegerToFloatingPointOperator (transform_operators.hpp:121-131): (T)x / 10^p in double.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_int_to_float(const int32_t* __restrict__ x, size_t n, int p,
                                                         T* __restrict__ out) {
    const double f = kPow10[p];
    const size_t stride = (size_t)gridDim.x * kBlock;
    auto conv = [&](int32_t v) -> T { return (T)__ddiv_rn((double)(T)v, f); };
    // (4-byte T: measured slower for double, whose row loop already stores
    // 8 bytes per row)
    constexpr int L = sizeof(T) >= 4 ? (int)sizeof(T) / 4 : 1;  // 16-byte stores per 4 rows
    const bool vec = sizeof(T) == 4 && ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
    const size_t n4 = vec ? n / 4 : 0;
    for (size_t g = (size_t)blockIdx.x * kBlock + threadIdx.x; g < n4; g += stride) {
        const u32x4 a = *reinterpret_cast<const u32x4*>(x + 4 * g);
        int32_t v[4];
        __builtin_memcpy(v, &a, 16);
        T r[4];
#pragma unroll
        for (int k = 0; k < 4; k++) r[k] = conv(v[k]);
        u32x4 o[L] = {};
        if constexpr (sizeof(T) >= 4) __builtin_memcpy(o, r, 16 * L);
#pragma unroll
        for (int l = 0; l < L; l++) __builtin_nontemporal_store(o[l], reinterpret_cast<u32x4*>(out + 4 * g) + l);
    }
    for (size_t i = 4 * n4 + (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) out[i] = conv(x[i]);
}

__global__ __launch_bounds__(kBlock) void k_fill_zero(uint32_t* p, size_t words) {
    const size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < words; i += stride) p[i] = 0u;
}

// One range per blockIdx.y.  Destinations sit at arbitrary byte offsets of
// the blob: a byte head brings the destination to 16-byte alignment, the
// body stores 16 bytes per thread-iteration built from aligned source words
// by funnel shifts (an aligned word holding a valid byte never crosses a
// page), and a byte tail finishes.
__global__ __launch_bounds__(kBlock) void k_gather_copy(const CopyRange* __restrict__ r, uint8_t* __restrict__ dst) {
    const CopyRange cr = r[blockIdx.y];
    const uint8_t* src = cr.src;
    uint8_t* d = dst + cr.dst_off;
    const uint64_t nb = cr.bytes;
    const size_t tid = (size_t)blockIdx.x * kBlock + threadIdx.x, stride = (size_t)gridDim.x * kBlock;
    uint64_t head = (16 - (reinterpret_cast<uintptr_t>(d) & 15)) & 15;
    head = head < nb ? head : nb;
    if (tid < head) d[tid] = src[tid];
    const uint64_t body = (nb - head) / 16;
    const uintptr_t sa = reinterpret_cast<uintptr_t>(src + head);
    const uint32_t sh = (uint32_t)(sa & 3) * 8;
    const uint32_t* w = reinterpret_cast<const uint32_t*>(sa & ~(uintptr_t)3);
    u32x4* d4 = reinterpret_cast<u32x4*>(d + head);
    if (sh == 0) {
        for (size_t i = tid; i < body; i += stride) {
            const uint32_t* p = w + 4 * i;
            u32x4 o = {p[0], p[1], p[2], p[3]};
            __builtin_nontemporal_store(o, d4 + i);
        }
    } else {
        for (size_t i = tid; i < body; i += stride) {
            const uint32_t* p = w + 4 * i;
            const uint32_t a0 = p[0], a1 = p[1], a2 = p[2], a3 = p[3], a4 = p[4];
            u32x4 o = {__builtin_amdgcn_alignbit(a1, a0, sh), __builtin_amdgcn_alignbit(a2, a1, sh),
                       __builtin_amdgcn_alignbit(a3, a2, sh), __builtin_amdgcn_alignbit(a4, a3, sh)};
            __builtin_nontemporal_store(o, d4 + i);
        }
    }
    const uint64_t t0 = head + body * 16;
    if (tid < nb - t0) d[t0 + tid] = src[t0 + tid];
}

__global__ __launch_bounds__(kBlock) void k_copy_pairs(const CopyPair* __restrict__ r) {
    const CopyPair cp = r[blockIdx.y];
    const size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < cp.bytes; i += stride) cp.dst[i] = cp.src[i];
}

// ------------------------------------------------------------- delta decode
// out[0] = x0; out[i+1] = x0 + d[0] + ... + d[i] (wrapping).  Tiles of 4096
// deltas: k_delta_sums writes each tile's sum, k_scan_tiles their prefixes,
// and k_delta_decode scans its tile on top of that prefix.  Sums run over
// the zero-extended bits of T in 64 bits; only their low 8*sizeof(T) bits
// are kept, which is the wrapping sum in T.
constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;

// Row i of a T stream at any byte address.
template <typename T> __device__ __forceinline__ T load_row(const uint8_t* p, size_t i) {
    T t;
    __builtin_memcpy(&t, p + i * sizeof(T), sizeof(T));
    return t;
}

// Small read-backs: the ranges go straight into the context's pinned
// landing area through its device address, all of them in one launch of
// one workgroup (the runtime's copy takes a launch per range, or for larger
// ranges a DMA transfer).  Every thread releases its stores at system scope
// (a release alone: the full fence would also invalidate the L2), then,
// after the barrier, thread 0 stores `seq` to the flag the host polls.
__global__ __launch_bounds__(kBlock) void k_to_host(ToHostPieces t, uint8_t* __restrict__ dst, uint32_t* flag,
                                                    uint32_t seq) {
    const int tid = threadIdx.x;
    for (int p = 0; p < t.count; p++) {  // uniform
        const uint8_t* s = t.src[p];
        uint8_t* d = dst + t.off[p];
        const uint32_t nb = t.bytes[p];
        const uintptr_t al = reinterpret_cast<uintptr_t>(s) | t.off[p] | nb;
        if ((al & 15) == 0) {
            for (uint32_t i = tid; i < nb / 16; i += kBlock)
                reinterpret_cast<u32x4*>(d)[i] = reinterpret_cast<const u32x4*>(s)[i];
        } else if ((al & 7) == 0) {
            for (uint32_t i = tid; i < nb / 8; i += kBlock)
                reinterpret_cast<uint64_t*>(d)[i] = reinterpret_cast<const uint64_t*>(s)[i];
        } else {
            for (uint32_t i = tid; i < nb; i += kBlock) d[i] = s[i];
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    __syncthreads();
    if (tid == 0) __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_delta_sums(const uint8_t* __restrict__ din, size_t m,
                                                       uint64_t* __restrict__ vals) {
    __shared__ uint64_t red[kBlock / kWave];
    const size_t base = (size_t)blockIdx.x * kScanTile;
    const int cnt = (int)((m - base) < (size_t)kScanTile ? (m - base) : (size_t)kScanTile);
    const uint8_t* tp = din + base * sizeof(T);
    uint64_t s = 0;
    if (cnt == kScanTile && (reinterpret_cast<uintptr_t>(tp) & 15) == 0) {
        constexpr int V = 16 / sizeof(T);
        const u32x4* s4 = reinterpret_cast<const u32x4*>(tp);
#pragma unroll
        for (int q = 0; q < kScanTile / V / kBlock; q++) {
            const u32x4 w = s4[threadIdx.x + kBlock * q];
            T v[V];
            __builtin_memcpy(v, &w, 16);
#pragma unroll
            for (int k = 0; k < V; k++) s += ubits(v[k]);
        }
    } else {
        for (int i = threadIdx.x; i < cnt; i += kBlock) s += ubits(load_row<T>(tp, i));
    }
    uint64_t total;
    block_exclusive_scan<uint64_t>(s, &total, red);
    if (threadIdx.x == 0) vals[blockIdx.x] = total;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_delta_decode(const uint8_t* first_bytes, const uint8_t* __restrict__ din,
                                                         size_t m, T* __restrict__ out,
                                                         const uint64_t* __restrict__ prefix) {
    T first;
    __builtin_memcpy(&first, first_bytes, sizeof(T));
    __shared__ uint64_t buf[kScanTile + kScanTile / 32];
    __shared__ uint64_t red[kBlock / kWave];
    const uint32_t tile = blockIdx.x;
    const size_t base = (size_t)tile * kScanTile;
    const int cnt = (int)((m - base) < (size_t)kScanTile ? (m - base) : (size_t)kScanTile);
    auto pad = [](int i) { return i + (i >> 5); };
    const uint8_t* tp = din + base * sizeof(T);
    if (cnt == kScanTile && (reinterpret_cast<uintptr_t>(tp) & 15) == 0) {
        constexpr int V = 16 / sizeof(T);
        const u32x4* s4 = reinterpret_cast<const u32x4*>(tp);
#pragma unroll
        for (int q = 0; q < kScanTile / V / kBlock; q++) {
            const int i = threadIdx.x + kBlock * q;
            const u32x4 w = s4[i];
            T v[V];
            __builtin_memcpy(v, &w, 16);
#pragma unroll
            for (int k = 0; k < V; k++) buf[pad(V * i + k)] = ubits(v[k]);
        }
    } else {
        for (int i = threadIdx.x; i < kScanTile; i += kBlock) buf[pad(i)] = i < cnt ? ubits(load_row<T>(tp, i)) : 0;
    }
    __syncthreads();
    uint64_t local[kScanItems];
    uint64_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        local[k] = buf[pad(threadIdx.x * kScanItems + k)];
        sum += local[k];
    }
    uint64_t total;
    const uint64_t excl = block_exclusive_scan<uint64_t>(sum, &total, red);  // a barrier: buf is free again
    uint64_t run = ubits(first) + prefix[tile] + excl;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        run += local[k];
        buf[pad(threadIdx.x * kScanItems + k)] = run;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += kBlock) out[base + i + 1] = (T)buf[pad(i)];
    if (tile == 0 && threadIdx.x == 0) out[0] = first;
}

// Exclusive prefixes of per-tile values; workgroup 0 of kScanThreads.
// A step covers kScanRows rows of kScanThreads consecutive values: thread t
// holds value t of every row (all loads and stores in flight together),
// each row is scanned across the workgroup (wave scans, then the waves'
// totals in LDS), and the rows are chained by their totals.
// With `keys`, workgroups 1..F fold per-tile {min, max, min, max} order keys
// into *mm_a and *mm_b while workgroup 0 scans: each folds a slice into
// key_part, and the last to finish (a ticket on `ticket`, F arrivals, reset
// by that workgroup) folds the F partials.
__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(const uint64_t* __restrict__ vals, size_t tiles,
                                                             uint64_t* __restrict__ prefix,
                                                             const uint64_t* __restrict__ keys, MinMaxResult* mm_a,
                                                             MinMaxResult* mm_b, uint64_t* key_part, uint32_t* ticket) {
    constexpr int R = kScanRows, NW = kScanThreads / kWave;
    __shared__ uint64_t wsum[R][NW];  // per row: the waves' totals, then their exclusive offsets
    __shared__ uint64_t rowtot[R];
    __shared__ uint64_t rowoff[R + 1];
    __shared__ unsigned long long mm[4];
    __shared__ bool last;
    const int t = threadIdx.x, lane = t & (kWave - 1), wid = t / kWave;
    if (blockIdx.x > 0) {  // key folding (launched only with keys)
        const uint32_t F = gridDim.x - 1, b = blockIdx.x - 1;
        const size_t per = (tiles + F - 1) / F, t0 = (size_t)b * per, t1 = t0 + per < tiles ? t0 + per : tiles;
        if (t < 4) mm[t] = (t & 1) ? 0ull : ~0ull;
        __syncthreads();
        uint64_t r[4] = {~0ull, 0, ~0ull, 0};
        for (size_t i = t0 + t; i < t1; i += kScanThreads) {
            const u32x4* k4 = reinterpret_cast<const u32x4*>(keys + 4 * i);
            const u32x4 lo = k4[0], hi = k4[1];
            const uint64_t v[4] = {lo[0] | (uint64_t)lo[1] << 32, lo[2] | (uint64_t)lo[3] << 32,
                                   hi[0] | (uint64_t)hi[1] << 32, hi[2] | (uint64_t)hi[3] << 32};
#pragma unroll
            for (int j = 0; j < 4; j++) r[j] = (j & 1) ? (v[j] > r[j] ? v[j] : r[j]) : (v[j] < r[j] ? v[j] : r[j]);
        }
        r[0] = wave_min(r[0]);
        r[1] = wave_max(r[1]);
        r[2] = wave_min(r[2]);
        r[3] = wave_max(r[3]);
        if (lane == 0) {
            atomicMin(&mm[0], (unsigned long long)r[0]);
            atomicMax(&mm[1], (unsigned long long)r[1]);
            atomicMin(&mm[2], (unsigned long long)r[2]);
            atomicMax(&mm[3], (unsigned long long)r[3]);
        }
        __syncthreads();
        if (t < 4) st_agent(&key_part[4 * (size_t)b + t], (uint64_t)mm[t]);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (t == 0) last = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == F - 1;
        __syncthreads();
        if (!last) return;
        if (t < 4) {
            uint64_t f = (t & 1) ? 0ull : ~0ull;
            for (uint32_t q = 0; q < F; q++) {
                const uint64_t v = __hip_atomic_load(&key_part[4 * (size_t)q + t], __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
                f = (t & 1) ? (v > f ? v : f) : (v < f ? v : f);
            }
            MinMaxResult* o = t < 2 ? mm_a : mm_b;
            if (t & 1)
                o->max_key = f;
            else
                o->min_key = f;
        }
        if (t == 0) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    uint64_t carry = 0;
    for (size_t c0 = 0; c0 < tiles; c0 += (size_t)R * kScanThreads) {  // uniform
        uint64_t v[R], inc[R];
#pragma unroll
        for (int k = 0; k < R; k++) {
            const size_t i = c0 + (size_t)k * kScanThreads + t;
            v[k] = i < tiles ? vals[i] : 0;
        }
#pragma unroll
        for (int k = 0; k < R; k++) inc[k] = wave_inclusive_scan(v[k]);
        if (lane == kWave - 1) {
#pragma unroll
            for (int k = 0; k < R; k++) wsum[k][wid] = inc[k];
        }
        __syncthreads();
        if (t < R) {
            uint64_t run = 0;
            for (int w = 0; w < NW; w++) {
                const uint64_t x = wsum[t][w];
                wsum[t][w] = run;
                run += x;
            }
            rowtot[t] = run;
        }
        __syncthreads();
        if (t == 0) {
            uint64_t run = carry;
            for (int k = 0; k < R; k++) {
                rowoff[k] = run;
                run += rowtot[k];
            }
            rowoff[R] = run;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < R; k++) {
            const size_t i = c0 + (size_t)k * kScanThreads + t;
            if (i < tiles) prefix[i] = rowoff[k] + wsum[k][wid] + inc[k] - v[k];
        }
        carry = rowoff[R];
        __syncthreads();  // wsum / rowoff are rewritten by the next step
    }
    if (t == 0) prefix[tiles] = carry;
}

}  // namespace

void launch_delta_encode(Ctx* c, int elem_size, const void* x, size_t n, void* out) {
    if (n < 2) return;
    dispatch_int(elem_size, [&](auto tag) {
        using T = decltype(tag);
        { KernelTimer kt_(c, "k_delta_encode"); hipLaunchKernelGGL(k_delta_encode<T>, dim3(grid_for(c, n, 16 / sizeof(T))), dim3(kBlock), 0, c->stream,
                           static_cast<const T*>(x), n, static_cast<T*>(out)); }
        HIPC(hipGetLastError());
    });
}

void launch_add_scalar(Ctx* c, int elem_size, const void* x, size_t n, uint64_t s, void* out) {
    if (!n) return;
    dispatch_int(elem_size, [&](auto tag) {
        using T = decltype(tag);
        { KernelTimer kt_(c, "k_add_scalar"); hipLaunchKernelGGL((k_add_scalar<T, true>), dim3(grid_for(c, n, 16 / sizeof(T))), dim3(kBlock), 0, c->stream,
                           static_cast<const T*>(x), n, s, static_cast<T*>(out)); }
        HIPC(hipGetLastError());
    });
}

void launch_sub_scalar(Ctx* c, int elem_size, const void* x, size_t n, uint64_t s, void* out) {
    if (!n) return;
    dispatch_int(elem_size, [&](auto tag) {
        using T = decltype(tag);
        { KernelTimer kt_(c, "k_add_scalar"); hipLaunchKernelGGL((k_add_scalar<T, false>), dim3(grid_for(c, n, 16 / sizeof(T))), dim3(kBlock), 0,
                           c->stream, static_cast<const T*>(x), n, s, static_cast<T*>(out)); }
        HIPC(hipGetLastError());
    });
}

void launch_float_split(Ctx* c, const float* x, size_t n, int32_t* mant, int32_t* expo, int32_t* sign) {
    if (!n) return;
    { KernelTimer kt_(c, "k_float_split"); hipLaunchKernelGGL(k_float_split, dim3(grid_for(c, n)), dim3(kBlock), 0, c->stream, x, n, mant, expo, sign); }
    HIPC(hipGetLastError());
}

void launch_float_compose(Ctx* c, const int32_t* mant, const int32_t* expo, const uint8_t* bits, int sign_const,
                          size_t n, float* out) {
    if (!n) return;
    { KernelTimer kt_(c, "k_float_compose"); hipLaunchKernelGGL(k_float_compose, dim3(grid_for(c, n)), dim3(kBlock), 0, c->stream, mant, expo, bits,
                       sign_const, n, out); }
    HIPC(hipGetLastError());
}

void launch_float_to_int(Ctx* c, int dtype, const void* x, size_t n, int precision, int32_t* out) {
    if (!n) return;
    dispatch_native(dtype, [&](auto tag) {
        using T = decltype(tag);
        { KernelTimer kt_(c, "k_float_to_int"); hipLaunchKernelGGL(k_float_to_int<T>, dim3(grid_for(c, n, 4)), dim3(kBlock), 0, c->stream,
                           static_cast<const T*>(x), n, precision, out); }
        HIPC(hipGetLastError());
    });
}

void launch_int_to_float(Ctx* c, int dtype, const uint8_t* x, size_t n, int precision, void* out) {
    if (!n) return;
    dispatch_native(dtype, [&](auto tag) {
        using T = decltype(tag);
        { KernelTimer kt_(c, "k_int_to_float"); hipLaunchKernelGGL(k_int_to_float<T>, dim3(grid_for(c, n, 4)), dim3(kBlock), 0, c->stream,
                           reinterpret_cast<const int32_t*>(x), n, precision, static_cast<T*>(out)); }
        HIPC(hipGetLastError());
    });
}

TileScan::TileScan(Ctx* c, size_t n_tiles) : mem(c, 8 * (2 * n_tiles + 1)), tiles(n_tiles) {
    vals = mem.get<uint64_t>();
    prefix = vals + n_tiles;
}

void TileScan::scan(Ctx* c, const uint64_t* keys, MinMaxResult* mm_a, MinMaxResult* mm_b) {
    if (keys && (!mm_a || !mm_b)) fail(DDJ_EINVAL, "tile key fold without outputs");
    if (!keys && tiles > (size_t)kScanRows * kScanThreads) {  // past one step of the one-workgroup scan
        launch_scan_large(c, vals, tiles, prefix);
        return;
    }
    kernel_bytes(c, "k_scan_tiles", 8.0 * (double)(2 * tiles + 1) + (keys ? 32.0 * (double)tiles : 0.0));
    // with keys: one folding workgroup per 1024 tiles (at most 64) beside the scan
    const unsigned F = keys ? (unsigned)std::min<size_t>(64, std::max<size_t>(1, (tiles + kScanThreads - 1) / kScanThreads)) : 0;
    DevBuf part;
    if (F) part = DevBuf(c, 32 * (size_t)F);
    { KernelTimer kt_(c, "k_scan_tiles"); hipLaunchKernelGGL(k_scan_tiles, dim3(1 + F), dim3(kScanThreads), 0, c->stream, vals, tiles, prefix, keys, mm_a, mm_b,
                                                           F ? part.get<uint64_t>() : nullptr, &c->control->counters[3]); }
    HIPC(hipGetLastError());
}

void launch_scan_tiles(Ctx* c, const uint64_t* vals, size_t tiles, uint64_t* prefix) {
    kernel_bytes(c, "k_scan_tiles", 8.0 * (double)(2 * tiles + 1));
    { KernelTimer kt_(c, "k_scan_tiles"); hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanThreads), 0, c->stream, vals, tiles, prefix,
                                                           (const uint64_t*)nullptr, (MinMaxResult*)nullptr,
                                                           (MinMaxResult*)nullptr, (uint64_t*)nullptr, (uint32_t*)nullptr); }
    HIPC(hipGetLastError());
}

void arm_error_flag(Ctx* c) { c->error_flag_armed = true; }

void launch_fill_zero(Ctx* c, void* p, size_t bytes) {
    if (!bytes) return;
    size_t words = (bytes + 3) / 4;  // allocations are padded, whole words are safe
    { KernelTimer kt_(c, "k_fill_zero"); hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(c, words)), dim3(kBlock), 0, c->stream, static_cast<uint32_t*>(p),
                       words); }
    HIPC(hipGetLastError());
}

void launch_gather_copy(Ctx* c, const CopyRange* d_ranges, int n_ranges, uint8_t* dst, uint64_t max_bytes) {
    if (n_ranges <= 0 || max_bytes == 0) return;
    unsigned gx = grid_for(c, max_bytes, 64);  // 4 x 16 bytes per thread
    gx = gx > 2048 ? 2048 : gx;
    { KernelTimer kt_(c, "k_gather_copy"); hipLaunchKernelGGL(k_gather_copy, dim3(gx, n_ranges), dim3(kBlock), 0, c->stream, d_ranges, dst); }
    HIPC(hipGetLastError());
}

void launch_to_host(Ctx* c, const ToHostPieces& t, uint8_t* host_dev, uint32_t* flag_dev, uint32_t seq) {
    if (t.count < 0 || t.count > kToHostPieces) fail(DDJ_EINVAL, "too many read-back ranges");
    { KernelTimer kt_(c, "k_to_host"); hipLaunchKernelGGL(k_to_host, dim3(1), dim3(kBlock), 0, c->stream, t, host_dev, flag_dev, seq); }
    HIPC(hipGetLastError());
}

void launch_copy_pairs(Ctx* c, const CopyPair* d_pairs, int n_pairs, uint64_t max_bytes) {
    if (n_pairs <= 0 || max_bytes == 0) return;
    unsigned gx = grid_for(c, max_bytes, 4);
    gx = gx > 1024 ? 1024 : gx;
    { KernelTimer kt_(c, "k_copy_pairs"); hipLaunchKernelGGL(k_copy_pairs, dim3(gx, n_pairs), dim3(kBlock), 0, c->stream, d_pairs); }
    HIPC(hipGetLastError());
}

void launch_delta_decode(Ctx* c, int elem_size, const void* first, const uint8_t* d, size_t m, void* out) {
    dispatch_int(elem_size, [&](auto tag) {
        using T = decltype(tag);
        if (m == 0) {
            d2d(c, out, first, sizeof(T));
            return;
        }
        const size_t tiles = (m + kScanTile - 1) / kScanTile;
        TileScan ts(c, tiles);
        kernel_bytes(c, "k_delta_sums", (double)m * sizeof(T));
        { KernelTimer kt_(c, "k_delta_sums"); hipLaunchKernelGGL(k_delta_sums<T>, dim3((unsigned)tiles), dim3(kBlock), 0, c->stream, d, m, ts.vals); }
        HIPC(hipGetLastError());
        ts.scan(c);
        { KernelTimer kt_(c, "k_delta_decode"); hipLaunchKernelGGL(k_delta_decode<T>, dim3((unsigned)tiles), dim3(kBlock), 0, c->stream,
                           static_cast<const uint8_t*>(first), d, m, static_cast<T*>(out), (const uint64_t*)ts.prefix); }
        HIPC(hipGetLastError());
    });
}

}  // namespace ddj
