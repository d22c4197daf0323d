// Single-pass reductions: min/max (FL bit width, scale, patch thresholds)
// and the fused statistics pass the planner reads.
//
// Every workgroup reduces a grid-stride share with 16-byte loads and writes a
// partial.  Min/max partials are folded by a one-workgroup launch
// (k_minmax_fold); the statistics pass folds its own in the last workgroup
// to finish (ticket on a control counter that it resets).
#include <algorithm>
#include <vector>
#include <limits>

#include "dispatch.h"
#include "k_common.h"
#include "kernels.h"

namespace ddj {

namespace {

constexpr int kMaxReduceBlocks = 2048;

__device__ __forceinline__ uint64_t block_reduce_min(uint64_t v, uint64_t* sm) {
    v = wave_min(v);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (lane == 0) sm[wid] = v;
    __syncthreads();
    uint64_t r = sm[0];
    for (int w = 1; w < kBlock / kWave; w++) r = sm[w] < r ? sm[w] : r;
    __syncthreads();
    return r;
}
__device__ __forceinline__ uint64_t block_reduce_max(uint64_t v, uint64_t* sm) {
    v = wave_max(v);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (lane == 0) sm[wid] = v;
    __syncthreads();
    uint64_t r = sm[0];
    for (int w = 1; w < kBlock / kWave; w++) r = sm[w] > r ? sm[w] : r;
    __syncthreads();
    return r;
}
template <typename A> __device__ __forceinline__ A block_reduce_sum(A v, A* sm) {
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (lane == 0) sm[wid] = v;
    __syncthreads();
    A r = 0;
    for (int w = 0; w < kBlock / kWave; w++) r += sm[w];
    __syncthreads();
    return r;
}

__device__ __forceinline__ uint64_t rd(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t rd32(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------------------ minmax
// Keys of the reference's minimum and maximum: thrust::minmax_element
// (util/statistics/cuda_array_statistics.cu:20-28) under operator<, taken as
// the sequential fold m = x[0]; m = x[i] < m ? x[i] : m.  For floats that
// means: NaN rows never win, so a NaN in row 0 is both the minimum and the
// maximum; -0.0 and +0.0 tie, and the first zero in row order gives the sign.
// Rows feed add(); `zf` keeps (row << 1 | sign) of the first zero.
template <typename T, bool FLOAT = std::is_floating_point<T>::value> struct FoldKeys {
    uint64_t zf = ~0ull;
    uint64_t kmn = ~0ull, kmx = 0;
    __device__ __forceinline__ void add(T v, size_t row) {
        if (v != v) return;
        if (v == T(0)) {
            const uint64_t z = ((uint64_t)row << 1) | (__builtin_signbit(v) ? 1ull : 0ull);
            zf = z < zf ? z : zf;
            v = T(0);
        }
        const uint64_t key = order_key(v);
        kmn = key < kmn ? key : kmn;
        kmx = key > kmx ? key : kmx;
    }
    __device__ __forceinline__ uint64_t min_key() const { return kmn; }
    __device__ __forceinline__ uint64_t max_key() const { return kmx; }
};
// Integers: native compares per row (order keys are monotone in the value),
// keys only for the result; with no rows the keys are ~0 / 0 as above.
template <typename T> struct FoldKeys<T, false> {
    uint64_t zf = ~0ull;
    T mn = std::numeric_limits<T>::max(), mx = std::numeric_limits<T>::min();
    bool any = false;
    __device__ __forceinline__ void add(T v, size_t) {
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
        any = true;
    }
    __device__ __forceinline__ uint64_t min_key() const { return any ? order_key(mn) : ~0ull; }
    __device__ __forceinline__ uint64_t max_key() const { return any ? order_key(mx) : 0ull; }
};

// Final keys from the reduced ones and row 0.
template <typename T> __device__ __forceinline__ void fold_fixup(const T* x, uint64_t& mn, uint64_t& mx, uint64_t zf) {
    if constexpr (std::is_floating_point<T>::value) {
        const T x0 = x[0];
        if (x0 != x0) {
            mn = mx = order_key(x0);
            return;
        }
        if (zf != ~0ull && (zf & 1)) {
            const uint64_t kz = order_key(T(0)), knz = order_key(-T(0));
            if (mn == kz) mn = knz;
            if (mx == kz) mx = knz;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_minmax(const T* __restrict__ x, size_t n, uint64_t* __restrict__ partials,
                                                   const uint64_t* d_rows) {
    __shared__ uint64_t sm[kBlock / kWave];
    if (d_rows) n = *d_rows < n ? *d_rows : n;  // rows counted on the device (uniform)
    FoldKeys<T> f;
    const size_t tid = (size_t)blockIdx.x * kBlock + threadIdx.x, stride = (size_t)gridDim.x * kBlock;
    constexpr int V = 16 / sizeof(T);
    size_t done = 0;
    if ((reinterpret_cast<uintptr_t>(x) & 15) == 0) {
        const u32x4* xv = reinterpret_cast<const u32x4*>(x);
        const size_t nv = n / V;
        constexpr int U = 4;  // loads in flight per thread
        for (size_t i0 = tid; i0 < nv; i0 += U * stride) {
            u32x4 v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const size_t i = i0 + u * stride;
                if (i < nv) v[u] = xv[i];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const size_t i = i0 + u * stride;
                if (i >= nv) break;
                T vals[V];
                __builtin_memcpy(vals, &v[u], 16);
#pragma unroll
                for (int k = 0; k < V; k++) f.add(vals[k], i * V + k);
            }
        }
        done = nv * V;
    }
    for (size_t i = done + tid; i < n; i += stride) f.add(x[i], i);
    uint64_t mn = block_reduce_min(f.min_key(), sm);
    uint64_t mx = block_reduce_max(f.max_key(), sm);
    uint64_t zf = block_reduce_min(f.zf, sm);
    if (threadIdx.x == 0) {
        st_agent(partials + 3 * blockIdx.x, mn);
        st_agent(partials + 3 * blockIdx.x + 1, mx);
        st_agent(partials + 3 * blockIdx.x + 2, zf);
    }
}

// Min/max of the wrapping differences d[i] = x[i+1] - x[i], i < n - 1, read
// straight from x (the delta encoder's output is never written when an FL
// node consumes it).  Integer types only.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_minmax_delta(const T* __restrict__ x, size_t n, uint64_t* __restrict__ partials) {
    __shared__ uint64_t sm[kBlock / kWave];
    using U = typename UnsignedOf<T>::type;
    T dmn = std::numeric_limits<T>::max(), dmx = std::numeric_limits<T>::min();
    bool any = false;
    auto add = [&](T a, T b) {  // d = b - a, wrapping in T
        const T d = (T)(U)((U)b - (U)a);
        dmn = d < dmn ? d : dmn;
        dmx = d > dmx ? d : dmx;
        any = true;
    };
    const size_t m = n - 1;
    const size_t tid = (size_t)blockIdx.x * kBlock + threadIdx.x, stride = (size_t)gridDim.x * kBlock;
    constexpr int V = 16 / sizeof(T);
    size_t done = 0;
    if ((reinterpret_cast<uintptr_t>(x) & 15) == 0) {
        const u32x4* xv = reinterpret_cast<const u32x4*>(x);
        const size_t nv = m / V;  // vectors whose V deltas are all in range (x[i*V+V] exists)
        constexpr int Uf = 4;
        const int lane = threadIdx.x & (kWave - 1);
        for (size_t i0 = tid; i0 < nv; i0 += Uf * stride) {
            u32x4 v[Uf];
            T nx[Uf];
            // The row after vector i starts vector i + 1, which the next lane
            // holds (lanes take consecutive vectors); the last lane of the
            // wave, and the thread with the last vector, load it with the
            // vectors so that the loop waits on memory once.
#pragma unroll
            for (int u = 0; u < Uf; u++) {
                const size_t i = i0 + u * stride;
                if (i < nv) {
                    v[u] = xv[i];
                    if (lane == kWave - 1 || i + 1 >= nv) nx[u] = x[(i + 1) * V];  // (i + 1) * V <= m < n
                }
            }
#pragma unroll
            for (int u = 0; u < Uf; u++) {
                const size_t i = i0 + u * stride;
                // first 8 bytes of the next lane's vector (all lanes take part)
                const uint32_t nb0 = (uint32_t)__shfl_down((int)v[u].x, 1, kWave);
                const uint32_t nb1 = (uint32_t)__shfl_down((int)v[u].y, 1, kWave);
                if (i >= nv) continue;
                T vals[V + 1];
                __builtin_memcpy(vals, &v[u], 16);
                if (lane == kWave - 1 || i + 1 >= nv) {
                    vals[V] = nx[u];
                } else {
                    const uint64_t nb = ((uint64_t)nb1 << 32) | nb0;
                    __builtin_memcpy(&vals[V], &nb, sizeof(T));
                }
#pragma unroll
                for (int k = 0; k < V; k++) add(vals[k], vals[k + 1]);
            }
        }
        done = nv * V;
    }
    for (size_t i = done + tid; i < m; i += stride) add(x[i], x[i + 1]);
    uint64_t mn = block_reduce_min(any ? order_key(dmn) : ~0ull, sm);
    uint64_t mx = block_reduce_max(any ? order_key(dmx) : 0ull, sm);
    if (threadIdx.x == 0) {
        st_agent(partials + 2 * blockIdx.x, mn);
        st_agent(partials + 2 * blockIdx.x + 1, mx);
    }
}

// Folds the per-workgroup partials of k_minmax (STRIDE 3: min, max, first
// zero) or k_minmax_delta (STRIDE 2) into the result, in a launch of its own:
// a "last workgroup" ticket would take one device-scope atomic per
// workgroup on a single address, which costs more than this launch.
template <typename T, int STRIDE>
__global__ __launch_bounds__(kBlock) void k_minmax_fold(const uint64_t* __restrict__ partials, int blocks, const T* x,
                                                        MinMaxResult* out, uint64_t* precision_out) {
    __shared__ uint64_t sm[kBlock / kWave];
    uint64_t mn = ~0ull, mx = 0, zf = ~0ull, pr = 0;
    for (int b = threadIdx.x; b < blocks; b += kBlock) {
        const uint64_t a = partials[STRIDE * b], z = partials[STRIDE * b + 1];
        mn = a < mn ? a : mn;
        mx = z > mx ? z : mx;
        if (STRIDE >= 3) {
            const uint64_t q = partials[STRIDE * b + 2];
            zf = q < zf ? q : zf;
        }
        if (STRIDE == 4) {
            const uint64_t q = partials[STRIDE * b + 3];
            pr = q > pr ? q : pr;
        }
    }
    mn = block_reduce_min(mn, sm);
    mx = block_reduce_max(mx, sm);
    if (STRIDE >= 3) zf = block_reduce_min(zf, sm);
    if (STRIDE == 4) pr = block_reduce_max(pr, sm);
    if (threadIdx.x == 0) {
        if (STRIDE >= 3) fold_fixup(x, mn, mx, zf);
        out->min_key = mn;
        out->max_key = mx;
        if (STRIDE == 4) *precision_out = pr;
    }
}

// ------------------------------------------------------------------- stats
// _getFloatPrecision (reference util/statistics/cuda_array_statistics.cu:37-44)
// as the loop count k of round(x*10^k)/10^k != x in float arithmetic.
__device__ __forceinline__ uint32_t float_precision(float number) {
    float e = 1.0f;
    uint32_t k = 0;
    while (k < 10) {
        float r = roundf(number * e);
        if (__fdiv_rn(r, e) == number) break;
        e *= 10.0f;
        k++;
    }
    return k;
}

// Min/max keys and the float precision (the maximum of float_precision
// over the rows, as the statistics pass computes it) in one pass: the two
// inputs of the floatToInt encoder (float_to_int_encoding.cu).  Partials:
// {min, max, first zero, precision} per workgroup.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_minmax_prec(const T* __restrict__ x, size_t n,
                                                        uint64_t* __restrict__ partials) {
    __shared__ uint64_t sm[kBlock / kWave];
    FoldKeys<T> f;
    uint32_t prec = 0;
    // The precision is a maximum capped at 10: a lane that has seen 10
    // (noisy or NaN data) only folds min/max from then on, and once every
    // lane of the wave has, the precision loop is skipped by the whole wave.
    auto add = [&](T v, size_t row) {
        f.add(v, row);
        if (prec < 10) {
            const uint32_t p = float_precision((float)v);
            prec = p > prec ? p : prec;
        }
    };
    const size_t tid = (size_t)blockIdx.x * kBlock + threadIdx.x, stride = (size_t)gridDim.x * kBlock;
    constexpr int V = 16 / sizeof(T);
    size_t done = 0;
    if ((reinterpret_cast<uintptr_t>(x) & 15) == 0) {
        const u32x4* xv = reinterpret_cast<const u32x4*>(x);
        const size_t nv = n / V;
        constexpr int U = 4;  // loads in flight per thread
        for (size_t i0 = tid; i0 < nv; i0 += U * stride) {
            u32x4 v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const size_t i = i0 + u * stride;
                if (i < nv) v[u] = xv[i];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const size_t i = i0 + u * stride;
                if (i >= nv) break;
                T vals[V];
                __builtin_memcpy(vals, &v[u], 16);
#pragma unroll
                for (int k = 0; k < V; k++) add(vals[k], i * V + k);
            }
        }
        done = nv * V;
    }
    for (size_t i = done + tid; i < n; i += stride) add(x[i], i);
    const uint64_t mn = block_reduce_min(f.min_key(), sm);
    const uint64_t mx = block_reduce_max(f.max_key(), sm);
    const uint64_t zf = block_reduce_min(f.zf, sm);
    const uint64_t pr = block_reduce_max((uint64_t)prec, sm);
    if (threadIdx.x == 0) {
        st_agent(partials + 4 * blockIdx.x, mn);
        st_agent(partials + 4 * blockIdx.x + 1, mx);
        st_agent(partials + 4 * blockIdx.x + 2, zf);
        st_agent(partials + 4 * blockIdx.x + 3, pr);
    }
}

template <typename T> __device__ __forceinline__ int32_t to_int32_sat(T v) {
    if constexpr (std::is_floating_point<T>::value) {
        double d = (double)v;
        if (d != d) return 0;
        if (d >= 2147483647.0) return 2147483647;
        if (d <= -2147483648.0) return (-2147483647 - 1);
        return (int32_t)d;
    } else {
        return (int32_t)(int64_t)v;
    }
}

template <typename T> __device__ __forceinline__ bool rl_equal(T a, T b) {
    // _rlMetricKernel (cuda_array_statistics.cu:97-115): int value = data[idx]; value == data[idx+1]
    int32_t v = to_int32_sat(a);
    if constexpr (std::is_floating_point<T>::value)
        return (T)v == b;
    else
        return (int64_t)v == (int64_t)b;
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

template <typename T> __device__ __forceinline__ uint64_t zigzag_bits(T v) {
    if constexpr (std::is_floating_point<T>::value) {
        return ubits(v);
    } else {
        int64_t s = (int64_t)v;
        return ((uint64_t)s << 1) ^ (uint64_t)(s >> 63);
    }
}
__device__ __forceinline__ uint32_t bitwidth64(uint64_t v) { return v ? 64 - __clzll(v) : 0; }

struct StatsAcc {
    uint64_t mn, mx, zf, unsorted, rl_eq, runs, sum_bits;
    double sum_f;
    uint32_t prec, dmax;
};

template <typename T>
__device__ __forceinline__ void stats_elem(StatsAcc& a, T v, T next, bool has_next, bool has_next2, size_t row,
                                           uint32_t* sm_hist, uint32_t* sm_hll) {
    if constexpr (std::is_floating_point<T>::value) {
        FoldKeys<T> fk;
        fk.kmn = a.mn;
        fk.kmx = a.mx;
        fk.zf = a.zf;
        fk.add(v, row);
        a.mn = fk.kmn;
        a.mx = fk.kmx;
        a.zf = fk.zf;
    } else {
        const uint64_t key = order_key(v);
        a.mn = key < a.mn ? key : a.mn;
        a.mx = key > a.mx ? key : a.mx;
    }
    if constexpr (std::is_floating_point<T>::value) {
        a.sum_f += (double)v;
        uint32_t p = float_precision((float)v);
        a.prec = p > a.prec ? p : a.prec;
    } else {
        a.sum_bits += (uint64_t)(int64_t)v;
    }
    if (has_next) {
        if (v > next) a.unsorted = 1;
        if (!(next == v)) a.runs++;
        if constexpr (!std::is_floating_point<T>::value) {
            constexpr int W = 8 * sizeof(T);
            const uint64_t d = ubits(next) - ubits(v);
            const int64_t sd = (int64_t)(d << (64 - W)) >> (64 - W);  // wrap to W bits, sign-extend
            const uint64_t z = ((uint64_t)sd << 1) ^ (uint64_t)(sd >> 63);
            const uint32_t bw = bitwidth64(z);
            a.dmax = bw > a.dmax ? bw : a.dmax;
        }
    }
    if (has_next2 && rl_equal(v, next)) a.rl_eq++;
    atomicAdd(&sm_hist[bitwidth64(zigzag_bits(v))], 1u);
    const uint64_t h = mix64(ubits(v) + 0x9e3779b97f4a7c15ull);
    const uint32_t reg = (uint32_t)(h >> 58);
    const uint32_t rank = (uint32_t)__clzll((h << 6) | (1ull << 5)) + 1;
    atomicMax(&sm_hll[reg], rank);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_stats(const T* __restrict__ x, size_t n, DevStats* __restrict__ partials,
                                                  uint32_t* counter, DevStats* out) {
    __shared__ uint64_t sm[kBlock / kWave];
    __shared__ double smd[kBlock / kWave];
    __shared__ uint32_t sm_hist[65];
    __shared__ uint32_t sm_hll[64];
    __shared__ uint32_t fin_hist[65];
    __shared__ uint32_t fin_hll[64];
    __shared__ bool last;
    for (int i = threadIdx.x; i < 65; i += kBlock) sm_hist[i] = 0;
    for (int i = threadIdx.x; i < 64; i += kBlock) sm_hll[i] = 0;
    __syncthreads();
    StatsAcc a{~0ull, 0, ~0ull, 0, 0, 0, 0, 0.0, 0, 0};
    const size_t tid = (size_t)blockIdx.x * kBlock + threadIdx.x, stride = (size_t)gridDim.x * kBlock;
    constexpr int V = 16 / sizeof(T);
    size_t done = 0;
    if ((reinterpret_cast<uintptr_t>(x) & 15) == 0) {
        const u32x4* xv = reinterpret_cast<const u32x4*>(x);
        const size_t nv = n / V;
        for (size_t i = tid; i < nv; i += stride) {
            u32x4 v = xv[i];
            T vals[V + 1];
            __builtin_memcpy(vals, &v, 16);
            const size_t base = i * V;
            vals[V] = (base + V < n) ? x[base + V] : T(0);
#pragma unroll
            for (int k = 0; k < V; k++)
                stats_elem(a, vals[k], vals[k + 1], base + k + 1 < n, base + k + 2 < n, base + k, sm_hist, sm_hll);
        }
        done = nv * V;
    }
    for (size_t i = done + tid; i < n; i += stride) {
        T v = x[i];
        T nx = (i + 1 < n) ? x[i + 1] : T(0);
        stats_elem(a, v, nx, i + 1 < n, i + 2 < n, i, sm_hist, sm_hll);
    }
    uint64_t mn = block_reduce_min(a.mn, sm);
    uint64_t mx = block_reduce_max(a.mx, sm);
    uint64_t zf = block_reduce_min(a.zf, sm);
    uint64_t uns = block_reduce_max(a.unsorted, sm);
    uint64_t rl = block_reduce_sum<uint64_t>(a.rl_eq, sm);
    uint64_t runs = block_reduce_sum<uint64_t>(a.runs, sm);
    uint64_t sb = block_reduce_sum<uint64_t>(a.sum_bits, sm);
    double sf = block_reduce_sum<double>(a.sum_f, smd);
    uint64_t pr = block_reduce_max((uint64_t)a.prec, sm);
    uint64_t dm = block_reduce_max((uint64_t)a.dmax, sm);
    DevStats* p = partials + blockIdx.x;
    if (threadIdx.x == 0) {
        st_agent(&p->min_key, mn);
        st_agent(&p->max_key, mx);
        st_agent(&p->first_zero, zf);
        st_agent(&p->unsorted, uns);
        st_agent(&p->rl_eq, rl);
        st_agent(&p->runs, runs);
        st_agent(&p->sum_bits, sb);
        uint64_t sfb;
        __builtin_memcpy(&sfb, &sf, 8);
        st_agent(reinterpret_cast<uint64_t*>(&p->sum_f), sfb);
        st_agent(&p->precision, (uint32_t)pr);
        st_agent(&p->delta_bits_max, (uint32_t)dm);
    }
    for (int i = threadIdx.x; i < 65; i += kBlock) st_agent(&p->bit_hist[i], sm_hist[i]);
    for (int i = threadIdx.x; i < 64; i += kBlock) st_agent(&p->hll[i], sm_hll[i]);
    if (!last_block_arrives(counter, &last)) return;
    StatsAcc f{~0ull, 0, ~0ull, 0, 0, 0, 0, 0.0, 0, 0};
    for (uint32_t b = threadIdx.x; b < gridDim.x; b += kBlock) {
        const DevStats* q = partials + b;
        const uint64_t v0 = rd(&q->min_key), v1 = rd(&q->max_key);
        f.mn = v0 < f.mn ? v0 : f.mn;
        f.mx = v1 > f.mx ? v1 : f.mx;
        const uint64_t z1 = rd(&q->first_zero);
        f.zf = z1 < f.zf ? z1 : f.zf;
        f.unsorted |= rd(&q->unsorted);
        f.rl_eq += rd(&q->rl_eq);
        f.runs += rd(&q->runs);
        f.sum_bits += rd(&q->sum_bits);
        const uint64_t sfb = rd(reinterpret_cast<const uint64_t*>(&q->sum_f));
        double sfd;
        __builtin_memcpy(&sfd, &sfb, 8);
        f.sum_f += sfd;
        const uint32_t pp = rd32(&q->precision), dd = rd32(&q->delta_bits_max);
        f.prec = pp > f.prec ? pp : f.prec;
        f.dmax = dd > f.dmax ? dd : f.dmax;
    }
    mn = block_reduce_min(f.mn, sm);
    mx = block_reduce_max(f.mx, sm);
    zf = block_reduce_min(f.zf, sm);
    uns = block_reduce_max(f.unsorted, sm);
    rl = block_reduce_sum<uint64_t>(f.rl_eq, sm);
    runs = block_reduce_sum<uint64_t>(f.runs, sm);
    sb = block_reduce_sum<uint64_t>(f.sum_bits, sm);
    sf = block_reduce_sum<double>(f.sum_f, smd);
    pr = block_reduce_max((uint64_t)f.prec, sm);
    dm = block_reduce_max((uint64_t)f.dmax, sm);
    for (int i = threadIdx.x; i < 65; i += kBlock) fin_hist[i] = 0;
    for (int i = threadIdx.x; i < 64; i += kBlock) fin_hll[i] = 0;
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < gridDim.x; b += kBlock) {
        const DevStats* q = partials + b;
        for (int e = 0; e < 65; e++) atomicAdd(&fin_hist[e], rd32(&q->bit_hist[e]));
        for (int e = 0; e < 64; e++) atomicMax(&fin_hll[e], rd32(&q->hll[e]));
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 65; i += kBlock) out->bit_hist[i] = fin_hist[i];
    for (int i = threadIdx.x; i < 64; i += kBlock) out->hll[i] = fin_hll[i];
    if (threadIdx.x == 0) {
        fold_fixup(x, mn, mx, zf);
        out->min_key = mn;
        out->max_key = mx;
        out->first_zero = zf;
        out->unsorted = uns;
        out->rl_eq = rl;
        out->runs = runs + (n > 0 ? 1 : 0);
        out->sum_bits = sb;
        out->sum_f = sf;
        out->precision = (uint32_t)pr;
        out->delta_bits_max = (uint32_t)dm;
        __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

int reduce_grid(Ctx* c, size_t n, int elem) {
    const size_t per_block = (size_t)kBlock * (16 / elem) * 4;
    const size_t g = (n + per_block - 1) / per_block;
    const size_t cap = std::min<size_t>(kMaxReduceBlocks, (size_t)c->cu_count * 4);
    return (int)std::max<size_t>(1, std::min(g, cap));
}

}  // namespace

void launch_minmax(Ctx* c, int dtype, const void* x, size_t n, MinMaxResult* d_result, const uint64_t* d_rows) {
    if (n == 0) fail(DDJ_EINVAL, "minmax of an empty array");
    dispatch_native(dtype, [&](auto tag) {
        using T = decltype(tag);
        const int grid = reduce_grid(c, n, sizeof(T));
        DevBuf partials(c, 24 * (size_t)grid);
        { KernelTimer kt_(c, "k_minmax"); hipLaunchKernelGGL(k_minmax<T>, dim3(grid), dim3(kBlock), 0, c->stream, static_cast<const T*>(x), n,
                           partials.get<uint64_t>(), d_rows); }
        HIPC(hipGetLastError());
        { KernelTimer kt_(c, "k_minmax_fold"); hipLaunchKernelGGL((k_minmax_fold<T, 3>), dim3(1), dim3(kBlock), 0, c->stream,
                           partials.get<uint64_t>(), grid, static_cast<const T*>(x), d_result, nullptr); }
        HIPC(hipGetLastError());
    });
}

void launch_float_scan(Ctx* c, int dtype, const void* x, size_t n, FloatScanResult* d_result) {
    if (n == 0) fail(DDJ_EINVAL, "float scan of an empty array");
    dispatch_native(dtype, [&](auto tag) {
        using T = decltype(tag);
        if constexpr (std::is_floating_point<T>::value) {
            const int grid = reduce_grid(c, n, sizeof(T));
            DevBuf partials(c, 32 * (size_t)grid);
            kernel_bytes(c, "k_minmax_prec", (double)n * sizeof(T));
            { KernelTimer kt_(c, "k_minmax_prec"); hipLaunchKernelGGL(k_minmax_prec<T>, dim3(grid), dim3(kBlock), 0, c->stream,
                               static_cast<const T*>(x), n, partials.get<uint64_t>()); }
            HIPC(hipGetLastError());
            { KernelTimer kt_(c, "k_minmax_fold"); hipLaunchKernelGGL((k_minmax_fold<T, 4>), dim3(1), dim3(kBlock), 0, c->stream,
                               partials.get<uint64_t>(), grid, static_cast<const T*>(x), &d_result->mm,
                               &d_result->precision); }
            HIPC(hipGetLastError());
        } else {
            fail(DDJ_EINVAL, "float scan of an integer column");
        }
    });
}

// Workgroups of a reduction whose partials go to the host (6 KB of them):
// RLE's values and lengths, whose row count (the run count) is only known
// on the device and usually far below the bound the grid is sized for.
// Against a full column's 1024 workgroups and a device fold, the 64M-row
// step-function tree compresses in 0.145 instead of 0.158 ms.
constexpr int kPartialsGrid = 256;

int launch_minmax_partials(Ctx* c, int dtype, const void* x, size_t n, const uint64_t* d_rows, DevBuf& partials) {
    if (n == 0) fail(DDJ_EINVAL, "minmax of an empty array");
    int grid = 0;
    dispatch_native(dtype, [&](auto tag) {
        using T = decltype(tag);
        if constexpr (std::is_floating_point<T>::value) {
            fail(DDJ_EINVAL, "host-folded min/max: integer columns only");
        } else {
            grid = std::min(reduce_grid(c, n, sizeof(T)), kPartialsGrid);
            partials = DevBuf(c, 24 * (size_t)grid);
            KernelTimer kt_(c, "k_minmax");
            hipLaunchKernelGGL(k_minmax<T>, dim3(grid), dim3(kBlock), 0, c->stream, static_cast<const T*>(x), n,
                               partials.get<uint64_t>(), d_rows);
            HIPC(hipGetLastError());
        }
    });
    return grid;
}

MinMaxResult fold_minmax_partials(const uint64_t* h, int grid, int stride) {
    MinMaxResult r{~0ull, 0ull};
    for (int b = 0; b < grid; b++) {  // order keys: the min of the minima, the max of the maxima
        r.min_key = std::min(r.min_key, h[(size_t)stride * b]);
        r.max_key = std::max(r.max_key, h[(size_t)stride * b + 1]);
    }
    return r;
}

MinMaxResult minmax_to_host(Ctx* c, int dtype, const void* x, size_t n, bool delta) {
    if (n == 0 || (delta && n < 2)) fail(DDJ_EINVAL, "min/max of too few rows");
    MinMaxResult r{~0ull, 0ull};
    dispatch_native(dtype, [&](auto tag) {
        using T = decltype(tag);
        if constexpr (std::is_floating_point<T>::value) {
            fail(DDJ_EINVAL, "host-folded min/max: integer columns only");
        } else {
            const int grid = reduce_grid(c, n, sizeof(T));
            const int stride = delta ? 2 : 3;
            DevBuf partials(c, 8 * (size_t)stride * (size_t)grid);
            {
                KernelTimer kt_(c, delta ? "k_minmax_delta" : "k_minmax");
                if (delta)
                    hipLaunchKernelGGL(k_minmax_delta<T>, dim3(grid), dim3(kBlock), 0, c->stream, static_cast<const T*>(x),
                                       n, partials.get<uint64_t>());
                else
                    hipLaunchKernelGGL(k_minmax<T>, dim3(grid), dim3(kBlock), 0, c->stream, static_cast<const T*>(x), n,
                                       partials.get<uint64_t>(), nullptr);
            }
            HIPC(hipGetLastError());
            std::vector<uint64_t> h((size_t)stride * (size_t)grid);
            d2h(c, h.data(), partials.get(), 8 * h.size());
            r = fold_minmax_partials(h.data(), grid, stride);
        }
    });
    return r;
}

void launch_minmax_delta(Ctx* c, int elem_size, const void* x, size_t n, MinMaxResult* d_result) {
    if (n < 2) fail(DDJ_EINVAL, "delta min/max needs two rows");
    dispatch_int(elem_size, [&](auto tag) {
        using T = decltype(tag);
        const int grid = reduce_grid(c, n, sizeof(T));
        DevBuf partials(c, 16 * (size_t)grid);
        { KernelTimer kt_(c, "k_minmax_delta"); hipLaunchKernelGGL(k_minmax_delta<T>, dim3(grid), dim3(kBlock), 0, c->stream, static_cast<const T*>(x), n,
                           partials.get<uint64_t>()); }
        HIPC(hipGetLastError());
        { KernelTimer kt_(c, "k_minmax_fold"); hipLaunchKernelGGL((k_minmax_fold<T, 2>), dim3(1), dim3(kBlock), 0, c->stream,
                           partials.get<uint64_t>(), grid, static_cast<const T*>(x), d_result, nullptr); }
        HIPC(hipGetLastError());
    });
}

void decode_minmax(int dtype, const MinMaxResult& r, void* mn, void* mx) {
    dispatch_native(dtype, [&](auto tag) {
        using T = decltype(tag);
        T a = from_order_key<T>(r.min_key), b = from_order_key<T>(r.max_key);
        memcpy(mn, &a, sizeof(T));
        memcpy(mx, &b, sizeof(T));
    });
}

void launch_stats(Ctx* c, int dtype, const void* x, size_t n, DevStats* d_out) {
    if (n == 0) fail(DDJ_EINVAL, "stats of an empty array");
    dispatch_native(dtype, [&](auto tag) {
        using T = decltype(tag);
        const int grid = std::min(reduce_grid(c, n, sizeof(T)), 2 * c->cu_count);
        DevBuf partials(c, sizeof(DevStats) * (size_t)grid);
        { KernelTimer kt_(c, "k_stats"); hipLaunchKernelGGL(k_stats<T>, dim3(grid), dim3(kBlock), 0, c->stream, static_cast<const T*>(x), n,
                           partials.get<DevStats>(), &c->control->counters[1], d_out); }
        HIPC(hipGetLastError());
    });
}

}  // namespace ddj
