// Most-frequent values (Histogram::GetMostFrequent, reference
// util/histogram/histogram.cpp:29-41 -> thrust_sparse_histogram.cu:20-69 and
// most_frequent.cu:10-26): exact counts of every distinct value, then the K
// largest counts, equal counts ordered by the smaller value (what the
// reference's stable sort of key-sorted bins yields).
//
// Counting: each workgroup aggregates its rows in an LDS hash table (plain
// LDS atomics; the 16 rows a thread loads at once keep HBM busy while they
// wait) and flushes it into a global open-addressing table (64-bit keys,
// ~0 = empty);
// rows that find no LDS slot go straight to the global table.  A 64-bit
// value equal to the empty marker is counted on the side.  Floats count
// -0.0 with +0.0 (operator==), reporting -0.0 when one was seen (it sorts
// first in the reference).  A NaN equals nothing, so each NaN row is a bin
// of count 1: they stay out of the table, and every workgroup keeps the four
// smallest (radix order) NaN rows, which is all the top-K can ever take from
// them.  The global table starts at 2n slots (at least 1024, at most 64K);
// if it ends up more than 70% full the host retries once with a table of
// twice the keys seen (the exact count, or with a full table the keys seen
// plus the rows it dropped, each possibly a new key), and a second retry,
// never needed in practice, sizes it for every row distinct.
#include "dispatch.h"
#include "k_common.h"
#include "kernels.h"

namespace ddj {

namespace {

constexpr uint64_t kEmpty = ~0ull;
constexpr int kLdsSlots = 2048;
constexpr int kMaxK = 4;

struct HistCtl {
    unsigned long long distinct;
    unsigned long long special;  // rows whose 64-bit pattern equals kEmpty
    unsigned long long nans;     // NaN rows (floats)
    unsigned long long dropped;  // rows that found no slot in a full table
    uint32_t overflow;
    uint32_t negzero;
};

__device__ __forceinline__ uint64_t hmix(uint64_t z) {
    z ^= z >> 33;
    z *= 0xff51afd7ed558ccdull;
    z ^= z >> 33;
    z *= 0xc4ceb9fe1a85ec53ull;
    z ^= z >> 33;
    return z;
}

template <typename T> __device__ __forceinline__ uint64_t hist_key(T v, bool& negzero) {
    if constexpr (std::is_floating_point<T>::value) {
        if (v == T(0)) {
            if (__builtin_signbit(v)) negzero = true;
            return 0;
        }
    }
    return ubits(v);
}

__device__ bool global_insert(unsigned long long* keys, unsigned long long* counts, uint64_t cap_mask, uint64_t key,
                              unsigned long long cnt, HistCtl* ctl) {
    uint64_t h = hmix(key) & cap_mask;
    for (uint64_t p = 0; p <= cap_mask && p < 4096; p++) {
        const uint64_t s = (h + p) & cap_mask;
        unsigned long long k = __hip_atomic_load(&keys[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (k == key) {
            atomicAdd(&counts[s], cnt);
            return true;
        }
        if (k == kEmpty) {
            unsigned long long prev = atomicCAS(&keys[s], kEmpty, (unsigned long long)key);
            if (prev == kEmpty || prev == key) {
                atomicAdd(&counts[s], cnt);
                if (prev == kEmpty) atomicAdd(&ctl->distinct, 1ull);
                return true;
            }
        }
    }
    atomicOr(&ctl->overflow, 1u);
    atomicAdd(&ctl->dropped, cnt);
    return false;
}

// Ascending list of the (up to) four smallest keys seen, with multiplicity.
struct Small4 {
    uint64_t l[kMaxK];
    int c = 0;
    __device__ __forceinline__ void insert(uint64_t k) {
#pragma unroll
        for (int j = 0; j < kMaxK; j++) {
            if (j < c) {
                if (k < l[j]) {
                    const uint64_t t = l[j];
                    l[j] = k;
                    k = t;
                }
            } else if (j == c) {
                l[j] = k;
            }
        }
        c = c < kMaxK ? c + 1 : kMaxK;
    }
    __device__ __forceinline__ void pop() {
#pragma unroll
        for (int j = 0; j + 1 < kMaxK; j++) l[j] = l[j + 1];
        c--;
    }
};

// Four smallest keys of the workgroup (each thread holds a Small4) into
// out_keys[0..*out_n).  Wave rounds take the smallest head, then thread 0
// merges the waves' lists.
template <int NT>
__device__ void block_small4(Small4& s, uint64_t* sm_keys /* NT/kWave * kMaxK */, int* sm_cnt /* NT/kWave */,
                             uint64_t* out_keys, uint32_t* out_n) {
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    int got = 0;
    for (int r = 0; r < kMaxK; r++) {
        const bool has = s.c > 0;
        const uint64_t h = has ? s.l[0] : ~0ull;
        const uint64_t m = wave_min(h);
        const uint64_t who = __ballot(has && h == m);
        if (who == 0) break;
        if (lane == __builtin_ctzll(who)) s.pop();
        if (lane == 0) sm_keys[wid * kMaxK + r] = m;
        got++;
    }
    if (lane == 0) sm_cnt[wid] = got;
    __syncthreads();
    if (threadIdx.x == 0) {
        Small4 b;
        for (int w = 0; w < NT / kWave; w++)
            for (int r = 0; r < sm_cnt[w]; r++) b.insert(sm_keys[w * kMaxK + r]);
#pragma unroll
        for (int r = 0; r < kMaxK; r++)
            if (r < b.c) out_keys[r] = b.l[r];
        *out_n = (uint32_t)b.c;
    }
}

// LDS slot of a key: a multiplicative hash of its folded halves (cheaper
// than hmix's 64-bit multiplies, which the global table keeps).
__device__ __forceinline__ uint32_t lds_slot(uint64_t key) {
    const uint32_t a = (uint32_t)key ^ ((uint32_t)(key >> 32) * 0x85ebca6bu);
    return (a * 0x9E3779B1u) >> (32 - 11);
}
static_assert(kLdsSlots == 1 << 11, "lds_slot yields 11 bits");

// The LDS table has 2048 slots.  (8192, a quarter of the CU's LDS, was
// slower on a long-tailed column of 10^5 distinct values: 420 against 310
// us for 16M rows, as every workgroup then flushes more keys into the
// global table, whose device-scope atomics are the cost.)
// Workgroups of 1024 threads, one per CU, stride over tiles of 16384 rows,
// each thread taking 16 rows with 16-byte loads issued together.  Few, big
// workgroups keep the final flush short: every workgroup adds each of its
// distinct keys to the global table with a device-scope atomic, and those
// serialize per key.
constexpr int kHistThreads = 1024;
template <typename T>
__global__ __launch_bounds__(kHistThreads) void k_hist_count(const T* __restrict__ x, size_t n, unsigned long long* keys,
                                                       unsigned long long* counts, uint64_t cap_mask, HistCtl* ctl,
                                                       uint64_t* nan_keys /* [grid * 4] */,
                                                       uint32_t* nan_n /* [grid] */) {
    __shared__ unsigned long long skey[kLdsSlots];
    __shared__ uint32_t scnt[kLdsSlots];
    __shared__ uint64_t s_nk[kHistThreads / kWave * kMaxK];
    __shared__ int s_nc[kHistThreads / kWave];
    Small4 nan4;
    unsigned long long nans = 0;
    for (int i = threadIdx.x; i < kLdsSlots; i += kHistThreads) {
        skey[i] = kEmpty;
        scnt[i] = 0;
    }
    __syncthreads();
    bool negzero = false;
    unsigned long long special = 0;
    // Adds `add` rows of `key` to the LDS table (or, when its probe run is
    // full, to the global table).
    auto lds_add = [&](uint64_t key, uint32_t add) {
        const uint32_t h = lds_slot(key);
        for (int p = 0; p < 16; p++) {
            const uint32_t s = (h + p) & (kLdsSlots - 1);
            unsigned long long k = skey[s];
            if (k == kEmpty) k = atomicCAS(&skey[s], kEmpty, (unsigned long long)key);
            if (k == kEmpty || k == key) {
                atomicAdd(&scnt[s], add);
                return;
            }
        }
        global_insert(keys, counts, cap_mask, key, (unsigned long long)add, ctl);
    };
    // Two keys per thread counted in registers (the first two distinct keys
    // the thread meets): on skewed columns most rows hit them and skip the
    // LDS atomics, whose same-address conflicts are the cost.
    uint64_t rk0 = kEmpty, rk1 = kEmpty;
    uint32_t rc0 = 0, rc1 = 0;
    auto count = [&](T v) {
        if constexpr (std::is_floating_point<T>::value) {
            if (v != v) {
                nan4.insert(order_key(v));
                nans++;
                return;
            }
        }
        const uint64_t key = hist_key(v, negzero);
        if (key == kEmpty) {
            special++;
            return;
        }
        if (key == rk0) {
            rc0++;
        } else if (key == rk1) {
            rc1++;
        } else if (rk0 == kEmpty) {
            rk0 = key;
            rc0 = 1;
        } else if (rk1 == kEmpty) {
            rk1 = key;
            rc1 = 1;
        } else {
            lds_add(key, 1u);
        }
    };
    constexpr int kRows = 16 * kHistThreads;
    constexpr int V = 16 / sizeof(T);
    constexpr int Q = kRows / V / kHistThreads;  // 16-byte loads per thread and tile
    const size_t tiles = (n + kRows - 1) / kRows;
    const bool aligned = (reinterpret_cast<uintptr_t>(x) & 15) == 0;
    for (size_t t = blockIdx.x; t < tiles; t += gridDim.x) {  // uniform per workgroup
        const size_t base = t * kRows;
        if (aligned && base + kRows <= n) {
            const u32x4* s4 = reinterpret_cast<const u32x4*>(x + base);
            u32x4 w[Q];
#pragma unroll
            for (int q = 0; q < Q; q++) w[q] = s4[threadIdx.x + kHistThreads * q];
#pragma unroll
            for (int q = 0; q < Q; q++) {
                T v[V];
                __builtin_memcpy(v, &w[q], 16);
#pragma unroll
                for (int k = 0; k < V; k++) count(v[k]);
            }
        } else {
            const int cnt = (int)((n - base) < (size_t)kRows ? (n - base) : (size_t)kRows);
            for (int i = threadIdx.x; i < cnt; i += kHistThreads) count(x[base + i]);
        }
    }
    if (rk0 != kEmpty) lds_add(rk0, rc0);
    if (rk1 != kEmpty) lds_add(rk1, rc1);
    __syncthreads();
    for (int s = threadIdx.x; s < kLdsSlots; s += kHistThreads)
        if (skey[s] != kEmpty) global_insert(keys, counts, cap_mask, skey[s], scnt[s], ctl);
    const uint64_t nz = wave_or((uint64_t)negzero);
    const unsigned long long sp = wave_sum(special);
    const unsigned long long nn = wave_sum(nans);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        if (nz) atomicOr(&ctl->negzero, 1u);
        if (sp) atomicAdd(&ctl->special, sp);
        if (nn) atomicAdd(&ctl->nans, nn);
    }
    if constexpr (std::is_floating_point<T>::value)
        block_small4<kHistThreads>(nan4, s_nk, s_nc, nan_keys + (size_t)blockIdx.x * kMaxK, nan_n + blockIdx.x);
}

struct Cand {
    unsigned long long count;
    uint64_t key;
};

template <typename T> __device__ __forceinline__ bool better(const Cand& a, const Cand& b) {
    if (a.count != b.count) return a.count > b.count;
    T va, vb;
    using U = typename UnsignedOf<T>::type;
    U ua = (U)a.key, ub = (U)b.key;
    __builtin_memcpy(&va, &ua, sizeof(T));
    __builtin_memcpy(&vb, &ub, sizeof(T));
    return order_key(va) < order_key(vb);
}

// Inserts c into the sorted list top[0..K) (K <= kMaxK).  Loops run to
// kMaxK with guards so that a register-resident list keeps constant indices.
template <typename T> __device__ __forceinline__ void push(Cand* top, int K, Cand c) {
    if (c.count == 0) return;
    int pos = 0;  // entries that stay ahead of c
#pragma unroll
    for (int i = 0; i < kMaxK; i++)
        if (i < K && top[i].count != 0 && !better<T>(c, top[i])) pos = i + 1;
    if (pos >= K) return;
#pragma unroll
    for (int j = kMaxK - 1; j > 0; j--)
        if (j < K && j > pos) top[j] = top[j - 1];
#pragma unroll
    for (int i = 0; i < kMaxK; i++)
        if (i == pos) top[i] = c;
}

// a = the best K of the sorted lists a and b (empty entries have count 0).
template <typename T> __device__ __forceinline__ void merge_lists(Cand* a, const Cand* b, int K) {
    Cand o[kMaxK];
    int i = 0, j = 0;
#pragma unroll
    for (int k = 0; k < kMaxK; k++) {
        if (k >= K) break;
        const bool ta = i < K && a[i].count != 0, tb = j < K && b[j].count != 0;
        if (ta && (!tb || better<T>(a[i], b[j])))
            o[k] = a[i++];
        else if (tb)
            o[k] = b[j++];
        else
            o[k] = Cand{0, 0};
    }
#pragma unroll
    for (int k = 0; k < kMaxK; k++)
        if (k < K) a[k] = o[k];
}

// Pairwise tree merge of nl sorted lists of K in LDS into list 0.
template <typename T> __device__ void block_merge(Cand* sc, int nl, int K) {
    for (int s = 1; s < nl; s <<= 1) {
        __syncthreads();
        const int t = threadIdx.x;
        if (t < nl && t % (2 * s) == 0 && t + s < nl) merge_lists<T>(sc + t * K, sc + (t + s) * K, K);
    }
    __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_hist_topk(const unsigned long long* keys, const unsigned long long* counts,
                                                      uint64_t cap, int K, Cand* partials, uint32_t* counter,
                                                      const HistCtl* ctl, const uint64_t* nan_keys,
                                                      const uint32_t* nan_n, uint32_t nan_blocks, Cand* out) {
    __shared__ Cand sc[kBlock * kMaxK];
    __shared__ bool last;
    Cand top[kMaxK];
    for (int i = 0; i < kMaxK; i++) top[i] = Cand{0, 0};
    const size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t s = (size_t)blockIdx.x * kBlock + threadIdx.x; s < cap; s += stride) {
        const uint64_t k = keys[s];
        if (k != kEmpty) push<T>(top, K, Cand{counts[s], k});
    }
#pragma unroll
    for (int i = 0; i < kMaxK; i++)
        if (i < K) sc[threadIdx.x * K + i] = top[i];
    block_merge<T>(sc, kBlock, K);
    if (threadIdx.x == 0)
        for (int i = 0; i < K; i++) {
            st_agent(reinterpret_cast<uint64_t*>(&partials[blockIdx.x * K + i].count), (uint64_t)sc[i].count);
            st_agent(&partials[blockIdx.x * K + i].key, sc[i].key);
        }
    if (!last_block_arrives(counter, &last)) return;
    // the last workgroup gathers every workgroup's list into LDS and merges
    for (uint32_t i = threadIdx.x; i < gridDim.x * (uint32_t)K; i += kBlock) {
        Cand c;
        c.count = __hip_atomic_load(&partials[i].count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        c.key = __hip_atomic_load(&partials[i].key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        sc[i] = c;
    }
    block_merge<T>(sc, (int)gridDim.x, K);
    if (threadIdx.x != 0) return;
    Cand b[kMaxK];
    for (int i = 0; i < kMaxK; i++) b[i] = i < K ? sc[i] : Cand{0, 0};
    if (ctl->special) push<T>(b, K, Cand{ctl->special, kEmpty});
    if constexpr (std::is_floating_point<T>::value) {
        if (ctl->nans)  // every NaN row is a bin of count 1
            for (uint32_t g = 0; g < nan_blocks; g++)
                for (uint32_t r = 0; r < nan_n[g]; r++) {
                    const T v = from_order_key<T>(nan_keys[(size_t)g * kMaxK + r]);
                    push<T>(b, K, Cand{1ull, ubits(v)});
                }
    }
#pragma unroll
    for (int i = 0; i < kMaxK; i++)
        if (i < K) out[i] = b[i];
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The histogram's global table emptied (keys ~0, counts 0) and its control
// block cleared, in one launch.
__global__ void k_hist_init(unsigned long long* keys, unsigned long long* counts, size_t cap, HistCtl* ctl) {
    const size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < cap; i += stride) {
        keys[i] = kEmpty;
        counts[i] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *ctl = HistCtl{};
}

uint64_t pow2_at_least(uint64_t v) {
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

}  // namespace

void topk_most_frequent(Ctx* c, int dtype, const void* x, size_t n, int K, uint64_t* keys_bits_host,
                        int* n_distinct_host) {
    if (K < 1 || K > kMaxK) fail(DDJ_EINVAL, "top-K supports K in 1..4");
    if (n == 0) fail(DDJ_EINVAL, "top-K of an empty array");
    dispatch_native(dtype, [&](auto tag) {
        using T = decltype(tag);
        const uint64_t all_distinct = pow2_at_least(2 * (uint64_t)n + 2);
        uint64_t cap = std::min<uint64_t>(std::max<uint64_t>(1024, all_distinct), 1u << 16);
        for (int attempt = 0; attempt < 3; attempt++) {
            DevBuf keys(c, 8 * cap), counts(c, 8 * cap), ctl(c, sizeof(HistCtl));
            const unsigned fg = (unsigned)std::min<uint64_t>((cap + kBlock - 1) / kBlock, (uint64_t)c->cu_count * 16);
            {
                KernelTimer kt_(c, "k_hist_init");
                hipLaunchKernelGGL(k_hist_init, dim3(fg), dim3(kBlock), 0, c->stream, keys.get<unsigned long long>(),
                                   counts.get<unsigned long long>(), cap, ctl.get<HistCtl>());
            }
            size_t g = (n + 16 * kHistThreads - 1) / (16 * kHistThreads);  // at most one workgroup per tile and per CU
            g = std::max<size_t>(1, std::min<size_t>(g, (size_t)c->cu_count));
            DevBuf nan_keys(c, 8 * kMaxK * g), nan_n(c, 4 * g);
            { KernelTimer kt_(c, "k_hist_count"); hipLaunchKernelGGL(k_hist_count<T>, dim3((unsigned)g), dim3(kHistThreads), 0, c->stream,
                               static_cast<const T*>(x), n, keys.get<unsigned long long>(),
                               counts.get<unsigned long long>(), cap - 1, ctl.get<HistCtl>(),
                               nan_keys.get<uint64_t>(), nan_n.get<uint32_t>()); }
            HIPC(hipGetLastError());
            // the top-K runs before the table's state is read back, so that one
            // read-back brings both (its result is dropped if the table filled)
            const unsigned tg = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((cap + kBlock - 1) / kBlock, 128));
            DevBuf partials(c, sizeof(Cand) * K * tg), out(c, sizeof(Cand) * K);
            { KernelTimer kt_(c, "k_hist_topk"); hipLaunchKernelGGL(k_hist_topk<T>, dim3(tg), dim3(kBlock), 0, c->stream, keys.get<unsigned long long>(),
                               counts.get<unsigned long long>(), cap, K, partials.get<Cand>(), &c->control->counters[2],
                               ctl.get<HistCtl>(), nan_keys.get<uint64_t>(), nan_n.get<uint32_t>(), (uint32_t)g,
                               out.get<Cand>()); }
            HIPC(hipGetLastError());
            HistCtl h;
            Cand res[kMaxK];
            const ReadBack rb[2] = {{&h, ctl.get(), sizeof h}, {res, out.get(), sizeof(Cand) * K}};
            d2h_many(c, rb, 2);
            if (h.overflow || h.distinct * 10 > cap * 7) {
                if (attempt == 2 || cap >= all_distinct) fail(DDJ_ENOMEM, "histogram table overflow");
                const uint64_t seen = h.distinct + (h.overflow ? h.dropped : 0);
                const uint64_t next = attempt == 0 ? std::min(all_distinct, pow2_at_least(2 * seen + 2)) : all_distinct;
                cap = std::max(next, std::min(all_distinct, 2 * cap));  // always grows
                continue;
            }
            int nd = (int)(h.distinct + (h.special ? 1 : 0) + h.nans);
            if (n_distinct_host) *n_distinct_host = nd;
            // fewer than K distinct values: repeat the last real key
            int real = std::min(K, nd);
            for (int i = 0; i < K; i++) {
                const Cand& r = res[i < real ? i : real - 1];
                uint64_t kb = r.key;
                if (std::is_floating_point<T>::value && kb == 0 && h.negzero) kb = (sizeof(T) == 8) ? (1ull << 63) : (1ull << 31);
                keys_bits_host[i] = kb;
            }
            return;
        }
    });
}

}  // namespace ddj
