// Standalone unique encoding (UniqueEncoding, reference
// compression/unique/unique_encoding.cu:35-300).
//
// FindUnique (:35-46) sorts the column and keeps the first of each run of
// equal neighbours; every row is then coded as the index of its value in
// that list, ALT_BITLEN(U - 1) bits a code, packed 32 / b codes to a 32-bit
// word from the low bits up (_compressUniqueKernel, :48-76).
//
// Here: the column's order keys (the radix order device thrust sorts
// floats in: -0.0 just before +0.0, sign-set NaNs first, other NaNs last)
// are radix-sorted; a counting pass marks each sorted row that differs from
// its predecessor under T's == (so -0.0 and +0.0 merge into the first, and
// every NaN row stays), a tile scan gives each kept row its index, and the
// writing pass stores the values straight into the metadata after its two
// size_t counts, and their keys for the coder.  The coder takes one output
// word per thread and finds each code by binary search over the keys (in
// LDS when they fit).  A NaN row equals no listed value: code 0, as in the
// reference.  Decoding is a gather, one row per thread.
#include "dispatch.h"
#include "k_common.h"
#include "kernels.h"

namespace ddj {

namespace {

constexpr int kUniqTile = 4096;  // sorted rows per workgroup of the counting and writing passes
constexpr int kUniqPer = kUniqTile / kBlock;
constexpr int kLdsKeyBytes = 32 << 10;

template <typename T> using KeyOf = typename std::conditional<sizeof(T) == 8, uint64_t, uint32_t>::type;

template <typename T>
__global__ __launch_bounds__(kBlock) void k_order_keys(const T* __restrict__ x, size_t n, KeyOf<T>* __restrict__ keys) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        keys[i] = (KeyOf<T>)order_key(x[i]);
}

// Sorted row i starts a new listed value: it is the first row, or its value
// differs (under T's ==) from row i - 1's.
template <typename T> __device__ __forceinline__ bool starts_value(const KeyOf<T>* keys, size_t i) {
    if (i == 0) return true;
    const T a = from_order_key<T>(keys[i - 1]), b = from_order_key<T>(keys[i]);
    return !(a == b);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_unique_count(const KeyOf<T>* __restrict__ keys, size_t n,
                                                         uint64_t* __restrict__ vals) {
    __shared__ uint64_t red[kBlock / kWave];
    const size_t base = (size_t)blockIdx.x * kUniqTile;
    uint64_t s = 0;
#pragma unroll
    for (int r = 0; r < kUniqPer; r++) {
        const size_t i = base + (size_t)r * kBlock + threadIdx.x;
        if (i < n) s += starts_value<T>(keys, i);
    }
    uint64_t total;
    block_exclusive_scan<uint64_t>(s, &total, red);
    if (threadIdx.x == 0) vals[blockIdx.x] = total;
}

// Each thread takes 16 consecutive sorted rows, so that the kept rows'
// order follows the thread order of the block scan.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_unique_write(const KeyOf<T>* __restrict__ keys, size_t n,
                                                         const uint64_t* __restrict__ prefix, uint8_t* __restrict__ vals,
                                                         KeyOf<T>* __restrict__ ukeys) {
    __shared__ uint64_t red[kBlock / kWave];
    const size_t first = (size_t)blockIdx.x * kUniqTile + (size_t)threadIdx.x * kUniqPer;
    uint32_t flags = 0, cnt = 0;
#pragma unroll
    for (int k = 0; k < kUniqPer; k++) {
        const size_t i = first + k;
        if (i < n && starts_value<T>(keys, i)) {
            flags |= 1u << k;
            cnt++;
        }
    }
    uint64_t total;
    uint64_t at = prefix[blockIdx.x] + block_exclusive_scan<uint64_t>(cnt, &total, red);
#pragma unroll
    for (int k = 0; k < kUniqPer; k++) {
        if (flags & (1u << k)) {
            const KeyOf<T> key = keys[first + k];
            const T v = from_order_key<T>(key);
            __builtin_memcpy(vals + at * sizeof(T), &v, sizeof(T));  // vals sits 16 bytes into the metadata
            ukeys[at] = key;
            at++;
        }
    }
}

// Index of v's listed value: the first key not below v's (zeros search from
// -0.0's key, so +0.0 rows find a listed -0.0).  NaN: no listed value, 0.
template <typename T>
__device__ __forceinline__ uint32_t code_of(T v, const KeyOf<T>* uk, uint32_t U) {
    KeyOf<T> key;
    if constexpr (std::is_floating_point<T>::value) {
        if (v != v) return 0;
        key = (KeyOf<T>)order_key(v == T(0) ? -T(0) : v);
    } else {
        key = (KeyOf<T>)order_key(v);
    }
    uint32_t lo = 0, hi = U;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (uk[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

template <typename T, bool LDS>
__global__ __launch_bounds__(kBlock) void k_unique_encode(const T* __restrict__ x, size_t n,
                                                          const KeyOf<T>* __restrict__ ukeys, uint32_t U, int b, int P,
                                                          uint32_t* __restrict__ words, size_t nwords) {
    __shared__ KeyOf<T> sk[LDS ? kLdsKeyBytes / sizeof(KeyOf<T>) : 1];
    const KeyOf<T>* uk = ukeys;
    if (LDS) {
        for (uint32_t j = threadIdx.x; j < U; j += kBlock) sk[j] = ukeys[j];
        __syncthreads();
        uk = sk;
    }
    for (size_t w = (size_t)blockIdx.x * kBlock + threadIdx.x; w < nwords; w += (size_t)gridDim.x * kBlock) {
        uint32_t word = 0;
        const size_t r0 = w * (size_t)P;
        for (int i = 0; i < P; i++) {
            if (r0 + i >= n) break;
            word |= code_of<T>(x[r0 + i], uk, U) << (b * i);
        }
        words[w] = word;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_unique_decode(const uint32_t* __restrict__ words, const uint8_t* uvals,
                                                          uint32_t U, int b, int P, size_t n, T* __restrict__ out,
                                                          uint32_t* err) {
    const uint32_t mask = b >= 32 ? ~0u : ((1u << b) - 1u);
    bool bad = false;
    for (size_t r = (size_t)blockIdx.x * kBlock + threadIdx.x; r < n; r += (size_t)gridDim.x * kBlock) {
        const uint32_t code = (words[r / P] >> (b * (int)(r % P))) & mask;
        T v = T(0);
        if (code < U)
            __builtin_memcpy(&v, uvals + (size_t)code * sizeof(T), sizeof(T));
        else
            bad = true;
        out[r] = v;
    }
    if (bad) __hip_atomic_store(err, 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

unsigned grid_cap(Ctx* c, size_t threads) {
    const size_t g = (threads + kBlock - 1) / kBlock;
    const size_t cap = (size_t)c->cu_count * 16;
    return (unsigned)(g < cap ? (g ? g : 1) : cap);
}

// ALT_BITLEN (core/cuda_macros.cuh:191-205) of U - 1 >= 0: bits of the
// largest code, at least 1.
int code_bits(uint64_t U) {
    uint64_t v = U - 1;
    int r = 0;
    while (v >>= 1) r++;
    return r + 1;
}

// The sorted keys of the column and, per 4096 sorted rows, how many listed
// values begin there (scanned).
template <typename T> struct SortedUnique {
    DevBuf a, b;
    KeyOf<T>* sorted = nullptr;
    std::unique_ptr<TileScan> ts;
    size_t tiles = 0;
    uint64_t U = 0;
};

template <typename T> void sort_and_count(Ctx* c, const T* x, size_t n, SortedUnique<T>& s) {
    using K = KeyOf<T>;
    s.a = DevBuf(c, n * sizeof(K));
    s.b = DevBuf(c, n * sizeof(K));
    kernel_bytes(c, "k_order_keys", (double)n * (sizeof(T) + sizeof(K)));
    {
        KernelTimer kt_(c, "k_order_keys");
        hipLaunchKernelGGL(k_order_keys<T>, dim3(grid_cap(c, n)), dim3(kBlock), 0, c->stream, x, n, s.a.template get<K>());
    }
    HIPC(hipGetLastError());
    s.sorted = radix_sort<K>(c, s.a.template get<K>(), s.b.template get<K>(), n, 8 * (int)sizeof(T));
    s.tiles = (n + kUniqTile - 1) / kUniqTile;
    s.ts = std::make_unique<TileScan>(c, s.tiles);
    kernel_bytes(c, "k_unique_count", (double)n * sizeof(K));
    {
        KernelTimer kt_(c, "k_unique_count");
        hipLaunchKernelGGL(k_unique_count<T>, dim3((unsigned)s.tiles), dim3(kBlock), 0, c->stream,
                           (const K*)s.sorted, n, s.ts->vals);
    }
    HIPC(hipGetLastError());
    s.ts->scan(c);
    d2h(c, &s.U, s.ts->prefix + s.tiles, sizeof(uint64_t));
}

}  // namespace

uint64_t unique_count(Ctx* c, int dtype, const void* x, size_t n) {
    if (n == 0) return 0;
    return dispatch_native(dtype, [&](auto tag) -> uint64_t {
        using T = decltype(tag);
        SortedUnique<T> s;
        sort_and_count<T>(c, static_cast<const T*>(x), n, s);
        return s.U;
    });
}

void unique_encode_dev(Ctx* c, int dtype, const void* x, size_t n, DevBuf& meta, DevBuf& data) {
    if (n == 0) return;
    dispatch_native(dtype, [&](auto tag) {
        using T = decltype(tag);
        using K = KeyOf<T>;
        SortedUnique<T> s;
        sort_and_count<T>(c, static_cast<const T*>(x), n, s);
        const uint64_t U = s.U;
        if (U == 0 || U > 0xffffffffull) fail(DDJ_EUNSUPPORTED, "unique: more than 2^32 - 1 distinct values");
        const int b = code_bits(U);
        const int P = 32 / b;
        const size_t nwords = (n + P - 1) / P;
        meta = DevBuf(c, 16 + U * sizeof(T));
        const uint64_t head[2] = {U, (uint64_t)n};
        h2d(c, meta.get(), head, sizeof head);
        DevBuf ukeys(c, U * sizeof(K));
        kernel_bytes(c, "k_unique_write", (double)n * sizeof(K) + (double)U * (sizeof(T) + sizeof(K)));
        {
            KernelTimer kt_(c, "k_unique_write");
            hipLaunchKernelGGL(k_unique_write<T>, dim3((unsigned)s.tiles), dim3(kBlock), 0, c->stream,
                               (const K*)s.sorted, n, (const uint64_t*)s.ts->prefix, meta.bytes_ptr() + 16,
                               ukeys.get<K>());
        }
        HIPC(hipGetLastError());
        data = DevBuf(c, nwords * 4);
        kernel_bytes(c, "k_unique_encode", (double)n * sizeof(T) + (double)nwords * 4);
        KernelTimer kt_(c, "k_unique_encode");
        if (U * sizeof(K) <= (uint64_t)kLdsKeyBytes)
            hipLaunchKernelGGL((k_unique_encode<T, true>), dim3(grid_cap(c, nwords)), dim3(kBlock), 0, c->stream,
                               static_cast<const T*>(x), n, (const K*)ukeys.get<K>(), (uint32_t)U, b, P,
                               data.get<uint32_t>(), nwords);
        else
            hipLaunchKernelGGL((k_unique_encode<T, false>), dim3(grid_cap(c, nwords)), dim3(kBlock), 0, c->stream,
                               static_cast<const T*>(x), n, (const K*)ukeys.get<K>(), (uint32_t)U, b, P,
                               data.get<uint32_t>(), nwords);
        HIPC(hipGetLastError());
    });
}

void unique_decode_dev(Ctx* c, int dtype, const uint8_t* uvals, uint64_t U, uint64_t n, const uint8_t* words,
                       size_t nwords, void* out) {
    if (n == 0) return;
    if (U == 0 || U > 0xffffffffull) fail(DDJ_ECORRUPT, "unique: bad value count");
    const int b = code_bits(U);
    const int P = 32 / b;
    if ((uint64_t)nwords * P < n) fail(DDJ_ECORRUPT, "unique: fewer code words than rows");
    if (reinterpret_cast<uintptr_t>(words) & 3) fail(DDJ_EINVAL, "unique: code words not 4-byte aligned");
    arm_error_flag(c);
    dispatch_native(dtype, [&](auto tag) {
        using T = decltype(tag);
        kernel_bytes(c, "k_unique_decode", (double)n * sizeof(T) + (double)nwords * 4);
        KernelTimer kt_(c, "k_unique_decode");
        hipLaunchKernelGGL(k_unique_decode<T>, dim3(grid_cap(c, n)), dim3(kBlock), 0, c->stream,
                           reinterpret_cast<const uint32_t*>(words), uvals, (uint32_t)U, b, P, (size_t)n,
                           static_cast<T*>(out), &c->control->error_flag);
    });
    HIPC(hipGetLastError());
}

}  // namespace ddj
