/*
 * ddj_oracle.cpp -- CPU restatement of the reference codecs (see ddj_oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
 *
 * Written in plain sequential style: one loop per reference kernel or Thrust
 * call, in the order the reference performs them, with the reference's
 * integer widths and wrap-around.  Citations are file:line in the reference's
 * src/ tree.  Where the reference reads uninitialised or out-of-bounds memory
 * this restatement writes zeros (documented per site).
 */
#include "ddj_oracle.h"
#include "fl_loops.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <type_traits>
#include <vector>

namespace {

using namespace orc_fl;

typedef std::vector<uint8_t> Bytes;
typedef std::vector<Bytes> Outs;

enum { OK = 0, EINVAL_ = -1, EUNSUP = -2, ECORRUPT = -3 };

template <typename T> const T* as(const Bytes& b) { return reinterpret_cast<const T*>(b.data()); }
template <typename T> size_t count(const Bytes& b) { return b.size() / sizeof(T); }
template <typename T> Bytes from(const T* p, size_t n) {
    Bytes b(n * sizeof(T));
    if (n) memcpy(b.data(), p, n * sizeof(T));
    return b;
}
template <typename T> Bytes from(const std::vector<T>& v) { return from(v.data(), v.size()); }
template <typename T> std::vector<T> vec(const Bytes& b) {
    std::vector<T> v(count<T>(b));
    if (!v.empty()) memcpy(v.data(), b.data(), v.size() * sizeof(T));
    return v;
}
template <typename T> Bytes scalar(T v) { return from(&v, 1); }

/* BITLEN (core/cuda_macros.cuh:168-175): bit width of max, or all bits when
 * min is negative.  The float/double specialisations return 32/64 (:184-189). */
template <typename T> int bitlen(T mn, T mx) {
    if constexpr (std::is_floating_point<T>::value) {
        return 8 * sizeof(T);
    } else {
        if (mn < 0) return 8 * sizeof(T);
        int ret = 0;
        while (mx >>= 1) ret++;
        return ret + 1;
    }
}

/* ALT_BITLEN (core/cuda_macros.cuh:191-205) */
int alt_bitlen(int v) {
    if (v < 0) return 32;
    unsigned r = 0;
    unsigned u = (unsigned)v;
    while (u >>= 1) r++;
    return (int)r + 1;
}

/* --------------------------------------------------------------- statistics */
template <typename T> void minmax(const T* x, size_t n, T& mn, T& mx) {
    /* thrust::minmax_element (cuda_array_statistics.cu:20-28): first minimum
     * and first maximum under operator<. */
    mn = x[0];
    mx = x[0];
    for (size_t i = 1; i < n; i++) {
        if (x[i] < mn) mn = x[i];
        if (mx < x[i]) mx = x[i];
    }
}

/* _getFloatPrecision (cuda_array_statistics.cu:37-44): multiply by 10^k
 * until round(x*10^k)/10^k == x in float arithmetic, k <= MAX_PRECISION (10).
 * The reference returns logf(10^k)/logf(10) truncated to int; on the host
 * (where its known-answer tests run) that equals k for every k in 0..10. */
int float_precision(float number) {
    long e = 1;
    const long maxPrecision = 10000000000L;
    int k = 0;
    while (roundf(number * (float)e) / (float)e != number && e < maxPrecision) {
        e *= 10;
        k++;
    }
    return k;
}

/* float -> int conversion as the GPU performs it (saturating, NaN -> 0); used
 * where the reference converts T to int inside a kernel (_rlMetricKernel). */
template <typename T> int32_t to_int_sat(T v) {
    if (std::is_floating_point<T>::value) {
        double d = (double)v;
        if (d != d) return 0;
        if (d >= 2147483647.0) return 2147483647;
        if (d <= -2147483648.0) return (-2147483647 - 1);
        return (int32_t)d;
    }
    return (int32_t)(int64_t)v; /* integral: wrap to 32 bits */
}

template <typename T> T mean_of(const T* x, size_t n) {
    /* CudaArrayStatistics::Mean (cuda_array_statistics.cu:134-139): sum in T
     * (wrapping for integers), divided in double, converted back to T.  The
     * floating-point sum order of thrust::reduce is not pinned. */
    T s = 0;
    for (size_t i = 0; i < n; i++) s = (T)(s + x[i]);
    return (T)(s / (double)n);
}

template <typename T> orc_stats stats_of(const T* x, size_t n) {
    /* getStatistics (cuda_array_statistics.cu:141-156) */
    orc_stats st;
    memset(&st, 0, sizeof(st));
    st.size = n;
    if (n == 0) return st;
    T mn, mx;
    minmax(x, n, mn, mx);
    st.min = (double)mn;
    st.max = (double)mx;
    st.minBitCnt = 64; /* BITLEN(double, double) -> 64 (cuda_macros.cuh:187-189) */
    int prec = 0;
    for (size_t i = 0; i < n; i++) prec = std::max(prec, float_precision((float)x[i]));
    st.precision = prec;
    int unsorted = 0; /* _sortedKernel (:69-75) */
    for (size_t i = 0; i + 1 < n; i++)
        if (x[i] > x[i + 1]) unsorted = 1;
    st.sorted = !unsorted;
    /* RlMetric<T,2> (:97-132): mean over idx < n-2 of 1 + [int(x[idx]) == x[idx+1]] */
    if (2 >= n) {
        st.rlMetric = 1.0f;
    } else {
        size_t m = n - 2;
        int32_t sum = 0;
        for (size_t i = 0; i < m; i++) {
            int32_t v = to_int_sat(x[i]);
            int eq;
            if (std::is_floating_point<T>::value)
                eq = ((T)v == x[i + 1]);
            else
                eq = ((int64_t)v == (int64_t)x[i + 1]);
            sum += 1 + eq;
        }
        st.rlMetric = (float)(sum / (double)m);
    }
    st.mean = (double)mean_of(x, n);
    return st;
}

/* -------------------------------------------------------------- histograms
 * ThrustSparseHistogram (util/histogram/thrust_sparse_histogram.cu:20-69):
 * thrust::sort of the keys, then reduce_by_key with equal_to.  Device thrust
 * sorts floating-point keys by radix order, i.e. by the bit pattern mapped to
 * an unsigned key (negatives reversed): -0.0 sorts just before +0.0, NaNs
 * with the sign bit set before everything, other NaNs after everything.  The
 * equal_to grouping then merges -0.0 with +0.0 (the bin keeps the first,
 * -0.0) and never merges a NaN with anything, so every NaN row is a bin of
 * its own with count 1.  GetMostFrequentSparse (most_frequent.cu:10-26):
 * sort bins by count, descending; equal counts keep key order (the
 * reference's sort_by_key is not guaranteed stable, so this tie order is
 * the documented choice).  When fewer than K bins exist the reference leaves
 * the tail uninitialised; this restatement repeats the last real key there
 * (both are outside what its tests pin). */
template <typename T> uint64_t radix_key(T v) {
    if (std::is_floating_point<T>::value) {
        uint64_t b = 0;
        memcpy(&b, &v, sizeof(T));
        const int W = 8 * (int)sizeof(T);
        const uint64_t sign = 1ull << (W - 1);
        const uint64_t mask = W == 64 ? ~0ull : ((1ull << W) - 1);
        return (b & sign) ? (~b & mask) : (b | sign);
    }
    const int W = 8 * (int)sizeof(T);
    return W == 64 ? ((uint64_t)(int64_t)v ^ (1ull << 63)) : (uint64_t)((int64_t)v + (1ll << (W - 1)));
}

template <typename T> std::vector<T> most_frequent(const T* x, size_t n, int K, int* ndistinct) {
    std::vector<T> keys(x, x + n);
    std::stable_sort(keys.begin(), keys.end(), [](T a, T b) { return radix_key(a) < radix_key(b); });
    std::vector<std::pair<T, int32_t>> bins;
    for (size_t i = 0; i < n; i++) {
        if (i > 0 && keys[i] == bins.back().first)
            bins.back().second++;
        else
            bins.push_back(std::make_pair(keys[i], 1));
    }
    std::stable_sort(bins.begin(), bins.end(),
                     [](const std::pair<T, int32_t>& a, const std::pair<T, int32_t>& b) { return a.second > b.second; });
    if (ndistinct) *ndistinct = (int)bins.size();
    std::vector<T> out((size_t)K);
    for (int i = 0; i < K; i++) {
        if ((size_t)i < bins.size())
            out[i] = bins[i].first;
        else
            out[i] = bins.empty() ? (T)0 : bins.back().first;
    }
    return out;
}

/* ----------------------------------------------------------------- stencil
 * Stencil::pack (util/stencil/stencil.cu:14-33, :52-67): byte 0 = n % 8, then
 * ceil(n/8) bytes, bit j of byte q = s[8q+j]; rows past n repeat s[n-1]. */
Bytes stencil_pack(const std::vector<int32_t>& s) {
    size_t n = s.size();
    size_t chars = (n + 7) / 8;
    Bytes out(chars + 1, 0);
    out[0] = (uint8_t)(n % 8);
    for (size_t q = 0; q < chars; q++) {
        uint8_t part = 0;
        for (int i = 0; i < 8; i++) {
            size_t idx = 8 * q + i;
            if (idx >= n) idx = n - 1;
            if (s[idx] & 1) part |= (uint8_t)(1u << i);
        }
        out[1 + q] = part;
    }
    return out;
}

/* Stencil::unpack (stencil.cu:69-96): n = 8*(len-1) - (rest ? 8-rest : 0) */
int stencil_unpack(const uint8_t* p, size_t len, std::vector<int32_t>& s) {
    if (len < 1) return ECORRUPT;
    int rest = (int8_t)p[0];
    long long n = (long long)(len - 1) * 8;
    if (rest) n -= 8 - rest;
    if (n < 0) return ECORRUPT;
    s.assign((size_t)n, 0);
    for (size_t i = 0; i < (size_t)n; i++) s[i] = (p[1 + i / 8] >> (i % 8)) & 1;
    return OK;
}

/* Splitter::Split (util/splitter/splitter.cu:67-100): stable partition */
template <typename T> void split(const T* x, const std::vector<int32_t>& s, std::vector<T>& yes, std::vector<T>& no) {
    yes.clear();
    no.clear();
    for (size_t i = 0; i < s.size(); i++) (s[i] ? yes : no).push_back(x[i]);
}

/* Splitter::Merge (splitter.cu:159-184) */
template <typename T> int merge(const std::vector<int32_t>& s, const Bytes& yesb, const Bytes& nob, std::vector<T>& out) {
    size_t ny = count<T>(yesb), nn = count<T>(nob), iy = 0, in = 0;
    out.assign(s.size(), (T)0);
    for (size_t i = 0; i < s.size(); i++) {
        if (s[i]) {
            if (iy >= ny) return ECORRUPT;
            memcpy(&out[i], yesb.data() + sizeof(T) * iy++, sizeof(T));
        } else {
            if (in >= nn) return ECORRUPT;
            memcpy(&out[i], nob.data() + sizeof(T) * in++, sizeof(T));
        }
    }
    return OK;
}

/* AflEncoding::Encode<T> (compression/afl/afl_encoding.cu:15-56).  Metadata
 * bytes 2..3 are uninitialised pinned memory in the reference; written 0. */
template <typename T> int afl_encode(const T* x, size_t n, Outs& outs) {
    outs.assign(2, Bytes());
    if (n == 0) return OK;
    T mn, mx;
    minmax(x, n, mn, mx);
    int minBit = bitlen(mn, mx);
    const uint64_t W = 8 * sizeof(T);
    uint64_t words = (minBit * (uint64_t)n + W - 1) / W;
    uint64_t bytes = words * sizeof(T);
    int rest = (int)((bytes * 8 - (uint64_t)n * minBit) & 0xff);
    std::vector<T> packed(words, (T)0);
    fl_compress<T>(minBit, x, packed.data(), n, words);
    outs[0] = Bytes{(uint8_t)minBit, (uint8_t)rest, 0, 0};
    outs[1] = from(packed);
    return OK;
}

/* DecodeAfl / AflEncoding::Decode<T> (afl_encoding.cu:74-113) */
template <typename T> int afl_decode_raw(const uint8_t* data, size_t size, int minBit, int rest, std::vector<T>& out) {
    if (minBit <= 0 || minBit > (int)(8 * sizeof(T))) return ECORRUPT;
    long long comprBits = (long long)size * 8 - rest;
    if (comprBits < 0) return ECORRUPT;
    long long length = comprBits / minBit;
    out.assign((size_t)length, (T)0);
    std::vector<T> words(size / sizeof(T));
    if (!words.empty()) memcpy(words.data(), data, words.size() * sizeof(T));
    fl_decompress<T>(minBit, words.data(), out.data(), length, words.size());
    return OK;
}

template <typename T> int afl_decode(const Outs& ins, std::vector<T>& out) {
    out.clear();
    if (ins.size() < 2) return EINVAL_;
    if (ins[1].empty()) return OK;
    if (ins[0].size() < 2) return ECORRUPT;
    int minBit = (int8_t)ins[0][0];
    int rest = (int8_t)ins[0][1];
    return afl_decode_raw<T>(ins[1].data(), ins[1].size(), minBit, rest, out);
}

/* AflEncoding::Encode<float> (afl_encoding.cu:115-167): split into
 * (mantissa, exponent, sign) int arrays (_splitFloatKernel :58-72), FL-encode
 * mantissa and exponent; metadata = size_t |FL(mantissa)| + char sign
 * (1: all >= 0, 2: all < 0, 0: mixed, then a packed sign stencil). */
int afl_float_encode(const float* x, size_t n, Outs& outs) {
    outs.assign(2, Bytes());
    if (n == 0) return OK;
    float mn, mx;
    minmax(x, n, mn, mx);
    char sign = (char)((mn >= 0 ? 1 : 0) + (mx < 0 ? 2 : 0));
    std::vector<int32_t> mant(n), expo(n), sgn(n);
    for (size_t i = 0; i < n; i++) {
        uint32_t u;
        memcpy(&u, &x[i], 4);
        mant[i] = (int32_t)(u & 0x7fffffu);
        expo[i] = (int32_t)((u >> 23) & 0xffu);
        sgn[i] = (int32_t)(u >> 31);
    }
    Outs a, b;
    afl_encode(mant.data(), n, a);
    afl_encode(expo.data(), n, b);
    size_t msize = a[1].size();
    Bytes meta(sizeof(size_t) + 1);
    memcpy(meta.data(), &msize, sizeof(size_t));
    meta[sizeof(size_t)] = (uint8_t)sign;
    if (sign == 0) {
        Bytes st = stencil_pack(sgn);
        meta.insert(meta.end(), st.begin(), st.end());
    }
    Bytes data;
    for (auto* p : {&a[0], &a[1], &b[0], &b[1]}) data.insert(data.end(), p->begin(), p->end());
    outs[0] = meta;
    outs[1] = data;
    return OK;
}

/* AflEncoding::Decode<float> (afl_encoding.cu:185-254) */
int afl_float_decode(const Outs& ins, std::vector<float>& out) {
    out.clear();
    if (ins.size() < 2) return EINVAL_;
    const Bytes& meta = ins[0];
    const Bytes& data = ins[1];
    if (data.empty()) return OK;
    if (meta.size() < sizeof(size_t) + 1 || data.size() < 8) return ECORRUPT;
    size_t msize;
    memcpy(&msize, meta.data(), sizeof(size_t));
    int sign = (int8_t)meta[sizeof(size_t)];
    if (msize + 8 > data.size()) return ECORRUPT;
    std::vector<int32_t> mant, expo;
    int rc = afl_decode_raw<int32_t>(data.data() + 4, msize, (int8_t)data[0], (int8_t)data[1], mant);
    if (rc) return rc;
    size_t esize = data.size() - msize - 8;
    const uint8_t* eb = data.data() + 4 + msize;
    rc = afl_decode_raw<int32_t>(eb + 4, esize, (int8_t)eb[0], (int8_t)eb[1], expo);
    if (rc) return rc;
    size_t n = mant.size();
    std::vector<int32_t> sgn;
    if (sign) {
        sgn.assign(n, sign - 1);
    } else {
        rc = stencil_unpack(meta.data() + sizeof(size_t) + 1, meta.size() - sizeof(size_t) - 1, sgn);
        if (rc) return rc;
    }
    if (expo.size() < n || sgn.size() < n) return ECORRUPT;
    out.resize(n);
    for (size_t i = 0; i < n; i++) {
        uint32_t u = ((uint32_t)(sgn[i] & 1) << 31) | (((uint32_t)expo[i] & 0xffu) << 23) | ((uint32_t)mant[i] & 0x7fffffu);
        memcpy(&out[i], &u, 4);
    }
    return OK;
}

/* ------------------------------------------------------------------- delta
 * DeltaEncoding (compression/delta/delta_encoding.cu:19-102): metadata =
 * x[0]; data[i] = x[i+1] - x[i] wrapping in T; decode = x[0] + inclusive scan. */
template <typename T> int delta_encode(const T* x, size_t n, Outs& outs) {
    outs.assign(2, Bytes());
    if (n == 0) return OK;
    std::vector<T> d(n - 1);
    for (size_t i = 0; i + 1 < n; i++) d[i] = (T)(ubits(x[i + 1]) - ubits(x[i]));
    outs[0] = scalar(x[0]);
    outs[1] = from(d);
    return OK;
}
template <typename T> int delta_decode(const Outs& ins, std::vector<T>& out) {
    out.clear();
    if (ins.size() < 2) return EINVAL_;
    if (ins[0].empty()) return OK;
    if (ins[0].size() < sizeof(T)) return ECORRUPT;
    T x0 = as<T>(ins[0])[0];
    size_t m = count<T>(ins[1]);
    out.resize(m + 1);
    uint64_t acc = 0;
    out[0] = x0;
    for (size_t i = 0; i < m; i++) {
        T di;
        memcpy(&di, ins[1].data() + i * sizeof(T), sizeof(T));
        acc += ubits(di);
        out[i + 1] = (T)(ubits(x0) + acc);
    }
    return OK;
}

/* ------------------------------------------------------------------- scale
 * ScaleEncoding (compression/scale/scale_encoding.cu:30-108): m = min; the
 * overflow guard (:49-50) compares numeric_limits<T>::max() - max (wrapping for
 * int/long; promoted to int for short/char) with std::abs<T>(min), which keeps
 * T's minimum unchanged; when it fires m = 0.  data = x - m (wrapping). */
template <typename T> T scale_min(const T* x, size_t n) {
    T mn, mx;
    minmax(x, n, mn, mx);
    T m = mn;
    if (mn < 0) {
        long long tmax = (long long)std::numeric_limits<T>::max();
        long long diff;
        if (sizeof(T) >= 4)
            diff = (long long)(T)(ubits((T)tmax) - ubits(mx)); /* wraps in T */
        else
            diff = tmax - (long long)mx;
        long long absmin = (mn == std::numeric_limits<T>::min()) ? (long long)mn : -(long long)mn;
        if (diff < absmin) m = 0;
    }
    return m;
}
template <typename T> int scale_encode(const T* x, size_t n, Outs& outs) {
    outs.assign(2, Bytes());
    if (n == 0) return OK;
    T m = scale_min(x, n);
    std::vector<T> d(n);
    for (size_t i = 0; i < n; i++) d[i] = (T)(ubits(x[i]) - ubits(m));
    outs[0] = scalar(m);
    outs[1] = from(d);
    return OK;
}
template <typename T> int scale_decode(const Outs& ins, std::vector<T>& out) {
    out.clear();
    if (ins.size() < 2) return EINVAL_;
    if (ins[1].empty()) return OK;
    if (ins[0].size() < sizeof(T)) return ECORRUPT;
    T m = as<T>(ins[0])[0];
    size_t k = count<T>(ins[1]);
    out.resize(k);
    for (size_t i = 0; i < k; i++) {
        T v;
        memcpy(&v, ins[1].data() + i * sizeof(T), sizeof(T));
        out[i] = (T)(ubits(v) + ubits(m));
    }
    return OK;
}

/* --------------------------------------------------------------------- rle
 * RleEncoding (compression/rle/rle_encode.cu:18-62, rle_decode.cu:19-73):
 * thrust::reduce_by_key with operator== (so +0.0 and -0.0 join one run whose
 * value is the run's first element); metadata = int32 R, out1 = int32
 * lengths[R], out2 = T values[R]. */
template <typename T> int rle_encode(const T* x, size_t n, Outs& outs) {
    outs.assign(3, Bytes());
    std::vector<int32_t> lens;
    std::vector<T> vals;
    for (size_t i = 0; i < n; i++) {
        if (i > 0 && x[i] == vals.back())
            lens.back()++;
        else {
            vals.push_back(x[i]);
            lens.push_back(1);
        }
    }
    int32_t R = (int32_t)vals.size();
    outs[0] = scalar(R);
    outs[1] = from(lens);
    outs[2] = from(vals);
    return OK;
}
template <typename T> int rle_decode(const Outs& ins, std::vector<T>& out) {
    out.clear();
    if (ins.size() < 3) return EINVAL_;
    if (ins[1].empty()) return OK;
    if (ins[0].size() < 4) return ECORRUPT;
    int32_t R = as<int32_t>(ins[0])[0];
    if (R < 0 || (size_t)R > count<int32_t>(ins[1]) || (size_t)R > count<T>(ins[2])) return ECORRUPT;
    for (int32_t r = 0; r < R; r++) {
        int32_t len = as<int32_t>(ins[1])[r];
        T v;
        memcpy(&v, ins[2].data() + r * sizeof(T), sizeof(T));
        if (len < 0) return ECORRUPT;
        out.insert(out.end(), (size_t)len, v);
    }
    return OK;
}

/* ------------------------------------------------------------------ unique
 * UniqueEncoding::CompressUnique (compression/unique/unique_encoding.cu:88-131)
 * and _compressUniqueKernel (:48-76): [size_t U][size_t m][U x T values] then
 * ceil(m/P) uint32 words, P = 32 / ALT_BITLEN(U-1); slot i of word w holds the
 * OR of the indices j with x[w*P+i] == unique[j].  The reference reads past the
 * end of x for the last word's spare slots; they are 0 here. */
template <typename T> void pack_codes(const std::vector<T>& x, const std::vector<T>& uniq, std::vector<uint32_t>& words) {
    int bits = alt_bitlen((int)uniq.size() - 1);
    int P = 32 / bits;
    size_t m = x.size();
    words.assign((m + P - 1) / P, 0u);
    for (size_t w = 0; w < words.size(); w++) {
        uint32_t r = 0;
        for (int i = 0; i < P; i++) {
            size_t idx = w * P + i;
            if (idx >= m) break;
            for (size_t j = 0; j < uniq.size(); j++)
                if (x[idx] == uniq[j]) r |= (uint32_t)j << (bits * i);
        }
        words[w] = r;
    }
}
template <typename T> Bytes compress_unique(const std::vector<T>& x, const std::vector<T>& uniq) {
    if (x.empty()) return Bytes();
    std::vector<uint32_t> words;
    pack_codes(x, uniq, words);
    size_t sizes[2] = {uniq.size(), x.size()};
    Bytes out(2 * sizeof(size_t));
    memcpy(out.data(), sizes, sizeof(sizes));
    Bytes u = from(uniq), wb = from(words);
    out.insert(out.end(), u.begin(), u.end());
    out.insert(out.end(), wb.begin(), wb.end());
    return out;
}
/* _decompressUniqueKernel (:183-204): code at bits [i*b, i*b+b) of word w */
template <typename T>
int decode_codes(const uint8_t* words, size_t nwords, const T* uniq, size_t U, size_t m, std::vector<T>& out) {
    int bits = alt_bitlen((int)U - 1);
    int P = 32 / bits;
    out.assign(m, (T)0);
    for (size_t idx = 0; idx < m; idx++) {
        size_t w = idx / P;
        if (w >= nwords) return ECORRUPT;
        uint32_t word;
        memcpy(&word, words + 4 * w, 4);
        uint32_t code = (uint32_t)getnpbits(word, bits, (idx % P) * bits);
        if (code >= U) return ECORRUPT;
        out[idx] = uniq[code];
    }
    return OK;
}
/* DecompressUnique (:215-259) */
template <typename T> int decompress_unique(const Bytes& b, std::vector<T>& out) {
    out.clear();
    if (b.empty()) return OK;
    if (b.size() < 16) return ECORRUPT;
    size_t sizes[2];
    memcpy(sizes, b.data(), 16);
    size_t U = sizes[0], m = sizes[1];
    if (U == 0 || 16 + U * sizeof(T) > b.size()) return ECORRUPT;
    std::vector<T> uniq(U);
    memcpy(uniq.data(), b.data() + 16, U * sizeof(T));
    size_t off = 16 + U * sizeof(T);
    return decode_codes(b.data() + off, (b.size() - off) / 4, uniq.data(), U, m, out);
}
/* UniqueEncoding::Encode / Decode (:133-181, :261-300).  FindUnique
 * (:35-46): thrust::sort (device radix order, as for the histogram above:
 * -0.0 just before +0.0, sign-set NaNs first, other NaNs last) then
 * thrust::unique with equal_to, which keeps the first of each run of equal
 * neighbours: -0.0 and +0.0 merge into whichever sorts first (-0.0), and
 * no NaN equals anything, so every NaN row stays in the list.  A NaN row
 * matches no listed value and is coded 0. */
template <typename T> int unique_encode(const T* x, size_t n, Outs& outs) {
    outs.assign(2, Bytes());
    if (n == 0) return OK;
    std::vector<T> uniq(x, x + n);
    std::stable_sort(uniq.begin(), uniq.end(), [](T a, T b) { return radix_key(a) < radix_key(b); });
    uniq.erase(std::unique(uniq.begin(), uniq.end(), [](T a, T b) { return a == b; }), uniq.end());
    std::vector<T> xv(x, x + n);
    std::vector<uint32_t> words;
    pack_codes(xv, uniq, words);
    size_t sizes[2] = {uniq.size(), n};
    Bytes meta(16);
    memcpy(meta.data(), sizes, 16);
    Bytes u = from(uniq);
    meta.insert(meta.end(), u.begin(), u.end());
    outs[0] = meta;
    outs[1] = from(words);
    return OK;
}
template <typename T> int unique_decode(const Outs& ins, std::vector<T>& out) {
    out.clear();
    if (ins.size() < 2) return EINVAL_;
    if (ins[1].empty()) return OK;
    if (ins[0].size() < 16) return ECORRUPT;
    size_t sizes[2];
    memcpy(sizes, ins[0].data(), 16);
    size_t U = sizes[0], m = sizes[1];
    if (U == 0 || 16 + U * sizeof(T) > ins[0].size()) return ECORRUPT;
    return decode_codes(ins[1].data(), ins[1].size() / 4, (const T*)(ins[0].data() + 16), U, m, out);
}

/* -------------------------------------------------------------------- dict
 * DictEncoding (compression/dict/dict_encoding.cu:66-130), K = 4
 * (dict_encoding.hpp:30): stencil s = x in topK (getMostFrequentStencilKernel
 * :15-34); out0 = packed s; out1 = CompressUnique(frequent rows, topK);
 * out2 = the other rows raw. */
template <typename T> int dict_encode(const T* x, size_t n, Outs& outs) {
    outs.assign(3, Bytes());
    if (n == 0) return OK;
    const int K = 4;
    std::vector<T> top = most_frequent(x, n, K, nullptr);
    std::vector<int32_t> s(n);
    for (size_t i = 0; i < n; i++) {
        s[i] = 0;
        for (int k = 0; k < K; k++)
            if (x[i] == top[k]) s[i] = 1;
    }
    std::vector<T> yes, no;
    split(x, s, yes, no);
    outs[0] = stencil_pack(s);
    outs[1] = compress_unique(yes, top);
    outs[2] = from(no);
    return OK;
}
template <typename T> int dict_decode(const Outs& ins, std::vector<T>& out) {
    out.clear();
    if (ins.size() < 3) return EINVAL_;
    if (ins[1].empty() && ins[2].empty()) return OK;
    std::vector<int32_t> s;
    int rc = stencil_unpack(ins[0].data(), ins[0].size(), s);
    if (rc) return rc;
    std::vector<T> freq;
    rc = decompress_unique<T>(ins[1], freq);
    if (rc) return rc;
    return merge<T>(s, from(freq), ins[2], out);
}

/* ------------------------------------------------------------------- const
 * ConstEncoding (compression/const/const_encoding.cu:43-131): c = top-1
 * value; metadata = c ++ packed (x != c); data = x[x != c]. */
template <typename T> int const_encode(const T* x, size_t n, Outs& outs) {
    outs.assign(2, Bytes());
    if (n == 0) return OK;
    T c = most_frequent(x, n, 1, nullptr)[0];
    std::vector<int32_t> s(n);
    for (size_t i = 0; i < n; i++) s[i] = (x[i] != c) ? 1 : 0;
    std::vector<T> yes, no;
    split(x, s, yes, no);
    Bytes meta = scalar(c), st = stencil_pack(s);
    meta.insert(meta.end(), st.begin(), st.end());
    outs[0] = meta;
    outs[1] = from(yes);
    return OK;
}
template <typename T> int const_decode(const Outs& ins, std::vector<T>& out) {
    out.clear();
    if (ins.size() < 2) return EINVAL_;
    if (ins[0].empty()) return OK;
    if (ins[0].size() < sizeof(T) + 1) return ECORRUPT;
    T c;
    memcpy(&c, ins[0].data(), sizeof(T));
    std::vector<int32_t> s;
    int rc = stencil_unpack(ins[0].data() + sizeof(T), ins[0].size() - sizeof(T), s);
    if (rc) return rc;
    size_t k = 0, nd = count<T>(ins[1]);
    out.assign(s.size(), c);
    for (size_t i = 0; i < s.size(); i++)
        if (s[i]) {
            if (k >= nd) return ECORRUPT;
            memcpy(&out[i], ins[1].data() + sizeof(T) * k++, sizeof(T));
        }
    return OK;
}

/* ------------------------------------------------------------------- patch
 * PatchEncodingFactory::Get(data) (compression/patch/patch_encoding_factory.hpp:40-57)
 * computes min/max of T; LowerPatchEncoding::GetOperator
 * (lower_patch_encoding.cu:22-31): dist = T(max - min), thr = T(max - 0.1*dist)
 * in double; s = x < thr.  OutsidePatchEncoding (outside_patch_encoding.cu:22-32):
 * s = low <= x <= high with low = T(min + 0.1*dist), high = T(max - 0.1*dist).
 * out0 = packed s, out1 = x[s], out2 = x[!s]. */
template <typename T> int patch_encode(const T* x, size_t n, Outs& outs, bool outside) {
    outs.assign(3, Bytes());
    if (n == 0) return OK;
    T mnT, mxT;
    minmax(x, n, mnT, mxT);
    volatile double mn = (double)mnT, mx = (double)mxT, factor = 0.1;
    T dist = (T)(mx - mn);
    std::vector<int32_t> s(n);
    if (!outside) {
        volatile double prod = factor * (double)dist;
        T thr = (T)(mx - prod);
        for (size_t i = 0; i < n; i++) s[i] = x[i] < thr;
    } else {
        volatile double prod = factor * (double)dist;
        T low = (T)(mn + prod), high = (T)(mx - prod);
        for (size_t i = 0; i < n; i++) s[i] = (x[i] > high || x[i] < low) ? 0 : 1;
    }
    std::vector<T> yes, no;
    split(x, s, yes, no);
    outs[0] = stencil_pack(s);
    outs[1] = from(yes);
    outs[2] = from(no);
    return OK;
}
template <typename T> int patch_decode(const Outs& ins, std::vector<T>& out) {
    out.clear();
    if (ins.size() < 3) return EINVAL_;
    if (ins[1].empty() && ins[2].empty()) return OK;
    std::vector<int32_t> s;
    int rc = stencil_unpack(ins[0].data(), ins[0].size(), s);
    if (rc) return rc;
    return merge<T>(s, ins[1], ins[2], out);
}

/* -------------------------------------------------------------- floatToInt
 * FloatEncoding (compression/float/float_encoding.cu:17-88) with operators
 * FloatingPointToIntegerOperator / IntegerToFloatingPointOperator
 * (util/transform/transform_operators.hpp:109-131): p = max precision of
 * float(x); if p < 10 and min*10^p > INT_MIN and max*10^p < INT_MAX (in T
 * arithmetic) the data becomes int32(lrint(x * 10^p)) (in double), otherwise
 * p = 10 and the raw bytes are kept. */
static const double kPow10[11] = {1e0, 1e1, 1e2, 1e3, 1e4, 1e5, 1e6, 1e7, 1e8, 1e9, 1e10};
template <typename T> int float_encode(const T* x, size_t n, Outs& outs) {
    outs.assign(2, Bytes());
    if (n == 0) return OK;
    int precision = 0;
    for (size_t i = 0; i < n; i++) precision = std::max(precision, float_precision((float)x[i]));
    bool transform = false;
    if (precision < 10) {
        T mn, mx;
        minmax(x, n, mn, mx);
        int scaleFactor = (int)kPow10[precision];
        typedef typename std::conditional<std::is_floating_point<T>::value, T, long long>::type A;
        A lo = (A)mn * (A)scaleFactor, hi = (A)mx * (A)scaleFactor;
        if (std::is_same<T, int32_t>::value) { /* int*int wraps in the reference */
            lo = (A)(int32_t)(uint32_t)((uint64_t)(int64_t)mn * (uint64_t)(int64_t)scaleFactor);
            hi = (A)(int32_t)(uint32_t)((uint64_t)(int64_t)mx * (uint64_t)(int64_t)scaleFactor);
        }
        if (lo > (A)std::numeric_limits<int32_t>::min() && hi < (A)std::numeric_limits<int32_t>::max())
            transform = true;
        else
            precision = 10;
    }
    if (transform) {
        std::vector<int32_t> y(n);
        for (size_t i = 0; i < n; i++) {
            volatile double prod = (double)x[i] * kPow10[precision];
            y[i] = (int32_t)(long)lrint(prod);
        }
        outs[1] = from(y);
    } else {
        outs[1] = from(x, n);
    }
    outs[0] = scalar((int32_t)precision);
    return OK;
}
template <typename T> int float_decode(const Outs& ins, std::vector<T>& out) {
    out.clear();
    if (ins.size() < 2) return EINVAL_;
    if (ins[1].empty()) return OK;
    if (ins[0].size() < 4) return ECORRUPT;
    int32_t p = as<int32_t>(ins[0])[0];
    if (p < 10) {
        if (p < 0) return ECORRUPT;
        size_t k = count<int32_t>(ins[1]);
        out.resize(k);
        for (size_t i = 0; i < k; i++) {
            int32_t v;
            memcpy(&v, ins[1].data() + 4 * i, 4);
            volatile double q = (double)(T)v / kPow10[p];
            out[i] = (T)q;
        }
    } else {
        out = vec<T>(ins[1]);
    }
    return OK;
}

/* --------------------------------------------------------------------- gfc
 * GfcEncoding (compression/gfc/gfc_encoding.cpp:16-108) with CompressDouble /
 * CompressFloat and their kernels (gfc_encoding_impl_double.cu:31-303,
 * gfc_encoding_impl_float.cu:25-289).  Chunking: N' = roundup32(n) (zero
 * padded), blocks = ceil(n / 32768), Wc = blocks * warpsPerBlock (16 for
 * double, 32 for float), per = roundup32(max(32, ceil(N'/Wc))),
 * cut[i] = min((i+1)*per, N').  In a chunk every 32-row subchunk is coded
 * against last = 0 (first subchunk) or the chunk's element 31 (later ones);
 * diff = v - last, sign = MSB, |diff|; byte count nb (double: z = clz/8,
 * nb = z == 6 ? 3 : 8 - z, code = nb >= 3 ? nb - 1 : nb; float: nb = code =
 * 4 - clz/8).  Subchunk = 16 nibble bytes (lane 2j low, 2j+1 high;
 * nibble = sign<<3 | code) then each lane's low nb bytes.  out0 = int32
 * {blocks, warpsPerBlock, n}; out1 = int32 per chunk: scratch start
 * ((start+1)/2 * (S+1)) + chunk bytes; out2 = concatenated chunk payloads. */
template <typename U> int gfc_encode(const U* bits, size_t n, Outs& outs) {
    const bool dbl = sizeof(U) == 8;
    outs.assign(3, Bytes());
    if (n == 0) return OK;
    const long long warpsperblock = dbl ? 16 : 32;
    const long long blocks = ((long long)n + 32768 - 1) / 32768;
    const long long Wc = blocks * warpsperblock;
    const long long N = ((long long)n + 31) & ~31LL;
    std::vector<U> v((size_t)N, 0);
    memcpy(v.data(), bits, n * sizeof(U));
    long long per = (N + Wc - 1) / Wc;
    if (per < 32) per = 32;
    per = (per + 31) & ~31LL;
    const long long S = dbl ? 16 : 8; /* scratch bytes per 2 values minus the nibble byte */
    std::vector<int32_t> offs((size_t)Wc);
    Bytes payload;
    long long curr = 0, start = 0;
    for (long long w = 0; w < Wc; w++) {
        curr += per;
        long long end = std::min(curr, N);
        long long chunkBytes = 0;
        U last = 0;
        for (long long sb = start; sb < end; sb += 32) {
            uint8_t nib[32];
            Bytes data;
            for (int lane = 0; lane < 32; lane++) {
                U diff = (U)(v[(size_t)(sb + lane)] - last);
                int sign = (int)(diff >> (8 * sizeof(U) - 1)) & 1;
                if (sign) diff = (U)(0 - diff);
                int lz = 0;
                for (int b = 8 * (int)sizeof(U) - 1; b >= 0 && !((diff >> b) & 1); b--) lz++;
                int z = lz >> 3;
                int nb, code;
                if (dbl) {
                    nb = (z == 6) ? 3 : 8 - z;
                    code = nb >= 3 ? nb - 1 : nb;
                } else {
                    nb = 4 - z;
                    code = nb;
                }
                nib[lane] = (uint8_t)((sign << 3) | code);
                for (int b = 0; b < nb; b++) data.push_back((uint8_t)(diff >> (8 * b)));
            }
            for (int j = 0; j < 16; j++) payload.push_back((uint8_t)(nib[2 * j] | (nib[2 * j + 1] << 4)));
            payload.insert(payload.end(), data.begin(), data.end());
            chunkBytes += 16 + (long long)data.size();
            last = v[(size_t)(start + 31)];
        }
        offs[(size_t)w] = (int32_t)((start + 1) / 2 * (S + 1) + chunkBytes);
        start = end;
    }
    int32_t meta[3] = {(int32_t)blocks, (int32_t)warpsperblock, (int32_t)n};
    outs[0] = from(meta, 3);
    outs[1] = from(offs);
    outs[2] = payload;
    return OK;
}
template <typename U> int gfc_decode(const Outs& ins, std::vector<U>& out) {
    const bool dbl = sizeof(U) == 8;
    out.clear();
    if (ins.size() < 3) return EINVAL_;
    if (ins[2].empty()) return OK;
    if (ins[0].size() < 12) return ECORRUPT;
    const int32_t* meta = as<int32_t>(ins[0]);
    long long blocks = meta[0], wpb = meta[1], n = meta[2];
    if (blocks <= 0 || wpb <= 0 || n <= 0) return ECORRUPT;
    long long Wc = blocks * wpb;
    if ((size_t)Wc > count<int32_t>(ins[1])) return ECORRUPT;
    const long long N = (n + 31) & ~31LL;
    long long per = (N + Wc - 1) / Wc;
    if (per < 32) per = 32;
    per = (per + 31) & ~31LL;
    const long long S = dbl ? 16 : 8;
    std::vector<U> v((size_t)N, 0);
    const int32_t* offs = as<int32_t>(ins[1]);
    const Bytes& pay = ins[2];
    long long curr = 0, start = 0, pos = 0;
    for (long long w = 0; w < Wc; w++) {
        curr += per;
        long long end = std::min(curr, N);
        long long chunkBytes = offs[w] - (start + 1) / 2 * (S + 1);
        long long cpos = pos;
        U last = 0;
        for (long long sb = start; sb < end; sb += 32) {
            if (cpos + 16 > (long long)pay.size()) return ECORRUPT;
            int nib[32];
            for (int j = 0; j < 16; j++) {
                nib[2 * j] = pay[(size_t)(cpos + j)] & 15;
                nib[2 * j + 1] = pay[(size_t)(cpos + j)] >> 4;
            }
            cpos += 16;
            for (int lane = 0; lane < 32; lane++) {
                int code = nib[lane] & 7;
                int nb = dbl ? (code >= 2 ? code + 1 : code) : code;
                if (cpos + nb > (long long)pay.size()) return ECORRUPT;
                U diff = 0;
                for (int b = nb - 1; b >= 0; b--) diff = (U)((diff << 8) | pay[(size_t)(cpos + b)]);
                cpos += nb;
                if (nib[lane] & 8) diff = (U)(0 - diff);
                v[(size_t)(sb + lane)] = (U)(last + diff);
            }
            last = v[(size_t)(start + 31)];
        }
        if (cpos - pos != chunkBytes) return ECORRUPT;
        pos = cpos;
        start = end;
    }
    v.resize((size_t)n);
    out = v;
    return OK;
}

/* -------------------------------------------------------------------- none
 * NoneEncoding (compression/none/none_encoding.cu:13-49): metadata = size_t
 * byte length, data = a copy. */
int none_encode(const uint8_t* x, size_t nbytes, Outs& outs) {
    outs.assign(2, Bytes());
    outs[0] = scalar((size_t)nbytes);
    outs[1] = Bytes(x, x + nbytes);
    return OK;
}

/* ------------------------------------------------------------ dispatching */
template <typename T> int enc_typed(int enc, const T* x, size_t n, Outs& outs) {
    switch (enc) {
    case ORC_DICT: return dict_encode(x, n, outs);
    case ORC_UNIQUE: return unique_encode(x, n, outs);
    case ORC_CONST: return const_encode(x, n, outs);
    case ORC_PATCH: return patch_encode(x, n, outs, false);
    case ORC_OUTSIDE_PATCH: return patch_encode(x, n, outs, true);
    case ORC_RLE: return rle_encode(x, n, outs);
    case ORC_FLOAT2INT: return float_encode(x, n, outs);
    default: return EUNSUP;
    }
}
template <typename T> int dec_typed(int enc, const Outs& ins, std::vector<T>& out) {
    switch (enc) {
    case ORC_DICT: return dict_decode(ins, out);
    case ORC_UNIQUE: return unique_decode(ins, out);
    case ORC_CONST: return const_decode(ins, out);
    case ORC_PATCH:
    case ORC_OUTSIDE_PATCH: return patch_decode(ins, out);
    case ORC_RLE: return rle_decode(ins, out);
    case ORC_FLOAT2INT: return float_decode(ins, out);
    default: return EUNSUP;
    }
}

/* Integer view used by delta/scale/afl for each DataType
 * (delta_encoding.hpp:57-64, scale_encoding.hpp:52-59). */
template <typename F> int by_int_view(int dtype, F f) {
    switch (dtype) {
    case ORC_D_INT:
    case ORC_D_FLOAT: return f((int32_t*)nullptr);
    case ORC_D_TIME:
    case ORC_D_DOUBLE: return f((int64_t*)nullptr);
    case ORC_D_SHORT: return f((int16_t*)nullptr);
    case ORC_D_CHAR: return f((int8_t*)nullptr);
    default: return EINVAL_;
    }
}
template <typename F> int by_native(int dtype, F f) {
    switch (dtype) {
    case ORC_D_INT: return f((int32_t*)nullptr);
    case ORC_D_TIME: return f((int64_t*)nullptr);
    case ORC_D_FLOAT: return f((float*)nullptr);
    case ORC_D_DOUBLE: return f((double*)nullptr);
    case ORC_D_SHORT: return f((int16_t*)nullptr);
    case ORC_D_CHAR: return f((int8_t*)nullptr);
    default: return EINVAL_;
    }
}

int encode_any(int enc, int dtype, const uint8_t* in, size_t nbytes, Outs& outs) {
    if (orc_dtype_size(dtype) == 0) return EINVAL_;
    size_t tsz = orc_dtype_size(dtype);
    if (enc == ORC_NONE) return none_encode(in, nbytes, outs);
    /* A CudaPtr<T> built over a char buffer sees floor(bytes / sizeof(T))
     * elements (core/cuda_ptr.hpp:60): trailing bytes are not encoded. */
    nbytes -= nbytes % tsz;
    if (enc == ORC_DELTA || enc == ORC_SCALE) {
        return by_int_view(dtype, [&](auto* tag) {
            typedef typename std::remove_pointer<decltype(tag)>::type T;
            std::vector<T> x = vec<T>(Bytes(in, in + nbytes));
            return enc == ORC_DELTA ? delta_encode(x.data(), x.size(), outs) : scale_encode(x.data(), x.size(), outs);
        });
    }
    if (enc == ORC_AFL || (enc == ORC_GFC && dtype != ORC_D_FLOAT && dtype != ORC_D_DOUBLE)) {
        int rc;
        if (dtype == ORC_D_DOUBLE) return EUNSUP; /* AflEncoding::EncodeDouble returns no outputs (afl_encoding.cu:268-271) */
        if (dtype == ORC_D_FLOAT) {
            std::vector<float> x = vec<float>(Bytes(in, in + nbytes));
            rc = afl_float_encode(x.data(), x.size(), outs);
        } else {
            rc = by_int_view(dtype, [&](auto* tag) {
                typedef typename std::remove_pointer<decltype(tag)>::type T;
                std::vector<T> x = vec<T>(Bytes(in, in + nbytes));
                return afl_encode(x.data(), x.size(), outs);
            });
        }
        if (enc == ORC_GFC && rc == OK) outs.push_back(Bytes()); /* gfc_encoding.cpp:119-166 */
        return rc;
    }
    if (enc == ORC_GFC) {
        if (dtype == ORC_D_DOUBLE) {
            std::vector<uint64_t> x = vec<uint64_t>(Bytes(in, in + nbytes));
            return gfc_encode(x.data(), x.size(), outs);
        }
        std::vector<uint32_t> x = vec<uint32_t>(Bytes(in, in + nbytes));
        return gfc_encode(x.data(), x.size(), outs);
    }
    return by_native(dtype, [&](auto* tag) {
        typedef typename std::remove_pointer<decltype(tag)>::type T;
        std::vector<T> x = vec<T>(Bytes(in, in + nbytes));
        return enc_typed(enc, x.data(), x.size(), outs);
    });
}

int decode_any(int enc, int dtype, const Outs& ins, Bytes& out) {
    if (orc_dtype_size(dtype) == 0) return EINVAL_;
    if (enc == ORC_NONE) {
        if (ins.size() < 2) return EINVAL_;
        out = ins[1];
        return OK;
    }
    if (enc == ORC_DELTA || enc == ORC_SCALE) {
        return by_int_view(dtype, [&](auto* tag) {
            typedef typename std::remove_pointer<decltype(tag)>::type T;
            std::vector<T> v;
            int rc = enc == ORC_DELTA ? delta_decode(ins, v) : scale_decode(ins, v);
            out = from(v);
            return rc;
        });
    }
    if (enc == ORC_AFL || (enc == ORC_GFC && dtype != ORC_D_FLOAT && dtype != ORC_D_DOUBLE)) {
        if (dtype == ORC_D_DOUBLE) return EUNSUP;
        if (dtype == ORC_D_FLOAT) {
            std::vector<float> v;
            int rc = afl_float_decode(ins, v);
            out = from(v);
            return rc;
        }
        return by_int_view(dtype, [&](auto* tag) {
            typedef typename std::remove_pointer<decltype(tag)>::type T;
            std::vector<T> v;
            int rc = afl_decode(ins, v);
            out = from(v);
            return rc;
        });
    }
    if (enc == ORC_GFC) {
        if (dtype == ORC_D_DOUBLE) {
            std::vector<uint64_t> v;
            int rc = gfc_decode(ins, v);
            out = from(v);
            return rc;
        }
        std::vector<uint32_t> v;
        int rc = gfc_decode(ins, v);
        out = from(v);
        return rc;
    }
    return by_native(dtype, [&](auto* tag) {
        typedef typename std::remove_pointer<decltype(tag)>::type T;
        std::vector<T> v;
        int rc = dec_typed(enc, ins, v);
        out = from(v);
        return rc;
    });
}

void to_orc(const Bytes& b, orc_buf* o) {
    o->n = b.size();
    o->p = (uint8_t*)malloc(b.size() ? b.size() : 1);
    if (!b.empty()) memcpy(o->p, b.data(), b.size());
}

} // namespace

extern "C" {

size_t orc_dtype_size(int dtype) {
    switch (dtype) {
    case ORC_D_INT: return 4;
    case ORC_D_TIME: return 8;
    case ORC_D_FLOAT: return 4;
    case ORC_D_DOUBLE: return 8;
    case ORC_D_SHORT: return 2;
    case ORC_D_CHAR: return 1;
    default: return 0;
    }
}

int orc_bitlen(int dtype, const void* minp, const void* maxp) {
    return by_native(dtype, [&](auto* tag) {
        typedef typename std::remove_pointer<decltype(tag)>::type T;
        return bitlen(*(const T*)minp, *(const T*)maxp);
    });
}

int orc_alt_bitlen(int v) { return alt_bitlen(v); }
int orc_float_precision(float x) { return float_precision(x); }

int orc_encoding_info(int enc, int dtype, int* n_results, int* rt) {
    if (orc_dtype_size(dtype) == 0) return EINVAL_;
    rt[0] = dtype;
    rt[1] = dtype;
    switch (enc) {
    case ORC_NONE: *n_results = 0; break;
    case ORC_DELTA:
    case ORC_SCALE:
    case ORC_AFL:
    case ORC_UNIQUE:
    case ORC_CONST: *n_results = 1; break;
    case ORC_FLOAT2INT: *n_results = 1; rt[0] = ORC_D_INT; break;   /* float_encoding.hpp:31 */
    case ORC_RLE: *n_results = 2; rt[0] = ORC_D_INT; break;         /* rle_encoding.hpp:35 */
    case ORC_DICT:
    case ORC_PATCH:
    case ORC_OUTSIDE_PATCH:
    case ORC_GFC: *n_results = 2; break;
    default: return EINVAL_;
    }
    return OK;
}

int orc_encode(int enc, int dtype, const void* in, size_t in_bytes, orc_buf* outs, int* n_outs) {
    Outs o;
    int rc = encode_any(enc, dtype, (const uint8_t*)in, in_bytes, o);
    if (rc) return rc;
    for (size_t i = 0; i < o.size(); i++) to_orc(o[i], &outs[i]);
    *n_outs = (int)o.size();
    return OK;
}

int orc_decode(int enc, int dtype, const orc_buf* ins, int n_ins, orc_buf* out) {
    Outs o((size_t)n_ins);
    for (int i = 0; i < n_ins; i++) o[i] = Bytes(ins[i].p, ins[i].p + ins[i].n);
    Bytes r;
    int rc = decode_any(enc, dtype, o, r);
    if (rc) return rc;
    to_orc(r, out);
    return OK;
}

/* GetMetadataSize / GetCompressedSize of each encoding (its .hpp/.cu). */
int orc_predict_sizes(int enc, int dtype, const void* in, size_t nbytes, size_t* meta, size_t* data) {
    size_t tsz = orc_dtype_size(dtype);
    if (!tsz) return EINVAL_;
    size_t n = nbytes / tsz;
    const uint8_t* x = (const uint8_t*)in;
    switch (enc) {
    case ORC_NONE: *meta = sizeof(size_t); *data = nbytes; return OK;
    case ORC_DELTA: *meta = n ? tsz : 0; *data = n ? nbytes - tsz : 0; return OK;
    case ORC_SCALE: *meta = n ? tsz : 0; *data = n ? nbytes : 0; return OK;
    case ORC_FLOAT2INT: *meta = n ? 4 : 0; *data = nbytes; return OK;
    case ORC_GFC: *meta = 12; *data = 0; return OK;
    case ORC_PATCH:
    case ORC_OUTSIDE_PATCH: *meta = n ? (n + 7) / 8 + 1 : 0; *data = nbytes; return OK;
    case ORC_RLE: {
        *meta = n ? 4 : 0;
        if (!n) { *data = 0; return OK; }
        if (dtype != ORC_D_INT && dtype != ORC_D_FLOAT) return EUNSUP;
        Outs o;
        encode_any(ORC_RLE, dtype, x, nbytes, o);
        *data = o[1].size() + o[2].size();
        return OK;
    }
    case ORC_CONST: {
        *meta = n ? (n + 7) / 8 + tsz + 1 : 0;
        if (!n) { *data = 0; return OK; }
        if (dtype != ORC_D_INT && dtype != ORC_D_FLOAT) return EUNSUP;
        Outs o;
        encode_any(ORC_CONST, dtype, x, nbytes, o);
        *data = o[1].size();
        return OK;
    }
    case ORC_DICT: {
        *meta = (n + 7) / 8 + 1;
        if (!n) { *data = 0; return OK; }
        if (dtype != ORC_D_INT && dtype != ORC_D_FLOAT) return EUNSUP;
        Outs o;
        encode_any(ORC_DICT, dtype, x, nbytes, o);
        /* dict_encoding.cu:156-172: sized with K = 4 keys whether or not 4 are distinct */
        size_t m = 0;
        if (o[1].size() >= 16) memcpy(&m, o[1].data() + 8, 8);
        *data = 16 + 4 * tsz + (m + 15) / 16 * 4 + o[2].size();
        return OK;
    }
    case ORC_UNIQUE: {
        if (!n) { *meta = 0; *data = 0; return OK; }
        if (dtype != ORC_D_INT && dtype != ORC_D_FLOAT) return EUNSUP;
        Outs o;
        encode_any(ORC_UNIQUE, dtype, x, nbytes, o);
        *meta = o[0].size();
        *data = o[1].size();
        return OK;
    }
    case ORC_AFL: {
        if (!n) { *meta = 0; *data = 0; return OK; }
        if (dtype == ORC_D_INT) {
            Outs o;
            encode_any(ORC_AFL, dtype, x, nbytes, o);
            *meta = 4;
            *data = o[1].size();
            return OK;
        }
        if (dtype == ORC_D_FLOAT) {
            Outs o;
            encode_any(ORC_AFL, dtype, x, nbytes, o);
            *meta = sizeof(size_t) + 1;
            *data = o[1].size();
            return OK;
        }
        return EUNSUP;
    }
    default: return EINVAL_;
    }
}

int orc_stats_compute(int dtype, const void* in, size_t nbytes, orc_stats* out) {
    return by_native(dtype, [&](auto* tag) {
        typedef typename std::remove_pointer<decltype(tag)>::type T;
        std::vector<T> x = vec<T>(Bytes((const uint8_t*)in, (const uint8_t*)in + nbytes));
        *out = stats_of(x.data(), x.size());
        return OK;
    });
}

int orc_most_frequent(int dtype, const void* in, size_t nbytes, int k, void* keys_out, int* n_distinct) {
    return by_native(dtype, [&](auto* tag) {
        typedef typename std::remove_pointer<decltype(tag)>::type T;
        std::vector<T> x = vec<T>(Bytes((const uint8_t*)in, (const uint8_t*)in + nbytes));
        std::vector<T> top = most_frequent(x.data(), x.size(), k, n_distinct);
        memcpy(keys_out, top.data(), top.size() * sizeof(T));
        return OK;
    });
}

int orc_stencil_pack(const int32_t* s, size_t n, orc_buf* out) {
    if (n == 0) { to_orc(Bytes{0}, out); return OK; }
    to_orc(stencil_pack(std::vector<int32_t>(s, s + n)), out);
    return OK;
}

int orc_stencil_unpack(const uint8_t* p, size_t nbytes, orc_buf* out) {
    std::vector<int32_t> s;
    int rc = stencil_unpack(p, nbytes, s);
    if (rc) return rc;
    to_orc(from(s), out);
    return OK;
}

int orc_fl_pack(int dtype, int bits, const void* in, size_t n, void* out, size_t out_words) {
    return by_int_view(dtype, [&](auto* tag) {
        typedef typename std::remove_pointer<decltype(tag)>::type T;
        if (bits <= 0 || bits > (int)(8 * sizeof(T))) return EINVAL_;
        memset(out, 0, out_words * sizeof(T));
        fl_compress<T>(bits, (const T*)in, (T*)out, n, out_words);
        return OK;
    });
}

int orc_fl_unpack(int dtype, int bits, const void* in, size_t n, void* out) {
    return by_int_view(dtype, [&](auto* tag) {
        typedef typename std::remove_pointer<decltype(tag)>::type T;
        if (bits <= 0 || bits > (int)(8 * sizeof(T))) return EINVAL_;
        const uint64_t W = 8 * sizeof(T);
        size_t words = (size_t)((bits * (uint64_t)n + W - 1) / W);
        fl_decompress<T>(bits, (const T*)in, (T*)out, n, words);
        return OK;
    });
}

/* run_afl_compress_cpu<T,32> / run_afl_decompress_cpu<T,32> over whole
 * 32-lane blocks; out_words = ceil(n / (32*W)) * 32 * b. */
int orc_fl_pack_lanes(int dtype, int bits, const void* in, size_t n, void* out, size_t out_words) {
    return by_int_view(dtype, [&](auto* tag) {
        typedef typename std::remove_pointer<decltype(tag)>::type T;
        if (bits <= 0 || bits > (int)(8 * sizeof(T))) return EINVAL_;
        memset(out, 0, out_words * sizeof(T));
        fl_compress_lanes<T>(bits, (const T*)in, (T*)out, n, out_words);
        return OK;
    });
}

int orc_fl_unpack_lanes(int dtype, int bits, const void* in, size_t in_words, size_t n, void* out) {
    return by_int_view(dtype, [&](auto* tag) {
        typedef typename std::remove_pointer<decltype(tag)>::type T;
        if (bits <= 0 || bits > (int)(8 * sizeof(T))) return EINVAL_;
        fl_decompress_lanes<T>(bits, (const T*)in, (T*)out, n, in_words);
        return OK;
    });
}

/* afl_decompress_base_value_gpu for every row in rows[0..count): lanes 1 or 32. */
int orc_fl_fetch(int dtype, int bits, int lanes, const void* in, size_t n, const uint64_t* rows, size_t count,
                 void* out) {
    return by_int_view(dtype, [&](auto* tag) {
        typedef typename std::remove_pointer<decltype(tag)>::type T;
        if (bits <= 0 || bits > (int)(8 * sizeof(T)) || (lanes != 1 && lanes != 32)) return EINVAL_;
        for (size_t i = 0; i < count; i++) {
            if (rows[i] >= n) return EINVAL_;
            ((T*)out)[i] = lanes == 1 ? afl_value<T, 1>(bits, (const T*)in, rows[i])
                                      : afl_value<T, 32>(bits, (const T*)in, rows[i]);
        }
        return OK;
    });
}

void orc_free(orc_buf* b) {
    if (b && b->p) free(b->p);
    if (b) { b->p = nullptr; b->n = 0; }
}

} /* extern "C" */
