// GFC float/double codec (reference compression/gfc/gfc_encoding.cpp:16-108,
// gfc_encoding_impl_double.cu:31-303, gfc_encoding_impl_float.cu:25-289).
//
// Stream semantics kept from the reference: rows are zero-padded to a
// multiple of 32 and cut into Wc = blocks * warpsPerBlock chunks of `per`
// rows; inside a chunk every 32-row subchunk codes diff = v - last, with
// last = 0 for the chunk's first subchunk and the chunk's row 31 for the
// others; a subchunk is 16 bytes of sign|code nibbles then each row's low
// nb bytes of |diff|.  out[1] holds, per chunk, the end offset in the
// reference's scratch layout; out[2] is the concatenated chunk payloads.
//
// MI355X mapping: one workgroup per chunk (<= 2048 rows).  A size pass
// computes every chunk's byte count (and the reference's offsets table),
// k_scan_tiles places the chunks, and the encode pass recomputes the byte
// counts, scans them in the workgroup, assembles the chunk payload in LDS and
// writes it at its place.  Decode takes the chunk sizes from the offsets
// table (scanned the same way) and stages each chunk's payload in LDS; the
// walk over the subchunk headers (the format's only sequential step) runs
// for double chunks as a kernel of its own, one lane per chunk, and inside
// the one-wave decoder for float chunks; all rows then decode in parallel
// and the chunk goes out through LDS as consecutive 16-byte vectors.
#include "dispatch.h"
#include "k_common.h"
#include "kernels.h"

namespace ddj {

namespace {

constexpr int kGfcMaxPer = 2048;  // rows per chunk is at most 2048 (double) / 1024 (float)

struct GfcGeom {
    long long n, N, blocks, wpb, Wc, per, S;
};

GfcGeom gfc_geom(bool dbl, long long n) {
    GfcGeom g;
    g.n = n;
    g.wpb = dbl ? 16 : 32;
    g.blocks = (n + 32768 - 1) / 32768;
    g.Wc = g.blocks * g.wpb;
    g.N = (n + 31) & ~31LL;
    long long per = (g.N + g.Wc - 1) / g.Wc;
    if (per < 32) per = 32;
    g.per = (per + 31) & ~31LL;
    g.S = dbl ? 16 : 8;
    return g;
}

constexpr int kGfcRpt = kGfcMaxPer / kBlock;  // rows per thread (8)

// Residual bytes announced by one 32-bit word of a subchunk header (the
// sign|code nibbles of 8 rows): the sum of the codes, plus one for every
// double code >= 2 (double codes 2..7 stand for 3..8 bytes).
template <bool DBL> __device__ __forceinline__ uint32_t gfc_word_bytes(uint32_t w) {
    const uint32_t m = w & 0x77777777u;
    uint32_t s = (m & 0x0f0f0f0fu) + ((m >> 4) & 0x0f0f0f0fu);
    s = (s * 0x01010101u) >> 24;
    if (DBL) s += (uint32_t)__builtin_popcount(((m >> 1) | (m >> 2)) & 0x11111111u);
    return s;
}

// Thread t's rows [8t, 8t+8) of the chunk starting at row c0: |diff| against
// the predictor, byte counts and the sign|code nibbles.
template <typename U> struct GfcRows {
    U d[kGfcRpt];
    int nbs[kGfcRpt];
    uint32_t nibs = 0;   // 4 bits per row
    uint32_t nbsum = 0;  // residual bytes of the thread's rows
};

// Whether thread t's 8 rows of the chunk at c0 are all inside the column
// and 16-byte aligned at p, so that they move as whole 16-byte vectors.
template <typename U>
__device__ __forceinline__ bool gfc_rows_whole(const U* p, long long n, long long c0, int rows) {
    const int r0 = threadIdx.x * kGfcRpt;
    return r0 + kGfcRpt <= rows && c0 + r0 + kGfcRpt <= n && (reinterpret_cast<uintptr_t>(p + c0 + r0) & 15) == 0;
}

template <typename U>
__device__ __forceinline__ void gfc_rows(const U* __restrict__ x, long long n, long long c0, int rows, U last_c,
                                         GfcRows<U>& g) {
    constexpr bool dbl = sizeof(U) == 8;
    constexpr int V = 16 / (int)sizeof(U);
    const int r0 = threadIdx.x * kGfcRpt;
    U v8[kGfcRpt];
    if (gfc_rows_whole(x, n, c0, rows)) {
        const u32x4* s4 = reinterpret_cast<const u32x4*>(x + c0 + r0);
#pragma unroll
        for (int q = 0; q < kGfcRpt / V; q++) {
            const u32x4 t = s4[q];
            __builtin_memcpy(&v8[q * V], &t, 16);
        }
    } else {
#pragma unroll
        for (int k = 0; k < kGfcRpt; k++) v8[k] = (r0 + k < rows && c0 + r0 + k < n) ? x[c0 + r0 + k] : U(0);
    }
#pragma unroll
    for (int k = 0; k < kGfcRpt; k++) {
        const int r = r0 + k;
        g.nbs[k] = 0;
        g.d[k] = 0;
        if (r < rows) {
            const U v = v8[k];
            const U last = (r < 32) ? U(0) : last_c;
            U diff = (U)(v - last);
            const uint32_t sign = (uint32_t)(diff >> (8 * sizeof(U) - 1)) & 1u;
            if (sign) diff = (U)(U(0) - diff);
            int nb, code;
            if (dbl) {
                const int z = (diff == 0) ? 8 : (__clzll((unsigned long long)diff) >> 3);
                nb = (z == 6) ? 3 : 8 - z;
                code = nb >= 3 ? nb - 1 : nb;
            } else {
                const int z = (diff == 0) ? 4 : (__clz((unsigned)diff) >> 3);
                nb = 4 - z;
                code = nb;
            }
            g.nbs[k] = nb;
            g.d[k] = diff;
            g.nibs |= (uint32_t)((sign << 3) | code) << (4 * k);
            g.nbsum += nb;
        }
    }
}

// The chunk's rows and its predictor (row 31 of the chunk).
struct GfcChunk {
    long long c0;
    int rows;
};
__device__ __forceinline__ GfcChunk gfc_chunk(uint32_t chunk, long long N, long long per) {
    const long long c0 = (long long)chunk * per;
    long long c1 = c0 + per;
    c1 = c1 < N ? c1 : N;
    return GfcChunk{c0, c0 < N ? (int)(c1 - c0) : 0};
}

// The reference's offsets-table entry: the chunk's end in its scratch layout.
__device__ __forceinline__ long long gfc_scratch_start(long long c0, long long N, int S) {
    return (c0 < N ? (c0 + 1) / 2 : (N + 1) / 2) * (S + 1);
}

// Size pass: payload bytes of every chunk, and the offsets table.
template <typename U>
__global__ __launch_bounds__(kBlock) void k_gfc_sizes(const U* __restrict__ x, long long n, long long N, long long per,
                                                      int S, uint64_t* __restrict__ vals,
                                                      int32_t* __restrict__ offsets) {
    __shared__ U s_last;
    __shared__ uint64_t red[kBlock / kWave];
    const uint32_t chunk = blockIdx.x;
    const GfcChunk ch = gfc_chunk(chunk, N, per);
    if (threadIdx.x == 0) s_last = (ch.rows > 0 && ch.c0 + 31 < n) ? x[ch.c0 + 31] : U(0);
    __syncthreads();
    GfcRows<U> g;
    gfc_rows(x, n, ch.c0, ch.rows, s_last, g);
    uint64_t total_nb;
    block_exclusive_scan<uint64_t>(g.nbsum, &total_nb, red);
    const uint64_t chunk_bytes = 16ull * (ch.rows / 32) + total_nb;
    if (threadIdx.x == 0) {
        vals[chunk] = chunk_bytes;
        offsets[chunk] = (int32_t)(gfc_scratch_start(ch.c0, N, S) + (long long)chunk_bytes);
    }
}

template <typename U>
__global__ __launch_bounds__(kBlock) void k_gfc_encode(const U* __restrict__ x, long long n, long long N, long long per,
                                                       uint8_t* __restrict__ payload,
                                                       const uint64_t* __restrict__ prefix) {
    constexpr int RPT = kGfcRpt;
    __shared__ __attribute__((aligned(16))) uint8_t out[kGfcMaxPer / 32 * (16 + 32 * (int)sizeof(U)) + 16];
    __shared__ U s_last;
    __shared__ uint64_t red[kBlock / kWave];
    const uint32_t chunk = blockIdx.x;
    const GfcChunk ch = gfc_chunk(chunk, N, per);
    const int rows = ch.rows;
    if (threadIdx.x == 0) s_last = (rows > 0 && ch.c0 + 31 < n) ? x[ch.c0 + 31] : U(0);
    __syncthreads();
    GfcRows<U> g;
    gfc_rows(x, n, ch.c0, rows, s_last, g);
    const uint32_t nibs = g.nibs, nbsum = g.nbsum;
    const int* nbs = g.nbs;
    const U* d = g.d;
    uint64_t total_nb;
    const uint64_t excl = block_exclusive_scan<uint64_t>(nbsum, &total_nb, red);
    const int nsub = rows / 32;
    const uint64_t chunk_bytes = 16ull * nsub + total_nb;
    // thread t covers rows [8t, 8t+8) of subchunk s = t/4, i.e. header bytes
    // 4*(t%4) .. +3; the subchunk starts after 16 header bytes per earlier
    // subchunk plus the byte counts of all earlier rows (the exclusive scan
    // value of the subchunk's first thread).
    const int sub = threadIdx.x / 4;
    __shared__ uint32_t sub_excl[kGfcMaxPer / 32];
    if ((threadIdx.x & 3) == 0 && threadIdx.x * RPT < rows) sub_excl[sub] = (uint32_t)excl;
    // the chunk is assembled by OR-ing each thread's bytes into zeroed words
    // (neighbouring threads share the words at the ends of their bytes)
    uint32_t* ow = reinterpret_cast<uint32_t*>(out);
    constexpr uint32_t kOutWords = sizeof(out) / 4;
    const uint32_t nwords = ((uint32_t)chunk_bytes + 7) / 4;  // one spare word for the funnel shift below
    for (uint32_t k = threadIdx.x; k < nwords && k < kOutWords; k += kBlock) ow[k] = 0;
    __syncthreads();
    auto or_word = [&](uint32_t at, uint32_t v) {
        if (v) __hip_atomic_fetch_or(ow + at, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    if (threadIdx.x * RPT < rows) {
        const uint32_t sub_off = 16u * sub + sub_excl[sub];
        const uint32_t h = sub_off + 4u * (threadIdx.x & 3);
        const uint32_t hs = 8u * (h & 3);
        or_word(h >> 2, nibs << hs);
        if (hs) or_word((h >> 2) + 1, nibs >> (32 - hs));
        // the rows' low bytes, LSB first, through a 64-bit accumulator that
        // starts at the word holding the first byte
        const uint32_t pos = 16u * (sub + 1) + (uint32_t)excl;
        uint32_t wi = pos >> 2;
        uint64_t acc = 0;
        uint32_t nbits = 8u * (pos & 3);
        auto push = [&](uint32_t v, uint32_t len) {  // len <= 32 bits, v < 2^len
            acc |= (uint64_t)v << nbits;
            nbits += len;
            if (nbits >= 32) {
                or_word(wi++, (uint32_t)acc);
                acc >>= 32;
                nbits -= 32;
            }
        };
#pragma unroll
        for (int k = 0; k < RPT; k++) {
            const uint64_t dv = (uint64_t)d[k];
            const uint32_t nb = (uint32_t)nbs[k];
            const uint32_t lo = nb < 4 ? nb : 4u;
            push((uint32_t)dv & (lo == 4 ? ~0u : (1u << (8 * lo)) - 1u), 8 * lo);
            if (sizeof(U) == 8 && nb > 4) {
                const uint32_t hb = nb - 4;
                push((uint32_t)(dv >> 32) & (hb == 4 ? ~0u : (1u << (8 * hb)) - 1u), 8 * hb);
            }
        }
        if (nbits) or_word(wi, (uint32_t)acc);
    }
    __syncthreads();
    // out at payload + O (any byte address): single bytes up to a 16-byte
    // boundary, then 16-byte stores of funnel-shifted words, single bytes at
    // the end; the chunks on either side own the other bytes of the edge
    // vectors
    const uint64_t O = prefix[chunk];
    uint8_t* dst = payload + O;
    const int L = (int)chunk_bytes;
    int head = (int)((16 - (reinterpret_cast<uintptr_t>(dst) & 15)) & 15);
    head = head < L ? head : L;
    if ((int)threadIdx.x < head) dst[threadIdx.x] = out[threadIdx.x];
    const int r = head & 3;
    auto src_word = [&](int s) -> uint32_t {  // chunk bytes [s, s + 4), s + 4 <= L
        const int q = s >> 2;
        return r ? __builtin_amdgcn_alignbyte(ow[q + 1], ow[q], r) : ow[q];
    };
    const int nq = (L - head) / 16;
    u32x4* d4 = reinterpret_cast<u32x4*>(dst + head);
    for (int i = threadIdx.x; i < nq; i += kBlock) {
        const int s = head + 16 * i;
        __builtin_nontemporal_store(u32x4{src_word(s), src_word(s + 4), src_word(s + 8), src_word(s + 12)}, d4 + i);
    }
    for (int j = head + 16 * nq + threadIdx.x; j < L; j += kBlock) dst[j] = out[j];
}

// Payload bytes of a chunk from the offsets table; out-of-range entries
// (corrupt input) count as empty chunks and are caught by the header walk.
template <typename U>
__device__ __forceinline__ uint64_t gfc_chunk_bytes(const uint8_t* offsets, uint32_t chunk, long long N, long long per,
                                                    int S) {
    constexpr int kMaxBytes = kGfcMaxPer / 32 * (16 + 32 * (int)sizeof(U));
    int32_t end;
    __builtin_memcpy(&end, offsets + 4 * (size_t)chunk, 4);
    const long long cb = (long long)end - gfc_scratch_start((long long)chunk * per, N, S);
    return (cb < 0 || cb > kMaxBytes) ? 0 : (uint64_t)cb;
}

template <typename U>
__global__ __launch_bounds__(kBlock) void k_gfc_chunk_bytes(const uint8_t* __restrict__ offsets, size_t chunks,
                                                            long long N, long long per, int S,
                                                            uint64_t* __restrict__ vals) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < chunks) vals[i] = gfc_chunk_bytes<U>(offsets, (uint32_t)i, N, per, S);
}

// The header walk of every chunk, one lane per chunk, reading the headers
// straight from the payload (each step's four header words are one round
// trip; the lanes of a wave walk 64 chunks side by side):
// sub_off[chunk * stride + s] = subchunk s's offset in the chunk, s <= nsub.
// A walk that does not end at the chunk's size raises the error flag.
template <typename U>
__global__ __launch_bounds__(kBlock) void k_gfc_walk(const uint8_t* __restrict__ payload, size_t payload_bytes,
                                                     const uint8_t* __restrict__ offsets, size_t chunks, long long N,
                                                     long long per, int S, const uint64_t* __restrict__ prefix,
                                                     uint32_t* __restrict__ sub_off, int stride, uint32_t* error) {
    constexpr bool dbl = sizeof(U) == 8;
    const size_t chunk = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (chunk >= chunks) return;
    const int nsub = gfc_chunk((uint32_t)chunk, N, per).rows / 32;
    const uint64_t cb = gfc_chunk_bytes<U>(offsets, (uint32_t)chunk, N, per, S);
    const uint64_t P = prefix[chunk];
    const uint64_t lim = P >= payload_bytes ? 0 : (payload_bytes - P < cb ? payload_bytes - P : cb);  // readable
    const uint8_t* base = payload + (P < payload_bytes ? P : 0);
    uint32_t* so = sub_off + chunk * (size_t)stride;
    uint32_t off = 0;
    for (int s = 0; s < nsub; s++) {
        so[s] = off;
        uint32_t tot = 0;
        if ((uint64_t)off + 16 <= lim) {
            // the aligned words holding header bytes [off, off + 16): each
            // holds at least one of them, so none leaves the payload's pages
            const uintptr_t a = reinterpret_cast<uintptr_t>(base + off);
            const uint32_t* w = reinterpret_cast<const uint32_t*>(a & ~(uintptr_t)3);
            const uint32_t r = (uint32_t)(a & 3);
            uint32_t h[5];
#pragma unroll
            for (int k = 0; k < 4; k++) h[k] = w[k];
            h[4] = r ? w[4] : 0u;
#pragma unroll
            for (int k = 0; k < 4; k++) tot += gfc_word_bytes<dbl>(r ? __builtin_amdgcn_alignbyte(h[k + 1], h[k], r) : h[k]);
        }
        off += 16 + tot;
    }
    so[nsub] = off;
    if (off != cb || P + cb > payload_bytes) __hip_atomic_store(error, 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Decode of chunks of up to 2048 rows, one 256-thread workgroup each, the
// subchunk offsets taken from k_gfc_walk (walked, stride per chunk).
template <typename U>
__global__ __launch_bounds__(kBlock) void k_gfc_decode(const uint8_t* __restrict__ payload, size_t payload_bytes,
                                                       const uint8_t* __restrict__ offsets, long long n, long long N,
                                                       long long per, int S, U* __restrict__ out,
                                                       const uint64_t* __restrict__ prefix, uint32_t* error,
                                                       const uint32_t* __restrict__ walked, int stride) {
    constexpr int RPT = kGfcMaxPer / kBlock;
    constexpr bool dbl = sizeof(U) == 8;
    constexpr int kMaxBytes = kGfcMaxPer / 32 * (16 + 32 * (int)sizeof(U));
    // a chunk at any 16-byte phase, in whole 16-byte vectors, plus one
    // spare word for the funnel shift of the last row
    constexpr int kVec = (kMaxBytes + 15 + 15) / 16;
    constexpr int kVR = (kVec + kBlock - 1) / kBlock;
    __shared__ __attribute__((aligned(16))) uint8_t buf[16 * kVec + 16];
    __shared__ uint32_t sub_off[kGfcMaxPer / 32 + 1];
    __shared__ U s_last;
    static_assert(sizeof(buf) >= kGfcMaxPer * sizeof(U), "the decoded chunk fits the staging buffer");
    __shared__ int s_bad;
    const uint32_t chunk = blockIdx.x;
    const long long c0 = (long long)chunk * per;
    const int rows = gfc_chunk(chunk, N, per).rows;
    const int nsub = rows / 32;
    const uint64_t chunk_bytes = gfc_chunk_bytes<U>(offsets, chunk, N, per, S);
    if (threadIdx.x == 0) s_bad = 0;
    __syncthreads();
    const uint64_t P = prefix[chunk];
    // the aligned 16-byte vectors that hold the chunk's bytes, all loads in
    // flight together; chunk byte i lands at buf[base + i].  A vector that
    // starts before the payload's end never crosses a page; its bytes past
    // the payload are never decoded (a chunk that claims them is an error).
    const uintptr_t a0 = reinterpret_cast<uintptr_t>(payload + P);
    const uint32_t base = (uint32_t)(a0 & 15);
    {
        const u32x4* gv = reinterpret_cast<const u32x4*>(a0 - base);
        const uintptr_t pend = reinterpret_cast<uintptr_t>(payload) + payload_bytes;
        const uint32_t nvec = (base + (uint32_t)chunk_bytes + 15) / 16;
        u32x4 t[kVR];
#pragma unroll
        for (int r = 0; r < kVR; r++) {
            const uint32_t k = threadIdx.x + r * kBlock;
            t[r] = (k < nvec && reinterpret_cast<uintptr_t>(gv + k) < pend) ? gv[k] : u32x4{0u, 0u, 0u, 0u};
        }
        u32x4* b4 = reinterpret_cast<u32x4*>(buf);
#pragma unroll
        for (int r = 0; r < kVR; r++) {
            const uint32_t k = threadIdx.x + r * kBlock;
            if (k < nvec) b4[k] = t[r];
        }
        if (threadIdx.x == 0 && P + chunk_bytes > payload_bytes) s_bad = 1;
    }
    __syncthreads();
    const uint32_t* bw = reinterpret_cast<const uint32_t*>(buf);
    constexpr uint32_t kBufWords = sizeof(buf) / 4;
    // chunk bytes [a, a + 4); words past the buffer (corrupt input only)
    // read as zero
    auto lds_word = [&](uint32_t a) -> uint32_t {
        a += base;
        const uint32_t q = a >> 2, s = a & 3;
        const uint32_t lo = q < kBufWords ? bw[q] : 0u;
        if (!s) return lo;
        const uint32_t hi = q + 1 < kBufWords ? bw[q + 1] : 0u;
        return __builtin_amdgcn_alignbyte(hi, lo, s);
    };
    // subchunk offsets from the header walk (k_gfc_walk)
    {
        const uint32_t* so = walked + (size_t)chunk * stride;
        for (int s = threadIdx.x; s <= nsub; s += kBlock) sub_off[s] = so[s];
    }
    __syncthreads();
    const int sub = threadIdx.x / 4;
    auto nb_of = [](uint32_t nib) -> uint32_t {
        const uint32_t code = nib & 7;
        return dbl ? (code >= 2 ? code + 1 : code) : code;
    };
    // this thread's rows are nibbles 8*(t%4) .. +7 of its subchunk's header:
    // one header word; the rows of the threads before it in the subchunk
    // come from their words by lane shuffles
    const bool mine = threadIdx.x * RPT < rows;
    const uint32_t so = mine ? sub_off[sub] : 0u;
    const uint32_t hw = mine ? lds_word(so + 4u * (threadIdx.x & 3)) : 0u;
    uint32_t own = 0;
#pragma unroll
    for (int k = 0; k < RPT; k++) own += nb_of(hw >> (4 * k));
    const int j = threadIdx.x & 3;
    const uint32_t up1 = __shfl_up(own, 1, 4), up2 = __shfl_up(own, 2, 4), up3 = __shfl_up(own, 3, 4);
    const uint32_t pre = (j >= 1 ? up1 : 0u) + (j >= 2 ? up2 : 0u) + (j >= 3 ? up3 : 0u);
    U dec[RPT];
    if (mine) {
        uint32_t pos = so + 16 + pre;
#pragma unroll
        for (int k = 0; k < RPT; k++) {
            const uint32_t nib = (hw >> (4 * k)) & 15;
            const uint32_t nb = nb_of(nib);
            uint64_t dv = lds_word(pos);
            if (dbl && nb > 4) dv |= (uint64_t)lds_word(pos + 4) << 32;
            dv &= nb >= 8 ? ~0ull : (1ull << (8 * nb)) - 1ull;
            pos += nb;
            U diff = (U)dv;
            if (nib & 8) diff = (U)(U(0) - diff);
            dec[k] = diff;
        }
    }
    // row 31 of the chunk is the predictor of every later subchunk
    if (threadIdx.x == 3 && rows > 0) s_last = dec[RPT - 1];
    __syncthreads();  // also: every read of the staged payload is done
    if (c0 + rows <= n && (reinterpret_cast<uintptr_t>(out + c0) & 15) == 0) {  // uniform
        // the decoded rows go through LDS so that consecutive lanes store
        // consecutive 16-byte vectors of the chunk (a lane's own 8 rows,
        // stored straight, put 64 scattered pieces in each store)
        U* ob = reinterpret_cast<U*>(buf);
        if (mine) {
            const U last = sub == 0 ? U(0) : s_last;
#pragma unroll
            for (int k = 0; k < RPT; k++) ob[threadIdx.x * RPT + k] = (U)(last + dec[k]);
        }
        __syncthreads();
        const u32x4* b4 = reinterpret_cast<const u32x4*>(buf);
        u32x4* d4 = reinterpret_cast<u32x4*>(out + c0);
        const int nvec = rows * (int)sizeof(U) / 16;  // rows is a multiple of 32
        for (int i = threadIdx.x; i < nvec; i += kBlock) d4[i] = b4[i];
    } else if (mine) {
        const U last = sub == 0 ? U(0) : s_last;
        U v8[RPT];
#pragma unroll
        for (int k = 0; k < RPT; k++) v8[k] = (U)(last + dec[k]);
        const long long g0 = c0 + threadIdx.x * RPT;
        if (gfc_rows_whole(out, n, c0, rows)) {
            constexpr int V = 16 / (int)sizeof(U);
            u32x4* d4 = reinterpret_cast<u32x4*>(out + g0);
#pragma unroll
            for (int q = 0; q < RPT / V; q++) {
                u32x4 t;
                __builtin_memcpy(&t, &v8[q * V], 16);
                d4[q] = t;
            }
        } else {
#pragma unroll
            for (int k = 0; k < RPT; k++)
                if (g0 + k < n) out[g0 + k] = v8[k];
        }
    }
    if (threadIdx.x == 0 && s_bad) __hip_atomic_store(error, 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One wave per chunk (chunks of at most MAXPER rows): the chunk is staged
// in LDS, the wave walks the subchunk headers, then each lane decodes
// MAXPER / 64 consecutive rows (16 for float, 32 for double) and stores them
// as whole vectors.  A workgroup of one wave needs a quarter of the LDS of a
// chunk per 256-thread workgroup's idle waves, so four times as many chunks
// are in flight on a CU while their headers are walked.
template <typename U, int MAXPER>
__global__ __launch_bounds__(kWave) void k_gfc_decode_wave(const uint8_t* __restrict__ payload, size_t payload_bytes,
                                                           const uint8_t* __restrict__ offsets, long long n,
                                                           long long N, long long per, int S, U* __restrict__ out,
                                                           const uint64_t* __restrict__ prefix, uint32_t* error) {
    constexpr bool dbl = sizeof(U) == 8;
    constexpr int RPL = MAXPER / kWave;  // rows per lane
    constexpr int LPS = 32 / RPL;        // lanes per subchunk
    static_assert(LPS == 1 || LPS == 2, "one or two lanes per 32-row subchunk");
    constexpr int kMaxBytes = MAXPER / 32 * (16 + 32 * (int)sizeof(U));
    constexpr int kVec = (kMaxBytes + 15 + 15) / 16;
    constexpr int kVR = (kVec + kWave - 1) / kWave;
    __shared__ __attribute__((aligned(16))) uint8_t buf[16 * kVec + 16];
    __shared__ uint32_t sub_off[MAXPER / 32 + 1];
    static_assert(sizeof(buf) >= MAXPER * sizeof(U), "the decoded chunk fits the staging buffer");
    const int lane = threadIdx.x;
    const uint32_t chunk = blockIdx.x;
    const long long c0 = (long long)chunk * per;
    const int rows = gfc_chunk(chunk, N, per).rows;
    const int nsub = rows / 32;
    uint64_t chunk_bytes = gfc_chunk_bytes<U>(offsets, chunk, N, per, S);
    bool bad = false;
    if (chunk_bytes > (uint64_t)kMaxBytes) {  // more bytes than MAXPER rows can hold: corrupt
        bad = true;
        chunk_bytes = 0;
    }
    const uint64_t P = prefix[chunk];
    const uintptr_t a0 = reinterpret_cast<uintptr_t>(payload + P);
    const uint32_t base = (uint32_t)(a0 & 15);
    {
        const u32x4* gv = reinterpret_cast<const u32x4*>(a0 - base);
        const uintptr_t pend = reinterpret_cast<uintptr_t>(payload) + payload_bytes;
        const uint32_t nvec = (base + (uint32_t)chunk_bytes + 15) / 16;
        u32x4 t[kVR];
#pragma unroll
        for (int r = 0; r < kVR; r++) {
            const uint32_t k = lane + r * kWave;
            t[r] = (k < nvec && reinterpret_cast<uintptr_t>(gv + k) < pend) ? gv[k] : u32x4{0u, 0u, 0u, 0u};
        }
        u32x4* b4 = reinterpret_cast<u32x4*>(buf);
#pragma unroll
        for (int r = 0; r < kVR; r++) {
            const uint32_t k = lane + r * kWave;
            if (k < nvec) b4[k] = t[r];
        }
        if (P + chunk_bytes > payload_bytes) bad = true;
    }
    __syncthreads();
    const uint32_t* bw = reinterpret_cast<const uint32_t*>(buf);
    constexpr uint32_t kBufWords = sizeof(buf) / 4;
    auto lds_word = [&](uint32_t a) -> uint32_t {  // chunk bytes [a, a + 4); past the buffer: zero
        a += base;
        const uint32_t q = a >> 2, sh = a & 3;
        const uint32_t lo = q < kBufWords ? bw[q] : 0u;
        if (!sh) return lo;
        const uint32_t hi = q + 1 < kBufWords ? bw[q + 1] : 0u;
        return __builtin_amdgcn_alignbyte(hi, lo, sh);
    };
    // header walk, lanes 0..31 reading one nibble each
    {
        uint32_t off = 0;
        for (int s = 0; s < nsub; s++) {
            int nb = 0;
            if (lane < 32 && off + 16 <= chunk_bytes) {
                const uint8_t byte = buf[base + off + (lane >> 1)];
                const int nib = (lane & 1) ? (byte >> 4) : (byte & 15);
                const int code = nib & 7;
                nb = dbl ? (code >= 2 ? code + 1 : code) : code;
            }
            const int tot = __builtin_amdgcn_readlane(wave_inclusive_scan(nb), 31);
            if (lane == 0) sub_off[s] = off;
            off += 16 + tot;
        }
        if (off != chunk_bytes) bad = true;
    }
    __syncthreads();
    auto nb_of = [](uint32_t nib) -> uint32_t {
        const uint32_t code = nib & 7;
        return dbl ? (code >= 2 ? code + 1 : code) : code;
    };
    const int sub = lane / LPS, part = lane % LPS;
    const int r0 = sub * 32 + part * RPL;  // first row of this lane in the chunk
    const bool mine = r0 < rows;
    const uint32_t so = mine ? sub_off[sub] : 0u;
    uint32_t hw[RPL / 8];
    uint32_t own = 0;
#pragma unroll
    for (int w = 0; w < RPL / 8; w++) {
        hw[w] = mine ? lds_word(so + (uint32_t)(part * RPL / 2 + 4 * w)) : 0u;
#pragma unroll
        for (int k = 0; k < 8; k++) own += nb_of(hw[w] >> (4 * k));
    }
    uint32_t pre = 0;
    if (LPS == 2) {
        const uint32_t up = __shfl_up(own, 1, 2);
        pre = part ? up : 0u;
    }
    U dec[RPL];
    uint32_t pos = so + 16 + pre;
#pragma unroll
    for (int k = 0; k < RPL; k++) {
        const uint32_t nib = (hw[k / 8] >> (4 * (k % 8))) & 15;
        const uint32_t nb = mine ? nb_of(nib) : 0u;
        uint64_t dv = mine ? lds_word(pos) : 0u;
        if (dbl && nb > 4) dv |= (uint64_t)lds_word(pos + 4) << 32;
        dv &= nb >= 8 ? ~0ull : (1ull << (8 * nb)) - 1ull;
        pos += nb;
        U diff = (U)dv;
        if (nib & 8) diff = (U)(U(0) - diff);
        dec[k] = diff;
    }
    // row 31 of the chunk (the last row of lane LPS - 1) predicts every later subchunk
    U last;
    if constexpr (dbl) {
        const uint64_t v = (uint64_t)dec[RPL - 1];
        const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, LPS - 1);
        const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), LPS - 1);
        last = (U)(((uint64_t)hi << 32) | lo);
    } else {
        last = (U)__builtin_amdgcn_readlane((uint32_t)dec[RPL - 1], LPS - 1);
    }
    if (sub == 0) last = U(0);
    __syncthreads();  // every read of the staged payload is done
    if (c0 + rows <= n && (reinterpret_cast<uintptr_t>(out + c0) & 15) == 0) {  // uniform
        // through LDS, so that consecutive lanes store consecutive 16-byte
        // vectors of the chunk
        U* ob = reinterpret_cast<U*>(buf);
        if (mine) {
#pragma unroll
            for (int k = 0; k < RPL; k++) ob[r0 + k] = (U)(last + dec[k]);
        }
        __syncthreads();
        const u32x4* b4 = reinterpret_cast<const u32x4*>(buf);
        u32x4* d4 = reinterpret_cast<u32x4*>(out + c0);
        const int nvec = rows * (int)sizeof(U) / 16;  // rows is a multiple of 32
        for (int i = lane; i < nvec; i += kWave) d4[i] = b4[i];
    } else if (mine) {
        const long long g0 = c0 + r0;
        U v[RPL];
#pragma unroll
        for (int k = 0; k < RPL; k++) v[k] = (U)(last + dec[k]);
        if (r0 + RPL <= rows && g0 + RPL <= n && (reinterpret_cast<uintptr_t>(out + g0) & 15) == 0) {
            constexpr int V = 16 / (int)sizeof(U);
            u32x4* d4 = reinterpret_cast<u32x4*>(out + g0);
#pragma unroll
            for (int q = 0; q < RPL / V; q++) {
                u32x4 t;
                __builtin_memcpy(&t, &v[q * V], 16);
                d4[q] = t;
            }
        } else {
#pragma unroll
            for (int k = 0; k < RPL; k++)
                if (g0 + k < n && r0 + k < rows) out[g0 + k] = v[k];
        }
    }
    if (lane == 0 && bad) __hip_atomic_store(error, 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace

GfcPlan gfc_plan(Ctx* c, bool dbl, const void* x, size_t n, DevBuf& meta, DevBuf& offsets) {
    const GfcGeom g = gfc_geom(dbl, (long long)n);
    GfcPlan p;
    p.dbl = dbl;
    p.x = x;
    p.n = n;
    p.N = g.N;
    p.per = g.per;
    p.chunks = (size_t)g.Wc;
    p.scan = std::make_shared<TileScan>(c, p.chunks);
    offsets = DevBuf(c, 4 * p.chunks);
    auto run = [&](auto tag) {
        using U = decltype(tag);
        kernel_bytes(c, "k_gfc_sizes", (double)n * sizeof(U) + 12.0 * (double)p.chunks);
        { KernelTimer kt_(c, "k_gfc_sizes"); hipLaunchKernelGGL(k_gfc_sizes<U>, dim3((unsigned)p.chunks), dim3(kBlock), 0, c->stream,
                           static_cast<const U*>(x), g.n, g.N, g.per, (int)g.S, p.scan->vals, offsets.get<int32_t>()); }
        HIPC(hipGetLastError());
    };
    if (dbl)
        run(uint64_t());
    else
        run(uint32_t());
    p.scan->scan(c);
    int32_t h_meta[3] = {(int32_t)g.blocks, (int32_t)g.wpb, (int32_t)n};
    meta = make_buf(c, h_meta, 12);
    d2h(c, &p.payload_bytes, p.scan->prefix + p.chunks, 8);
    return p;
}

void gfc_write(Ctx* c, const GfcPlan& p, uint8_t* payload) {
    auto run = [&](auto tag) {
        using U = decltype(tag);
        kernel_bytes(c, "k_gfc_encode", (double)p.n * sizeof(U) + (double)p.payload_bytes);
        { KernelTimer kt_(c, "k_gfc_encode"); hipLaunchKernelGGL(k_gfc_encode<U>, dim3((unsigned)p.chunks), dim3(kBlock), 0, c->stream,
                           static_cast<const U*>(p.x), (long long)p.n, p.N, p.per, payload, (const uint64_t*)p.scan->prefix); }
        HIPC(hipGetLastError());
    };
    if (p.dbl)
        run(uint64_t());
    else
        run(uint32_t());
}

void gfc_encode(Ctx* c, bool dbl, const void* x, size_t n, DevBuf& meta, DevBuf& offsets, DevBuf& payload) {
    const GfcPlan p = gfc_plan(c, dbl, x, n, meta, offsets);
    payload = DevBuf(c, std::max<size_t>(p.payload_bytes, 1));
    payload.set_size(p.payload_bytes);
    gfc_write(c, p, payload.get<uint8_t>());
}

void gfc_decode(Ctx* c, bool dbl, const uint8_t* meta, size_t meta_bytes, const uint8_t* offsets, size_t off_bytes,
                const uint8_t* payload, size_t payload_bytes, DevBuf& out) {
    if (meta_bytes < 12) fail(DDJ_ECORRUPT, "GFC metadata shorter than 12 bytes");
    int32_t h_meta[3];
    read_small(c, h_meta, meta, 12);
    if (h_meta[0] <= 0 || h_meta[1] <= 0 || h_meta[2] <= 0) fail(DDJ_ECORRUPT, "GFC metadata out of range");
    const long long n = h_meta[2];
    GfcGeom g = gfc_geom(dbl, n);
    if (h_meta[0] != g.blocks || h_meta[1] != g.wpb) {
        // other launch shapes are legal in the format: recompute the chunking from them
        g.blocks = h_meta[0];
        g.wpb = h_meta[1];
        g.Wc = g.blocks * g.wpb;
        long long per = (g.N + g.Wc - 1) / g.Wc;
        if (per < 32) per = 32;
        g.per = (per + 31) & ~31LL;
        if (g.per > kGfcMaxPer) fail(DDJ_EUNSUPPORTED, "GFC chunks longer than 2048 rows");
    }
    const size_t tiles = (size_t)g.Wc;
    if (off_bytes < 4 * tiles) fail(DDJ_ECORRUPT, "GFC offsets shorter than the chunk count");
    TileScan ts(c, tiles);
    out = DevBuf(c, (size_t)n * (dbl ? 8 : 4));
    uint32_t* err_flag = &c->control->error_flag;  // zero between calls
    auto run = [&](auto tag) {
        using U = decltype(tag);
        { KernelTimer kt_(c, "k_gfc_chunk_bytes"); hipLaunchKernelGGL(k_gfc_chunk_bytes<U>, dim3((unsigned)((tiles + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           c->stream, offsets, tiles, g.N, g.per, (int)g.S, ts.vals); }
        HIPC(hipGetLastError());
        ts.scan(c);
        kernel_bytes(c, "k_gfc_chunk_bytes", 12.0 * (double)tiles);
        kernel_bytes(c, "k_gfc_decode", (double)payload_bytes + 4.0 * (double)tiles + (double)n * sizeof(U));
        // float chunks (at most 1024 rows, 4.6 KB staged) decode one wave
        // each, walk included: 2.1x faster than four waves a chunk (float32
        // walk, 2^26 rows, 273 -> 132 us).
        if constexpr (sizeof(U) == 4) {
            if (g.per <= 1024) {
                KernelTimer kt_(c, "k_gfc_decode");
                hipLaunchKernelGGL((k_gfc_decode_wave<U, 1024>), dim3((unsigned)tiles), dim3(kWave), 0, c->stream,
                                   payload, payload_bytes, offsets, g.n, g.N, g.per, (int)g.S, out.get<U>(),
                                   (const uint64_t*)ts.prefix, err_flag);
                HIPC(hipGetLastError());
                return;
            }
        }
        // Double chunks stage 17 KB (nine a CU): the header walks run first,
        // one lane per chunk (k_gfc_walk), so that a decoding workgroup holds
        // its LDS only while it loads, decodes and stores (float64 walk, 2^26
        // rows, decode end to end 348 -> 326 us).
        const int stride = (int)(g.per / 32 + 1);
        DevBuf subs(c, 4 * tiles * (size_t)stride);
        kernel_bytes(c, "k_gfc_walk", 20.0 * (double)(g.N / 32) + 4.0 * (double)tiles);
        {
            KernelTimer kw_(c, "k_gfc_walk");
            hipLaunchKernelGGL(k_gfc_walk<U>, dim3((unsigned)((tiles + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                               c->stream, payload, payload_bytes, offsets, tiles, g.N, g.per, (int)g.S,
                               (const uint64_t*)ts.prefix, subs.get<uint32_t>(), stride, err_flag);
            HIPC(hipGetLastError());
        }
        KernelTimer kt_(c, "k_gfc_decode");
        hipLaunchKernelGGL(k_gfc_decode<U>, dim3((unsigned)tiles), dim3(kBlock), 0, c->stream, payload,
                           payload_bytes, offsets, g.n, g.N, g.per, (int)g.S, out.get<U>(),
                           (const uint64_t*)ts.prefix, err_flag, subs.get<uint32_t>(), stride);
        HIPC(hipGetLastError());
    };
    if (dbl)
        run(uint64_t());
    else
        run(uint32_t());
    uint32_t err = 0;
    d2h(c, &err, err_flag, 4);
    if (err) {
        const uint32_t zero = 0;
        h2d(c, err_flag, &zero, 4);
        fail(DDJ_ECORRUPT, "GFC chunk sizes disagree with their headers");
    }
}

}  // namespace ddj
