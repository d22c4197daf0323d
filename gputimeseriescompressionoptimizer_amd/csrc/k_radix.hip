// Exclusive scans of long value arrays and an LSD radix sort of unsigned
// keys: the sort behind the standalone unique encoding (FindUnique's
// thrust::sort, reference compression/unique/unique_encoding.cu:35-46).
//
// Scan: k_scan_tiles runs in one workgroup, which is right for the few
// thousand tile values most launches have but not for millions.  Long
// arrays go reduce-then-scan over 4096-value blocks: block sums, one
// k_scan_tiles over them, then every block rescans its values on top of its
// block's prefix.  No workgroup waits on another.
//
// Radix sort: 8-bit digits, least significant first, stable.  Per pass a
// counting kernel writes each 4096-key tile's digit histogram digit-major
// (counts[d * tiles + t]), so that one exclusive scan of that array gives
// every (digit, tile) its output offset; the scatter kernel then places the
// tile's keys in 16 rounds of 256, ranking keys of equal digit inside a
// round by wave ballots (the lanes that agree on all eight digit bits) and
// across the round's four waves by per-wave digit counts in LDS.
#include "k_common.h"
#include "kernels.h"

namespace ddj {

namespace {

constexpr int kScanBlockVals = 4096;  // values per block of the long scan, 16 per thread
constexpr int kPer = kScanBlockVals / kBlock;

__global__ __launch_bounds__(kBlock) void k_scan_block_sums(const uint64_t* __restrict__ vals, size_t count,
                                                            uint64_t* __restrict__ sums) {
    __shared__ uint64_t red[kBlock / kWave];
    const size_t base = (size_t)blockIdx.x * kScanBlockVals;
    uint64_t s = 0;
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        const size_t i = base + (size_t)r * kBlock + threadIdx.x;
        s += i < count ? vals[i] : 0;
    }
    uint64_t total;
    block_exclusive_scan<uint64_t>(s, &total, red);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// prefix[i] = block_prefix[block] + the block's values before i; the last
// block also writes prefix[count].
__global__ __launch_bounds__(kBlock) void k_scan_block_apply(const uint64_t* __restrict__ vals, size_t count,
                                                             const uint64_t* __restrict__ block_prefix,
                                                             uint64_t* __restrict__ prefix) {
    __shared__ uint64_t buf[kScanBlockVals + kScanBlockVals / 16];
    __shared__ uint64_t red[kBlock / kWave];
    auto pad = [](int i) { return i + (i >> 4); };
    const size_t base = (size_t)blockIdx.x * kScanBlockVals;
#pragma unroll
    for (int r = 0; r < kPer; r++) {  // coalesced loads, staged in LDS
        const int j = r * kBlock + threadIdx.x;
        buf[pad(j)] = base + j < count ? vals[base + j] : 0;
    }
    __syncthreads();
    uint64_t v[kPer], s = 0;
#pragma unroll
    for (int k = 0; k < kPer; k++) {  // each thread scans 16 consecutive values
        v[k] = buf[pad(threadIdx.x * kPer + k)];
        s += v[k];
    }
    uint64_t total;
    uint64_t run = block_prefix[blockIdx.x] + block_exclusive_scan<uint64_t>(s, &total, red);
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        buf[pad(threadIdx.x * kPer + k)] = run;
        run += v[k];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        const int j = r * kBlock + threadIdx.x;
        if (base + j < count) prefix[base + j] = buf[pad(j)];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) prefix[count] = block_prefix[blockIdx.x] + total;
}

constexpr int kRadixTile = 4096;  // keys per workgroup
constexpr int kRadixRounds = kRadixTile / kBlock;
static_assert(kBlock == 256, "one thread per digit value");

template <typename K>
__global__ __launch_bounds__(kBlock) void k_radix_count(const K* __restrict__ keys, size_t n, int shift,
                                                        uint64_t* __restrict__ counts, size_t tiles) {
    __shared__ uint32_t hist[256];
    hist[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kRadixTile;
    K k[kRadixRounds];
#pragma unroll
    for (int r = 0; r < kRadixRounds; r++) {
        const size_t i = base + (size_t)r * kBlock + threadIdx.x;
        k[r] = i < n ? keys[i] : K(0);
    }
#pragma unroll
    for (int r = 0; r < kRadixRounds; r++) {
        const size_t i = base + (size_t)r * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&hist[(uint32_t)(k[r] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[(size_t)threadIdx.x * tiles + blockIdx.x] = hist[threadIdx.x];
}

template <typename K>
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const K* __restrict__ keys, size_t n, int shift,
                                                          const uint64_t* __restrict__ offsets, size_t tiles,
                                                          K* __restrict__ out) {
    constexpr int NW = kBlock / kWave;
    __shared__ uint64_t base_of[256];      // next output row of each digit for this tile
    __shared__ uint32_t wcnt[NW][256];     // keys of each digit per wave in the current round
    const int t = threadIdx.x, lane = t & (kWave - 1), wid = t / kWave;
    base_of[t] = offsets[(size_t)t * tiles + blockIdx.x];
#pragma unroll
    for (int w = 0; w < NW; w++) wcnt[w][t] = 0;
    const size_t base = (size_t)blockIdx.x * kRadixTile;
    K k[kRadixRounds];
#pragma unroll
    for (int r = 0; r < kRadixRounds; r++) {
        const size_t i = base + (size_t)r * kBlock + t;
        k[r] = i < n ? keys[i] : K(0);
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
    for (int r = 0; r < kRadixRounds; r++) {
        const size_t i = base + (size_t)r * kBlock + t;
        const bool valid = i < n;
        const uint32_t d = (uint32_t)(k[r] >> shift) & 255u;
        uint64_t same = __ballot(valid);  // lanes whose digit equals this lane's
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            const bool on = (d >> bit) & 1u;
            const uint64_t b = __ballot(on);
            same &= on ? b : ~b;
        }
        const uint32_t before = (uint32_t)__builtin_popcountll(same & below);
        if (valid && before == 0) wcnt[wid][d] = (uint32_t)__builtin_popcountll(same);
        __syncthreads();
        if (valid) {
            uint64_t pos = base_of[d] + before;
            for (int w = 0; w < wid; w++) pos += wcnt[w][d];
            out[pos] = k[r];
        }
        __syncthreads();
        uint64_t add = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) {
            add += wcnt[w][t];
            wcnt[w][t] = 0;
        }
        base_of[t] += add;
        __syncthreads();
    }
}

}  // namespace

void launch_scan_large(Ctx* c, const uint64_t* vals, size_t count, uint64_t* prefix) {
    if (count <= (size_t)kScanRows * kScanThreads) {  // one step of the one-workgroup scan
        launch_scan_tiles(c, vals, count, prefix);
        return;
    }
    const size_t blocks = (count + kScanBlockVals - 1) / kScanBlockVals;
    TileScan bs(c, blocks);
    kernel_bytes(c, "k_scan_block_sums", 8.0 * (double)count);
    {
        KernelTimer kt_(c, "k_scan_block_sums");
        hipLaunchKernelGGL(k_scan_block_sums, dim3((unsigned)blocks), dim3(kBlock), 0, c->stream, vals, count, bs.vals);
    }
    HIPC(hipGetLastError());
    bs.scan(c);
    kernel_bytes(c, "k_scan_block_apply", 16.0 * (double)count);
    {
        KernelTimer kt_(c, "k_scan_block_apply");
        hipLaunchKernelGGL(k_scan_block_apply, dim3((unsigned)blocks), dim3(kBlock), 0, c->stream, vals, count,
                           (const uint64_t*)bs.prefix, prefix);
    }
    HIPC(hipGetLastError());
}

template <typename K> K* radix_sort(Ctx* c, K* keys, K* tmp, size_t n, int key_bits) {
    if (n <= 1) return keys;
    const size_t tiles = (n + kRadixTile - 1) / kRadixTile;
    DevBuf counts(c, sizeof(uint64_t) * (256 * tiles + 1) * 2);
    uint64_t* cnt = counts.get<uint64_t>();
    uint64_t* off = cnt + 256 * tiles + 1;
    K* src = keys;
    K* dst = tmp;
    for (int shift = 0; shift < key_bits; shift += 8) {
        kernel_bytes(c, "k_radix_count", (double)n * sizeof(K));
        {
            KernelTimer kt_(c, "k_radix_count");
            hipLaunchKernelGGL(k_radix_count<K>, dim3((unsigned)tiles), dim3(kBlock), 0, c->stream, src, n, shift, cnt,
                               tiles);
        }
        HIPC(hipGetLastError());
        launch_scan_large(c, cnt, 256 * tiles, off);
        kernel_bytes(c, "k_radix_scatter", 2.0 * (double)n * sizeof(K));
        {
            KernelTimer kt_(c, "k_radix_scatter");
            hipLaunchKernelGGL(k_radix_scatter<K>, dim3((unsigned)tiles), dim3(kBlock), 0, c->stream, src, n, shift,
                               (const uint64_t*)off, tiles, dst);
        }
        HIPC(hipGetLastError());
        K* t = src;
        src = dst;
        dst = t;
    }
    return src;
}

template uint32_t* radix_sort<uint32_t>(Ctx*, uint32_t*, uint32_t*, size_t, int);
template uint64_t* radix_sort<uint64_t>(Ctx*, uint64_t*, uint64_t*, size_t, int);

}  // namespace ddj
