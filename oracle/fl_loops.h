/*
 * fl_loops.h -- the reference's FL (AFL) host loops, restated (CWARP_SIZE 1 and 32).
 *
 * TEST INFRASTRUCTURE ONLY: shared by the CPU oracle (ddj_oracle.cpp) and the
 * CPU baseline timing (fl_cpu.cpp).  Citations are file:line in the
 * reference's src/ tree.
 */
#ifndef ORC_FL_LOOPS_H_
#define ORC_FL_LOOPS_H_

#include <cstddef>
#include <cstdint>
#include <type_traits>

namespace orc_fl {

/* ---------------------------------------------------------------- bit macros
 * core/cuda_macros.cuh:51-99 (GETNPBITS = bfe: bits [start, start+n) of the
 * W-bit pattern; n == 0 gives 0) and :101-160 (GETNBITS = low n bits). */
template <typename T> uint64_t ubits(T v) {
    typedef typename std::make_unsigned<T>::type U;
    return (uint64_t)(U)v;
}
inline uint64_t lowmask(unsigned n) { return n >= 64 ? ~0ull : ((1ull << n) - 1ull); }
template <typename T> uint64_t getnpbits(T v, unsigned n, unsigned start) {
    const unsigned W = 8 * sizeof(T);
    if (n == 0 || start >= W) return 0;
    return (ubits(v) >> start) & lowmask(n);
}
template <typename T> uint64_t getnbits(T v, unsigned n) { return getnpbits(v, n, 0); }
template <typename T> T shl(uint64_t v, unsigned s) { return (T)(s >= 64 ? 0 : (v << s)); }

/* ------------------------------------------------------------------ FL (AFL)
 * afl_compress_base_gpu / afl_decompress_base_gpu (compression/afl/afl_gpu.cu:124-195)
 * for CWARP_SIZE = LANES, driven by the host loops of run_afl_compress_cpu /
 * run_afl_decompress_cpu (afl_gpu.cu:6-65).  LANES = 1: "thread" t packs
 * values [t*W, t*W+W) into words [t*b, t*b+b).  LANES = 32: lane l of block
 * k packs values k*32*W + l + 32*i into words k*32*b + l + 32*j. */
template <typename T, unsigned LANES = 1>
void afl_compress_base(unsigned bit_length, unsigned long data_id, unsigned long comp_data_id, const T* data,
                       T* compressed_data, unsigned long length, unsigned long comprLength) {
    const unsigned W = 8 * sizeof(T);
    T v1, value = 0;
    unsigned v1_pos = 0, v1_len;
    unsigned long pos = comp_data_id, pos_data = data_id;
    for (unsigned i = 0; (i < W) && (pos_data < length); ++i) {
        v1 = data[pos_data];
        pos_data += LANES;
        if (v1_pos >= W - bit_length && (pos < comprLength)) {
            v1_len = W - v1_pos;
            value = (T)(ubits(value) | (getnbits(v1, v1_len) << v1_pos));
            compressed_data[pos] = value;
            v1_pos = bit_length - v1_len;
            value = (T)getnpbits(v1, v1_pos, v1_len);
            pos += LANES;
        } else {
            v1_len = bit_length;
            value = (T)(ubits(value) | shl<T>(getnbits(v1, v1_len), v1_pos));
            v1_pos += v1_len;
        }
    }
    if ((pos_data >= length) && (pos_data < length + LANES) && (pos < comprLength)) compressed_data[pos] = value;
}

template <typename T, unsigned LANES = 1>
void afl_decompress_base(unsigned bit_length, unsigned long comp_data_id, unsigned long data_id, const T* compressed_data,
                         T* data, unsigned long length, unsigned long words) {
    const unsigned W = 8 * sizeof(T);
    unsigned long pos = comp_data_id, pos_decomp = data_id;
    unsigned v1_pos = 0, v1_len;
    T v1, ret;
    if (pos_decomp >= length) return; /* the reference tests '>' and then reads one word past the end */
    v1 = pos < words ? compressed_data[pos] : (T)0;
    for (unsigned i = 0; i < W && pos_decomp < length; ++i) {
        if (v1_pos > W - bit_length) {
            v1_len = W - v1_pos;
            ret = (T)getnpbits(v1, v1_len, v1_pos);
            pos += LANES;
            v1 = pos < words ? compressed_data[pos] : (T)0;
            v1_pos = bit_length - v1_len;
            ret = (T)(ubits(ret) | shl<T>(getnbits(v1, v1_pos), v1_len));
        } else {
            v1_len = bit_length;
            ret = (T)getnpbits(v1, v1_len, v1_pos);
            v1_pos += v1_len;
        }
        data[pos_decomp] = ret;
        pos_decomp += LANES;
    }
}

template <typename T> void fl_compress(unsigned bits, const T* data, T* out, unsigned long n, unsigned long words) {
    const unsigned long W = 8 * sizeof(T);
    unsigned long threads = (n + W - 1) / W;
    for (unsigned long t = 0; t < threads; t++) afl_compress_base<T>(bits, t * W, t * bits, data, out, n, words);
}

template <typename T> void fl_decompress(unsigned bits, const T* in, T* out, unsigned long n, unsigned long words) {
    const unsigned long W = 8 * sizeof(T);
    unsigned long threads = (n + W - 1) / W;
    for (unsigned long t = 0; t < threads; t++) afl_decompress_base<T>(bits, t * bits, t * W, in, out, n, words);
}

/* run_afl_compress_cpu<T,32> / run_afl_decompress_cpu<T,32> over whole
 * blocks of 32 lanes x W values: data_block = 32*k for lane group k
 * (afl_gpu.cu:13-28), comprLength = the whole blocks' 32*b words per block,
 * output words no lane writes left as the caller cleared them. */
template <typename T> void fl_compress_lanes(unsigned bits, const T* data, T* out, unsigned long n, unsigned long words) {
    const unsigned long W = 8 * sizeof(T);
    const unsigned long groups = (n + 32 * W - 1) / (32 * W);
    for (unsigned long g = 0; g < groups; g++)
        for (unsigned long l = 0; l < 32; l++)
            afl_compress_base<T, 32>(bits, 32 * g * W + l, 32 * g * bits + l, data, out, n, words);
}
template <typename T> void fl_decompress_lanes(unsigned bits, const T* in, T* out, unsigned long n, unsigned long words) {
    const unsigned long W = 8 * sizeof(T);
    const unsigned long groups = (n + 32 * W - 1) / (32 * W);
    for (unsigned long g = 0; g < groups; g++)
        for (unsigned long l = 0; l < 32; l++)
            afl_decompress_base<T, 32>(bits, 32 * g * bits + l, 32 * g * W + l, in, out, n, words);
}

/* afl_decompress_base_value_gpu (afl_gpu.cu:197-223): value `pos` of a
 * stream in LANES layout from its two candidate words. */
template <typename T, unsigned LANES>
T afl_value(unsigned bit_length, const T* compressed_data, unsigned long pos) {
    const unsigned long W = 8 * sizeof(T);
    const unsigned long data_block = pos / (LANES * W);
    const unsigned long pos_in_block = pos % (LANES * W);
    const unsigned long pos_in_warp_lane = pos_in_block % LANES;
    const unsigned long pos_in_warp_comp_block = pos_in_block / LANES;
    const unsigned long cblock_id = data_block * (LANES * bit_length) + pos_in_warp_lane +
                                    ((pos_in_warp_comp_block * bit_length) / W) * LANES;
    const unsigned bit_pos = (unsigned)(pos_in_warp_comp_block * bit_length % W);
    const unsigned bit_ret = bit_pos <= W - bit_length ? bit_length : (unsigned)(W - bit_pos);
    uint64_t ret = getnpbits(compressed_data[cblock_id], bit_ret, bit_pos);
    if (bit_ret < bit_length) ret |= getnbits(compressed_data[cblock_id + LANES], bit_length - bit_ret) << bit_ret;
    return (T)ret;
}

} /* namespace orc_fl */
#endif
