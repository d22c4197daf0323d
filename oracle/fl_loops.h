/*
 * fl_loops.h -- the reference's FL (AFL, CWARP_SIZE = 1) host loops, restated.
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

/* ------------------------------------------------------------- FL (AFL w=1)
 * afl_compress_base_gpu / afl_decompress_base_gpu with CWARP_SIZE = 1
 * (compression/afl/afl_gpu.cu:124-195), driven by the host loops of
 * run_afl_compress_cpu / run_afl_decompress_cpu (afl_gpu.cu:6-65): "thread" t
 * packs values [t*W, t*W+W) into words [t*b, t*b+b). */
template <typename T>
void afl_compress_base(unsigned bit_length, unsigned long data_id, unsigned long comp_data_id, const T* data,
                       T* compressed_data, unsigned long length, unsigned long comprLength) {
    const unsigned W = 8 * sizeof(T);
    T v1, value = 0;
    unsigned v1_pos = 0, v1_len;
    unsigned long pos = comp_data_id, pos_data = data_id;
    for (unsigned i = 0; (i < W) && (pos_data < length); ++i) {
        v1 = data[pos_data];
        pos_data += 1;
        if (v1_pos >= W - bit_length && (pos < comprLength)) {
            v1_len = W - v1_pos;
            value = (T)(ubits(value) | (getnbits(v1, v1_len) << v1_pos));
            compressed_data[pos] = value;
            v1_pos = bit_length - v1_len;
            value = (T)getnpbits(v1, v1_pos, v1_len);
            pos += 1;
        } else {
            v1_len = bit_length;
            value = (T)(ubits(value) | shl<T>(getnbits(v1, v1_len), v1_pos));
            v1_pos += v1_len;
        }
    }
    if ((pos_data >= length) && (pos_data < length + 1) && (pos < comprLength)) compressed_data[pos] = value;
}

template <typename T>
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
            pos += 1;
            v1 = pos < words ? compressed_data[pos] : (T)0;
            v1_pos = bit_length - v1_len;
            ret = (T)(ubits(ret) | shl<T>(getnbits(v1, v1_pos), v1_len));
        } else {
            v1_len = bit_length;
            ret = (T)getnpbits(v1, v1_len, v1_pos);
            v1_pos += v1_len;
        }
        data[pos_decomp] = ret;
        pos_decomp += 1;
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

} /* namespace orc_fl */
#endif
