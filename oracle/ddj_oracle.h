/*
 * ddj_oracle.h -- CPU restatement of the reference column-compression codecs.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  It is never linked into, called by or shipped with the product
 * library (libddj_hip.so).
 *
 * Every function restates, in plain sequential host code, what the reference
 * computes (citations are to dzitkowskik/GpuTimeSeriesCompressionOptimizer,
 * paths relative to its src/ directory).  Buffers are host memory.
 *
 * Output buffers are allocated with malloc() by the oracle and released with
 * orc_free().  Status codes: 0 = ok, -1 = invalid argument, -2 = unsupported
 * (the reference throws or returns nothing), -3 = corrupt input.
 */
#ifndef DDJ_ORACLE_H_
#define DDJ_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* EncodingType values: compression/encoding_type.hpp:17-30 */
enum { ORC_DELTA = 0, ORC_AFL, ORC_SCALE, ORC_RLE, ORC_DICT, ORC_UNIQUE, ORC_PATCH,
       ORC_NONE, ORC_FLOAT2INT, ORC_CONST, ORC_GFC, ORC_OUTSIDE_PATCH = 100 };

/* DataType codes: the reference's DataType enum lives in the absent
 * time-series-data-manager submodule, so its numeric values are unknown.
 * The order below is the order of Encoding::Encode's switch
 * (compression/encoding.hpp:30-46). */
enum { ORC_D_INT = 0, ORC_D_TIME = 1, ORC_D_FLOAT = 2, ORC_D_DOUBLE = 3,
       ORC_D_SHORT = 4, ORC_D_CHAR = 5 };

typedef struct { uint8_t* p; size_t n; } orc_buf;

/* DataStatistics: util/statistics/cuda_array_statistics.hpp:24-34 */
typedef struct {
    double min;
    double max;
    int32_t minBitCnt;
    int32_t precision;
    int32_t sorted;
    float rlMetric;
    double mean;
    uint64_t size;
} orc_stats;

size_t orc_dtype_size(int dtype);
int orc_bitlen(int dtype, const void* minp, const void* maxp);   /* BITLEN, core/cuda_macros.cuh:168-189 */
int orc_alt_bitlen(int v);                                        /* ALT_BITLEN, core/cuda_macros.cuh:191-205 */
int orc_float_precision(float x);                                 /* _getFloatPrecision, cuda_array_statistics.cu:37-44 */

/* n_results / return types: Encoding::GetNumberOfResults / GetReturnTypes */
int orc_encoding_info(int enc, int dtype, int* n_results, int* return_types /* [2] */);

/* Encoding::Encode: outs[0] = metadata, outs[1..n_results] = outputs */
int orc_encode(int enc, int dtype, const void* in, size_t in_bytes, orc_buf* outs, int* n_outs);
/* Encoding::Decode: ins[0] = metadata, ins[1..] = children outputs */
int orc_decode(int enc, int dtype, const orc_buf* ins, int n_ins, orc_buf* out);
/* Encoding::GetMetadataSize / GetCompressedSize */
int orc_predict_sizes(int enc, int dtype, const void* in, size_t in_bytes, size_t* meta, size_t* data);
/* CudaArrayStatistics::GenerateStatistics */
int orc_stats_compute(int dtype, const void* in, size_t in_bytes, orc_stats* out);
/* Histogram::GetMostFrequent: K most frequent keys (count desc, key asc) */
int orc_most_frequent(int dtype, const void* in, size_t in_bytes, int k, void* keys_out, int* n_distinct);
/* Stencil::pack of an int 0/1 stencil; Stencil::unpack */
int orc_stencil_pack(const int32_t* s, size_t n, orc_buf* out);
int orc_stencil_unpack(const uint8_t* p, size_t nbytes, orc_buf* out);

/* run_afl_compress_cpu<T,1> / run_afl_decompress_cpu<T,1> (compression/afl/afl_gpu.cu:6-65):
 * the reference's only CPU path.  words = ceil(n*b/W) words of T. */
int orc_fl_pack(int dtype, int bits, const void* in, size_t n, void* out, size_t out_words);
int orc_fl_unpack(int dtype, int bits, const void* in, size_t n, void* out);

/* The CWARP_SIZE = 32 instantiation of the same loops (afl_gpu.cu:67-111,
 * afl_encoding_benchmark.cpp:44) over whole blocks of 32 lanes x W values:
 * out_words = ceil(n / (32*W)) * 32 * b, words no lane writes stay zero. */
int orc_fl_pack_lanes(int dtype, int bits, const void* in, size_t n, void* out, size_t out_words);
int orc_fl_unpack_lanes(int dtype, int bits, const void* in, size_t in_words, size_t n, void* out);
/* afl_decompress_base_value_gpu (afl_gpu.cu:197-223) at each of rows[0..count). */
int orc_fl_fetch(int dtype, int bits, int lanes, const void* in, size_t n, const uint64_t* rows, size_t count,
                 void* out);

void orc_free(orc_buf* b);

#ifdef __cplusplus
}
#endif
#endif
