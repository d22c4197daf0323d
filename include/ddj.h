/*
 * ddj.h -- C ABI of the MI355X column-compression engine (libddj_hip.so).
 *
 * This is the drop-in boundary for the reference's lightweight lossless
 * column-compression path (dzitkowskik/GpuTimeSeriesCompressionOptimizer,
 * src/).  Every entry point takes plain pointers and sizes; device buffers
 * are ddj_buf handles allocated from the context's caching allocator.  All
 * GPU work of a context runs on its own HIP stream; calls on one context
 * are serialised by the caller, separate contexts are independent.
 *
 * Status codes replace the reference's exit(1) / print-and-continue error
 * handling (core/macros.h:8-17).
 */
#ifndef DDJ_H_
#define DDJ_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    DDJ_OK = 0,
    DDJ_EINVAL = 1,       /* bad argument (unknown encoding, type, size) */
    DDJ_EUNSUPPORTED = 2, /* the reference throws NotImplementedException or returns nothing */
    DDJ_ECORRUPT = 3,     /* encoded input fails validation */
    DDJ_ENOMEM = 4,       /* device allocation failed */
    DDJ_EHIP = 5,         /* HIP runtime error */
    DDJ_ETIMEOUT = 6      /* reserved: no kernel waits on another, so none is raised */
} ddj_status;

/* EncodingType (reference: compression/encoding_type.hpp:17-30) */
enum {
    DDJ_ENC_DELTA = 0,
    DDJ_ENC_AFL = 1,
    DDJ_ENC_SCALE = 2,
    DDJ_ENC_RLE = 3,
    DDJ_ENC_DICT = 4,
    DDJ_ENC_UNIQUE = 5,
    DDJ_ENC_PATCH = 6,
    DDJ_ENC_NONE = 7,
    DDJ_ENC_FLOAT_TO_INT = 8,
    DDJ_ENC_CONST = 9,
    DDJ_ENC_GFC = 10,
    DDJ_ENC_OUTSIDE_PATCH = 100 /* OutsidePatchEncoding (compression/patch/outside_patch_encoding.cu) */
};

/* DataType.  The reference's enum lives in its absent data-manager
 * submodule; these values follow the order of Encoding::Encode's switch
 * (compression/encoding.hpp:30-46) and are what tree headers carry. */
enum { DDJ_D_INT = 0, DDJ_D_TIME = 1, DDJ_D_FLOAT = 2, DDJ_D_DOUBLE = 3, DDJ_D_SHORT = 4, DDJ_D_CHAR = 5 };

typedef struct ddj_ctx ddj_ctx;

/* A device buffer: `dev` points to at least `bytes` bytes (0 bytes -> NULL). */
typedef struct {
    void* dev;
    size_t bytes;
} ddj_buf;

/* DataStatistics (util/statistics/cuda_array_statistics.hpp:24-34) plus the
 * planner features the reference does not have. */
typedef struct {
    double min;
    double max;
    int32_t minBitCnt;   /* BITLEN(double, double): 64 for n > 0, as the reference */
    int32_t precision;   /* max _getFloatPrecision over rows */
    int32_t sorted;      /* non-decreasing */
    float rlMetric;      /* RlMetric<T,2> */
    double mean;         /* Mean<T>: sum in T / n, converted to T */
    uint64_t size;       /* rows */
    /* features beyond the reference (learned planner) */
    uint64_t runs;            /* number of maximal runs of equal neighbours */
    uint64_t distinct_est;    /* distinct-value estimate (64-register HyperLogLog) */
    uint32_t bit_hist[65];    /* rows per bit width 0..64 of the zig-zag value */
    uint32_t delta_bits_max;  /* bit width of the largest |x[i+1]-x[i]| */
} ddj_stats_t;

/* ---------------------------------------------------------------- context */
ddj_status ddj_ctx_create(ddj_ctx** out, int device);
void ddj_ctx_destroy(ddj_ctx* ctx);
ddj_status ddj_ctx_sync(ddj_ctx* ctx);
const char* ddj_last_error(ddj_ctx* ctx);
/* Returns cached device memory to the driver. */
void ddj_ctx_trim(ddj_ctx* ctx);
/* The HIP stream of the context (hipStream_t), for callers that order their own work. */
void* ddj_ctx_stream(ddj_ctx* ctx);

/* Per-kernel timing with HIP events recorded on the context stream around
 * every launch while enabled (no counterpart in the reference; used by the
 * benchmark to time one kernel inside a timed region). */
ddj_status ddj_profile_enable(ddj_ctx* ctx, int on);
ddj_status ddj_profile_kernel(ddj_ctx* ctx, const char* kernel, double* total_ms, uint64_t* launches,
                              double* algo_bytes);
ddj_status ddj_profile_reset(ddj_ctx* ctx);

/* ----------------------------------------------------------------- memory */
ddj_status ddj_buf_alloc(ddj_ctx* ctx, size_t bytes, ddj_buf* out);
void ddj_buf_free(ddj_ctx* ctx, ddj_buf* buf);
/* Transfers ownership of a library buffer to the caller, who releases it
 * with hipFree (the reference's CudaPtr(T*, size) adopts such pointers,
 * core/cuda_ptr.hpp:49-56).  The pointer must not be passed to ddj_buf_free. */
ddj_status ddj_buf_detach(ddj_ctx* ctx, ddj_buf* buf);
ddj_status ddj_memcpy_h2d(ddj_ctx* ctx, void* dst_dev, const void* src_host, size_t bytes);
ddj_status ddj_memcpy_d2h(ddj_ctx* ctx, void* dst_host, const void* src_dev, size_t bytes);
ddj_status ddj_memcpy_d2d(ddj_ctx* ctx, void* dst_dev, const void* src_dev, size_t bytes);

/* -------------------------------------------------------------- encodings */
/* Encoding::GetNumberOfResults / GetReturnTypes (compression/encoding.hpp:74-75) */
ddj_status ddj_encoding_info(int enc, int dtype, uint32_t* n_results, int* return_dtypes /* [2] */);

/* Encoding::Encode(SharedCudaPtr<char>, DataType) (compression/encoding.hpp:28-47).
 * outs[0] = metadata, outs[1..n_results] = outputs; the caller frees them.
 * `outs` holds 1 + max(n_results, 1) entries: `none` (0 results) returns its
 * data copy as outs[1] (compression/none/none_encoding.cu:13-30). */
ddj_status ddj_encode(ddj_ctx* ctx, int enc, int dtype, const void* d_in, size_t in_bytes, ddj_buf* outs);

/* Encoding::Decode(SharedCudaPtrVector<char>, DataType) (compression/encoding.hpp:49-67).
 * ins[0] = metadata, ins[1..] = decoded children outputs. */
ddj_status ddj_decode(ddj_ctx* ctx, int enc, int dtype, const ddj_buf* ins, uint32_t n_ins, ddj_buf* out);

/* Encoding::GetMetadataSize / GetCompressedSize (compression/encoding.hpp:78-79) */
ddj_status ddj_predict_sizes(ddj_ctx* ctx, int enc, int dtype, const void* d_in, size_t in_bytes,
                             size_t* meta_bytes, size_t* data_bytes);

/* CudaArrayStatistics::GenerateStatistics (util/statistics/cuda_array_statistics.cu:159-179) */
ddj_status ddj_stats(ddj_ctx* ctx, int dtype, const void* d_in, size_t in_bytes, ddj_stats_t* out);

/* FL bit-packing alone, with a given bit width (run_afl_compress_gpu<T,1> /
 * run_afl_decompress_gpu<T,1>, compression/afl/afl_gpu.cu:67-81). */
ddj_status ddj_fl_pack(ddj_ctx* ctx, int dtype, int bits, const void* d_in, size_t n, void* d_out);
ddj_status ddj_fl_unpack(ddj_ctx* ctx, int dtype, int bits, const void* d_in, size_t in_bytes, size_t n, void* d_out);

/* Random access into an FL stream: afl_decompress_value_gpu /
 * afl_decompress_base_value_gpu (compression/afl/afl_gpu.cu:113-121, :197-223),
 * generalised from "rows 0..length" to any set of rows.  d_out[i] = value
 * d_rows[i] (uint64, device) of the n-value b-bit stream at d_stream (any
 * byte address), in `lanes` layout: 1 = the contiguous stream AflEncoding
 * writes (and ddj_fl_pack), 32 = the interleaved stream of ddj_fl_pack_lanes.
 * A row >= n returns DDJ_EINVAL (its output slot reads 0). */
ddj_status ddj_fl_fetch(ddj_ctx* ctx, int dtype, int bits, int lanes, const void* d_stream, size_t stream_bytes,
                        size_t n, const uint64_t* d_rows, size_t count, void* d_out);

/* The interleaved FL layout: run_afl_compress_gpu<T,32> /
 * run_afl_decompress_gpu<T,32> (compression/afl/afl_gpu.cu:67-111, the
 * CWARP_SIZE = 32 instantiation of afl_encoding_benchmark.cpp:44).  Blocks
 * of 32 lanes x W values (W = bits of T); lane l of block k packs values
 * k*32*W + l + 32*i LSB first into b words stored at k*32*b + l + 32*j.
 * The last block is stored whole with its missing values packed as zero:
 * ddj_fl_lanes_bytes(dtype, bits, n) = ceil(n / (32*W)) * 32 * b words of T
 * (0 for a bad type or width).  For n a multiple of 32*W the stream equals
 * the reference's byte for byte; for other n the reference keeps only its
 * first ceil(b*n/W) words and so loses rows of the last block.  Column and
 * stream are aligned to the row type. */
size_t ddj_fl_lanes_bytes(int dtype, int bits, size_t n);
ddj_status ddj_fl_pack_lanes(ddj_ctx* ctx, int dtype, int bits, const void* d_in, size_t n, void* d_out);
ddj_status ddj_fl_unpack_lanes(ddj_ctx* ctx, int dtype, int bits, const void* d_in, size_t in_bytes, size_t n,
                               void* d_out);

/* ------------------------------------------------------------------- tree */
/* CompressionTree built from a pre-order path of encodings
 * (PathGenerator::GenerateTree, optimizer/path_generator.cpp:70-91); leaves
 * missing from the path are completed with `none` (CompressionNode::Fix). */
ddj_status ddj_tree_compress(ddj_ctx* ctx, const int* preorder, size_t path_len, int dtype, const void* d_in,
                             size_t in_bytes, ddj_buf* blob, double* node_ratios /* [path_len] or NULL */);
/* CompressionTree::Decompress (tree/compression_tree.cpp:148-155) */
ddj_status ddj_tree_decompress(ddj_ctx* ctx, const void* d_blob, size_t blob_bytes, ddj_buf* out, int* dtype);
/* Partial-column decode: rows [row_begin, row_begin + row_count) of the
 * blob's column (the thesis' random-access use of FL streams,
 * afl_gpu.cu:113-121, :197-223, carried through the tree).  none leaves,
 * afl -> none and delta -> afl -> none (integer) decode only what the range
 * needs (delta sums the packed deltas before the range without writing
 * them); scale maps its child's rows; other nodes decode whole and are
 * sliced.  Rows past the column's end return DDJ_EINVAL. */
ddj_status ddj_tree_decompress_rows(ddj_ctx* ctx, const void* d_blob, size_t blob_bytes, size_t row_begin,
                                   size_t row_count, ddj_buf* out, int* dtype);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  Decompresses a blob whose tree is none, afl -> none or
 * delta -> afl -> none over an integer column into the caller's d_out
 * (out_cap bytes, aligned to the row type) without allocating: the delta
 * decode's per-tile scratch lives in a fixed 1 MiB device array per GPU, so
 * a delta column holds at most 268,431,361 rows (longer ones: DDJ_EINVAL).
 * Other trees return DDJ_EUNSUPPORTED (use ddj_tree_decompress).  The
 * kernels are enqueued on the context's stream.  All calls on one GPU share
 * the scratch array: calls that overlap on one GPU (other contexts, other
 * host threads) race and must be serialised by the caller. */
ddj_status ddj_tree_decompress_noalloc(ddj_ctx* ctx, const void* d_blob, size_t blob_bytes, void* d_out,
                                      size_t out_cap, size_t* out_bytes, int* dtype);
/* CompressionTree::GetPredictedSizeAfterCompression (tree/compression_tree.cpp:157-160) */
ddj_status ddj_tree_predict_size(ddj_ctx* ctx, const int* preorder, size_t path_len, int dtype, const void* d_in,
                                 size_t in_bytes, size_t* size);

/* ---------------------------------------------------------------- planner */
enum { DDJ_PLAN_EXHAUSTIVE = 0, DDJ_PLAN_OPTIMIZE30 = 1, DDJ_PLAN_LEARNED = 2 };
/* EXHAUSTIVE: phase 1 of CompressionOptimizer::CompressData (optimizer/compression_optimizer.cpp:156-182);
 * OPTIMIZE30: CompressionOptimizer::OptimizeTree (:16-36).  The chosen tree
 * is written as a pre-order path (fixed, leaves included). */
ddj_status ddj_plan(ddj_ctx* ctx, int mode, int dtype, const void* d_in, size_t in_bytes, int* path_out,
                    size_t path_cap, size_t* path_len, double* ratio, uint64_t* n_candidates);

/* Stateful optimizer: CompressionOptimizer::CompressData over successive
 * parts of one column (re-plans on the first part and every 10th). */
typedef struct ddj_optimizer ddj_optimizer;
ddj_status ddj_optimizer_create(ddj_ctx* ctx, ddj_optimizer** out);
void ddj_optimizer_destroy(ddj_optimizer* opt);
ddj_status ddj_optimizer_compress(ddj_optimizer* opt, int dtype, const void* d_in, size_t in_bytes, ddj_buf* blob);
ddj_status ddj_optimizer_tree(ddj_optimizer* opt, int* path_out, size_t path_cap, size_t* path_len);

/* --------------------------------------------------------- batched trees */
/* Many columns, each with its own tree, compressed together: worker threads
 * (one per worker context, each with its own HIP stream) take columns from a
 * shared queue, so one column's host round trips (size read-backs) overlap
 * other columns' kernels.  Blobs are owned by `ctx` (ddj_buf_free).  No
 * counterpart in the reference, whose column API compresses one column at
 * a time on one worker (parallel_ts_compressor.cpp:31-36). */
typedef struct {
    int dtype;
    const void* d_in;      /* device data */
    size_t in_bytes;
    const int* preorder;   /* the column's tree, as for ddj_tree_compress */
    size_t path_len;
} ddj_column;
ddj_status ddj_batch_compress(ddj_ctx* ctx, const ddj_column* cols, uint32_t n_cols, uint32_t workers /* 0: 4 */,
                              ddj_buf* blobs);
/* The knobs of a batched call.  Columns whose tree is afl -> none or
 * delta -> afl -> none on int32/int64 go through one captured graph of
 * persistent kernels (min/max over work items of 2^range_rows_log2 rows,
 * then FL packing over 4096-value tiles, wg_per_cu workgroups per CU);
 * the others through `workers` worker contexts.  Zero fields take the
 * defaults (4 workers, 2^18 rows, 4 workgroups per CU). */
typedef struct {
    uint32_t workers;          /* 1..16 */
    uint32_t range_rows_log2;  /* 12..24 */
    uint32_t wg_per_cu;        /* 1..16 */
} ddj_batch_config;
ddj_status ddj_batch_compress_config(ddj_ctx* ctx, const ddj_column* cols, uint32_t n_cols,
                                     const ddj_batch_config* config, ddj_buf* blobs);
/* Autotuner: compresses the batch under candidate configurations (worker
 * counts first, then the graph lane's chunk length and workgroups per CU
 * with the best worker count), timing each after an untimed run, and
 * returns the fastest in *best with its time.  The columns are read only. */
ddj_status ddj_batch_autotune(ddj_ctx* ctx, const ddj_column* cols, uint32_t n_cols, ddj_batch_config* best,
                              double* best_ms);

/* -------------------------------------------------------- learned planner */
/* No counterpart in the reference.  A 32-128-128-48 network (ReLU, BF16
 * matrix-core products, FP32 accumulation, kernels in k_mlp.hip) maps one
 * statistics pass over a block to the predicted log2(compressed bytes /
 * input bytes) of 48 candidate trees.  Two column classes, each with its own
 * candidates and weights: integer columns (class 0) and float / double
 * columns (class 1).  Labels are exact serialized sizes from the library's
 * encoders; a candidate a type cannot take (AFL on a double output) is
 * labelled 8 (2^8 times the input) and never chosen. */
enum { DDJ_PLANNER_INT = 0, DDJ_PLANNER_FLOAT = 1 };
typedef struct ddj_planner ddj_planner;
/* The 48 candidate trees as pre-order paths, concatenated; offsets[49]
 * (ddj_planner_candidates: the integer class). */
ddj_status ddj_planner_candidates(int* paths, size_t cap, size_t* offsets);
ddj_status ddj_planner_candidates_class(int cls, int* paths, size_t cap, size_t* offsets);
/* 32 features of one block (one statistics pass). */
ddj_status ddj_planner_features(ddj_ctx* ctx, int dtype, const void* d_in, size_t in_bytes, float* out /* [32] */);
/* log2(exact blob bytes / input bytes) of every candidate tree on a block. */
ddj_status ddj_planner_labels(ddj_ctx* ctx, int dtype, const void* d_in, size_t in_bytes, float* out /* [48] */);
/* ddj_planner_create: the integer class. */
ddj_status ddj_planner_create(ddj_ctx* ctx, uint64_t seed, ddj_planner** out);
ddj_status ddj_planner_create_class(ddj_ctx* ctx, int cls, uint64_t seed, ddj_planner** out);
void ddj_planner_destroy(ddj_planner* planner);
/* Parameters (FP32, W1[32][128] b1[128] W2[128][128] b2[128] W3[128][64] b3[64]):
 * copied out to get_out and/or in from set_in (either may be NULL). */
ddj_status ddj_planner_params(ddj_planner* planner, float* get_out, const float* set_in);
/* Mean squared error over n rows and its gradient (no update). */
ddj_status ddj_planner_gradient(ddj_planner* planner, const float* X, const float* T, size_t n, float* grad_out,
                                float* loss);
/* Adam over `epochs` shuffled passes of minibatches; *loss = final MSE. */
ddj_status ddj_planner_train(ddj_planner* planner, const float* X, const float* T, size_t n, int epochs, int batch,
                             float lr, uint64_t seed, float* loss);
ddj_status ddj_planner_predict(ddj_planner* planner, const float* X, size_t n, float* Y /* [n][48] */);
/* Chooses a tree for a block: the smallest predicted size, or the exhaustive
 * planner's choice when the runner-up is within min_margin (log2 units). */
ddj_status ddj_planner_plan(ddj_planner* planner, int dtype, const void* d_in, size_t in_bytes, float min_margin,
                            int* path_out, size_t path_cap, size_t* path_len, int* used_fallback, float* margin);
/* No counterpart in the reference (its planner is the exhaustive search):
 * for callers that want the learned planner's rate with exact sizes where
 * the network is unsure.  The candidates predicted within `within` (log2
 * units) of the best, at most max_checked of them, are compressed exactly
 * (shared prefixes once) and the smallest is chosen; *checked = how many
 * were compressed (0 when the best prediction stood alone). */
ddj_status ddj_planner_plan_checked(ddj_planner* planner, int dtype, const void* d_in, size_t in_bytes,
                                    float within, int max_checked, int* path_out, size_t path_cap,
                                    size_t* path_len, int* checked);

/* ------------------------------------------------------------- column API */
/* ParallelTSCompressor (parallel_ts_compressor.cpp:17-102): rows of several
 * host columns are cut into batches of `batch_rows`; for every batch, each
 * column in order goes through its own stateful optimizer
 * (CompressionOptimizer::CompressData) and is written as [size_t bytes][blob]
 * (optimizer/compression_task.cpp:35-40).  Successive ddj_ts_compress calls
 * continue the same per-column optimizers (pass append=1 to add to a file). */
typedef struct ddj_ts_compressor ddj_ts_compressor;
ddj_status ddj_ts_create(ddj_ctx* ctx, uint32_t n_columns, const int* dtypes, size_t batch_rows,
                         ddj_ts_compressor** out);
void ddj_ts_destroy(ddj_ts_compressor* ts);
ddj_status ddj_ts_compress(ddj_ts_compressor* ts, const void* const* host_columns, size_t rows, const char* out_path,
                           int append);
/* ParallelTSCompressor::Decompress (:63-100) with DecompressionTask
 * (optimizer/decompression_task.cpp:19-46): one batch per call into
 * caller buffers of `cap_rows` rows per column; *rows = 0 at end of file. */
typedef struct ddj_ts_reader ddj_ts_reader;
ddj_status ddj_ts_reader_open(ddj_ctx* ctx, const char* in_path, uint32_t n_columns, const int* dtypes,
                              ddj_ts_reader** out);
ddj_status ddj_ts_reader_next(ddj_ts_reader* reader, void* const* host_columns, size_t cap_rows, size_t* rows);
void ddj_ts_reader_close(ddj_ts_reader* reader);

/* ------------------------------------------------ time-series record files */
/* The record files the reference's column API and CLI read and write: its
 * File / FileDefinition / TimeSeriesReaderBinary / TimeSeriesReaderCSV come
 * from the data-manager submodule, which is absent from the reference tree
 * (used by main.cpp:59-76, parallel_ts_compressor.cpp:39-102).  Host code
 * only, no context.  Formats:
 *   header  one column per line, "name,type,decimals" (the .header files of sample_data/),
 *           type int | time | float | double | short | char;
 *   binary  records of every column's native little-endian value in order,
 *           then `padding` bytes (main.cpp's -p option);
 *   CSV     one record per line, comma-separated; floats written with the
 *           column's decimals.
 * Errors carry a message in ddj_io_last_error() (per thread). */
enum { DDJ_FORMAT_BIN = 0, DDJ_FORMAT_CSV = 1 };
typedef struct ddj_file_def ddj_file_def;
const char* ddj_io_last_error(void);
/* TimeSeriesReader::ReadFileDefinition (main.cpp:92-94) */
ddj_status ddj_file_def_read(const char* header_path, ddj_file_def** out);
ddj_status ddj_file_def_create(uint32_t n_columns, const char* const* names, const int* dtypes, const int* decimals,
                               ddj_file_def** out);
ddj_status ddj_file_def_write(const ddj_file_def* def, const char* header_path);
uint32_t ddj_file_def_columns(const ddj_file_def* def);
ddj_status ddj_file_def_column(const ddj_file_def* def, uint32_t i, const char** name, int* dtype, int* decimals);
void ddj_file_def_free(ddj_file_def* def);
/* TimeSeriesReader::Read(file, batch) / Write(file, ts) (parallel_ts_compressor.cpp:46, :93):
 * up to max_rows records per call into one host array per column; *rows = 0
 * at end of file. */
typedef struct ddj_series_file ddj_series_file;
ddj_status ddj_series_open(const ddj_file_def* def, const char* path, int format, int padding, int write,
                           ddj_series_file** out);
ddj_status ddj_series_read(ddj_series_file* file, size_t max_rows, void* const* host_columns, size_t* rows);
ddj_status ddj_series_write(ddj_series_file* file, const void* const* host_columns, size_t rows);
void ddj_series_close(ddj_series_file* file);

#ifdef __cplusplus
}
#endif
#endif
