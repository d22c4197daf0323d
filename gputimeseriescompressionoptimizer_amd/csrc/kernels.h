// Host-side launchers of the HIP kernels.  Every launcher enqueues work on
// ctx->stream and returns without synchronising unless stated otherwise.
#pragma once

#include <memory>

#include "ddj_internal.h"

namespace ddj {

// Tile offsets of a variable-length output, computed ahead of the output
// pass: a counting kernel fills vals[0..tiles) (one value per tile), scan()
// launches k_scan_tiles, which writes prefix[t] = vals[0] + ... + vals[t-1]
// for t <= tiles (prefix[tiles] is the total).  Sums wrap at 64 bits.
struct MinMaxResult;
struct TileScan {
    DevBuf mem;
    uint64_t* vals = nullptr;
    uint64_t* prefix = nullptr;
    size_t tiles = 0;
    TileScan(Ctx* c, size_t tiles);
    // keys (optional): per-tile {min, max, min, max} order keys of two
    // outputs, folded in the same launch into *mm_a and *mm_b.
    void scan(Ctx* c, const uint64_t* keys = nullptr, MinMaxResult* mm_a = nullptr, MinMaxResult* mm_b = nullptr);
};
// k_scan_tiles over caller memory: prefix[t] for t <= tiles.
void launch_scan_tiles(Ctx* c, const uint64_t* vals, size_t tiles, uint64_t* prefix);
// The same for arrays of any length: one workgroup up to 16K values, else
// block sums, their scan and a per-block rescan (three launches).
void launch_scan_large(Ctx* c, const uint64_t* vals, size_t count, uint64_t* prefix);

// Standalone unique encoding (k_unique.hip): the number of listed values U;
// the metadata [size_t U][size_t n][U values] and the code words of n rows
// (aligned x); and the decode of n rows from the listed values (any
// address) and 4-byte aligned code words.
uint64_t unique_count(Ctx* c, int dtype, const void* x, size_t n);
void unique_encode_dev(Ctx* c, int dtype, const void* x, size_t n, DevBuf& meta, DevBuf& data);
void unique_decode_dev(Ctx* c, int dtype, const uint8_t* uvals, uint64_t U, uint64_t n, const uint8_t* words,
                       size_t nwords, void* out);
// Stable LSD radix sort (8-bit digits) of n unsigned keys whose low
// key_bits bits vary; keys and tmp (n keys each) are used in turn; returns
// the one holding the sorted keys.
template <typename K> K* radix_sort(Ctx* c, K* keys, K* tmp, size_t n, int key_bits);

// Kernels that can raise the device error flag (encoded input failing a
// device-side check) arm it first; the call then reads and clears it.
void arm_error_flag(Ctx* c);

// Scratch of a single-pass reduction (per-block partials + result).
struct MinMaxResult {
    uint64_t min_key;  // order_key of the minimum
    uint64_t max_key;  // order_key of the maximum
};

// The floatToInt encoder's inputs for a float / double column: min/max keys
// and the float precision (maximum over the rows, as DevStats::precision).
struct FloatScanResult {
    MinMaxResult mm;
    uint64_t precision;
};
void launch_float_scan(Ctx* c, int dtype, const void* x, size_t n, FloatScanResult* d_result);

// min/max of n values of `dtype` into d_result (device).  One launch.
void launch_minmax(Ctx* c, int dtype, const void* x, size_t n, MinMaxResult* d_result,
                   const uint64_t* d_rows = nullptr /* device row count (<= n) read by the kernel, or null */);
// min/max of the wrapping differences x[i+1] - x[i] of n integer rows.
void launch_minmax_delta(Ctx* c, int elem_size, const void* x, size_t n, MinMaxResult* d_result);
// Integer columns whose min/max the host needs: the reduction's partials
// are read back and folded on the host (no fold launch).  delta: of the
// differences x[i+1] - x[i].
MinMaxResult minmax_to_host(Ctx* c, int dtype, const void* x, size_t n, bool delta);
// The reduction alone (integer columns, rows from *d_rows when given):
// `partials` receives 3 order keys per workgroup, of which the returned
// count; fold_minmax_partials folds them (stride 3; 2 for delta) on the host.
int launch_minmax_partials(Ctx* c, int dtype, const void* x, size_t n, const uint64_t* d_rows, DevBuf& partials);
MinMaxResult fold_minmax_partials(const uint64_t* h, int grid, int stride);
// FL pack of the m = n - 1 differences of n integer rows (bits from mm).
// With mm null the width is `bits`; no byte at or past out + out_bytes is
// written (0: room for the widest b, rounded up to whole words).
void launch_fl_pack_delta(Ctx* c, int elem_size, const void* x, size_t m, const MinMaxResult* mm, void* out,
                          int bits = 0, size_t out_bytes = 0);
// Converts order keys back to the typed min/max (host).
void decode_minmax(int dtype, const MinMaxResult& r, void* mn, void* mx);

// Device statistics record filled by launch_stats.
struct DevStats {
    uint64_t min_key, max_key;
    uint64_t first_zero;   // floats: row << 1 | sign of the first zero (~0: none)
    uint64_t unsorted;     // any x[i] > x[i+1]
    uint64_t rl_eq;        // #{i < n-2 : int(x[i]) == x[i+1]}
    uint64_t runs;         // #{i : i == 0 || x[i] != x[i-1]}
    uint64_t sum_bits;     // integer types: wrapping sum
    double sum_f;          // float types: sum (order not pinned)
    uint32_t precision;    // max _getFloatPrecision
    uint32_t delta_bits_max;
    uint32_t bit_hist[65];
    uint32_t hll[64];
};
void launch_stats(Ctx* c, int dtype, const void* x, size_t n, DevStats* d_out);

// FL pack: bits from `mm` (device min/max) when mm is set, otherwise `bits`.
// Writes the ceil(n*b/W)*sizeof(T) stream bytes to out (any byte address),
// and zeros after them up to the end of the last 32-bit word, but no byte at
// or past out + out_bytes (0: room for the widest b, rounded up to whole
// words).
void launch_fl_pack(Ctx* c, int dtype, const void* x, size_t n, const MinMaxResult* mm, int bits, void* out,
                    size_t out_bytes = 0);
// FL unpack of m b-bit deltas fused with the delta decode: out[0] = first
// (device, one value), out[i+1] = out[i] + d[i] (wrapping).  The per-tile
// sums and prefixes go to tile_scratch (fl_unpack_delta_scratch_words(m)
// 64-bit words, every one written before it is read) or, when it is null,
// to memory from the context's allocator.
size_t fl_unpack_delta_scratch_words(size_t m);
void launch_fl_unpack_delta(Ctx* c, int elem_size, const uint8_t* first, const uint8_t* in, size_t in_bytes, size_t m,
                           int bits, void* out, uint64_t* tile_scratch = nullptr);
// FL unpack of the n values [first_row, first_row + n) of a b-bit stream at
// an arbitrary byte address.
void launch_fl_unpack(Ctx* c, int dtype, const uint8_t* in, size_t in_bytes, size_t n, int bits, void* out,
                      size_t first_row = 0);
// Rows [begin, begin + count) of the column that delta + FL encoded as
// `first` and a stream of b-bit deltas (out[0] = first + d[0] + ... +
// d[begin - 1]); reads the stream up to the range's end, twice.
void launch_fl_unpack_delta_rows(Ctx* c, int elem_size, const uint8_t* first, const uint8_t* in, size_t in_bytes,
                                 int bits, size_t begin, size_t count, void* out);
// Random access (afl_decompress_value_gpu, compression/afl/afl_gpu.cu:113-121,
// 197-223): out[i] = value rows[i] of an n-value b-bit stream in `lanes`
// layout (1: contiguous; 32: interleaved blocks of 32 lanes x W values).  A
// row >= n reads as 0 and raises the device error flag with code 3.
void launch_fl_fetch(Ctx* c, int dtype, int bits, int lanes, const uint8_t* in, size_t in_bytes, size_t n,
                     const uint64_t* rows, size_t count, void* out);
// The interleaved layout (run_afl_compress_gpu<T,32> /
// run_afl_decompress_gpu<T,32>, afl_gpu.cu:67-111): stream bytes for n
// values, and pack / unpack.
size_t fl_lanes_bytes(int dtype, int bits, size_t n);
void launch_fl_pack_lanes(Ctx* c, int dtype, int bits, const void* x, size_t n, void* out);
void launch_fl_unpack_lanes(Ctx* c, int dtype, int bits, const void* in, size_t n, void* out);

// Elementwise kernels.
void launch_delta_encode(Ctx* c, int elem_size, const void* x, size_t n, void* out /* n-1 */);
void launch_delta_decode(Ctx* c, int elem_size, const void* first /* device, 1 elem */, const uint8_t* d,
                         size_t m /* deltas */, void* out /* m+1 */);
void launch_add_scalar(Ctx* c, int elem_size, const void* x, size_t n, uint64_t s, void* out);  // wrapping
void launch_sub_scalar(Ctx* c, int elem_size, const void* x, size_t n, uint64_t s, void* out);
void launch_float_split(Ctx* c, const float* x, size_t n, int32_t* mant, int32_t* expo, int32_t* sign);
void launch_float_compose(Ctx* c, const int32_t* mant, const int32_t* expo, const uint8_t* stencil_bits,
                          int sign_const, size_t n, float* out);
void launch_precision(Ctx* c, int dtype, const void* x, size_t n, uint32_t* d_max);
void launch_float_to_int(Ctx* c, int dtype, const void* x, size_t n, int precision, int32_t* out);
void launch_int_to_float(Ctx* c, int dtype, const uint8_t* x, size_t n, int precision, void* out);

// Stable partition by a predicate (patch/const/dict) with the packed
// stencil.  counts[0] = #yes (device).
enum PredKind { PRED_LOWER = 0, PRED_INSIDE = 1, PRED_NOT_EQUAL = 2, PRED_IN_SET = 3 };
struct PredArgs {
    int kind;
    uint64_t a_bits, b_bits;  // thresholds / constant, as T bit patterns
    uint64_t set_bits[4];     // PRED_IN_SET keys
    int set_n;
};
// mm_yes / mm_no (integer types, optional): the order keys of each output's
// minimum and maximum, computed by the counting pass.
// codes (dictionary split, 4-byte aligned): the 2-bit codes packed 16 to a
// little-endian 32-bit word; every word holding a code is written (room for
// ceil(n / 16) words).
// `no` may sit at any byte address (rows stored unchanged into a blob).
void launch_split(Ctx* c, int dtype, const void* x, size_t n, const PredArgs& p, uint8_t* stencil /* 1+ceil(n/8) */,
                  void* yes, void* no, uint8_t* codes /* packed dict codes or null */, uint64_t* d_counts,
                  MinMaxResult* mm_yes = nullptr, MinMaxResult* mm_no = nullptr);
// The same in two steps: the counting pass and scan (the yes count is then
// prefix[tiles] of the plan's scan, on the device), and the writing pass.
struct SplitPlan {
    int dtype = 0;
    const void* x = nullptr;
    size_t n = 0;
    PredArgs p{};
    std::shared_ptr<TileScan> ts;
};
SplitPlan launch_split_count(Ctx* c, int dtype, const void* x, size_t n, const PredArgs& p,
                             MinMaxResult* mm_yes = nullptr, MinMaxResult* mm_no = nullptr);
void launch_split_write(Ctx* c, const SplitPlan& plan, uint8_t* stencil, void* yes, uint8_t* no, uint8_t* codes,
                        uint64_t* d_counts);
// One output of a patch split FL-encoded straight into its stream: `bytes`
// stream bytes at dst (any byte address), values of `bits` bits.
// An FL stream read by a merge: its bytes and bit width.
struct FlSrc {
    size_t bytes;
    int bits;
};
struct FlSide {
    uint8_t* dst;
    uint64_t bytes;
    int bits;
    uint32_t* edges;  // set by the launcher
};
// The writing pass of a patch split whose two outputs are FL leaves: the
// stencil, and both sides packed into their streams (k_split_fl, then
// k_fl_edges for the words neighbouring tiles share).
void launch_split_fl(Ctx* c, const SplitPlan& plan, uint8_t* stencil, FlSide yes, FlSide no, uint64_t* d_counts);
// Merge: out[i] = bit i ? yes[k++] : no[j++] (const_no: `no_const` for every
// clear bit).  yes / no may sit at any byte address (outputs stored
// unchanged in a blob); n_yes / n_no rows are available; a stencil asking
// for more raises the device error flag (corrupt input).
void launch_merge(Ctx* c, int dtype, const uint8_t* stencil_bits /* packed bits, no header */, size_t n,
                  const uint8_t* yes, size_t n_yes, const uint8_t* no, size_t n_no, bool const_no, uint64_t no_const,
                  void* out);
// Merge whose two sides are FL streams (n_yes rows of `by` bits in
// yes_bytes at yes, n_no rows of `bn` bits at no; any byte addresses):
// each tile's rows of either stream are unpacked in LDS.  Integer columns.
void launch_merge_fl(Ctx* c, int dtype, const uint8_t* stencil_bits, size_t n, const uint8_t* yes, size_t yes_bytes,
                     size_t n_yes, int by, const uint8_t* no, size_t no_bytes, size_t n_no, int bn, void* out);
// Merge whose set-bit rows are dict[code j] for the m `code_bits`-wide codes
// in `words` (the dictionary decoder without materialising those rows).
void launch_merge_codes(Ctx* c, int dtype, const uint8_t* stencil_bits, size_t n, const uint8_t* words, size_t m,
                        int code_bits, const uint8_t* dict, uint32_t K, const uint8_t* no, size_t n_no, void* out);

// RLE.
// Returns the run count (one read-back, which also brings the min/max of the
// values and of the lengths into *h_values / *h_lengths when those are set;
// values: integer types only).
uint64_t launch_rle_encode(Ctx* c, int dtype, const void* x, size_t n, int32_t* lengths, void* values, uint64_t* d_runs,
                           MinMaxResult* h_values = nullptr, MinMaxResult* h_lengths = nullptr);
DevBuf rle_decode_dev(Ctx* c, int dtype, const uint8_t* lengths, const uint8_t* values, size_t runs);

// Top-K most frequent keys (count desc, key asc), K <= 4.
void topk_most_frequent(Ctx* c, int dtype, const void* x, size_t n, int K, uint64_t* keys_bits_host,
                        int* n_distinct_host);

// GFC (float: 32-bit words, double: 64-bit words).
// A GFC encode whose size pass ran: every chunk's place in the payload and
// the payload's size are known, the chunks are not written yet.
struct GfcPlan {
    bool dbl = false;
    const void* x = nullptr;
    size_t n = 0;
    long long N = 0, per = 0;
    size_t chunks = 0;
    std::shared_ptr<TileScan> scan;  // prefix[chunk]: the chunk's payload offset
    uint64_t payload_bytes = 0;
};
GfcPlan gfc_plan(Ctx* c, bool dbl, const void* x, size_t n, DevBuf& meta, DevBuf& offsets);
// Writes the planned chunks at payload (payload_bytes bytes, any byte address).
void gfc_write(Ctx* c, const GfcPlan& p, uint8_t* payload);
// gfc_plan, then gfc_write into a payload buffer of its own.
void gfc_encode(Ctx* c, bool dbl, const void* x, size_t n, DevBuf& meta, DevBuf& offsets, DevBuf& payload);
void gfc_decode(Ctx* c, bool dbl, const uint8_t* meta, size_t meta_bytes, const uint8_t* offsets, size_t off_bytes,
                const uint8_t* payload, size_t payload_bytes, DevBuf& out);

// Planner network (32 -> 128 -> 128 -> 48, ReLU), FP32 parameters.
constexpr int kMlpFeatures = 32, kMlpHidden = 128, kMlpOutputs = 48, kMlpOutputsPadded = 64;
struct MlpParams {
    float* W1;  // [32][128]
    float* b1;  // [128]
    float* W2;  // [128][128]
    float* b2;  // [128]
    float* W3;  // [128][64], columns 48..63 zero
    float* b3;  // [64]
};
constexpr size_t kMlpParamCount = 32 * 128 + 128 + 128 * 128 + 128 + 128 * 64 + 64;
inline MlpParams mlp_params_at(float* base) {
    MlpParams p;
    p.W1 = base;
    p.b1 = p.W1 + 32 * 128;
    p.W2 = p.b1 + 128;
    p.b2 = p.W2 + 128 * 128;
    p.W3 = p.b2 + 128;
    p.b3 = p.W3 + 128 * 64;
    return p;
}
void launch_mlp_forward(Ctx* c, const float* X, int B, const MlpParams& p, float* Y, uint16_t* H1s, uint16_t* H2s);
void launch_mlp_backward(Ctx* c, const float* X, const uint16_t* H1s, const uint16_t* H2s, const float* Y,
                         const float* T, int B, const MlpParams& p, const MlpParams& g);
void launch_adam(Ctx* c, float* w, float* grad, float* m, float* v, int n, float lr, int step);

// Batched FL-leaf trees (afl -> none, delta -> afl -> none; int32 / int64):
// one descriptor per column, work items of (column, row range) for the
// min/max pass and (column, FL tile) for the pack pass.
struct BatchColumn {
    const uint8_t* x;  // 16-byte aligned column
    uint64_t n;        // rows
    uint32_t es;       // 4 or 8
    uint32_t delta;    // 1: delta -> afl -> none
    uint32_t dtype;
    uint32_t bits;     // written by the layout pass
    uint64_t min_key, max_key;
    uint64_t out_off, blob_bytes, data_off, data_bytes;  // written by the layout pass
};
struct BatchItem {
    uint32_t col;
    uint32_t pad;
    uint64_t start, end;  // row range, or (tile, unused) for the pack pass
};
constexpr uint64_t kFlTileRows = 4096;  // values per FL pack tile
// The four launches (init, min/max, layout, pack); no host round trip, so
// the sequence can be captured into a graph.
void launch_batch_fl(Ctx* c, BatchColumn* d_cols, int ncol, const BatchItem* d_ranges, int nranges,
                     const BatchItem* d_tiles, int ntiles, uint8_t* arena, unsigned wg_per_cu);

// Zero-fill of device memory by a kernel (graph-capture safe).
void launch_fill_zero(Ctx* c, void* p, size_t bytes);
// Gather of several device ranges into one buffer (tree blob assembly).
struct CopyRange {
    const uint8_t* src;
    uint64_t dst_off;
    uint64_t bytes;
};
void launch_gather_copy(Ctx* c, const CopyRange* d_ranges, int n_ranges, uint8_t* dst, uint64_t max_bytes);
// Copies of several device ranges, each to its own destination.
struct CopyPair {
    const uint8_t* src;
    uint8_t* dst;
    uint64_t bytes;
};
void launch_copy_pairs(Ctx* c, const CopyPair* d_pairs, int n_pairs, uint64_t max_bytes);
// Up to kToHostPieces device ranges copied by one launch into pinned host
// memory that the device addresses directly (the context's read-back area).
constexpr int kToHostPieces = 8;
struct ToHostPieces {
    const uint8_t* src[kToHostPieces];
    uint32_t off[kToHostPieces];
    uint32_t bytes[kToHostPieces];
    int count;
};
// After the data, the launch stores `seq` at flag_dev (host memory too), so
// the host can poll for completion instead of waiting on the stream.
void launch_to_host(Ctx* c, const ToHostPieces& t, uint8_t* host_dev, uint32_t* flag_dev, uint32_t seq);

}  // namespace ddj
