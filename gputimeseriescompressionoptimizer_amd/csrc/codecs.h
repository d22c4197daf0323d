// Encodings: Encode / Decode / size predictors of every reference encoding,
// orchestrated on the host, computed by the HIP kernels.
#pragma once

#include <functional>
#include <memory>
#include <vector>

#include "ddj_internal.h"
#include "kernels.h"

namespace ddj {

struct FlJob;
struct SplitJob;

// A device byte range that keeps its storage alive.
// Device bytes, or a few host bytes (metadata computed on the host, which
// blob assembly uploads together with the node headers instead of one copy
// per node).  Only encoder metadata (out[0]) is ever host-resident.
// A part with `fl` set holds no bytes yet: its FL pack runs when the blob
// is assembled, straight into the blob (p is null, n is the exact size).
struct Part {
    std::shared_ptr<DevBuf> owner;  // may be null for caller-owned memory
    const uint8_t* p = nullptr;
    size_t n = 0;
    std::shared_ptr<const std::vector<uint8_t>> host;  // set: the bytes live here, p is null
    // Optional: a MinMaxResult (device) of the part's values, computed by the
    // encoder that produced them, so that an FL child skips its own pass.
    std::shared_ptr<DevBuf> minmax;
    std::shared_ptr<const MinMaxResult> minmax_host;  // the same keys on the host, when read back already
    std::shared_ptr<const FlJob> fl;
    // set: rows a deferred split writes straight into the blob (p is null,
    // n is the exact size); the split also fills the other outputs of its
    // encoder, so blob assembly runs it before gathering those
    std::shared_ptr<const SplitJob> split;
    bool on_host() const { return host != nullptr; }
    DView view() const { return DView{p, n}; }
};
// An FL pack whose bit width is known (its min/max pass and read-back ran)
// and whose output goes straight into a blob being assembled; or another
// stream of known size written the same way (`write`).
struct FlJob {
    int dtype = 0;
    bool delta = false;  // value i is x[i+1] - x[i]; x holds n + 1 rows
    int bits = 0;
    size_t n = 0;  // values packed
    Part x;        // the rows, kept alive until the pack runs
    // set: writes the stream itself (a GFC payload) instead of an FL pack
    std::function<void(Ctx*, uint8_t*, size_t)> write;
};
// Runs the pack into the `bytes` stream bytes at out (any byte address).
void run_fl_job(Ctx* c, const FlJob& j, void* out, size_t bytes);

// A dictionary split whose counting pass ran (the yes count is known): its
// writing pass fills the stencil and code buffers (parts of its encoder's
// output) and writes the other rows at the blob offset assembly gives it.
struct SplitJob {
    SplitPlan plan;
    Part x, stencil, freq;  // kept alive until the split runs
    uint8_t* codes = nullptr;
    std::shared_ptr<DevBuf> count;
};
void run_split_job(Ctx* c, const SplitJob& j, void* no_rows);

Part own(DevBuf&& b);
Part borrow(DView v, std::shared_ptr<DevBuf> keep = nullptr);
Part host_part(const void* h, size_t n);

int n_results(int enc);
// GetReturnTypes (each encoding's .hpp)
void return_types(int enc, int dtype, int out[2]);
bool enc_ok(int enc);

// Encoding::Encode: out[0] = metadata, out[1..n_results(enc)] = outputs.
// `minmax` (optional, device): the input's min/max keys, when known.
// `defer_fl`: an integer AFL encoder returns its data as an FlJob part, a
// float GFC encoder its payload (for a blob being assembled, when that
// output goes to the blob unchanged).
std::vector<Part> encode(Ctx* c, int enc, int dtype, DView in, const MinMaxResult* minmax = nullptr,
                         bool defer_fl = false, const MinMaxResult* minmax_host = nullptr, bool defer_split = false);
// Encoding::Decode: in[0] = metadata, in[1..] = children outputs.
Part decode(Ctx* c, int enc, int dtype, const std::vector<DView>& in);
// GetMetadataSize / GetCompressedSize.
void predict_sizes(Ctx* c, int enc, int dtype, DView in, size_t* meta, size_t* data);
// CudaArrayStatistics::GenerateStatistics plus planner features.
ddj_stats_t stats(Ctx* c, int dtype, DView in);

// DELTA followed by AFL on an integer column of >= 2 rows, without writing
// the deltas: {delta metadata, AFL metadata, AFL data}, byte-identical to
// delta_encode followed by the AFL encoder on its output.
std::vector<Part> delta_afl_encode(Ctx* c, int dtype, DView in, bool defer_fl = false);
// The inverse: AFL data (non-empty) decoded and prefix-summed in one pass.
Part delta_afl_decode(Ctx* c, int dtype, DView delta_meta, DView afl_meta, DView afl_data);

// Patch (lower, or outside with `outside`) whose two outputs each go to an
// AFL node stored unchanged, on an integer column of >= 1 row: the patch
// metadata (the stencil) and both FL data streams are FlJob parts written
// together by one fused split when the blob is assembled; each side's AFL
// metadata is host-resident and its input size (rows * sizeof(T)) is the
// size of side_in (no bytes).
struct PatchAflParts {
    Part stencil;
    Part side_in[2];
    Part side_meta[2];
    Part side_data[2];
};
PatchAflParts patch_afl_encode(Ctx* c, int dtype, DView in, bool outside);

// Rows [begin, begin + count) of what an integer AFL stream (metadata,
// data) decodes to, unpacking only those rows.
// patch -> (afl -> none), (afl -> none) decoded in one merge over both FL
// streams (the patch's stencil, each AFL node's metadata and leaf data).
Part patch_afl_decode(Ctx* c, int dtype, DView stencil, DView ymeta, DView ydata, DView nmeta, DView ndata);
Part afl_decode_rows(Ctx* c, int dtype, DView meta, DView data, size_t begin, size_t count);
// The same for a column stored as delta -> AFL: the deltas before the range
// are summed (per tile, never written), the range unpacked and scanned.
Part delta_afl_decode_rows(Ctx* c, int dtype, DView dmeta, DView ameta, DView data, size_t begin, size_t count);

}  // namespace ddj
