// Device-side parse of a serialized compression tree (reference
// DecompressNodes, tree/compression_tree.cpp:105-146): one thread walks the
// pre-order node records and writes a table of (encoding, type, metadata and
// leaf-data ranges, plus each node's first metadata bytes), so the host needs
// a single read-back instead of one per node header or small metadata.
#include "k_common.h"
#include "tree.h"

namespace ddj {

namespace {

__device__ int dev_n_results(int enc) {
    switch (enc) {
    case DDJ_ENC_NONE: return 0;
    case DDJ_ENC_DELTA:
    case DDJ_ENC_SCALE:
    case DDJ_ENC_AFL:
    case DDJ_ENC_UNIQUE:
    case DDJ_ENC_CONST:
    case DDJ_ENC_FLOAT_TO_INT: return 1;
    case DDJ_ENC_RLE:
    case DDJ_ENC_DICT:
    case DDJ_ENC_PATCH:
    case DDJ_ENC_OUTSIDE_PATCH:
    case DDJ_ENC_GFC: return 2;
    default: return -1;
    }
}

__device__ uint64_t rd_le(const uint8_t* p, int n) {
    uint64_t v = 0;
    for (int i = n - 1; i >= 0; i--) v = (v << 8) | p[i];
    return v;
}

__device__ int walk_tree(const uint8_t* blob, uint64_t bytes, ParsedNode* out, int max_nodes, int* status);

// Thread 0 walks the records; then every thread copies one byte of the
// nodes' metadata / leaf-data heads (kMetaHead bytes each).
__global__ __launch_bounds__(kBlock) void k_parse_tree(const uint8_t* blob, uint64_t bytes, ParsedNode* out,
                                                       int max_nodes, int* status) {
    __shared__ int s_count;
    if (threadIdx.x == 0) s_count = walk_tree(blob, bytes, out, max_nodes, status);
    __syncthreads();
    const int count = s_count;
    for (int i = threadIdx.x; i < count * 2 * kMetaHead; i += kBlock) {
        ParsedNode& nd = out[i / (2 * kMetaHead)];
        const int k = i % (2 * kMetaHead);
        if (k < kMetaHead)
            nd.meta_head[k] = (uint64_t)k < nd.meta_len ? blob[nd.meta_off + k] : 0;
        else
            nd.data_head[k - kMetaHead] = (uint64_t)(k - kMetaHead) < nd.data_len ? blob[nd.data_off + k - kMetaHead] : 0;
    }
}

__device__ int walk_tree(const uint8_t* blob, uint64_t bytes, ParsedNode* out, int max_nodes, int* status) {
    int stack_node[kMaxTreeDepth];
    int stack_left[kMaxTreeDepth];
    int sp = 0, count = 0;
    uint64_t pos = 0;
    int st = 0;
    do {
        if (count >= max_nodes) { st = 1; break; }
        if (pos + 8 > bytes) { st = 2; break; }
        const int enc = (int16_t)rd_le(blob + pos, 2);
        const int dt = (int16_t)rd_le(blob + pos + 2, 2);
        const int64_t mlen = (int32_t)rd_le(blob + pos + 4, 4);
        pos += 8;
        const int nres = dev_n_results(enc);
        if (nres < 0 || dt < 0 || dt > DDJ_D_CHAR || mlen < 0 || pos + (uint64_t)mlen > bytes) { st = 3; break; }
        ParsedNode& nd = out[count];
        nd.enc = enc;
        nd.dtype = dt;
        nd.nchild = nres;
        nd.parent = sp > 0 ? stack_node[sp - 1] : -1;
        nd.meta_off = pos;
        nd.meta_len = (uint64_t)mlen;
        nd.data_off = 0;
        nd.data_len = 0;
        pos += (uint64_t)mlen;
        if (enc == DDJ_ENC_NONE) {
            if (mlen < 8) { st = 3; break; }
            const uint64_t dlen = rd_le(blob + nd.meta_off, 8);
            if (pos + dlen > bytes) { st = 4; break; }
            nd.data_off = pos;
            nd.data_len = dlen;
            pos += dlen;
        }
        const int me = count++;
        if (sp > 0) stack_left[sp - 1]--;
        if (nres > 0) {
            if (sp >= kMaxTreeDepth) { st = 5; break; }
            stack_node[sp] = me;
            stack_left[sp] = nres;
            sp++;
        }
        while (sp > 0 && stack_left[sp - 1] == 0) sp--;
    } while (sp > 0);
    status[0] = st;
    status[1] = count;
    status[2] = (int)(pos & 0x7fffffff);
    return count;
}

}  // namespace

void launch_parse_tree(Ctx* c, const uint8_t* blob, uint64_t bytes, ParsedNode* d_out, int max_nodes, int* d_status) {
    { KernelTimer kt_(c, "k_parse_tree"); hipLaunchKernelGGL(k_parse_tree, dim3(1), dim3(kBlock), 0, c->stream, blob, bytes, d_out, max_nodes, d_status); }
    HIPC(hipGetLastError());
}

}  // namespace ddj
