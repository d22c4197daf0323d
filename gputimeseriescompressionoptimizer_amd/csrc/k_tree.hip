// Device-side parse of a serialized compression tree (reference
// DecompressNodes, tree/compression_tree.cpp:105-146): one thread walks the
// pre-order node records and writes a table of (encoding, type, metadata and
// leaf-data ranges), so the host needs a single read-back instead of one per
// node header.
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

__global__ void k_parse_tree(const uint8_t* blob, uint64_t bytes, ParsedNode* out, int max_nodes, int* status) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
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
}

}  // namespace

void launch_parse_tree(Ctx* c, const uint8_t* blob, uint64_t bytes, ParsedNode* d_out, int max_nodes, int* d_status) {
    { KernelTimer kt_(c, "k_parse_tree"); hipLaunchKernelGGL(k_parse_tree, dim3(1), dim3(kWave), 0, c->stream, blob, bytes, d_out, max_nodes, d_status); }
    HIPC(hipGetLastError());
}

}  // namespace ddj
