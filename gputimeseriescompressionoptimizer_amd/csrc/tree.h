// CompressionNode / CompressionTree: the cascade, its pre-order serialized
// format and size prediction (reference tree/compression_node.cpp,
// tree/compression_tree.cpp).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "codecs.h"

namespace ddj {

constexpr int kMaxTreeDepth = 64;
constexpr int kMaxTreeNodes = 1024;

struct TNode;
using NodeP = std::shared_ptr<TNode>;

struct TNode {
    int enc = DDJ_ENC_NONE;
    int dtype = DDJ_D_INT;
    std::vector<NodeP> ch;
    double ratio = 1.0;  // CompressionNode::_compressionRatio
};

NodeP make_node(int enc, int dtype);
NodeP copy_tree(const NodeP& n);
// PathGenerator::GenerateTree / CreateTree (optimizer/path_generator.cpp:70-91);
// a path that ends on a non-none node leaves it a leaf.
NodeP tree_from_path(const int* path, size_t len, int dtype);
// CompressionNode::Fix (compression_node.cpp:209-223)
void fix_tree(const NodeP& n);
std::vector<int> preorder(const NodeP& n);
std::string tree_string(const NodeP& n);

// CompressionTree::Compress (compression_tree.cpp:98-103): updates the
// nodes' ratios and returns the blob.
Part compress_tree(Ctx* c, const NodeP& root, DView data);
// CompressionTree::Decompress (compression_tree.cpp:148-155)
Part decompress_tree(Ctx* c, DView blob, int* dtype_out, NodeP* tree_out = nullptr);
// Partial-column decode: rows [begin, begin + count) of the blob's column
// (random access, afl_gpu.cu:113-121, 197-223, extended to trees).
Part decompress_tree_rows(Ctx* c, DView blob, size_t begin, size_t count, int* dtype_out);
// CompressionNode::PredictCompressionSize (compression_node.cpp:163-177)
size_t predict_tree(Ctx* c, const NodeP& n, DView data);

constexpr int kMetaHead = 32;  // leading metadata bytes copied into the parse table
struct ParsedNode {
    int32_t enc, dtype, nchild, parent;
    uint64_t meta_off, meta_len, data_off, data_len;
    uint8_t meta_head[kMetaHead];  // metadata bytes [0, min(meta_len, kMetaHead))
    uint8_t data_head[kMetaHead];  // leaves: data bytes [0, min(data_len, kMetaHead))
};
void launch_parse_tree(Ctx* c, const uint8_t* blob, uint64_t bytes, ParsedNode* d_out, int max_nodes, int* d_status);

// Decompression of none, afl -> none and delta -> afl -> none blobs (integer
// columns) into caller memory with no device allocation (noalloc.hip).
void tree_decompress_noalloc(Ctx* c, const uint8_t* blob, size_t blob_bytes, void* out, size_t out_cap,
                             size_t* out_bytes, int* dtype_out);
// Takes the device's address of that decoder's scratch (once per device).
hipError_t noalloc_init_device(int device);

}  // namespace ddj
