// The compression cascade: node records, blob assembly, parse and decode.
#include "tree.h"

#include <cstring>
#include <functional>
#include <sstream>

namespace ddj {

namespace {

const char* enc_name(int e) {
    switch (e) {
    case DDJ_ENC_DELTA: return "delta";
    case DDJ_ENC_AFL: return "afl";
    case DDJ_ENC_SCALE: return "scale";
    case DDJ_ENC_RLE: return "rle";
    case DDJ_ENC_DICT: return "dict";
    case DDJ_ENC_UNIQUE: return "unique";
    case DDJ_ENC_PATCH: return "patch";
    case DDJ_ENC_NONE: return "none";
    case DDJ_ENC_FLOAT_TO_INT: return "floatToInt";
    case DDJ_ENC_CONST: return "constData";
    case DDJ_ENC_GFC: return "gfc";
    case DDJ_ENC_OUTSIDE_PATCH: return "patch(outside)";
    default: return "?";
    }
}

// Encoding::GetCompressionRatio (compression/encoding.hpp:78-84)
double ratio_of(size_t in, size_t out) {
    if (out == 0 || out > in) return 1.0;
    return (double)in / (double)out;
}

// One piece of a serialized blob: an 8-byte header (host) or a part.  A
// keep-alive piece holds a node's input until the blob is assembled
// (encoders may hand out views into their input, e.g. delta's first row)
// and contributes no bytes.
struct Piece {
    bool is_header = false;
    bool keep_alive = false;
    uint8_t hdr[8];
    Part part;
};

bool is_int_type(int dt) { return dt == DDJ_D_INT || dt == DDJ_D_TIME || dt == DDJ_D_SHORT || dt == DDJ_D_CHAR; }

// The node's single output goes to the blob unchanged: no child, or one
// none leaf.
bool output_is_leaf(const TNode& node) {
    return node.ch.empty() || (node.ch.size() == 1 && node.ch[0]->enc == DDJ_ENC_NONE && node.ch[0]->ch.empty());
}

// DELTA -> AFL (-> none) on an integer column: the deltas feed the FL pass
// straight from the column (delta_afl_encode) and are never written.
bool fusable_delta_afl(const TNode& node, const Part& data) {
    if (node.enc != DDJ_ENC_DELTA || !is_int_type(node.dtype) || node.ch.size() != 1) return false;
    const TNode& a = *node.ch[0];
    if (a.enc != DDJ_ENC_AFL || a.dtype != node.dtype || !output_is_leaf(a)) return false;
    return data.n / dtype_size(node.dtype) >= 2;
}

// PATCH -> (AFL -> none), (AFL -> none) on an integer column: the split
// FL-packs both sides straight into the blob (patch_afl_encode).
bool fusable_patch_afl(const TNode& node, const Part& data) {
    if ((node.enc != DDJ_ENC_PATCH && node.enc != DDJ_ENC_OUTSIDE_PATCH) || !is_int_type(node.dtype) ||
        node.ch.size() != 2)
        return false;
    for (const NodeP& a : node.ch)
        if (a->enc != DDJ_ENC_AFL || a->dtype != node.dtype || !output_is_leaf(*a)) return false;
    return data.n / dtype_size(node.dtype) >= 1;
}

// CompressionNode::Compress (compression_node.cpp:103-138) and PrepareMetadata (:84-101).
// `pre` holds this node's encoder outputs when a parent already produced them.
size_t compress_node(Ctx* c, TNode& node, const Part& data, std::vector<Piece>& out,
                     const std::vector<Part>* pre = nullptr) {
    std::vector<Part> res;
    std::vector<std::vector<Part>> child_pre;  // per child: its outputs when this node produced them
    if (data.owner) {
        Piece k;
        k.keep_alive = true;
        k.part = data;
        out.push_back(k);
    }
    if (pre) {
        res = *pre;
    } else if (node.enc == DDJ_ENC_NONE) {
        size_t len = data.n;  // the leaf refers to its input instead of copying it
        res.push_back(host_part(&len, sizeof len));
        res.push_back(data);
    } else if (fusable_delta_afl(node, data)) {
        auto f = delta_afl_encode(c, node.dtype, data.view(), true);
        const size_t es = dtype_size(node.dtype);
        Part deltas;  // never materialised: only its size is read
        deltas.n = (data.n / es - 1) * es;
        res = {f[0], deltas};
        child_pre = {{f[1], f[2]}};
    } else if (fusable_patch_afl(node, data)) {
        auto f = patch_afl_encode(c, node.dtype, data.view(), node.enc == DDJ_ENC_OUTSIDE_PATCH);
        res = {f.stencil, f.side_in[0], f.side_in[1]};
        child_pre = {{f.side_meta[0], f.side_data[0]}, {f.side_meta[1], f.side_data[1]}};
    } else {
        // an integer AFL whose data goes to the blob unchanged packs it there
        // when the blob is assembled; a dictionary split whose two outputs
        // are stored unchanged runs its writing pass then
        auto none_leaf = [](const NodeP& ch) { return ch->enc == DDJ_ENC_NONE && ch->ch.empty(); };
        // so does a float GFC payload stored unchanged
        const bool gfc_leaf = node.enc == DDJ_ENC_GFC && !is_int_type(node.dtype) &&
                              (node.ch.empty() || (node.ch.size() == 2 && none_leaf(node.ch[1])));
        const bool defer = (node.enc == DDJ_ENC_AFL && is_int_type(node.dtype) && output_is_leaf(node)) || gfc_leaf;
        const bool defer_split =
            node.enc == DDJ_ENC_DICT && node.ch.size() == 2 && none_leaf(node.ch[0]) && none_leaf(node.ch[1]);
        res = encode(c, node.enc, node.dtype, data.view(), data.minmax ? data.minmax->get<MinMaxResult>() : nullptr,
                     defer, data.minmax_host.get(), defer_split);
    }
    Piece h;
    h.is_header = true;
    const int16_t e = (int16_t)node.enc, d = (int16_t)node.dtype;
    const int32_t ml = (int32_t)res[0].n;
    memcpy(h.hdr, &e, 2);
    memcpy(h.hdr + 2, &d, 2);
    memcpy(h.hdr + 4, &ml, 4);
    size_t total = 8 + res[0].n;
    out.push_back(h);
    Piece m;
    m.part = res[0];
    out.push_back(m);
    if (node.ch.empty()) {
        Piece p;
        p.part = res.size() > 1 ? res[1] : Part{};
        total += p.part.n;
        out.push_back(p);
    } else {
        for (size_t i = 0; i < node.ch.size(); i++) {
            if (i + 1 >= res.size()) fail(DDJ_EINVAL, "tree node has more children than its encoding has outputs");
            total += compress_node(c, *node.ch[i], res[i + 1], out, i < child_pre.size() ? &child_pre[i] : nullptr);
        }
    }
    node.ratio = ratio_of(data.n, total);
    return total;
}

}  // namespace

NodeP make_node(int enc, int dtype) {
    auto n = std::make_shared<TNode>();
    n->enc = enc;
    n->dtype = dtype;
    return n;
}

NodeP copy_tree(const NodeP& n) {
    auto r = std::make_shared<TNode>();
    r->enc = n->enc;
    r->dtype = n->dtype;
    r->ratio = n->ratio;
    for (auto& c : n->ch) r->ch.push_back(copy_tree(c));
    return r;
}

NodeP tree_from_path(const int* path, size_t len, int dtype) {
    if (len == 0) fail(DDJ_EINVAL, "empty tree path");
    size_t pos = 0;
    std::function<NodeP(int, int)> make = [&](int dt, int depth) -> NodeP {
        if (depth > kMaxTreeDepth) fail(DDJ_EINVAL, "tree path too deep");
        const int et = path[pos];
        if (!enc_ok(et)) fail(DDJ_EINVAL, "bad encoding in tree path");
        NodeP node = make_node(et, dt);
        if (et == DDJ_ENC_NONE) return node;
        int rt[2];
        return_types(et, dt, rt);
        const int k = n_results(et);
        for (int i = 0; i < k; i++) {
            if (pos + 1 >= len) break;  // path ended: the node stays a leaf
            pos++;
            node->ch.push_back(make(rt[i], depth + 1));
        }
        return node;
    };
    NodeP root = make(dtype, 0);
    if (pos + 1 != len) fail(DDJ_EINVAL, "tree path longer than the tree it describes");
    return root;
}

void fix_tree(const NodeP& n) {
    if (n->ch.empty()) {
        if (n->enc == DDJ_ENC_NONE) return;
        const int k = n_results(n->enc);
        for (int i = 0; i < k; i++) {
            NodeP leaf = make_node(DDJ_ENC_NONE, n->dtype);
            leaf->ratio = 1.0;
            n->ch.push_back(leaf);
        }
    } else {
        for (auto& c : n->ch) fix_tree(c);
    }
}

std::vector<int> preorder(const NodeP& n) {
    std::vector<int> out{n->enc};
    for (auto& c : n->ch) {
        auto s = preorder(c);
        out.insert(out.end(), s.begin(), s.end());
    }
    return out;
}

std::string tree_string(const NodeP& n) {
    // CompressionNode::Print (compression_node.cpp:180-186)
    std::ostringstream os;
    std::function<void(const NodeP&)> rec = [&](const NodeP& x) {
        os << enc_name(x->enc) << "[" << x->ratio << "],";
        for (auto& c : x->ch) rec(c);
    };
    rec(n);
    return os.str();
}

Part compress_tree(Ctx* c, const NodeP& root, DView data) {
    std::vector<Piece> pieces;
    size_t total = compress_node(c, *root, borrow(data), pieces);
    // Concatenate (util/copy/cuda_array_copy.cu:14-33): deferred FL packs
    // write straight into their blob ranges, everything else moves in one
    // gather kernel.  The copy table, the node headers and the host-resident
    // metadata form one table, read by the kernel from the pinned staging
    // ring (or, without a device address for it, uploaded in one copy):
    // their ranges point into the table itself.
    DevBuf blob(c, total);
    size_t n_ranges = 0, inline_bytes = 0;
    for (auto& p : pieces) {
        if (p.keep_alive) continue;
        if (p.is_header)
            inline_bytes += 8;
        else if (p.part.on_host())
            inline_bytes += p.part.n;
        if (p.is_header || (p.part.n && !p.part.fl && !p.part.split)) n_ranges++;
    }
    const size_t table_bytes = sizeof(CopyRange) * n_ranges + inline_bytes;
    std::vector<uint8_t> table(table_bytes);
    std::vector<size_t> inline_ranges;  // their src holds the offset in the inline bytes until placed
    CopyRange* ranges = reinterpret_cast<CopyRange*>(table.data());
    uint8_t* inl = table.data() + sizeof(CopyRange) * n_ranges;
    size_t off = 0, io = 0, ri = 0;
    uint64_t max_bytes = 0, copied = 0;
    struct JobAt {
        const FlJob* job;
        uint64_t off, bytes;
    };
    std::vector<JobAt> jobs;
    std::vector<std::pair<const SplitJob*, uint64_t>> splits;
    for (auto& p : pieces) {
        if (p.keep_alive) continue;
        if (!p.is_header && p.part.split) {  // rows written in place by the split (also when there are none)
            splits.emplace_back(p.part.split.get(), off);
            off += p.part.n;
            continue;
        }
        if (p.is_header) {
            memcpy(inl + io, p.hdr, 8);
            inline_ranges.push_back(ri);
            ranges[ri++] = CopyRange{reinterpret_cast<const uint8_t*>(io), off, 8};
            io += 8;
            off += 8;
            copied += 8;
        } else if (p.part.n) {
            const size_t len = p.part.n;
            if (p.part.fl) {  // packed in place below
                jobs.push_back(JobAt{p.part.fl.get(), off, len});
            } else if (p.part.on_host()) {
                memcpy(inl + io, p.part.host->data(), len);
                inline_ranges.push_back(ri);
                ranges[ri++] = CopyRange{reinterpret_cast<const uint8_t*>(io), off, len};
                io += len;
            } else {
                ranges[ri++] = CopyRange{p.part.p, off, len};
            }
            if (!p.part.fl) {
                max_bytes = std::max<uint64_t>(max_bytes, len);
                copied += len;
            }
            off += len;
        }
    }
    if (off != total || ri != n_ranges) fail(DDJ_EHIP, "blob size mismatch");
    // deferred splits first: they also fill outputs that the gather copies
    for (const auto& sj : splits) run_split_job(c, *sj.first, blob.bytes_ptr() + sj.second);
    // the table's device address: a slot of the staging ring, taken now (no
    // staged copy may follow before the launch), else a device buffer
    DevBuf rdev;
    const uint8_t* tdev = staging_slot(c, table_bytes);
    if (!tdev) {
        rdev = DevBuf(c, table_bytes);
        tdev = rdev.bytes_ptr();
    }
    for (size_t k : inline_ranges)
        ranges[k].src = tdev + sizeof(CopyRange) * n_ranges + reinterpret_cast<uintptr_t>(ranges[k].src);
    if (rdev.get())
        h2d(c, rdev.get(), table.data(), table.size());
    else
        memcpy(c->staging + (tdev - c->staging_dev), table.data(), table.size());
    kernel_bytes(c, "k_gather_copy", 2.0 * (double)copied);
    launch_gather_copy(c, reinterpret_cast<const CopyRange*>(tdev), (int)n_ranges, blob.bytes_ptr(),
                       std::max<uint64_t>(max_bytes, 8));
    // the FL streams go straight to their blob ranges (any byte offset)
    for (const JobAt& j : jobs) run_fl_job(c, *j.job, blob.bytes_ptr() + j.off, j.bytes);
    return own(std::move(blob));
}

namespace {

// A parsed blob: the node table read back from k_parse_tree, each node's
// children, and while alive the host copies of every node's leading metadata
// (and leaf data) bytes registered as shadows, from which decoders read
// their small metadata without a device round trip.
struct ParsedBlob {
    Ctx* c;
    DView blob;
    std::vector<ParsedNode> nodes;
    std::vector<std::vector<int>> kids;
    size_t shadows_before = 0;

    ParsedBlob(Ctx* cc, DView b) : c(cc), blob(b) {
        if (blob.n < 8) fail(DDJ_ECORRUPT, "blob shorter than one node header");
        // status words and node table in one buffer, read back with one copy
        // (trees have a few nodes; 64 table entries cover every planner tree)
        constexpr int kHead = 64;  // bytes before the table
        DevBuf parsed(c, kHead + sizeof(ParsedNode) * kMaxTreeNodes);
        ParsedNode* table = reinterpret_cast<ParsedNode*>(parsed.bytes_ptr() + kHead);
        launch_parse_tree(c, blob.p, blob.n, table, kMaxTreeNodes, parsed.get<int>());
        constexpr int kFirst = 64;
        std::vector<uint8_t> host(kHead + sizeof(ParsedNode) * kFirst);
        d2h(c, host.data(), parsed.get(), host.size());
        int st[3];
        memcpy(st, host.data(), sizeof st);
        if (st[0] != 0) fail(DDJ_ECORRUPT, "malformed tree blob (parse status " + std::to_string(st[0]) + ")");
        const int count = st[1];
        if (count <= 0 || count > kMaxTreeNodes) fail(DDJ_ECORRUPT, "malformed tree blob (node count)");
        nodes.resize((size_t)count);
        if (count <= kFirst)
            memcpy(nodes.data(), host.data() + kHead, sizeof(ParsedNode) * count);
        else
            d2h(c, nodes.data(), table, sizeof(ParsedNode) * count);
        kids.resize((size_t)count);
        for (int i = 1; i < count; i++) kids[(size_t)nodes[i].parent].push_back(i);
        shadows_before = c->shadows.size();
        for (int i = 0; i < count; i++) {
            const ParsedNode& nd = nodes[(size_t)i];
            c->shadows.push_back(
                {blob.p + nd.meta_off, (size_t)std::min<uint64_t>(nd.meta_len, kMetaHead), nd.meta_head});
            if (nd.enc == DDJ_ENC_NONE)
                c->shadows.push_back(
                    {blob.p + nd.data_off, (size_t)std::min<uint64_t>(nd.data_len, kMetaHead), nd.data_head});
        }
    }
    ~ParsedBlob() { c->shadows.resize(shadows_before); }
    ParsedBlob(const ParsedBlob&) = delete;
    ParsedBlob& operator=(const ParsedBlob&) = delete;

    DView meta(int i) const { return DView{blob.p + nodes[(size_t)i].meta_off, (size_t)nodes[(size_t)i].meta_len}; }
    DView leaf_data(int i) const {
        return DView{blob.p + nodes[(size_t)i].data_off, (size_t)nodes[(size_t)i].data_len};
    }
    bool is_leaf(int i) const { return nodes[(size_t)i].enc == DDJ_ENC_NONE && kids[(size_t)i].empty(); }
    // The integer node i is DELTA over AFL over a none leaf: the AFL node.
    int delta_afl_child(int i) const {
        const ParsedNode& nd = nodes[(size_t)i];
        if (nd.enc != DDJ_ENC_DELTA || !is_int_type(nd.dtype) || nd.nchild != 1 || kids[(size_t)i].size() != 1)
            return -1;
        const int a = kids[(size_t)i][0];
        const ParsedNode& an = nodes[(size_t)a];
        if (an.enc != DDJ_ENC_AFL || an.dtype != nd.dtype || an.nchild != 1 || kids[(size_t)a].size() != 1) return -1;
        return is_leaf(kids[(size_t)a][0]) ? a : -1;
    }

    // The integer node i is a patch whose two children are AFL over a none
    // leaf: whether it is.
    bool patch_afl_kids(int i) const {
        const ParsedNode& nd = nodes[(size_t)i];
        if ((nd.enc != DDJ_ENC_PATCH && nd.enc != DDJ_ENC_OUTSIDE_PATCH) || !is_int_type(nd.dtype) || nd.nchild != 2 ||
            kids[(size_t)i].size() != 2)
            return false;
        for (int a : kids[(size_t)i]) {
            const ParsedNode& an = nodes[(size_t)a];
            if (an.enc != DDJ_ENC_AFL || an.dtype != nd.dtype || an.nchild != 1 || kids[(size_t)a].size() != 1 ||
                !is_leaf(kids[(size_t)a][0]))
                return false;
        }
        return true;
    }

    // CompressionNode::Decompress (compression_node.cpp:140-161): children first.
    Part decode_node(int i) const {
        const ParsedNode& nd = nodes[(size_t)i];
        const DView m = meta(i);
        if (nd.enc == DDJ_ENC_NONE) return borrow(leaf_data(i));
        if ((int)kids[(size_t)i].size() != nd.nchild) fail(DDJ_ECORRUPT, "node child count mismatch");
        // PATCH over two AFL -> none (integer): merge straight from both FL streams
        if (patch_afl_kids(i)) {
            const int y = kids[(size_t)i][0], z = kids[(size_t)i][1];
            return patch_afl_decode(c, nd.dtype, m, meta(y), leaf_data(kids[(size_t)y][0]), meta(z),
                                    leaf_data(kids[(size_t)z][0]));
        }
        // DELTA over AFL over none (integer): unpack and prefix-sum in one pass
        const int a = delta_afl_child(i);
        if (a >= 0) {
            const int leaf = kids[(size_t)a][0];
            if (nodes[(size_t)leaf].data_len > 0 && m.n >= dtype_size(nd.dtype))
                return delta_afl_decode(c, nd.dtype, m, meta(a), leaf_data(leaf));
        }
        std::vector<Part> parts;
        std::vector<DView> in{m};
        for (int k : kids[(size_t)i]) {
            parts.push_back(decode_node(k));
            in.push_back(parts.back().view());
        }
        return decode(c, nd.enc, nd.dtype, in);
    }

    // Rows [begin, begin + count) of node i's decoded output.  Leaves, FL
    // streams and delta -> FL columns decode only those rows; scale maps
    // them from its child's; any other node decodes whole and is sliced.
    Part decode_rows(int i, size_t begin, size_t count) const {
        const ParsedNode& nd = nodes[(size_t)i];
        const size_t es = dtype_size(nd.dtype);
        auto check = [&](size_t total_rows) {
            if (begin > total_rows || count > total_rows - begin) fail(DDJ_EINVAL, "rows beyond the end of the column");
        };
        if (nd.enc == DDJ_ENC_NONE) {
            const DView d = leaf_data(i);
            check(d.n / es);
            return borrow(DView{d.p + begin * es, count * es});
        }
        if ((int)kids[(size_t)i].size() != nd.nchild) fail(DDJ_ECORRUPT, "node child count mismatch");
        if (nd.enc == DDJ_ENC_AFL && is_int_type(nd.dtype) && nd.nchild == 1 && is_leaf(kids[(size_t)i][0]))
            return afl_decode_rows(c, nd.dtype, meta(i), leaf_data(kids[(size_t)i][0]), begin, count);
        const int a = delta_afl_child(i);
        if (a >= 0 && meta(i).n >= es)
            return delta_afl_decode_rows(c, nd.dtype, meta(i), meta(a), leaf_data(kids[(size_t)a][0]), begin, count);
        if (nd.enc == DDJ_ENC_SCALE && is_int_type(nd.dtype) && nd.nchild == 1 && meta(i).n > 0) {
            Part ch = decode_rows(kids[(size_t)i][0], begin, count);
            if (ch.n == 0) return ch;
            return decode(c, DDJ_ENC_SCALE, nd.dtype, {meta(i), ch.view()});
        }
        Part whole = decode_node(i);
        check(whole.n / es);
        return borrow(DView{whole.p + begin * es, count * es}, whole.owner);
    }
};

// A part the caller may own: borrowed bytes (a view into the blob, or into
// a larger decoded buffer) are copied into a buffer of their own size.
Part owned_copy(Ctx* c, Part r, bool exact) {
    if (r.owner && !exact) return r;
    if (r.owner && r.owner->get() == r.p && r.owner->size() == r.n) return r;
    DevBuf b(c, r.n);
    if (r.n) d2d(c, b.get(), r.p, r.n);
    return own(std::move(b));
}

}  // namespace

Part decompress_tree(Ctx* c, DView blob, int* dtype_out, NodeP* tree_out) {
    ParsedBlob pb(c, blob);
    const int count = (int)pb.nodes.size();
    if (tree_out) {
        std::vector<NodeP> ns((size_t)count);
        for (int i = 0; i < count; i++) ns[i] = make_node(pb.nodes[i].enc, pb.nodes[i].dtype);
        for (int i = 1; i < count; i++) ns[(size_t)pb.nodes[i].parent]->ch.push_back(ns[i]);
        *tree_out = ns[0];
    }
    if (dtype_out) *dtype_out = pb.nodes[0].dtype;
    Part r = pb.decode_node(0);
    if (pb.nodes[0].enc == DDJ_ENC_NONE) r = owned_copy(c, r, true);  // a bare leaf: hand out a copy
    return r;
}

Part decompress_tree_rows(Ctx* c, DView blob, size_t begin, size_t count, int* dtype_out) {
    ParsedBlob pb(c, blob);
    if (dtype_out) *dtype_out = pb.nodes[0].dtype;
    return owned_copy(c, pb.decode_rows(0, begin, count), true);
}

size_t predict_tree(Ctx* c, const NodeP& n, DView data) {
    size_t meta = 0, dsz = 0;
    predict_sizes(c, n->enc, n->dtype, data, &meta, &dsz);
    size_t size = meta + 8;
    if (n->ch.empty()) return size + dsz;
    auto outs = encode(c, n->enc, n->dtype, data);
    for (size_t i = 0; i < n->ch.size(); i++) size += predict_tree(c, n->ch[i], outs[i + 1].view());
    return size;
}

}  // namespace ddj
