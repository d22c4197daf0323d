// The statistics-driven planner.
#include "planner.h"

#include <deque>

namespace ddj {

namespace {

double ratio_of(size_t in, size_t out) {
    if (out == 0 || out > in) return 1.0;
    return (double)in / (double)out;
}

constexpr int kMaxLevel = 3;  // path_generator.cpp:15

}  // namespace

std::vector<int> get_continuations(int et, int dt, const ddj_stats_t& st, int level) {
    std::vector<int> r;
    if (level != 0 && (et == DDJ_ENC_NONE || st.size <= 0)) return r;
    r.push_back(DDJ_ENC_NONE);
    if (et == DDJ_ENC_AFL || et == DDJ_ENC_GFC) return r;
    if (level < kMaxLevel) {
        if (et != DDJ_ENC_DELTA) r.push_back(DDJ_ENC_DELTA);
        if (et != DDJ_ENC_SCALE && st.min != 0) r.push_back(DDJ_ENC_SCALE);
        if (et != DDJ_ENC_FLOAT_TO_INT && dt == DDJ_D_FLOAT) r.push_back(DDJ_ENC_FLOAT_TO_INT);
        if (et != DDJ_ENC_PATCH && et != DDJ_ENC_DICT) r.push_back(DDJ_ENC_PATCH);
        if (et != DDJ_ENC_DICT && et != DDJ_ENC_PATCH) r.push_back(DDJ_ENC_DICT);
        if (st.sorted || st.rlMetric >= 1.5f) {
            if (et != DDJ_ENC_CONST) r.push_back(DDJ_ENC_CONST);
            if (et != DDJ_ENC_RLE) r.push_back(DDJ_ENC_RLE);
        }
    }
    r.push_back((dt == DDJ_D_FLOAT || dt == DDJ_D_DOUBLE) ? DDJ_ENC_GFC : DDJ_ENC_AFL);
    return r;
}

std::vector<std::vector<int>> generate_paths() {
    const int D = DDJ_ENC_DELTA, F = DDJ_ENC_FLOAT_TO_INT, S = DDJ_ENC_SCALE, A = DDJ_ENC_AFL, N = DDJ_ENC_NONE;
    std::vector<std::vector<int>> first{{}, {F}, {D}, {S}, {D, F}, {S, F}};
    std::vector<std::vector<int>> second{{DDJ_ENC_DICT, A, N, A, N},
                                         {DDJ_ENC_CONST, A, N},
                                         {DDJ_ENC_RLE, A, N, A, N},
                                         {A, N},
                                         {DDJ_ENC_PATCH, A, N, A, N}};
    std::vector<std::vector<int>> out;
    for (auto& a : first)
        for (auto& b : second) {
            auto t = a;
            t.insert(t.end(), b.begin(), b.end());
            out.push_back(t);
        }
    return out;
}

std::vector<Candidate> full_statistics_update(Ctx* c, const Part& data, int et, int dt, int level) {
    const ddj_stats_t st = stats(c, dt, data.view());
    std::vector<Candidate> result, part1;
    for (int cont : get_continuations(et, dt, st, level)) {
        std::vector<Part> compr;
        if (cont == DDJ_ENC_NONE) {
            size_t len = data.n;
            compr.push_back(host_part(&len, sizeof len));
            compr.push_back(data);
        } else {
            compr = encode(c, cont, dt, data.view(), data.minmax ? data.minmax->get<MinMaxResult>() : nullptr, false,
                           data.minmax_host.get());
        }
        size_t psize = 8;
        for (auto& p : compr) psize += p.n;
        NodeP parent = make_node(cont, dt);
        parent->ratio = ratio_of(data.n, psize);
        int rt[2];
        return_types(cont, dt, rt);
        const int k = n_results(cont);
        if (k == 1) {
            auto kids = full_statistics_update(c, compr[1], cont, rt[0], level + 1);
            part1.clear();
            for (auto& ch : kids) {
                const size_t out = ch.size + compr[0].n + 8;
                NodeP t = make_node(cont, dt);
                t->ch.push_back(copy_tree(ch.tree));
                t->ratio = ratio_of(data.n, out);
                part1.push_back(Candidate{t, out});
            }
        } else if (k == 2) {
            auto left = full_statistics_update(c, compr[1], cont, rt[0], level + 1);
            auto right = full_statistics_update(c, compr[2], cont, rt[1], level + 1);
            part1.clear();
            for (auto& l : left)
                for (auto& r : right) {
                    const size_t out = l.size + r.size + compr[0].n + 8;
                    NodeP t = make_node(cont, dt);
                    t->ch.push_back(copy_tree(l.tree));
                    t->ch.push_back(copy_tree(r.tree));
                    t->ratio = ratio_of(data.n, out);
                    part1.push_back(Candidate{t, out});
                }
        }
        // The reference keeps part1 across iterations (compression_optimizer.cpp:96-137);
        // `none` comes first, so it only ever sees an empty list here.
        if (part1.empty()) part1.push_back(Candidate{parent, psize});
        result.insert(result.end(), part1.begin(), part1.end());
    }
    return result;
}

NodeP plan_exhaustive(Ctx* c, DView data, int dt, size_t* best_size, uint64_t* n_candidates) {
    // GetSampleDataForFullUpdateSize (compression_optimizer.cpp:148-154)
    const size_t es = dtype_size(dt);
    const size_t rows = data.n / es;
    const size_t sample = rows <= 10000 ? data.n : 10000 * es;
    DevBuf s(c, sample);
    d2d(c, s.get(), data.p, sample);
    Part sp = own(std::move(s));
    auto cands = full_statistics_update(c, sp, DDJ_ENC_NONE, dt, 0);
    if (n_candidates) *n_candidates = cands.size();
    NodeP best;
    size_t bsize = 0;
    double bratio = -1;
    for (auto& cd : cands) {
        fix_tree(cd.tree);
        if (cd.tree->ratio > bratio) {  // std::max_element keeps the first maximum
            bratio = cd.tree->ratio;
            best = cd.tree;
            bsize = cd.size;
        }
    }
    if (best_size) *best_size = bsize;
    return best;
}

NodeP optimize_tree(Ctx* c, DView data, int dt) {
    NodeP best = make_node(DDJ_ENC_NONE, dt);
    size_t min_size = data.n;
    for (auto& path : generate_paths()) {
        NodeP t = tree_from_path(path.data(), path.size(), dt);
        const size_t s = predict_tree(c, t, data);
        if (s < min_size) {
            min_size = s;
            best = t;
        }
    }
    return best;
}

// ------------------------------------------------------------ edge stats
EdgeStats::EdgeStats(int height) : stat_((size_t)((1 << (height + 1)) - 2)) {}

double* EdgeStats::find(int no, std::pair<int, int> e, bool create) {
    if (no < 0 || no >= (int)stat_.size()) fail(DDJ_EINVAL, "edge number out of range");
    for (auto& kv : stat_[(size_t)no])
        if (kv.first == e) return &kv.second;
    if (!create) return nullptr;
    stat_[(size_t)no].push_back({e, 0.0});
    return &stat_[(size_t)no].back().second;
}

void EdgeStats::update(int no, std::pair<int, int> e, double ratio) {
    double* v = find(no, e, true);  // compression_statistics.cpp:29-37
    if (*v < 1.0) *v = 1.0;
    if (ratio < 1.0) ratio = 1.0;
    *v = (*v + ratio) / 2.0;
}

void EdgeStats::set(int no, std::pair<int, int> e, double v) { *find(no, e, true) = v; }

double EdgeStats::get(int no, std::pair<int, int> e) { return *find(no, e, true); }

std::pair<std::pair<int, int>, double> EdgeStats::get_best(int no, std::pair<int, int> old, bool same) {
    std::pair<int, int> be{old.first, DDJ_ENC_NONE};  // compression_statistics.cpp:52-71
    double bv = 1;
    if (no < 0 || no >= (int)stat_.size()) return {be, bv};
    for (auto& kv : stat_[(size_t)no]) {
        if (same && kv.first.first != old.first) continue;
        if (kv.first.second != old.second && kv.second > bv) {
            be = kv.first;
            bv = kv.second;
        }
    }
    return {be, bv};
}

std::vector<Edge> get_edges(const NodeP& root, int max_height) {
    std::vector<Edge> out;
    const int maxno = (1 << (max_height + 1)) - 2;
    TNode fake;
    std::deque<TNode*> fifo{root.get()};
    int no = 0;
    while (!fifo.empty() && no < maxno) {
        TNode* el = fifo.front();
        fifo.pop_front();
        if (el->ch.size() >= 1) {
            out.push_back(Edge{el, el->ch[0].get(), no});
            fifo.push_back(el->ch[0].get());
        } else {
            fifo.push_back(&fake);
        }
        no++;
        if (el->ch.size() >= 2) {
            out.push_back(Edge{el, el->ch[1].get(), no});
            fifo.push_back(el->ch[1].get());
        } else {
            fifo.push_back(&fake);
        }
        no++;
    }
    return out;
}

void update_statistics(const NodeP& root, EdgeStats& st) {
    for (auto& e : get_edges(root, 5)) st.update(e.no, {e.from->enc, e.to->enc}, (e.from->ratio + e.to->ratio) / 2.0);
}

// ------------------------------------------------------------- optimizer
Part Optimizer::compress_data(DView part, int dt) {
    // IsFullUpdateNeeded (compression_optimizer.cpp:142-146)
    if (!tree_ || parts_ % 10 == 9) {
        const size_t es = dtype_size(dt);
        const size_t rows = part.n / es;
        const size_t sample = rows <= 10000 ? part.n : 10000 * es;
        DevBuf s(c_, sample);
        d2d(c_, s.get(), part.p, sample);
        auto cands = full_statistics_update(c_, own(std::move(s)), DDJ_ENC_NONE, dt, 0);
        NodeP best;
        double bratio = -1;
        for (auto& cd : cands) {
            fix_tree(cd.tree);
            update_statistics(cd.tree, stats_);
            if (cd.tree->ratio > bratio) {
                bratio = cd.tree->ratio;
                best = cd.tree;
            }
        }
        tree_ = best;
        old_ratio_ = best->ratio;  // OptimalTree::OptimalTree (optimal_tree.hpp:26-33)
        old_stats_ = stats_;
        aliased_ = false;
    }
    Part out = compress_tree(c_, tree_, part);
    update_statistics(tree_, stats_);  // CompressionTree::Compress with statistics set
    try_correct_tree();
    parts_++;
    return out;
}

// OptimalTree::Replace (optimal_tree.cpp:14-37)
void Optimizer::replace(TNode* node, EdgeStats& stats, int edge_no) {
    if (edge_no > max_edge_no_) return;
    const int k = n_results(node->enc);
    int rt[2];
    return_types(node->enc, node->dtype, rt);
    for (int i = 0; i < k; i++) {
        const int t = node->enc;
        auto best = stats.get_best(edge_no + i, {t, t});
        if (best.second > 1) {
            NodeP ch = make_node(best.first.second, rt[i]);
            node->ch.push_back(ch);
            replace(ch.get(), stats, 2 * (edge_no + i) + 2);
        } else {
            node->ch.push_back(make_node(DDJ_ENC_NONE, rt[i]));
        }
    }
}

// OptimalTree::TryCorrectTree (optimal_tree.cpp:39-93)
bool Optimizer::try_correct_tree() {
    const double old_ratio = old_ratio_, new_ratio = tree_->ratio;
    if (old_ratio <= new_ratio && old_ratio > 1) return false;
    auto edges = get_edges(tree_, 5);
    for (auto& e : edges) {
        const std::pair<int, int> type{e.from->enc, e.to->enc};
        EdgeStats& olds = aliased_ ? stats_ : old_stats_;
        const double old_value = olds.get(e.no, type);
        const double new_value = stats_.get(e.no, type);
        auto best = stats_.get_best(e.no, type);
        if (new_value < old_value || best.second > new_value) {
            auto same = stats_.get_best(e.no, type, true);
            if (same.second > new_value) {
                TNode* to = e.to;  // CompressionNode::Reset (compression_node.cpp:225-240)
                to->ch.clear();
                to->ratio = 1.0;
                to->enc = same.first.second;
                replace(to, stats_, 2 * (e.no + 1));
            } else {
                TNode* from = e.from;
                from->ch.clear();
                from->ratio = 1.0;
                from->enc = best.first.first;
                replace(from, stats_, e.no % 2 ? e.no - 1 : e.no);
            }
            old_ratio_ = tree_->ratio;
            aliased_ = true;  // the reference now aliases the live statistics (optimal_tree.cpp:88)
            return true;
        }
    }
    return false;
}

}  // namespace ddj
