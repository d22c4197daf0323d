// CompressionOptimizer: statistics-driven candidate enumeration (phase 1),
// the 30-path brute force, the per-edge ratio statistics (phase 2) and tree
// repair (phase 3).  Reference: optimizer/path_generator.cpp,
// optimizer/compression_optimizer.cpp, optimizer/optimal_tree.cpp,
// tree/compression_statistics.cpp.
#pragma once

#include <map>
#include <utility>
#include <vector>

#include "tree.h"

namespace ddj {

// PathGenerator::GetContinuations (path_generator.cpp:93-146)
std::vector<int> get_continuations(int et, int dt, const ddj_stats_t& st, int level);
// PathGenerator::GeneratePaths (path_generator.cpp:20-68)
std::vector<std::vector<int>> generate_paths();

struct Candidate {
    NodeP tree;
    size_t size;
};
// CompressionOptimizer::FullStatisticsUpdate (compression_optimizer.cpp:88-140)
std::vector<Candidate> full_statistics_update(Ctx* c, const Part& data, int et, int dt, int level);
// Phase 1: candidates on the first 10,000 rows, best = first largest root ratio.
NodeP plan_exhaustive(Ctx* c, DView data, int dt, size_t* best_size, uint64_t* n_candidates);
// CompressionOptimizer::OptimizeTree (compression_optimizer.cpp:16-36)
NodeP optimize_tree(Ctx* c, DView data, int dt);

// CompressionStatistics: per edge slot a map (from, to) -> running ratio.
// Slots keep insertion order (the reference iterates an unordered_map).
class EdgeStats {
public:
    explicit EdgeStats(int height = 5);
    void update(int no, std::pair<int, int> e, double ratio);
    void set(int no, std::pair<int, int> e, double v);
    double get(int no, std::pair<int, int> e);
    std::pair<std::pair<int, int>, double> get_best(int no, std::pair<int, int> old, bool same_beginning = false);
    int slots() const { return (int)stat_.size(); }

private:
    std::vector<std::vector<std::pair<std::pair<int, int>, double>>> stat_;
    double* find(int no, std::pair<int, int> e, bool create);
};

struct Edge {
    TNode* from;
    TNode* to;
    int no;
};
// CompressionTree::GetEdges (compression_tree.cpp:201-230)
std::vector<Edge> get_edges(const NodeP& root, int max_height = 5);
void update_statistics(const NodeP& root, EdgeStats& st);

// CompressionOptimizer::CompressData with OptimalTree (optimal_tree.cpp)
class Optimizer {
public:
    explicit Optimizer(Ctx* c) : c_(c) {}
    Part compress_data(DView part, int dt);
    NodeP tree() const { return tree_; }

private:
    bool try_correct_tree();
    void replace(TNode* node, EdgeStats& stats, int edge_no);
    Ctx* c_;
    NodeP tree_;
    EdgeStats stats_;      // the optimizer's statistics, shared with the tree
    EdgeStats old_stats_;  // OptimalTree::_statistics (a copy at selection / last repair)
    bool aliased_ = false;  // after a repair the reference's OptimalTree shares the live statistics
    double old_ratio_ = 0;
    size_t parts_ = 0;
    int max_edge_no_ = (1 << 6) - 2;
};

}  // namespace ddj
