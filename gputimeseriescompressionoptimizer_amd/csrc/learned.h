// Learned planner: a 32-128-128-48 network over one pass of block
// statistics predicts log2(compressed bytes / input bytes) for each of 48
// candidate trees; the planner picks the smallest prediction and falls back
// to the exhaustive search when the margin to the runner-up is small.
// Training labels are exact serialized sizes computed by the library's own
// encoders (the same sizes the exhaustive planner compares).
#pragma once

#include <array>
#include <vector>

#include "kernels.h"
#include "tree.h"

namespace ddj {

constexpr int kCandidates = kMlpOutputs;

// Two networks of the same shape: one for integer columns, one for float
// and double columns, each with its own 48 candidate trees.
constexpr int kClassInt = 0, kClassFloat = 1;
int planner_class(int dtype);
// The candidate trees of a class (pre-order paths).
const std::vector<std::vector<int>>& candidate_paths(int cls = kClassInt);
// A candidate the type cannot take (AFL on a double output) is never
// chosen; its training label is kInvalidLabel (2^8 times the input).
constexpr float kInvalidLabel = 8.0f;
bool candidate_valid(int idx, int dtype);

// 32 features of one block from its statistics pass.
std::array<float, kMlpFeatures> block_features(Ctx* c, int dtype, DView data);

// Exact blob size of every candidate tree on `data` (shared prefixes are
// encoded once), as log2(size / input bytes).
std::array<float, kCandidates> candidate_labels(Ctx* c, int dtype, DView data);
// Exact blob sizes of the listed candidates on `data` (SIZE_MAX for a tree
// the type cannot take).
std::vector<size_t> candidate_sizes(Ctx* c, int dtype, DView data, const std::vector<int>& idx);

class PlannerNet {
public:
    PlannerNet(Ctx* c, uint64_t seed);
    void get(float* host) const;
    void set(const float* host);
    // Mean squared error and its gradient (host copies) on n rows.
    float gradient(const float* X, const float* T, size_t n, float* grad_host);
    // Adam over `epochs` passes of shuffled minibatches; returns the last
    // epoch's mean loss.
    float train(const float* X, const float* T, size_t n, int epochs, int batch, float lr, uint64_t seed);
    void predict(const float* X, size_t n, float* Y);
    // Device-resident variants (rows already on the device).
    void predict_dev(const float* dX, int n, float* dY);

private:
    Ctx* c_;
    DevBuf params_, grads_, m_, v_;
    int step_ = 0;
};

}  // namespace ddj
