// ipt_select.h — candidate configurations, shape features and the MLP cost
// model shared by the autotuner, the selector and the dispatcher.
#pragma once

#include <stdint.h>

#include "../../include/inplace_transpose.h"

namespace ipt {

// The configurations the autotuner times and the selector chooses among.
struct Candidate {
    int variant;
    ipt_tune tune;  // zero fields = library default
};

constexpr int kNumCandidates = IPT_NUM_CANDIDATES;
extern const Candidate kCandidates[kNumCandidates];

// The selector decides only inside the domain its timing sweep covers:
// payloads up to kSelectorMaxBytes and both sides up to kSelectorMaxSide;
// other shapes keep the heuristic (the network does not extrapolate).
constexpr uint64_t kSelectorMaxBytes = (uint64_t)1 << 30;
constexpr int64_t kSelectorMaxSide = 65536;
inline bool selector_domain(int64_t d1, int64_t d2, int64_t d3, int esz) {
    return d1 <= kSelectorMaxSide && d2 <= kSelectorMaxSide &&
           (uint64_t)d1 * d2 * d3 * esz <= kSelectorMaxBytes;
}

// Candidate the plain heuristic picks for a shape (maps choose_variant onto
// the candidate list).
int heuristic_candidate(int variant);

// MLP 8 -> 64 -> 64 -> K with ReLU, fp32, row-major weights:
//   h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = W3 h2 + b3
// y[k] estimates log2(time in microseconds) of candidate k.
constexpr int kFeat = IPT_MLP_IN, kHid = IPT_MLP_HIDDEN, kOut = IPT_NUM_CANDIDATES;
constexpr int kW1 = kHid * kFeat, kB1 = kHid, kW2 = kHid * kHid, kB2 = kHid, kW3 = kOut * kHid,
              kB3 = kOut;
constexpr int kParams = kW1 + kB1 + kW2 + kB2 + kW3 + kB3;

// Shape features (all O(1) in magnitude):
//   log2 R / 16, log2 C / 16, log2 gcd(R,C) / 16, esz == 8, (log2 R - log2 C) / 16,
//   log2(bytes of one slice) / 32, R == C, log2 d3 / 16
// with R = d2 (rows) and C = d1 (columns) of the 213 problem.
void shape_features(int esz, int64_t d1, int64_t d2, int64_t d3, float* f);

// Host evaluation of the MLP (same arithmetic order as the GPU kernels).
void mlp_forward_host(const float* params, const float* x, float* y);

}  // namespace ipt
