// Host-side part of the HighQuality prior: the support selection and grid
// statistics (a few thousand supports on a 16-pixel grid: small, sort-heavy,
// order-sensitive).  Everything per pixel or per disparity state runs on the
// device (k_prior.hip, k_refine.hip).
#pragma once

#include <stdint.h>

#include <vector>

#include "../../include/apg_sgm_hip.h"

namespace apgk {

constexpr int kPriorCell = 16;   // prior grid cell (reference src/prior_estimator.cpp:22)
constexpr int16_t kNoCandidate = INT16_MIN;

// Samples of the prior: even x in [2, w-2), even y in [2, h-2).
inline int prior_samples_x(int w) { return w >= 3 ? (w - 3) / 2 : 0; }
inline int prior_samples_y(int h) { return h >= 3 ? (h - 3) / 2 : 0; }

// Per-cell prior grid after hole filling: disparity (-1 = none),
// confidence, spread, lowest and highest support disparity.
struct PriorGrid {
    int gw = 0, gh = 0;
    std::vector<float> d, c, s, lo, hi;
};

// PriorEstimator::extract_supports' selection and interpolate_prior's grid
// stages (reference src/prior_estimator.cpp:29-166, 181-310) from the
// device's per-sample candidates (cand_d = kNoCandidate where the sample is
// not a candidate).  Returns the number of supports kept.
int prior_build_grid(const apg_hip_params& p, int w, int h, const int16_t* cand_d, const float* cand_u,
                     PriorGrid& grid, std::vector<apg_hip_support>* supports = nullptr);

}  // namespace apgk
