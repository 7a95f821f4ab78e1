// Layout constants of the HighQuality prior shared by the device kernels
// (k_prior.hip) and the host code that sizes their buffers.
#pragma once

#include <stdint.h>

#include "../../include/apg_sgm_hip.h"

namespace apgk {

constexpr int kPriorCell = 16;   // prior grid cell (reference src/prior_estimator.cpp:22)
constexpr int16_t kNoCandidate = INT16_MIN;

// Samples of the prior: even x in [2, w-2), even y in [2, h-2).
inline int prior_samples_x(int w) { return w >= 3 ? (w - 3) / 2 : 0; }
inline int prior_samples_y(int h) { return h >= 3 ? (h - 3) / 2 : 0; }

}  // namespace apgk
