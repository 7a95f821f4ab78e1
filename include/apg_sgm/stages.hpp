// Stage classes with the reference's interface, backed by the MI355X device
// layer (include/apg_sgm_hip.h):
//   CostComputer    (reference include/apg_sgm/cost_computer.hpp:10-32)
//   CostAggregator  (reference include/apg_sgm/cost_aggregator.hpp:9-30)
//   SgmOptimizer    (reference include/apg_sgm/sgm_optimizer.hpp:20-61)
//   ConfidenceEstimator (reference include/apg_sgm/confidence_estimator.hpp)
//   PriorEstimator  (reference include/apg_sgm/prior_estimator.hpp)
//   Refiner         (reference include/apg_sgm/refiner.hpp)
//   PostProcessor   (reference include/apg_sgm/postprocess.hpp:8-13)
// They keep the reference's host-buffer contract (PipelineBuffers in, volumes
// out on the host), so every call uploads its inputs, runs the device
// kernels on the calling thread's default device context and downloads its
// outputs.  The packed (ragged) methods run the dense kernels over the
// layout's per-pixel ranges and pack the result, which the reference's own
// dense == packed tests show to be the same values (reference
// tests/test_sanity.cpp:83-476).  They are the parity boundary; throughput is measured at the
// session / C-ABI level where volumes stay in HBM.  Failures throw
// std::runtime_error / std::bad_alloc like the reference's stages.
#pragma once

#include "apg_sgm/buffers.hpp"
#include "apg_sgm/config.hpp"
#include "apg_sgm/image_view.hpp"

namespace apg {

inline bool can_use_u16_sgm_accumulator(const PipelineConfig& cfg, int npath) {
    const uint64_t p2_max = std::max<uint64_t>(static_cast<uint64_t>(cfg.sgm.P2_base),
                                               static_cast<uint64_t>(cfg.sgm.P1) + 1);
    return static_cast<uint64_t>(npath) * (static_cast<uint64_t>(cfg.cost.cost_max) + p2_max) < kInvalidCost;
}

class CostComputer {
public:
    void compute_aux(const ImageView8& left, const ImageView8& right, const PipelineConfig& cfg,
                     PipelineBuffers& buf) const;
    void compute_aux(const Image8& left, const Image8& right, const PipelineConfig& cfg, PipelineBuffers& buf) const;
    void compute_volume(const PipelineConfig& cfg, PipelineBuffers& buf) const;
    // ragged volume over packed_cost's layout (allocated from buf.range when
    // it has none); reference src/cost_computer.cpp:205-268
    void compute_volume_packed(const PipelineConfig& cfg, const PipelineBuffers& buf,
                               PackedCostVolume16& packed_cost) const;

    static uint32_t symmetric_census9x7(const Image8& gray, int x, int y);
    static uint64_t census9x7(const Image8& gray, int x, int y);
    static int popcount32(uint32_t x) { return __builtin_popcount(x); }
    static int popcount64(uint64_t x) { return __builtin_popcountll(x); }
};

// Search ranges (reference src/prior_estimator.cpp:414-435): the dense init
// [min, min(max, x+1)), narrowed by the sparse prior when cfg.prior.enable.
// extract_supports (reference :29-166) returns the support matches the prior
// would use, from the buffers' census and gradients.  The reference's two
// middle steps (interpolate_prior, apply_search_range) run fused on the device
// inside estimate and are not exposed separately.
class PriorEstimator {
public:
    void estimate(const PipelineConfig& cfg, PipelineBuffers& buf) const;
    std::vector<SupportMatch> extract_supports(const PipelineConfig& cfg, const PipelineBuffers& buf) const;
};

class CostAggregator {
public:
    void aggregate(const PipelineConfig& cfg, PipelineBuffers& buf) const;
    // reference src/cost_aggregator.cpp:151-403
    void aggregate_packed(const PipelineConfig& cfg, const Image8& left_gray, PackedCostVolume16& packed_cost) const;
};

class SgmOptimizer {
public:
    void optimize(const PipelineConfig& cfg, PipelineBuffers& buf) const;
    void aggregate_path(const PipelineConfig& cfg, const Image8& gray, const CostVolume& base, CostVolume32& acc,
                        int dx, int dy) const;
    void winner_take_all(const PipelineConfig& cfg, const CostVolume32& vol, const SearchRange& range,
                         Image32f& disp_out, Image32f* best_cost = nullptr, Image32f* second_cost = nullptr) const;
    void winner_take_all(const PipelineConfig& cfg, const CostVolume& vol, const SearchRange& range,
                         Image32f& disp_out, Image32f* best_cost = nullptr, Image32f* second_cost = nullptr) const;
    // ragged forms over the cost volume's layout (reference
    // src/sgm_optimizer.cpp:83-182, 276-404, 557-710); the outputs take that layout
    void optimize_packed(const PipelineConfig& cfg, const Image8& gray, const PackedCostVolume16& packed_cost,
                         PackedCostVolume16& packed_sgm16) const;
    void optimize_packed(const PipelineConfig& cfg, const Image8& gray, const PackedCostVolume16& packed_cost,
                         PackedCostVolume32& packed_cost32) const;
    void winner_take_all_packed(const PipelineConfig& cfg, const PackedCostVolume32& vol, Image32f& disp_out,
                                Image32f* best_cost = nullptr, Image32f* second_cost = nullptr) const;
    void winner_take_all_packed(const PipelineConfig& cfg, const PackedCostVolume16& vol, Image32f& disp_out,
                                Image32f* best_cost = nullptr, Image32f* second_cost = nullptr) const;
    void aggregate_path_packed(const PipelineConfig& cfg, const Image8& gray, const PackedCostVolume16& base,
                               PackedCostVolume16& acc, int dx, int dy) const;
    void aggregate_path_packed(const PipelineConfig& cfg, const Image8& gray, const PackedCostVolume16& base,
                               PackedCostVolume32& acc, int dx, int dy) const;
};

class ConfidenceEstimator {
public:
    void estimate(const PipelineConfig& cfg, PipelineBuffers& buf, const Image32f& best_cost,
                  const Image32f& second_cost) const;
};

// Randomised raster refinement of the unreliable pixels (reference
// src/refiner.cpp:42-98) over buf.disparity, using the buffers' gray,
// gradients, census, search range and reliable mask.
class Refiner {
public:
    void refine(const PipelineConfig& cfg, PipelineBuffers& buf) const;
};

class PostProcessor {
public:
    void left_right_check(const PipelineConfig& cfg, PipelineBuffers& buf) const;
    void fill_holes(const PipelineConfig& cfg, PipelineBuffers& buf) const;
    void median(const PipelineConfig& cfg, PipelineBuffers& buf) const;
};

}  // namespace apg
