// Accuracy metrics of a disparity map against ground truth, with the
// reference's interface and definitions (reference include/apg_sgm/metrics.hpp:9-146,
// src/metrics.cpp:8-505): end-point error, bad-pixel rates, KITTI D1, the
// search-range recall of a prior, support precision, region breakdowns
// (visible / occluded, disparity boundary, low texture) and refiner deltas.
#pragma once

#include <stddef.h>

#include <vector>

#include "apg_sgm/buffers.hpp"
#include "apg_sgm/image.hpp"
#include "apg_sgm/search_range.hpp"
#include "apg_sgm/types.hpp"

namespace apg {

struct SupportMetrics {
    int total = 0, gt_valid = 0, visible = 0;
    int correct_05 = 0, correct_1 = 0, correct_2 = 0;
    int visible_correct_05 = 0, visible_correct_1 = 0, visible_correct_2 = 0;
    int visible_gt_cells = 0, correct_visible_cells = 0;
    float precision_05 = 0.f, precision_1 = 0.f, precision_2 = 0.f;   // = precision_all_*
    float precision_all_05 = 0.f, precision_all_1 = 0.f, precision_all_2 = 0.f;
    float precision_vis_05 = 0.f, precision_vis_1 = 0.f, precision_vis_2 = 0.f;
    float mean_abs_error = 0.f, mae_all = 0.f, mae_visible = 0.f;
    // visible GT cells holding a visible support within 1 px, over visible GT cells
    float grid_recall_1 = 0.f, grid_coverage = 0.f, support_grid_recall_vis_1 = 0.f;
};

struct RegionMetrics {
    int total_pixels = 0;       // GT-valid pixels of the region
    int valid_est_pixels = 0;   // of them, with an estimate (finite, >= 0)
    float valid_ratio = 0.f;
    float epe = 0.f;            // mean |e - g| over valid estimates
    float bad_0_5 = 0.f, bad_1_0 = 0.f, bad_2_0 = 0.f, bad_3_0 = 0.f;   // % of valid estimates
    float d1 = 0.f;             // KITTI D1 % of total_pixels (missing estimates count)
};

struct QualityThresholds {
    float boundary_disparity_gradient = 1.0f;   // GT 4-neighbour difference marking a boundary
    float low_texture_gradient = 4.0f;          // left-image luma gradient below which texture is low
    float max_valid_gt = 1e5f;
};

struct StereoMetrics {
    int total_pixels = 0, evaluated_pixels = 0, valid_pixels = 0;
    float valid_ratio = 0.f;
    float epe = 0.f, bad_0_5 = 0.f, bad_1_0 = 0.f, bad_2_0 = 0.f, bad_3_0 = 0.f;
    float kitti_d1_all = 0.f, kitti_d1_noc = 0.f;
    int edge_pixels = 0;
    float edge_epe = 0.f;
    int nonedge_pixels = 0;
    float nonedge_epe = 0.f;
    float lr_fail_ratio = 0.f;

    bool has_range = false;
    float range_recall_all = 0.f, range_recall_matchable = 0.f, range_recall_visible = 0.f;
    size_t range_visible_eval_pixels = 0, range_visible_in_pixels = 0;
    size_t range_matchable_eval_pixels = 0, range_matchable_in_pixels = 0;
    int prior_miss_visible_count = 0, prior_miss_edge = 0, prior_miss_nonedge = 0;
    float prior_miss_edge_ratio = 0.f, prior_miss_nonedge_ratio = 0.f;

    bool has_supports = false;
    SupportMetrics support_metrics;

    RegionMetrics reg_all, reg_noc, reg_occ, reg_boundary, reg_non_boundary, reg_low_texture, reg_high_texture;

    bool has_refine_stats = false;
    int refine_eval_pixels = 0;
    float refine_epe_delta = 0.f, post_epe_delta = 0.f, total_epe_delta = 0.f;
    float refine_improved_ratio = 0.f, refine_worsened_ratio = 0.f, refine_unchanged_ratio = 0.f;
    float refine_changed_ratio = 0.f;
};

SupportMetrics evaluate_supports(const std::vector<SupportMatch>& supports, const Image32f& gt,
                                 const Image8* vis_mask = nullptr, float max_valid_gt = 1e5f, int cell_size = 16);

StereoMetrics evaluate_stereo(const Image32f& est, const Image32f& gt, const SearchRange* range = nullptr,
                              const Image32f* d_before_refine = nullptr, const Image32f* d_after_refine = nullptr,
                              const Image8* vis_mask = nullptr, float max_valid_gt = 1e5f,
                              const std::vector<SupportMatch>* supports = nullptr, const Image8* left_image = nullptr,
                              const QualityThresholds& thresholds = QualityThresholds{});

}  // namespace apg
