// Test harness around a build of the reference's own sources (oracle/_ref).
// TEST INFRASTRUCTURE ONLY: compiled together with the reference sources by
// oracle/Makefile (target `ref`); used to write golden vectors and to time the
// reference CPU path as a reported baseline. It exposes the reference's stage
// classes through a C ABI that ctypes can call; it contains no algorithm.
#include "apg_sgm/config.hpp"
#include "apg_sgm/cost_aggregator.hpp"
#include "apg_sgm/cost_computer.hpp"
#include "apg_sgm/pipeline.hpp"
#include "apg_sgm/prior_estimator.hpp"
#include "apg_sgm/sgm_optimizer.hpp"

#include "../oracle/oracle.h"

#include <chrono>
#include <cstring>
#include <vector>
#include <algorithm>
#if defined(_OPENMP)
#include <omp.h>
#endif

using namespace apg;

static PipelineConfig make_cfg(const orc_cfg* c, int backend) {
    PipelineConfig cfg = PipelineConfig::from_mode(static_cast<QualityMode>(c->mode), c->max_disparity);
    cfg.min_disparity = c->min_disparity;
    cfg.max_disparity = c->max_disparity;
    cfg.volume_backend = static_cast<VolumeBackend>(backend);
    cfg.cost.census = c->census_type == 1 ? CensusType::SymmetricCensus9x7 : CensusType::Census9x7;
    cfg.cost.lambda_census = c->lambda_census;
    cfg.cost.lambda_ad = c->lambda_ad;
    cfg.cost.lambda_grad = c->lambda_grad;
    cfg.cost.eta_ad = c->eta_ad;
    cfg.cost.mu_grad = c->mu_grad;
    cfg.cost.use_ad = c->use_ad != 0;
    cfg.cost.use_grad = c->use_grad != 0;
    cfg.cost.cost_max = c->cost_max;
    cfg.aggregation.enable = c->agg_enable != 0;
    cfg.aggregation.max_arm_length = c->max_arm_length;
    cfg.aggregation.color_threshold = c->color_threshold;
    cfg.aggregation.iterations = c->agg_iterations;
    cfg.sgm.paths = c->paths == 8 ? PathType::Path8 : PathType::Path4;
    cfg.sgm.P1 = c->P1;
    cfg.sgm.P2_base = c->P2_base;
    cfg.sgm.adaptive_penalties = c->adaptive_penalties != 0;
    cfg.sgm.grad_t1 = c->grad_t1;
    cfg.sgm.grad_t2 = c->grad_t2;
    cfg.sgm.p2_alpha = c->p2_alpha;
    cfg.sgm.piecewise_p2 = c->piecewise_p2 != 0;
    cfg.confidence.uniqueness_ratio = c->conf_uniqueness_ratio;
    cfg.confidence.lr_max_diff = c->conf_lr_max_diff;
    cfg.confidence.min_texture = c->conf_min_texture;
    cfg.confidence.reliable_threshold = c->conf_reliable_threshold;
    cfg.post.subpixel = c->post_subpixel != 0;
    cfg.post.lr_check = c->post_lr_check != 0;
    cfg.post.lr_max_diff = c->post_lr_max_diff;
    cfg.post.median_radius = c->post_median_radius;
    cfg.post.max_fill_gap = c->post_max_fill_gap;
    cfg.prior.enable = c->prior_enable != 0;
    cfg.prior.uniqueness_ratio = c->prior_uniqueness_ratio;
    cfg.prior.texture_threshold = c->prior_texture_threshold;
    cfg.prior.lr_max_diff = c->prior_lr_max_diff;
    cfg.prior.search_radius = c->prior_search_radius;
    cfg.prior.min_supports = c->prior_min_supports;
    cfg.prior.max_supports = c->prior_max_supports;
    cfg.refine.enable = c->refine_enable != 0;
    cfg.refine.iterations = c->refine_iterations;
    cfg.refine.random_radius = c->refine_random_radius;
    return cfg;
}

static Image8 to_image(const uint8_t* p, int w, int h, int stride, int ch) {
    Image8 img(w, h, ch);
    for (int y = 0; y < h; ++y) std::memcpy(img.data() + static_cast<size_t>(y) * w * ch, p + static_cast<size_t>(y) * stride, static_cast<size_t>(w) * ch);
    return img;
}

extern "C" {

// Stage-by-stage run of the dense path through the reference classes
// (the order of StereoMatcher::compute_impl's dense branch). Any output may be NULL.
int ref_stages(const orc_cfg* c, const uint8_t* left, const uint8_t* right, int w, int h, int stride, int ch,
               uint16_t* out_cost, uint16_t* out_cost_agg, uint32_t* out_cost32,
               float* out_disp, float* out_best, float* out_second,
               uint32_t* out_census_l, uint8_t* out_gray_l, int16_t* out_dmin, int16_t* out_dmax) {
    PipelineConfig cfg = make_cfg(c, 1);
    PipelineBuffers buf;
    CostComputer cost;
    PriorEstimator prior;
    CostAggregator agg;
    SgmOptimizer sgm;
    ImageView8 lv(left, w, h, stride, ch), rv(right, w, h, stride, ch);
    cost.compute_aux(lv, rv, cfg, buf);
    prior.estimate(cfg, buf);
    cost.compute_volume(cfg, buf);
    const size_t nv = static_cast<size_t>(w) * h * (cfg.max_disparity - cfg.min_disparity);
    if (out_cost) std::memcpy(out_cost, buf.cost.data(), nv * 2);
    agg.aggregate(cfg, buf);
    if (out_cost_agg) std::memcpy(out_cost_agg, buf.cost.data(), nv * 2);
    sgm.optimize(cfg, buf);
    if (out_cost32) std::memcpy(out_cost32, buf.cost32.data(), nv * 4);
    Image32f disp, best, second;
    sgm.winner_take_all(cfg, buf.cost32, buf.range, disp, &best, &second);
    const size_t n = static_cast<size_t>(w) * h;
    if (out_disp) std::memcpy(out_disp, disp.data(), n * 4);
    if (out_best) std::memcpy(out_best, best.data(), n * 4);
    if (out_second) std::memcpy(out_second, second.data(), n * 4);
    if (out_census_l && !buf.census_left.empty()) std::memcpy(out_census_l, buf.census_left.data(), n * 4);
    if (out_gray_l) std::memcpy(out_gray_l, buf.left_gray.data(), n);
    if (out_dmin) std::memcpy(out_dmin, buf.range.dmin.data(), n * 2);
    if (out_dmax) std::memcpy(out_dmax, buf.range.dmax.data(), n * 2);
    return 0;
}

// Full StereoMatcher::compute into PipelineBuffers (any mode and backend).
int ref_pipeline(const orc_cfg* c, int backend, const uint8_t* left, const uint8_t* right,
                 int w, int h, int stride, int ch,
                 float* disparity, float* confidence, uint8_t* invalid_reason, uint8_t* reliable,
                 float* disp_right, float* d_before_refine, int16_t* dmin, int16_t* dmax) {
    PipelineConfig cfg = make_cfg(c, backend);
    StereoMatcher m(cfg);
    PipelineBuffers buf;
    Image8 li = to_image(left, w, h, stride, ch), ri = to_image(right, w, h, stride, ch);
    if (!m.compute(li, ri, buf)) return 8;
    const size_t n = static_cast<size_t>(w) * h;
    if (disparity) std::memcpy(disparity, buf.disparity.data(), n * 4);
    if (confidence) std::memcpy(confidence, buf.confidence.data(), n * 4);
    if (invalid_reason) std::memcpy(invalid_reason, buf.invalid_reason.data(), n);
    if (reliable) std::memcpy(reliable, buf.reliable_mask.data(), n);
    if (disp_right && !buf.disparity_right.empty()) std::memcpy(disp_right, buf.disparity_right.data(), n * 4);
    if (d_before_refine && !buf.d_before_refine.empty()) std::memcpy(d_before_refine, buf.d_before_refine.data(), n * 4);
    if (dmin) std::memcpy(dmin, buf.range.dmin.data(), n * 2);
    if (dmax) std::memcpy(dmax, buf.range.dmax.data(), n * 2);
    return 0;
}

// One SGM path over a caller volume into a zeroed u32 accumulator
// (SgmOptimizer::aggregate_path) -- used for the 1-D KAT.
void ref_sgm_path(const orc_cfg* c, const uint8_t* gray, int w, int h, int D,
                  const uint16_t* cost, uint32_t* acc, int dx, int dy) {
    PipelineConfig cfg = make_cfg(c, 1);
    CostVolume base(w, h, 0, D, 0);
    std::memcpy(base.data(), cost, static_cast<size_t>(w) * h * D * 2);
    CostVolume32 a(w, h, 0, D, 0);
    std::memcpy(a.data(), acc, static_cast<size_t>(w) * h * D * 4);
    Image8 g(w, h, 1);
    std::memcpy(g.data(), gray, static_cast<size_t>(w) * h);
    SgmOptimizer().aggregate_path(cfg, g, base, a, dx, dy);
    std::memcpy(acc, a.data(), static_cast<size_t>(w) * h * D * 4);
}

// Median wall time (ms) of `reps` StereoMatcher::compute calls after `warmup`
// calls (apps/benchmark_stereo.cpp:537-575 protocol). threads<=0: OpenMP default.
double ref_time_pipeline(const orc_cfg* c, int backend, const uint8_t* left, const uint8_t* right,
                         int w, int h, int threads, int warmup, int reps) {
    PipelineConfig cfg = make_cfg(c, backend);
    cfg.num_threads = threads > 0 ? threads : 0;
    StereoMatcher m(cfg);
    Image8 li = to_image(left, w, h, w, 1), ri = to_image(right, w, h, w, 1);
    std::vector<double> t;
    for (int i = 0; i < warmup + reps; ++i) {
        PipelineBuffers buf;
        auto t0 = std::chrono::steady_clock::now();
        if (!m.compute(li, ri, buf)) return -1.0;
        auto t1 = std::chrono::steady_clock::now();
        if (i >= warmup) t.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
    }
    std::sort(t.begin(), t.end());
    return t.empty() ? -1.0 : t[t.size() / 2];
}

int ref_max_threads(void) {
#if defined(_OPENMP)
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
