/*
 * CPU oracle for the dense-stereo hot path: a plain-C restatement of the
 * apg-sgm reference algorithm (cost -> right WTA -> Cross -> SGM -> WTA ->
 * confidence -> LR check -> fill -> median).
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / CPU baseline.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the reference repository root).  The restatement keeps the reference's
 * arithmetic exactly: integer work is bit-exact, the float cost expression is
 * evaluated in the same order with the same libm expf and no contraction
 * (build with -ffp-contract=off, no -mfma).
 *
 * Pinning (see tests/test_oracle_*.py):
 *   - the 13 frozen FNV-1a pipeline fingerprints of
 *     tests/test_api_contract.cpp:243-257 (full pipeline, all stages);
 *   - the 1-D SGM KAT of tests/test_synthetic.cpp:26-71;
 *   - the WTA earliest-tie KAT of tests/test_sanity.cpp:1904-1938;
 *   - subpixel sign/magnitude cases of tests/test_subpixel.cpp:10-99;
 *   - golden vectors written from a build of the reference's own sources
 *     (oracle/_ref, tests/golden/make_golden.py) for every stage output;
 *   - brute-force / closed-form checks (census bit definition, windowed mean,
 *     Hirschmueller recurrence on tiny volumes, path-order independence).
 */
#ifndef APG_ORACLE_H
#define APG_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors PipelineConfig (include/apg_sgm/config.hpp:7-112); plain ints for
 * enums and bools so that ctypes can fill it. */
typedef struct {
    int mode;               /* 0 Fast, 1 Balanced, 2 HighQuality */
    int min_disparity;
    int max_disparity;
    /* CostParams */
    int census_type;        /* 0 Census9x7 (64-bit), 1 SymmetricCensus9x7 */
    float lambda_census, lambda_ad, lambda_grad, eta_ad, mu_grad;
    int use_ad, use_grad, cost_max;
    /* AggregationParams */
    int agg_enable, max_arm_length, color_threshold, agg_iterations;
    /* SgmParams */
    int paths;              /* 4 or 8 */
    int P1, P2_base, adaptive_penalties, grad_t1, grad_t2;
    float p2_alpha;
    int piecewise_p2;
    /* ConfidenceParams */
    float conf_uniqueness_ratio;
    int conf_lr_max_diff;
    float conf_min_texture, conf_reliable_threshold;
    /* PostParams */
    int post_subpixel, post_lr_check, post_lr_max_diff, post_median_radius, post_max_fill_gap;
    /* PriorParams (include/apg_sgm/config.hpp:37-45) */
    int prior_enable;
    float prior_uniqueness_ratio;
    int prior_texture_threshold, prior_lr_max_diff, prior_search_radius, prior_min_supports, prior_max_supports;
    /* RefineParams (include/apg_sgm/config.hpp:54-58) */
    int refine_enable, refine_iterations, refine_random_radius;
} orc_cfg;

/* PipelineConfig::from_mode (src/config.cpp:5-45) */
void orc_config_from_mode(int mode, int max_disparity, orc_cfg* out);

/* gray + |dI/dx| + |dI/dy| (src/cost_computer.cpp:63-88, image_view.hpp:50-56) */
void orc_gray_grad(const uint8_t* src, int w, int h, int stride, int channels,
                   uint8_t* gray, uint8_t* gx, uint8_t* gy);
/* symmetric census 9x7, 31 bits (src/cost_computer.cpp:33-47) */
void orc_census_sym(const uint8_t* gray, int w, int h, uint32_t* out);
/* centre-compare census 9x7, 63 bits (src/cost_computer.cpp:49-61) */
void orc_census64(const uint8_t* gray, int w, int h, uint64_t* out);
/* search range with the prior off (src/prior_estimator.cpp:414-424,
 * include/apg_sgm/search_range.hpp:29-35) */
void orc_range_init(const orc_cfg* cfg, int w, int h, int16_t* dmin, int16_t* dmax);

/* dense cost volume [y][x][d], per-state expf (src/cost_computer.cpp:139-203).
 * census64_* are used when census_type == 0, census32_* otherwise. */
void orc_cost_volume(const orc_cfg* cfg, int w, int h,
                     const uint8_t* gray_l, const uint8_t* gx_l, const uint8_t* gy_l,
                     const uint8_t* gray_r, const uint8_t* gx_r, const uint8_t* gy_r,
                     const uint32_t* census32_l, const uint32_t* census32_r,
                     const uint64_t* census64_l, const uint64_t* census64_r,
                     const int16_t* dmin, const int16_t* dmax, uint16_t* cost);

/* right-view WTA on the raw volume (src/pipeline.cpp:24-85) */
void orc_right_wta(const orc_cfg* cfg, int w, int h, const uint16_t* cost, float* disp_right);

/* Cross arms (src/cost_aggregator.cpp:17-63) */
void orc_cross_arms(const orc_cfg* cfg, const uint8_t* gray, int w, int h,
                    int16_t* L, int16_t* R, int16_t* U, int16_t* Dn);
/* Cross H+V, brute-force window mean, `iterations` passes, in place
 * (src/cost_aggregator.cpp:65-149). No-op when agg_enable == 0. */
void orc_cross_aggregate(const orc_cfg* cfg, const uint8_t* gray, int w, int h,
                         const int16_t* dmin, const int16_t* dmax, uint16_t* cost);
/* the H pass alone, C -> T (src/cost_aggregator.cpp:74-105) */
void orc_cross_h(const orc_cfg* cfg, int w, int h, const int16_t* dmin, const int16_t* dmax,
                 const int16_t* L, const int16_t* R, const uint16_t* cost, uint16_t* tmp);
/* the V pass alone, T -> C' (src/cost_aggregator.cpp:107-138) */
void orc_cross_v(const orc_cfg* cfg, int w, int h, const int16_t* dmin, const int16_t* dmax,
                 const int16_t* U, const int16_t* Dn, const uint16_t* tmp, uint16_t* cost);

/* adaptive P2 (src/sgm_common.hpp:25-39) */
int orc_adaptive_p2(const orc_cfg* cfg, int dI);
/* one SGM path added into acc (u32, zero-initialised by the caller)
 * (src/sgm_optimizer.cpp:21-81, 186-259) */
void orc_sgm_path(const orc_cfg* cfg, const uint8_t* gray, int w, int h, int D,
                  const uint16_t* cost, uint32_t* acc, int dx, int dy);
/* full optimize: zero acc then 4 or 8 paths (src/sgm_optimizer.cpp:261-274) */
void orc_sgm_optimize(const orc_cfg* cfg, const uint8_t* gray, int w, int h,
                      const uint16_t* cost, uint32_t* acc);

/* legacy parabola (src/subpixel_estimator.hpp:115-124) */
float orc_subpixel_legacy(float cm, float c0, float cp);
/* WTA over a u32 volume (src/sgm_optimizer.cpp:436-555); best/second may be NULL */
void orc_wta_u32(const orc_cfg* cfg, int w, int h, const uint32_t* vol,
                 const int16_t* dmin, const int16_t* dmax,
                 float* disp, float* best, float* second);
/* WTA over a u16 volume (same function, TCost = uint16_t) */
void orc_wta_u16(const orc_cfg* cfg, int w, int h, const uint16_t* vol,
                 const int16_t* dmin, const int16_t* dmax,
                 float* disp, float* best, float* second);

/* confidence (src/confidence_estimator.cpp:7-63) */
void orc_confidence(const orc_cfg* cfg, int w, int h, const float* disp, const float* disp_right,
                    const uint8_t* gx_l, const uint8_t* gy_l, const float* best, const float* second,
                    float* confidence, uint8_t* reliable, uint8_t* invalid_reason);
/* LR check, in place (src/postprocess.cpp:10-40) */
void orc_lr_check(const orc_cfg* cfg, int w, int h, float* disp, const float* disp_right,
                  uint8_t* invalid_reason);
/* legacy hole fill (src/postprocess.cpp:42-253, Legacy route); d_prior may be NULL */
void orc_fill_holes(const orc_cfg* cfg, int w, int h, float* disp, const uint8_t* invalid_reason,
                    const uint8_t* gray_l, const float* d_prior);
/* valid-only median (src/postprocess.cpp:255-309) */
void orc_median(const orc_cfg* cfg, int w, int h, float* disp);

/* ---- HighQuality stages (oracle_hq.cpp) ---------------------------- */

typedef struct {
    int x, y;
    float disparity, confidence;
} orc_support;

/* PriorEstimator::extract_supports (src/prior_estimator.cpp:13-166); writes at
 * most `cap` supports, returns how many the reference keeps */
int orc_prior_supports(const orc_cfg* cfg, int w, int h, const uint8_t* gx_l, const uint8_t* gy_l,
                       const uint32_t* c32l, const uint32_t* c32r, const uint64_t* c64l, const uint64_t* c64r,
                       orc_support* out, int cap);
/* PriorEstimator::interpolate_prior (src/prior_estimator.cpp:168-375) */
void orc_prior_interpolate(const orc_support* sup, int nsup, int w, int h, float* d_prior, float* conf,
                           float* spread, float* pmin, float* pmax);
/* PriorEstimator::apply_search_range (src/prior_estimator.cpp:377-412) */
void orc_prior_apply_range(const orc_cfg* cfg, int w, int h, const float* d_prior, const float* conf,
                           const float* spread, const float* pmin, const float* pmax, int16_t* dmin,
                           int16_t* dmax);
/* PriorEstimator::estimate (src/prior_estimator.cpp:414-435); returns the
 * support count */
int orc_prior_estimate(const orc_cfg* cfg, int w, int h, const uint8_t* gx_l, const uint8_t* gy_l,
                       const uint32_t* c32l, const uint32_t* c32r, const uint64_t* c64l, const uint64_t* c64r,
                       int16_t* dmin, int16_t* dmax, float* d_prior);
/* Refiner::refine, scalar route (src/refiner.cpp:14-98), in place on disp */
void orc_refine(const orc_cfg* cfg, int w, int h, const uint8_t* gl, const uint8_t* gxl, const uint8_t* gyl,
                const uint8_t* gr, const uint8_t* gxr, const uint8_t* gyr, const uint32_t* c32l,
                const uint32_t* c32r, const uint64_t* c64l, const uint64_t* c64r, const int16_t* dmin,
                const int16_t* dmax, const uint8_t* reliable, float* disp);
/* draws of std::uniform_int_distribution<int>(-r, r) over a replayed 32-bit
 * engine stream; the number of stream outputs consumed in *consumed */
void orc_uniform_draws(const uint32_t* stream, size_t n, int r, int ndraw, int* out, size_t* consumed);

/* Pipeline (src/pipeline.cpp:435-605): aux, prior (or the dense range
 * init), cost, right WTA, Cross, SGM, WTA, confidence, refine, LR check,
 * fill (with the prior), median.  The dense branch with the per-pixel ranges
 * is also what the packed branch computes (the reference's own dense ==
 * packed tests, test_sanity.cpp:83-476).  Any intermediate pointer may be NULL.
 * Returns 0 on success or the reference's StatusCode on validation failure. */
typedef struct {
    uint16_t* cost;          /* C  (W*H*D), raw cost volume */
    uint16_t* cost_agg;      /* C' (W*H*D), after Cross */
    uint32_t* sgm;           /* S  (W*H*D), path sums */
    float* disp_right;       /* W*H */
    float* disp_wta;         /* W*H, WTA output before post (d_before_refine) */
    float* best;             /* W*H */
    float* second;           /* W*H */
    int16_t* dmin;           /* W*H search range (prior-narrowed in HighQuality) */
    int16_t* dmax;           /* W*H */
    float* d_prior;          /* W*H, -1 where there is no prior */
    float* after_refine;     /* W*H, disparity after the refiner (d_after_refine,
                                src/pipeline.cpp:597-599) */
} orc_intermediates;

int orc_pipeline(const orc_cfg* cfg,
                 const uint8_t* left, const uint8_t* right, int w, int h, int stride, int channels,
                 float* disparity, float* confidence, uint8_t* invalid_reason, uint8_t* reliable,
                 orc_intermediates* inter);

/* FNV-1a 64 over bytes (tests/test_api_contract.cpp:14-32) */
uint64_t orc_fnv1a64(const void* data, size_t bytes);

/* threads the OpenMP regions use (0 = library default) */
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif

#endif
