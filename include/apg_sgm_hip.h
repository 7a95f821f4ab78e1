/*
 * apg_sgm_hip.h -- thin C-ABI device layer of the MI355X (gfx950) build of the
 * apg-sgm dense-stereo hot path.
 *
 * Every entry point takes plain pointers and sizes, returns an
 * apg_status_code_t (0 = OK), never throws, and is stream-ordered on the
 * context's HIP stream.  Pointers named d_* are device (HBM) pointers owned
 * by the caller or by the context's workspace.  Volumes use the reference's
 * [y][x][d] layout (include/apg_sgm/buffers.hpp:45-47 in the reference):
 * element (x, y, d) lives at ((y * w + x) * D + (d - min_disparity)).
 *
 * Each stage function cites the reference interface it replaces; the C++
 * stage classes (include/apg_sgm/ headers) and the session C ABI
 * (include/apg_sgm_c_api.h) are built on top of these calls.
 */
#ifndef APG_SGM_HIP_H
#define APG_SGM_HIP_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__) && __GNUC__ >= 4
#define APG_HIP_API __attribute__((visibility("default")))
#else
#define APG_HIP_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Stage parameters: the fields of PipelineConfig (reference
 * include/apg_sgm/config.hpp:7-112) that the hot path reads, as plain ints
 * and floats. Fill with apg_hip_params_from_mode() and override. */
typedef struct {
    int mode;               /* 0 Fast, 1 Balanced, 2 HighQuality */
    int min_disparity;
    int max_disparity;
    int census_type;        /* 0 Census9x7 (64-bit), 1 SymmetricCensus9x7 */
    float lambda_census, lambda_ad, lambda_grad, eta_ad, mu_grad;
    int use_ad, use_grad, cost_max;
    int agg_enable, max_arm_length, color_threshold, agg_iterations;
    int paths;              /* 4 or 8 */
    int P1, P2_base, adaptive_penalties, grad_t1, grad_t2;
    float p2_alpha;
    int piecewise_p2;
    float conf_uniqueness_ratio;
    int conf_lr_max_diff;
    float conf_min_texture, conf_reliable_threshold;
    int post_subpixel, post_lr_check, post_lr_max_diff, post_median_radius, post_max_fill_gap;
    /* PriorParams (reference include/apg_sgm/config.hpp:37-45): sparse census
     * supports -> per-pixel search ranges (HighQuality) */
    int prior_enable;
    float prior_uniqueness_ratio;
    int prior_texture_threshold, prior_lr_max_diff, prior_search_radius, prior_min_supports, prior_max_supports;
    /* RefineParams (:54-58): randomised raster refinement (HighQuality) */
    int refine_enable, refine_iterations, refine_random_radius;
} apg_hip_params;

typedef struct apg_hip_ctx_s* apg_hip_ctx;

/* Per-stage device time of the last pipeline call, milliseconds (HIP events
 * on the context stream; zero unless stage timing is enabled). Field names
 * follow StageTiming (reference include/apg_sgm/pipeline_stats.hpp:10-22). */
typedef struct {
    double aux_ms, prior_ms, cost_ms, right_wta_ms, cross_ms, sgm_ms, wta_ms,
           confidence_ms, refine_ms, post_ms, total_ms;
} apg_hip_timing;

/* PipelineConfig::from_mode (reference src/config.cpp:5-45). */
APG_HIP_API void apg_hip_params_from_mode(int mode, int max_disparity, apg_hip_params* out);

/* validate_config (reference src/pipeline.cpp:171-213) plus the device
 * path's own representability limits; returns a status code. */
APG_HIP_API int apg_hip_validate_params(const apg_hip_params* p);

/* Context: one HIP stream at default priority on `device`, a device
 * workspace and the LUTs.  No counterpart in the reference (its state lives
 * in PipelineBuffers / StereoWorkspace, reference include/apg_sgm/workspace.hpp). */
APG_HIP_API int apg_hip_ctx_create(int device, apg_hip_ctx* out);
APG_HIP_API void apg_hip_ctx_destroy(apg_hip_ctx ctx);
APG_HIP_API int apg_hip_synchronize(apg_hip_ctx ctx);
APG_HIP_API void* apg_hip_stream(apg_hip_ctx ctx);          /* hipStream_t */
APG_HIP_API size_t apg_hip_workspace_bytes(apg_hip_ctx ctx);
APG_HIP_API void apg_hip_release_workspace(apg_hip_ctx ctx);
APG_HIP_API void apg_hip_enable_stage_timing(apg_hip_ctx ctx, int on);
APG_HIP_API int apg_hip_last_timing(apg_hip_ctx ctx, apg_hip_timing* out);
APG_HIP_API const char* apg_hip_last_error(apg_hip_ctx ctx);

/* Device memory plumbing (synchronous on the context stream). */
APG_HIP_API int apg_hip_malloc(apg_hip_ctx ctx, size_t bytes, void** d_out);
APG_HIP_API int apg_hip_free(apg_hip_ctx ctx, void* d_ptr);
APG_HIP_API int apg_hip_upload(apg_hip_ctx ctx, void* d_dst, const void* h_src, size_t bytes);
APG_HIP_API int apg_hip_download(apg_hip_ctx ctx, void* h_dst, const void* d_src, size_t bytes);
APG_HIP_API int apg_hip_memset(apg_hip_ctx ctx, void* d_dst, int value, size_t bytes);
/* Pinned (page-locked) host memory: uploads and downloads from it run at the
 * link's full rate.  No counterpart in the reference (its buffers are host
 * memory already); the session path stages its images and results here. */
APG_HIP_API int apg_hip_host_malloc(apg_hip_ctx ctx, size_t bytes, void** h_out);
APG_HIP_API int apg_hip_host_free(apg_hip_ctx ctx, void* h_ptr);

/* ---- stages --------------------------------------------------------- */

/* gray, |gx|, |gy| and census of one image: replaces
 * CostComputer::build_gray_and_grad + build_census (reference
 * src/cost_computer.cpp:63-122) as called by compute_aux (:124-130).
 * d_src: strided u8 image (stride >= w*channels, channels 1 or 3).
 * d_census32 (census_type 1) or d_census64 (census_type 0); the other may be NULL. */
APG_HIP_API int apg_hip_aux(apg_hip_ctx ctx, const apg_hip_params* p,
                            const uint8_t* d_src, int w, int h, int stride, int channels,
                            uint8_t* d_gray, uint8_t* d_gx, uint8_t* d_gy,
                            uint32_t* d_census32, uint64_t* d_census64);

/* search range with the prior off: replaces PriorEstimator::estimate's range
 * init (reference src/prior_estimator.cpp:414-424). */
APG_HIP_API int apg_hip_range_init(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h,
                                   int16_t* d_dmin, int16_t* d_dmax);

/* dense u16 cost volume: replaces CostComputer::compute_volume (reference
 * src/cost_computer.cpp:139-203). */
APG_HIP_API int apg_hip_cost_volume(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h,
                                    const uint8_t* d_gray_l, const uint8_t* d_gx_l, const uint8_t* d_gy_l,
                                    const uint8_t* d_gray_r, const uint8_t* d_gx_r, const uint8_t* d_gy_r,
                                    const uint32_t* d_census32_l, const uint32_t* d_census32_r,
                                    const uint64_t* d_census64_l, const uint64_t* d_census64_r,
                                    const int16_t* d_dmin, const int16_t* d_dmax, uint16_t* d_cost);

/* right-view WTA on the raw volume: replaces wta_right_from_left_volume
 * (reference src/pipeline.cpp:24-85). */
APG_HIP_API int apg_hip_right_wta(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h,
                                  const uint16_t* d_cost, float* d_disp_right);

/* Cross arms from the left gray: replaces CostAggregator::build_cross_arms
 * (reference src/cost_aggregator.cpp:17-63). Arms are packed one byte each:
 * left | right << 8 | up << 16 | down << 24. */
APG_HIP_API int apg_hip_cross_arms(apg_hip_ctx ctx, const apg_hip_params* p,
                                   const uint8_t* d_gray, int w, int h, uint32_t* d_arms);

/* Cross H+V, in place on d_cost: replaces CostAggregator::aggregate
 * (reference src/cost_aggregator.cpp:141-149 -> aggregate_hv :65-139).
 * d_tmp: caller scratch of w*h*D u16 (the H-pass output T). No-op when
 * p->agg_enable == 0. */
APG_HIP_API int apg_hip_cross_aggregate(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h,
                                        const uint8_t* d_gray, const int16_t* d_dmin, const int16_t* d_dmax,
                                        uint16_t* d_cost, uint16_t* d_tmp);

/* SGM path sums: replaces SgmOptimizer::optimize (reference
 * src/sgm_optimizer.cpp:261-274). Writes a u16 sum volume (d_sum16, invalid =
 * 65535) and, when d_sum32 is not NULL, the reference's u32 layout
 * (invalid = 0xFFFFFFFF) as well. */
APG_HIP_API int apg_hip_sgm_optimize(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h,
                                     const uint8_t* d_gray, const uint16_t* d_cost,
                                     uint16_t* d_sum16, uint32_t* d_sum32);

/* one SGM path added into a u32 accumulator: replaces
 * SgmOptimizer::aggregate_path (reference src/sgm_optimizer.cpp:186-259).
 * D is the volume depth (the accumulator must already hold the previous paths). */
APG_HIP_API int apg_hip_sgm_path(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h, int D,
                                 const uint8_t* d_gray, const uint16_t* d_cost, uint32_t* d_acc,
                                 int dx, int dy);

/* WTA + second best + legacy subpixel: replaces SgmOptimizer::winner_take_all
 * (reference src/sgm_optimizer.cpp:436-555). vol_bits = 16 or 32.
 * d_best / d_second may be NULL. */
APG_HIP_API int apg_hip_wta(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h,
                            const void* d_vol, int vol_bits,
                            const int16_t* d_dmin, const int16_t* d_dmax,
                            float* d_disp, float* d_best, float* d_second);

/* ConfidenceEstimator::estimate (reference src/confidence_estimator.cpp:7-63). */
APG_HIP_API int apg_hip_confidence(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h,
                                   const float* d_disp, const float* d_disp_right,
                                   const uint8_t* d_gx, const uint8_t* d_gy,
                                   const float* d_best, const float* d_second,
                                   float* d_confidence, uint8_t* d_reliable, uint8_t* d_invalid_reason);

/* PostProcessor::left_right_check, in place (reference src/postprocess.cpp:10-40). */
APG_HIP_API int apg_hip_lr_check(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h,
                                 float* d_disp, const float* d_disp_right, uint8_t* d_invalid_reason);

/* PostProcessor::fill_holes, Legacy route (reference src/postprocess.cpp:42-253):
 * reads d_disp, writes d_out (distinct buffers). d_prior may be NULL. */
APG_HIP_API int apg_hip_fill_holes(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h,
                                   const float* d_disp, const uint8_t* d_invalid_reason,
                                   const uint8_t* d_gray, const float* d_prior, float* d_out);

/* PostProcessor::median, valid-only (reference src/postprocess.cpp:255-309):
 * reads d_disp, writes d_out (distinct buffers). */
APG_HIP_API int apg_hip_median(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h,
                               const float* d_disp, float* d_out);

/* PriorEstimator::estimate (reference src/prior_estimator.cpp:414-435):
 * the dense range init, then, when p->prior_enable and enough supports are
 * found, the prior-narrowed per-pixel ranges.  The census-only candidates of
 * extract_supports and the per-pixel interpolation / range of
 * interpolate_prior + apply_search_range run on the device; the support
 * selection and the cell-grid statistics run on the host in between, so the
 * call synchronises the context stream.  d_prior: w*h floats, -1 where there
 * is no prior.  *n_supports (may be NULL): supports kept. */
APG_HIP_API int apg_hip_prior_estimate(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h,
                                       const uint8_t* d_gx_l, const uint8_t* d_gy_l,
                                       const uint32_t* d_census32_l, const uint32_t* d_census32_r,
                                       const uint64_t* d_census64_l, const uint64_t* d_census64_r,
                                       int16_t* d_dmin, int16_t* d_dmax, float* d_prior, int* n_supports);

/* Refiner::refine (reference src/refiner.cpp:14-98), in place on
 * d_disparity, on the device and stream-ordered: the reference's raster sweeps
 * with the std::mt19937(12345) jitter, bit-exact.  The first call on a context
 * generates the engine outputs on the host and uploads them (synchronises);
 * heights up to 4096.  No-op unless p->refine_enable. */
APG_HIP_API int apg_hip_refine(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h,
                               const uint8_t* d_gray_l, const uint8_t* d_gx_l, const uint8_t* d_gy_l,
                               const uint8_t* d_gray_r, const uint8_t* d_gx_r, const uint8_t* d_gy_r,
                               const uint32_t* d_census32_l, const uint32_t* d_census32_r,
                               const uint64_t* d_census64_l, const uint64_t* d_census64_r,
                               const int16_t* d_dmin, const int16_t* d_dmax, const uint8_t* d_reliable,
                               float* d_disparity);

/* ---- whole pipeline -------------------------------------------------- */

/* Optional device intermediates of the pipeline (any may be NULL). Setting
 * d_cost / d_cost_agg / d_sum32 selects the stage-by-stage (materialising)
 * schedule; otherwise the fused schedule runs. */
typedef struct {
    uint16_t* d_cost;        /* raw C  (w*h*D) */
    uint16_t* d_cost_agg;    /* C' after Cross (w*h*D) */
    uint32_t* d_sum32;       /* S in the reference's u32 layout (w*h*D) */
    float* d_disp_right;     /* w*h */
    float* d_disp_wta;       /* w*h, WTA output before post-processing */
    float* d_best;           /* w*h */
    float* d_second;         /* w*h */
    int16_t* d_dmin;         /* w*h search range (prior-narrowed in HighQuality) */
    int16_t* d_dmax;         /* w*h */
    float* d_prior;          /* w*h prior disparity, -1 where there is none */
    float* d_after_refine;   /* w*h disparity after the refiner (= WTA output when
                                the refiner is off), before the LR check */
} apg_hip_intermediates;

/* One support match of the HighQuality prior (reference
 * include/apg_sgm/types.hpp SupportMatch): pixel, disparity, uniqueness. */
typedef struct {
    int x, y;
    float disparity, confidence;
} apg_hip_support;

/* The supports the last prior (apg_hip_pipeline or apg_hip_prior_estimate on
 * this context) selected, in cell order: copies min(count, capacity) entries
 * to out (out may be NULL with capacity 0) and stores the total in *count.
 * No counterpart in the reference's C API; it backs PipelineBuffers::supports
 * (reference include/apg_sgm/buffers.hpp) for the benchmark's metrics. */
APG_HIP_API int apg_hip_last_supports(apg_hip_ctx ctx, apg_hip_support* out, int capacity, int* count);

/* Schedule of apg_hip_pipeline on this context: 0 (default) runs the fused
 * schedule whenever it is eligible (no raw-volume intermediates requested,
 * D even and <= 256, max(P2_base, P1+1) <= 254, npath*(cost bound +
 * P2max) < 65535, at most one Cross iteration, volume < 2 GiB) and the
 * stage-by-stage schedule otherwise; 1 always runs stage by stage.  Both
 * produce identical results. */
APG_HIP_API void apg_hip_set_schedule(apg_hip_ctx ctx, int schedule);

/* StereoMatcher::compute_impl (reference src/pipeline.cpp:435-605) on device
 * inputs and outputs: aux, range (prior when p->prior_enable), cost, right
 * WTA, Cross, SGM, WTA, confidence, refiner (p->refine_enable), LR check,
 * fill (with the prior), median.  The HighQuality preset's packed volume is
 * computed as the dense volume over the prior-narrowed per-pixel ranges,
 * which the reference's own dense == packed tests show to be identical
 * (tests/test_sanity.cpp:83-476).  With the prior or the refiner the call
 * synchronises the context stream (their host steps) and cannot be
 * captured.  Intermediate buffers come from the context workspace.
 * d_left / d_right: strided u8 images. */
APG_HIP_API int apg_hip_pipeline(apg_hip_ctx ctx, const apg_hip_params* p,
                                 const uint8_t* d_left, const uint8_t* d_right,
                                 int w, int h, int stride, int channels,
                                 float* d_disparity, float* d_confidence,
                                 uint8_t* d_invalid_reason, uint8_t* d_reliable,
                                 const apg_hip_intermediates* inter);

/* Timing harness (no counterpart in the reference; its benchmark times
 * StereoMatcher::compute with steady_clock, apps/benchmark_stereo.cpp:537-575):
 * `warmup` untimed pipeline calls, then `steps` calls bracketed by HIP events
 * on the context stream.  *ms_total = device time of the timed steps.  When
 * stage_avg is not NULL, a further `steps` calls run with per-stage events and
 * their per-stage averages are returned there. */
APG_HIP_API int apg_hip_bench_pipeline(apg_hip_ctx ctx, const apg_hip_params* p,
                                       const uint8_t* d_left, const uint8_t* d_right,
                                       int w, int h, int stride, int channels,
                                       float* d_disparity, float* d_confidence,
                                       uint8_t* d_invalid_reason, uint8_t* d_reliable,
                                       int warmup, int steps, double* ms_total,
                                       apg_hip_timing* stage_avg);

/* Per-pair timing (no counterpart in the reference): `warmup` untimed calls,
 * then `steps` back-to-back calls with a HIP event before each and after the
 * last; step_ms[i] = device time of call i on the context stream.  Their
 * percentiles are the per-pair latency at full load. */
APG_HIP_API int apg_hip_bench_pipeline_steps(apg_hip_ctx ctx, const apg_hip_params* p,
                                             const uint8_t* d_left, const uint8_t* d_right,
                                             int w, int h, int stride, int channels,
                                             float* d_disparity, float* d_confidence,
                                             uint8_t* d_invalid_reason, uint8_t* d_reliable,
                                             int warmup, int steps, float* step_ms);

/* Device bytes the workspace of one apg_hip_pipeline call on a (w, h) pair
 * with parameters p needs (fused schedule when eligible); 0 for invalid
 * parameters.  Used for HBM admission.  No counterpart in the reference
 * (StereoWorkspace::allocated_bytes, include/apg_sgm/workspace.hpp:10-75,
 * reports it after the fact). */
APG_HIP_API size_t apg_hip_pipeline_workspace_bytes(const apg_hip_params* p, int w, int h);

/* ---- allocation-free entry point -------------------------------------- */

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): apg_hip_pipeline's fused schedule (Balanced / Fast: no
 * prior or refiner, one Cross iteration, even D <= 256) on the caller's
 * stream, with the whole workspace in a device array the library holds on
 * each GPU it runs on until the process ends: apg_hip_static_scratch_bytes()
 * = 768 MiB (1280x720 D=128 Balanced needs about 705 MiB).  A pair that needs
 * more returns ResourceLimitExceeded (6), any other configuration
 * UnsupportedBackend (5), before anything is launched.  No hipMalloc,
 * hipFree or other allocating call is made.  Inputs and outputs are the
 * caller's, as for apg_hip_pipeline; stream may be NULL (the null stream).
 * Every call on one GPU shares that GPU's array, which keeps the last call's
 * contents: calls that overlap on one GPU (other threads, other streams)
 * race, so serialise them.  Under stream capture the cost LUT of the
 * parameter set must already be cached by an earlier eager call (eight sets
 * are kept).  Errors: the status code, with a message from
 * apg_hip_static_last_error() (per calling thread). */
APG_HIP_API int apg_hip_pipeline_static(const apg_hip_params* p,
                                        const uint8_t* d_left, const uint8_t* d_right,
                                        int w, int h, int stride, int channels,
                                        float* d_disparity, float* d_confidence,
                                        uint8_t* d_invalid_reason, uint8_t* d_reliable,
                                        void* stream);
/* Bytes of the static array one apg_hip_pipeline_static call on (w, h, p)
 * uses; 0 when p is invalid or not on the fused schedule. */
APG_HIP_API size_t apg_hip_pipeline_static_bytes(const apg_hip_params* p, int w, int h);
APG_HIP_API size_t apg_hip_static_scratch_bytes(void);
APG_HIP_API const char* apg_hip_static_last_error(void);

/* ---- launch shapes ----------------------------------------------------- */

/* Launch shapes of the fused schedule; 0 keeps the built-in choice.  They
 * change how work is split across waves, never the arithmetic: outputs are
 * identical for every setting.  No counterpart in the reference (its CPU
 * stages have no launch shapes). */
typedef struct {
    int cost_segment;    /* output columns per wave of the cost + Cross-H pass: 32..1024 (64) */
    int sgm_dual;        /* SGM wave shape: 0 built-in (D <= 128: both directions of a
                            scanline per wave, 32 lanes each; else one direction per
                            wave), 1 both directions over 64 lanes each, 2 one
                            direction per wave, 3 both directions over 32 lanes each,
                            4 (4 paths, D <= 128) both directions of two scanlines
                            per wave, 16 lanes each */
    int cross_v_waves;   /* target wave count of the Cross-V pass: 256..65536 (10240) */
} apg_hip_tuning;

/* NULL restores the built-in choice.  Applies to later pipeline calls. */
APG_HIP_API int apg_hip_set_tuning(apg_hip_ctx ctx, const apg_hip_tuning* t);

/* Hand-off wait of the refiner's banded sweep, in wall-clock ticks of the
 * device (0: 2 seconds; negative: give up at the first wait).  A band whose
 * wait gives up ends its sweep, the bands after it give up in turn, and the
 * pipeline call returns APG_STATUS_INTERNAL_ERROR (8).  No counterpart in
 * the reference (its refiner runs on one host thread); for tests of the
 * give-up path. */
APG_HIP_API int apg_hip_set_refine_wait(apg_hip_ctx ctx, long long ticks);
APG_HIP_API int apg_hip_get_tuning(apg_hip_ctx ctx, apg_hip_tuning* out);

/* ---- batched form ------------------------------------------------------ */

/* One stereo pair of a batch: device inputs and outputs, as for
 * apg_hip_pipeline.  No counterpart in the reference, which processes one
 * pair per StereoMatcher::compute call (include/apg_sgm/pipeline.hpp). */
typedef struct {
    apg_hip_params params;
    const uint8_t* d_left;
    const uint8_t* d_right;
    int width, height, stride, channels;
    float* d_disparity;
    float* d_confidence;
    uint8_t* d_invalid_reason;
    uint8_t* d_reliable;
} apg_hip_pair;

typedef struct apg_hip_batch_s* apg_hip_batch;

/* A batch executor on `device` with `slots` concurrent pipelines (1..16;
 * 0 = 4), each with its own stream and workspace, joined to one origin
 * stream. */
APG_HIP_API int apg_hip_batch_create(int device, int slots, apg_hip_batch* out);
APG_HIP_API void apg_hip_batch_destroy(apg_hip_batch b);
APG_HIP_API void* apg_hip_batch_stream(apg_hip_batch b);   /* origin hipStream_t */
APG_HIP_API int apg_hip_batch_synchronize(apg_hip_batch b);
APG_HIP_API size_t apg_hip_batch_workspace_bytes(apg_hip_batch b);
APG_HIP_API const char* apg_hip_batch_last_error(apg_hip_batch b);

#define APG_HIP_BATCH_GRAPH 1      /* replay the batch as one captured hipGraph */
#define APG_HIP_BATCH_PER_PAIR 2   /* no group launches: each pair its own kernels */

/* Runs every pair, stream-ordered after earlier work on the origin stream;
 * later work on it waits for all pairs.  Pairs are spread over the slots by
 * longest-processing-time order of W*H*D, each with the launch shapes its
 * size bucket was tuned to.  Unless APG_HIP_BATCH_PER_PAIR is set, the pairs
 * of a slot that take the fused column-range schedule (Fast, Balanced) run
 * as group launches: each step of the schedule is one launch per kernel
 * instance whose grid spans every pair of the slot's chunk that uses that
 * instance (chunks of up to 1 GiB of workspace, pairs of up to 32 M states
 * by default; each pair of a chunk has its own workspace, kept by the
 * batch); the other pairs run one at a time on the slot before them.  With APG_HIP_BATCH_GRAPH the whole batch is one
 * hipGraphLaunch whenever the same pair list ran before on this batch: a new
 * list runs directly once (sizing workspaces and LUTs), is then captured,
 * and replays from the next call on.  HighQuality pairs run one at a time on
 * their slot; in a replayed graph a refiner hand-off wait that gave up is
 * reported by apg_hip_batch_synchronize.  Results are identical to
 * apg_hip_pipeline on each pair alone. */
APG_HIP_API int apg_hip_batch_run(apg_hip_batch b, const apg_hip_pair* pairs, int n, int flags);
/* Group launches: workspace bytes of the pairs one slot runs as one chunk
 * (0: 1 GiB), and the largest pair (W*H*D states) that joins them; larger
 * pairs fill the device alone and run with their own kernels (0: 32 M states).
 * Applies from the next new pair list. */
APG_HIP_API int apg_hip_batch_set_grouping(apg_hip_batch b, size_t chunk_bytes, size_t max_pair_states);
/* Of the last pair list run with group launches: pairs in group launches,
 * the group launches, and the pairs that ran one at a time. */
APG_HIP_API int apg_hip_batch_group_stats(apg_hip_batch b, int* grouped_pairs, int* group_launches,
                                          int* single_pairs);

/* Autotuner: for each size bucket (W/256, H/128, D, paths, Cross on/off,
 * census) of `pairs` not yet tuned, times the candidate launch shapes of
 * each stage on the bucket's first pair (its outputs are overwritten with
 * the same values) and keeps the fastest.  With cache_path, known buckets
 * are read from that JSON file first and every bucket is written back. */
APG_HIP_API int apg_hip_batch_autotune(apg_hip_batch b, const apg_hip_pair* pairs, int n,
                                       const char* cache_path);
/* Launch shapes the batch uses for a pair of this size and mode. */
APG_HIP_API int apg_hip_batch_tuning_for(apg_hip_batch b, const apg_hip_params* p, int w, int h,
                                         apg_hip_tuning* out);

/* ---- serving form ------------------------------------------------------ */

/* One request in host memory: inputs are copied in and results copied out
 * by the server.  Field meaning as apg_hip_pair. */
typedef struct {
    apg_hip_params params;
    const uint8_t* left;
    const uint8_t* right;
    int width, height, stride, channels;
    float* disparity;            /* w*h, required */
    float* confidence;           /* w*h or NULL */
    uint8_t* invalid_reason;     /* w*h or NULL */
    uint8_t* reliable;           /* w*h or NULL */
} apg_hip_request;

typedef struct apg_hip_server_s* apg_hip_server;

/* Continuous-batching server: `slots` pipelines (1..16; 0 = 4) on their own
 * streams, each with a device-resident workspace that is kept between
 * requests.  A request is admitted into a free slot when the device bytes
 * of all slot workspaces stay within hbm_budget_bytes (0 = 64 GiB). */
APG_HIP_API int apg_hip_server_create(int device, int slots, size_t hbm_budget_bytes, apg_hip_server* out);
APG_HIP_API void apg_hip_server_destroy(apg_hip_server s);
APG_HIP_API const char* apg_hip_server_last_error(apg_hip_server s);

/* Queues a request (FIFO).  ResourceLimitExceeded when the pair alone does
 * not fit the budget.  Buffers must stay valid until its tag is returned. */
APG_HIP_API int apg_hip_server_submit(apg_hip_server s, const apg_hip_request* r, uint64_t tag);
/* Non-blocking: retires finished requests (results are then in the host
 * buffers) and admits queued ones into free slots.  Writes up to max_done
 * finished tags and statuses; returns their count, or -1 on error. */
APG_HIP_API int apg_hip_server_poll(apg_hip_server s, uint64_t* done_tags, int* done_status, int max_done);
/* Requests queued or in flight. */
APG_HIP_API int apg_hip_server_pending(apg_hip_server s);
/* Device bytes held by the slot workspaces and staging buffers. */
APG_HIP_API size_t apg_hip_server_resident_bytes(apg_hip_server s);

/* Open-loop replay: request i is submitted when arrival_s[i] seconds have
 * passed since the call began, and the loop polls until all are done.
 * latency_s[i] = completion observed by the host - arrival.  Returns 0 or
 * the first failing request's status. */
APG_HIP_API int apg_hip_server_replay(apg_hip_server s, const apg_hip_request* reqs, const double* arrival_s, int n,
                                      double* latency_s);

/* ---- synthetic inputs ------------------------------------------------ */

/* Seeded synthetic stereo pair (host memory).  kind 0: noise texture, flat
 * background + 4 rectangles; kind 1: + background ramp, 8 rectangles and a
 * right-image gain/bias; kind 2: smooth slanted planes + sinusoid.
 * gt may be NULL. Integer-only PRNG (splitmix64). */
APG_HIP_API int apg_synth_pair(int kind, int w, int h, int max_disparity, uint64_t seed,
                               uint8_t* left, uint8_t* right, float* gt);

APG_HIP_API const char* apg_hip_version_string(void);

#ifdef __cplusplus
}
#endif

#endif /* APG_SGM_HIP_H */
