// Host-side launchers of the gfx950 kernels (internal; the exported C ABI is
// include/apg_sgm_hip.h). All launchers enqueue on `s` and return
// hipSuccess or the launch error.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/apg_sgm_hip.h"
#include "group.hpp"

// Library-internal: marks a context whose stream is being captured into a
// hipGraph, so the pipeline refuses workspace growth and LUT uploads.
void apgk_ctx_set_capturing(apg_hip_ctx ctx, bool on);
// Library-internal: device bytes a pipeline call on (p, w, h) would add to
// the context's workspace.
size_t apgk_ctx_growth_bytes(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h);

// Library-internal, the batched form's group launches (launches that span
// pairs): a context on an existing stream, which it does not own and which
// has no stage timing; sizing a context's workspace and cost LUT for a pair
// of the fused column-range schedule (not eligible: nonzero status); and the
// launches of one of the schedule's kGroupSteps steps for that pair,
// recorded into a sink (a launch with no group form: nonzero status).
constexpr int kGroupSteps = 7;
namespace apgk {
struct LaunchSink;
}
int apgk_ctx_create_on(int device, hipStream_t st, apg_hip_ctx* out);
// Library-internal: reads and clears the context's sticky refiner status
// (after its stream's work is done): 0, or 8 when a banded sweep's hand-off
// wait gave up since the last read.
int apgk_ctx_take_status(apg_hip_ctx ctx);
int apgk_group_prepare(apg_hip_ctx ctx, const apg_hip_pair& q);
int apgk_group_record(apg_hip_ctx ctx, const apg_hip_pair& q, int step, apgk::LaunchSink* sink);

namespace apgk {

// Cost LUTs built on the host with the same libm expf as the reference
// (reference src/discrete_cost_lut.hpp:10-38): census[65], ad[256], grad[511].
struct CostLut {
    float census[65];
    float ad[256];
    float grad[511];
    float scale;
    int pad[1];
    // ceil(2^31 / n) for the window sizes n = 1 .. 2 * kMaxArm + 1 (entry 0 unused):
    // floor(s / n) = umulhi(2s, mag31[n]) for s * n < 2^31
    uint32_t mag31[2 * 62 + 2];
};

// Launch-shape choices of the fused schedule (0 = default): output columns per
// wave of the cost + Cross-H pass, one or both directions of a scanline per
// SGM wave, target wave count of the Cross-V pass.
struct KTune {
    int cost_segment = 0;
    int sgm_dual = 0;   // 0 built-in, 1 both chains on 64 lanes each, 2 one chain per wave, 3 on 32 lanes
                        // each, 4 both chains of two scanlines on 16 lanes each
    int cross_v_waves = 0;
};

void build_cost_lut(const apg_hip_params& p, CostLut* out);

// Device-side limits of this build (checked by apg_hip_validate_params).
constexpr int kMaxArm = 62;           // Cross arm length; window <= 125 columns/rows
static_assert(sizeof(CostLut::mag31) / sizeof(uint32_t) == 2 * kMaxArm + 2, "one magic per window size");
constexpr int kMaxD = 1024;           // disparity levels
constexpr int kMaxSgmK = 4;           // packed u16 pairs per lane in the SGM scan (D <= 512)

hipError_t launch_aux(hipStream_t s, const apg_hip_params& p, const uint8_t* src, int w, int h, int stride,
                      int ch, uint8_t* gray, uint8_t* gx, uint8_t* gy, uint32_t* c32, uint64_t* c64,
                      uint32_t* keys = nullptr);
// one view's inputs and aux outputs
struct AuxViews {
    const uint8_t* src;
    uint8_t *gray, *gx, *gy;
    uint32_t* c32;
    uint64_t* c64;
};
// both views in one launch; also the dense range init (left), the cleared
// right-WTA keys (rkeys may be null) and, where arms is not null, the left
// view's Cross arms (only where aux_pair_builds_arms(p))
hipError_t launch_aux_pair(hipStream_t s, const apg_hip_params& p, int w, int h, int stride, int ch,
                           const AuxViews& L, const AuxViews& R, uint32_t* rkeys, int16_t* dmin, int16_t* dmax,
                           uint32_t* arms, LaunchSink* sink = nullptr);
bool aux_pair_builds_arms(const apg_hip_params& p);
hipError_t launch_range_init(hipStream_t s, const apg_hip_params& p, int w, int h, int16_t* dmin, int16_t* dmax);
hipError_t launch_cost(hipStream_t s, const apg_hip_params& p, const CostLut* d_lut, int w, int h,
                       const uint8_t* gl, const uint8_t* gxl, const uint8_t* gyl,
                       const uint8_t* gr, const uint8_t* gxr, const uint8_t* gyr,
                       const uint32_t* c32l, const uint32_t* c32r, const uint64_t* c64l, const uint64_t* c64r,
                       const int16_t* dmin, const int16_t* dmax, uint16_t* cost);
hipError_t launch_right_wta(hipStream_t s, const apg_hip_params& p, int w, int h, const uint16_t* cost, float* dr);
hipError_t launch_cross_arms(hipStream_t s, const apg_hip_params& p, const uint8_t* gray, int w, int h, uint32_t* arms);
hipError_t launch_cross_h(hipStream_t s, const apg_hip_params& p, int w, int h, const int16_t* dmin,
                          const int16_t* dmax, const uint32_t* arms, const uint16_t* cost, uint16_t* tmp);
hipError_t launch_cross_v(hipStream_t s, const apg_hip_params& p, int w, int h, const int16_t* dmin,
                          const int16_t* dmax, const uint32_t* arms, const uint16_t* tmp, uint16_t* cost);
// One SGM direction over every scanline of that direction. mode 0: S = L
// (first path), 1: S = sat(S + L).
hipError_t launch_sgm_dir(hipStream_t s, const apg_hip_params& p, int w, int h, int D, const uint8_t* gray,
                          const uint16_t* cost, uint16_t* sum, int dx, int dy, int mode);
// One SGM direction added into a u32 accumulator with the reference's
// invalid rule (used by the per-path class API).
// One path added into a u32 accumulator, the path values computed into the
// caller's u16 scratch L (w * h * D); needs cost_max + max(P2, P1 + 1) < 65535.
hipError_t launch_sgm_dir32_ws(hipStream_t s, const apg_hip_params& p, int w, int h, int D, const uint8_t* gray,
                               const uint16_t* cost, uint32_t* acc, int dx, int dy, uint16_t* L);
hipError_t launch_sgm_dir32(hipStream_t s, const apg_hip_params& p, int w, int h, int D, const uint8_t* gray,
                            const uint16_t* cost, uint32_t* acc, int dx, int dy);
hipError_t launch_widen16(hipStream_t s, size_t n, const uint16_t* in, uint32_t* out);
hipError_t launch_wta(hipStream_t s, const apg_hip_params& p, int w, int h, const void* vol, int bits,
                      const int16_t* dmin, const int16_t* dmax, float* disp, float* best, float* second);
hipError_t launch_confidence(hipStream_t s, const apg_hip_params& p, int w, int h, const float* disp,
                             const float* dr, const uint8_t* gx, const uint8_t* gy, const float* best,
                             const float* second, float* conf, uint8_t* reliable, uint8_t* reason);
hipError_t launch_lr_check(hipStream_t s, const apg_hip_params& p, int w, int h, float* disp, const float* dr,
                           uint8_t* reason);
hipError_t launch_fill(hipStream_t s, const apg_hip_params& p, int w, int h, const float* disp, const uint8_t* reason,
                       const uint8_t* gray, const float* prior, float* out);
// median_radius 0 copies disp to out.
constexpr int kMaxMedianRadius = 64;
hipError_t launch_median(hipStream_t s, const apg_hip_params& p, int w, int h, const float* disp, float* out);
// fill_holes then median in one launch (disp and out distinct), for median
// radius 0 or 1 and max_fill_gap <= 62 (fill_median_fusable)
bool fill_median_fusable(const apg_hip_params& p);
hipError_t launch_fill_median(hipStream_t s, const apg_hip_params& p, int w, int h, const float* disp,
                              const uint8_t* reason, const uint8_t* gray, const float* prior, float* out,
                              LaunchSink* sink = nullptr);

// ---- fused schedule ----
// Launchers given a LaunchSink record the launches that group launches
// support (the column-range forms) instead of enqueuing them; the others are
// enqueued on s as without one.
// Largest quantised cost any state can take (bound from the LUT maxima);
// <= 254 allows u8 storage of T and C' with 255 as the invalid value.
int cost_upper_bound(const CostLut& lut, int cost_max);
// Every LUT term >= 0 and the largest quantised sum <= cost_max: the clamp of
// the quantised cost to [0, cost_max] never changes a value.
bool cost_clamp_free(const CostLut& lut, int cost_max);
// column_ranges: every pixel's range depends on its column only (dense range init)
hipError_t launch_cross_v_t(hipStream_t s, const apg_hip_params& p, int w, int h, const int16_t* dmin,
                            const int16_t* dmax, const uint32_t* arms, const void* tmp, void* cost, bool u8,
                            bool column_ranges, const KTune& tn = KTune(), LaunchSink* sink = nullptr);
hipError_t launch_fill_u32(hipStream_t s, size_t n, uint32_t v, uint32_t* p);
hipError_t launch_cost_crossh(hipStream_t s, const apg_hip_params& p, const CostLut* d_lut, int w, int h,
                              const uint8_t* gl, const uint8_t* gxl, const uint8_t* gyl, const uint8_t* gr,
                              const uint8_t* gxr, const uint8_t* gyr, const uint32_t* c32l, const uint32_t* c32r,
                              const uint64_t* c64l, const uint64_t* c64r, const int16_t* dmin, const int16_t* dmax,
                              const uint32_t* arms, void* out, bool out_u8, uint32_t* rkeys,
                              bool column_ranges, bool clamp_free = false,
                              const KTune& tn = KTune(), LaunchSink* sink = nullptr);
hipError_t launch_rwta_finalize(hipStream_t s, const apg_hip_params& p, const CostLut* d_lut, int w, int h,
                                const uint8_t* gl, const uint8_t* gxl, const uint8_t* gyl, const uint8_t* gr,
                                const uint8_t* gxr, const uint8_t* gyr, const uint32_t* c32l, const uint32_t* c32r,
                                const uint64_t* c64l, const uint64_t* c64r, const int16_t* dmin, const int16_t* dmax,
                                const uint32_t* rkeys, float* dr, LaunchSink* sink = nullptr);
// One u8 delta plane per SGM direction (2 per scanline family: forward, backward)
struct PlaneSet {
    uint8_t* p[8];
    // 0: the dense range init, C'(x, d) valid iff d <= x < W + d, so whole
    // batches of columns are known valid; 1: per-pixel ranges (the prior's),
    // validity only from C'
    int general_ranges;
    // bit f set: scanline family f (0 rows, 1 columns, 2 and 3 the diagonals)
    // has one plane, the sum of its two directions' deltas (both <= P2 <= 127,
    // so the sum fits u8), followed by kPlaneSpare bytes the SGM may
    // overwrite; otherwise two, one per direction.  Planes are listed by
    // family, the two of a plain family forward then backward.
    int paired;
};
template <>
struct GroupByRef<PlaneSet> : std::true_type {};
constexpr size_t kPlaneSpare = 256;
// Refiner (reference src/refiner.cpp:14-98) on the device.  The engine is
// the std::mt19937(12345) output stream; it must hold (iterations + 1) *
// (w * h + kRefineSlack) outputs.
constexpr size_t kRefineSlack = 4096;   // engine outputs a sweep may reject beyond its draws
constexpr int kMaxRefineIterations = 64;
constexpr int kMaxRefineRadius = 1024;
struct RefineInputs {
    const uint8_t *gl, *gxl, *gyl, *gr, *gxr, *gyr;
    const uint32_t *c32l, *c32r;
    const uint64_t *c64l, *c64r;
};
size_t refine_scratch_bytes(int w, int h);
// engine outputs the refiner of p needs at w x h
inline size_t refine_engine_len(const apg_hip_params& p, int w, int h) {
    return (size_t)(p.refine_iterations + 1) * ((size_t)w * h + kRefineSlack);
}
// The banded sweep (64-row bands, one wave each, hand-offs through HBM with
// a bounded wait) when enable is set and it applies; wait_ticks: wall-clock
// ticks a hand-off may wait, < 0 gives up at the first wait (for tests).  A
// wait that gives up sets *status to 1 (never clears it; the caller reads and
// clears it); *banded_out tells whether the banded sweep ran.
struct RefineBanding {
    int enable;
    int max_bands;          // bands that can be resident together (one per CU)
    long long wait_ticks;
};
hipError_t launch_refine(hipStream_t st, const apg_hip_params& p, const CostLut* d_lut, int w, int h,
                         const RefineInputs& in, const int16_t* dmin, const int16_t* dmax, const uint8_t* reliable,
                         float* disp, const uint32_t* eng, size_t eng_len, void* scratch, size_t scratch_bytes,
                         const RefineBanding& bd, uint32_t* status, bool* banded_out);
// the scanline families (bit f: 0 rows, 1 columns, 2 and 3 the diagonals)
// whose two directions launch_sgm_lines sums into one plane (the half-wave
// form with max(P2, P1 + 1) <= 127), 0 for none
int sgm_pair_mask(const apg_hip_params& p, int D, bool cost_u8, const KTune& tn);
hipError_t launch_sgm_lines(hipStream_t s, const apg_hip_params& p, int w, int h, int D, const uint8_t* gray,
                            const void* cost, bool cost_u8, const PlaneSet& planes, const KTune& tn = KTune(),
                            LaunchSink* sink = nullptr);
hipError_t launch_wta_fused(hipStream_t s, const apg_hip_params& p, int w, int h, const void* cost, bool cost_u8,
                            const PlaneSet& planes, const int16_t* dmin, const int16_t* dmax, const float* dr,
                            const uint8_t* gx, const uint8_t* gy, float* disp, float* conf, uint8_t* reason,
                            uint8_t* reliable, float* disp_wta, float* best, float* second, uint32_t* sum32,
                            LaunchSink* sink = nullptr);

// HighQuality prior (k_prior.hip): per-sample census candidates; the
// support selection, cell statistics and hole filling; the per-pixel prior
// and search range.  The work buffer holds, per cell, its candidates (at
// most per_cell, raster order: disparity, weight, sample index), their count
// and how many became supports; the grid (5 planes of gw*gh floats:
// disparity, confidence, spread, lowest, highest) and its second buffer for
// the filling sweeps; sort scratch; the number of supports.
struct KeyW;
struct PriorLayout {
    int gw, gh, ncell, per_cell;
    float *grid, *grid2;
    int *cnt, *taken;
    float *key, *wt;
    int* sidx;
    KeyW* tmp;
    int* nsel;
    size_t bytes;
};
int prior_per_cell(int ncell, int budget);
PriorLayout prior_layout(int w, int h, int budget, void* work);
size_t prior_work_bytes(int w, int h, int budget);
hipError_t launch_prior_candidates(hipStream_t s, const apg_hip_params& p, int w, int h, const uint8_t* gx,
                                   const uint8_t* gy, const uint32_t* c32l, const uint32_t* c32r,
                                   const uint64_t* c64l, const uint64_t* c64r, int16_t* cand_d, float* cand_u);
hipError_t launch_prior_grid(hipStream_t s, const apg_hip_params& p, int w, int h, const int16_t* cand_d,
                             const float* cand_u, const PriorLayout& L);
hipError_t launch_prior_ranges(hipStream_t s, const apg_hip_params& p, int w, int h, const PriorLayout& L,
                               int16_t* dmin, int16_t* dmax, float* d_prior);

}  // namespace apgk
