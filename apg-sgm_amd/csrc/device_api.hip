// extern "C" device layer (include/apg_sgm_hip.h): context, workspace,
// stage entry points and the device pipeline schedule.
#include <string.h>

#include <mutex>
#include <random>

#include <string>
#include <vector>

#include "device_common.hpp"
#include "hq_host.hpp"
#include "kernels.hpp"

using namespace apgk;

namespace {

struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

enum BufId {
    B_SRC_L, B_SRC_R, B_GRAY_L, B_GX_L, B_GY_L, B_GRAY_R, B_GX_R, B_GY_R, B_CEN_L, B_CEN_R,
    B_DMIN, B_DMAX, B_ARMS, B_COST, B_TMP, B_SUM, B_DR, B_BEST, B_SECOND, B_SCRATCH, B_LUT,
    B_RKEYS, B_DS0, B_DS1, B_DS2, B_DS3, B_DS4, B_DS5, B_DS6, B_DS7, B_PRIOR, B_CAND_D, B_CAND_U, B_GRID,
    B_REFINE, B_ENGINE, B_COUNT
};

enum StageId { S_AUX, S_PRIOR, S_COST, S_RWTA, S_CROSS, S_SGM, S_WTA, S_CONF, S_REFINE, S_POST, S_COUNT };

// The parameters build_cost_lut reads.
struct LutKey {
    float lc, la, lg, eta, mu;
    int use_ad, use_grad, cost_max;
    static LutKey of(const apg_hip_params& p) {
        LutKey k;
        memset(&k, 0, sizeof k);
        k.lc = p.lambda_census; k.la = p.lambda_ad; k.lg = p.lambda_grad; k.eta = p.eta_ad; k.mu = p.mu_grad;
        k.use_ad = p.use_ad; k.use_grad = p.use_grad; k.cost_max = p.cost_max;
        return k;
    }
    bool operator==(const LutKey& o) const { return memcmp(this, &o, sizeof o) == 0; }
};
constexpr int kLutSlots = 8;

}  // namespace

struct apg_hip_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = true;   // false: the stream belongs to a batch slot
    DevBuf bufs[B_COUNT];
    std::string last_error;
    bool timing = false;
    hipEvent_t ev[S_COUNT + 1] = {};
    apg_hip_timing last{};
    // cost LUTs of the last kLutSlots distinct LUT parameter sets: device
    // copies side by side in B_LUT, host copies kept for the cost bound
    CostLut lut_host[kLutSlots];
    LutKey lut_key[kLutSlots];
    bool lut_used[kLutSlots] = {};
    int lut_next = 0, lut_cur = -1;
    int schedule = 0;   // 0 automatic (fused when eligible), 1 stage-by-stage
    KTune tune;
    bool capturing = false;   // stream capture in progress: no allocation, no host copy
    // the last prior's supports: read back from its device layout on the
    // first request after it ran
    std::vector<apg_hip_support> last_supports;
    bool supports_pending = false;
    PriorLayout prior_lay{};
    int prior_sw = 0;
    size_t engine_len = 0;   // std::mt19937(12345) outputs held in B_ENGINE
    // the refiner's banded sweep: resident bands (CUs), wall-clock ticks per
    // second, and the test override of its hand-off wait (0: 2 s)
    int cu_count = 0;
    long long wall_hz = 0;
    long long refine_wait = 0;
    // sticky status of the refiner's banded sweep: 1 once a hand-off wait gave
    // up; read and cleared after a direct run, or by the batch's synchronize
    // after a replayed graph
    uint32_t* d_status = nullptr;
};

namespace {

int fail(apg_hip_ctx ctx, int code, const std::string& msg) {
    if (ctx) ctx->last_error = msg;
    return code;
}

int hip_fail(apg_hip_ctx ctx, hipError_t e, const char* where) {
    if (e == hipSuccess) return 0;
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? 7 : 8;
    return fail(ctx, code, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIPCHK(expr, where)                              \
    do {                                                 \
        hipError_t _e = (expr);                          \
        if (_e != hipSuccess) return hip_fail(ctx, _e, where); \
    } while (0)

// Grow-only device buffer of the workspace.
hipError_t ensure(apg_hip_ctx ctx, BufId id, size_t bytes) {
    DevBuf& b = ctx->bufs[id];
    if (b.bytes >= bytes && b.ptr) return hipSuccess;
    if (ctx->capturing) return hipErrorStreamCaptureUnsupported;
    if (b.ptr) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return e;
        (void)hipFree(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    hipError_t e = hipMalloc(&b.ptr, bytes < 256 ? 256 : bytes);
    if (e == hipSuccess) b.bytes = bytes;
    return e;
}

template <typename T>
T* buf(apg_hip_ctx ctx, BufId id) { return (T*)ctx->bufs[id].ptr; }

// Selects the device LUT of p's cost parameters, building and uploading it
// into the least recently built slot when it is not cached.  The upload is
// stream-ordered, so kernels still reading the slot's previous table finish
// first.
int ensure_lut(apg_hip_ctx ctx, const apg_hip_params* p) {
    const LutKey key = LutKey::of(*p);
    for (int i = 0; i < kLutSlots; ++i)
        if (ctx->lut_used[i] && ctx->lut_key[i] == key) {
            ctx->lut_cur = i;
            return 0;
        }
    if (ctx->capturing) return fail(ctx, 8, "cost LUT not cached before stream capture");
    HIPCHK(ensure(ctx, B_LUT, kLutSlots * sizeof(CostLut)), "lut alloc");
    const int i = ctx->lut_next;
    ctx->lut_next = (i + 1) % kLutSlots;
    // an evicted slot's host copy may still be the source of a queued upload
    if (ctx->lut_used[i]) HIPCHK(hipStreamSynchronize(ctx->stream), "lut sync");
    build_cost_lut(*p, &ctx->lut_host[i]);
    HIPCHK(hipMemcpyAsync(buf<CostLut>(ctx, B_LUT) + i, &ctx->lut_host[i], sizeof(CostLut), hipMemcpyHostToDevice,
                          ctx->stream),
           "lut upload");
    ctx->lut_key[i] = key;
    ctx->lut_used[i] = true;
    ctx->lut_cur = i;
    return 0;
}

const CostLut* cur_lut(apg_hip_ctx ctx) { return buf<CostLut>(ctx, B_LUT) + ctx->lut_cur; }

void mark(apg_hip_ctx ctx, int stage) {
    if (ctx->timing) (void)hipEventRecord(ctx->ev[stage], ctx->stream);
}

// Per-stage times from the stage events of the last pipeline call.
int finish_timing(apg_hip_ctx ctx) {
    if (ctx->timing) {
        HIPCHK(hipEventSynchronize(ctx->ev[S_COUNT]), "event sync");
        float ms[S_COUNT];
        for (int i = 0; i < S_COUNT; ++i) (void)hipEventElapsedTime(&ms[i], ctx->ev[i], ctx->ev[i + 1]);
        float tot = 0.f;
        (void)hipEventElapsedTime(&tot, ctx->ev[0], ctx->ev[S_COUNT]);
        apg_hip_timing& t = ctx->last;
        t.aux_ms = ms[S_AUX]; t.prior_ms = ms[S_PRIOR]; t.cost_ms = ms[S_COST]; t.right_wta_ms = ms[S_RWTA];
        t.cross_ms = ms[S_CROSS]; t.sgm_ms = ms[S_SGM]; t.wta_ms = ms[S_WTA]; t.confidence_ms = ms[S_CONF];
        t.refine_ms = ms[S_REFINE]; t.post_ms = ms[S_POST]; t.total_ms = tot;
    }
    return 0;
}

// The path sums fit u16 (reference include/apg_sgm/sgm_optimizer.hpp:9-18);
// beyond that bound they are accumulated in u32, as the reference's dense
// optimize always does (src/sgm_optimizer.cpp:261-274).
bool u16_path_sums(const apg_hip_params& p) {
    const long p2max = p.P2_base > p.P1 + 1 ? p.P2_base : p.P1 + 1;
    const long npath = p.paths == 8 ? 8 : 4;
    return npath * ((long)p.cost_max + p2max) < 65535;
}

// Workspace of one apg_hip_pipeline call: which schedule runs and the bytes
// of each buffer it needs.
struct WsPlan {
    bool fused = false, u8 = false;
    // stage-by-stage Cost and Cross, then the fused schedule's SGM (delta
    // planes) and WTA kernels over the per-pixel ranges: the HighQuality
    // preset, whose refiner sits between the confidence and the LR check
    bool planes_sgm = false;
    // planes_sgm with u8 T / C': the general-range forms of the fused Cost and
    // Cross kernels (right-view keys in the cost pass, no u16 volumes)
    bool planes_fused = false;
    // scanline families whose SGM planes hold the sum of their two
    // directions' deltas (one plane instead of two), as PlaneSet::paired
    int pair_mask = 0;
    size_t bytes[B_COUNT] = {};
};

WsPlan plan_workspace(const apg_hip_params& p, int w, int h, const CostLut& lut, int schedule,
                      const apg_hip_intermediates* inter, const KTune& tn = KTune()) {
    WsPlan pl;
    const size_t n = (size_t)w * h;
    const int D = p.max_disparity - p.min_disparity;
    const size_t nv = n * (size_t)D;
    const size_t cen = p.census_type == 1 ? 4 : 8;
    for (BufId b : {B_GRAY_L, B_GX_L, B_GY_L, B_GRAY_R, B_GX_R, B_GY_R}) pl.bytes[b] = n;
    pl.bytes[B_CEN_L] = pl.bytes[B_CEN_R] = n * cen;
    pl.bytes[B_DMIN] = pl.bytes[B_DMAX] = n * 2;
    pl.bytes[B_ARMS] = pl.bytes[B_DR] = pl.bytes[B_SCRATCH] = n * 4;
    pl.bytes[B_LUT] = kLutSlots * sizeof(CostLut);
    // Fused schedule: the raw cost volume never reaches HBM, C' is stored as
    // u8 when every reachable cost is <= 254, and the SGM keeps one u8 delta
    // plane per direction (needs P2max <= 254) instead of u16 path sums.
    const int p2max = p.P2_base > p.P1 + 1 ? p.P2_base : p.P1 + 1;
    const bool want_raw = inter && (inter->d_cost || inter->d_cost_agg);
    const int cbound = cost_upper_bound(lut, p.cost_max);
    pl.u8 = cbound <= 254;
    // the WTA stage rebuilds S in u16: npath * (C' + P2max) must stay below 65535
    const int npaths = p.paths == 8 ? 8 : 4;
    if (p.refine_enable) {
        // the refiner's random stream and its sweep scratch (run_refine)
        pl.bytes[B_ENGINE] = refine_engine_len(p, w, h) * 4;
        pl.bytes[B_REFINE] = refine_scratch_bytes(w, h);
    }
    if (p.prior_enable) {
        const size_t ns = (size_t)prior_samples_x(w) * prior_samples_y(h);
        pl.bytes[B_PRIOR] = n * 4;
        pl.bytes[B_CAND_D] = ns * 2;
        pl.bytes[B_CAND_U] = ns * 4;
        pl.bytes[B_GRID] = prior_work_bytes(w, h, p.prior_max_supports);
    }
    // the fused schedule assumes the dense range init and runs confidence and
    // the LR check in one kernel, so the prior and the refiner (which sits
    // between them) take the stage-by-stage schedule
    pl.fused = schedule != 1 && !want_raw && !p.prior_enable && !p.refine_enable && D % 2 == 0 && D <= 256 &&
               p2max <= 254 &&
               npaths * (cbound + p2max) < 65535 && (!p.agg_enable || p.agg_iterations <= 1) &&
               nv * (pl.u8 ? 1 : 2) < (1ull << 31);
    pl.planes_sgm = !pl.fused && schedule != 1 && !want_raw && (p.prior_enable || p.refine_enable) && D % 2 == 0 &&
                    D <= 256 && p2max <= 254 && npaths * (cbound + p2max) < 65535 && nv * 2 < (1ull << 31);
    pl.planes_fused = pl.planes_sgm && pl.u8 && (!p.agg_enable || p.agg_iterations <= 1) && nv < (1ull << 31);
    pl.pair_mask = (pl.fused || pl.planes_fused) ? sgm_pair_mask(p, D, pl.u8, tn) : 0;
    const int nplanes = npaths - __builtin_popcount((unsigned)pl.pair_mask);
    const size_t spare = pl.pair_mask ? kPlaneSpare : 0;
    if (pl.planes_fused) {
        pl.bytes[B_COST] = nv;
        if (p.agg_enable) pl.bytes[B_TMP] = nv;
        pl.bytes[B_RKEYS] = n * 4;
        for (int r = 0; r < nplanes; ++r) pl.bytes[B_DS0 + r] = nv + spare;
    } else if (pl.fused) {
        pl.bytes[B_COST] = nv * (pl.u8 ? 1 : 2);
        if (p.agg_enable) pl.bytes[B_TMP] = nv * (pl.u8 ? 1 : 2);
        pl.bytes[B_RKEYS] = n * 4;
        for (int r = 0; r < nplanes; ++r) pl.bytes[B_DS0 + r] = nv + spare;
    } else {
        pl.bytes[B_COST] = pl.bytes[B_TMP] = nv * 2;
        if (pl.planes_sgm)
            for (int r = 0; r < npaths; ++r) pl.bytes[B_DS0 + r] = nv;
        else
            pl.bytes[B_SUM] = nv * (u16_path_sums(p) ? 2 : 4);
        pl.bytes[B_BEST] = pl.bytes[B_SECOND] = n * 4;
    }
    return pl;
}

int check_ctx(apg_hip_ctx ctx, const apg_hip_params* p) {
    if (!ctx) return 8;
    if (!p) return fail(ctx, 4, "null params");
    const int s = apg_hip_validate_params(p);
    if (s) return s;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    return 0;
}

// One pair's images and outputs.
struct PairIO {
    const uint8_t *left, *right;
    int w, h, stride, ch;
    float* disp;
    float* conf;
    uint8_t *reason, *reliable;
};

enum FusedStep { F_AUX, F_COST, F_RWTA, F_CROSS, F_SGM, F_WTA, F_POST, F_STEPS };

// Launch step `step` of the fused schedules on the context's workspace
// (allocated for `plan`): F_AUX both views' aux pass (every schedule), then
// for the fused schedules the Cost + Cross-H pass, the right-view WTA,
// Cross-V, the SGM planes (column ranges when plan.fused, the prior's
// per-pixel ranges otherwise), the WTA with confidence (and the LR check when
// plan.fused) and fill + median (plan.fused).  With a sink every launch is
// recorded there and nothing is enqueued; a launch with no group form fails
// the step.
int fused_step(apg_hip_ctx ctx, hipStream_t st, const apg_hip_params* p, const PairIO& io, const WsPlan& plan,
               int step, const apg_hip_intermediates* inter, LaunchSink* sink) {
    if (sink && inter) return fail(ctx, 4, "intermediates have no group launch");
    const int w = io.w, h = io.h;
    const size_t n = (size_t)w * h;
    const int D = p->max_disparity - p->min_disparity;
    const bool fused = plan.fused, u8 = plan.u8, sym = p->census_type == 1;
    const int npaths = p->paths == 8 ? 8 : 4;
    uint8_t *gl = buf<uint8_t>(ctx, B_GRAY_L), *gxl = buf<uint8_t>(ctx, B_GX_L), *gyl = buf<uint8_t>(ctx, B_GY_L);
    uint8_t *gr = buf<uint8_t>(ctx, B_GRAY_R), *gxr = buf<uint8_t>(ctx, B_GX_R), *gyr = buf<uint8_t>(ctx, B_GY_R);
    void *cl = ctx->bufs[B_CEN_L].ptr, *cr = ctx->bufs[B_CEN_R].ptr;
    const uint32_t *c32l = sym ? (const uint32_t*)cl : nullptr, *c32r = sym ? (const uint32_t*)cr : nullptr;
    const uint64_t *c64l = sym ? nullptr : (const uint64_t*)cl, *c64r = sym ? nullptr : (const uint64_t*)cr;
    int16_t *dmin = buf<int16_t>(ctx, B_DMIN), *dmax = buf<int16_t>(ctx, B_DMAX);
    uint32_t* arms = buf<uint32_t>(ctx, B_ARMS);
    float* dr = buf<float>(ctx, B_DR);
    float* scratch = buf<float>(ctx, B_SCRATCH);
    // the fused schedules' right-view WTA keys start cleared, by the aux pass
    uint32_t* rkeys = (fused || plan.planes_fused) ? buf<uint32_t>(ctx, B_RKEYS) : nullptr;
    // the left view's Cross arms come from the aux pass's gray tiles where the arms are short enough
    uint32_t* arms_aux = aux_pair_builds_arms(*p) ? arms : nullptr;
    // fused Cost + Cross: dense range init (column ranges) or, for the
    // HighQuality preset, the prior's per-pixel ranges
    const bool colr = fused;
    PlaneSet planes{};
    planes.general_ranges = colr ? 0 : 1;
    planes.paired = plan.pair_mask;
    for (int r = 0; r < npaths - __builtin_popcount((unsigned)plan.pair_mask); ++r)
        planes.p[r] = buf<uint8_t>(ctx, (BufId)(B_DS0 + r));
    void* cp = ctx->bufs[B_COST].ptr;   // C' (u8 or u16)
    void* tt = ctx->bufs[B_TMP].ptr;    // T  (u8 or u16)
    switch (step) {
    case F_AUX: {
        // both views and the dense range init in one launch
        const AuxViews av_l{io.left, gl, gxl, gyl, sym ? (uint32_t*)cl : nullptr, sym ? nullptr : (uint64_t*)cl};
        const AuxViews av_r{io.right, gr, gxr, gyr, sym ? (uint32_t*)cr : nullptr, sym ? nullptr : (uint64_t*)cr};
        HIPCHK(launch_aux_pair(st, *p, w, h, io.stride, io.ch, av_l, av_r, rkeys, dmin, dmax, arms_aux, sink),
               "aux");
        return 0;
    }
    case F_COST:
        if (p->agg_enable && !arms_aux) {
            if (sink) return fail(ctx, 4, "Cross arms longer than the aux pass builds have no group launch");
            HIPCHK(launch_cross_arms(st, *p, gl, w, h, arms), "arms");
        }
        HIPCHK(launch_cost_crossh(st, *p, cur_lut(ctx), w, h, gl, gxl, gyl, gr, gxr, gyr, c32l, c32r, c64l, c64r,
                                  dmin, dmax, arms, p->agg_enable ? tt : cp, u8, rkeys, colr,
                                  cost_clamp_free(ctx->lut_host[ctx->lut_cur], p->cost_max), ctx->tune, sink),
               "cost_crossh");
        return 0;
    case F_RWTA:
        HIPCHK(launch_rwta_finalize(st, *p, cur_lut(ctx), w, h, gl, gxl, gyl, gr, gxr, gyr, c32l, c32r, c64l, c64r,
                                    dmin, dmax, rkeys, dr, sink),
               "rwta_finalize");
        return 0;
    case F_CROSS:
        if (p->agg_enable)
            HIPCHK(launch_cross_v_t(st, *p, w, h, dmin, dmax, arms, tt, cp, u8, colr, ctx->tune, sink), "cross_v");
        return 0;
    case F_SGM:
        HIPCHK(launch_sgm_lines(st, *p, w, h, D, gl, cp, u8, planes, ctx->tune, sink), "sgm_lines");
        return 0;
    case F_WTA:
        if (fused) {
            // without the refiner, the disparity after it is the WTA output
            float* wta_out = inter ? (inter->d_disp_wta ? inter->d_disp_wta : inter->d_after_refine) : nullptr;
            // fill + median in one pass reads the WTA's disparity from scratch
            float* wta_disp = fill_median_fusable(*p) ? scratch : io.disp;
            HIPCHK(launch_wta_fused(st, *p, w, h, cp, u8, planes, dmin, dmax, dr, gxl, gyl, wta_disp, io.conf,
                                    io.reason, io.reliable, wta_out, inter ? inter->d_best : nullptr,
                                    inter ? inter->d_second : nullptr, inter ? inter->d_sum32 : nullptr, sink),
                   "wta_fused");
            if (inter && inter->d_after_refine && wta_out != inter->d_after_refine)
                HIPCHK(hipMemcpyAsync(inter->d_after_refine, wta_out, n * 4, hipMemcpyDeviceToDevice, st),
                       "copy refined");
        } else {
            // WTA + confidence in one kernel; the LR check follows the refiner
            apg_hip_params pw = *p;
            pw.post_lr_check = 0;
            HIPCHK(launch_wta_fused(st, pw, w, h, cp, u8, planes, dmin, dmax, dr, gxl, gyl, io.disp, io.conf,
                                    io.reason, io.reliable, inter ? inter->d_disp_wta : nullptr,
                                    inter ? inter->d_best : nullptr, inter ? inter->d_second : nullptr,
                                    inter ? inter->d_sum32 : nullptr, sink),
                   "wta_fused");
        }
        if (inter && inter->d_disp_right)
            HIPCHK(hipMemcpyAsync(inter->d_disp_right, dr, n * 4, hipMemcpyDeviceToDevice, st), "copy dr");
        return 0;
    case F_POST:
        if (fill_median_fusable(*p)) {
            HIPCHK(launch_fill_median(st, *p, w, h, scratch, io.reason, gl, nullptr, io.disp, sink), "fill_median");
        } else {
            if (sink) return fail(ctx, 4, "separate fill and median have no group launch");
            HIPCHK(launch_fill(st, *p, w, h, io.disp, io.reason, gl, nullptr, scratch), "fill");
            HIPCHK(launch_median(st, *p, w, h, scratch, io.disp), "median");
        }
        return 0;
    default:
        return fail(ctx, 4, "invalid step");
    }
}

}  // namespace

extern "C" {

void apg_hip_params_from_mode(int mode, int max_disparity, apg_hip_params* c) {
    memset(c, 0, sizeof(*c));
    c->mode = mode;
    c->min_disparity = 0;
    c->max_disparity = max_disparity;
    c->census_type = 1;
    c->lambda_census = 16.f; c->lambda_ad = 30.f; c->lambda_grad = 20.f;
    c->eta_ad = 0.4f; c->mu_grad = 0.3f;
    c->use_ad = 1; c->use_grad = 1; c->cost_max = 255;
    c->agg_enable = 1; c->max_arm_length = 17; c->color_threshold = 12; c->agg_iterations = 1;
    c->paths = 4; c->P1 = 10; c->P2_base = 120; c->adaptive_penalties = 1;
    c->grad_t1 = 8; c->grad_t2 = 24; c->p2_alpha = 0.15f; c->piecewise_p2 = 1;
    c->conf_uniqueness_ratio = 0.15f; c->conf_lr_max_diff = 1;
    c->conf_min_texture = 2.f; c->conf_reliable_threshold = 0.25f;
    c->post_subpixel = 1; c->post_lr_check = 1; c->post_lr_max_diff = 1;
    c->post_median_radius = 1; c->post_max_fill_gap = 32;
    c->prior_enable = 0; c->prior_uniqueness_ratio = 0.4f; c->prior_texture_threshold = 4;
    c->prior_lr_max_diff = 1; c->prior_search_radius = 16; c->prior_min_supports = 32;
    c->prior_max_supports = 4000;
    c->refine_enable = 0; c->refine_iterations = 3; c->refine_random_radius = 4;
    if (mode == 0) {
        c->use_ad = 0; c->use_grad = 0; c->agg_enable = 0; c->paths = 4; c->adaptive_penalties = 0;
    } else if (mode == 1) {
        c->agg_iterations = 1; c->paths = 4; c->adaptive_penalties = 1;
    } else {
        c->paths = 8; c->adaptive_penalties = 1;
        c->prior_enable = 1; c->refine_enable = 1; c->refine_iterations = 3;
    }
}

int apg_hip_validate_params(const apg_hip_params* c) {
    if (!c) return 4;
    // validate_config (reference src/pipeline.cpp:171-186)
    if (c->max_disparity <= c->min_disparity) return 3;
    if (c->cost_max < 0 || c->cost_max > 65535) return 4;
    if (c->P1 < 0) return 4;
    if (c->P2_base <= c->P1) return 4;
    if (c->grad_t1 < 0 || c->grad_t2 < c->grad_t1) return 4;
    // device representability: u16 path sums need the reference's own
    // can_use_u16_sgm_accumulator bound (include/apg_sgm/sgm_optimizer.hpp:9-18);
    // Cross windows and D must fit the LDS layouts.
    const long D = (long)c->max_disparity - c->min_disparity;
    const long p2max = c->P2_base > c->P1 + 1 ? c->P2_base : c->P1 + 1;
    const long npath = c->paths == 8 ? 8 : 4;
    (void)npath;
    // one path's values (at most cost_max + P2max) are computed in u16; their
    // sums go to u32 past the u16 bound
    if ((long)c->cost_max + p2max >= 65535) return 5;
    if (D > kMaxD || D > 2 * 64 * kMaxSgmK) return 5;
    if (c->agg_enable && c->max_arm_length > kMaxArm) return 5;
    // the median's window tile (16 + 2r)^2 floats stays within 160 KB of LDS
    if (c->post_median_radius > kMaxMedianRadius) return 5;
    // the refiner's engine outputs are held on the device: (iterations + 1) per pixel
    if (c->refine_enable && c->refine_iterations > kMaxRefineIterations) return 5;
    // each sweep's window holds kRefineSlack rejected engine outputs beyond its draws;
    // a rejection has probability < (2r + 1) / 2^32, so r <= 1024 keeps a window of
    // 2^31 draws far inside it (about 1000 rejections expected)
    if (c->refine_enable && c->refine_random_radius > kMaxRefineRadius) return 5;
    if (c->paths != 4 && c->paths != 8) return 4;
    return 0;
}

int apg_hip_ctx_create(int device, apg_hip_ctx* out) {
    if (!out) return 8;
    *out = nullptr;
    apg_hip_ctx ctx = new (std::nothrow) apg_hip_ctx_s();
    if (!ctx) return 7;
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    for (int i = 0; e == hipSuccess && i <= S_COUNT; ++i) e = hipEventCreate(&ctx->ev[i]);
    if (e != hipSuccess) {
        apg_hip_ctx_destroy(ctx);
        return (e == hipErrorOutOfMemory) ? 7 : 8;
    }
    *out = ctx;
    return 0;
}

void apg_hip_ctx_destroy(apg_hip_ctx ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& b : ctx->bufs)
        if (b.ptr) (void)hipFree(b.ptr);
    if (ctx->d_status) (void)hipFree(ctx->d_status);
    for (int i = 0; i <= S_COUNT; ++i)
        if (ctx->ev[i]) (void)hipEventDestroy(ctx->ev[i]);
    if (ctx->stream && ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int apg_hip_synchronize(apg_hip_ctx ctx) {
    if (!ctx) return 8;
    HIPCHK(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return 0;
}

void* apg_hip_stream(apg_hip_ctx ctx) { return ctx ? (void*)ctx->stream : nullptr; }

size_t apg_hip_workspace_bytes(apg_hip_ctx ctx) {
    if (!ctx) return 0;
    size_t t = 0;
    for (auto& b : ctx->bufs) t += b.bytes;
    return t;
}

void apg_hip_release_workspace(apg_hip_ctx ctx) {
    if (!ctx) return;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& b : ctx->bufs) {
        if (b.ptr) (void)hipFree(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    for (bool& u : ctx->lut_used) u = false;
    ctx->lut_cur = -1;
    ctx->engine_len = 0;   // B_ENGINE is gone with the rest
}

void apg_hip_enable_stage_timing(apg_hip_ctx ctx, int on) {
    if (ctx) ctx->timing = on != 0;
}

int apg_hip_last_timing(apg_hip_ctx ctx, apg_hip_timing* out) {
    if (!ctx || !out) return 8;
    *out = ctx->last;
    return 0;
}

const char* apg_hip_last_error(apg_hip_ctx ctx) { return ctx ? ctx->last_error.c_str() : "invalid context"; }

int apg_hip_malloc(apg_hip_ctx ctx, size_t bytes, void** d_out) {
    if (!ctx || !d_out) return 8;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    HIPCHK(hipMalloc(d_out, bytes ? bytes : 1), "hipMalloc");
    return 0;
}

int apg_hip_free(apg_hip_ctx ctx, void* d_ptr) {
    if (!ctx) return 8;
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    HIPCHK(hipFree(d_ptr), "hipFree");
    return 0;
}

int apg_hip_host_malloc(apg_hip_ctx ctx, size_t bytes, void** h_out) {
    if (!ctx || !h_out) return 8;
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    HIPCHK(hipHostMalloc(h_out, bytes ? bytes : 1, hipHostMallocDefault), "hipHostMalloc");
    return 0;
}

int apg_hip_host_free(apg_hip_ctx ctx, void* h_ptr) {
    if (!ctx) return 8;
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    HIPCHK(hipHostFree(h_ptr), "hipHostFree");
    return 0;
}

int apg_hip_upload(apg_hip_ctx ctx, void* d_dst, const void* h_src, size_t bytes) {
    if (!ctx) return 8;
    HIPCHK(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream), "upload");
    HIPCHK(hipStreamSynchronize(ctx->stream), "upload sync");
    return 0;
}

int apg_hip_download(apg_hip_ctx ctx, void* h_dst, const void* d_src, size_t bytes) {
    if (!ctx) return 8;
    HIPCHK(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream), "download");
    HIPCHK(hipStreamSynchronize(ctx->stream), "download sync");
    return 0;
}

int apg_hip_memset(apg_hip_ctx ctx, void* d_dst, int value, size_t bytes) {
    if (!ctx) return 8;
    HIPCHK(hipMemsetAsync(d_dst, value, bytes, ctx->stream), "memset");
    HIPCHK(hipStreamSynchronize(ctx->stream), "memset sync");
    return 0;
}

// ---------------------------------------------------------------- stages ----

int apg_hip_aux(apg_hip_ctx ctx, const apg_hip_params* p, const uint8_t* d_src, int w, int h, int stride,
                int channels, uint8_t* d_gray, uint8_t* d_gx, uint8_t* d_gy, uint32_t* d_census32,
                uint64_t* d_census64) {
    if (int s = check_ctx(ctx, p)) return s;
    if (!d_src || w <= 0 || h <= 0) return fail(ctx, 1, "empty input");
    if ((channels != 1 && channels != 3) || stride < w * channels) return fail(ctx, 2, "bad stride/channels");
    if ((p->census_type == 1 && !d_census32) || (p->census_type != 1 && !d_census64)) return fail(ctx, 4, "census buffer");
    HIPCHK(launch_aux(ctx->stream, *p, d_src, w, h, stride, channels, d_gray, d_gx, d_gy, d_census32, d_census64), "aux");
    return 0;
}

int apg_hip_range_init(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h, int16_t* d_dmin, int16_t* d_dmax) {
    if (int s = check_ctx(ctx, p)) return s;
    HIPCHK(launch_range_init(ctx->stream, *p, w, h, d_dmin, d_dmax), "range");
    return 0;
}

int apg_hip_cost_volume(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h, const uint8_t* gl, const uint8_t* gxl,
                        const uint8_t* gyl, const uint8_t* gr, const uint8_t* gxr, const uint8_t* gyr,
                        const uint32_t* c32l, const uint32_t* c32r, const uint64_t* c64l, const uint64_t* c64r,
                        const int16_t* dmin, const int16_t* dmax, uint16_t* d_cost) {
    if (int s = check_ctx(ctx, p)) return s;
    if (int s = ensure_lut(ctx, p)) return s;
    HIPCHK(launch_cost(ctx->stream, *p, cur_lut(ctx), w, h, gl, gxl, gyl, gr, gxr, gyr, c32l, c32r, c64l,
                       c64r, dmin, dmax, d_cost),
           "cost");
    return 0;
}

int apg_hip_right_wta(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h, const uint16_t* d_cost, float* d_dr) {
    if (int s = check_ctx(ctx, p)) return s;
    HIPCHK(launch_right_wta(ctx->stream, *p, w, h, d_cost, d_dr), "right_wta");
    return 0;
}

int apg_hip_cross_arms(apg_hip_ctx ctx, const apg_hip_params* p, const uint8_t* d_gray, int w, int h, uint32_t* d_arms) {
    if (int s = check_ctx(ctx, p)) return s;
    HIPCHK(launch_cross_arms(ctx->stream, *p, d_gray, w, h, d_arms), "arms");
    return 0;
}

int apg_hip_cross_aggregate(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h, const uint8_t* d_gray,
                            const int16_t* dmin, const int16_t* dmax, uint16_t* d_cost, uint16_t* d_tmp) {
    if (int s = check_ctx(ctx, p)) return s;
    if (!p->agg_enable) return 0;
    HIPCHK(ensure(ctx, B_ARMS, (size_t)w * h * 4), "arms alloc");
    uint32_t* arms = buf<uint32_t>(ctx, B_ARMS);
    HIPCHK(launch_cross_arms(ctx->stream, *p, d_gray, w, h, arms), "arms");
    const int iters = p->agg_iterations > 1 ? p->agg_iterations : 1;
    for (int i = 0; i < iters; ++i) {
        HIPCHK(launch_cross_h(ctx->stream, *p, w, h, dmin, dmax, arms, d_cost, d_tmp), "cross_h");
        HIPCHK(launch_cross_v(ctx->stream, *p, w, h, dmin, dmax, arms, d_tmp, d_cost), "cross_v");
    }
    return 0;
}

static const int kDirs[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {1, -1}, {-1, 1}, {-1, -1}};

int apg_hip_sgm_optimize(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h, const uint8_t* d_gray,
                         const uint16_t* d_cost, uint16_t* d_sum16, uint32_t* d_sum32) {
    if (int s = check_ctx(ctx, p)) return s;
    const int D = p->max_disparity - p->min_disparity;
    const int npath = p->paths == 8 ? 8 : 4;
    if (!u16_path_sums(*p)) {
        // u32 accumulation; d_sum16 (w * h * D u16) holds each path's values
        if (!d_sum32 || !d_sum16) return fail(ctx, 4, "u32 path sums need d_sum32 and a u16 scratch in d_sum16");
        const size_t nv = (size_t)w * h * D;
        HIPCHK(launch_fill_u32(ctx->stream, nv, 0u, d_sum32), "sgm clear");
        for (int k = 0; k < npath; ++k)
            HIPCHK(launch_sgm_dir32_ws(ctx->stream, *p, w, h, D, d_gray, d_cost, d_sum32, kDirs[k][0], kDirs[k][1],
                                       d_sum16),
                   "sgm");
        return 0;
    }
    for (int k = 0; k < npath; ++k)
        HIPCHK(launch_sgm_dir(ctx->stream, *p, w, h, D, d_gray, d_cost, d_sum16, kDirs[k][0], kDirs[k][1], k == 0 ? 0 : 1),
               "sgm");
    if (d_sum32) HIPCHK(launch_widen16(ctx->stream, (size_t)w * h * D, d_sum16, d_sum32), "widen");
    return 0;
}

int apg_hip_sgm_path(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h, int D, const uint8_t* d_gray,
                     const uint16_t* d_cost, uint32_t* d_acc, int dx, int dy) {
    if (!ctx) return 8;
    if (!p) return fail(ctx, 4, "null params");
    if (D <= 0 || D > 2 * 64 * kMaxSgmK) return fail(ctx, 5, "unsupported D");
    if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || (dx == 0 && dy == 0)) return fail(ctx, 4, "bad direction");
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    HIPCHK(launch_sgm_dir32(ctx->stream, *p, w, h, D, d_gray, d_cost, d_acc, dx, dy), "sgm_path");
    return 0;
}

int apg_hip_wta(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h, const void* d_vol, int vol_bits,
                const int16_t* dmin, const int16_t* dmax, float* d_disp, float* d_best, float* d_second) {
    if (int s = check_ctx(ctx, p)) return s;
    if (vol_bits != 16 && vol_bits != 32) return fail(ctx, 4, "vol_bits must be 16 or 32");
    HIPCHK(launch_wta(ctx->stream, *p, w, h, d_vol, vol_bits, dmin, dmax, d_disp, d_best, d_second), "wta");
    return 0;
}

int apg_hip_confidence(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h, const float* d_disp,
                       const float* d_dr, const uint8_t* d_gx, const uint8_t* d_gy, const float* d_best,
                       const float* d_second, float* d_conf, uint8_t* d_rel, uint8_t* d_reason) {
    if (int s = check_ctx(ctx, p)) return s;
    HIPCHK(launch_confidence(ctx->stream, *p, w, h, d_disp, d_dr, d_gx, d_gy, d_best, d_second, d_conf, d_rel, d_reason),
           "confidence");
    return 0;
}

int apg_hip_lr_check(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h, float* d_disp, const float* d_dr,
                     uint8_t* d_reason) {
    if (int s = check_ctx(ctx, p)) return s;
    HIPCHK(launch_lr_check(ctx->stream, *p, w, h, d_disp, d_dr, d_reason), "lr_check");
    return 0;
}

int apg_hip_fill_holes(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h, const float* d_disp,
                       const uint8_t* d_reason, const uint8_t* d_gray, const float* d_prior, float* d_out) {
    if (int s = check_ctx(ctx, p)) return s;
    HIPCHK(launch_fill(ctx->stream, *p, w, h, d_disp, d_reason, d_gray, d_prior, d_out), "fill_holes");
    return 0;
}

int apg_hip_median(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h, const float* d_disp, float* d_out) {
    if (int s = check_ctx(ctx, p)) return s;
    HIPCHK(launch_median(ctx->stream, *p, w, h, d_disp, d_out), "median");
    return 0;
}

}  // extern "C"

// defined with the pipeline below
int run_prior(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h, const uint8_t* gxl, const uint8_t* gyl,
              const void* cl, const void* cr, int16_t* dmin, int16_t* dmax, float* d_prior, int* n_supports);
int run_refine(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h, const uint8_t* gl, const uint8_t* gxl,
               const uint8_t* gyl, const uint8_t* gr, const uint8_t* gxr, const uint8_t* gyr, const void* cl,
               const void* cr, const int16_t* dmin, const int16_t* dmax, const uint8_t* reliable, float* disp);

extern "C" {

int apg_hip_prior_estimate(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h, const uint8_t* d_gx_l,
                           const uint8_t* d_gy_l, const uint32_t* d_c32l, const uint32_t* d_c32r,
                           const uint64_t* d_c64l, const uint64_t* d_c64r, int16_t* d_dmin, int16_t* d_dmax,
                           float* d_prior, int* n_supports) {
    if (int s = check_ctx(ctx, p)) return s;
    if (w <= 0 || h <= 0 || !d_dmin || !d_dmax || !d_prior) return fail(ctx, 1, "empty input");
    const bool sym = p->census_type == 1;
    if (p->prior_enable && (sym ? (!d_c32l || !d_c32r) : (!d_c64l || !d_c64r)))
        return fail(ctx, 4, "census buffers");
    HIPCHK(launch_range_init(ctx->stream, *p, w, h, d_dmin, d_dmax), "range");
    if (int s = run_prior(ctx, p, w, h, d_gx_l, d_gy_l, sym ? (const void*)d_c32l : (const void*)d_c64l,
                          sym ? (const void*)d_c32r : (const void*)d_c64r, d_dmin, d_dmax, d_prior, n_supports))
        return s;
    HIPCHK(hipStreamSynchronize(ctx->stream), "prior sync");
    return 0;
}

int apg_hip_last_supports(apg_hip_ctx ctx, apg_hip_support* out, int capacity, int* count) {
    if (!ctx) return 8;
    if (ctx->supports_pending) {
        // each cell's first taken[] candidates, cells in index order
        if (ctx->capturing) return fail(ctx, 8, "the prior's supports cannot be read back under stream capture");
        const PriorLayout& L = ctx->prior_lay;
        const size_t nc = (size_t)L.ncell, ns = nc * (size_t)L.per_cell;
        std::vector<int> taken(nc), sidx(ns);
        std::vector<float> key(ns), wt(ns);
        hipStream_t st = ctx->stream;
        HIPCHK(hipMemcpyAsync(taken.data(), L.taken, nc * 4, hipMemcpyDeviceToHost, st), "supports download");
        HIPCHK(hipMemcpyAsync(sidx.data(), L.sidx, ns * 4, hipMemcpyDeviceToHost, st), "supports download");
        HIPCHK(hipMemcpyAsync(key.data(), L.key, ns * 4, hipMemcpyDeviceToHost, st), "supports download");
        HIPCHK(hipMemcpyAsync(wt.data(), L.wt, ns * 4, hipMemcpyDeviceToHost, st), "supports download");
        HIPCHK(hipStreamSynchronize(st), "supports download");
        ctx->last_supports.clear();
        for (size_t i = 0; i < nc; ++i)
            for (int k = 0; k < taken[i]; ++k) {
                const size_t o = i * (size_t)L.per_cell + (size_t)k;
                ctx->last_supports.push_back(apg_hip_support{2 + 2 * (sidx[o] % ctx->prior_sw),
                                                             2 + 2 * (sidx[o] / ctx->prior_sw), key[o], wt[o]});
            }
        ctx->supports_pending = false;
    }
    const int n = (int)ctx->last_supports.size();
    if (count) *count = n;
    if (out && capacity > 0) std::copy_n(ctx->last_supports.begin(), std::min(n, capacity), out);
    return 0;
}

int apg_hip_refine(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h, const uint8_t* d_gray_l,
                   const uint8_t* d_gx_l, const uint8_t* d_gy_l, const uint8_t* d_gray_r, const uint8_t* d_gx_r,
                   const uint8_t* d_gy_r, const uint32_t* d_c32l, const uint32_t* d_c32r, const uint64_t* d_c64l,
                   const uint64_t* d_c64r, const int16_t* d_dmin, const int16_t* d_dmax, const uint8_t* d_reliable,
                   float* d_disparity) {
    if (int s = check_ctx(ctx, p)) return s;
    if (w <= 0 || h <= 0) return fail(ctx, 1, "empty input");
    if (!p->refine_enable) return 0;
    const bool sym = p->census_type == 1;
    if (sym ? (!d_c32l || !d_c32r) : (!d_c64l || !d_c64r)) return fail(ctx, 4, "census buffers");
    if (int s = ensure_lut(ctx, p)) return s;
    return run_refine(ctx, p, w, h, d_gray_l, d_gx_l, d_gy_l, d_gray_r, d_gx_r, d_gy_r,
                      sym ? (const void*)d_c32l : (const void*)d_c64l, sym ? (const void*)d_c32r : (const void*)d_c64r,
                      d_dmin, d_dmax, d_reliable, d_disparity);
}

// -------------------------------------------------------------- pipeline ----

}  // extern "C"

namespace {
template <typename T>
hipError_t download_vec(hipStream_t st, std::vector<T>& v, const void* d, size_t count) {
    v.resize(count);
    return count ? hipMemcpyAsync(v.data(), d, count * sizeof(T), hipMemcpyDeviceToHost, st) : hipSuccess;
}

}  // namespace

// PriorEstimator::estimate (reference src/prior_estimator.cpp:414-435) after
// the dense range init, on the device: per-sample candidates, support
// selection, cell grid, per-pixel prior and ranges.  d_prior is -1 where
// there is no prior.  Synchronises only when n_supports is requested.
int run_prior(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h, const uint8_t* gxl, const uint8_t* gyl,
              const void* cl, const void* cr, int16_t* dmin, int16_t* dmax, float* d_prior, int* n_supports) {
    hipStream_t st = ctx->stream;
    const size_t n = (size_t)w * h;
    if (n_supports) *n_supports = 0;
    ctx->last_supports.clear();
    ctx->supports_pending = false;
    HIPCHK(launch_fill_u32(st, n, 0xBF800000u /* -1.0f */, (uint32_t*)d_prior), "prior init");
    const int sw = prior_samples_x(w), sh = prior_samples_y(h);
    if (!p->prior_enable || sw <= 0 || sh <= 0) return 0;
    if (ctx->capturing && n_supports) return fail(ctx, 8, "the support count cannot be read back under stream capture");
    const bool sym = p->census_type == 1;
    HIPCHK(ensure(ctx, B_CAND_D, (size_t)sw * sh * 2), "prior alloc");
    HIPCHK(ensure(ctx, B_CAND_U, (size_t)sw * sh * 4), "prior alloc");
    HIPCHK(ensure(ctx, B_GRID, prior_work_bytes(w, h, p->prior_max_supports)), "prior alloc");
    const PriorLayout L = prior_layout(w, h, p->prior_max_supports, ctx->bufs[B_GRID].ptr);
    int16_t* cd = buf<int16_t>(ctx, B_CAND_D);
    float* cu = buf<float>(ctx, B_CAND_U);
    HIPCHK(launch_prior_candidates(st, *p, w, h, gxl, gyl, sym ? (const uint32_t*)cl : nullptr,
                                   sym ? (const uint32_t*)cr : nullptr, sym ? nullptr : (const uint64_t*)cl,
                                   sym ? nullptr : (const uint64_t*)cr, cd, cu),
           "prior candidates");
    HIPCHK(launch_prior_grid(st, *p, w, h, cd, cu, L), "prior grid");
    HIPCHK(launch_prior_ranges(st, *p, w, h, L, dmin, dmax, d_prior), "prior ranges");
    ctx->prior_lay = L;
    ctx->prior_sw = sw;
    ctx->supports_pending = true;
    if (n_supports) {
        int ns = 0;
        HIPCHK(hipMemcpyAsync(&ns, L.nsel, 4, hipMemcpyDeviceToHost, st), "support count");
        HIPCHK(hipStreamSynchronize(st), "support count");
        *n_supports = ns;
    }
    return 0;
}

// Refiner::refine (reference src/refiner.cpp:90-98) between the confidence
// and the LR check, on the device.  The engine outputs are generated once on
// the host with std::mt19937 and kept for later frames.
int run_refine(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h, const uint8_t* gl, const uint8_t* gxl,
               const uint8_t* gyl, const uint8_t* gr, const uint8_t* gxr, const uint8_t* gyr, const void* cl,
               const void* cr, const int16_t* dmin, const int16_t* dmax, const uint8_t* reliable, float* disp) {
    if (h > 4096) return fail(ctx, 6, "the refiner supports heights up to 4096");
    hipStream_t st = ctx->stream;
    const size_t n = (size_t)w * h;
    const size_t need = refine_engine_len(*p, w, h);
    // draw offsets are 32-bit on the device
    if (need > 0xFFFFFFFFull) return fail(ctx, 5, "the refiner's random stream exceeds 2^32 outputs");
    (void)n;
    if (ctx->engine_len < need || !ctx->bufs[B_ENGINE].ptr) {
        if (ctx->capturing) return fail(ctx, 8, "the refiner's random stream was not uploaded before stream capture");
        std::vector<uint32_t> e(need);
        std::mt19937 gen(12345);
        for (auto& v : e) v = (uint32_t)gen();
        HIPCHK(ensure(ctx, B_ENGINE, need * 4), "alloc");
        HIPCHK(hipMemcpyAsync(ctx->bufs[B_ENGINE].ptr, e.data(), need * 4, hipMemcpyHostToDevice, st), "engine upload");
        HIPCHK(hipStreamSynchronize(st), "engine upload");
        ctx->engine_len = need;
    }
    const size_t sb = refine_scratch_bytes(w, h);
    HIPCHK(ensure(ctx, B_REFINE, sb), "alloc");
    const bool sym = p->census_type == 1;
    const RefineInputs in{gl, gxl, gyl, gr, gxr, gyr, sym ? (const uint32_t*)cl : nullptr,
                          sym ? (const uint32_t*)cr : nullptr, sym ? nullptr : (const uint64_t*)cl,
                          sym ? nullptr : (const uint64_t*)cr};
    if (!ctx->cu_count) {
        int cus = 0, khz = 0;
        HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device), "device attribute");
        HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device), "device attribute");
        ctx->cu_count = cus > 0 ? cus : 1;
        ctx->wall_hz = khz > 0 ? (long long)khz * 1000 : 100000000LL;
    }
    if (!ctx->d_status) {
        if (ctx->capturing) return fail(ctx, 8, "the refiner's status word was not allocated before stream capture");
        HIPCHK(hipMalloc((void**)&ctx->d_status, 256), "alloc");
        HIPCHK(launch_fill_u32(st, 1, 0u, ctx->d_status), "status clear");
    }
    // the banded sweep's hand-offs: a bounded wait (2 s of wall clock); a
    // give-up sets the context's sticky status, read below after a direct
    // run, or by apgk_ctx_take_status after a replayed graph
    const RefineBanding bd{1, ctx->cu_count, ctx->refine_wait ? ctx->refine_wait : 2 * ctx->wall_hz};
    bool banded = false;
    HIPCHK(launch_refine(st, *p, cur_lut(ctx), w, h, in, dmin, dmax, reliable, disp,
                         buf<uint32_t>(ctx, B_ENGINE), ctx->engine_len, ctx->bufs[B_REFINE].ptr,
                         ctx->bufs[B_REFINE].bytes, bd, ctx->d_status, &banded),
           "refine");
    if (banded && !ctx->capturing) return apgk_ctx_take_status(ctx);
    return 0;
}

extern "C" {

// Dense branch of StereoMatcher::compute_impl (reference src/pipeline.cpp:435-605)
// on the device: aux -> range -> cost -> right WTA (raw volume) -> Cross ->
// SGM -> WTA -> confidence -> LR -> fill -> median.
int apg_hip_pipeline(apg_hip_ctx ctx, const apg_hip_params* p, const uint8_t* d_left, const uint8_t* d_right, int w,
                     int h, int stride, int channels, float* d_disparity, float* d_confidence, uint8_t* d_reason,
                     uint8_t* d_reliable, const apg_hip_intermediates* inter) {
    if (!ctx) return 8;
    if (!d_left || !d_right || w <= 0 || h <= 0) return fail(ctx, 1, "empty input image(left or right)");
    if (stride < w * channels || (channels != 1 && channels != 3))
        return fail(ctx, 2, "invalid row stride (stride less than row byte length)");
    if (int s = check_ctx(ctx, p)) return s;
    if (int s = ensure_lut(ctx, p)) return s;
    hipStream_t st = ctx->stream;
    const size_t n = (size_t)w * h;
    const int D = p->max_disparity - p->min_disparity;
    const size_t nv = n * (size_t)D;
    const WsPlan plan = plan_workspace(*p, w, h, ctx->lut_host[ctx->lut_cur], ctx->schedule, inter, ctx->tune);
    for (int b = 0; b < B_COUNT; ++b)
        if (b != B_LUT && plan.bytes[b]) HIPCHK(ensure(ctx, (BufId)b, plan.bytes[b]), "alloc");
    const bool fused = plan.fused;
    const int npaths = p->paths == 8 ? 8 : 4;
    const PairIO io{d_left, d_right, w, h, stride, channels, d_disparity, d_confidence, d_reason, d_reliable};
    uint8_t *gl = buf<uint8_t>(ctx, B_GRAY_L), *gxl = buf<uint8_t>(ctx, B_GX_L), *gyl = buf<uint8_t>(ctx, B_GY_L);
    uint8_t *gr = buf<uint8_t>(ctx, B_GRAY_R), *gxr = buf<uint8_t>(ctx, B_GX_R), *gyr = buf<uint8_t>(ctx, B_GY_R);
    void *cl = ctx->bufs[B_CEN_L].ptr, *cr = ctx->bufs[B_CEN_R].ptr;
    int16_t *dmin = buf<int16_t>(ctx, B_DMIN), *dmax = buf<int16_t>(ctx, B_DMAX);
    uint32_t* arms = buf<uint32_t>(ctx, B_ARMS);
    uint16_t *cost = buf<uint16_t>(ctx, B_COST), *tmp = buf<uint16_t>(ctx, B_TMP), *sum = buf<uint16_t>(ctx, B_SUM);
    float *dr = buf<float>(ctx, B_DR), *best = buf<float>(ctx, B_BEST), *second = buf<float>(ctx, B_SECOND);
    float* scratch = buf<float>(ctx, B_SCRATCH);
    const bool sym = p->census_type == 1;
    const bool arms_aux = aux_pair_builds_arms(*p);

    mark(ctx, S_AUX);
    if (int s = fused_step(ctx, st, p, io, plan, F_AUX, inter, nullptr)) return s;
    mark(ctx, S_PRIOR);
    float* d_prior = p->prior_enable ? buf<float>(ctx, B_PRIOR) : nullptr;
    if (d_prior) {
        if (int s = run_prior(ctx, p, w, h, gxl, gyl, cl, cr, dmin, dmax, d_prior, nullptr)) return s;
    }
    if (inter && inter->d_dmin) HIPCHK(hipMemcpyAsync(inter->d_dmin, dmin, n * 2, hipMemcpyDeviceToDevice, st), "copy dmin");
    if (inter && inter->d_dmax) HIPCHK(hipMemcpyAsync(inter->d_dmax, dmax, n * 2, hipMemcpyDeviceToDevice, st), "copy dmax");
    if (inter && inter->d_prior) {
        if (d_prior) HIPCHK(hipMemcpyAsync(inter->d_prior, d_prior, n * 4, hipMemcpyDeviceToDevice, st), "copy prior");
        else HIPCHK(launch_fill_u32(st, n, 0xBF800000u, (uint32_t*)inter->d_prior), "fill prior");
    }

    if (fused || plan.planes_fused) {
        const int marks[F_STEPS] = {S_AUX, S_COST, S_RWTA, S_CROSS, S_SGM, S_WTA, S_POST};
        for (int k = F_COST; k <= F_WTA || (fused && k == F_POST); ++k) {
            if (k == F_POST) {
                mark(ctx, S_CONF);
                mark(ctx, S_REFINE);
            }
            mark(ctx, marks[k]);
            if (int s = fused_step(ctx, st, p, io, plan, k, inter, nullptr)) return s;
        }
        if (fused) {
            mark(ctx, S_COUNT);
            return finish_timing(ctx);
        }
        mark(ctx, S_CONF);
    } else {
    mark(ctx, S_COST);
    HIPCHK(launch_cost(st, *p, cur_lut(ctx), w, h, gl, gxl, gyl, gr, gxr, gyr,
                       sym ? (const uint32_t*)cl : nullptr, sym ? (const uint32_t*)cr : nullptr,
                       sym ? nullptr : (const uint64_t*)cl, sym ? nullptr : (const uint64_t*)cr, dmin, dmax, cost),
           "cost");
    if (inter && inter->d_cost) HIPCHK(hipMemcpyAsync(inter->d_cost, cost, nv * 2, hipMemcpyDeviceToDevice, st), "copy C");
    mark(ctx, S_RWTA);
    HIPCHK(launch_right_wta(st, *p, w, h, cost, dr), "right_wta");
    mark(ctx, S_CROSS);
    if (p->agg_enable) {
        if (!arms_aux) HIPCHK(launch_cross_arms(st, *p, gl, w, h, arms), "arms");
        const int iters = p->agg_iterations > 1 ? p->agg_iterations : 1;
        for (int i = 0; i < iters; ++i) {
            HIPCHK(launch_cross_h(st, *p, w, h, dmin, dmax, arms, cost, tmp), "cross_h");
            HIPCHK(launch_cross_v(st, *p, w, h, dmin, dmax, arms, tmp, cost), "cross_v");
        }
    }
    if (inter && inter->d_cost_agg)
        HIPCHK(hipMemcpyAsync(inter->d_cost_agg, cost, nv * 2, hipMemcpyDeviceToDevice, st), "copy C'");
    mark(ctx, S_SGM);
    if (plan.planes_sgm) {
        PlaneSet planes{};
        planes.general_ranges = 1;
        for (int r = 0; r < npaths; ++r) planes.p[r] = buf<uint8_t>(ctx, (BufId)(B_DS0 + r));
        HIPCHK(launch_sgm_lines(st, *p, w, h, D, gl, cost, false, planes, ctx->tune), "sgm_lines");
        mark(ctx, S_WTA);
        // WTA + confidence in one kernel; the LR check follows the refiner
        apg_hip_params pw = *p;
        pw.post_lr_check = 0;
        HIPCHK(launch_wta_fused(st, pw, w, h, cost, false, planes, dmin, dmax, dr, gxl, gyl, d_disparity, d_confidence,
                                d_reason, d_reliable, inter ? inter->d_disp_wta : nullptr,
                                inter && inter->d_best ? inter->d_best : best,
                                inter && inter->d_second ? inter->d_second : second, inter ? inter->d_sum32 : nullptr),
               "wta_fused");
        if (inter && inter->d_disp_right)
            HIPCHK(hipMemcpyAsync(inter->d_disp_right, dr, n * 4, hipMemcpyDeviceToDevice, st), "copy dr");
        mark(ctx, S_CONF);
    } else {
        const int npath = p->paths == 8 ? 8 : 4;
        const bool s16 = u16_path_sums(*p);
        if (s16) {
            for (int k = 0; k < npath; ++k)
                HIPCHK(launch_sgm_dir(st, *p, w, h, D, gl, cost, sum, kDirs[k][0], kDirs[k][1], k == 0 ? 0 : 1),
                       "sgm");
            if (inter && inter->d_sum32) HIPCHK(launch_widen16(st, nv, sum, inter->d_sum32), "widen");
        } else {
            // u32 sums: each path into the free T buffer, added to a zeroed accumulator
            uint32_t* sum32 = (uint32_t*)sum;
            HIPCHK(launch_fill_u32(st, nv, 0u, sum32), "sgm clear");
            for (int k = 0; k < npath; ++k)
                HIPCHK(launch_sgm_dir32_ws(st, *p, w, h, D, gl, cost, sum32, kDirs[k][0], kDirs[k][1], tmp), "sgm");
            if (inter && inter->d_sum32)
                HIPCHK(hipMemcpyAsync(inter->d_sum32, sum32, nv * 4, hipMemcpyDeviceToDevice, st), "copy S");
        }
        mark(ctx, S_WTA);
        HIPCHK(launch_wta(st, *p, w, h, sum, s16 ? 16 : 32, dmin, dmax, d_disparity, best, second), "wta");
        if (inter && inter->d_disp_wta)
            HIPCHK(hipMemcpyAsync(inter->d_disp_wta, d_disparity, n * 4, hipMemcpyDeviceToDevice, st), "copy wta");
        if (inter && inter->d_disp_right)
            HIPCHK(hipMemcpyAsync(inter->d_disp_right, dr, n * 4, hipMemcpyDeviceToDevice, st), "copy dr");
        if (inter && inter->d_best)
            HIPCHK(hipMemcpyAsync(inter->d_best, best, n * 4, hipMemcpyDeviceToDevice, st), "copy best");
        if (inter && inter->d_second)
            HIPCHK(hipMemcpyAsync(inter->d_second, second, n * 4, hipMemcpyDeviceToDevice, st), "copy second");
        mark(ctx, S_CONF);
        HIPCHK(launch_confidence(st, *p, w, h, d_disparity, dr, gxl, gyl, best, second, d_confidence, d_reliable,
                                 d_reason),
               "confidence");
    }
    }   // stage-by-stage Cost, Cross, SGM, WTA, confidence
    mark(ctx, S_REFINE);
    if (p->refine_enable) {
        if (int s = run_refine(ctx, p, w, h, gl, gxl, gyl, gr, gxr, gyr, cl, cr, dmin, dmax, d_reliable, d_disparity))
            return s;
    }
    if (inter && inter->d_after_refine)
        HIPCHK(hipMemcpyAsync(inter->d_after_refine, d_disparity, n * 4, hipMemcpyDeviceToDevice, st), "copy refined");
    mark(ctx, S_POST);
    HIPCHK(launch_lr_check(st, *p, w, h, d_disparity, dr, d_reason), "lr");
    HIPCHK(launch_fill(st, *p, w, h, d_disparity, d_reason, gl, d_prior, scratch), "fill");
    HIPCHK(launch_median(st, *p, w, h, scratch, d_disparity), "median");
    mark(ctx, S_COUNT);
    return finish_timing(ctx);
}

void apg_hip_set_schedule(apg_hip_ctx ctx, int schedule) {
    if (ctx) ctx->schedule = schedule;
}

size_t apg_hip_pipeline_workspace_bytes(const apg_hip_params* p, int w, int h) {
    if (!p || w <= 0 || h <= 0 || apg_hip_validate_params(p)) return 0;
    CostLut lut;
    build_cost_lut(*p, &lut);
    const WsPlan pl = plan_workspace(*p, w, h, lut, 0, nullptr);
    size_t t = 0;
    for (size_t b : pl.bytes)
        if (b) t += b < 256 ? 256 : b;
    return t;
}

int apg_hip_set_refine_wait(apg_hip_ctx ctx, long long ticks) {
    if (!ctx) return 8;
    ctx->refine_wait = ticks;
    return 0;
}

int apg_hip_set_tuning(apg_hip_ctx ctx, const apg_hip_tuning* t) {
    if (!ctx) return 8;
    if (!t) {
        ctx->tune = KTune();
        return 0;
    }
    if (t->cost_segment != 0 && (t->cost_segment < 32 || t->cost_segment > 1024))
        return fail(ctx, 4, "cost_segment must be 0 or 32..1024");
    if (t->sgm_dual < 0 || t->sgm_dual > 4) return fail(ctx, 4, "sgm_dual must be 0..4");
    if (t->cross_v_waves != 0 && (t->cross_v_waves < 256 || t->cross_v_waves > 65536))
        return fail(ctx, 4, "cross_v_waves must be 0 or 256..65536");
    ctx->tune.cost_segment = t->cost_segment;
    ctx->tune.sgm_dual = t->sgm_dual;
    ctx->tune.cross_v_waves = t->cross_v_waves;
    return 0;
}

int apg_hip_get_tuning(apg_hip_ctx ctx, apg_hip_tuning* out) {
    if (!ctx || !out) return 8;
    out->cost_segment = ctx->tune.cost_segment;
    out->sgm_dual = ctx->tune.sgm_dual;
    out->cross_v_waves = ctx->tune.cross_v_waves;
    return 0;
}

}  // extern "C"

// Library-internal (hidden): stream-capture guard used by the batch executor.
void apgk_ctx_set_capturing(apg_hip_ctx ctx, bool on) {
    if (ctx) ctx->capturing = on;
}

int apgk_ctx_take_status(apg_hip_ctx ctx) {
    if (!ctx || !ctx->d_status) return 0;
    uint32_t gave_up = 0;
    HIPCHK(hipMemcpyAsync(&gave_up, ctx->d_status, 4, hipMemcpyDeviceToHost, ctx->stream), "refine status");
    HIPCHK(hipStreamSynchronize(ctx->stream), "refine status");
    if (!gave_up) return 0;
    HIPCHK(launch_fill_u32(ctx->stream, 1, 0u, ctx->d_status), "status clear");
    HIPCHK(hipStreamSynchronize(ctx->stream), "status clear");
    return fail(ctx, 8, "refiner: a band's hand-off wait gave up");
}

size_t apgk_ctx_growth_bytes(apg_hip_ctx ctx, const apg_hip_params* p, int w, int h) {
    if (!ctx || !p || w <= 0 || h <= 0 || apg_hip_validate_params(p)) return 0;
    CostLut lut;
    build_cost_lut(*p, &lut);
    const WsPlan pl = plan_workspace(*p, w, h, lut, ctx->schedule, nullptr, ctx->tune);
    size_t g = 0;
    for (int b = 0; b < B_COUNT; ++b) {
        const size_t want = pl.bytes[b] ? (pl.bytes[b] < 256 ? 256 : pl.bytes[b]) : 0;
        const size_t have = ctx->bufs[b].ptr ? ctx->bufs[b].bytes : 0;
        if (want > have) g += want - have;
    }
    return g;
}

int apgk_ctx_create_on(int device, hipStream_t st, apg_hip_ctx* out) {
    if (!out || !st) return 8;
    *out = nullptr;
    apg_hip_ctx ctx = new (std::nothrow) apg_hip_ctx_s();
    if (!ctx) return 7;
    ctx->device = device;
    ctx->stream = st;
    ctx->own_stream = false;
    *out = ctx;
    return 0;
}

int apgk_group_prepare(apg_hip_ctx ctx, const apg_hip_pair& q) {
    if (!ctx) return 8;
    if (!q.d_left || !q.d_right || q.width <= 0 || q.height <= 0) return fail(ctx, 1, "empty input image(left or right)");
    if (q.stride < q.width * q.channels || (q.channels != 1 && q.channels != 3))
        return fail(ctx, 2, "invalid row stride (stride less than row byte length)");
    if (int s = check_ctx(ctx, &q.params)) return s;
    if (ctx->capturing) return fail(ctx, 8, "group workspace not prepared before stream capture");
    if (int s = ensure_lut(ctx, &q.params)) return s;
    const WsPlan plan = plan_workspace(q.params, q.width, q.height, ctx->lut_host[ctx->lut_cur], ctx->schedule,
                                       nullptr, ctx->tune);
    if (!plan.fused) return fail(ctx, 4, "the pair does not take the fused column-range schedule");
    for (int b = 0; b < B_COUNT; ++b)
        if (b != B_LUT && plan.bytes[b]) HIPCHK(ensure(ctx, (BufId)b, plan.bytes[b]), "alloc");
    return 0;
}

int apgk_group_record(apg_hip_ctx ctx, const apg_hip_pair& q, int step, apgk::LaunchSink* sink) {
    if (!ctx || !sink) return 8;
    if (step < 0 || step >= kGroupSteps) return fail(ctx, 4, "invalid step");
    // the LUT must already be resident: an upload now could evict the
    // table an earlier recorded launch of this context reads
    const LutKey key = LutKey::of(q.params);
    int li = -1;
    for (int i = 0; i < kLutSlots; ++i)
        if (ctx->lut_used[i] && ctx->lut_key[i] == key) li = i;
    if (li < 0) return fail(ctx, 4, "cost LUT not resident");
    ctx->lut_cur = li;
    const WsPlan plan = plan_workspace(q.params, q.width, q.height, ctx->lut_host[li], ctx->schedule, nullptr,
                                       ctx->tune);
    if (!plan.fused) return fail(ctx, 4, "the pair does not take the fused column-range schedule");
    for (int b = 0; b < B_COUNT; ++b)
        if (b != B_LUT && plan.bytes[b] > ctx->bufs[b].bytes) return fail(ctx, 4, "workspace not prepared");
    const PairIO io{q.d_left, q.d_right, q.width, q.height, q.stride, q.channels,
                    q.d_disparity, q.d_confidence, q.d_invalid_reason, q.d_reliable};
    static const int kSteps[kGroupSteps] = {F_AUX, F_COST, F_RWTA, F_CROSS, F_SGM, F_WTA, F_POST};
    return fused_step(ctx, ctx->stream, &q.params, io, plan, kSteps[step], nullptr, sink);
}

extern "C" {

int apg_hip_bench_pipeline(apg_hip_ctx ctx, const apg_hip_params* p, const uint8_t* d_left, const uint8_t* d_right,
                           int w, int h, int stride, int channels, float* d_disp, float* d_conf, uint8_t* d_reason,
                           uint8_t* d_rel, int warmup, int steps, double* ms_total, apg_hip_timing* stage_avg) {
    if (!ctx) return 8;
    const bool was_timing = ctx->timing;
    ctx->timing = false;
    for (int i = 0; i < warmup; ++i) {
        const int s = apg_hip_pipeline(ctx, p, d_left, d_right, w, h, stride, channels, d_disp, d_conf, d_reason, d_rel,
                                       nullptr);
        if (s) { ctx->timing = was_timing; return s; }
    }
    HIPCHK(hipStreamSynchronize(ctx->stream), "bench sync");
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0), "event");
    HIPCHK(hipEventCreate(&e1), "event");
    (void)hipEventRecord(e0, ctx->stream);
    int status = 0;
    for (int i = 0; i < steps && !status; ++i)
        status = apg_hip_pipeline(ctx, p, d_left, d_right, w, h, stride, channels, d_disp, d_conf, d_reason, d_rel,
                                  nullptr);
    (void)hipEventRecord(e1, ctx->stream);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms_total) *ms_total = ms;
    if (!status && stage_avg) {
        apg_hip_timing acc{};
        ctx->timing = true;
        for (int i = 0; i < steps && !status; ++i) {
            status = apg_hip_pipeline(ctx, p, d_left, d_right, w, h, stride, channels, d_disp, d_conf, d_reason, d_rel,
                                      nullptr);
            const apg_hip_timing& t = ctx->last;
            acc.aux_ms += t.aux_ms; acc.prior_ms += t.prior_ms; acc.cost_ms += t.cost_ms;
            acc.right_wta_ms += t.right_wta_ms; acc.cross_ms += t.cross_ms; acc.sgm_ms += t.sgm_ms;
            acc.wta_ms += t.wta_ms; acc.confidence_ms += t.confidence_ms; acc.refine_ms += t.refine_ms;
            acc.post_ms += t.post_ms;
            acc.total_ms += t.total_ms;
        }
        const double k = steps > 0 ? 1.0 / steps : 0.0;
        stage_avg->aux_ms = acc.aux_ms * k; stage_avg->prior_ms = acc.prior_ms * k; stage_avg->cost_ms = acc.cost_ms * k;
        stage_avg->right_wta_ms = acc.right_wta_ms * k; stage_avg->cross_ms = acc.cross_ms * k;
        stage_avg->sgm_ms = acc.sgm_ms * k; stage_avg->wta_ms = acc.wta_ms * k;
        stage_avg->confidence_ms = acc.confidence_ms * k; stage_avg->refine_ms = acc.refine_ms * k;
        stage_avg->post_ms = acc.post_ms * k;
        stage_avg->total_ms = acc.total_ms * k;
    }
    ctx->timing = was_timing;
    return status;
}

int apg_hip_bench_pipeline_steps(apg_hip_ctx ctx, const apg_hip_params* p, const uint8_t* d_left,
                                 const uint8_t* d_right, int w, int h, int stride, int channels, float* d_disp,
                                 float* d_conf, uint8_t* d_reason, uint8_t* d_rel, int warmup, int steps,
                                 float* step_ms) {
    if (!ctx) return 8;
    if (steps <= 0 || !step_ms) return fail(ctx, 4, "steps and step_ms required");
    const bool was_timing = ctx->timing;
    ctx->timing = false;
    int status = 0;
    for (int i = 0; i < warmup && !status; ++i)
        status = apg_hip_pipeline(ctx, p, d_left, d_right, w, h, stride, channels, d_disp, d_conf, d_reason, d_rel,
                                  nullptr);
    std::vector<hipEvent_t> ev((size_t)steps + 1, nullptr);
    hipError_t e = hipSuccess;
    for (auto& x : ev)
        if (e == hipSuccess) e = hipEventCreate(&x);
    if (!status && e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (!status && e == hipSuccess) {
        // one event before each call and one after the last: back-to-back calls,
        // each interval is one pair's device time on the context stream
        (void)hipEventRecord(ev[0], ctx->stream);
        for (int i = 0; i < steps && !status; ++i) {
            status = apg_hip_pipeline(ctx, p, d_left, d_right, w, h, stride, channels, d_disp, d_conf, d_reason,
                                      d_rel, nullptr);
            (void)hipEventRecord(ev[(size_t)i + 1], ctx->stream);
        }
        e = hipEventSynchronize(ev[(size_t)steps]);
        for (int i = 0; i < steps && e == hipSuccess && !status; ++i)
            e = hipEventElapsedTime(&step_ms[i], ev[(size_t)i], ev[(size_t)i + 1]);
    }
    for (auto& x : ev)
        if (x) (void)hipEventDestroy(x);
    ctx->timing = was_timing;
    if (status) return status;
    if (e != hipSuccess) return hip_fail(ctx, e, "bench steps");
    return 0;
}

const char* apg_hip_version_string(void) { return "1.0.0-gfx950"; }

}  // extern "C"

// ------------------------------------------------ allocation-free entry ----
//
// apg_hip_pipeline_static: the fused schedule with its whole workspace in a
// file-scope device array instead of context buffers.  Each GPU the library
// runs on holds its own copy of `scratch` from the moment the code object is
// loaded until the process ends.  Host code never passes the array's name to
// a kernel: on a device's first call the array's device address is taken
// with hipGetSymbolAddress and kept per device, and every buffer handed to the
// kernels is an offset into that address.

namespace {

constexpr size_t kStaticScratchBytes = size_t(768) << 20;
constexpr int kStaticMaxDevices = 64;

// [0, kStaticLutBytes): the cost LUT slots (their host-side cache survives
// between calls); the rest: one call's workspace, laid out per call
constexpr size_t kStaticLutBytes = (kLutSlots * sizeof(CostLut) + 255) & ~size_t(255);

static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

struct StaticDevice {
    std::mutex mu;          // serialises the host side of calls on one device
    bool ready = false;
    uint8_t* ptr = nullptr;   // device address of `scratch` on this device
    apg_hip_ctx_s ctx;        // LUT cache and launch shapes; no owned buffers
};
StaticDevice g_static[kStaticMaxDevices];
thread_local std::string g_static_error;

int static_fail(int code, const std::string& msg) {
    g_static_error = msg;
    return code;
}

}  // namespace

extern "C" {

size_t apg_hip_static_scratch_bytes(void) { return kStaticScratchBytes; }

const char* apg_hip_static_last_error(void) { return g_static_error.c_str(); }

size_t apg_hip_pipeline_static_bytes(const apg_hip_params* p, int w, int h) {
    if (!p || w <= 0 || h <= 0 || apg_hip_validate_params(p)) return 0;
    CostLut lut;
    build_cost_lut(*p, &lut);
    const WsPlan pl = plan_workspace(*p, w, h, lut, 0, nullptr);
    if (!pl.fused) return 0;
    size_t t = kStaticLutBytes;
    for (int b = 0; b < B_COUNT; ++b)
        if (b != B_LUT && pl.bytes[b]) t += (pl.bytes[b] + 255) & ~size_t(255);
    return t;
}

int apg_hip_pipeline_static(const apg_hip_params* p, const uint8_t* d_left, const uint8_t* d_right, int w, int h,
                            int stride, int channels, float* d_disparity, float* d_confidence, uint8_t* d_reason,
                            uint8_t* d_reliable, void* stream) {
    g_static_error.clear();
    if (!p) return static_fail(4, "null params");
    if (!d_left || !d_right || w <= 0 || h <= 0) return static_fail(1, "empty input image(left or right)");
    if (!d_disparity || !d_confidence || !d_reason || !d_reliable) return static_fail(4, "null output");
    if (stride < w * channels || (channels != 1 && channels != 3))
        return static_fail(2, "invalid row stride (stride less than row byte length)");
    if (int s = apg_hip_validate_params(p)) return static_fail(s, "invalid parameters");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return static_fail(8, "no current HIP device");
    if (dev >= kStaticMaxDevices) return static_fail(5, "device index beyond the static entry point's table");
    StaticDevice& sd = g_static[dev];
    std::lock_guard<std::mutex> lock(sd.mu);
    apg_hip_ctx ctx = &sd.ctx;
    if (!sd.ready) {
        // once per device, with that device current
        void* addr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !addr)
            return static_fail(8, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        sd.ptr = static_cast<uint8_t*>(addr);
        ctx->device = dev;
        sd.ready = true;
    }
    ctx->stream = static_cast<hipStream_t>(stream);
    ctx->timing = false;
    ctx->schedule = 0;
    // under capture the LUT upload would record a host copy: the parameter
    // set's LUT must already be cached by an earlier eager call
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(ctx->stream, &cs) != hipSuccess) return static_fail(8, "hipStreamIsCapturing");
    ctx->capturing = cs != hipStreamCaptureStatusNone;
    ctx->bufs[B_LUT] = DevBuf{sd.ptr, kLutSlots * sizeof(CostLut)};
    if (int s = ensure_lut(ctx, p)) return static_fail(s, ctx->last_error);
    const WsPlan plan = plan_workspace(*p, w, h, ctx->lut_host[ctx->lut_cur], 0, nullptr, ctx->tune);
    if (!plan.fused)
        return static_fail(5, "the static entry point runs the fused schedule only (no prior or refiner, even D <= 256, "
                              "one Cross iteration)");
    // this call's workspace, carved from the array after the LUT slots
    size_t off = kStaticLutBytes;
    for (int b = 0; b < B_COUNT; ++b) {
        if (b == B_LUT) continue;
        if (!plan.bytes[b]) {
            ctx->bufs[b] = DevBuf{};
            continue;
        }
        const size_t need = (plan.bytes[b] + 255) & ~size_t(255);
        if (off + need > kStaticScratchBytes) {
            for (int k = 0; k < B_COUNT; ++k)
                if (k != B_LUT) ctx->bufs[k] = DevBuf{};
            return static_fail(6, "pair needs " + std::to_string(apg_hip_pipeline_static_bytes(p, w, h)) +
                                      " bytes of workspace; the static scratch holds " +
                                      std::to_string(kStaticScratchBytes));
        }
        ctx->bufs[b] = DevBuf{sd.ptr + off, plan.bytes[b]};
        off += need;
    }
    // every buffer fits, so the pipeline's ensure() allocates nothing
    const int s = apg_hip_pipeline(ctx, p, d_left, d_right, w, h, stride, channels, d_disparity, d_confidence,
                                   d_reason, d_reliable, nullptr);
    if (s) g_static_error = ctx->last_error;
    return s;
}

}  // extern "C"
