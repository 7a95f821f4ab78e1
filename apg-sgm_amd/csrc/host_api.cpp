// C++ host layer: the reference's stage classes, StereoMatcher, workspace and
// session, implemented on the device C ABI (include/apg_sgm_hip.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "apg_sgm/pipeline.hpp"
#include "apg_sgm/stages.hpp"
#include "apg_sgm_hip.h"
#include "copy_pool.hpp"

namespace apg {

// PipelineConfig -> device stage parameters.
apg_hip_params to_device_params(const PipelineConfig& c) {
    apg_hip_params p;
    apg_hip_params_from_mode(static_cast<int>(c.mode), c.max_disparity, &p);
    p.min_disparity = c.min_disparity;
    p.max_disparity = c.max_disparity;
    p.census_type = c.cost.census == CensusType::SymmetricCensus9x7 ? 1 : 0;
    p.lambda_census = c.cost.lambda_census;
    p.lambda_ad = c.cost.lambda_ad;
    p.lambda_grad = c.cost.lambda_grad;
    p.eta_ad = c.cost.eta_ad;
    p.mu_grad = c.cost.mu_grad;
    p.use_ad = c.cost.use_ad;
    p.use_grad = c.cost.use_grad;
    p.cost_max = c.cost.cost_max;
    p.agg_enable = c.aggregation.enable;
    p.max_arm_length = c.aggregation.max_arm_length;
    p.color_threshold = c.aggregation.color_threshold;
    p.agg_iterations = c.aggregation.iterations;
    p.paths = c.sgm.paths == PathType::Path8 ? 8 : 4;
    p.P1 = c.sgm.P1;
    p.P2_base = c.sgm.P2_base;
    p.adaptive_penalties = c.sgm.adaptive_penalties;
    p.grad_t1 = c.sgm.grad_t1;
    p.grad_t2 = c.sgm.grad_t2;
    p.p2_alpha = c.sgm.p2_alpha;
    p.piecewise_p2 = c.sgm.piecewise_p2;
    p.conf_uniqueness_ratio = c.confidence.uniqueness_ratio;
    p.conf_lr_max_diff = c.confidence.lr_max_diff;
    p.conf_min_texture = c.confidence.min_texture;
    p.conf_reliable_threshold = c.confidence.reliable_threshold;
    p.post_subpixel = c.post.subpixel;
    p.post_lr_check = c.post.lr_check;
    p.post_lr_max_diff = c.post.lr_max_diff;
    p.post_median_radius = c.post.median_radius;
    p.post_max_fill_gap = c.post.max_fill_gap;
    p.prior_enable = c.prior.enable ? 1 : 0;
    p.prior_uniqueness_ratio = c.prior.uniqueness_ratio;
    p.prior_texture_threshold = c.prior.texture_threshold;
    p.prior_lr_max_diff = c.prior.lr_max_diff;
    p.prior_search_radius = c.prior.search_radius;
    p.prior_min_supports = c.prior.min_supports;
    p.prior_max_supports = c.prior.max_supports;
    p.refine_enable = c.refine.enable ? 1 : 0;
    p.refine_iterations = c.refine.iterations;
    p.refine_random_radius = c.refine.random_radius;
    return p;
}

namespace {

// Per-thread device context used by the stage classes (they are stateless
// const objects in the reference, so the device state lives beside them).
struct ThreadCtx {
    apg_hip_ctx ctx = nullptr;
    ~ThreadCtx() {
        if (ctx) apg_hip_ctx_destroy(ctx);
    }
};

apg_hip_ctx stage_ctx() {
    thread_local ThreadCtx t;
    if (!t.ctx) {
        const int rc = apg_hip_ctx_create(0, &t.ctx);
        if (rc == 7) throw std::bad_alloc();
        if (rc) throw std::runtime_error("apg: no usable HIP device for the stage classes");
    }
    return t.ctx;
}

void check(apg_hip_ctx ctx, int rc, const char* what) {
    if (rc == 0) return;
    if (rc == 7) throw std::bad_alloc();
    throw std::runtime_error(std::string(what) + ": " + apg_hip_last_error(ctx));
}

// RAII device allocation of the stage classes.
class Dev {
public:
    Dev(apg_hip_ctx ctx, size_t bytes) : ctx_(ctx) { check(ctx, apg_hip_malloc(ctx, bytes ? bytes : 1, &p_), "malloc"); }
    Dev(apg_hip_ctx ctx, const void* host, size_t bytes) : Dev(ctx, bytes) {
        if (bytes) check(ctx, apg_hip_upload(ctx, p_, host, bytes), "upload");
    }
    ~Dev() {
        if (p_) apg_hip_free(ctx_, p_);
    }
    Dev(const Dev&) = delete;
    Dev& operator=(const Dev&) = delete;
    template <typename T>
    T* as() const { return static_cast<T*>(p_); }
    void get(void* host, size_t bytes) const {
        if (bytes) check(ctx_, apg_hip_download(ctx_, host, p_, bytes), "download");
    }

private:
    apg_hip_ctx ctx_;
    void* p_ = nullptr;
};

size_t npx(int w, int h) { return static_cast<size_t>(w) * static_cast<size_t>(h); }

// Packs a strided view into a contiguous w*channels image.
std::vector<uint8_t> pack(const ImageView8& v) {
    std::vector<uint8_t> out(npx(v.width(), v.height()) * v.channels());
    for (int y = 0; y < v.height(); ++y)
        std::memcpy(out.data() + static_cast<size_t>(y) * v.width() * v.channels(), v.row(y),
                    static_cast<size_t>(v.width()) * v.channels());
    return out;
}

void aux_one(apg_hip_ctx ctx, const apg_hip_params& p, const ImageView8& v, Image8& gray, Image8& gx, Image8& gy,
             std::vector<uint32_t>& c32, std::vector<uint64_t>& c64) {
    const int w = v.width(), h = v.height();
    const size_t n = npx(w, h);
    const std::vector<uint8_t> src = pack(v);
    Dev ds(ctx, src.data(), src.size()), dg(ctx, n), dgx(ctx, n), dgy(ctx, n), dc(ctx, n * 8);
    const bool sym = p.census_type == 1;
    check(ctx,
          apg_hip_aux(ctx, &p, ds.as<uint8_t>(), w, h, w * v.channels(), v.channels(), dg.as<uint8_t>(),
                      dgx.as<uint8_t>(), dgy.as<uint8_t>(), sym ? dc.as<uint32_t>() : nullptr,
                      sym ? nullptr : dc.as<uint64_t>()),
          "aux");
    gray = Image8(w, h, 1);
    gx = Image8(w, h, 1);
    gy = Image8(w, h, 1);
    dg.get(gray.data(), n);
    dgx.get(gx.data(), n);
    dgy.get(gy.data(), n);
    if (sym) {
        c32.assign(n, 0);
        c64.clear();
        dc.get(c32.data(), n * 4);
    } else {
        c64.assign(n, 0);
        c32.clear();
        dc.get(c64.data(), n * 8);
    }
}

}  // namespace

// ------------------------------------------------------------ CostComputer --
void CostComputer::compute_aux(const ImageView8& left, const ImageView8& right, const PipelineConfig& cfg,
                               PipelineBuffers& buf) const {
    apg_hip_ctx ctx = stage_ctx();
    const apg_hip_params p = to_device_params(cfg);
    aux_one(ctx, p, left, buf.left_gray, buf.left_gx, buf.left_gy, buf.census_left, buf.census_left64);
    aux_one(ctx, p, right, buf.right_gray, buf.right_gx, buf.right_gy, buf.census_right, buf.census_right64);
}

void CostComputer::compute_aux(const Image8& left, const Image8& right, const PipelineConfig& cfg,
                               PipelineBuffers& buf) const {
    buf.left = left;
    buf.right = right;
    compute_aux(ImageView8(left), ImageView8(right), cfg, buf);
}

void CostComputer::compute_volume(const PipelineConfig& cfg, PipelineBuffers& buf) const {
    apg_hip_ctx ctx = stage_ctx();
    const apg_hip_params p = to_device_params(cfg);
    const int w = buf.left_gray.width(), h = buf.left_gray.height();
    const size_t n = npx(w, h);
    const bool sym = cfg.cost.census == CensusType::SymmetricCensus9x7;
    Dev gl(ctx, buf.left_gray.data(), n), gxl(ctx, buf.left_gx.data(), n), gyl(ctx, buf.left_gy.data(), n);
    Dev gr(ctx, buf.right_gray.data(), n), gxr(ctx, buf.right_gx.data(), n), gyr(ctx, buf.right_gy.data(), n);
    Dev cl(ctx, sym ? (const void*)buf.census_left.data() : (const void*)buf.census_left64.data(), n * (sym ? 4 : 8));
    Dev cr(ctx, sym ? (const void*)buf.census_right.data() : (const void*)buf.census_right64.data(), n * (sym ? 4 : 8));
    Dev dmin(ctx, buf.range.dmin.data(), n * 2), dmax(ctx, buf.range.dmax.data(), n * 2);
    buf.cost.allocate(w, h, cfg.min_disparity, cfg.max_disparity, kInvalidCost);
    Dev out(ctx, buf.cost.bytes());
    check(ctx,
          apg_hip_cost_volume(ctx, &p, w, h, gl.as<uint8_t>(), gxl.as<uint8_t>(), gyl.as<uint8_t>(), gr.as<uint8_t>(),
                              gxr.as<uint8_t>(), gyr.as<uint8_t>(), sym ? cl.as<uint32_t>() : nullptr,
                              sym ? cr.as<uint32_t>() : nullptr, sym ? nullptr : cl.as<uint64_t>(),
                              sym ? nullptr : cr.as<uint64_t>(), dmin.as<int16_t>(), dmax.as<int16_t>(),
                              out.as<uint16_t>()),
          "compute_volume");
    out.get(buf.cost.data(), buf.cost.bytes());
}

// ------------------------------------------------------------ packed forms --
namespace {

// The dense span [d0, d1) that covers every pixel's range of a layout, and
// the layout's ranges as a SearchRange over it.
struct DenseSpan {
    int d0 = 0, d1 = 0;
    SearchRange range;
    bool empty() const { return d1 <= d0; }
};

DenseSpan span_of(const PackedVolumeLayout& L) {
    DenseSpan s;
    const size_t n = npx(L.width, L.height);
    bool any = false;
    for (size_t q = 0; q < n; ++q)
        if (L.disp_width(q) > 0) {
            s.d0 = any ? std::min(s.d0, (int)L.dmin[q]) : L.dmin[q];
            s.d1 = any ? std::max(s.d1, L.dmax(q)) : L.dmax(q);
            any = true;
        }
    s.range.allocate(L.width, L.height, s.d0, s.d1);
    for (size_t q = 0; q < n; ++q) {
        s.range.dmin.data()[q] = L.dmin[q];
        s.range.dmax.data()[q] = (int16_t)L.dmax(q);
    }
    return s;
}

// packed -> dense over the span (states outside a pixel's range: `fill`),
// with an optional value map (u16 -> u32 sentinel widening)
template <typename TS, typename TD, typename Map>
void unpack(const PackedCostVolume<TS>& v, const DenseSpan& s, CostVolumeT<TD>& out, TD fill, Map map) {
    const PackedVolumeLayout& L = *v.layout();
    out.allocate(L.width, L.height, s.d0, s.d1, fill);
    const size_t n = npx(L.width, L.height);
    const int D = s.d1 - s.d0;
    for (size_t q = 0; q < n; ++q) {
        const TS* src = v.data() + L.offsets[q];
        TD* dst = out.data() + q * D + (L.dmin[q] - s.d0);
        for (int k = 0, e = L.disp_width(q); k < e; ++k) dst[k] = map(src[k]);
    }
}

template <typename TS, typename TD, typename Map>
void pack(const CostVolumeT<TS>& in, PackedCostVolume<TD>& v, Map map) {
    const PackedVolumeLayout& L = *v.layout();
    const size_t n = npx(L.width, L.height);
    const int D = in.D();
    for (size_t q = 0; q < n; ++q) {
        const TS* src = in.data() + q * D + (L.dmin[q] - in.d0());
        TD* dst = v.data() + L.offsets[q];
        for (int k = 0, e = L.disp_width(q); k < e; ++k) dst[k] = map(src[k]);
    }
}

template <typename T>
T same(T v) { return v; }
uint32_t widen(uint16_t v) { return v == kInvalidCost ? kInvalidCost32 : v; }
uint16_t narrow(uint32_t v) { return v == kInvalidCost32 ? kInvalidCost : static_cast<uint16_t>(v); }

PipelineConfig over_span(const PipelineConfig& cfg, const DenseSpan& s) {
    PipelineConfig c = cfg;
    c.min_disparity = s.d0;
    c.max_disparity = s.d1;
    return c;
}

}  // namespace

void CostComputer::compute_volume_packed(const PipelineConfig& cfg, const PipelineBuffers& buf,
                                         PackedCostVolume16& packed_cost) const {
    if (!packed_cost.layout()) packed_cost.allocate_for_overwrite(buf.range);
    const DenseSpan s = span_of(*packed_cost.layout());
    if (s.empty()) return;
    // the dense kernel over the span with the layout's ranges
    PipelineBuffers b;
    b.left_gray = buf.left_gray;
    b.left_gx = buf.left_gx;
    b.left_gy = buf.left_gy;
    b.right_gray = buf.right_gray;
    b.right_gx = buf.right_gx;
    b.right_gy = buf.right_gy;
    b.census_left = buf.census_left;
    b.census_right = buf.census_right;
    b.census_left64 = buf.census_left64;
    b.census_right64 = buf.census_right64;
    b.range = s.range;
    compute_volume(over_span(cfg, s), b);
    pack(b.cost, packed_cost, same<uint16_t>);
}

void CostAggregator::aggregate_packed(const PipelineConfig& cfg, const Image8& left_gray,
                                      PackedCostVolume16& packed_cost) const {
    if (!cfg.aggregation.enable || !packed_cost.layout()) return;
    const DenseSpan s = span_of(*packed_cost.layout());
    if (s.empty()) return;
    PipelineBuffers b;
    b.left_gray = left_gray;
    b.range = s.range;
    unpack(packed_cost, s, b.cost, kInvalidCost, same<uint16_t>);
    aggregate(over_span(cfg, s), b);
    pack(b.cost, packed_cost, same<uint16_t>);
}

void SgmOptimizer::optimize_packed(const PipelineConfig& cfg, const Image8& gray, const PackedCostVolume16& packed_cost,
                                   PackedCostVolume16& packed_sgm16) const {
    packed_sgm16.allocate_for_overwrite(packed_cost.layout());
    if (!packed_cost.layout()) return;
    const DenseSpan s = span_of(*packed_cost.layout());
    if (s.empty()) return;
    PipelineBuffers b;
    b.left_gray = gray;
    unpack(packed_cost, s, b.cost, kInvalidCost, same<uint16_t>);
    optimize(over_span(cfg, s), b);
    pack(b.cost32, packed_sgm16, narrow);
}

void SgmOptimizer::optimize_packed(const PipelineConfig& cfg, const Image8& gray, const PackedCostVolume16& packed_cost,
                                   PackedCostVolume32& packed_cost32) const {
    packed_cost32.allocate_for_overwrite(packed_cost.layout());
    if (!packed_cost.layout()) return;
    const DenseSpan s = span_of(*packed_cost.layout());
    if (s.empty()) return;
    PipelineBuffers b;
    b.left_gray = gray;
    unpack(packed_cost, s, b.cost, kInvalidCost, same<uint16_t>);
    optimize(over_span(cfg, s), b);
    pack(b.cost32, packed_cost32, same<uint32_t>);
}

void SgmOptimizer::winner_take_all_packed(const PipelineConfig& cfg, const PackedCostVolume32& vol, Image32f& disp,
                                          Image32f* best, Image32f* second) const {
    if (!vol.layout()) return;
    const DenseSpan s = span_of(*vol.layout());
    CostVolume32 dense;
    unpack(vol, s, dense, kInvalidCost32, same<uint32_t>);
    winner_take_all(over_span(cfg, s), dense, s.range, disp, best, second);
}

void SgmOptimizer::winner_take_all_packed(const PipelineConfig& cfg, const PackedCostVolume16& vol, Image32f& disp,
                                          Image32f* best, Image32f* second) const {
    if (!vol.layout()) return;
    const DenseSpan s = span_of(*vol.layout());
    CostVolume dense;
    unpack(vol, s, dense, kInvalidCost, same<uint16_t>);
    winner_take_all(over_span(cfg, s), dense, s.range, disp, best, second);
}

void SgmOptimizer::aggregate_path_packed(const PipelineConfig& cfg, const Image8& gray, const PackedCostVolume16& base,
                                         PackedCostVolume16& acc, int dx, int dy) const {
    if (!base.layout() || !acc.layout()) return;
    const DenseSpan s = span_of(*base.layout());
    if (s.empty()) return;
    CostVolume dense;
    CostVolume32 dacc;
    unpack(base, s, dense, kInvalidCost, same<uint16_t>);
    unpack(acc, s, dacc, 0u, widen);
    aggregate_path(over_span(cfg, s), gray, dense, dacc, dx, dy);
    pack(dacc, acc, narrow);
}

void SgmOptimizer::aggregate_path_packed(const PipelineConfig& cfg, const Image8& gray, const PackedCostVolume16& base,
                                         PackedCostVolume32& acc, int dx, int dy) const {
    if (!base.layout() || !acc.layout()) return;
    const DenseSpan s = span_of(*base.layout());
    if (s.empty()) return;
    CostVolume dense;
    CostVolume32 dacc;
    unpack(base, s, dense, kInvalidCost, same<uint16_t>);
    unpack(acc, s, dacc, 0u, same<uint32_t>);
    aggregate_path(over_span(cfg, s), gray, dense, dacc, dx, dy);
    pack(dacc, acc, same<uint32_t>);
}

// Direct definitions of the descriptors (reference src/cost_computer.cpp:33-61).
uint32_t CostComputer::symmetric_census9x7(const Image8& gray, int x, int y) {
    uint32_t bits = 0;
    int b = 0;
    for (int dy = -3; dy <= 3; ++dy)
        for (int dx = -4; dx <= 4; ++dx) {
            if (!(dy < 0 || (dy == 0 && dx < 0))) continue;
            if (gray.sample(x + dx, y + dy) > gray.sample(x - dx, y - dy)) bits |= 1u << b;
            ++b;
        }
    return bits;
}

uint64_t CostComputer::census9x7(const Image8& gray, int x, int y) {
    const uint8_t c = gray.sample(x, y);
    uint64_t bits = 0;
    int b = 0;
    for (int dy = -3; dy <= 3; ++dy)
        for (int dx = -4; dx <= 4; ++dx) {
            if (dx == 0 && dy == 0) continue;
            if (gray.sample(x + dx, y + dy) > c) bits |= 1ull << b;
            ++b;
        }
    return bits;
}

// ---------------------------------------------------------- PriorEstimator --
void PriorEstimator::estimate(const PipelineConfig& cfg, PipelineBuffers& buf) const {
    const int w = buf.left_gray.width(), h = buf.left_gray.height();
    buf.range.allocate(w, h, cfg.min_disparity, cfg.max_disparity);
    buf.d_prior = Image32f(w, h, -1.f);
    buf.supports.clear();
    buf.support_count = 0;
    if (!cfg.prior.enable) {
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x)
                buf.range.set_pixel(x, y, cfg.min_disparity, std::min(cfg.max_disparity, x + 1));
        return;
    }
    // the prior on the device; the support list itself stays there (its
    // count is reported)
    apg_hip_ctx ctx = stage_ctx();
    const apg_hip_params p = to_device_params(cfg);
    const size_t n = npx(w, h);
    const bool sym = cfg.cost.census == CensusType::SymmetricCensus9x7;
    Dev gx(ctx, buf.left_gx.data(), n), gy(ctx, buf.left_gy.data(), n);
    Dev cl(ctx, sym ? (const void*)buf.census_left.data() : (const void*)buf.census_left64.data(), n * (sym ? 4 : 8));
    Dev cr(ctx, sym ? (const void*)buf.census_right.data() : (const void*)buf.census_right64.data(), n * (sym ? 4 : 8));
    Dev dmin(ctx, n * 2), dmax(ctx, n * 2), dp(ctx, n * 4);
    int nsup = 0;
    check(ctx,
          apg_hip_prior_estimate(ctx, &p, w, h, gx.as<uint8_t>(), gy.as<uint8_t>(), sym ? cl.as<uint32_t>() : nullptr,
                                 sym ? cr.as<uint32_t>() : nullptr, sym ? nullptr : cl.as<uint64_t>(),
                                 sym ? nullptr : cr.as<uint64_t>(), dmin.as<int16_t>(), dmax.as<int16_t>(),
                                 dp.as<float>(), &nsup),
          "prior");
    dmin.get(buf.range.dmin.data(), n * 2);
    dmax.get(buf.range.dmax.data(), n * 2);
    dp.get(buf.d_prior.data(), n * 4);
    buf.support_count = static_cast<size_t>(nsup);
}

std::vector<SupportMatch> PriorEstimator::extract_supports(const PipelineConfig& cfg,
                                                           const PipelineBuffers& buf) const {
    std::vector<SupportMatch> out;
    if (!cfg.prior.enable) return out;
    const int w = buf.left_gray.width(), h = buf.left_gray.height();
    apg_hip_ctx ctx = stage_ctx();
    // the device candidates and the host selection of estimate()
    const apg_hip_params p = to_device_params(cfg);
    const size_t n = npx(w, h);
    const bool sym = cfg.cost.census == CensusType::SymmetricCensus9x7;
    Dev gx(ctx, buf.left_gx.data(), n), gy(ctx, buf.left_gy.data(), n);
    Dev cl(ctx, sym ? (const void*)buf.census_left.data() : (const void*)buf.census_left64.data(), n * (sym ? 4 : 8));
    Dev cr(ctx, sym ? (const void*)buf.census_right.data() : (const void*)buf.census_right64.data(), n * (sym ? 4 : 8));
    Dev dmin(ctx, n * 2), dmax(ctx, n * 2), dp(ctx, n * 4);
    int nsup = 0;
    check(ctx,
          apg_hip_prior_estimate(ctx, &p, w, h, gx.as<uint8_t>(), gy.as<uint8_t>(), sym ? cl.as<uint32_t>() : nullptr,
                                 sym ? cr.as<uint32_t>() : nullptr, sym ? nullptr : cl.as<uint64_t>(),
                                 sym ? nullptr : cr.as<uint64_t>(), dmin.as<int16_t>(), dmax.as<int16_t>(),
                                 dp.as<float>(), &nsup),
          "prior");
    std::vector<apg_hip_support> sup(static_cast<size_t>(nsup));
    int cnt = 0;
    check(ctx, apg_hip_last_supports(ctx, sup.data(), nsup, &cnt), "supports");
    out.reserve(sup.size());
    for (const apg_hip_support& s : sup) {
        SupportMatch m;
        m.x = s.x;
        m.y = s.y;
        m.disparity = s.disparity;
        m.confidence = s.confidence;
        out.push_back(m);
    }
    return out;
}

// ----------------------------------------------------------------- Refiner --
void Refiner::refine(const PipelineConfig& cfg, PipelineBuffers& buf) const {
    if (!cfg.refine.enable) return;
    const int w = buf.disparity.width(), h = buf.disparity.height();
    apg_hip_ctx ctx = stage_ctx();
    const apg_hip_params p = to_device_params(cfg);
    const size_t n = npx(w, h);
    const bool sym = cfg.cost.census == CensusType::SymmetricCensus9x7;
    Dev gl(ctx, buf.left_gray.data(), n), gxl(ctx, buf.left_gx.data(), n), gyl(ctx, buf.left_gy.data(), n);
    Dev gr(ctx, buf.right_gray.data(), n), gxr(ctx, buf.right_gx.data(), n), gyr(ctx, buf.right_gy.data(), n);
    Dev cl(ctx, sym ? (const void*)buf.census_left.data() : (const void*)buf.census_left64.data(), n * (sym ? 4 : 8));
    Dev cr(ctx, sym ? (const void*)buf.census_right.data() : (const void*)buf.census_right64.data(), n * (sym ? 4 : 8));
    Dev dmin(ctx, buf.range.dmin.data(), n * 2), dmax(ctx, buf.range.dmax.data(), n * 2);
    Dev rel(ctx, buf.reliable_mask.data(), n), disp(ctx, buf.disparity.data(), n * 4);
    check(ctx,
          apg_hip_refine(ctx, &p, w, h, gl.as<uint8_t>(), gxl.as<uint8_t>(), gyl.as<uint8_t>(), gr.as<uint8_t>(),
                         gxr.as<uint8_t>(), gyr.as<uint8_t>(), sym ? cl.as<uint32_t>() : nullptr,
                         sym ? cr.as<uint32_t>() : nullptr, sym ? nullptr : cl.as<uint64_t>(),
                         sym ? nullptr : cr.as<uint64_t>(), dmin.as<int16_t>(), dmax.as<int16_t>(),
                         rel.as<uint8_t>(), disp.as<float>()),
          "refine");
    disp.get(buf.disparity.data(), n * 4);
}

// ---------------------------------------------------------- CostAggregator --
void CostAggregator::aggregate(const PipelineConfig& cfg, PipelineBuffers& buf) const {
    if (!cfg.aggregation.enable) return;
    apg_hip_ctx ctx = stage_ctx();
    const apg_hip_params p = to_device_params(cfg);
    const int w = buf.cost.width(), h = buf.cost.height();
    const size_t n = npx(w, h);
    Dev g(ctx, buf.left_gray.data(), n), dmin(ctx, buf.range.dmin.data(), n * 2), dmax(ctx, buf.range.dmax.data(), n * 2);
    Dev c(ctx, buf.cost.data(), buf.cost.bytes()), t(ctx, buf.cost.bytes());
    check(ctx,
          apg_hip_cross_aggregate(ctx, &p, w, h, g.as<uint8_t>(), dmin.as<int16_t>(), dmax.as<int16_t>(),
                                  c.as<uint16_t>(), t.as<uint16_t>()),
          "aggregate");
    c.get(buf.cost.data(), buf.cost.bytes());
}

// ------------------------------------------------------------ SgmOptimizer --
void SgmOptimizer::optimize(const PipelineConfig& cfg, PipelineBuffers& buf) const {
    apg_hip_ctx ctx = stage_ctx();
    const apg_hip_params p = to_device_params(cfg);
    const int w = buf.cost.width(), h = buf.cost.height();
    const size_t n = npx(w, h);
    Dev g(ctx, buf.left_gray.data(), n), c(ctx, buf.cost.data(), buf.cost.bytes());
    Dev s16(ctx, buf.cost.bytes());
    buf.cost32.allocate(w, h, buf.cost.d0(), buf.cost.d0() + buf.cost.D(), 0);
    Dev s32(ctx, buf.cost32.bytes());
    check(ctx, apg_hip_sgm_optimize(ctx, &p, w, h, g.as<uint8_t>(), c.as<uint16_t>(), s16.as<uint16_t>(), s32.as<uint32_t>()),
          "optimize");
    s32.get(buf.cost32.data(), buf.cost32.bytes());
}

void SgmOptimizer::aggregate_path(const PipelineConfig& cfg, const Image8& gray, const CostVolume& base,
                                  CostVolume32& acc, int dx, int dy) const {
    apg_hip_ctx ctx = stage_ctx();
    const apg_hip_params p = to_device_params(cfg);
    const int w = base.width(), h = base.height();
    Dev g(ctx, gray.data(), npx(w, h)), c(ctx, base.data(), base.bytes()), a(ctx, acc.data(), acc.bytes());
    check(ctx, apg_hip_sgm_path(ctx, &p, w, h, base.D(), g.as<uint8_t>(), c.as<uint16_t>(), a.as<uint32_t>(), dx, dy),
          "aggregate_path");
    a.get(acc.data(), acc.bytes());
}

namespace {
template <typename Vol>
void wta_common(const PipelineConfig& cfg, const Vol& vol, int bits, const SearchRange& range, Image32f& disp,
                Image32f* best, Image32f* second) {
    apg_hip_ctx ctx = stage_ctx();
    apg_hip_params p = to_device_params(cfg);
    p.min_disparity = vol.d0();
    p.max_disparity = vol.d0() + vol.D();
    const int w = vol.width(), h = vol.height();
    const size_t n = npx(w, h);
    Dev v(ctx, vol.data(), vol.bytes()), dmin(ctx, range.dmin.data(), n * 2), dmax(ctx, range.dmax.data(), n * 2);
    Dev dd(ctx, n * 4), db(ctx, n * 4), ds(ctx, n * 4);
    check(ctx,
          apg_hip_wta(ctx, &p, w, h, v.as<void>(), bits, dmin.as<int16_t>(), dmax.as<int16_t>(), dd.as<float>(),
                      db.as<float>(), ds.as<float>()),
          "winner_take_all");
    disp = Image32f(w, h, -1.f);
    dd.get(disp.data(), n * 4);
    if (best) {
        *best = Image32f(w, h, 0.f);
        db.get(best->data(), n * 4);
    }
    if (second) {
        *second = Image32f(w, h, 0.f);
        ds.get(second->data(), n * 4);
    }
}
}  // namespace

void SgmOptimizer::winner_take_all(const PipelineConfig& cfg, const CostVolume32& vol, const SearchRange& range,
                                   Image32f& disp, Image32f* best, Image32f* second) const {
    wta_common(cfg, vol, 32, range, disp, best, second);
}

void SgmOptimizer::winner_take_all(const PipelineConfig& cfg, const CostVolume& vol, const SearchRange& range,
                                   Image32f& disp, Image32f* best, Image32f* second) const {
    wta_common(cfg, vol, 16, range, disp, best, second);
}

// ----------------------------------------------------- ConfidenceEstimator --
void ConfidenceEstimator::estimate(const PipelineConfig& cfg, PipelineBuffers& buf, const Image32f& best,
                                   const Image32f& second) const {
    apg_hip_ctx ctx = stage_ctx();
    const apg_hip_params p = to_device_params(cfg);
    const int w = buf.disparity.width(), h = buf.disparity.height();
    const size_t n = npx(w, h);
    Dev d(ctx, buf.disparity.data(), n * 4), gx(ctx, buf.left_gx.data(), n), gy(ctx, buf.left_gy.data(), n);
    Dev b(ctx, best.data(), n * 4), s(ctx, second.data(), n * 4);
    const bool has_dr = !buf.disparity_right.empty();
    Dev dr(ctx, has_dr ? n * 4 : 4);
    if (has_dr) check(ctx, apg_hip_upload(ctx, dr.as<void>(), buf.disparity_right.data(), n * 4), "upload");
    Dev conf(ctx, n * 4), rel(ctx, n), rea(ctx, n);
    check(ctx,
          apg_hip_confidence(ctx, &p, w, h, d.as<float>(), has_dr ? dr.as<float>() : nullptr, gx.as<uint8_t>(),
                             gy.as<uint8_t>(), b.as<float>(), s.as<float>(), conf.as<float>(), rel.as<uint8_t>(),
                             rea.as<uint8_t>()),
          "confidence");
    buf.confidence = Image32f(w, h, 0.f);
    buf.reliable_mask = Image8u1(w, h, 0);
    buf.invalid_reason = Image8u1(w, h, 0);
    conf.get(buf.confidence.data(), n * 4);
    rel.get(buf.reliable_mask.data(), n);
    rea.get(buf.invalid_reason.data(), n);
}

// ----------------------------------------------------------- PostProcessor --
void PostProcessor::left_right_check(const PipelineConfig& cfg, PipelineBuffers& buf) const {
    if (!cfg.post.lr_check || buf.disparity_right.empty()) return;
    apg_hip_ctx ctx = stage_ctx();
    const apg_hip_params p = to_device_params(cfg);
    const int w = buf.disparity.width(), h = buf.disparity.height();
    const size_t n = npx(w, h);
    Dev d(ctx, buf.disparity.data(), n * 4), dr(ctx, buf.disparity_right.data(), n * 4);
    Dev r(ctx, buf.invalid_reason.data(), n);
    check(ctx, apg_hip_lr_check(ctx, &p, w, h, d.as<float>(), dr.as<float>(), r.as<uint8_t>()), "left_right_check");
    d.get(buf.disparity.data(), n * 4);
    r.get(buf.invalid_reason.data(), n);
}

void PostProcessor::fill_holes(const PipelineConfig& cfg, PipelineBuffers& buf) const {
    apg_hip_ctx ctx = stage_ctx();
    const apg_hip_params p = to_device_params(cfg);
    const int w = buf.disparity.width(), h = buf.disparity.height();
    const size_t n = npx(w, h);
    Dev d(ctx, buf.disparity.data(), n * 4), r(ctx, buf.invalid_reason.data(), n), g(ctx, buf.left_gray.data(), n);
    const bool has_prior = !buf.d_prior.empty();
    Dev pr(ctx, has_prior ? n * 4 : 4);
    if (has_prior) check(ctx, apg_hip_upload(ctx, pr.as<void>(), buf.d_prior.data(), n * 4), "upload");
    Dev out(ctx, n * 4);
    check(ctx,
          apg_hip_fill_holes(ctx, &p, w, h, d.as<float>(), r.as<uint8_t>(), g.as<uint8_t>(),
                             has_prior ? pr.as<float>() : nullptr, out.as<float>()),
          "fill_holes");
    out.get(buf.disparity.data(), n * 4);
}

void PostProcessor::median(const PipelineConfig& cfg, PipelineBuffers& buf) const {
    if (cfg.post.median_radius <= 0) return;
    apg_hip_ctx ctx = stage_ctx();
    const apg_hip_params p = to_device_params(cfg);
    const int w = buf.disparity.width(), h = buf.disparity.height();
    const size_t n = npx(w, h);
    Dev d(ctx, buf.disparity.data(), n * 4), out(ctx, n * 4);
    check(ctx, apg_hip_median(ctx, &p, w, h, d.as<float>(), out.as<float>()), "median");
    out.get(buf.disparity.data(), n * 4);
}

// -------------------------------------------------------- validation/budget --
Status validate_config(const PipelineConfig& cfg) {
    if (cfg.max_disparity <= cfg.min_disparity)
        return Status(StatusCode::InvalidDisparityRange, "max_disparity must be strictly greater than min_disparity");
    if (cfg.cost.cost_max < 0 || cfg.cost.cost_max > 65535)
        return Status(StatusCode::InvalidConfig, "cost_max out of valid uint16 representation range [0, 65535]");
    if (cfg.sgm.P1 < 0) return Status(StatusCode::InvalidConfig, "SGM P1 penalty must be non-negative");
    if (cfg.sgm.P2_base <= cfg.sgm.P1) return Status(StatusCode::InvalidConfig, "SGM P2_base must be strictly greater than P1");
    if (cfg.sgm.grad_t1 < 0 || cfg.sgm.grad_t2 < cfg.sgm.grad_t1)
        return Status(StatusCode::InvalidConfig, "SGM gradient thresholds must satisfy 0 <= grad_t1 <= grad_t2");
    if (cfg.prior.enable && cfg.prior.search_radius < 0)
        return Status(StatusCode::InvalidConfig, "prior search_radius must be non-negative");
    if (cfg.refine.enable && cfg.refine.iterations < 0)
        return Status(StatusCode::InvalidConfig, "refine iterations must be non-negative");
    return Status::Ok();
}

Status validate_inputs(const ImageView8& left, const ImageView8& right) {
    if (left.empty() || right.empty()) return Status(StatusCode::EmptyInput, "empty input image(left or right)");
    if (left.width() <= 0 || left.height() <= 0)
        return Status(StatusCode::EmptyInput, "image dimensions must be positive non-zero");
    if (left.width() != right.width() || left.height() != right.height())
        return Status(StatusCode::SizeMismatch, "left/right size mismatch (images must be rectified and aligned)");
    if (left.channels() != right.channels()) return Status(StatusCode::SizeMismatch, "left/right channel count mismatch");
    if (left.stride() < left.width() * left.channels() || right.stride() < right.width() * right.channels())
        return Status(StatusCode::SizeMismatch, "invalid row stride (stride less than row byte length)");
    if (left.channels() != 1 && left.channels() != 3)
        return Status(StatusCode::SizeMismatch, "only 1- and 3-channel images are supported");
    return Status::Ok();
}

size_t estimate_workspace_bytes(int width, int height, const PipelineConfig& cfg) {
    if (width <= 0 || height <= 0) return 0;
    const unsigned __int128 n = static_cast<unsigned __int128>(width) * static_cast<unsigned>(height);
    const int dr = cfg.max_disparity - cfg.min_disparity;
    const unsigned __int128 vox = n * static_cast<unsigned>(dr > 0 ? dr : 128);
    unsigned __int128 est = n * 6;                                                  // gray + gradients
    est += n * (cfg.cost.census == CensusType::Census9x7 ? 8u : 4u) * 2;            // census L/R
    est += n * 4;                                                                   // search range
    if (cfg.resolves_to_packed()) {
        est += n * 6;                                                               // packed layout
        est += vox * 2;                                                             // packed cost
        const int npath = cfg.sgm.paths == PathType::Path8 ? 8 : 4;
        est += vox * (can_use_u16_sgm_accumulator(cfg, npath) ? 2 : 4);             // SGM accumulator
    } else {
        est += vox * 2 + vox * 4;                                                   // cost + cost32
    }
    if (cfg.aggregation.enable) est += vox * 2;                                     // Cross tmp
    est += n * 48 + n * 2;                                                          // float planes + masks
    if (est > static_cast<unsigned __int128>(SIZE_MAX)) return SIZE_MAX;
    return static_cast<size_t>(est);
}

// ----------------------------------------------------------- workspace ------
StereoWorkspace::StereoWorkspace(ExecutionOptions options) : options_(options) {}

StereoWorkspace::~StereoWorkspace() {
    if (ctx_) {
        if (dev_io_) apg_hip_free(ctx_, dev_io_);
        if (host_io_) apg_hip_host_free(ctx_, host_io_);
        apg_hip_ctx_destroy(ctx_);
    }
}

apg_hip_ctx_s* StereoWorkspace::device_context() {
    if (!ctx_) {
        const int rc = apg_hip_ctx_create(options_.device, &ctx_);
        if (rc == 7) throw std::bad_alloc();
        if (rc) throw std::runtime_error("apg: cannot create a HIP context on the requested device");
    }
    return ctx_;
}

size_t StereoWorkspace::allocated_bytes() const { return ctx_ ? apg_hip_workspace_bytes(ctx_) + dev_io_bytes_ : 0; }

void StereoWorkspace::record_frame(int w, int h, int d_max, bool was_reuse) {
    if (was_reuse) ++frame_reuse_count_;
    else ++frame_allocation_count_;
    last_w_ = w;
    last_h_ = h;
    last_dmax_ = d_max;
}

bool StereoWorkspace::can_reuse(int w, int h, int d_max) const {
    return options_.memory_policy == MemoryPolicy::BoundedReuse && w == last_w_ && h == last_h_ && d_max == last_dmax_ &&
           allocated_bytes() > 0;
}

void StereoWorkspace::release() {
    if (ctx_) {
        if (dev_io_) apg_hip_free(ctx_, dev_io_);
        dev_io_ = nullptr;
        dev_io_bytes_ = 0;
        if (host_io_) apg_hip_host_free(ctx_, host_io_);
        host_io_ = nullptr;
        host_io_bytes_ = 0;
        apg_hip_release_workspace(ctx_);
    }
    last_w_ = last_h_ = last_dmax_ = 0;
}

// ------------------------------------------------------------ StereoMatcher --
namespace {
// shared by every session of the process (copies of one call at a time
// finish before it returns)
apgk::CopyPool& copy_pool() {
    static apgk::CopyPool pool(apgk::CopyPool::default_threads());
    return pool;
}

Status status_from(int rc, const char* msg) {
    if (rc == 0) return Status::Ok();
    return Status(static_cast<StatusCode>(rc), msg ? msg : "");
}
}  // namespace

Status StereoMatcher::compute_impl(const ImageView8& left, const ImageView8& right, StereoWorkspace& ws,
                                   StereoResult& result, PipelineBuffers* extra, PipelineStats* stats) const {
    last_error_.clear();
    Status s = validate_inputs(left, right);
    if (!s.ok()) return s;
    s = validate_config(cfg_);
    if (!s.ok()) return s;
    if (cfg_.hierarchical.enable)
        return Status(StatusCode::UnsupportedBackend, "the hierarchical range estimator is not part of this device build");
    const apg_hip_params p = to_device_params(cfg_);
    const int vs = apg_hip_validate_params(&p);
    if (vs) return Status(static_cast<StatusCode>(vs), "configuration outside the device path's supported range");
    try {
        apg_hip_ctx ctx = ws.device_context();
        const int w = left.width(), h = left.height(), ch = left.channels();
        const size_t n = npx(w, h);
        const size_t in_l = static_cast<size_t>(h - 1) * left.stride() + static_cast<size_t>(w) * ch;
        const size_t in_r = static_cast<size_t>(h - 1) * right.stride() + static_cast<size_t>(w) * ch;
        // one staging block: left | right | disparity | confidence | reason | reliable
        auto al = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
        const size_t o_r = al(in_l), o_d = o_r + al(in_r), o_c = o_d + al(n * 4), o_re = o_c + al(n * 4),
                     o_m = o_re + al(n), total = o_m + al(n);
        if (ws.dev_io_bytes_ < total) {
            if (ws.dev_io_) apg_hip_free(ctx, ws.dev_io_);
            ws.dev_io_ = nullptr;
            ws.dev_io_bytes_ = 0;
            const int rc = apg_hip_malloc(ctx, total, &ws.dev_io_);
            if (rc) return status_from(rc, "device staging allocation failed");
            ws.dev_io_bytes_ = total;
        }
        uint8_t* base = static_cast<uint8_t*>(ws.dev_io_);
        // A workspace that keeps its memory between frames also keeps a pinned
        // host block of the same layout: the images are copied into it and
        // the results out of it over a few threads, and each direction is one
        // transfer at the link's full rate.  Per-frame release copies directly
        // (a pinned allocation per frame would cost more than it saves).
        uint8_t* hbase = nullptr;
        if (ws.options().memory_policy != MemoryPolicy::ReleasePerFrame) {
            if (ws.host_io_bytes_ < total) {
                if (ws.host_io_) apg_hip_host_free(ctx, ws.host_io_);
                ws.host_io_ = nullptr;
                ws.host_io_bytes_ = 0;
                if (apg_hip_host_malloc(ctx, total, &ws.host_io_) == 0) ws.host_io_bytes_ = total;
            }
            hbase = static_cast<uint8_t*>(ws.host_io_);
        }
        // Equal strides: the strided rows are uploaded as-is and the kernels
        // read through the stride. Different strides: both views are packed.
        int rc;
        if (right.stride() == left.stride()) {
            if (hbase) {
                copy_pool().copy({{hbase, left.data(), in_l}, {hbase + o_r, right.data(), in_r}});
                rc = apg_hip_upload(ctx, base, hbase, o_r + in_r);
            } else {
                rc = apg_hip_upload(ctx, base, left.data(), in_l);
                if (!rc) rc = apg_hip_upload(ctx, base + o_r, right.data(), in_r);
            }
        } else {
            const std::vector<uint8_t> lp = pack(left), rp = pack(right);
            rc = apg_hip_upload(ctx, base, lp.data(), lp.size());
            if (!rc) rc = apg_hip_upload(ctx, base + o_r, rp.data(), rp.size());
        }
        if (rc) return status_from(rc, apg_hip_last_error(ctx));
        const int stride = right.stride() != left.stride() ? w * ch : left.stride();
        apg_hip_enable_stage_timing(ctx, stats != nullptr);
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<float> dr_host;
        apg_hip_intermediates inter{};
        // the legacy buffers take the right disparity, the ranges and the
        // disparities around the refiner; the stats need the ranges when a
        // prior narrowed them and the refiner's input and output when it ran
        const bool want_range = extra || (stats && p.prior_enable);
        const bool want_refine = extra || (stats && p.refine_enable);
        std::unique_ptr<Dev> dr_dev, dmin_dev, dmax_dev, db_dev, da_dev;
        if (extra) {
            dr_dev.reset(new Dev(ctx, n * 4));
            inter.d_disp_right = dr_dev->as<float>();
        }
        if (want_range) {
            dmin_dev.reset(new Dev(ctx, n * 2));
            dmax_dev.reset(new Dev(ctx, n * 2));
            inter.d_dmin = dmin_dev->as<int16_t>();
            inter.d_dmax = dmax_dev->as<int16_t>();
        }
        if (want_refine) {
            db_dev.reset(new Dev(ctx, n * 4));
            da_dev.reset(new Dev(ctx, n * 4));
            inter.d_disp_wta = db_dev->as<float>();
            inter.d_after_refine = da_dev->as<float>();
        }
        const bool any_inter = extra || want_range || want_refine;
        rc = apg_hip_pipeline(ctx, &p, base, base + o_r, w, h, stride, ch, reinterpret_cast<float*>(base + o_d),
                              reinterpret_cast<float*>(base + o_c), base + o_re, base + o_m,
                              any_inter ? &inter : nullptr);
        if (rc) return status_from(rc, apg_hip_last_error(ctx));
        // every pixel of the four planes is written below: planes of the
        // right size are reused as they are
        if (result.disparity.width() != w || result.disparity.height() != h) result.disparity = Image32f(w, h, -1.f);
        if (result.confidence.width() != w || result.confidence.height() != h)
            result.confidence = Image32f(w, h, 0.f);
        if (result.invalid_reason.width() != w || result.invalid_reason.height() != h)
            result.invalid_reason = Image8u1(w, h, 0);
        if (result.reliable_mask.width() != w || result.reliable_mask.height() != h)
            result.reliable_mask = Image8u1(w, h, 0);
        if (hbase) {
            rc = apg_hip_download(ctx, hbase + o_d, base + o_d, o_m + n - o_d);
            if (!rc)
                copy_pool().copy({{result.disparity.data(), hbase + o_d, n * 4},
                                  {result.confidence.data(), hbase + o_c, n * 4},
                                  {result.invalid_reason.data(), hbase + o_re, n},
                                  {result.reliable_mask.data(), hbase + o_m, n}});
        } else {
            rc = apg_hip_download(ctx, result.disparity.data(), base + o_d, n * 4);
            if (!rc) rc = apg_hip_download(ctx, result.confidence.data(), base + o_c, n * 4);
            if (!rc) rc = apg_hip_download(ctx, result.invalid_reason.data(), base + o_re, n);
            if (!rc) rc = apg_hip_download(ctx, result.reliable_mask.data(), base + o_m, n);
        }
        if (!rc && extra) {
            extra->disparity_right = Image32f(w, h, -1.f);
            rc = apg_hip_download(ctx, extra->disparity_right.data(), dr_dev->as<void>(), n * 4);
        }
        SearchRange range;
        if (!rc && want_range) {
            range.allocate(w, h, cfg_.min_disparity, cfg_.max_disparity);
            rc = apg_hip_download(ctx, range.dmin.data(), dmin_dev->as<void>(), n * 2);
            if (!rc) rc = apg_hip_download(ctx, range.dmax.data(), dmax_dev->as<void>(), n * 2);
        }
        Image32f before, after;
        if (!rc && want_refine) {
            before = Image32f(w, h, -1.f);
            after = Image32f(w, h, -1.f);
            rc = apg_hip_download(ctx, before.data(), db_dev->as<void>(), n * 4);
            if (!rc) rc = apg_hip_download(ctx, after.data(), da_dev->as<void>(), n * 4);
        }
        std::vector<SupportMatch> supports;
        if (!rc && p.prior_enable) {
            int ns = 0;
            apg_hip_last_supports(ctx, nullptr, 0, &ns);
            std::vector<apg_hip_support> hs(ns);
            apg_hip_last_supports(ctx, hs.data(), ns, &ns);
            supports.resize(hs.size());
            for (size_t i = 0; i < hs.size(); ++i)
                supports[i] = SupportMatch{hs[i].x, hs[i].y, hs[i].disparity, hs[i].confidence};
        }
        if (rc) return status_from(rc, apg_hip_last_error(ctx));
        const auto t1 = std::chrono::steady_clock::now();
        if (stats) {
            apg_hip_timing t{};
            apg_hip_last_timing(ctx, &t);
            stats->timing.aux_ms = t.aux_ms;
            stats->timing.prior_ms = t.prior_ms;
            stats->timing.cost_ms = t.cost_ms;
            stats->timing.right_wta_ms = t.right_wta_ms;
            stats->timing.cross_ms = t.cross_ms;
            stats->timing.sgm_ms = t.sgm_ms;
            stats->timing.wta_ms = t.wta_ms;
            stats->timing.confidence_ms = t.confidence_ms;
            stats->timing.refine_ms = t.refine_ms;
            stats->timing.post_ms = t.post_ms;
            stats->timing.total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
            const size_t vox = n * static_cast<size_t>(cfg_.max_disparity - cfg_.min_disparity);
            stats->cost_bytes = vox * 2;
            stats->aggregated_cost_bytes = vox * 2;
            stats->estimated_peak_bytes = apg_hip_workspace_bytes(ctx);
            size_t valid = 0, reliable = 0, lr_fail = 0;
            for (size_t i = 0; i < n; ++i) {
                valid += result.disparity.data()[i] >= 0.f;
                reliable += result.reliable_mask.data()[i] != 0;
                const uint8_t r = result.invalid_reason.data()[i];
                lr_fail += r == static_cast<uint8_t>(InvalidReason::Mismatch) ||
                           r == static_cast<uint8_t>(InvalidReason::Occlusion);
            }
            // range statistics (reference src/pipeline.cpp:640-695); without a
            // prior the ranges are the dense geometric init itself
            double geom = 0.0, search = 0.0;
            for (int x = 0; x < w; ++x) {
                const int hi = std::min(cfg_.max_disparity, x + 1), lo = cfg_.min_disparity;
                geom += (hi > lo ? hi - lo : 0) * static_cast<double>(h);
            }
            if (want_range) {
                for (size_t i = 0; i < n; ++i) {
                    const int lo = range.dmin.data()[i], hi = range.dmax.data()[i];
                    search += hi > lo ? hi - lo : 0;
                }
            } else {
                search = geom;
            }
            size_t refine_changed = 0;
            if (want_refine)
                for (size_t i = 0; i < n; ++i) {
                    const float db = before.data()[i], da = after.data()[i];
                    refine_changed += db >= 0.f && da >= 0.f && std::abs(da - db) > 1e-4f;
                }
            const int global_D = cfg_.max_disparity - cfg_.min_disparity;
            stats->prior_support_count = supports.size();
            stats->mean_search_width = n ? search / static_cast<double>(n) : 0.0;
            stats->mean_geometry_width = n ? geom / static_cast<double>(n) : 0.0;
            stats->geometry_reduction_ratio = global_D > 0 ? 1.0 - stats->mean_geometry_width / global_D : 0.0;
            stats->prior_incremental_reduction_ratio =
                stats->mean_geometry_width > 0.0 ? 1.0 - stats->mean_search_width / stats->mean_geometry_width : 0.0;
            stats->search_reduction_ratio = global_D > 0 ? 1.0 - stats->mean_search_width / global_D : 0.0;
            stats->valid_pixels = valid;
            stats->reliable_pixels = reliable;
            stats->unreliable_pixels = n - reliable;
            stats->refine_changed_pixels = refine_changed;
            stats->lr_fail_pixels = lr_fail;
            stats->valid_disparity_ratio = n ? static_cast<float>(valid) / n : 0.f;
            stats->reliable_ratio = n ? static_cast<float>(reliable) / n : 0.f;
            stats->unreliable_ratio = n ? static_cast<float>(n - reliable) / n : 0.f;
            stats->refine_changed_ratio = n ? static_cast<float>(refine_changed) / n : 0.f;
            stats->lr_fail_ratio = n ? static_cast<float>(lr_fail) / n : 0.f;
        }
        if (extra) {
            extra->range = std::move(range);
            extra->supports = std::move(supports);
            extra->support_count = extra->supports.size();
            extra->d_before_refine = std::move(before);
            extra->d_after_refine = std::move(after);
        }
        return Status::Ok();
    } catch (const std::bad_alloc& e) {
        return Status(StatusCode::AllocationFailure, std::string("memory allocation failure: ") + e.what());
    } catch (const std::exception& e) {
        return Status(StatusCode::InternalError, std::string("internal pipeline error: ") + e.what());
    }
}

Status StereoMatcher::compute_status(const ImageView8& left, const ImageView8& right, StereoResult& result,
                                    StereoWorkspace& workspace, PipelineStats* stats) const {
    const int w = left.width(), h = left.height();
    if (workspace.options().max_retained_bytes > 0 &&
        estimate_workspace_bytes(w, h, cfg_) > workspace.options().max_retained_bytes) {
        result.clear();
        return Status(StatusCode::ResourceLimitExceeded, "workspace memory budget exceeded before compute");
    }
    const bool was_reuse = workspace.can_reuse(w, h, cfg_.max_disparity);
    Status s = compute_impl(left, right, workspace, result, nullptr, stats);
    if (!s.ok()) {
        last_error_ = s.message;
        result.clear();
        if (workspace.options().memory_policy == MemoryPolicy::ReleasePerFrame) workspace.release();
        return s;
    }
    workspace.record_frame(w, h, cfg_.max_disparity, was_reuse);
    if (workspace.options().memory_policy == MemoryPolicy::ReleasePerFrame) {
        workspace.release();
    } else if (workspace.options().max_retained_bytes > 0 &&
               workspace.allocated_bytes() > workspace.options().max_retained_bytes) {
        workspace.release();
        result.clear();
        if (stats) stats->retained_workspace_bytes = 0;
        return Status(StatusCode::ResourceLimitExceeded, "workspace memory budget exceeded after compute");
    }
    if (stats) {
        stats->retained_workspace_bytes = workspace.allocated_bytes();
        stats->frame_allocation_count = stats->allocation_count = workspace.frame_allocation_count();
        stats->frame_reuse_count = stats->reuse_count = workspace.frame_reuse_count();
        stats->requested_threads = stats->effective_threads = workspace.options().num_threads;
        stats->requested_backend = cfg_.volume_backend == VolumeBackend::Packed ? "Packed"
                                   : cfg_.volume_backend == VolumeBackend::Dense ? "Dense" : "Auto";
        stats->resolved_backend = cfg_.resolves_to_packed() ? "Packed" : "Dense";
        stats->memory_policy = workspace.options().memory_policy == MemoryPolicy::BoundedReuse ? "BoundedReuse"
                                                                                              : "ReleasePerFrame";
    }
    return Status::Ok();
}

Status StereoMatcher::compute_status(const ImageView8& left, const ImageView8& right, StereoResult& result,
                                    PipelineStats* stats) const {
    StereoWorkspace ws;
    return compute_status(left, right, result, ws, stats);
}

bool StereoMatcher::compute(const ImageView8& l, const ImageView8& r, StereoResult& res, StereoWorkspace& ws,
                            PipelineStats* stats) const {
    const Status s = compute_status(l, r, res, ws, stats);
    if (!s.ok()) last_error_ = s.message;
    return s.ok();
}

bool StereoMatcher::compute(const ImageView8& l, const ImageView8& r, StereoResult& res, PipelineStats* stats) const {
    const Status s = compute_status(l, r, res, stats);
    if (!s.ok()) last_error_ = s.message;
    return s.ok();
}

bool StereoMatcher::compute(const Image8& left, const Image8& right, PipelineBuffers& out, PipelineStats* stats) const {
    StereoWorkspace ws;
    StereoResult res;
    const Status s = compute_impl(ImageView8(left), ImageView8(right), ws, res, &out, stats);
    if (!s.ok()) {
        last_error_ = s.message;
        return false;
    }
    out.left = left;
    out.right = right;
    out.disparity = res.disparity;
    out.confidence = res.confidence;
    out.invalid_reason = res.invalid_reason;
    out.reliable_mask = res.reliable_mask;
    return true;
}

Status StereoSession::process_status(const ImageView8& left, const ImageView8& right, StereoResult& result,
                                     PipelineStats* stats) {
    last_error_.clear();
    const auto t0 = std::chrono::steady_clock::now();
    const Status s = matcher_.compute_status(left, right, result, workspace_, stats);
    const auto t1 = std::chrono::steady_clock::now();
    if (!s.ok()) {
        last_error_ = s.message;
        return s;
    }
    ++frame_count_;
    total_time_ms_ += std::chrono::duration<double, std::milli>(t1 - t0).count();
    return Status::Ok();
}

std::string PipelineStats::to_json() const {
    std::ostringstream o;
    o << "{\"timing\": {\"total_ms\": " << timing.total_ms << ", \"aux_ms\": " << timing.aux_ms
      << ", \"prior_ms\": " << timing.prior_ms << ", \"cost_ms\": " << timing.cost_ms
      << ", \"cross_ms\": " << timing.cross_ms << ", \"right_wta_ms\": " << timing.right_wta_ms
      << ", \"sgm_ms\": " << timing.sgm_ms << ", \"wta_ms\": " << timing.wta_ms
      << ", \"confidence_ms\": " << timing.confidence_ms << ", \"post_ms\": " << timing.post_ms << "}"
      << ", \"memory\": {\"retained_workspace_bytes\": " << retained_workspace_bytes << "}"
      << ", \"quality\": {\"valid_disparity_ratio\": " << valid_disparity_ratio
      << ", \"reliable_ratio\": " << reliable_ratio << ", \"lr_fail_ratio\": " << lr_fail_ratio << "}}";
    return o.str();
}

}  // namespace apg
