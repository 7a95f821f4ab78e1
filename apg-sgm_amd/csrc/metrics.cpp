// Accuracy metrics (definitions of reference src/metrics.cpp:8-505): the
// same pixel sets, tests and arithmetic types (double sums, float ratios in
// the same operation order), so the printed values agree.
#include "apg_sgm/metrics.hpp"

#include <math.h>

#include <algorithm>
#include <cmath>

namespace apg {

namespace {

bool gt_ok(float g, float max_valid) { return std::isfinite(g) && g >= 0.f && g < max_valid; }
bool est_ok(float e) { return std::isfinite(e) && e >= 0.f; }

// x - round(g) inside the image: the pixel is matchable; visible is the mask
// (>= 200) when one is given, else matchable
struct Visibility {
    const Image8* mask;
    int w;
    bool operator()(int x, int y, float g) const {
        const int xr = x - static_cast<int>(std::round(g));
        const bool matchable = xr >= 0 && xr < w;
        return (mask && !mask->empty()) ? mask->at(x, y) >= 200 : matchable;
    }
    bool matchable(int x, float g) const {
        const int xr = x - static_cast<int>(std::round(g));
        return xr >= 0 && xr < w;
    }
};

struct RegionSum {
    int total = 0, valid = 0, c05 = 0, c10 = 0, c20 = 0, c30 = 0, d1 = 0;
    double sum = 0.0;
    void add(float e, float g) {
        ++total;
        const bool ok = est_ok(e);
        const float diff = ok ? std::abs(e - g) : 1e9f;
        if (!ok || (diff > 3.0f && (g > 1e-3f ? diff / g > 0.05f : true))) ++d1;
        if (!ok) return;
        ++valid;
        sum += diff;
        c05 += diff > 0.5f;
        c10 += diff > 1.0f;
        c20 += diff > 2.0f;
        c30 += diff > 3.0f;
    }
    RegionMetrics result() const {
        RegionMetrics r;
        r.total_pixels = total;
        r.valid_est_pixels = valid;
        if (total > 0) {
            r.valid_ratio = static_cast<float>(valid) / static_cast<float>(total);
            r.d1 = static_cast<float>(d1) * 100.0f / static_cast<float>(total);
        }
        if (valid > 0) {
            r.epe = static_cast<float>(sum / valid);
            r.bad_0_5 = static_cast<float>(c05) * 100.0f / static_cast<float>(valid);
            r.bad_1_0 = static_cast<float>(c10) * 100.0f / static_cast<float>(valid);
            r.bad_2_0 = static_cast<float>(c20) * 100.0f / static_cast<float>(valid);
            r.bad_3_0 = static_cast<float>(c30) * 100.0f / static_cast<float>(valid);
        }
        return r;
    }
};

}  // namespace

SupportMetrics evaluate_supports(const std::vector<SupportMatch>& supports, const Image32f& gt, const Image8* vis_mask,
                                 float max_valid_gt, int cell_size) {
    SupportMetrics sm;
    sm.total = static_cast<int>(supports.size());
    if (supports.empty() || gt.empty()) return sm;
    const int w = gt.width(), h = gt.height();
    const int gw = (w + cell_size - 1) / cell_size, gh = (h + cell_size - 1) / cell_size;
    const Visibility vis{vis_mask, w};
    std::vector<uint8_t> cell_vis(static_cast<size_t>(gw) * gh, 0), cell_hit(static_cast<size_t>(gw) * gh, 0);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const float g = gt.at(x, y);
            if (gt_ok(g, max_valid_gt) && vis(x, y, g)) cell_vis[static_cast<size_t>(y / cell_size) * gw + x / cell_size] = 1;
        }
    double sum_all = 0.0, sum_vis = 0.0;
    for (const SupportMatch& s : supports) {
        if (s.x < 0 || s.x >= w || s.y < 0 || s.y >= h) continue;
        const float g = gt.at(s.x, s.y);
        if (!gt_ok(g, max_valid_gt)) continue;
        ++sm.gt_valid;
        const bool v = vis(s.x, s.y, g);
        sm.visible += v;
        const float err = std::abs(s.disparity - g);
        sum_all += err;
        if (v) sum_vis += err;
        if (err <= 0.5f) {
            ++sm.correct_05;
            sm.visible_correct_05 += v;
        }
        if (err <= 1.0f) {
            ++sm.correct_1;
            if (v) {
                ++sm.visible_correct_1;
                cell_hit[static_cast<size_t>(s.y / cell_size) * gw + s.x / cell_size] = 1;
            }
        }
        if (err <= 2.0f) {
            ++sm.correct_2;
            sm.visible_correct_2 += v;
        }
    }
    if (sm.gt_valid > 0) {
        const float n = static_cast<float>(sm.gt_valid);
        sm.precision_all_05 = sm.precision_05 = static_cast<float>(sm.correct_05) / n;
        sm.precision_all_1 = sm.precision_1 = static_cast<float>(sm.correct_1) / n;
        sm.precision_all_2 = sm.precision_2 = static_cast<float>(sm.correct_2) / n;
        sm.mae_all = sm.mean_abs_error = static_cast<float>(sum_all / sm.gt_valid);
    }
    if (sm.visible > 0) {
        const float n = static_cast<float>(sm.visible);
        sm.precision_vis_05 = static_cast<float>(sm.visible_correct_05) / n;
        sm.precision_vis_1 = static_cast<float>(sm.visible_correct_1) / n;
        sm.precision_vis_2 = static_cast<float>(sm.visible_correct_2) / n;
        sm.mae_visible = static_cast<float>(sum_vis / sm.visible);
    }
    for (size_t i = 0; i < cell_vis.size(); ++i) {
        sm.visible_gt_cells += cell_vis[i];
        sm.correct_visible_cells += cell_vis[i] && cell_hit[i];
    }
    if (sm.visible_gt_cells > 0)
        sm.grid_recall_1 = sm.grid_coverage = sm.support_grid_recall_vis_1 =
            static_cast<float>(sm.correct_visible_cells) / static_cast<float>(sm.visible_gt_cells);
    return sm;
}

StereoMetrics evaluate_stereo(const Image32f& est, const Image32f& gt, const SearchRange* range,
                              const Image32f* d_before_refine, const Image32f* d_after_refine, const Image8* vis_mask,
                              float max_valid_gt, const std::vector<SupportMatch>* supports, const Image8* left_image,
                              const QualityThresholds& th) {
    StereoMetrics m;
    const int w = est.width(), h = est.height();
    if (w <= 0 || h <= 0 || gt.width() != w || gt.height() != h) return m;
    m.total_pixels = w * h;
    const size_t n = static_cast<size_t>(w) * h;
    const Visibility vis{vis_mask, w};

    // disparity boundaries: a GT-valid pixel whose valid 4-neighbours differ by >= the threshold
    std::vector<uint8_t> edge(n, 0);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const float g = gt.at(x, y);
            if (!gt_ok(g, max_valid_gt)) continue;
            float md = 0.f;
            const int nb[4][2] = {{x + 1, y}, {x - 1, y}, {x, y + 1}, {x, y - 1}};
            for (const auto& q : nb) {
                if (q[0] < 0 || q[0] >= w || q[1] < 0 || q[1] >= h) continue;
                const float ng = gt.at(q[0], q[1]);
                if (gt_ok(ng, max_valid_gt)) md = std::max(md, std::abs(g - ng));
            }
            edge[static_cast<size_t>(y) * w + x] = md >= th.boundary_disparity_gradient;
        }
    // low texture: central-difference luma gradient of the left image (one-sided at the borders)
    const bool tex = left_image && !left_image->empty() && left_image->width() == w && left_image->height() == h;
    std::vector<uint8_t> lowtex(tex ? n : 0, 0);
    if (tex) {
        const Image8& L = *left_image;
        auto luma = [&](int x, int y) -> float {
            if (L.channels() == 1) return static_cast<float>(L.at(x, y));
            return 0.299f * L.at(x, y, 0) + 0.587f * L.at(x, y, 1) + 0.114f * L.at(x, y, 2);
        };
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                const float sx = (x == 0 || x == w - 1) ? 1.0f : 2.0f, sy = (y == 0 || y == h - 1) ? 1.0f : 2.0f;
                const float dx = (luma(std::min(w - 1, x + 1), y) - luma(std::max(0, x - 1), y)) / sx;
                const float dy = (luma(x, std::min(h - 1, y + 1)) - luma(x, std::max(0, y - 1))) / sy;
                lowtex[static_cast<size_t>(y) * w + x] = std::sqrt(dx * dx + dy * dy) < th.low_texture_gradient;
            }
    }

    double sum = 0.0, sum_edge = 0.0, sum_nonedge = 0.0;
    int c05 = 0, c10 = 0, c20 = 0, c30 = 0, d1_all = 0, d1_noc = 0, noc = 0, lr_fail = 0;
    int r_all = 0, r_in = 0, r_mat = 0, r_mat_in = 0, r_vis = 0, r_vis_in = 0;
    RegionSum a_all, a_noc, a_occ, a_bnd, a_nbnd, a_low, a_high;
    const bool refine = d_before_refine && !d_before_refine->empty() && d_after_refine && !d_after_refine->empty();
    double rb = 0.0, ra = 0.0, tb = 0.0, ta = 0.0, tf = 0.0;
    int rn = 0, r_impr = 0, r_worse = 0, r_same = 0, r_changed = 0, tn = 0;

    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const float g = gt.at(x, y);
            if (!gt_ok(g, max_valid_gt)) continue;
            ++m.evaluated_pixels;
            const size_t i = static_cast<size_t>(y) * w + x;
            const bool v = vis(x, y, g), mt = vis.matchable(x, g), eg = edge[i] != 0;
            if (range) {
                m.has_range = true;
                const bool in = g >= static_cast<float>(range->dmin.at(x, y)) && g < static_cast<float>(range->dmax.at(x, y));
                ++r_all;
                r_in += in;
                if (mt) {
                    ++r_mat;
                    r_mat_in += in;
                }
                if (v) {
                    ++r_vis;
                    if (in) ++r_vis_in;
                    else {
                        ++m.prior_miss_visible_count;
                        if (eg) ++m.prior_miss_edge;
                        else ++m.prior_miss_nonedge;
                    }
                }
            }
            const float e = est.at(x, y);
            if (!est_ok(e)) {
                ++lr_fail;
            } else {
                ++m.valid_pixels;
                const float diff = std::abs(e - g);
                sum += diff;
                c05 += diff > 0.5f;
                c10 += diff > 1.0f;
                c20 += diff > 2.0f;
                c30 += diff > 3.0f;
                if (eg) {
                    ++m.edge_pixels;
                    sum_edge += diff;
                } else {
                    ++m.nonedge_pixels;
                    sum_nonedge += diff;
                }
            }
            const bool d1 = !est_ok(e) || (std::abs(e - g) > 3.0f && (g > 1e-3f ? std::abs(e - g) / g > 0.05f : true));
            d1_all += d1;
            if (v) {
                ++noc;
                d1_noc += d1;
            }
            a_all.add(e, g);
            (v ? a_noc : a_occ).add(e, g);
            (eg ? a_bnd : a_nbnd).add(e, g);
            if (tex) (lowtex[i] ? a_low : a_high).add(e, g);
            if (refine) {
                const float eb = d_before_refine->at(x, y), ea = d_after_refine->at(x, y);
                if (est_ok(eb) && est_ok(ea)) {
                    ++rn;
                    const float db = std::abs(eb - g), da = std::abs(ea - g);
                    rb += db;
                    ra += da;
                    if (da < db - 1e-4f) ++r_impr;
                    else if (da > db + 1e-4f) ++r_worse;
                    else ++r_same;
                    r_changed += std::abs(ea - eb) > 1e-4f;
                    if (est_ok(e)) {
                        ++tn;
                        tb += db;
                        ta += da;
                        tf += std::abs(e - g);
                    }
                }
            }
        }

    if (m.evaluated_pixels > 0) {
        m.valid_ratio = static_cast<float>(m.valid_pixels) / static_cast<float>(m.evaluated_pixels);
        m.lr_fail_ratio = static_cast<float>(lr_fail) / static_cast<float>(m.evaluated_pixels);
        m.kitti_d1_all = static_cast<float>(d1_all) / static_cast<float>(m.evaluated_pixels) * 100.0f;
    }
    if (range) {
        m.has_range = true;
        m.range_recall_all = r_all > 0 ? static_cast<float>(r_in) / r_all : 0.f;
        m.range_recall_matchable = r_mat > 0 ? static_cast<float>(r_mat_in) / r_mat : 0.f;
        m.range_recall_visible = r_vis > 0 ? static_cast<float>(r_vis_in) / r_vis : 0.f;
        m.range_visible_eval_pixels = static_cast<size_t>(r_vis);
        m.range_visible_in_pixels = static_cast<size_t>(r_vis_in);
        m.range_matchable_eval_pixels = static_cast<size_t>(r_mat);
        m.range_matchable_in_pixels = static_cast<size_t>(r_mat_in);
        if (m.prior_miss_visible_count > 0) {
            m.prior_miss_edge_ratio = static_cast<float>(m.prior_miss_edge) / m.prior_miss_visible_count;
            m.prior_miss_nonedge_ratio = static_cast<float>(m.prior_miss_nonedge) / m.prior_miss_visible_count;
        }
    }
    if (supports && !supports->empty()) {
        m.has_supports = true;
        m.support_metrics = evaluate_supports(*supports, gt, vis_mask, max_valid_gt);
    }
    if (m.valid_pixels > 0) {
        const float nv = static_cast<float>(m.valid_pixels);
        m.epe = static_cast<float>(sum / m.valid_pixels);
        m.bad_0_5 = static_cast<float>(c05) / nv * 100.0f;
        m.bad_1_0 = static_cast<float>(c10) / nv * 100.0f;
        m.bad_2_0 = static_cast<float>(c20) / nv * 100.0f;
        m.bad_3_0 = static_cast<float>(c30) / nv * 100.0f;
    }
    if (noc > 0) m.kitti_d1_noc = static_cast<float>(d1_noc) / static_cast<float>(noc) * 100.0f;
    if (m.edge_pixels > 0) m.edge_epe = static_cast<float>(sum_edge / m.edge_pixels);
    if (m.nonedge_pixels > 0) m.nonedge_epe = static_cast<float>(sum_nonedge / m.nonedge_pixels);
    if (rn > 0) {
        m.has_refine_stats = true;
        m.refine_eval_pixels = rn;
        m.refine_epe_delta = static_cast<float>((ra - rb) / rn);
        m.refine_improved_ratio = static_cast<float>(r_impr) / static_cast<float>(rn);
        m.refine_worsened_ratio = static_cast<float>(r_worse) / static_cast<float>(rn);
        m.refine_unchanged_ratio = static_cast<float>(r_same) / static_cast<float>(rn);
        m.refine_changed_ratio = static_cast<float>(r_changed) / static_cast<float>(rn);
    }
    if (tn > 0) {
        m.post_epe_delta = static_cast<float>((tf - ta) / tn);
        m.total_epe_delta = static_cast<float>((tf - tb) / tn);
    }
    m.reg_all = a_all.result();
    m.reg_noc = a_noc.result();
    m.reg_occ = a_occ.result();
    m.reg_boundary = a_bnd.result();
    m.reg_non_boundary = a_nbnd.result();
    m.reg_low_texture = a_low.result();
    m.reg_high_texture = a_high.result();
    return m;
}

}  // namespace apg
