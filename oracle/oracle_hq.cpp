// CPU oracle, HighQuality stages: the sparse-support prior that narrows the
// per-pixel search ranges, and the randomised raster refiner.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).  C++ rather than C because the
// reference's results depend on two libstdc++ behaviours that a restatement
// must share exactly: std::sort's order of equal keys (the weighted medians
// of the prior) and the std::mt19937 + std::uniform_int_distribution<int>
// stream of the refiner.  Float expressions are evaluated in the reference's
// order; build with -ffp-contract=off and no -mfma.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <random>
#include <utility>
#include <vector>

#include "oracle.h"

namespace {

int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

constexpr int kCell = 16;   // prior grid cell (src/prior_estimator.cpp:22, :183)

int pop_lr(const uint32_t* c32l, const uint32_t* c32r, const uint64_t* c64l, const uint64_t* c64r, size_t ql,
           size_t qr) {
    if (c32l) return __builtin_popcount(c32l[ql] ^ c32r[qr]);
    return __builtin_popcountll(c64l[ql] ^ c64r[qr]);
}

}  // namespace

extern "C" {

// PriorEstimator::extract_supports (src/prior_estimator.cpp:13-166): census-only
// WTA with a uniqueness test and a right-view check on every second pixel of
// each 16x16 cell, the first candidates of each cell in raster order, then the
// stratified selection under max_supports.
int orc_prior_supports(const orc_cfg* cfg, int w, int h, const uint8_t* gx_l, const uint8_t* gy_l,
                       const uint32_t* c32l, const uint32_t* c32r, const uint64_t* c64l, const uint64_t* c64r,
                       orc_support* out, int cap) {
    const int d0 = cfg->min_disparity, d1 = cfg->max_disparity;
    if (cfg->census_type != 1) c32l = c32r = nullptr;
    const int gw = (w + kCell - 1) / kCell, gh = (h + kCell - 1) / kCell, ncell = gw * gh;
    const int hard_max = cfg->prior_max_supports;
    const int per_cell = std::max(8, (hard_max * 2 + ncell - 1) / std::max(1, ncell));
    std::vector<std::vector<orc_support>> cand(ncell);
    for (int ci = 0; ci < ncell; ++ci) {
        const int cx = ci % gw, cy = ci / gw;
        const int ylo = std::max(2, cy * kCell), yhi = std::min(h - 2, (cy + 1) * kCell);
        const int xlo = std::max(2, cx * kCell), xhi = std::min(w - 2, (cx + 1) * kCell);
        auto& dst = cand[ci];
        for (int y = ylo + (ylo & 1); y < yhi; y += 2) {
            for (int x = xlo + (xlo & 1); x < xhi; x += 2) {
                if ((int)dst.size() >= per_cell) break;
                const size_t q = (size_t)y * w + x;
                if ((int)gx_l[q] + (int)gy_l[q] < cfg->prior_texture_threshold) continue;
                int bd = d0, b1 = 255, b2 = 255;
                for (int d = d0; d < d1; ++d) {
                    const int xr = x - d;
                    if (xr < 0 || xr >= w) continue;
                    const int c = pop_lr(c32l, c32r, c64l, c64r, q, (size_t)y * w + xr);
                    if (c < b1) {
                        b2 = b1;
                        b1 = c;
                        bd = d;
                    } else if (c < b2) {
                        b2 = c;
                    }
                }
                if (b2 <= 0) continue;
                const float uniq = (float)(b2 - b1) / (float)b2;
                if (uniq < cfg->prior_uniqueness_ratio) continue;
                const int xr = x - bd;
                if (xr < 1 || xr >= w - 1) continue;
                int rb = 255, rd = 0;
                for (int d = d0; d < d1; ++d) {
                    const int xl = xr + d;
                    if (xl < 0 || xl >= w) continue;
                    const int c = pop_lr(c32l, c32r, c64l, c64r, (size_t)y * w + xl, (size_t)y * w + xr);
                    if (c < rb) {
                        rb = c;
                        rd = d;
                    }
                }
                if (std::abs(rd - bd) > cfg->prior_lr_max_diff) continue;
                dst.push_back(orc_support{x, y, (float)bd, uniq});
            }
            if ((int)dst.size() >= per_cell) break;
        }
    }
    std::vector<int> cells;
    for (int i = 0; i < ncell; ++i)
        if (!cand[i].empty()) cells.push_back(i);
    int n = 0;
    auto emit = [&](const orc_support& s) {
        if (n < cap) out[n] = s;
        ++n;
    };
    if ((int)cells.size() <= hard_max) {
        // every non-empty cell gets a base quota, the rest round-robin
        std::vector<int> taken(ncell, 0);
        const int quota = std::max(1, hard_max / std::max(1, (int)cells.size()));
        for (int ci : cells) {
            const int take = std::min(quota, (int)cand[ci].size());
            for (int k = 0; k < take && n < hard_max; ++k) {
                emit(cand[ci][k]);
                ++taken[ci];
            }
        }
        bool more = true;
        while (more && n < hard_max) {
            more = false;
            for (int ci : cells) {
                if (n >= hard_max) break;
                if (taken[ci] < (int)cand[ci].size()) {
                    emit(cand[ci][taken[ci]++]);
                    more = true;
                }
            }
        }
    } else {
        // more non-empty cells than the budget: uniform subsampling of cells
        const size_t m = cells.size();
        for (int k = 0; k < hard_max; ++k) {
            const int ci = cells[(size_t)((uint64_t)k * m / (uint64_t)hard_max)];
            emit(cand[ci].front());
        }
    }
    return n;
}

// PriorEstimator::interpolate_prior (src/prior_estimator.cpp:168-375): per
// cell weighted median, spread and confidence of its supports, empty cells
// filled from their 3x3 neighbours until nothing changes, then bilinear
// interpolation to every pixel.
void orc_prior_interpolate(const orc_support* sup, int nsup, int w, int h, float* d_prior, float* conf,
                           float* spread, float* pmin, float* pmax) {
    const size_t npx = (size_t)w * h;
    std::fill(d_prior, d_prior + npx, -1.f);
    std::fill(conf, conf + npx, 0.f);
    std::fill(spread, spread + npx, 999.f);
    std::fill(pmin, pmin + npx, -1.f);
    std::fill(pmax, pmax + npx, -1.f);
    if (nsup <= 0) return;
    const int gw = (w + kCell - 1) / kCell, gh = (h + kCell - 1) / kCell, ncell = gw * gh;
    std::vector<std::vector<std::pair<float, float>>> per(ncell);
    for (int i = 0; i < nsup; ++i) {
        const int cx = clampi(sup[i].x / kCell, 0, gw - 1), cy = clampi(sup[i].y / kCell, 0, gh - 1);
        per[cy * gw + cx].push_back({sup[i].disparity, sup[i].confidence});
    }
    std::vector<float> gd(ncell, -1.f), gc(ncell, 0.f), gs(ncell, -1.f), gmn(ncell, -1.f), gmx(ncell, -1.f);
    auto by_first = [](const std::pair<float, float>& a, const std::pair<float, float>& b) {
        return a.first < b.first;
    };
    for (int i = 0; i < ncell; ++i) {
        auto& pts = per[i];
        if (pts.empty()) continue;
        std::sort(pts.begin(), pts.end(), by_first);
        gmn[i] = pts.front().first;
        gmx[i] = pts.back().first;
        const float span = gmx[i] - gmn[i];
        float tw = 0.f;
        for (const auto& p : pts) tw += p.second;
        if (tw <= 1e-6f) continue;
        const float half = tw * 0.5f;
        float acc = 0.f, med = pts.front().first;
        for (const auto& p : pts) {
            acc += p.second;
            if (acc >= half) {
                med = p.first;
                break;
            }
        }
        gd[i] = med;
        std::vector<std::pair<float, float>> dev;
        dev.reserve(pts.size());
        for (const auto& p : pts) dev.push_back({std::abs(p.first - med), p.second});
        std::sort(dev.begin(), dev.end(), by_first);
        acc = 0.f;
        float mad = dev.front().first;
        for (const auto& d : dev) {
            acc += d.second;
            if (acc >= half) {
                mad = d.first;
                break;
            }
        }
        gs[i] = std::max(mad * 1.4826f, span * 0.5f);
        const float mean_c = tw / (float)pts.size();
        const float count_f = clampf((float)pts.size() / 3.f, 0.3f, 1.f);
        gc[i] = clampf(mean_c * count_f, 0.f, 1.f);
    }
    // hole filling of the grid, Jacobi style (every sweep reads the previous one)
    for (int it = 0, iters = std::max(gw, gh); it < iters; ++it) {
        std::vector<float> nd = gd, nc = gc, ns = gs, nmn = gmn, nmx = gmx;
        bool changed = false;
        for (int cy = 0; cy < gh; ++cy)
            for (int cx = 0; cx < gw; ++cx) {
                const int i = cy * gw + cx;
                if (gd[i] >= 0.f) continue;
                float sd = 0.f, sc = 0.f, ss = 0.f, cnt = 0.f, lo = 1e9f, hi = -1e9f;
                for (int oy = -1; oy <= 1; ++oy)
                    for (int ox = -1; ox <= 1; ++ox) {
                        const int ux = cx + ox, uy = cy + oy;
                        if (ux < 0 || uy < 0 || ux >= gw || uy >= gh) continue;
                        const int j = uy * gw + ux;
                        const float v = gd[j];
                        if (v < 0.f) continue;
                        sd += v;
                        sc += gc[j];
                        ss += gs[j] >= 0.f ? gs[j] : 8.f;
                        lo = std::min(lo, gmn[j] >= 0.f ? gmn[j] : v);
                        hi = std::max(hi, gmx[j] >= 0.f ? gmx[j] : v);
                        cnt += 1.f;
                    }
                if (cnt > 0.f) {
                    nd[i] = sd / cnt;
                    nc[i] = (sc / cnt) * 0.85f;
                    ns[i] = ss / cnt;
                    nmn[i] = lo;
                    nmx[i] = hi;
                    changed = true;
                }
            }
        gd.swap(nd);
        gc.swap(nc);
        gs.swap(ns);
        gmn.swap(nmn);
        gmx.swap(nmx);
        if (!changed) break;
    }
    for (int y = 0; y < h; ++y) {
        const float fy = ((float)y + 0.5f) / kCell - 0.5f;
        const int y0 = clampi((int)std::floor(fy), 0, gh - 1), y1 = std::min(y0 + 1, gh - 1);
        const float ty = clampf(fy - (float)y0, 0.f, 1.f);
        for (int x = 0; x < w; ++x) {
            const float fx = ((float)x + 0.5f) / kCell - 0.5f;
            const int x0 = clampi((int)std::floor(fx), 0, gw - 1), x1 = std::min(x0 + 1, gw - 1);
            const float tx = clampf(fx - (float)x0, 0.f, 1.f);
            const int k[4] = {y0 * gw + x0, y0 * gw + x1, y1 * gw + x0, y1 * gw + x1};
            auto has = [&](int j) { return gd[j] < 0.f ? 0.f : 1.f; };
            auto pos = [](float v) { return v < 0.f ? 0.f : v; };
            const float w0 = (1.f - tx) * (1.f - ty) * has(k[0]);
            const float w1 = tx * (1.f - ty) * has(k[1]);
            const float w2 = (1.f - tx) * ty * has(k[2]);
            const float w3 = tx * ty * has(k[3]);
            const float ws = w0 + w1 + w2 + w3;
            const size_t q = (size_t)y * w + x;
            if (!(ws > 1e-6f)) continue;   // outputs keep their "no prior" values
            const float dp = (w0 * pos(gd[k[0]]) + w1 * pos(gd[k[1]]) + w2 * pos(gd[k[2]]) + w3 * pos(gd[k[3]])) / ws;
            d_prior[q] = dp;
            conf[q] = (w0 * gc[k[0]] + w1 * gc[k[1]] + w2 * gc[k[2]] + w3 * gc[k[3]]) / ws;
            float lo = 1e9f, hi = -1e9f;
            for (int j : k)
                if (gd[j] >= 0.f) {
                    lo = std::min(lo, gmn[j] >= 0.f ? gmn[j] : gd[j]);
                    hi = std::max(hi, gmx[j] >= 0.f ? gmx[j] : gd[j]);
                }
            spread[q] = (w0 * pos(gs[k[0]]) + w1 * pos(gs[k[1]]) + w2 * pos(gs[k[2]]) + w3 * pos(gs[k[3]])) / ws;
            pmin[q] = lo <= hi ? std::min(lo, dp) : dp;
            pmax[q] = lo <= hi ? std::max(hi, dp) : dp;
        }
    }
}

// PriorEstimator::apply_search_range (src/prior_estimator.cpp:377-412) with
// SearchRange::set_pixel's clamp to the global range (search_range.hpp:29-35).
void orc_prior_apply_range(const orc_cfg* cfg, int w, int h, const float* d_prior, const float* conf,
                           const float* spread, const float* pmin, const float* pmax, int16_t* dmin, int16_t* dmax) {
    const int gD = cfg->max_disparity - cfg->min_disparity;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const size_t q = (size_t)y * w + x;
            const float dp = d_prior[q];
            if (dp < 0.f) continue;
            const float c = conf[q], s = spread[q];
            const float lo_p = pmin[q] >= 0.f ? pmin[q] : dp, hi_p = pmax[q] >= 0.f ? pmax[q] : dp;
            int R;
            if (c < 0.25f) R = gD;
            else if (c > 0.85f && s < 1.5f) R = 8;
            else if (c > 0.60f && s < 3.0f) R = (int)(12.f + 1.5f * s + 0.5f);
            else if (c > 0.40f && s < 6.0f) R = (int)(18.f + 2.0f * s + 0.5f);
            else R = std::max(24, (int)(16.f + 2.0f * s + 0.5f));
            R = std::min(R, gD);
            const int blo = (int)std::floor(lo_p) - 4, bhi = (int)std::ceil(hi_p) + 4 + 1;
            int lo = std::min((int)std::floor(dp) - R, blo);
            int hi = std::max((int)std::ceil(dp) + R + 1, bhi);
            hi = std::min(hi, x + 1);
            lo = std::max(lo, cfg->min_disparity);
            hi = std::min(hi, cfg->max_disparity);
            if (hi < lo) hi = lo;
            dmin[q] = (int16_t)lo;
            dmax[q] = (int16_t)hi;
        }
}

// PriorEstimator::estimate (src/prior_estimator.cpp:414-435): dense range
// init, then (prior enabled and enough supports) the narrowed ranges.
// d_prior is -1 wherever there is no prior.  Returns the support count.
int orc_prior_estimate(const orc_cfg* cfg, int w, int h, const uint8_t* gx_l, const uint8_t* gy_l,
                       const uint32_t* c32l, const uint32_t* c32r, const uint64_t* c64l, const uint64_t* c64r,
                       int16_t* dmin, int16_t* dmax, float* d_prior) {
    const size_t npx = (size_t)w * h;
    orc_range_init(cfg, w, h, dmin, dmax);
    std::fill(d_prior, d_prior + npx, -1.f);
    if (!cfg->prior_enable) return 0;
    std::vector<orc_support> sup(std::max(1, cfg->prior_max_supports));
    const int n = orc_prior_supports(cfg, w, h, gx_l, gy_l, c32l, c32r, c64l, c64r, sup.data(), (int)sup.size());
    if (n < cfg->prior_min_supports) return n;
    std::vector<float> conf(npx), spread(npx), pmin(npx), pmax(npx);
    orc_prior_interpolate(sup.data(), n, w, h, d_prior, conf.data(), spread.data(), pmin.data(), pmax.data());
    orc_prior_apply_range(cfg, w, h, d_prior, conf.data(), spread.data(), pmin.data(), pmax.data(), dmin, dmax);
    return n;
}

// Refiner (src/refiner.cpp:14-98): up to `iterations` raster sweeps,
// alternating direction, over the unreliable pixels; each keeps the cheapest
// of its rounded disparity, the two already-visited neighbours and one random
// jitter, by the unquantised cost of the discrete LUT
// (src/discrete_cost_lut.hpp:10-38).  The jitter stream is std::mt19937(12345)
// through std::uniform_int_distribution<int>(-r, r), one draw per pixel whose
// current disparity is >= 0.
void orc_refine(const orc_cfg* cfg, int w, int h, const uint8_t* gl, const uint8_t* gxl, const uint8_t* gyl,
                const uint8_t* gr, const uint8_t* gxr, const uint8_t* gyr, const uint32_t* c32l,
                const uint32_t* c32r, const uint64_t* c64l, const uint64_t* c64r, const int16_t* dmin,
                const int16_t* dmax, const uint8_t* reliable, float* disp) {
    if (!cfg->refine_enable) return;
    if (cfg->census_type != 1) c32l = c32r = nullptr;
    float lc_[65], la_[256] = {}, lg_[511] = {};
    const float lc = std::max(cfg->lambda_census, 1e-3f), la = std::max(cfg->lambda_ad, 1e-3f),
                lg = std::max(cfg->lambda_grad, 1e-3f);
    for (int i = 0; i <= 64; ++i) lc_[i] = 1.f - std::exp(-(float)i / lc);
    if (cfg->use_ad)
        for (int i = 0; i < 256; ++i) la_[i] = cfg->eta_ad * (1.f - std::exp(-(float)i / la));
    if (cfg->use_grad)
        for (int i = 0; i < 511; ++i) lg_[i] = cfg->mu_grad * (1.f - std::exp(-(float)i / lg));
    auto cost = [&](int x, int y, int d) -> float {
        const int xr = x - d;
        const size_t q = (size_t)y * w + x, qr = (size_t)y * w + xr;
        if (xr < 0 || xr >= w || d < dmin[q] || d >= dmax[q]) return 1e6f;
        float c = lc_[pop_lr(c32l, c32r, c64l, c64r, q, qr)];
        if (cfg->use_ad) c += la_[std::abs((int)gl[q] - (int)gr[qr])];
        if (cfg->use_grad) c += lg_[std::abs((int)gxl[q] - (int)gxr[qr]) + std::abs((int)gyl[q] - (int)gyr[qr])];
        return c;
    };
    std::mt19937 rng(12345);
    for (int it = 0; it < cfg->refine_iterations; ++it) {
        const bool fwd = it % 2 == 0;
        const int r = std::max(1, cfg->refine_random_radius >> it);
        std::uniform_int_distribution<int> jitter(-r, r);
        for (int k = 0; k < h; ++k) {
            const int y = fwd ? k : h - 1 - k;
            for (int m = 0; m < w; ++m) {
                const int x = fwd ? m : w - 1 - m;
                const size_t q = (size_t)y * w + x;
                if (reliable[q]) continue;
                const int lo = dmin[q], hi = dmax[q];
                if (hi <= lo) continue;
                const float cur = disp[q];
                int cands[6], nc = 0;
                auto add = [&](float d) {
                    if (d >= 0.f) cands[nc++] = clampi((int)std::round(d), lo, hi - 1);
                };
                add(cur);
                if (fwd) {
                    if (x > 0) add(disp[q - 1]);
                    if (y > 0) add(disp[q - w]);
                } else {
                    if (x + 1 < w) add(disp[q + 1]);
                    if (y + 1 < h) add(disp[q + w]);
                }
                if (cur >= 0.f) add(cur + (float)jitter(rng));
                float bc = 1e9f;
                int bd = cur >= 0.f ? clampi((int)std::round(cur), lo, hi - 1) : lo;
                for (int i = 0; i < nc; ++i) {
                    const float c = cost(x, y, cands[i]);
                    if (c < bc) {
                        bc = c;
                        bd = cands[i];
                    }
                }
                disp[q] = (float)bd;
            }
        }
    }
}

// std::uniform_int_distribution<int>(-r, r) applied ndraw times to an engine
// that replays `stream` (32-bit outputs; past its end it repeats the last):
// the draws in out, the number of outputs consumed in *consumed.  The
// refiner's jitter is this with the std::mt19937(12345) stream.
void orc_uniform_draws(const uint32_t* stream, size_t n, int r, int ndraw, int* out, size_t* consumed) {
    struct Replay {
        using result_type = uint32_t;
        const uint32_t* s;
        size_t n, i;
        static constexpr result_type min() { return 0u; }
        static constexpr result_type max() { return 0xFFFFFFFFu; }
        result_type operator()() { return s[i < n ? i++ : (i++, n - 1)]; }
    } eng{stream, n, 0};
    std::uniform_int_distribution<int> dist(-r, r);
    for (int k = 0; k < ndraw; ++k) out[k] = dist(eng);
    *consumed = eng.i;
}

}  // extern "C"
