/*
 * CPU oracle: plain-C restatement of the apg-sgm dense hot path.
 * TEST INFRASTRUCTURE ONLY -- see oracle.h for scope and pinning.
 * Build: gcc -O2 -fopenmp -ffp-contract=off (no -mfma, no -ffast-math).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define INV16 ((uint16_t)65535u)           /* kInvalidCost   include/apg_sgm/types.hpp:44 */
#define INV32 ((uint32_t)0xFFFFFFFFu)      /* kInvalidCost32 include/apg_sgm/types.hpp:45 */
#define PATH_INF (1 << 28)                 /* kPathInf       include/apg_sgm/types.hpp:46 */

enum { R_VALID = 0, R_OCCLUSION = 1, R_MISMATCH = 2, R_LOWCONF = 3, R_OUTOFRANGE = 4 };

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline float fmaxf_(float a, float b) { return a < b ? b : a; } /* std::max(a,b) */

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* src/config.cpp:5-45 with the member defaults of include/apg_sgm/config.hpp:7-66 */
void orc_config_from_mode(int mode, int max_disparity, orc_cfg* c) {
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
    if (mode == 0) {            /* Fast */
        c->use_ad = 0; c->use_grad = 0; c->agg_enable = 0;
        c->paths = 4; c->adaptive_penalties = 0;
    } else if (mode == 1) {     /* Balanced */
        c->agg_iterations = 1; c->paths = 4; c->adaptive_penalties = 1;
    } else {                    /* HighQuality */
        c->paths = 8; c->adaptive_penalties = 1;
        c->prior_enable = 1; c->refine_enable = 1; c->refine_iterations = 3;
    }
}

/* ---------------------------------------------------------------- aux ---- */

/* ImageView8::gray (include/apg_sgm/image_view.hpp:50-56) and
 * Image8::sample clamp-to-edge (src/image.cpp:15-19);
 * build_gray_and_grad (src/cost_computer.cpp:63-88). */
void orc_gray_grad(const uint8_t* src, int w, int h, int stride, int ch,
                   uint8_t* gray, uint8_t* gx, uint8_t* gy) {
    for (int y = 0; y < h; ++y) {
        const uint8_t* row = src + (size_t)y * stride;
        for (int x = 0; x < w; ++x) {
            uint8_t v;
            if (ch == 1) {
                v = row[x];
            } else {
                const int r = row[x * ch + 0], g = row[x * ch + 1], b = row[x * ch + 2];
                v = (uint8_t)((77 * r + 150 * g + 29 * b) >> 8);
            }
            gray[(size_t)y * w + x] = v;
        }
    }
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const int xp = clampi(x + 1, 0, w - 1), xm = clampi(x - 1, 0, w - 1);
            const int yp = clampi(y + 1, 0, h - 1), ym = clampi(y - 1, 0, h - 1);
            const int gxv = (int)gray[(size_t)y * w + xp] - (int)gray[(size_t)y * w + xm];
            const int gyv = (int)gray[(size_t)yp * w + x] - (int)gray[(size_t)ym * w + x];
            gx[(size_t)y * w + x] = (uint8_t)clampi(abs(gxv), 0, 255);
            gy[(size_t)y * w + x] = (uint8_t)clampi(abs(gyv), 0, 255);
        }
    }
}

static inline uint8_t samp(const uint8_t* g, int w, int h, int x, int y) {
    return g[(size_t)clampi(y, 0, h - 1) * w + clampi(x, 0, w - 1)];
}

/* CostComputer::symmetric_census9x7 (src/cost_computer.cpp:33-47): pairs
 * (dx,dy) with dy<0 || (dy==0 && dx<0) in raster order, bit b = I(p+o) > I(p-o). */
void orc_census_sym(const uint8_t* gray, int w, int h, uint32_t* out) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            uint32_t bits = 0;
            int b = 0;
            for (int dy = -3; dy <= 3; ++dy)
                for (int dx = -4; dx <= 4; ++dx)
                    if (dy < 0 || (dy == 0 && dx < 0)) {
                        if (samp(gray, w, h, x + dx, y + dy) > samp(gray, w, h, x - dx, y - dy))
                            bits |= 1u << b;
                        ++b;
                    }
            out[(size_t)y * w + x] = bits;
        }
    }
}

/* CostComputer::census9x7 (src/cost_computer.cpp:49-61) */
void orc_census64(const uint8_t* gray, int w, int h, uint64_t* out) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const uint8_t c = samp(gray, w, h, x, y);
            uint64_t bits = 0;
            int b = 0;
            for (int dy = -3; dy <= 3; ++dy)
                for (int dx = -4; dx <= 4; ++dx) {
                    if (dx == 0 && dy == 0) continue;
                    if (samp(gray, w, h, x + dx, y + dy) > c) bits |= 1ull << b;
                    ++b;
                }
            out[(size_t)y * w + x] = bits;
        }
    }
}

/* PriorEstimator::estimate with the prior off (src/prior_estimator.cpp:414-424)
 * through SearchRange::set_pixel (include/apg_sgm/search_range.hpp:29-35). */
void orc_range_init(const orc_cfg* cfg, int w, int h, int16_t* dmin, int16_t* dmax) {
    const int gmin = cfg->min_disparity, gmax = cfg->max_disparity;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            int lo = cfg->min_disparity;
            int hi = imin(cfg->max_disparity, x + 1);
            lo = imax(lo, gmin);
            hi = imin(hi, gmax);
            if (hi < lo) hi = lo;
            dmin[(size_t)y * w + x] = (int16_t)lo;
            dmax[(size_t)y * w + x] = (int16_t)hi;
        }
}

/* --------------------------------------------------------------- cost ---- */

/* CostComputer::compute_volume (src/cost_computer.cpp:139-203). */
void orc_cost_volume(const orc_cfg* cfg, int w, int h,
                     const uint8_t* gray_l, const uint8_t* gx_l, const uint8_t* gy_l,
                     const uint8_t* gray_r, const uint8_t* gx_r, const uint8_t* gy_r,
                     const uint32_t* c32l, const uint32_t* c32r,
                     const uint64_t* c64l, const uint64_t* c64r,
                     const int16_t* dmin, const int16_t* dmax, uint16_t* cost) {
    const int d0 = cfg->min_disparity, d1 = cfg->max_disparity, D = d1 - d0;
    const int sym = cfg->census_type == 1;
    const int use_ad = cfg->use_ad, use_grad = cfg->use_grad;
    const float lc = fmaxf_(cfg->lambda_census, 1e-3f);
    const float la = fmaxf_(cfg->lambda_ad, 1e-3f);
    const float lg = fmaxf_(cfg->lambda_grad, 1e-3f);
    const float eta = cfg->eta_ad, mu = cfg->mu_grad;
    const float norm = 1.f + (use_ad ? eta : 0.f) + (use_grad ? mu : 0.f);
    const float scale = (float)cfg->cost_max / fmaxf_(norm, 1e-6f);
#pragma omp parallel for schedule(dynamic, 4)
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const size_t p = (size_t)y * w + x;
            const int lo = dmin[p], hi = dmax[p];
            uint16_t* s = cost + p * D;
            for (int d = d0; d < d1; ++d) {
                const int di = d - d0;
                if (d < lo || d >= hi) { s[di] = INV16; continue; }
                const int xr = x - d;
                if (xr < 0 || xr >= w) { s[di] = INV16; continue; }
                const size_t q = (size_t)y * w + xr;
                float cc;
                if (sym) cc = (float)__builtin_popcount(c32l[p] ^ c32r[q]);
                else cc = (float)__builtin_popcountll(c64l[p] ^ c64r[q]);
                float c = 1.f - expf(-cc / lc);
                if (use_ad) {
                    const int ad = abs((int)gray_l[p] - (int)gray_r[q]);
                    c += eta * (1.f - expf(-(float)ad / la));
                }
                if (use_grad) {
                    const int g = abs((int)gx_l[p] - (int)gx_r[q]) + abs((int)gy_l[p] - (int)gy_r[q]);
                    c += mu * (1.f - expf(-(float)g / lg));
                }
                const int qv = (int)(c * scale + 0.5f);
                s[di] = (uint16_t)clampi(qv, 0, cfg->cost_max);
            }
        }
    }
}

/* legacy parabola (src/subpixel_estimator.hpp:115-124) */
float orc_subpixel_legacy(float cm, float c0, float cp) {
    const float denom = cm - 2.0f * c0 + cp;
    if (fabsf(denom) > 1e-6f) {
        const float delta = 0.5f * (cm - cp) / denom;
        if (delta < -0.5f) return -0.5f;
        if (delta > 0.5f) return 0.5f;
        return delta;
    }
    return 0.0f;
}

/* wta_right_from_left_volume (src/pipeline.cpp:24-85) */
void orc_right_wta(const orc_cfg* cfg, int w, int h, const uint16_t* vol, float* dr) {
    const int d0 = cfg->min_disparity, D = cfg->max_disparity - cfg->min_disparity;
#pragma omp parallel for schedule(dynamic, 4)
    for (int y = 0; y < h; ++y) {
        for (int xr = 0; xr < w; ++xr) {
            int best_d = -1;
            uint32_t best = 0xFFFFFFFFu;
            for (int d = d0; d < d0 + D; ++d) {
                const int xl = xr + d;
                if (xl < 0 || xl >= w) continue;
                const uint16_t c = vol[((size_t)y * w + xl) * D + (d - d0)];
                if (c == INV16) continue;
                if (c < best) { best = c; best_d = d; }
            }
            float dval = -1.f;
            if (best_d >= 0) {
                dval = (float)best_d;
                if (cfg->post_subpixel) {
                    const int di = best_d - d0;
                    const int xm = xr + best_d - 1, x0 = xr + best_d, xp = xr + best_d + 1;
                    if (xm >= 0 && xm < w && x0 >= 0 && x0 < w && xp >= 0 && xp < w && di > 0 && di + 1 < D) {
                        const uint16_t cm = vol[((size_t)y * w + xm) * D + di - 1];
                        const uint16_t c0 = vol[((size_t)y * w + x0) * D + di];
                        const uint16_t cp = vol[((size_t)y * w + xp) * D + di + 1];
                        if (cm != INV16 && c0 != INV16 && cp != INV16)
                            dval += orc_subpixel_legacy((float)cm, (float)c0, (float)cp);
                    }
                }
            }
            dr[(size_t)y * w + xr] = dval;
        }
    }
}

/* -------------------------------------------------------------- cross ---- */

/* CostAggregator::build_cross_arms (src/cost_aggregator.cpp:17-63) */
void orc_cross_arms(const orc_cfg* cfg, const uint8_t* g, int w, int h,
                    int16_t* La, int16_t* Ra, int16_t* Ua, int16_t* Da) {
    const int Lmax = imax(1, cfg->max_arm_length);
    const int tau = cfg->color_threshold;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const int c0 = g[(size_t)y * w + x];
            int l = 0, r = 0, u = 0, v = 0;
            while (l < Lmax && x - l - 1 >= 0 && abs((int)g[(size_t)y * w + x - l - 1] - c0) <= tau) ++l;
            while (r < Lmax && x + r + 1 < w && abs((int)g[(size_t)y * w + x + r + 1] - c0) <= tau) ++r;
            while (u < Lmax && y - u - 1 >= 0 && abs((int)g[(size_t)(y - u - 1) * w + x] - c0) <= tau) ++u;
            while (v < Lmax && y + v + 1 < h && abs((int)g[(size_t)(y + v + 1) * w + x] - c0) <= tau) ++v;
            const size_t i = (size_t)y * w + x;
            La[i] = (int16_t)l; Ra[i] = (int16_t)r; Ua[i] = (int16_t)u; Da[i] = (int16_t)v;
        }
    }
}

/* horizontal pass of CostAggregator::aggregate_hv (src/cost_aggregator.cpp:74-105) */
void orc_cross_h(const orc_cfg* cfg, int w, int h, const int16_t* dmin, const int16_t* dmax,
                 const int16_t* La, const int16_t* Ra, const uint16_t* cost, uint16_t* tmp) {
    const int d0 = cfg->min_disparity, D = cfg->max_disparity - cfg->min_disparity;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const size_t i = (size_t)y * w + x;
            const int x0 = x - La[i], x1 = x + Ra[i];
            const int lo = dmin[i], hi = dmax[i];
            uint16_t* dst = tmp + i * D;
            for (int di = 0; di < D; ++di) {
                const int d = d0 + di;
                if (d < lo || d >= hi) { dst[di] = INV16; continue; }
                int acc = 0, cnt = 0;
                for (int xx = x0; xx <= x1; ++xx) {
                    const size_t j = (size_t)y * w + xx;
                    if (d < dmin[j] || d >= dmax[j]) continue;
                    const uint16_t c = cost[j * D + di];
                    if (c == INV16) continue;
                    acc += c; ++cnt;
                }
                dst[di] = cnt > 0 ? (uint16_t)(acc / cnt) : INV16;
            }
        }
    }
}

/* vertical pass of CostAggregator::aggregate_hv (src/cost_aggregator.cpp:107-138) */
void orc_cross_v(const orc_cfg* cfg, int w, int h, const int16_t* dmin, const int16_t* dmax,
                 const int16_t* Ua, const int16_t* Da, const uint16_t* tmp, uint16_t* cost) {
    const int d0 = cfg->min_disparity, D = cfg->max_disparity - cfg->min_disparity;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const size_t i = (size_t)y * w + x;
            const int y0 = y - Ua[i], y1 = y + Da[i];
            const int lo = dmin[i], hi = dmax[i];
            uint16_t* dst = cost + i * D;
            for (int di = 0; di < D; ++di) {
                const int d = d0 + di;
                if (d < lo || d >= hi) { dst[di] = INV16; continue; }
                int acc = 0, cnt = 0;
                for (int yy = y0; yy <= y1; ++yy) {
                    const size_t j = (size_t)yy * w + x;
                    if (d < dmin[j] || d >= dmax[j]) continue;
                    const uint16_t c = tmp[j * D + di];
                    if (c == INV16) continue;
                    acc += c; ++cnt;
                }
                dst[di] = cnt > 0 ? (uint16_t)(acc / cnt) : INV16;
            }
        }
    }
}

/* CostAggregator::aggregate (src/cost_aggregator.cpp:141-149) */
void orc_cross_aggregate(const orc_cfg* cfg, const uint8_t* gray, int w, int h,
                         const int16_t* dmin, const int16_t* dmax, uint16_t* cost) {
    if (!cfg->agg_enable) return;
    const size_t n = (size_t)w * h;
    const int D = cfg->max_disparity - cfg->min_disparity;
    int16_t* La = malloc(n * sizeof(int16_t));
    int16_t* Ra = malloc(n * sizeof(int16_t));
    int16_t* Ua = malloc(n * sizeof(int16_t));
    int16_t* Da = malloc(n * sizeof(int16_t));
    uint16_t* tmp = malloc(n * (size_t)D * sizeof(uint16_t));
    orc_cross_arms(cfg, gray, w, h, La, Ra, Ua, Da);
    const int iters = imax(1, cfg->agg_iterations);
    for (int it = 0; it < iters; ++it) {
        orc_cross_h(cfg, w, h, dmin, dmax, La, Ra, cost, tmp);
        orc_cross_v(cfg, w, h, dmin, dmax, Ua, Da, tmp, cost);
    }
    free(La); free(Ra); free(Ua); free(Da); free(tmp);
}

/* ---------------------------------------------------------------- SGM ---- */

/* detail::adaptive_p2 (src/sgm_common.hpp:25-39) */
int orc_adaptive_p2(const orc_cfg* c, int dI) {
    dI = abs(dI);
    int P2 = c->P2_base;
    if (!c->adaptive_penalties) return imax(P2, c->P1 + 1);
    if (c->piecewise_p2) {
        if (dI >= c->grad_t2) P2 = c->P2_base / 4;
        else if (dI >= c->grad_t1) P2 = c->P2_base / 2;
        else P2 = c->P2_base;
    } else {
        P2 = (int)((float)c->P2_base / (1.f + c->p2_alpha * (float)dI) + 0.5f);
    }
    return imax(P2, c->P1 + 1);
}

/* process_pixel (src/sgm_optimizer.cpp:21-81) */
static void sgm_pixel(const orc_cfg* cfg, int w, int D, const uint8_t* gray,
                      const uint16_t* base, uint32_t* acc, int x, int y, int has_prev,
                      int px, int py, int* prev, int* cur) {
    const uint16_t* c = base + ((size_t)y * w + x) * D;
    uint32_t* a = acc + ((size_t)y * w + x) * D;
    if (!has_prev) {
        for (int d = 0; d < D; ++d) {
            if (c[d] == INV16) { cur[d] = PATH_INF; a[d] = INV32; }
            else { cur[d] = c[d]; if (a[d] != INV32) a[d] += (uint32_t)cur[d]; }
        }
        memcpy(prev, cur, sizeof(int) * D);
        return;
    }
    const int dI = (int)gray[(size_t)y * w + x] - (int)gray[(size_t)py * w + px];
    const int P2 = orc_adaptive_p2(cfg, dI);
    const int P1 = cfg->P1;
    int min_prev = PATH_INF;
    for (int d = 0; d < D; ++d) if (prev[d] < min_prev) min_prev = prev[d];
    for (int d = 0; d < D; ++d) {
        if (c[d] == INV16) { cur[d] = PATH_INF; a[d] = INV32; continue; }
        int best = PATH_INF;
        if (prev[d] < PATH_INF) best = imin(best, prev[d]);
        if (d > 0 && prev[d - 1] < PATH_INF) best = imin(best, prev[d - 1] + P1);
        if (d + 1 < D && prev[d + 1] < PATH_INF) best = imin(best, prev[d + 1] + P1);
        if (min_prev < PATH_INF) best = imin(best, min_prev + P2);
        cur[d] = best >= PATH_INF ? (int)c[d] : (int)c[d] + best - min_prev;
        if (a[d] != INV32) a[d] += (uint32_t)cur[d];
    }
    memcpy(prev, cur, sizeof(int) * D);
}

static inline int inb(int x, int y, int w, int h) { return x >= 0 && x < w && y >= 0 && y < h; }

/* SgmOptimizer::aggregate_path (src/sgm_optimizer.cpp:186-259). Scanlines of
 * one direction are independent, so they run in parallel here too; the
 * diagonal order (ray start = first in-bounds pixel walking back) is kept. */
void orc_sgm_path(const orc_cfg* cfg, const uint8_t* gray, int w, int h, int D,
                  const uint16_t* cost, uint32_t* acc, int dx, int dy) {
    if (dy == 0) {
#pragma omp parallel
        {
            int* prev = malloc(sizeof(int) * (size_t)D);
            int* cur = malloc(sizeof(int) * (size_t)D);
#pragma omp for schedule(dynamic, 4)
            for (int y = 0; y < h; ++y) {
                if (dx == 1)
                    for (int x = 0; x < w; ++x) sgm_pixel(cfg, w, D, gray, cost, acc, x, y, x > 0, x - 1, y, prev, cur);
                else
                    for (int x = w - 1; x >= 0; --x) sgm_pixel(cfg, w, D, gray, cost, acc, x, y, x + 1 < w, x + 1, y, prev, cur);
            }
            free(prev); free(cur);
        }
    } else if (dx == 0) {
#pragma omp parallel
        {
            int* prev = malloc(sizeof(int) * (size_t)D);
            int* cur = malloc(sizeof(int) * (size_t)D);
#pragma omp for schedule(dynamic, 4)
            for (int x = 0; x < w; ++x) {
                if (dy == 1)
                    for (int y = 0; y < h; ++y) sgm_pixel(cfg, w, D, gray, cost, acc, x, y, y > 0, x, y - 1, prev, cur);
                else
                    for (int y = h - 1; y >= 0; --y) sgm_pixel(cfg, w, D, gray, cost, acc, x, y, y + 1 < h, x, y + 1, prev, cur);
            }
            free(prev); free(cur);
        }
    } else {
        /* diagonal rays: every ray starts on the image border it enters from */
        const int nstart = w + h;
#pragma omp parallel
        {
            int* prev = malloc(sizeof(int) * (size_t)D);
            int* cur = malloc(sizeof(int) * (size_t)D);
#pragma omp for schedule(dynamic, 4)
            for (int s = 0; s < nstart; ++s) {
                /* enumerate start pixels: those whose predecessor (x-dx, y-dy) is out of bounds */
                int x, y;
                if (s < w) { x = s; y = (dy == 1) ? 0 : h - 1; }
                else {
                    y = s - w; x = (dx == 1) ? 0 : w - 1;
                    /* skip the corner pixel already enumerated on the horizontal border */
                    if (y == ((dy == 1) ? 0 : h - 1)) continue;
                }
                int hp = 0, px = x, py = y;
                while (inb(x, y, w, h)) {
                    sgm_pixel(cfg, w, D, gray, cost, acc, x, y, hp, px, py, prev, cur);
                    hp = 1; px = x; py = y; x += dx; y += dy;
                }
            }
            free(prev); free(cur);
        }
    }
}

/* SgmOptimizer::optimize (src/sgm_optimizer.cpp:261-274) */
void orc_sgm_optimize(const orc_cfg* cfg, const uint8_t* gray, int w, int h,
                      const uint16_t* cost, uint32_t* acc) {
    static const int dirs[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {1, -1}, {-1, 1}, {-1, -1}};
    const int D = cfg->max_disparity - cfg->min_disparity;
    memset(acc, 0, (size_t)w * h * D * sizeof(uint32_t));
    const int npath = cfg->paths == 8 ? 8 : 4;
    for (int p = 0; p < npath; ++p) orc_sgm_path(cfg, gray, w, h, D, cost, acc, dirs[p][0], dirs[p][1]);
}

/* ---------------------------------------------------------------- WTA ---- */

/* wta_impl (src/sgm_optimizer.cpp:436-555), Legacy subpixel, no telemetry.
 * Generic over the element type via the INV sentinel and a loader. */
#define WTA_BODY(T, INV)                                                                      \
    const int d0 = cfg->min_disparity, D = cfg->max_disparity - cfg->min_disparity;           \
    _Pragma("omp parallel for schedule(dynamic, 4)")                                          \
    for (int y = 0; y < h; ++y) {                                                              \
        for (int x = 0; x < w; ++x) {                                                          \
            const size_t p = (size_t)y * w + x;                                                \
            const T* s = vol + p * D;                                                          \
            const int lo = dmin[p], hi = dmax[p];                                              \
            float dv = -1.f, bv = INFINITY, sv = INFINITY;                                     \
            if (hi > lo) {                                                                     \
                int best_d = -1;                                                               \
                uint64_t bst = UINT64_MAX;                                                     \
                for (int d = lo; d < hi; ++d) {                                                \
                    const int di = d - d0;                                                     \
                    if (di < 0 || di >= D) continue;                                           \
                    const T c = s[di];                                                         \
                    if (c == INV) continue;                                                    \
                    if ((uint64_t)c < bst) { bst = c; best_d = d; }                            \
                }                                                                              \
                if (best_d >= 0) {                                                             \
                    uint64_t sec = UINT64_MAX;                                                 \
                    for (int d = lo; d < hi; ++d) {                                            \
                        if (abs(d - best_d) <= 1) continue;                                    \
                        const int di = d - d0;                                                 \
                        if (di < 0 || di >= D) continue;                                       \
                        const T c = s[di];                                                     \
                        if (c == INV) continue;                                                \
                        if ((uint64_t)c < sec) sec = c;                                        \
                    }                                                                          \
                    float dval = (float)best_d;                                                \
                    if (cfg->post_subpixel && best_d - 1 >= lo && best_d + 1 < hi) {           \
                        const int di = best_d - d0;                                            \
                        if (di > 0 && di + 1 < D) {                                            \
                            const T cm = s[di - 1], c0 = s[di], cp = s[di + 1];                \
                            if (cm != INV && cp != INV)                                        \
                                dval += orc_subpixel_legacy((float)cm, (float)c0, (float)cp);  \
                        }                                                                      \
                    }                                                                          \
                    dv = dval;                                                                 \
                    bv = (float)bst;                                                           \
                    sv = sec < UINT64_MAX ? (float)sec : INFINITY;                             \
                }                                                                              \
            }                                                                                  \
            disp[p] = dv;                                                                      \
            if (best) best[p] = bv;                                                            \
            if (second) second[p] = sv;                                                        \
        }                                                                                      \
    }

void orc_wta_u32(const orc_cfg* cfg, int w, int h, const uint32_t* vol,
                 const int16_t* dmin, const int16_t* dmax, float* disp, float* best, float* second) {
    WTA_BODY(uint32_t, INV32)
}

void orc_wta_u16(const orc_cfg* cfg, int w, int h, const uint16_t* vol,
                 const int16_t* dmin, const int16_t* dmax, float* disp, float* best, float* second) {
    WTA_BODY(uint16_t, INV16)
}

/* --------------------------------------------------------------- post ---- */

/* ConfidenceEstimator::estimate (src/confidence_estimator.cpp:7-63) */
void orc_confidence(const orc_cfg* cfg, int w, int h, const float* disp, const float* dr,
                    const uint8_t* gx, const uint8_t* gy, const float* best, const float* second,
                    float* confidence, uint8_t* reliable, uint8_t* reason) {
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const size_t p = (size_t)y * w + x;
            const float c1 = best[p], c2 = second[p];
            float uniq = 0.f;
            if (isfinite(c1) && isfinite(c2)) {
                uniq = (c2 - c1) / fmaxf_(c2, 1e-3f);
                uniq = clampf(uniq, 0.f, 1.f);
            }
            float lr = 1.f;
            if (dr) {
                const float d = disp[p];
                const int xr = (int)roundf((float)x - d);
                if (xr < 0 || xr >= w || d < 0.f) lr = 0.f;
                else {
                    const float drv = dr[(size_t)y * w + xr];
                    if (drv < 0.f || fabsf(d - drv) > (float)cfg->conf_lr_max_diff + 0.5f) lr = 0.f;
                }
            }
            const float tex = (float)(gx[p] + gy[p]);
            const float tex_term = clampf(tex / 16.f, 0.f, 1.f);
            float conf = clampf(uniq, 0.f, 4.f);
            conf = conf * (0.5f + 0.5f * tex_term);
            if (lr < 0.5f) conf *= 0.25f;
            confidence[p] = conf;
            const int lr_ok = lr >= 0.5f;
            const int rel = lr_ok && conf >= cfg->conf_reliable_threshold &&
                            uniq >= cfg->conf_uniqueness_ratio && tex >= cfg->conf_min_texture &&
                            disp[p] >= 0.f;
            reliable[p] = rel ? 1 : 0;
            reason[p] = R_VALID;
            if (!rel) reason[p] = lr_ok ? R_LOWCONF : R_MISMATCH;
        }
    }
}

/* PostProcessor::left_right_check (src/postprocess.cpp:10-40) */
void orc_lr_check(const orc_cfg* cfg, int w, int h, float* disp, const float* dr, uint8_t* reason) {
    if (!cfg->post_lr_check || !dr) return;
    const float tol = (float)cfg->post_lr_max_diff + 0.5f;
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const size_t p = (size_t)y * w + x;
            const float d = disp[p];
            if (d < 0.f) continue;
            const int xr = (int)roundf((float)x - d);
            if (xr < 0 || xr >= w) { disp[p] = -1.f; reason[p] = R_OCCLUSION; continue; }
            const float drv = dr[(size_t)y * w + xr];
            if (drv < 0.f || fabsf(d - drv) > tol) {
                disp[p] = -1.f;
                reason[p] = (drv > d + tol) ? R_OCCLUSION : R_MISMATCH;
            }
        }
    }
}

/* PostProcessor::fill_holes, Legacy route with no telemetry
 * (src/postprocess.cpp:42-168) */
void orc_fill_holes(const orc_cfg* cfg, int w, int h, float* disp, const uint8_t* reason,
                    const uint8_t* gray, const float* d_prior) {
    const int max_gap = cfg->post_max_fill_gap;
    float* out = malloc(sizeof(float) * (size_t)w * h);
    memcpy(out, disp, sizeof(float) * (size_t)w * h);
    for (int y = 0; y < h; ++y) {
        const float* row = disp + (size_t)y * w;
        for (int x = 0; x < w; ++x) {
            if (row[x] >= 0.f) continue;
            const int r = reason[(size_t)y * w + x];
            int xl = x - 1;
            while (xl >= 0 && xl >= x - max_gap && row[xl] < 0.f) --xl;
            int xr = x + 1;
            while (xr < w && xr <= x + max_gap && row[xr] < 0.f) ++xr;
            const float dl = (xl >= 0 && row[xl] >= 0.f) ? row[xl] : -1.f;
            const float drr = (xr < w && row[xr] >= 0.f) ? row[xr] : -1.f;
            float* o = out + (size_t)y * w + x;
            if (r == R_OCCLUSION) {
                if (dl >= 0.f && drr >= 0.f) *o = dl < drr ? dl : drr; /* std::min */
                else if (dl >= 0.f) *o = dl;
                else if (drr >= 0.f) *o = drr;
                continue;
            }
            const int c0 = gray[(size_t)y * w + x];
            float best_d = -1.f;
            int best_abs = 999;
            const int cx[4] = {xl, xr, x, x};
            const int cy[4] = {y, y, y - 1, y + 1};
            for (int k = 0; k < 4; ++k) {
                if (k == 2 && !(y > 0)) continue;
                if (k == 3 && !(y + 1 < h)) continue;
                const int xx = cx[k], yy = cy[k];
                if (xx < 0 || yy < 0 || xx >= w || yy >= h) continue;
                const float d = disp[(size_t)yy * w + xx];
                if (d < 0.f) continue;
                const int ad = abs((int)gray[(size_t)yy * w + xx] - c0);
                if (ad < best_abs) { best_abs = ad; best_d = d; }
            }
            if (best_d >= 0.f && best_abs <= 18) *o = best_d;
            else if (d_prior && d_prior[(size_t)y * w + x] >= 0.f) *o = d_prior[(size_t)y * w + x];
            else if (dl >= 0.f && drr >= 0.f) {
                const float t = (float)(x - xl) / (float)(xr - xl);
                *o = dl * (1.f - t) + drr * t;
            } else if (dl >= 0.f) *o = dl;
            else if (drr >= 0.f) *o = drr;
        }
    }
    memcpy(disp, out, sizeof(float) * (size_t)w * h);
    free(out);
}

static void sort_small(float* a, int n) {
    for (int i = 1; i < n; ++i) {
        const float v = a[i];
        int j = i - 1;
        while (j >= 0 && a[j] > v) { a[j + 1] = a[j]; --j; }
        a[j + 1] = v;
    }
}

/* PostProcessor::median (src/postprocess.cpp:255-309): nth_element at size/2
 * equals the sorted element at size/2. */
void orc_median(const orc_cfg* cfg, int w, int h, float* disp) {
    const int r = cfg->post_median_radius;
    if (r <= 0) return;
    float* out = malloc(sizeof(float) * (size_t)w * h);
    memcpy(out, disp, sizeof(float) * (size_t)w * h);
    const int maxwin = (2 * r + 1) * (2 * r + 1);
#pragma omp parallel
    {
        float* win = malloc(sizeof(float) * (size_t)maxwin);
#pragma omp for schedule(static)
        for (int y = 0; y < h; ++y) {
            for (int x = 0; x < w; ++x) {
                int n = 0;
                for (int oy = -r; oy <= r; ++oy)
                    for (int ox = -r; ox <= r; ++ox) {
                        const int xx = x + ox, yy = y + oy;
                        if (xx < 0 || yy < 0 || xx >= w || yy >= h) continue;
                        const float d = disp[(size_t)yy * w + xx];
                        if (d >= 0.f) win[n++] = d;
                    }
                if (n == 0) continue;
                sort_small(win, n);
                out[(size_t)y * w + x] = win[n / 2];
            }
        }
        free(win);
    }
    memcpy(disp, out, sizeof(float) * (size_t)w * h);
    free(out);
}

/* ----------------------------------------------------------- pipeline ---- */

uint64_t orc_fnv1a64(const void* data, size_t bytes) {
    const uint8_t* b = (const uint8_t*)data;
    uint64_t hsh = 14695981039346656037ULL;
    for (size_t i = 0; i < bytes; ++i) { hsh ^= b[i]; hsh *= 1099511628211ULL; }
    return hsh;
}

/* validate_config (src/pipeline.cpp:171-213, prior/refine/hierarchical off) */
static int validate_cfg(const orc_cfg* c) {
    if (c->max_disparity <= c->min_disparity) return 3;
    if (c->cost_max < 0 || c->cost_max > 65535) return 4;
    if (c->P1 < 0) return 4;
    if (c->P2_base <= c->P1) return 4;
    if (c->grad_t1 < 0 || c->grad_t2 < c->grad_t1) return 4;
    return 0;
}

/* StereoMatcher::compute_impl dense branch (src/pipeline.cpp:435-605) */
int orc_pipeline(const orc_cfg* cfg, const uint8_t* left, const uint8_t* right, int w, int h,
                 int stride, int ch, float* disparity, float* confidence, uint8_t* reason,
                 uint8_t* reliable, orc_intermediates* inter) {
    if (!left || !right || w <= 0 || h <= 0) return 1;
    if (stride < w * ch) return 2;
    const int vs = validate_cfg(cfg);
    if (vs) return vs;
    const size_t n = (size_t)w * h;
    const int D = cfg->max_disparity - cfg->min_disparity;
    const size_t nv = n * (size_t)D;
    uint8_t *gl = malloc(n), *gxl = malloc(n), *gyl = malloc(n);
    uint8_t *gr = malloc(n), *gxr = malloc(n), *gyr = malloc(n);
    uint32_t *c32l = NULL, *c32r = NULL;
    uint64_t *c64l = NULL, *c64r = NULL;
    int16_t *dmin = malloc(n * 2), *dmax = malloc(n * 2);
    uint16_t* cost = malloc(nv * 2);
    uint32_t* acc = malloc(nv * 4);
    float *dr = malloc(n * 4), *best = malloc(n * 4), *second = malloc(n * 4);
    orc_gray_grad(left, w, h, stride, ch, gl, gxl, gyl);
    orc_gray_grad(right, w, h, stride, ch, gr, gxr, gyr);
    if (cfg->census_type == 1) {
        c32l = malloc(n * 4); c32r = malloc(n * 4);
        orc_census_sym(gl, w, h, c32l); orc_census_sym(gr, w, h, c32r);
    } else {
        c64l = malloc(n * 8); c64r = malloc(n * 8);
        orc_census64(gl, w, h, c64l); orc_census64(gr, w, h, c64r);
    }
    float* d_prior = malloc(n * 4);
    orc_prior_estimate(cfg, w, h, gxl, gyl, c32l, c32r, c64l, c64r, dmin, dmax, d_prior);
    if (inter && inter->dmin) memcpy(inter->dmin, dmin, n * 2);
    if (inter && inter->dmax) memcpy(inter->dmax, dmax, n * 2);
    if (inter && inter->d_prior) memcpy(inter->d_prior, d_prior, n * 4);
    orc_cost_volume(cfg, w, h, gl, gxl, gyl, gr, gxr, gyr, c32l, c32r, c64l, c64r, dmin, dmax, cost);
    if (inter && inter->cost) memcpy(inter->cost, cost, nv * 2);
    orc_right_wta(cfg, w, h, cost, dr);
    orc_cross_aggregate(cfg, gl, w, h, dmin, dmax, cost);
    if (inter && inter->cost_agg) memcpy(inter->cost_agg, cost, nv * 2);
    orc_sgm_optimize(cfg, gl, w, h, cost, acc);
    if (inter && inter->sgm) memcpy(inter->sgm, acc, nv * 4);
    orc_wta_u32(cfg, w, h, acc, dmin, dmax, disparity, best, second);
    if (inter && inter->disp_wta) memcpy(inter->disp_wta, disparity, n * 4);
    if (inter && inter->disp_right) memcpy(inter->disp_right, dr, n * 4);
    if (inter && inter->best) memcpy(inter->best, best, n * 4);
    if (inter && inter->second) memcpy(inter->second, second, n * 4);
    orc_confidence(cfg, w, h, disparity, dr, gxl, gyl, best, second, confidence, reliable, reason);
    orc_refine(cfg, w, h, gl, gxl, gyl, gr, gxr, gyr, c32l, c32r, c64l, c64r, dmin, dmax, reliable, disparity);
    if (inter && inter->after_refine) memcpy(inter->after_refine, disparity, n * 4);
    orc_lr_check(cfg, w, h, disparity, dr, reason);
    orc_fill_holes(cfg, w, h, disparity, reason, gl, d_prior);
    orc_median(cfg, w, h, disparity);
    free(gl); free(gxl); free(gyl); free(gr); free(gxr); free(gyr);
    free(c32l); free(c32r); free(c64l); free(c64r);
    free(dmin); free(dmax); free(cost); free(acc); free(dr); free(best); free(second); free(d_prior);
    return 0;
}
