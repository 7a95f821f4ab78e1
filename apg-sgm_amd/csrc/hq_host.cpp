// Host-side parts of the HighQuality stages (see hq_host.hpp).
#include "hq_host.hpp"

#include <math.h>
#include <stdlib.h>

#include <algorithm>
#include <utility>

namespace apgk {

namespace {

float fclamp(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

using Pt = std::pair<float, float>;   // (disparity or deviation, weight)

bool first_less(const Pt& a, const Pt& b) { return a.first < b.first; }

// Weighted median of pts sorted by first: the first element at which the
// running weight reaches half of the total.
float weighted_median(const std::vector<Pt>& pts, float half) {
    float run = 0.f;
    for (const Pt& q : pts) {
        run += q.second;
        if (run >= half) return q.first;
    }
    return pts.front().first;
}

}  // namespace

int prior_build_grid(const apg_hip_params& p, int w, int h, const int16_t* cand_d, const float* cand_u,
                     PriorGrid& g, std::vector<apg_hip_support>* supports) {
    const int sw = prior_samples_x(w);
    g.gw = (w + kPriorCell - 1) / kPriorCell;
    g.gh = (h + kPriorCell - 1) / kPriorCell;
    const int ncell = g.gw * g.gh;
    g.d.assign(ncell, -1.f);
    g.c.assign(ncell, 0.f);
    g.s.assign(ncell, -1.f);
    g.lo.assign(ncell, -1.f);
    g.hi.assign(ncell, -1.f);
    const int budget = p.prior_max_supports;
    const int per_cell = std::max(8, (budget * 2 + ncell - 1) / std::max(1, ncell));

    // candidates of each cell, raster order, at most per_cell (extract_supports :35-106)
    std::vector<std::vector<Pt>> cand(ncell);
    std::vector<std::vector<std::pair<int, int>>> cxy(supports ? ncell : 0);   // their pixels
    for (int cy = 0; cy < g.gh; ++cy)
        for (int cx = 0; cx < g.gw; ++cx) {
            std::vector<Pt>& dst = cand[cy * g.gw + cx];
            const int y0 = std::max(2, cy * kPriorCell), y1 = std::min(h - 2, (cy + 1) * kPriorCell);
            const int x0 = std::max(2, cx * kPriorCell), x1 = std::min(w - 2, (cx + 1) * kPriorCell);
            for (int y = y0 + (y0 & 1); y < y1 && (int)dst.size() < per_cell; y += 2)
                for (int x = x0 + (x0 & 1); x < x1 && (int)dst.size() < per_cell; x += 2) {
                    const size_t si = (size_t)((y - 2) / 2) * sw + (x - 2) / 2;
                    if (cand_d[si] != kNoCandidate) {
                        dst.push_back(Pt((float)cand_d[si], cand_u[si]));
                        if (supports) cxy[cy * g.gw + cx].emplace_back(x, y);
                    }
                }
        }

    // stratified selection under the budget (:108-166): how many of each
    // cell's candidates become supports (always a prefix)
    std::vector<int> cells;
    for (int i = 0; i < ncell; ++i)
        if (!cand[i].empty()) cells.push_back(i);
    std::vector<int> taken(ncell, 0);
    int n = 0;
    if ((int)cells.size() <= budget) {
        const int quota = std::max(1, budget / std::max(1, (int)cells.size()));
        for (int ci : cells) {
            const int t = std::min(quota, (int)cand[ci].size());
            for (int k = 0; k < t && n < budget; ++k) ++taken[ci], ++n;
        }
        for (bool more = true; more && n < budget;) {
            more = false;
            for (int ci : cells) {
                if (n >= budget) break;
                if (taken[ci] < (int)cand[ci].size()) ++taken[ci], ++n, more = true;
            }
        }
    } else {
        const uint64_t m = cells.size();
        for (int k = 0; k < budget; ++k) {
            ++taken[cells[(size_t)((uint64_t)k * m / (uint64_t)budget)]];
            ++n;
        }
    }
    if (supports) {   // the selected set: each cell's first taken[] candidates
        supports->clear();
        for (int i = 0; i < ncell; ++i)
            for (int k = 0; k < taken[i]; ++k)
                supports->push_back(apg_hip_support{cxy[i][k].first, cxy[i][k].second, cand[i][k].first,
                                                    cand[i][k].second});
    }
    if (n < p.prior_min_supports || n == 0) return n;

    // per-cell statistics (interpolate_prior :196-250)
    for (int i = 0; i < ncell; ++i) {
        if (!taken[i]) continue;
        std::vector<Pt> pts(cand[i].begin(), cand[i].begin() + taken[i]);
        std::sort(pts.begin(), pts.end(), first_less);
        g.lo[i] = pts.front().first;
        g.hi[i] = pts.back().first;
        float tot = 0.f;
        for (const Pt& q : pts) tot += q.second;
        if (tot <= 1e-6f) continue;
        const float half = tot * 0.5f;
        const float med = weighted_median(pts, half);
        g.d[i] = med;
        std::vector<Pt> dev;
        dev.reserve(pts.size());
        for (const Pt& q : pts) dev.push_back(Pt(fabsf(q.first - med), q.second));
        std::sort(dev.begin(), dev.end(), first_less);
        const float mad = weighted_median(dev, half);
        g.s[i] = std::max(mad * 1.4826f, (g.hi[i] - g.lo[i]) * 0.5f);
        const float mean = tot / (float)pts.size();
        g.c[i] = fclamp(mean * fclamp((float)pts.size() / 3.f, 0.3f, 1.f), 0.f, 1.f);
    }

    // empty cells take the mean of their filled 3x3 neighbours, one ring per
    // sweep, every sweep reading the previous one (:252-298)
    const int sweeps = std::max(g.gw, g.gh);
    for (int it = 0; it < sweeps; ++it) {
        std::vector<float> d = g.d, c = g.c, s = g.s, lo = g.lo, hi = g.hi;
        bool changed = false;
        for (int cy = 0; cy < g.gh; ++cy)
            for (int cx = 0; cx < g.gw; ++cx) {
                const int i = cy * g.gw + cx;
                if (g.d[i] >= 0.f) continue;
                float sd = 0.f, sc = 0.f, ss = 0.f, k = 0.f, mn = 1e9f, mx = -1e9f;
                for (int oy = -1; oy <= 1; ++oy)
                    for (int ox = -1; ox <= 1; ++ox) {
                        const int nx = cx + ox, ny = cy + oy;
                        if (nx < 0 || ny < 0 || nx >= g.gw || ny >= g.gh) continue;
                        const int j = ny * g.gw + nx;
                        const float v = g.d[j];
                        if (v < 0.f) continue;
                        sd += v;
                        sc += g.c[j];
                        ss += g.s[j] >= 0.f ? g.s[j] : 8.f;
                        mn = std::min(mn, g.lo[j] >= 0.f ? g.lo[j] : v);
                        mx = std::max(mx, g.hi[j] >= 0.f ? g.hi[j] : v);
                        k += 1.f;
                    }
                if (k > 0.f) {
                    d[i] = sd / k;
                    c[i] = (sc / k) * 0.85f;
                    s[i] = ss / k;
                    lo[i] = mn;
                    hi[i] = mx;
                    changed = true;
                }
            }
        g.d.swap(d);
        g.c.swap(c);
        g.s.swap(s);
        g.lo.swap(lo);
        g.hi.swap(hi);
        if (!changed) break;
    }
    return n;
}

}  // namespace apgk
