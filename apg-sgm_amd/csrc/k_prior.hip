// gfx950 kernels of the HighQuality prior (reference src/prior_estimator.cpp):
//   prior_cand_kernel : every prior sample (even x, y inside a 2-pixel border)
//                       runs the census-only WTA with best / second cost, the
//                       uniqueness test and the right-view check of
//                       extract_supports (:53-101); one thread per sample.
//   prior_range_kernel: bilinear interpolation of the filled cell grid to
//                       every pixel (interpolate_prior :312-374) and the
//                       per-pixel search range (apply_search_range :377-412),
//                       one thread per pixel.
//   prior_collect / select / stats / fill: the support selection under the
//                       budget, the per-cell statistics and the hole filling
//                       between them (extract_supports :35-166,
//                       interpolate_prior :181-310).
// Float expressions keep the reference's order; the library is built without
// contraction and with IEEE division.
#include <algorithm>

#include "device_common.hpp"
#include "hq_host.hpp"
#include "kernels.hpp"
#include "prior_sort.hpp"

namespace apgk {

template <bool SYM>
__global__ __launch_bounds__(256) void prior_cand_kernel(int w, int h, int d0, int d1, int tex_min, float uniq_min,
                                                         int lr_max, const uint8_t* __restrict__ gx,
                                                         const uint8_t* __restrict__ gy,
                                                         const uint32_t* __restrict__ c32l,
                                                         const uint32_t* __restrict__ c32r,
                                                         const uint64_t* __restrict__ c64l,
                                                         const uint64_t* __restrict__ c64r, int sw,
                                                         int16_t* __restrict__ cand_d, float* __restrict__ cand_u) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= sw) return;
    const int x = 2 + 2 * i, y = 2 + 2 * j;
    const size_t row = (size_t)y * w, q = row + x, s = (size_t)j * sw + i;
    auto pop = [&](size_t ql, size_t qr) {
        return SYM ? __popc(c32l[ql] ^ c32r[qr]) : __popcll(c64l[ql] ^ c64r[qr]);
    };
    int16_t outd = kNoCandidate;
    float outu = 0.f;
    if ((int)gx[q] + (int)gy[q] >= tex_min) {
        int bd = d0, b1 = 255, b2 = 255;
        const int dlo = max(d0, x - w + 1), dhi = min(d1, x + 1);   // x - d inside the image
        for (int d = dlo; d < dhi; ++d) {
            const int c = pop(q, row + x - d);
            if (c < b1) {
                b2 = b1;
                b1 = c;
                bd = d;
            } else if (c < b2) {
                b2 = c;
            }
        }
        if (b2 > 0) {
            const float uniq = (float)(b2 - b1) / (float)b2;
            const int xr = x - bd;
            if (!(uniq < uniq_min) && xr >= 1 && xr < w - 1) {
                int rb = 255, rd = 0;
                const int elo = max(d0, -xr), ehi = min(d1, w - xr);   // xr + d inside the image
                for (int d = elo; d < ehi; ++d) {
                    const int c = pop(row + xr + d, row + xr);
                    if (c < rb) {
                        rb = c;
                        rd = d;
                    }
                }
                if (abs(rd - bd) <= lr_max) {
                    outd = (int16_t)bd;
                    outu = uniq;
                }
            }
        }
    }
    cand_d[s] = outd;
    cand_u[s] = outu;
}

__device__ __forceinline__ float fmin_(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float fmax_(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ float fclamp_(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

// grid: 5 planes of gw*gh floats (disparity, confidence, spread, lowest, highest)
__global__ __launch_bounds__(256) void prior_range_kernel(int w, int h, int gw, int gh, int gmin, int gmax,
                                                          const float* __restrict__ grid,
                                                          const int* __restrict__ nsel, int min_sup,
                                                          int16_t* __restrict__ dmin, int16_t* __restrict__ dmax,
                                                          float* __restrict__ d_prior) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    // too few supports: no prior (d_prior stays -1), the dense range stays
    const int nsup = *nsel;
    if (nsup < min_sup || nsup == 0) return;
    const size_t ng = (size_t)gw * gh;
    const float *gd = grid, *gc = grid + ng, *gs = grid + 2 * ng, *glo = grid + 3 * ng, *ghi = grid + 4 * ng;
    const float fy = ((float)y + 0.5f) / (float)kPriorCell - 0.5f;
    const int y0 = clampi((int)floorf(fy), 0, gh - 1), y1 = min(y0 + 1, gh - 1);
    const float ty = fclamp_(fy - (float)y0, 0.f, 1.f);
    const float fx = ((float)x + 0.5f) / (float)kPriorCell - 0.5f;
    const int x0 = clampi((int)floorf(fx), 0, gw - 1), x1 = min(x0 + 1, gw - 1);
    const float tx = fclamp_(fx - (float)x0, 0.f, 1.f);
    const int k0 = y0 * gw + x0, k1 = y0 * gw + x1, k2 = y1 * gw + x0, k3 = y1 * gw + x1;
    const float v0 = gd[k0], v1 = gd[k1], v2 = gd[k2], v3 = gd[k3];
    const float w0 = (1.f - tx) * (1.f - ty) * (v0 < 0.f ? 0.f : 1.f);
    const float w1 = tx * (1.f - ty) * (v1 < 0.f ? 0.f : 1.f);
    const float w2 = (1.f - tx) * ty * (v2 < 0.f ? 0.f : 1.f);
    const float w3 = tx * ty * (v3 < 0.f ? 0.f : 1.f);
    const float ws = w0 + w1 + w2 + w3;
    const size_t q = (size_t)y * w + x;
    if (!(ws > 1e-6f)) {
        d_prior[q] = -1.f;   // no prior: the dense range stays
        return;
    }
    auto pos = [](float v) { return v < 0.f ? 0.f : v; };
    const float dp = (w0 * pos(v0) + w1 * pos(v1) + w2 * pos(v2) + w3 * pos(v3)) / ws;
    const float conf = (w0 * gc[k0] + w1 * gc[k1] + w2 * gc[k2] + w3 * gc[k3]) / ws;
    const float spread = (w0 * pos(gs[k0]) + w1 * pos(gs[k1]) + w2 * pos(gs[k2]) + w3 * pos(gs[k3])) / ws;
    float lo = 1e9f, hi = -1e9f;
    const int ks[4] = {k0, k1, k2, k3};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const float v = gd[ks[t]];
        if (v >= 0.f) {
            lo = fmin_(lo, glo[ks[t]] >= 0.f ? glo[ks[t]] : v);
            hi = fmax_(hi, ghi[ks[t]] >= 0.f ? ghi[ks[t]] : v);
        }
    }
    const float pmin = lo <= hi ? fmin_(lo, dp) : dp;
    const float pmax = lo <= hi ? fmax_(hi, dp) : dp;
    d_prior[q] = dp;
    // apply_search_range (its search_radius default is overwritten by every
    // branch); dp >= 0 here since the grid disparities are >= 0
    const int gD = gmax - gmin;
    int R;
    if (conf < 0.25f) R = gD;
    else if (conf > 0.85f && spread < 1.5f) R = 8;
    else if (conf > 0.60f && spread < 3.0f) R = (int)(12.f + 1.5f * spread + 0.5f);
    else if (conf > 0.40f && spread < 6.0f) R = (int)(18.f + 2.0f * spread + 0.5f);
    else R = max(24, (int)(16.f + 2.0f * spread + 0.5f));
    R = min(R, gD);
    const float plo = pmin >= 0.f ? pmin : dp, phi = pmax >= 0.f ? pmax : dp;
    const int blo = (int)floorf(plo) - 4, bhi = (int)ceilf(phi) + 4 + 1;
    int rlo = min((int)floorf(dp) - R, blo);
    int rhi = max((int)ceilf(dp) + R + 1, bhi);
    rhi = min(rhi, x + 1);
    // SearchRange::set_pixel (reference include/apg_sgm/search_range.hpp:29-35)
    rlo = max(rlo, gmin);
    rhi = min(rhi, gmax);
    if (rhi < rlo) rhi = rlo;
    dmin[q] = (int16_t)rlo;
    dmax[q] = (int16_t)rhi;
}

// ------------------------------------------------ support selection, grid ----
// PriorEstimator::extract_supports' per-cell candidates and stratified budget
// (reference src/prior_estimator.cpp:35-166) and interpolate_prior's grid
// statistics and hole filling (:181-310) on the device.  Every cell keeps its
// candidates in raster order, at most per_cell of them; a selection always
// takes a prefix of them, so a cell's supports are its first taken[] ones, in
// the order interpolate_prior bins them.

// One wavefront per cell: the cell's sample positions (at most 8 x 8 even
// pixels) in raster order, one per lane; candidates kept by ordered ballot.
__global__ __launch_bounds__(64) void prior_collect_kernel(int w, int h, int gw, int per_cell, int sw,
                                                           const int16_t* __restrict__ cand_d,
                                                           const float* __restrict__ cand_u, int* __restrict__ cnt,
                                                           float* __restrict__ key, float* __restrict__ wt,
                                                           int* __restrict__ sidx) {
    const int cell = blockIdx.x, lane = threadIdx.x;
    const int cx = cell % gw, cy = cell / gw;
    const int y0 = max(2, cy * kPriorCell), y1 = min(h - 2, (cy + 1) * kPriorCell);
    const int x0 = max(2, cx * kPriorCell), x1 = min(w - 2, (cx + 1) * kPriorCell);
    const int y = y0 + (y0 & 1) + 2 * (lane >> 3), x = x0 + (x0 & 1) + 2 * (lane & 7);
    const bool in = y < y1 && x < x1;
    const int si = in ? ((y - 2) / 2) * sw + (x - 2) / 2 : 0;
    const bool valid = in && cand_d[si] != kNoCandidate;
    const uint64_t m = __ballot(valid);
    const int pos = __popcll(m & ((1ull << lane) - 1ull));
    if (valid && pos < per_cell) {
        const size_t o = (size_t)cell * per_cell + pos;
        key[o] = (float)cand_d[si];
        wt[o] = cand_u[si];
        sidx[o] = si;
    }
    if (lane == 0) cnt[cell] = min(__popcll(m), per_cell);
}

constexpr int kSelThreads = 1024;

// exclusive prefix of the threads' v in thread order; total of all
__device__ int block_excl_scan(int v, int* s, int& total) {
    const int t = threadIdx.x;
    s[t] = v;
    __syncthreads();
    for (int off = 1; off < kSelThreads; off <<= 1) {
        const int a = t >= off ? s[t - off] : 0;
        __syncthreads();
        s[t] += a;
        __syncthreads();
    }
    total = s[kSelThreads - 1];
    const int incl = s[t];
    __syncthreads();
    return incl - v;
}

// The stratified budget (extract_supports :108-166) as closed forms over the
// cells in index order, thread t owning a contiguous run of cells:
//  case A (nonempty cells m <= budget): pass 1 gives each nonempty cell
//    t1 = min(quota, count) in order until the budget runs out, i.e.
//    clamp(budget - (sum of t1 before it), 0, t1); the round-robin passes then
//    add one per cell with candidates left, so after P full passes a cell
//    holds min(P, left) more; P* is the most passes that fit, and the
//    interrupted pass P* + 1 reaches the first R cells with more than P* left;
//  case B (m > budget): sample k takes the first candidate of nonempty cell
//    floor(k m / budget); cell j is hit iff k = ceil(j budget / m) < budget
//    maps back to j.
__global__ __launch_bounds__(kSelThreads) void prior_select_kernel(int ncell, int per_cell, int budget,
                                                                   const int* __restrict__ cnt,
                                                                   int* __restrict__ taken, int* __restrict__ nsel) {
    __shared__ int s[kSelThreads];
    const int t = threadIdx.x;
    const int run = (ncell + kSelThreads - 1) / kSelThreads;
    const int c0 = min(ncell, t * run), c1 = min(ncell, c0 + run);
    int ne = 0;
    for (int i = c0; i < c1; ++i) ne += cnt[i] > 0 ? 1 : 0;
    int m;
    const int ne_before = block_excl_scan(ne, s, m);
    if (m <= budget) {
        const int quota = max(1, budget / max(1, m));
        int t1 = 0;
        for (int i = c0; i < c1; ++i) t1 += cnt[i] > 0 ? min(quota, cnt[i]) : 0;
        int t1_all;
        int before = block_excl_scan(t1, s, t1_all);
        for (int i = c0; i < c1; ++i) {
            const int ti = cnt[i] > 0 ? min(quota, cnt[i]) : 0;
            taken[i] = min(max(budget - before, 0), ti);
            before += ti;
        }
        const int n1 = min(budget, t1_all);
        if (n1 < budget) {
            auto added = [&](int passes) {   // sum over cells of min(passes, left)
                int a = 0;
                for (int i = c0; i < c1; ++i) a += min(passes, cnt[i] - taken[i]);
                int all;
                (void)block_excl_scan(a, s, all);
                return all;
            };
            int ps = per_cell, R = 0;
            if (n1 + added(per_cell) > budget) {
                int lo = 0, hi = per_cell;   // n1 + added(lo) <= budget < n1 + added(hi)
                while (hi - lo > 1) {
                    const int mid = (lo + hi) / 2;
                    if (n1 + added(mid) <= budget) lo = mid;
                    else hi = mid;
                }
                ps = lo;
                R = budget - n1 - added(lo);
            }
            int e = 0;
            for (int i = c0; i < c1; ++i) e += cnt[i] - taken[i] > ps ? 1 : 0;
            int e_all;
            int rank = block_excl_scan(e, s, e_all);
            for (int i = c0; i < c1; ++i) {
                const int left = cnt[i] - taken[i];
                int add = min(ps, left);
                if (left > ps) {
                    if (rank < R) ++add;
                    ++rank;
                }
                taken[i] += add;
            }
        }
    } else {
        int j = ne_before;
        for (int i = c0; i < c1; ++i) {
            int sel = 0;
            if (cnt[i] > 0) {
                const long long k0 = ((long long)j * budget + m - 1) / m;
                sel = (k0 < budget && (k0 * m) / budget == j) ? 1 : 0;
                ++j;
            }
            taken[i] = sel;
        }
    }
    int nl = 0;
    for (int i = c0; i < c1; ++i) nl += taken[i];
    int n;
    (void)block_excl_scan(nl, s, n);
    if (t == 0) *nsel = n;
}

__device__ __forceinline__ float fclamp_s(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

// first element at which the running weight reaches half (the front's key if none)
__device__ float weighted_median(const KeyW* a, int n, float half) {
    float run = 0.f;
    for (int i = 0; i < n; ++i) {
        run += a[i].w;
        if (run >= half) return a[i].k;
    }
    return a[0].k;
}

// interpolate_prior's per-cell statistics (:196-250), one thread per cell,
// the cell's supports sorted in place in a scratch copy; grid planes
// (disparity, confidence, spread, lowest, highest) of nc floats each
__global__ __launch_bounds__(256) void prior_stats_kernel(int ncell, int per_cell, const int* __restrict__ taken,
                                                          const float* __restrict__ key,
                                                          const float* __restrict__ wt, KeyW* __restrict__ tmp,
                                                          float* __restrict__ grid) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncell) return;
    const size_t nc = (size_t)ncell;
    float d = -1.f, c = 0.f, sp = -1.f, lo = -1.f, hi = -1.f;
    const int n = taken[i];
    if (n > 0) {
        KeyW* a = tmp + (size_t)i * per_cell;
        for (int k = 0; k < n; ++k) a[k] = KeyW{key[(size_t)i * per_cell + k], wt[(size_t)i * per_cell + k]};
        sort_like_libstdcxx(a, n);
        lo = a[0].k;
        hi = a[n - 1].k;
        float tot = 0.f;
        for (int k = 0; k < n; ++k) tot += a[k].w;
        if (!(tot <= 1e-6f)) {
            const float half = tot * 0.5f;
            const float med = weighted_median(a, n, half);
            d = med;
            for (int k = 0; k < n; ++k) a[k].k = fabsf(a[k].k - med);   // deviations, in sorted order
            sort_like_libstdcxx(a, n);
            const float mad = weighted_median(a, n, half);
            const float s1 = mad * 1.4826f, s2 = (hi - lo) * 0.5f;
            sp = s1 < s2 ? s2 : s1;
            const float mean = tot / (float)n;
            c = fclamp_s(mean * fclamp_s((float)n / 3.f, 0.3f, 1.f), 0.f, 1.f);
        }
    }
    grid[i] = d;
    grid[nc + i] = c;
    grid[2 * nc + i] = sp;
    grid[3 * nc + i] = lo;
    grid[4 * nc + i] = hi;
}

// interpolate_prior's hole filling (:252-298): each sweep reads the previous
// grid and fills every empty cell with a filled 3 x 3 neighbour, until a
// sweep fills none or max(gw, gh) sweeps; one workgroup, grids g (result)
// and g2 (the other buffer)
__global__ __launch_bounds__(kSelThreads) void prior_fill_kernel(int gw, int gh, int min_sup,
                                                                 const int* __restrict__ nsel, float* g, float* g2) {
    const int n = *nsel;
    if (n < min_sup || n == 0) return;
    const int ncell = gw * gh;
    const size_t nc = (size_t)ncell;
    float* cur = g;
    float* nxt = g2;
    const int sweeps = max(gw, gh);
    for (int it = 0; it < sweeps; ++it) {
        int changed = 0;
        for (int i = threadIdx.x; i < ncell; i += kSelThreads) {
            float v[5];
            for (int p = 0; p < 5; ++p) v[p] = cur[p * nc + i];
            if (!(v[0] >= 0.f)) {
                const int cx = i % gw, cy = i / gw;
                float sd = 0.f, sc = 0.f, ss = 0.f, k = 0.f, mn = 1e9f, mx = -1e9f;
                for (int oy = -1; oy <= 1; ++oy)
                    for (int ox = -1; ox <= 1; ++ox) {
                        const int nx = cx + ox, ny = cy + oy;
                        if (nx < 0 || ny < 0 || nx >= gw || ny >= gh) continue;
                        const int j = ny * gw + nx;
                        const float dv = cur[j];
                        if (dv < 0.f) continue;
                        sd += dv;
                        sc += cur[nc + j];
                        const float sj = cur[2 * nc + j];
                        ss += sj >= 0.f ? sj : 8.f;
                        const float lj = cur[3 * nc + j], hj = cur[4 * nc + j];
                        const float l = lj >= 0.f ? lj : dv, u = hj >= 0.f ? hj : dv;
                        mn = l < mn ? l : mn;   // std::min(mn, l)
                        mx = mx < u ? u : mx;   // std::max(mx, u)
                        k += 1.f;
                    }
                if (k > 0.f) {
                    v[0] = sd / k;
                    v[1] = (sc / k) * 0.85f;
                    v[2] = ss / k;
                    v[3] = mn;
                    v[4] = mx;
                    changed = 1;
                }
            }
            for (int p = 0; p < 5; ++p) nxt[p * nc + i] = v[p];
        }
        __syncthreads();   // this sweep's grid is written before any thread reads it
        const int any = __syncthreads_or(changed);
        float* t = cur;
        cur = nxt;
        nxt = t;
        if (!any) break;
    }
    if (cur != g) {
        for (int i = threadIdx.x; i < 5 * ncell; i += kSelThreads) g[i] = cur[i];
    }
}

hipError_t launch_prior_candidates(hipStream_t s, const apg_hip_params& p, int w, int h, const uint8_t* gx,
                                   const uint8_t* gy, const uint32_t* c32l, const uint32_t* c32r,
                                   const uint64_t* c64l, const uint64_t* c64r, int16_t* cand_d, float* cand_u) {
    const int sw = prior_samples_x(w), sh = prior_samples_y(h);
    if (sw <= 0 || sh <= 0) return hipSuccess;
    const dim3 grid((sw + 255) / 256, sh);
    if (p.census_type == 1)
        hipLaunchKernelGGL(prior_cand_kernel<true>, grid, dim3(256), 0, s, w, h, p.min_disparity, p.max_disparity,
                           p.prior_texture_threshold, p.prior_uniqueness_ratio, p.prior_lr_max_diff, gx, gy, c32l,
                           c32r, c64l, c64r, sw, cand_d, cand_u);
    else
        hipLaunchKernelGGL(prior_cand_kernel<false>, grid, dim3(256), 0, s, w, h, p.min_disparity, p.max_disparity,
                           p.prior_texture_threshold, p.prior_uniqueness_ratio, p.prior_lr_max_diff, gx, gy, c32l,
                           c32r, c64l, c64r, sw, cand_d, cand_u);
    return hipGetLastError();
}

int prior_per_cell(int ncell, int budget) {
    return std::max(8, (budget * 2 + ncell - 1) / std::max(1, ncell));
}

PriorLayout prior_layout(int w, int h, int budget, void* work) {
    PriorLayout L{};
    L.gw = (w + kPriorCell - 1) / kPriorCell;
    L.gh = (h + kPriorCell - 1) / kPriorCell;
    L.ncell = L.gw * L.gh;
    L.per_cell = prior_per_cell(L.ncell, budget);
    const size_t nc = (size_t)L.ncell, ns = nc * (size_t)L.per_cell;
    auto r256 = [](size_t v) { return (v + 255) & ~(size_t)255; };
    uint8_t* p = (uint8_t*)work;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        uint8_t* q = p ? p + o : nullptr;
        o += r256(bytes);
        return q;
    };
    L.grid = (float*)take(5 * nc * 4);
    L.grid2 = (float*)take(5 * nc * 4);
    L.cnt = (int*)take(nc * 4);
    L.taken = (int*)take(nc * 4);
    L.key = (float*)take(ns * 4);
    L.wt = (float*)take(ns * 4);
    L.sidx = (int*)take(ns * 4);
    L.tmp = (KeyW*)take(ns * sizeof(KeyW));
    L.nsel = (int*)take(4);
    L.bytes = o;
    return L;
}

size_t prior_work_bytes(int w, int h, int budget) { return prior_layout(w, h, budget, nullptr).bytes; }

hipError_t launch_prior_grid(hipStream_t s, const apg_hip_params& p, int w, int h, const int16_t* cand_d,
                             const float* cand_u, const PriorLayout& L) {
    const int sw = prior_samples_x(w);
    hipLaunchKernelGGL(prior_collect_kernel, dim3(L.ncell), dim3(64), 0, s, w, h, L.gw, L.per_cell, sw, cand_d, cand_u,
                       L.cnt, L.key, L.wt, L.sidx);
    hipLaunchKernelGGL(prior_select_kernel, dim3(1), dim3(kSelThreads), 0, s, L.ncell, L.per_cell,
                       p.prior_max_supports, L.cnt, L.taken, L.nsel);
    hipLaunchKernelGGL(prior_stats_kernel, dim3((L.ncell + 255) / 256), dim3(256), 0, s, L.ncell, L.per_cell, L.taken,
                       L.key, L.wt, L.tmp, L.grid);
    hipLaunchKernelGGL(prior_fill_kernel, dim3(1), dim3(kSelThreads), 0, s, L.gw, L.gh, p.prior_min_supports, L.nsel,
                       L.grid, L.grid2);
    return hipGetLastError();
}

hipError_t launch_prior_ranges(hipStream_t s, const apg_hip_params& p, int w, int h, const PriorLayout& L,
                               int16_t* dmin, int16_t* dmax, float* d_prior) {
    hipLaunchKernelGGL(prior_range_kernel, dim3((w + 255) / 256, h), dim3(256), 0, s, w, h, L.gw, L.gh,
                       p.min_disparity, p.max_disparity, L.grid, L.nsel, p.prior_min_supports, dmin, dmax, d_prior);
    return hipGetLastError();
}

}  // namespace apgk
