// gfx950 kernels: winner-take-all with second best and legacy subpixel,
// confidence, left-right check, legacy hole fill and valid-only median.
#include <math.h>

#include "device_common.hpp"
#include "kernels.hpp"

namespace apgk {

constexpr int WTA_PB = 64;   // pixels per block (one per lane)
constexpr int WTA_DC = 64;   // disparities staged per chunk

// wta_impl (reference src/sgm_optimizer.cpp:436-555) as a single streaming
// pass per pixel.  Scanning d upward over [lo, hi):
//   best: earliest strict minimum (the reference's `c < best`);
//   second: min over valid d with |d - best_d| > 1.  When a new best appears
//   at d, every earlier index except d-1 is eligible, i.e. the prefix minimum
//   over indices <= d-2 (kept one step delayed); later indices d' >= best+2
//   feed a right-hand minimum.  second = min(left part, right part).
// Invalid entries (the volume's sentinel) and d outside [lo, hi) are skipped.
// The 64 pixels of a block are consecutive in memory, so each d-chunk is
// staged into LDS with row-contiguous loads and read back one pixel per lane.
template <typename T>
__global__ __launch_bounds__(64) void wta_kernel(int w, int h, int d0, int D, int subpixel, const T* __restrict__ vol,
                                                 const int16_t* __restrict__ dmin, const int16_t* __restrict__ dmax,
                                                 float* __restrict__ disp, float* __restrict__ best_out,
                                                 float* __restrict__ second_out) {
    __shared__ uint32_t st[WTA_PB][WTA_DC + 1];
    const T INV = (T)~(T)0;
    const int lane = threadIdx.x;
    const size_t npx = (size_t)w * h;
    const size_t p_base = (size_t)blockIdx.x * WTA_PB;
    const size_t p = p_base + lane;
    const bool act = p < npx;
    const int lo = act ? dmin[p] : 0, hi = act ? dmax[p] : 0;
    uint64_t bv = ~0ull;          // best value
    int bd = -1;                  // best d
    uint64_t pm = ~0ull;          // prefix min over scanned indices < d
    uint64_t pm_lag = ~0ull;      // prefix min over scanned indices < d-1
    uint64_t secL = ~0ull, secR = ~0ull;
    uint64_t vprev = ~0ull;       // value at d-1 (INF when skipped)
    uint64_t v_bm1 = ~0ull, v_bp1 = ~0ull;   // raw values at best-1 / best+1
    uint64_t raw_prev = ~0ull;
    for (int dc = 0; dc < D; dc += WTA_DC) {
        const int nd = min(WTA_DC, D - dc);
        // pixel rows in batches of 8: every load is issued (clamped inside the
        // volume), so a batch waits once
        const int lc = min(lane, nd - 1);
        for (int i0 = 0; i0 < WTA_PB; i0 += 8) {
            uint32_t v[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const size_t q = min(p_base + (size_t)(i0 + i), npx - 1);
                v[i] = (uint32_t)vol[q * D + dc + lc];
            }
#pragma unroll
            for (int i = 0; i < 8; ++i)
                if (p_base + (size_t)(i0 + i) < npx && lane < nd) st[i0 + i][lane] = v[i];
        }
        __syncthreads();
        if (act) {
            for (int j = 0; j < nd; ++j) {
                const int di = dc + j;
                const int d = d0 + di;
                const T raw = (T)st[lane][j];
                // raw neighbour for the subpixel step is taken regardless of range
                if (bd >= 0 && d == bd + 1) v_bp1 = (uint64_t)raw;
                const bool valid = d >= lo && d < hi && raw != INV;
                const uint64_t c = valid ? (uint64_t)raw : ~0ull;
                const uint64_t pm_new = pm < c ? pm : c;
                if (valid && c < bv) {
                    secL = pm_lag;        // min over indices <= d-2
                    secR = ~0ull;
                    bv = c; bd = d;
                    v_bm1 = raw_prev;
                    v_bp1 = ~0ull;
                } else if (valid && bd >= 0 && d >= bd + 2) {
                    secR = secR < c ? secR : c;
                }
                pm_lag = pm;
                pm = pm_new;
                vprev = c;
                raw_prev = (uint64_t)raw;
            }
        }
        __syncthreads();
    }
    (void)vprev;
    if (!act) return;
    float dv = -1.f, b = INFINITY, sc = INFINITY;
    if (hi > lo && bd >= 0) {
        const uint64_t sec = secL < secR ? secL : secR;
        float dval = (float)bd;
        if (subpixel && bd - 1 >= lo && bd + 1 < hi) {
            const int di = bd - d0;
            if (di > 0 && di + 1 < D) {
                const uint64_t INV64 = (uint64_t)INV;
                if (v_bm1 != INV64 && v_bp1 != INV64) dval += legacy_delta((float)v_bm1, (float)bv, (float)v_bp1);
            }
        }
        dv = dval;
        b = (float)bv;
        sc = sec != ~0ull ? (float)sec : INFINITY;
    }
    disp[p] = dv;
    if (best_out) best_out[p] = b;
    if (second_out) second_out[p] = sc;
}

hipError_t launch_wta(hipStream_t s, const apg_hip_params& p, int w, int h, const void* vol, int bits,
                      const int16_t* dmin, const int16_t* dmax, float* disp, float* best, float* second) {
    const size_t npx = (size_t)w * h;
    dim3 grid((unsigned)((npx + WTA_PB - 1) / WTA_PB));
    const int D = p.max_disparity - p.min_disparity;
    if (bits == 16)
        hipLaunchKernelGGL(wta_kernel<uint16_t>, grid, dim3(64), 0, s, w, h, p.min_disparity, D, p.post_subpixel,
                           (const uint16_t*)vol, dmin, dmax, disp, best, second);
    else
        hipLaunchKernelGGL(wta_kernel<uint32_t>, grid, dim3(64), 0, s, w, h, p.min_disparity, D, p.post_subpixel,
                           (const uint32_t*)vol, dmin, dmax, disp, best, second);
    return hipGetLastError();
}

// ------------------------------------------------------------ confidence ----
enum { R_VALID = 0, R_OCCLUSION = 1, R_MISMATCH = 2, R_LOWCONF = 3 };

__device__ __forceinline__ float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ConfidenceEstimator::estimate (reference src/confidence_estimator.cpp:7-63)
__global__ void confidence_kernel(apg_hip_params p, int w, int h, const float* __restrict__ disp,
                                  const float* __restrict__ dr, const uint8_t* __restrict__ gx,
                                  const uint8_t* __restrict__ gy, const float* __restrict__ best,
                                  const float* __restrict__ second, float* __restrict__ conf_out,
                                  uint8_t* __restrict__ reliable, uint8_t* __restrict__ reason) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t q = (size_t)y * w + x;
    const float c1 = best[q], c2 = second[q];
    float uniq = 0.f;
    if (isfinite(c1) && isfinite(c2)) {
        const float den = c2 < 1e-3f ? 1e-3f : c2;
        uniq = clampf((c2 - c1) / den, 0.f, 1.f);
    }
    float lr = 1.f;
    const float d = disp[q];
    if (dr) {
        const int xr = (int)roundf((float)x - d);
        if (xr < 0 || xr >= w || d < 0.f) lr = 0.f;
        else {
            const float drv = dr[(size_t)y * w + xr];
            if (drv < 0.f || fabsf(d - drv) > (float)p.conf_lr_max_diff + 0.5f) lr = 0.f;
        }
    }
    const float tex = (float)((int)gx[q] + (int)gy[q]);
    const float tex_term = clampf(tex / 16.f, 0.f, 1.f);
    float conf = clampf(uniq, 0.f, 4.f);
    conf = conf * (0.5f + 0.5f * tex_term);
    if (lr < 0.5f) conf *= 0.25f;
    conf_out[q] = conf;
    const bool lr_ok = lr >= 0.5f;
    const bool rel = lr_ok && conf >= p.conf_reliable_threshold && uniq >= p.conf_uniqueness_ratio &&
                     tex >= p.conf_min_texture && d >= 0.f;
    reliable[q] = rel ? 1 : 0;
    reason[q] = rel ? R_VALID : (lr_ok ? R_LOWCONF : R_MISMATCH);
}

hipError_t launch_confidence(hipStream_t s, const apg_hip_params& p, int w, int h, const float* disp,
                             const float* dr, const uint8_t* gx, const uint8_t* gy, const float* best,
                             const float* second, float* conf, uint8_t* reliable, uint8_t* reason) {
    hipLaunchKernelGGL(confidence_kernel, dim3((w + 127) / 128, h), dim3(128), 0, s, p, w, h, disp, dr, gx, gy, best,
                       second, conf, reliable, reason);
    return hipGetLastError();
}

// PostProcessor::left_right_check (reference src/postprocess.cpp:10-40)
__global__ void lr_kernel(int w, int h, int lr_max_diff, float* __restrict__ disp, const float* __restrict__ dr,
                          uint8_t* __restrict__ reason) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t q = (size_t)y * w + x;
    const float tol = (float)lr_max_diff + 0.5f;
    const float d = disp[q];
    if (d < 0.f) return;
    const int xr = (int)roundf((float)x - d);
    if (xr < 0 || xr >= w) { disp[q] = -1.f; reason[q] = R_OCCLUSION; return; }
    const float drv = dr[(size_t)y * w + xr];
    if (drv < 0.f || fabsf(d - drv) > tol) {
        disp[q] = -1.f;
        reason[q] = (drv > d + tol) ? R_OCCLUSION : R_MISMATCH;
    }
}

hipError_t launch_lr_check(hipStream_t s, const apg_hip_params& p, int w, int h, float* disp, const float* dr,
                           uint8_t* reason) {
    if (!p.post_lr_check || !dr) return hipSuccess;
    hipLaunchKernelGGL(lr_kernel, dim3((w + 127) / 128, h), dim3(128), 0, s, w, h, p.post_lr_max_diff, disp, dr, reason);
    return hipGetLastError();
}

// A hole pixel x of row y once the gap scans have stopped at xl (left) and
// xr (right), as the reference's loops leave them (reference
// src/postprocess.cpp:60-130).  in_at / gray_at read the input disparity and
// the gray image at an in-image pixel; cur is the pixel's own (invalid) value.
template <typename InAt, typename GrayAt>
__device__ __forceinline__ float fill_hole_at(int w, int h, int x, int y, int xl, int xr, float cur, bool occl, int c0,
                                              float prior_v, InAt in_at, GrayAt gray_at) {
    float o = cur;
    const float dl = (xl >= 0 && in_at(xl, y) >= 0.f) ? in_at(xl, y) : -1.f;
    const float drr = (xr < w && in_at(xr, y) >= 0.f) ? in_at(xr, y) : -1.f;
    if (occl) {
        if (dl >= 0.f && drr >= 0.f) o = drr < dl ? drr : dl;
        else if (dl >= 0.f) o = dl;
        else if (drr >= 0.f) o = drr;
    } else {
        float best_d = -1.f;
        int best_abs = 999;
        const int cx[4] = {xl, xr, x, x};
        const int cy[4] = {y, y, y - 1, y + 1};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int xx = cx[k], yy = cy[k];
            if (xx < 0 || yy < 0 || xx >= w || yy >= h) continue;
            const float dd = in_at(xx, yy);
            if (dd < 0.f) continue;
            const int ad = abs(gray_at(xx, yy) - c0);
            if (ad < best_abs) { best_abs = ad; best_d = dd; }
        }
        if (best_d >= 0.f && best_abs <= 18) o = best_d;
        else if (prior_v >= 0.f) o = prior_v;
        else if (dl >= 0.f && drr >= 0.f) {
            const float t = (float)(x - xl) / (float)(xr - xl);
            o = dl * (1.f - t) + drr * t;
        } else if (dl >= 0.f) o = dl;
        else if (drr >= 0.f) o = drr;
    }
    return o;
}

__device__ __forceinline__ float fill_hole(int w, int h, int x, int y, int xl, int xr, const float* row,
                                           const float* __restrict__ in, const uint8_t* __restrict__ reason,
                                           const uint8_t* __restrict__ gray, const float* __restrict__ prior) {
    const size_t q = (size_t)y * w + x;
    return fill_hole_at(
        w, h, x, y, xl, xr, row[x], reason[q] == R_OCCLUSION, (int)gray[q], prior ? prior[q] : -1.f,
        [&](int xx, int yy) { return yy == y ? row[xx] : in[(size_t)yy * w + xx]; },
        [&](int xx, int yy) { return (int)gray[(size_t)yy * w + xx]; });
}

__device__ __forceinline__ void fill_pixel(int w, int h, int max_gap, int x, int y, const float* row,
                                           const float* __restrict__ in, const uint8_t* __restrict__ reason,
                                           const uint8_t* __restrict__ gray, const float* __restrict__ prior,
                                           float* __restrict__ out) {
    const size_t q = (size_t)y * w + x;
    float o = row[x];
    if (o < 0.f) {
        int xl = x - 1;
        while (xl >= 0 && xl >= x - max_gap && row[xl] < 0.f) --xl;
        int xr = x + 1;
        while (xr < w && xr <= x + max_gap && row[xr] < 0.f) ++xr;
        o = fill_hole(w, h, x, y, xl, xr, row, in, reason, gray, prior);
    }
    out[q] = o;
}

// The same fill with the gap scans replaced by bit searches: one block per
// row, the row in LDS with one 64-bit mask of its valid pixels per 64
// columns (a wavefront ballot).  The left scan stops at the nearest valid
// pixel in [x - max_gap - 1, x - 1] if there is one, else at
// max(x - max_gap - 1, -1); the right scan likewise.
__global__ __launch_bounds__(256) void fill_mask_kernel(int w, int h, int max_gap, const float* __restrict__ in,
                                                        const uint8_t* __restrict__ reason,
                                                        const uint8_t* __restrict__ gray,
                                                        const float* __restrict__ prior, float* __restrict__ out) {
    extern __shared__ __align__(8) unsigned char fsh[];
    const int nseg = (w + 63) / 64;
    uint64_t* smask = (uint64_t*)fsh;
    float* srow = (float*)(fsh + (size_t)nseg * 8);
    const int y = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const float* row = in + (size_t)y * w;
    // four segments per wave per trip, their loads issued unconditionally
    // (clamped into the row) so that all four are in flight together
    for (int s0 = wave; s0 < nseg; s0 += 16) {
        float v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = row[min((s0 + 4 * j) * 64 + lane, w - 1)];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int s = s0 + 4 * j, x = s * 64 + lane;
            if (s < nseg) {   // uniform across the wave: the ballot sees every lane
                const float vv = x < w ? v[j] : -1.f;
                if (x < w) srow[x] = vv;
                const uint64_t m = __ballot(vv >= 0.f);
                if (lane == 0) smask[s] = m;
            }
        }
    }
    __syncthreads();
    for (int x = tid; x < w; x += 256) {
        float o = srow[x];
        if (o < 0.f) {
            const int lo = x - max_gap - 1, hi = x + max_gap + 1;
            // nearest valid pixel left of x, at or right of lo
            int L = -1;
            {
                int s = x >> 6;
                uint64_t m = smask[s] & ((1ull << (x & 63)) - 1ull);
                while (m == 0ull && s > 0 && s * 64 > lo) m = smask[--s];
                if (m != 0ull) L = s * 64 + 63 - __clzll((long long)m);
                if (L < lo) L = -1;
            }
            // nearest valid pixel right of x, at or left of hi
            int Rv = -1;
            {
                int s = x >> 6;
                const int b = x & 63;
                uint64_t m = b == 63 ? 0ull : smask[s] & ~((2ull << b) - 1ull);
                while (m == 0ull && s + 1 < nseg && (s + 1) * 64 <= hi) m = smask[++s];
                if (m != 0ull) Rv = s * 64 + __ffsll((long long)m) - 1;
                if (Rv > hi) Rv = -1;
            }
            const int xl = L >= 0 ? L : max(lo, -1);
            const int xr = Rv >= 0 ? Rv : min(hi, w);
            o = fill_hole(w, h, x, y, xl, xr, srow, in, reason, gray, prior);
        }
        out[(size_t)y * w + x] = o;
    }
}

// PostProcessor::fill_holes, Legacy route (reference src/postprocess.cpp:42-168):
// reads `in` (the LR-checked disparity), writes `out`.
// One block per row (ROW_LDS: the row is first copied to LDS, so the gap
// scans of hole pixels read LDS), otherwise one thread per pixel.
template <bool ROW_LDS>
__global__ void fill_kernel(int w, int h, int max_gap, const float* __restrict__ in, const uint8_t* __restrict__ reason,
                            const uint8_t* __restrict__ gray, const float* __restrict__ prior, float* __restrict__ out) {
    extern __shared__ float srow[];
    const int y = blockIdx.y;
    const float* row = in + (size_t)y * w;
    if (ROW_LDS) {
        for (int i = threadIdx.x; i < w; i += blockDim.x) srow[i] = row[i];
        __syncthreads();
        row = srow;
    }
    for (int x = ROW_LDS ? (int)threadIdx.x : (int)(blockIdx.x * blockDim.x + threadIdx.x); x < w;
         x += ROW_LDS ? (int)blockDim.x : w)
        fill_pixel(w, h, max_gap, x, y, row, in, reason, gray, prior, out);
}
// PostProcessor::median (reference src/postprocess.cpp:255-309): the element
// at position size/2 of the sorted valid window.
constexpr int kMaxMedianR = 3;
__global__ void median_kernel(int w, int h, int r, const float* __restrict__ in, float* __restrict__ out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    float win[(2 * kMaxMedianR + 1) * (2 * kMaxMedianR + 1)];
    int n = 0;
    for (int oy = -r; oy <= r; ++oy)
        for (int ox = -r; ox <= r; ++ox) {
            const int xx = x + ox, yy = y + oy;
            if (xx < 0 || yy < 0 || xx >= w || yy >= h) continue;
            const float d = in[(size_t)yy * w + xx];
            if (d >= 0.f) win[n++] = d;
        }
    const size_t q = (size_t)y * w + x;
    if (n == 0) { out[q] = in[q]; return; }
    for (int i = 1; i < n; ++i) {
        const float v = win[i];
        int j = i - 1;
        while (j >= 0 && win[j] > v) { win[j + 1] = win[j]; --j; }
        win[j + 1] = v;
    }
    out[q] = win[n / 2];
}

// Radius-1 median in registers: the 9 window values (invalid or outside the
// image as +inf) are sorted by a fixed compare-exchange network; the valid
// ones then lead the sorted order, so the result is sorted[n/2] as in the
// reference.
__device__ __forceinline__ void cswap(float& a, float& b) {
    const float lo = fminf(a, b), hi = fmaxf(a, b);
    a = lo;
    b = hi;
}

__global__ __launch_bounds__(256) void median3_kernel(int w, int h, const float* __restrict__ in,
                                                      float* __restrict__ out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    float v[9];
    int n = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int xx = x + k % 3 - 1, yy = y + k / 3 - 1;
        float d = -1.f;
        if (xx >= 0 && yy >= 0 && xx < w && yy < h) d = in[(size_t)yy * w + xx];
        const bool ok = d >= 0.f;
        n += ok ? 1 : 0;
        v[k] = ok ? d : INFINITY;
    }
    const size_t q = (size_t)y * w + x;
    if (n == 0) {
        out[q] = in[q];
        return;
    }
    // odd-even transposition sort: 9 rounds sort 9 inputs
#pragma unroll
    for (int r = 0; r < 9; ++r)
#pragma unroll
        for (int k = r & 1; k + 1 < 9; k += 2) cswap(v[k], v[k + 1]);
    float m = v[0];
#pragma unroll
    for (int k = 1; k < 9; ++k) m = (n / 2 == k) ? v[k] : m;
    out[q] = m;
}

// Fill and the radius-1 median in one pass over 64 x 16 output tiles.  The
// tile's input disparity and gray image are staged in LDS for rows y0-2 ..
// y0+17 and columns x0-64 .. x0+127 (outside the image: invalid), with one
// 64-bit mask of valid pixels per row and 64 columns (a wavefront ballot).
// The filled values of rows y0-1 .. y0+16, columns x0-1 .. x0+64 (every pixel
// the tile's 3x3 windows read) are then computed in LDS: a hole's gap scans
// stay inside [x - max_gap - 1, x + max_gap + 1], which the 64-column halo
// holds for max_gap <= 62.  The median of the filled values follows as in
// median3_kernel (radius 0: the filled value itself).
constexpr int FM_W = 64, FM_H = 16, FM_TC = 192, FM_TR = FM_H + 4, FM_FC = FM_W + 2, FM_FR = FM_H + 2;
constexpr int FM_GMAX = 62;

__device__ __forceinline__ void fill_median_body(const uint3 bid, int w, int h, int max_gap, int radius,
                                                 const float* __restrict__ in, const uint8_t* __restrict__ reason,
                                                 const uint8_t* __restrict__ gray, const float* __restrict__ prior,
                                                 float* __restrict__ out) {
    __shared__ float s_in[FM_TR * FM_TC];
    __shared__ uint8_t s_g[FM_TR * FM_TC];
    __shared__ uint64_t s_m[FM_TR * 3];
    __shared__ float s_f[FM_FR * FM_FC];
    const int x0 = bid.x * FM_W, y0 = bid.y * FM_H;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int xb = x0 - 64, yb = y0 - 2;   // image coordinates of tile cell (0, 0)
    // 60 row segments of 64 columns, 15 per wave; all loads issued first
    constexpr int NSEG = FM_TR * 3 / 4;
    float v[NSEG];
    uint32_t g[NSEG];
    uint32_t outside = 0u;   // bit i: segment i's cell lies outside the image
#pragma unroll
    for (int i = 0; i < NSEG; ++i) {
        const int sg = wave + 4 * i, r = sg / 3, c = (sg % 3) * 64 + lane;
        const int xx = xb + c, yy = yb + r;
        outside |= (uint32_t)!(xx >= 0 && xx < w && yy >= 0 && yy < h) << i;
        const size_t q = (size_t)clampi(yy, 0, h - 1) * w + clampi(xx, 0, w - 1);
        v[i] = in[q];
        g[i] = gray[q];
    }
    // the loaded values are first used here, after every load was issued
#pragma unroll
    for (int i = 0; i < NSEG; ++i) {
        const int sg = wave + 4 * i, r = sg / 3, c = (sg % 3) * 64 + lane;
        const float vv = (outside >> i) & 1u ? -1.f : v[i];
        s_in[r * FM_TC + c] = vv;
        s_g[r * FM_TC + c] = (uint8_t)g[i];
        const uint64_t m = __ballot(vv >= 0.f);
        if (lane == 0) s_m[sg] = m;
    }
    __syncthreads();
    auto in_at = [&](int xx, int yy) { return s_in[(yy - yb) * FM_TC + (xx - xb)]; };
    auto gray_at = [&](int xx, int yy) { return (int)s_g[(yy - yb) * FM_TC + (xx - xb)]; };
    // the reason codes (and prior values) of the thread's cells, all in flight
    // before the first hole needs one
    constexpr int NF = (FM_FR * FM_FC + 255) / 256;
    uint8_t rs[NF];
    float pv[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) {
        const int i = min(tid + 256 * k, FM_FR * FM_FC - 1);
        const int fr = i / FM_FC, fc = i - fr * FM_FC;
        const size_t q = (size_t)clampi(yb + fr + 1, 0, h - 1) * w + clampi(xb + fc + 63, 0, w - 1);
        rs[k] = reason[q];
        pv[k] = prior ? prior[q] : -1.f;
    }
#pragma unroll
    for (int k = 0; k < NF; ++k) {
        const int i = tid + 256 * k;
        if (i >= FM_FR * FM_FC) continue;
        const int fr = i / FM_FC, fc = i - fr * FM_FC;
        const int r = fr + 1, c = fc + 63;   // tile cell
        const int x = xb + c, y = yb + r;
        float o = -1.f;
        if (x >= 0 && x < w && y >= 0 && y < h) {
            o = s_in[r * FM_TC + c];
            if (o < 0.f) {
                const uint64_t* rm = s_m + r * 3;
                const int clo = c - max_gap - 1, chi = c + max_gap + 1;
                int Lc = -1;
                {
                    int sg = c >> 6;
                    uint64_t m = rm[sg] & ((1ull << (c & 63)) - 1ull);
                    while (m == 0ull && sg > 0 && sg * 64 > clo) m = rm[--sg];
                    if (m != 0ull) Lc = sg * 64 + 63 - __clzll((long long)m);
                    if (Lc < clo) Lc = -1;
                }
                int Rc = -1;
                {
                    int sg = c >> 6;
                    const int b = c & 63;
                    uint64_t m = b == 63 ? 0ull : rm[sg] & ~((2ull << b) - 1ull);
                    while (m == 0ull && sg < 2 && (sg + 1) * 64 <= chi) m = rm[++sg];
                    if (m != 0ull) Rc = sg * 64 + __ffsll((long long)m) - 1;
                    if (Rc > chi) Rc = -1;
                }
                const int xl = Lc >= 0 ? xb + Lc : max(x - max_gap - 1, -1);
                const int xr = Rc >= 0 ? xb + Rc : min(x + max_gap + 1, w);
                o = fill_hole_at(w, h, x, y, xl, xr, o, rs[k] == R_OCCLUSION, gray_at(x, y), pv[k], in_at, gray_at);
            }
        }
        s_f[i] = o;
    }
    __syncthreads();
    for (int i = tid; i < FM_W * FM_H; i += 256) {
        const int ry = i / FM_W, rx = i - ry * FM_W;
        const int x = x0 + rx, y = y0 + ry;
        if (x >= w || y >= h) continue;
        const float* fc = s_f + (ry + 1) * FM_FC + rx + 1;   // the pixel's filled value
        float res = fc[0];
        if (radius == 1) {
            float vv[9];
            int n = 0;
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                const int ox = k % 3 - 1, oy = k / 3 - 1;
                const int xx = x + ox, yy = y + oy;
                float d = -1.f;
                if (xx >= 0 && yy >= 0 && xx < w && yy < h) d = fc[oy * FM_FC + ox];
                const bool ok = d >= 0.f;
                n += ok ? 1 : 0;
                vv[k] = ok ? d : INFINITY;
            }
            if (n > 0) {
#pragma unroll
                for (int rr = 0; rr < 9; ++rr)
#pragma unroll
                    for (int k = rr & 1; k + 1 < 9; k += 2) cswap(vv[k], vv[k + 1]);
                float m = vv[0];
#pragma unroll
                for (int k = 1; k < 9; ++k) m = (n / 2 == k) ? vv[k] : m;
                res = m;
            }
        }
        out[(size_t)y * w + x] = res;
    }
}

__global__ __launch_bounds__(256) void fill_median_kernel(int w, int h, int max_gap, int radius,
                                                          const float* __restrict__ in,
                                                          const uint8_t* __restrict__ reason,
                                                          const uint8_t* __restrict__ gray,
                                                          const float* __restrict__ prior, float* __restrict__ out) {
    fill_median_body(block3(), w, h, max_gap, radius, in, reason, gray, prior, out);
}

struct FillMedianFn {
    static constexpr int kThreads = 256;
    template <typename... A>
    __device__ static void run(const uint3 b, const A&... a) { fill_median_body(b, a...); }
};

bool fill_median_fusable(const apg_hip_params& p) {
    return p.post_median_radius >= 0 && p.post_median_radius <= 1 && p.post_max_fill_gap >= 0 &&
           p.post_max_fill_gap <= FM_GMAX;
}

hipError_t launch_fill_median(hipStream_t s, const apg_hip_params& p, int w, int h, const float* disp,
                              const uint8_t* reason, const uint8_t* gray, const float* prior, float* out,
                              LaunchSink* sink) {
    if (!fill_median_fusable(p)) return hipErrorInvalidValue;
    if (sink)
        return sink->add<FillMedianFn>(dim3((w + FM_W - 1) / FM_W, (h + FM_H - 1) / FM_H), 0, w, h,
                                       p.post_max_fill_gap, p.post_median_radius, disp, reason, gray, prior, out);
    hipLaunchKernelGGL(fill_median_kernel, dim3((w + FM_W - 1) / FM_W, (h + FM_H - 1) / FM_H), dim3(256), 0, s, w, h,
                       p.post_max_fill_gap, p.post_median_radius, disp, reason, gray, prior, out);
    return hipGetLastError();
}

hipError_t launch_fill(hipStream_t s, const apg_hip_params& p, int w, int h, const float* disp, const uint8_t* reason,
                       const uint8_t* gray, const float* prior, float* out) {
    const size_t shm = (size_t)(w + 63) / 64 * 8 + (size_t)w * sizeof(float);
    if (shm <= 64 * 1024)
        hipLaunchKernelGGL(fill_mask_kernel, dim3(h), dim3(256), shm, s, w, h, p.post_max_fill_gap, disp, reason,
                           gray, prior, out);
    else if ((size_t)w * sizeof(float) <= 32 * 1024)
        hipLaunchKernelGGL(fill_kernel<true>, dim3(1, h), dim3(256), (size_t)w * sizeof(float), s, w, h,
                           p.post_max_fill_gap, disp, reason, gray, prior, out);
    else
        hipLaunchKernelGGL(fill_kernel<false>, dim3((w + 127) / 128, h), dim3(128), 0, s, w, h, p.post_max_fill_gap,
                           disp, reason, gray, prior, out);
    return hipGetLastError();
}

// Median of any radius: a 16 x 16 tile of pixels with its radius-r halo in
// LDS (invalid or outside the image as a key of all ones); each pixel selects
// the (n/2)-th smallest of its n valid window values, which is the value
// std::nth_element places at n/2.  Valid values are >= 0, so the low 31 bits
// of a float order like the float (-0 folds onto +0); the selection fixes the
// key one bit at a time from bit 30 down, counting the window values below
// the candidate prefix.
constexpr int kMedTile = 16;
__global__ __launch_bounds__(kMedTile * kMedTile) void median_select_kernel(int w, int h, int r,
                                                                             const float* __restrict__ in,
                                                                             float* __restrict__ out) {
    extern __shared__ uint32_t s_key[];
    const int tw = kMedTile + 2 * r;
    const int x0 = blockIdx.x * kMedTile - r, y0 = blockIdx.y * kMedTile - r;
    const int tid = threadIdx.y * kMedTile + threadIdx.x;
    for (int i = tid; i < tw * tw; i += kMedTile * kMedTile) {
        const int xx = x0 + i % tw, yy = y0 + i / tw;
        uint32_t k = 0xFFFFFFFFu;
        if (xx >= 0 && yy >= 0 && xx < w && yy < h) {
            const float d = in[(size_t)yy * w + xx];
            if (d >= 0.f) k = __float_as_uint(d) & 0x7FFFFFFFu;
        }
        s_key[i] = k;
    }
    __syncthreads();
    const int x = blockIdx.x * kMedTile + threadIdx.x, y = blockIdx.y * kMedTile + threadIdx.y;
    if (x >= w || y >= h) return;
    const int side = 2 * r + 1;
    const uint32_t* win = s_key + threadIdx.y * tw + threadIdx.x;
    int n = 0;
    for (int oy = 0; oy < side; ++oy)
        for (int ox = 0; ox < side; ++ox) n += win[oy * tw + ox] != 0xFFFFFFFFu ? 1 : 0;
    const size_t q = (size_t)y * w + x;
    if (n == 0) {
        out[q] = in[q];
        return;
    }
    int k = n / 2;   // rank among the valid keys
    uint32_t prefix = 0u;
    for (int b = 30; b >= 0; --b) {
        const uint32_t hi_mask = ~((2u << b) - 1u) & 0x7FFFFFFFu;   // bits above b
        int c = 0;   // valid keys matching the prefix above b with bit b clear
        for (int oy = 0; oy < side; ++oy)
            for (int ox = 0; ox < side; ++ox) {
                const uint32_t v = win[oy * tw + ox];
                c += (v != 0xFFFFFFFFu && (v & hi_mask) == prefix && !((v >> b) & 1u)) ? 1 : 0;
            }
        if (k >= c) {
            k -= c;
            prefix |= 1u << b;
        }
    }
    out[q] = __uint_as_float(prefix);
}

hipError_t launch_median(hipStream_t s, const apg_hip_params& p, int w, int h, const float* disp, float* out) {
    if (p.post_median_radius > kMaxMedianR) {
        const int r = p.post_median_radius;
        const size_t shm = (size_t)(kMedTile + 2 * r) * (kMedTile + 2 * r) * 4;
        if (shm > 64 * 1024) {
            hipError_t e = hipFuncSetAttribute((const void*)median_select_kernel,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(median_select_kernel, dim3((w + kMedTile - 1) / kMedTile, (h + kMedTile - 1) / kMedTile),
                           dim3(kMedTile, kMedTile), shm, s, w, h, r, disp, out);
        return hipGetLastError();
    }
    if (p.post_median_radius == 1) {
        hipLaunchKernelGGL(median3_kernel, dim3((w + 255) / 256, h), dim3(256), 0, s, w, h, disp, out);
        return hipGetLastError();
    }
    if (p.post_median_radius > 0) {
        hipLaunchKernelGGL(median_kernel, dim3((w + 127) / 128, h), dim3(128), 0, s, w, h, p.post_median_radius, disp,
                           out);
        return hipGetLastError();
    }
    return hipMemcpyAsync(out, disp, sizeof(float) * (size_t)w * h, hipMemcpyDeviceToDevice, s);
}

}  // namespace apgk
