// gfx950 kernels of the HighQuality prior (reference src/prior_estimator.cpp):
//   prior_cand_kernel : every prior sample (even x, y inside a 2-pixel border)
//                       runs the census-only WTA with best / second cost, the
//                       uniqueness test and the right-view check of
//                       extract_supports (:53-101); one thread per sample.
//   prior_range_kernel: bilinear interpolation of the filled cell grid to
//                       every pixel (interpolate_prior :312-374) and the
//                       per-pixel search range (apply_search_range :377-412),
//                       one thread per pixel.
// The support selection and the grid statistics between them run on the
// host (hq_host.cpp).  Float expressions keep the reference's order; the
// library is built without contraction and with IEEE division.
#include "device_common.hpp"
#include "hq_host.hpp"
#include "kernels.hpp"

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
                                                          int16_t* __restrict__ dmin, int16_t* __restrict__ dmax,
                                                          float* __restrict__ d_prior) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
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

hipError_t launch_prior_ranges(hipStream_t s, const apg_hip_params& p, int w, int h, int gw, int gh,
                               const float* grid, int16_t* dmin, int16_t* dmax, float* d_prior) {
    hipLaunchKernelGGL(prior_range_kernel, dim3((w + 255) / 256, h), dim3(256), 0, s, w, h, gw, gh, p.min_disparity,
                       p.max_disparity, grid, dmin, dmax, d_prior);
    return hipGetLastError();
}

}  // namespace apgk
