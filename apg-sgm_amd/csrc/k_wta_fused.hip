// gfx950 kernel of the fused schedule: path sums rebuilt from C' and the
// per-direction delta planes (S = npath * C' + sum of the deltas), winner-take-all + second best + legacy subpixel
// (reference src/sgm_optimizer.cpp:436-555), confidence (reference
// src/confidence_estimator.cpp:7-63) and the left-right check (reference
// src/postprocess.cpp:10-40), one lane per pixel.
#include <math.h>

#include "device_common.hpp"
#include "kernels.hpp"

namespace apgk {

constexpr int WF_PB = 64;    // pixels per block, one per lane
constexpr int WF_DC = 128;   // disparities per staged chunk
constexpr int WF_STRIDE = WF_DC / 2 + 1;   // u32 words per staged pixel row (odd: conflict-free)

__device__ __forceinline__ float clampf_(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

template <typename TC, int NP>
__global__ __launch_bounds__(64) void wta_fused_kernel(apg_hip_params p, int w, int h, int D,
                                                       const TC* __restrict__ cost, PlaneSet planes,
                                                       const int16_t* __restrict__ dmin,
                                                       const int16_t* __restrict__ dmax, const float* __restrict__ dr,
                                                       const uint8_t* __restrict__ gx, const uint8_t* __restrict__ gy,
                                                       float* __restrict__ disp, float* __restrict__ conf_out,
                                                       uint8_t* __restrict__ reason, uint8_t* __restrict__ reliable,
                                                       float* __restrict__ disp_wta, float* __restrict__ best_out,
                                                       float* __restrict__ second_out, uint32_t* __restrict__ sum32) {
    __shared__ uint32_t st[WF_PB * WF_STRIDE];
    constexpr uint32_t CINV = sizeof(TC) == 1 ? 0xFFu : 0xFFFFu;
    const int lane = threadIdx.x;
    const size_t npx = (size_t)w * h;
    const size_t p_base = (size_t)blockIdx.x * WF_PB;
    const int npb = (int)min((size_t)WF_PB, npx - p_base);
    const size_t pix = p_base + lane;
    const bool act = lane < npb;
    const int d0 = p.min_disparity;
    const int lo = act ? dmin[pix] : 0, hi = act ? dmax[pix] : 0;
    constexpr uint32_t npath = NP;
    uint32_t bv = 0xFFFFFFFFu, pm = 0xFFFFFFFFu, pm_lag = 0xFFFFFFFFu, secL = 0xFFFFFFFFu, secR = 0xFFFFFFFFu;
    uint32_t raw_prev = 0xFFFFu, v_bm1 = 0xFFFFu, v_bp1 = 0xFFFFu;
    int bd = -1;
    for (int dc = 0; dc < D; dc += WF_DC) {
        const int nd = min(WF_DC, D - dc);   // even
        // stage S for the block's pixels: element (i, j) of the chunk at st[i*STRIDE + j/2] (u16 halves)
        // one pixel's chunk per wave iteration: 64 lanes x 2 disparities,
        // contiguous; WF_G pixels of raw words are loaded before any is decoded
        // (lanes past the chunk read its first pair and store nothing)
        const int jw = lane;
        const bool jin = jw < nd / 2;
        const int jo = jin ? 2 * jw : 0;
        constexpr int WF_G = NP == 8 ? 4 : 8;
        for (int i0 = 0; i0 < npb; i0 += WF_G) {
            uint32_t rc[WF_G], rd[WF_G][NP];
#pragma unroll
            for (int g = 0; g < WF_G; ++g) {
                const size_t q = (p_base + min(i0 + g, npb - 1)) * (size_t)D + dc + jo;
                rc[g] = sizeof(TC) == 1 ? (uint32_t) * (const uint16_t*)((const uint8_t*)cost + q)
                                        : *(const uint32_t*)((const uint16_t*)cost + q);
#pragma unroll
                for (int f = 0; f < NP; ++f) rd[g][f] = (uint32_t) * (const uint16_t*)(planes.p[f] + q);
            }
#pragma unroll
            for (int g = 0; g < WF_G; ++g) {
                const int i = i0 + g;
                if (i >= npb || !jin) continue;
                const size_t q = (p_base + i) * (size_t)D + dc + 2 * jw;
                uint32_t c0, c1;
                if (sizeof(TC) == 1) { c0 = rc[g] & 0xFF; c1 = (rc[g] >> 8) & 0xFF; }
                else { c0 = rc[g] & 0xFFFF; c1 = rc[g] >> 16; }
                uint32_t a0 = npath * c0, a1 = npath * c1;
#pragma unroll
                for (int f = 0; f < NP; ++f) {
                    a0 += rd[g][f] & 0xFF;
                    a1 += (rd[g][f] >> 8) & 0xFF;
                }
                const uint32_t s0 = c0 == CINV ? 0xFFFFu : a0, s1 = c1 == CINV ? 0xFFFFu : a1;
                st[i * WF_STRIDE + jw] = s0 | (s1 << 16);
                if (sum32) {
                    sum32[q] = s0 == 0xFFFFu ? 0xFFFFFFFFu : s0;
                    sum32[q + 1] = s1 == 0xFFFFu ? 0xFFFFFFFFu : s1;
                }
            }
        }
        __syncthreads();
        if (act) {
            for (int jw = 0; jw < nd / 2; ++jw) {
                const uint32_t two = st[lane * WF_STRIDE + jw];
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const int di = dc + 2 * jw + e;
                    const int d = d0 + di;
                    const uint32_t raw = e ? (two >> 16) : (two & 0xFFFFu);
                    if (bd >= 0 && d == bd + 1) v_bp1 = raw;
                    const bool valid = d >= lo && d < hi && raw != 0xFFFFu;
                    const uint32_t c = valid ? raw : 0xFFFFFFFFu;
                    const uint32_t pm_new = min(pm, c);
                    if (valid && c < bv) {
                        secL = pm_lag;
                        secR = 0xFFFFFFFFu;
                        bv = c;
                        bd = d;
                        v_bm1 = raw_prev;
                        v_bp1 = 0xFFFFu;
                    } else if (valid && bd >= 0 && d >= bd + 2) {
                        secR = min(secR, c);
                    }
                    pm_lag = pm;
                    pm = pm_new;
                    raw_prev = raw;
                }
            }
        }
        __syncthreads();
    }
    if (!act) return;
    float dv = -1.f, b = INFINITY, sc = INFINITY;
    if (hi > lo && bd >= 0) {
        const uint32_t sec = min(secL, secR);
        float dval = (float)bd;
        if (p.post_subpixel && bd - 1 >= lo && bd + 1 < hi) {
            const int di = bd - d0;
            if (di > 0 && di + 1 < D && v_bm1 != 0xFFFFu && v_bp1 != 0xFFFFu)
                dval += legacy_delta((float)v_bm1, (float)bv, (float)v_bp1);
        }
        dv = dval;
        b = (float)bv;
        sc = sec != 0xFFFFFFFFu ? (float)sec : INFINITY;
    }
    if (disp_wta) disp_wta[pix] = dv;
    if (best_out) best_out[pix] = b;
    if (second_out) second_out[pix] = sc;
    // confidence on the WTA disparity
    const int x = (int)(pix % (size_t)w);
    const size_t row = pix - (size_t)x;
    float uniq = 0.f;
    if (isfinite(b) && isfinite(sc)) uniq = clampf_((sc - b) / (sc < 1e-3f ? 1e-3f : sc), 0.f, 1.f);
    float lr = 1.f;
    {
        const int xr = (int)roundf((float)x - dv);
        if (xr < 0 || xr >= w || dv < 0.f) lr = 0.f;
        else {
            const float drv = dr[row + xr];
            if (drv < 0.f || fabsf(dv - drv) > (float)p.conf_lr_max_diff + 0.5f) lr = 0.f;
        }
    }
    const float tex = (float)((int)gx[pix] + (int)gy[pix]);
    const float tex_term = clampf_(tex / 16.f, 0.f, 1.f);
    float conf = clampf_(uniq, 0.f, 4.f);
    conf = conf * (0.5f + 0.5f * tex_term);
    if (lr < 0.5f) conf *= 0.25f;
    const bool lr_ok = lr >= 0.5f;
    const bool rel = lr_ok && conf >= p.conf_reliable_threshold && uniq >= p.conf_uniqueness_ratio &&
                     tex >= p.conf_min_texture && dv >= 0.f;
    uint8_t why = rel ? 0 : (lr_ok ? 3 : 2);
    // left-right check on the WTA disparity
    float dout = dv;
    if (p.post_lr_check && dv >= 0.f) {
        const float tol = (float)p.post_lr_max_diff + 0.5f;
        const int xr = (int)roundf((float)x - dv);
        if (xr < 0 || xr >= w) {
            dout = -1.f;
            why = 1;
        } else {
            const float drv = dr[row + xr];
            if (drv < 0.f || fabsf(dv - drv) > tol) {
                dout = -1.f;
                why = (drv > dv + tol) ? 1 : 2;
            }
        }
    }
    conf_out[pix] = conf;
    reliable[pix] = rel ? 1 : 0;
    reason[pix] = why;
    disp[pix] = dout;
}

hipError_t launch_wta_fused(hipStream_t s, const apg_hip_params& p, int w, int h, const void* cost, bool cost_u8,
                            const PlaneSet& planes, const int16_t* dmin, const int16_t* dmax, const float* dr,
                            const uint8_t* gx, const uint8_t* gy, float* disp, float* conf, uint8_t* reason,
                            uint8_t* reliable, float* disp_wta, float* best, float* second, uint32_t* sum32) {
    const int D = p.max_disparity - p.min_disparity;
    if (D % 2) return hipErrorInvalidValue;
    const size_t npx = (size_t)w * h;
    dim3 grid((unsigned)((npx + WF_PB - 1) / WF_PB));
#define APG_WTAF(TC, NP)                                                                                          \
    hipLaunchKernelGGL((wta_fused_kernel<TC, NP>), grid, dim3(64), 0, s, p, w, h, D, (const TC*)cost, planes, dmin, \
                       dmax, dr, gx, gy, disp, conf, reason, reliable, disp_wta, best, second, sum32)
    if (p.paths == 8) {
        if (cost_u8) APG_WTAF(uint8_t, 8);
        else APG_WTAF(uint16_t, 8);
    } else {
        if (cost_u8) APG_WTAF(uint8_t, 4);
        else APG_WTAF(uint16_t, 4);
    }
#undef APG_WTAF
    return hipGetLastError();
}

}  // namespace apgk
