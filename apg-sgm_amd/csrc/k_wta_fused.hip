// gfx950 kernel of the fused schedule: path sums rebuilt from C' and the
// per-direction delta planes (S = npath * C' + sum of the deltas), winner-take-all + second best + legacy subpixel
// (reference src/sgm_optimizer.cpp:436-555), confidence (reference
// src/confidence_estimator.cpp:7-63) and the left-right check (reference
// src/postprocess.cpp:10-40), one lane per pixel.
#include <math.h>

#include "device_common.hpp"
#include "kernels.hpp"

namespace apgk {

constexpr int WF_PB = 64;   // pixels per block, one per lane
constexpr int WF_G = 4;     // pixels whose raw words are in flight together while staging

__device__ __forceinline__ float clampf_(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

// S of the block's 64 pixels is staged in LDS as u16 (row of a pixel: D
// values, 65535 = invalid, stride `rw` words, odd so that the lanes' rows
// fall in different banks).  C' is invalid outside the pixel's range, so S
// needs no separate range mask.  Then each lane scans its own row twice:
//   1. best = min over d of (S << 8 | d): the earliest minimum (D <= 256);
//   2. after S(best-1), S(best+1) are read and S at best-1..best+1 is set to
//      65535, second = min of the row (packed u16 minima).
// S of one (d, d+1) pair from C' and the planes' deltas, all as u16 halves
// (u8 C' undecoded: 255 = invalid): npath * C' + sum of the deltas (< 65535
// for valid C', checked by the host), 65535 where C' is invalid.
// ALLV: the caller knows no C' entry of the pair is invalid.  MUL is the
// number of paths: NP for one plane per direction, 2 NP where each plane
// holds the sum of a scanline's two directions.
template <typename TC, int NP, bool ALLV = false, int MUL = NP>
__device__ __forceinline__ uint32_t s_pair(uint32_t ch, const uint32_t (&ph)[NP]) {
    if (ALLV) {
        u16x2 ds = as_v2(ph[0]);
#pragma unroll
        for (int f = 1; f < NP; ++f) ds = ds + as_v2(ph[f]);
        return as_u32(as_v2(ch) * splat2(MUL) + ds);
    }
    const uint32_t c16 = sizeof(TC) == 1 ? as_u32(as_v2(ch) - (as_v2((ch + 0x00010001u) & 0xFF00FF00u)))   // 255 -> 65535
                                         : ch;
    u16x2 ds = as_v2(ph[0]);
#pragma unroll
    for (int f = 1; f < NP; ++f) ds = ds + as_v2(ph[f]);
    const uint32_t sv = as_u32(as_v2(c16) * splat2(MUL) + ds);
    const uint32_t inv = as_u32((as_v2(c16) + splat2(1)) == splat2(0) ? splat2(0xFFFF) : splat2(0));
    return sv | inv;
}

// WIDE (u8 C', D % 16 == 0, no sum32 output): the block's 64 pixels are one
// contiguous run of 64*D elements in C' and in every plane, and S is
// elementwise over it, so staging reads 16-byte chunks (WF_WG per lane in
// flight, all planes) and splits a chunk's (pixel, d) only for the LDS store.
constexpr int WF_WG = 4;

// LP lanes per pixel (1 or 2): with 2, a block holds 32 pixels, each lane
// scans half of its pixel's row and the two halves are combined by a DPP
// swap of adjacent lanes; half the LDS per wave, so twice the waves per CU
// keep loads in flight while others scan.
template <typename TC, int NP, bool WIDE, int LP, int MUL>
__device__ __forceinline__ void wta_fused_body(const uint3 bid, const apg_hip_params& p, int w, int h, int D, int rw,
                                               const TC* __restrict__ cost, const PlaneSet& planes,
                                               const int16_t* __restrict__ dmin, const int16_t* __restrict__ dmax,
                                               const float* __restrict__ dr, const uint8_t* __restrict__ gx,
                                               const uint8_t* __restrict__ gy, float* __restrict__ disp,
                                               float* __restrict__ conf_out, uint8_t* __restrict__ reason,
                                               uint8_t* __restrict__ reliable, float* __restrict__ disp_wta,
                                               float* __restrict__ best_out, float* __restrict__ second_out,
                                               uint32_t* __restrict__ sum32) {
    extern __shared__ uint32_t st[];   // [PB][rw] words
    constexpr int PB = WF_PB / LP;     // pixels per block
    const int lane = threadIdx.x;
    const int pi = lane / LP, hf = lane % LP;
    const size_t npx = (size_t)w * h;
    const size_t p_base = (size_t)bid.x * PB;
    const int npb = (int)min((size_t)PB, npx - p_base);
    const size_t pix = p_base + pi;
    const bool act = pi < npb;
    const int d0 = p.min_disparity;
    const int nw = D / 2;   // words per row (D even)
    // two lanes per pixel with D % 64 == 0: the row's second half starts one
    // word late and rows are rw = nw + 2 words apart (rw = 2 mod 32), so the
    // 32 lanes (pixel, half) of one scan read, and the staging writes, fall
    // in 32 different banks; logical word j sits at j + (j >= gapw)
    const int gapw = (LP == 2 && D % 64 == 0) ? nw / 2 : (1 << 30);
    const auto rc = raw_rsrc(cost);
    if (WIDE) {
        const uint32_t ebase = (uint32_t)(p_base * (size_t)D);
        const int nchunk = npb * D / 16;
        // every state of the block valid: its pixels lie in one row with x >= the
        // largest disparity (column ranges: C'(x, d) is valid iff d <= x)
        const int xb0 = (int)(p_base % (size_t)w);
        const bool allv = !planes.general_ranges && xb0 + npb <= w && xb0 >= d0 + D - 1 && d0 >= 0;
        auto stage = [&](int k0, auto allv_t) {
            constexpr bool AV = decltype(allv_t)::value;
            u32x4 cw[WF_WG], pw[WF_WG][NP];
#pragma unroll
            for (int g = 0; g < WF_WG; ++g) {
                const uint32_t off = (uint32_t)min(k0 + g * 64 + lane, nchunk - 1) * 16u;
                cw[g] = buf_ld128(rc, off, ebase);
#pragma unroll
                for (int f = 0; f < NP; ++f) pw[g][f] = buf_ld128(raw_rsrc(planes.p[f]), off, ebase);
            }
#pragma unroll
            for (int g = 0; g < WF_WG; ++g) {
                const int ch = k0 + g * 64 + lane;
                if (ch >= nchunk) continue;
                const int e = ch * 16;
                const int i = e / D;
                const int wo = (e - i * D) / 2;   // a chunk's 8 words lie in one half (nw / 2 % 8 == 0)
                uint32_t* dst = st + i * rw + wo + (wo >= gapw ? 1 : 0);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
#pragma unroll
                    for (int hh = 0; hh < 2; ++hh) {
                        const uint32_t sel = hh ? 0x0c030c02u : 0x0c010c00u;   // bytes -> u16 halves
                        uint32_t ph[NP];
#pragma unroll
                        for (int f = 0; f < NP; ++f) ph[f] = __builtin_amdgcn_perm(0u, pw[g][f][q], sel);
                        dst[q * 2 + hh] = s_pair<TC, NP, AV, MUL>(__builtin_amdgcn_perm(0u, cw[g][q], sel), ph);
                    }
                }
            }
        };
        for (int k0 = 0; k0 < nchunk; k0 += 64 * WF_WG) {
            if (allv) stage(k0, std::true_type{});
            else stage(k0, std::false_type{});
        }
    }
    // staging: one pixel per wave iteration, lane = pair index (two passes for D > 128)
    if (!WIDE)
    for (int j0 = 0; j0 < nw; j0 += 64) {
        const int jw = j0 + lane;
        const bool jin = jw < nw;
        const uint32_t cvo = (uint32_t)(jin ? jw : 0) * 2u * (uint32_t)sizeof(TC);
        const uint32_t pvo = (uint32_t)(jin ? jw : 0) * 2u;
        for (int i0 = 0; i0 < npb; i0 += WF_G) {
            uint32_t rcw[WF_G], rd[WF_G][NP];
#pragma unroll
            for (int g = 0; g < WF_G; ++g) {
                const uint32_t px = (uint32_t)(p_base + min(i0 + g, npb - 1));
                rcw[g] = sizeof(TC) == 1 ? buf_ld16(rc, cvo, px * (uint32_t)D) : buf_ld32(rc, cvo, px * 2u * (uint32_t)D);
#pragma unroll
                for (int f = 0; f < NP; ++f) rd[g][f] = buf_ld16(raw_rsrc(planes.p[f]), pvo, px * (uint32_t)D);
            }
#pragma unroll
            for (int g = 0; g < WF_G; ++g) {
                const int i = i0 + g;
                if (i >= npb || !jin) continue;
                uint32_t ph[NP];
#pragma unroll
                for (int f = 0; f < NP; ++f) ph[f] = bytes_to_halves(rd[g][f]);
                const uint32_t S = s_pair<TC, NP, false, MUL>(sizeof(TC) == 1 ? bytes_to_halves(rcw[g]) : rcw[g], ph);
                st[i * rw + jw] = S;
                if (sum32) {
                    const size_t q = (p_base + i) * (size_t)D + 2 * jw;
                    sum32[q] = (S & 0xFFFFu) == 0xFFFFu ? 0xFFFFFFFFu : (S & 0xFFFFu);
                    sum32[q + 1] = (S >> 16) == 0xFFFFu ? 0xFFFFFFFFu : (S >> 16);
                }
            }
        }
    }
    __syncthreads();
    if (!act) return;
    const int lo = dmin[pix], hi = dmax[pix];
    uint32_t* srow = st + pi * rw;
    const int j0 = hf * (nw / LP), j1 = j0 + nw / LP;   // this lane's words of the row
    const uint32_t* lrow = srow + (j0 >= gapw ? 1 : 0);   // them, past the gap
    auto p16 = [&](int u) { return u + ((u >> 1) >= gapw ? 2 : 0); };   // u16 entry u of the row
    // pass 1: earliest minimum as the key S << 8 | d-index
    uint32_t best = 0xFFFFFFFFu;
    for (int j = j0; j < j1; ++j) {
        const uint32_t v = lrow[j];
        const uint32_t idx = (uint32_t)(2 * j) | ((uint32_t)(2 * j + 1) << 8);
        const uint32_t kl = __builtin_amdgcn_perm(v, idx, 0x0c050400u);
        const uint32_t kh = __builtin_amdgcn_perm(v, idx, 0x0c070601u);
        best = min(best, min(kl, kh));   // v_min3_u32
    }
    if (LP == 2) best = min(best, (uint32_t)__builtin_amdgcn_mov_dpp((int)best, 0xB1, 0xF, 0xF, false));
    const uint32_t bv = best >> 8;
    const int bi = (int)(best & 0xFFu);
    // as in the reference, a negative winning disparity counts as no match
    // (best_d < 0 test, reference src/sgm_optimizer.cpp:491-496)
    const bool found = bv != 0xFFFFu && d0 + bi >= 0;
    // the row is only accessed as 32-bit words, so the compiler keeps these
    // reads, the poisoning and the second scan in program order (16-bit
    // accesses to it would not alias the 32-bit scan for the compiler)
    auto get16 = [&](int u) { const int q = p16(u); return (srow[q >> 1] >> (16 * (q & 1))) & 0xFFFFu; };
    auto poison = [&](int u) { const int q = p16(u); srow[q >> 1] |= 0xFFFFu << (16 * (q & 1)); };
    uint32_t v_m1 = 0xFFFFu, v_p1 = 0xFFFFu;
    if (found) {
        v_m1 = bi > 0 ? get16(bi - 1) : 0xFFFFu;
        v_p1 = bi + 1 < D ? get16(bi + 1) : 0xFFFFu;
        poison(bi);
        if (bi > 0) poison(bi - 1);
        if (bi + 1 < D) poison(bi + 1);
    }
    // pass 2: second best outside best-1..best+1 (both lanes of a pixel
    // poisoned the same words, and LDS operations of a wave are in order)
    uint32_t m2 = 0xFFFFFFFFu;
    for (int j = j0; j < j1; ++j) m2 = as_u32(__builtin_elementwise_min(as_v2(m2), as_v2(lrow[j])));
    if (LP == 2)
        m2 = as_u32(__builtin_elementwise_min(as_v2(m2), as_v2((uint32_t)__builtin_amdgcn_mov_dpp((int)m2, 0xB1, 0xF,
                                                                                                 0xF, false))));
    const uint32_t sec = min(m2 & 0xFFFFu, m2 >> 16);
    float dv = -1.f, b = INFINITY, sc = INFINITY;
    if (found) {
        const int bd = d0 + bi;
        float dval = (float)bd;
        if (p.post_subpixel && bd - 1 >= lo && bd + 1 < hi && bi > 0 && bi + 1 < D && v_m1 != 0xFFFFu &&
            v_p1 != 0xFFFFu)
            dval += legacy_delta((float)v_m1, (float)bv, (float)v_p1);
        dv = dval;
        b = (float)bv;
        sc = sec != 0xFFFFu ? (float)sec : INFINITY;
    }
    if (hf != 0) return;   // the pixel's first lane writes its outputs
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

template <typename TC, int NP, bool WIDE, int LP, int MUL>
__global__ __launch_bounds__(64) void wta_fused_kernel(apg_hip_params p, int w, int h, int D, int rw,
                                                       const TC* __restrict__ cost, PlaneSet planes,
                                                       const int16_t* __restrict__ dmin,
                                                       const int16_t* __restrict__ dmax, const float* __restrict__ dr,
                                                       const uint8_t* __restrict__ gx, const uint8_t* __restrict__ gy,
                                                       float* __restrict__ disp, float* __restrict__ conf_out,
                                                       uint8_t* __restrict__ reason, uint8_t* __restrict__ reliable,
                                                       float* __restrict__ disp_wta, float* __restrict__ best_out,
                                                       float* __restrict__ second_out, uint32_t* __restrict__ sum32) {
    wta_fused_body<TC, NP, WIDE, LP, MUL>(block3(), p, w, h, D, rw, cost, planes, dmin, dmax, dr, gx, gy, disp,
                                          conf_out, reason, reliable, disp_wta, best_out, second_out, sum32);
}

template <typename TC, int NP, bool WIDE, int LP, int MUL>
struct WtaFusedFn {
    static constexpr int kThreads = 64;
    template <typename... A>
    __device__ static void run(const uint3 b, const A&... a) { wta_fused_body<TC, NP, WIDE, LP, MUL>(b, a...); }
};

hipError_t launch_wta_fused(hipStream_t s, const apg_hip_params& p, int w, int h, const void* cost, bool cost_u8,
                            const PlaneSet& planes, const int16_t* dmin, const int16_t* dmax, const float* dr,
                            const uint8_t* gx, const uint8_t* gy, float* disp, float* conf, uint8_t* reason,
                            uint8_t* reliable, float* disp_wta, float* best, float* second, uint32_t* sum32,
                            LaunchSink* sink) {
    const int D = p.max_disparity - p.min_disparity;
    if (D % 2 || D > 256) return hipErrorInvalidValue;
    const size_t npx = (size_t)w * h;
    const bool wide = cost_u8 && D % 16 == 0 && !sum32;
    const int lp = wide ? 2 : 1;
    // row stride in words: odd for one lane per pixel; with two lanes and
    // D % 64 == 0, D/2 + 2 (2 mod 32) plus the one-word gap between the halves
    const int rw = (wide && D % 64 == 0) ? D / 2 + 2 : (D / 2) | 1;
    const size_t shm = (size_t)(WF_PB / lp) * rw * 4;
    dim3 grid((unsigned)((npx + WF_PB / lp - 1) / (WF_PB / lp)));
#define APG_WTAF(TC, NP, WD, MUL)                                                                              \
    do {                                                                                                       \
        if (sink)                                                                                              \
            return sink->add<WtaFusedFn<TC, NP, WD, (WD ? 2 : 1), MUL>>(                                       \
                grid, shm, p, w, h, D, rw, (const TC*)cost, planes, dmin, dmax, dr, gx, gy, disp, conf, reason,  \
                reliable, disp_wta, best, second, sum32);                                                      \
        hipLaunchKernelGGL((wta_fused_kernel<TC, NP, WD, (WD ? 2 : 1), MUL>), grid, dim3(64), shm, s, p, w, h, D, \
                           rw, (const TC*)cost, planes, dmin, dmax, dr, gx, gy, disp, conf, reason, reliable,     \
                           disp_wta, best, second, sum32);                                                     \
    } while (0)
    if (planes.paired) {
        // paired families keep one plane (u8 C' only): NP planes for MUL paths
        if (!cost_u8) return hipErrorInvalidValue;
        const int np = (p.paths == 8 ? 8 : 4) - __builtin_popcount((unsigned)planes.paired);
        if (p.paths == 8) {
            if (np == 5) {
                if (wide) APG_WTAF(uint8_t, 5, true, 8);
                else APG_WTAF(uint8_t, 5, false, 8);
            } else if (np == 4) {
                if (wide) APG_WTAF(uint8_t, 4, true, 8);
                else APG_WTAF(uint8_t, 4, false, 8);
            } else {
                return hipErrorInvalidValue;
            }
        } else {
            if (np == 3) {
                if (wide) APG_WTAF(uint8_t, 3, true, 4);
                else APG_WTAF(uint8_t, 3, false, 4);
            } else if (np == 2) {
                if (wide) APG_WTAF(uint8_t, 2, true, 4);
                else APG_WTAF(uint8_t, 2, false, 4);
            } else {
                return hipErrorInvalidValue;
            }
        }
    } else if (p.paths == 8) {
        if (wide) APG_WTAF(uint8_t, 8, true, 8);
        else if (cost_u8) APG_WTAF(uint8_t, 8, false, 8);
        else APG_WTAF(uint16_t, 8, false, 8);
    } else {
        if (wide) APG_WTAF(uint8_t, 4, true, 4);
        else if (cost_u8) APG_WTAF(uint8_t, 4, false, 4);
        else APG_WTAF(uint16_t, 4, false, 4);
    }
#undef APG_WTAF
    return hipGetLastError();
}

}  // namespace apgk
