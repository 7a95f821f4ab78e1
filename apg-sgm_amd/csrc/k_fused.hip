// gfx950 kernels of the fused (non-materialising) schedule:
//   cost_crossh_kernel : matching cost + right-view WTA candidates + Cross H
//                        pass in one pass over a row segment (the raw cost
//                        volume never reaches HBM);
//   rwta_finalize_kernel: right-view disparity with the legacy subpixel step
//                        from the per-pixel winning keys.
#include "device_common.hpp"
#include "kernels.hpp"

namespace apgk {

constexpr int F_TX = 64;   // output pixels per block

struct FusedShm {
    // byte offsets into the dynamic LDS block
    size_t lut, sl, slo, shi, sr, key, magic, P, total;
};

__host__ __device__ inline FusedShm fused_layout(int D, int Lmax) {
    FusedShm s;
    const int NC = F_TX + 2 * Lmax;
    size_t o = 0;
    s.lut = o; o += (65 + 256 + 511) * 4;
    s.sl = o; o += (size_t)NC * sizeof(PixDesc);
    s.slo = o; o += (size_t)NC * 4;
    s.shi = o; o += (size_t)NC * 4;
    s.sr = o; o += (size_t)(NC + D) * sizeof(PixDesc);
    s.key = o; o += (size_t)(F_TX + D) * 4;
    s.magic = o; o += (2 * kMaxArm + 2) * 4;
    o = (o + 15) & ~(size_t)15;
    s.P = o; o += (size_t)(NC + 1) * 64 * 4;
    s.total = o;
    return s;
}

__device__ __forceinline__ uint32_t magic_div(uint32_t c) {
    return c <= 1 ? 0u : (uint32_t)((0x100000000ull + c - 1) / c);
}

// Matching cost (reference src/cost_computer.cpp:139-203) for the columns of a
// row segment plus the Cross arm halo, the right-view WTA candidates of the
// segment's own columns (reference src/pipeline.cpp:24-52: for right pixel
// xr the minimum of C(xr+d, d), earliest d, as the key cost << 16 | d-index,
// combined with atomicMin first in LDS, then once per xr in HBM) and the
// Cross H pass (reference src/cost_aggregator.cpp:74-105) through packed
// (sum << 8 | count) prefix sums.  Writes T (u16) when CROSS, else the cost
// itself (Fast mode).
template <bool SYM, bool CROSS, typename TO>
__global__ __launch_bounds__(256) void cost_crossh_kernel(
    const CostLut* __restrict__ lut, int w, int h, int d0, int d1, int cost_max, int use_ad, int use_grad, int Lmax,
    const uint8_t* __restrict__ gl, const uint8_t* __restrict__ gxl, const uint8_t* __restrict__ gyl,
    const uint8_t* __restrict__ gr, const uint8_t* __restrict__ gxr, const uint8_t* __restrict__ gyr,
    const uint32_t* __restrict__ c32l, const uint32_t* __restrict__ c32r, const uint64_t* __restrict__ c64l,
    const uint64_t* __restrict__ c64r, const int16_t* __restrict__ dmin, const int16_t* __restrict__ dmax,
    const uint32_t* __restrict__ arms, TO* __restrict__ out, uint32_t* __restrict__ rkeys) {
    constexpr TO OINV = (TO)~(TO)0;
    extern __shared__ __align__(16) unsigned char shm[];
    const int D = d1 - d0;
    const FusedShm L = fused_layout(D, Lmax);
    float* s_lc = (float*)(shm + L.lut);
    float* s_la = s_lc + 65;
    float* s_lg = s_la + 256;
    PixDesc* s_l = (PixDesc*)(shm + L.sl);
    int* s_lo = (int*)(shm + L.slo);
    int* s_hi = (int*)(shm + L.shi);
    PixDesc* s_r = (PixDesc*)(shm + L.sr);
    uint32_t* s_key = (uint32_t*)(shm + L.key);
    uint32_t* magic = (uint32_t*)(shm + L.magic);
    uint32_t* P = (uint32_t*)(shm + L.P);

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int y = blockIdx.y;
    const int x0 = blockIdx.x * F_TX;
    const int NC = F_TX + 2 * Lmax;
    const int xb = x0 - Lmax;                 // column 0 of the segment
    const int xs = xb - (d1 - 1);             // right pixel of s_r[0]
    const int kbase = x0 - (d1 - 1);          // right pixel of s_key[0]
    const size_t row = (size_t)y * w;

    for (int i = tid; i < 65; i += 256) s_lc[i] = lut->census[i];
    for (int i = tid; i < 256; i += 256) s_la[i] = lut->ad[i];
    for (int i = tid; i < 511; i += 256) s_lg[i] = lut->grad[i];
    for (int i = tid; i < 2 * kMaxArm + 2; i += 256) magic[i] = magic_div((uint32_t)i);
    for (int i = tid; i < F_TX + D - 1; i += 256) s_key[i] = 0xFFFFFFFFu;
    for (int i = tid; i < NC; i += 256) {
        const int x = xb + i;
        PixDesc p{0u, 0u, 0u};
        int lo = 0, hi = 0;
        if (x >= 0 && x < w) {
            const size_t q = row + x;
            if (SYM) p.clo = c32l[q];
            else { const uint64_t c = c64l[q]; p.clo = (uint32_t)c; p.chi = (uint32_t)(c >> 32); }
            p.ggg = (uint32_t)gl[q] | ((uint32_t)gxl[q] << 8) | ((uint32_t)gyl[q] << 16);
            lo = dmin[q]; hi = dmax[q];
        }
        s_l[i] = p; s_lo[i] = lo; s_hi[i] = hi;
    }
    for (int i = tid; i < NC + D - 1; i += 256) {
        const int xr = xs + i;
        PixDesc p{0u, 0u, 1u << 24};
        if (xr >= 0 && xr < w) {
            const size_t q = row + xr;
            if (SYM) p.clo = c32r[q];
            else { const uint64_t c = c64r[q]; p.clo = (uint32_t)c; p.chi = (uint32_t)(c >> 32); }
            p.ggg = (uint32_t)gr[q] | ((uint32_t)gxr[q] << 8) | ((uint32_t)gyr[q] << 16);
        }
        s_r[i] = p;
    }
    __syncthreads();
    const float scale = lut->scale;
    const int seg = (NC + 3) / 4;
    const int c_beg = min(NC, wave * seg), c_end = min(NC, c_beg + seg);

    // this wave's column segment cached one column per lane (seg <= 64), read
    // back with readlane so the per-column values are scalars in the loop
    const int cl = min(c_beg + lane, NC - 1);
    const uint32_t col_clo = s_l[cl].clo, col_chi = s_l[cl].chi, col_ggg = s_l[cl].ggg;
    const uint32_t col_rng = (uint32_t)(s_lo[cl] & 0xFFFF) | ((uint32_t)s_hi[cl] << 16);
    for (int dc = 0; dc < D; dc += 64) {
        const int di = dc + lane;
        const int d = d0 + di;
        const bool dact = di < D;
        // phase A: costs of this wave's column segment, local prefix, WTA keys
        uint32_t acc = 0;
        for (int c = c_beg; c < c_end; ++c) {
            const int i = c - c_beg;
            const int x = xb + c;
            const int xr = x - d;
            PixDesc Lp;
            Lp.clo = (uint32_t)__builtin_amdgcn_readlane((int)col_clo, i);
            Lp.chi = SYM ? 0u : (uint32_t)__builtin_amdgcn_readlane((int)col_chi, i);
            Lp.ggg = (uint32_t)__builtin_amdgcn_readlane((int)col_ggg, i);
            const uint32_t rng = (uint32_t)__builtin_amdgcn_readlane((int)col_rng, i);
            const int lo = (int16_t)(rng & 0xFFFF), hi = (int16_t)(rng >> 16);
            const bool xin = x >= 0 && x < w;   // uniform
            const bool valid = dact && xin && d >= lo && d < hi && xr >= 0 && xr < w;
            const int ri = min(max(c + (d1 - 1 - d), 0), NC + D - 2);
            const uint16_t cost = state_cost<SYM>(Lp, s_r[ri], s_lc, s_la, s_lg, scale, cost_max, use_ad, use_grad);
            // a cost equal to the sentinel (cost_max 65535) is invalid downstream
            const bool ok = valid && cost != kInv16;
            const bool own = c >= Lmax && c < Lmax + F_TX;   // uniform
            if (own && ok) atomicMin(&s_key[xr - kbase], ((uint32_t)cost << 16) | (uint32_t)di);
            if (!CROSS && own && xin && dact) out[(row + x) * D + di] = ok ? (TO)cost : OINV;
            if (CROSS) {
                acc += ok ? ((uint32_t)cost << 8) | 1u : 0u;
                P[(c + 1) * 64 + lane] = acc;
            }
        }
        if (CROSS) {
            if (wave == 0) P[lane] = 0;
            __syncthreads();
            uint32_t off = 0;
            for (int ww = 0; ww < wave; ++ww) {
                const int e = min(NC, (ww + 1) * seg);
                if (e > ww * seg) off += P[e * 64 + lane];
            }
            __syncthreads();
            if (off)
                for (int c = c_beg; c < c_end; ++c) P[(c + 1) * 64 + lane] += off;
            __syncthreads();
            // phase B: windowed means of the segment's own pixels
            if (dact) {
                for (int px = wave; px < F_TX; px += 4) {
                    const int x = x0 + px;
                    if (x >= w) break;
                    const int c = px + Lmax;
                    const uint32_t a = arms[row + x];
                    const int La = a & 0xFF, Ra = (a >> 8) & 0xFF;
                    TO o = OINV;
                    if (d >= s_lo[c] && d < s_hi[c]) {
                        const uint32_t s = P[(c + Ra + 1) * 64 + lane] - P[(c - La) * 64 + lane];
                        const uint32_t cnt = s & 0xFF;
                        if (cnt) o = (TO)(cnt == 1 ? (s >> 8) : __umulhi(s >> 8, magic[cnt]));
                    }
                    out[(row + x) * D + di] = o;
                }
            }
            __syncthreads();
        }
    }
    __syncthreads();
    for (int i = tid; i < F_TX + D - 1; i += 256) {
        const int xr = kbase + i;
        const uint32_t k = s_key[i];
        if (xr >= 0 && xr < w && k != 0xFFFFFFFFu) atomicMin(&rkeys[row + xr], k);
    }
}

hipError_t launch_cost_crossh(hipStream_t s, const apg_hip_params& p, const CostLut* d_lut, int w, int h,
                              const uint8_t* gl, const uint8_t* gxl, const uint8_t* gyl, const uint8_t* gr,
                              const uint8_t* gxr, const uint8_t* gyr, const uint32_t* c32l, const uint32_t* c32r,
                              const uint64_t* c64l, const uint64_t* c64r, const int16_t* dmin, const int16_t* dmax,
                              const uint32_t* arms, void* out, bool out_u8, uint32_t* rkeys) {
    const int D = p.max_disparity - p.min_disparity;
    const bool cross = p.agg_enable != 0;
    const int Lmax = cross ? (p.max_arm_length > 1 ? p.max_arm_length : 1) : 0;
    const FusedShm L = fused_layout(D, Lmax);
    dim3 grid((w + F_TX - 1) / F_TX, h);
#define APG_LAUNCH_F(S, X, T)                                                                                      \
    do {                                                                                                           \
        auto k = cost_crossh_kernel<S, X, T>;                                                                      \
        if (L.total > 64 * 1024) {                                                                                 \
            hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize,         \
                                               (int)L.total);                                                      \
            if (e != hipSuccess) return e;                                                                         \
        }                                                                                                          \
        hipLaunchKernelGGL(k, grid, dim3(256), L.total, s, d_lut, w, h, p.min_disparity, p.max_disparity,         \
                           p.cost_max, p.use_ad, p.use_grad, Lmax, gl, gxl, gyl, gr, gxr, gyr, c32l, c32r, c64l,   \
                           c64r, dmin, dmax, arms, (T*)out, rkeys);                                               \
    } while (0)
    if (out_u8) {
        if (p.census_type == 1) {
            if (cross) APG_LAUNCH_F(true, true, uint8_t);
            else APG_LAUNCH_F(true, false, uint8_t);
        } else {
            if (cross) APG_LAUNCH_F(false, true, uint8_t);
            else APG_LAUNCH_F(false, false, uint8_t);
        }
    } else {
        if (p.census_type == 1) {
            if (cross) APG_LAUNCH_F(true, true, uint16_t);
            else APG_LAUNCH_F(true, false, uint16_t);
        } else {
            if (cross) APG_LAUNCH_F(false, true, uint16_t);
            else APG_LAUNCH_F(false, false, uint16_t);
        }
    }
#undef APG_LAUNCH_F
    return hipGetLastError();
}

// Right-view disparity from the winning keys (reference src/pipeline.cpp:53-82):
// best d plus the legacy parabola over C(xr+d-1, d-1), C(xr+d, d),
// C(xr+d+1, d+1), recomputed from the pixel descriptors.
template <bool SYM>
__global__ void rwta_finalize_kernel(const CostLut* __restrict__ lut, int w, int h, int d0, int D, int cost_max,
                                     int use_ad, int use_grad, int subpixel, const uint8_t* __restrict__ gl,
                                     const uint8_t* __restrict__ gxl, const uint8_t* __restrict__ gyl,
                                     const uint8_t* __restrict__ gr, const uint8_t* __restrict__ gxr,
                                     const uint8_t* __restrict__ gyr, const uint32_t* __restrict__ c32l,
                                     const uint32_t* __restrict__ c32r, const uint64_t* __restrict__ c64l,
                                     const uint64_t* __restrict__ c64r, const int16_t* __restrict__ dmin,
                                     const int16_t* __restrict__ dmax, const uint32_t* __restrict__ rkeys,
                                     float* __restrict__ dr) {
    const int xr = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (xr >= w) return;
    const size_t row = (size_t)y * w;
    const uint32_t k = rkeys[row + xr];
    float dval = -1.f;
    // as in the reference, a negative best disparity counts as "no match" (best_d >= 0 test)
    if (k != 0xFFFFFFFFu && d0 + (int)(k & 0xFFFF) >= 0) {
        const int di = (int)(k & 0xFFFF);
        const int bd = d0 + di;
        dval = (float)bd;
        const int xm = xr + bd - 1, xp = xr + bd + 1;
        if (subpixel && xm >= 0 && xm < w && xp >= 0 && xp < w && di > 0 && di + 1 < D) {
            auto desc = [&](const uint8_t* g, const uint8_t* gx, const uint8_t* gy, const uint32_t* c32,
                            const uint64_t* c64, int x) {
                const size_t q = row + x;
                PixDesc p;
                if (SYM) { p.clo = c32[q]; p.chi = 0; }
                else { const uint64_t c = c64[q]; p.clo = (uint32_t)c; p.chi = (uint32_t)(c >> 32); }
                p.ggg = (uint32_t)g[q] | ((uint32_t)gx[q] << 8) | ((uint32_t)gy[q] << 16);
                return p;
            };
            const PixDesc R = desc(gr, gxr, gyr, c32r, c64r, xr);
            // C(xl, d) is valid iff d lies in xl's range (xr = xl - d is inside the image here)
            const bool vm = (bd - 1) >= dmin[row + xm] && (bd - 1) < dmax[row + xm];
            const bool vp = (bd + 1) >= dmin[row + xp] && (bd + 1) < dmax[row + xp];
            if (vm && vp) {
                const uint16_t cm = state_cost<SYM>(desc(gl, gxl, gyl, c32l, c64l, xm), R, lut->census, lut->ad,
                                                    lut->grad, lut->scale, cost_max, use_ad, use_grad);
                const uint16_t cp = state_cost<SYM>(desc(gl, gxl, gyl, c32l, c64l, xp), R, lut->census, lut->ad,
                                                    lut->grad, lut->scale, cost_max, use_ad, use_grad);
                const uint16_t c0 = (uint16_t)(k >> 16);
                if (cm != kInv16 && cp != kInv16 && c0 != kInv16)
                    dval += legacy_delta((float)cm, (float)c0, (float)cp);
            }
        }
    }
    dr[row + xr] = dval;
}

hipError_t launch_rwta_finalize(hipStream_t s, const apg_hip_params& p, const CostLut* d_lut, int w, int h,
                                const uint8_t* gl, const uint8_t* gxl, const uint8_t* gyl, const uint8_t* gr,
                                const uint8_t* gxr, const uint8_t* gyr, const uint32_t* c32l, const uint32_t* c32r,
                                const uint64_t* c64l, const uint64_t* c64r, const int16_t* dmin, const int16_t* dmax,
                                const uint32_t* rkeys, float* dr) {
    const int D = p.max_disparity - p.min_disparity;
    dim3 grid((w + 127) / 128, h);
    if (p.census_type == 1)
        hipLaunchKernelGGL(rwta_finalize_kernel<true>, grid, dim3(128), 0, s, d_lut, w, h, p.min_disparity, D,
                           p.cost_max, p.use_ad, p.use_grad, p.post_subpixel, gl, gxl, gyl, gr, gxr, gyr, c32l, c32r,
                           c64l, c64r, dmin, dmax, rkeys, dr);
    else
        hipLaunchKernelGGL(rwta_finalize_kernel<false>, grid, dim3(128), 0, s, d_lut, w, h, p.min_disparity, D,
                           p.cost_max, p.use_ad, p.use_grad, p.post_subpixel, gl, gxl, gyl, gr, gxr, gyr, c32l, c32r,
                           c64l, c64r, dmin, dmax, rkeys, dr);
    return hipGetLastError();
}

// Fill of a u32 buffer (graph-capture safe replacement of a memset).
__global__ void fill_u32_kernel(size_t n, uint32_t v, uint32_t* __restrict__ p) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_fill_u32(hipStream_t s, size_t n, uint32_t v, uint32_t* p) {
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, v, p);
    return hipGetLastError();
}

}  // namespace apgk
