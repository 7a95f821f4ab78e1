// gfx950 kernels: gray/gradient/census (aux), search-range init, the dense
// u16 cost volume, right-view WTA on the raw volume and Cross arms.
#include <math.h>

#include "device_common.hpp"
#include "kernels.hpp"

namespace apgk {

// ------------------------------------------------------------------ LUT ----
// DiscreteCostLut (reference src/discrete_cost_lut.hpp:10-38) and the scale of
// CostComputer::compute_volume (reference src/cost_computer.cpp:149-155).
// Built with the host libm expf; the per-state expression of the dense
// reference (:176-198) evaluates exactly these float values, so the device
// only repeats the two float adds, the multiply and the +0.5 in order.
void build_cost_lut(const apg_hip_params& p, CostLut* out) {
    const float lc = p.lambda_census > 1e-3f ? p.lambda_census : 1e-3f;
    const float la = p.lambda_ad > 1e-3f ? p.lambda_ad : 1e-3f;
    const float lg = p.lambda_grad > 1e-3f ? p.lambda_grad : 1e-3f;
    for (int i = 0; i <= 64; ++i) out->census[i] = 1.f - expf(-(float)i / lc);
    for (int i = 0; i < 256; ++i) out->ad[i] = p.use_ad ? p.eta_ad * (1.f - expf(-(float)i / la)) : 0.f;
    for (int i = 0; i < 511; ++i) out->grad[i] = p.use_grad ? p.mu_grad * (1.f - expf(-(float)i / lg)) : 0.f;
    const float norm = 1.f + (p.use_ad ? p.eta_ad : 0.f) + (p.use_grad ? p.mu_grad : 0.f);
    out->scale = (float)p.cost_max / (norm > 1e-6f ? norm : 1e-6f);
    out->pad[0] = 0;
}

// Upper bound of every quantised cost: float addition and the final
// quantisation are monotone, so the sum of the three LUT maxima bounds c.
int cost_upper_bound(const CostLut& lut, int cost_max) {
    float mc = 0.f, ma = 0.f, mg = 0.f;
    for (float v : lut.census) mc = v > mc ? v : mc;
    for (float v : lut.ad) ma = v > ma ? v : ma;
    for (float v : lut.grad) mg = v > mg ? v : mg;
    const float c = mc + ma + mg;
    const int q = (int)(c * lut.scale + 0.5f);
    return q < 0 ? 0 : (q > cost_max ? cost_max : q);
}

bool cost_clamp_free(const CostLut& lut, int cost_max) {
    float mc = 0.f, ma = 0.f, mg = 0.f;
    for (float v : lut.census) {
        if (!(v >= 0.f)) return false;
        mc = v > mc ? v : mc;
    }
    for (float v : lut.ad) {
        if (!(v >= 0.f)) return false;
        ma = v > ma ? v : ma;
    }
    for (float v : lut.grad) {
        if (!(v >= 0.f)) return false;
        mg = v > mg ? v : mg;
    }
    if (!(lut.scale >= 0.f)) return false;
    const float c = (mc + ma) + mg;
    const float q = c * lut.scale + 0.5f;
    return q < 2147483520.f && (int)q <= cost_max;
}

// ------------------------------------------------------------------ aux ----
constexpr int AUX_TX = 64, AUX_TY = 4;

// build_gray_and_grad + census (reference src/cost_computer.cpp:33-122,
// include/apg_sgm/image_view.hpp:50-56). The LDS tile is loaded with
// clamp-to-edge coordinates, which is exactly Image8::sample.
// One launch can serve both views (blockIdx.z selects the image); the left
// view may also write the dense range init [min_disparity, min(max_disparity,
// x+1)) (range_kernel below) and the right view clear the right-WTA keys.
struct AuxImage {
    const uint8_t* src;
    uint8_t *gray, *gx, *gy;
    uint32_t* c32;
    uint64_t* c64;
    uint32_t* keys;
    int16_t *dmin, *dmax;
};
struct AuxPair {
    AuxImage img[2];
};

__device__ __forceinline__ void aux_pixel(const AuxImage& im, int w, int h, int stride, int ch, int census_type,
                                          int gmin, int gmax, uint8_t (&tile)[AUX_TY + 6][AUX_TX + 8]);

__global__ __launch_bounds__(256) void aux_pair_kernel(AuxPair ap, int w, int h, int stride, int ch, int census_type,
                                                       int gmin, int gmax) {
    __shared__ uint8_t tile[AUX_TY + 6][AUX_TX + 8];
    aux_pixel(ap.img[blockIdx.z], w, h, stride, ch, census_type, gmin, gmax, tile);
}

__global__ __launch_bounds__(256) void aux_kernel(const uint8_t* __restrict__ src, int w, int h, int stride, int ch,
                                                 int census_type, uint8_t* __restrict__ gray,
                                                 uint8_t* __restrict__ gx, uint8_t* __restrict__ gy,
                                                 uint32_t* __restrict__ c32, uint64_t* __restrict__ c64,
                                                 uint32_t* __restrict__ keys) {
    __shared__ uint8_t tile[AUX_TY + 6][AUX_TX + 8];
    aux_pixel(AuxImage{src, gray, gx, gy, c32, c64, keys, nullptr, nullptr}, w, h, stride, ch, census_type, 0, 0, tile);
}

__device__ __forceinline__ void aux_pixel(const AuxImage& im, int w, int h, int stride, int ch, int census_type,
                                          int gmin, int gmax, uint8_t (&tile)[AUX_TY + 6][AUX_TX + 8]) {
    const uint8_t* __restrict__ src = im.src;
    uint8_t* __restrict__ gray = im.gray;
    uint8_t* __restrict__ gx = im.gx;
    uint8_t* __restrict__ gy = im.gy;
    uint32_t* __restrict__ c32 = im.c32;
    uint64_t* __restrict__ c64 = im.c64;
    uint32_t* __restrict__ keys = im.keys;
    const int x0 = blockIdx.x * AUX_TX, y0 = blockIdx.y * AUX_TY;
    const int tid = threadIdx.y * AUX_TX + threadIdx.x;
    for (int i = tid; i < (AUX_TY + 6) * (AUX_TX + 8); i += 256) {
        const int ty = i / (AUX_TX + 8), tx = i % (AUX_TX + 8);
        const int sx = clampi(x0 - 4 + tx, 0, w - 1), sy = clampi(y0 - 3 + ty, 0, h - 1);
        const uint8_t* px = src + (size_t)sy * stride + (size_t)sx * ch;
        uint8_t v;
        if (ch == 1) v = px[0];
        else v = (uint8_t)((77 * (int)px[0] + 150 * (int)px[1] + 29 * (int)px[2]) >> 8);
        tile[ty][tx] = v;
    }
    __syncthreads();
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    if (x >= w || y >= h) return;
    const int cx = threadIdx.x + 4, cy = threadIdx.y + 3;
    const size_t pi = (size_t)y * w + x;
    gray[pi] = tile[cy][cx];
    const int gxv = (int)tile[cy][cx + 1] - (int)tile[cy][cx - 1];
    const int gyv = (int)tile[cy + 1][cx] - (int)tile[cy - 1][cx];
    gx[pi] = (uint8_t)min(abs(gxv), 255);
    gy[pi] = (uint8_t)min(abs(gyv), 255);
    // the fused schedule's right-view WTA key, cleared here to save a fill pass
    if (keys) keys[pi] = 0xFFFFFFFFu;
    if (im.dmin) {   // as range_kernel
        im.dmin[pi] = (int16_t)gmin;
        im.dmax[pi] = (int16_t)max(min(gmax, x + 1), gmin);
    }
    if (census_type == 1) {
        uint32_t bits = 0;
        int b = 0;
#pragma unroll
        for (int dy = -3; dy <= 3; ++dy)
#pragma unroll
            for (int dx = -4; dx <= 4; ++dx)
                if (dy < 0 || (dy == 0 && dx < 0)) {
                    bits |= (uint32_t)(tile[cy + dy][cx + dx] > tile[cy - dy][cx - dx]) << b;
                    ++b;
                }
        c32[pi] = bits;
    } else {
        const uint8_t c = tile[cy][cx];
        uint64_t bits = 0;
        int b = 0;
#pragma unroll
        for (int dy = -3; dy <= 3; ++dy)
#pragma unroll
            for (int dx = -4; dx <= 4; ++dx) {
                if (dx == 0 && dy == 0) continue;
                bits |= (uint64_t)(tile[cy + dy][cx + dx] > c) << b;
                ++b;
            }
        c64[pi] = bits;
    }
}

hipError_t launch_aux_pair(hipStream_t s, const apg_hip_params& p, int w, int h, int stride, int ch,
                           const AuxViews& L, const AuxViews& R, uint32_t* rkeys, int16_t* dmin, int16_t* dmax) {
    AuxPair ap;
    ap.img[0] = AuxImage{L.src, L.gray, L.gx, L.gy, L.c32, L.c64, nullptr, dmin, dmax};
    ap.img[1] = AuxImage{R.src, R.gray, R.gx, R.gy, R.c32, R.c64, rkeys, nullptr, nullptr};
    dim3 grid((w + AUX_TX - 1) / AUX_TX, (h + AUX_TY - 1) / AUX_TY, 2), block(AUX_TX, AUX_TY);
    hipLaunchKernelGGL(aux_pair_kernel, grid, block, 0, s, ap, w, h, stride, ch, p.census_type, p.min_disparity,
                       p.max_disparity);
    return hipGetLastError();
}

hipError_t launch_aux(hipStream_t s, const apg_hip_params& p, const uint8_t* src, int w, int h, int stride,
                      int ch, uint8_t* gray, uint8_t* gx, uint8_t* gy, uint32_t* c32, uint64_t* c64,
                      uint32_t* keys) {
    dim3 grid((w + AUX_TX - 1) / AUX_TX, (h + AUX_TY - 1) / AUX_TY), block(AUX_TX, AUX_TY);
    hipLaunchKernelGGL(aux_kernel, grid, block, 0, s, src, w, h, stride, ch, p.census_type, gray, gx, gy, c32, c64,
                       keys);
    return hipGetLastError();
}

// ---------------------------------------------------------------- range ----
// PriorEstimator::estimate, prior off (reference src/prior_estimator.cpp:414-424)
// through SearchRange::set_pixel (include/apg_sgm/search_range.hpp:29-35).
__global__ void range_kernel(int w, int h, int gmin, int gmax, int16_t* dmin, int16_t* dmax) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    int lo = max(gmin, gmin);
    int hi = min(min(gmax, x + 1), gmax);
    if (hi < lo) hi = lo;
    dmin[(size_t)y * w + x] = (int16_t)lo;
    dmax[(size_t)y * w + x] = (int16_t)hi;
}

hipError_t launch_range_init(hipStream_t s, const apg_hip_params& p, int w, int h, int16_t* dmin, int16_t* dmax) {
    hipLaunchKernelGGL(range_kernel, dim3((w + 255) / 256, h), dim3(256), 0, s, w, h, p.min_disparity,
                       p.max_disparity, dmin, dmax);
    return hipGetLastError();
}

// ----------------------------------------------------------------- cost ----
constexpr int COST_TX = 32;

struct PixStage {
    uint32_t clo, chi;   // census (chi only for the 64-bit census)
    uint32_t ggg;        // gray | gx << 8 | gy << 16 | (out of image) << 24
};

// CostComputer::compute_volume (reference src/cost_computer.cpp:139-203).
// Block = one row segment of COST_TX pixels x all D; left and right pixel
// descriptors for the segment (right span x0-(d1-1) .. x0+TX-1-d0) and the
// LUTs are staged in LDS. Wave w handles pixels w, w+4, ...; lanes run over d,
// so every wave store is a contiguous run of the pixel's d-vector.
template <bool SYM>
__global__ __launch_bounds__(256) void cost_kernel(const CostLut* __restrict__ lut, int w, int h, int d0, int d1,
                                                   int cost_max, int use_ad, int use_grad,
                                                   const uint8_t* __restrict__ gl, const uint8_t* __restrict__ gxl,
                                                   const uint8_t* __restrict__ gyl, const uint8_t* __restrict__ gr,
                                                   const uint8_t* __restrict__ gxr, const uint8_t* __restrict__ gyr,
                                                   const uint32_t* __restrict__ c32l, const uint32_t* __restrict__ c32r,
                                                   const uint64_t* __restrict__ c64l, const uint64_t* __restrict__ c64r,
                                                   const int16_t* __restrict__ dmin, const int16_t* __restrict__ dmax,
                                                   uint16_t* __restrict__ cost) {
    __shared__ float s_lc[65], s_la[256], s_lg[511];
    __shared__ PixStage s_r[COST_TX + kMaxD];
    __shared__ PixStage s_l[COST_TX];
    __shared__ int s_lo[COST_TX], s_hi[COST_TX];
    const int tid = threadIdx.x;
    const int D = d1 - d0;
    const int x0 = blockIdx.x * COST_TX, y = blockIdx.y;
    for (int i = tid; i < 65; i += 256) s_lc[i] = lut->census[i];
    for (int i = tid; i < 256; i += 256) s_la[i] = lut->ad[i];
    for (int i = tid; i < 511; i += 256) s_lg[i] = lut->grad[i];
    const int xs = x0 - (d1 - 1);
    const int nr = COST_TX + D - 1;
    for (int i = tid; i < nr; i += 256) {
        const int xr = xs + i;
        PixStage ps{0u, 0u, 1u << 24};
        if (xr >= 0 && xr < w) {
            const size_t q = (size_t)y * w + xr;
            if (SYM) ps.clo = c32r[q];
            else { const uint64_t c = c64r[q]; ps.clo = (uint32_t)c; ps.chi = (uint32_t)(c >> 32); }
            ps.ggg = (uint32_t)gr[q] | ((uint32_t)gxr[q] << 8) | ((uint32_t)gyr[q] << 16);
        }
        s_r[i] = ps;
    }
    for (int i = tid; i < COST_TX; i += 256) {
        const int x = x0 + i;
        PixStage ps{0u, 0u, 0u};
        int lo = 0, hi = 0;
        if (x < w) {
            const size_t q = (size_t)y * w + x;
            if (SYM) ps.clo = c32l[q];
            else { const uint64_t c = c64l[q]; ps.clo = (uint32_t)c; ps.chi = (uint32_t)(c >> 32); }
            ps.ggg = (uint32_t)gl[q] | ((uint32_t)gxl[q] << 8) | ((uint32_t)gyl[q] << 16);
            lo = dmin[q]; hi = dmax[q];
        }
        s_l[i] = ps; s_lo[i] = lo; s_hi[i] = hi;
    }
    __syncthreads();
    const float scale = lut->scale;
    const int lane = tid & 63, wave = tid >> 6;
    for (int px = wave; px < COST_TX; px += 4) {
        const int x = x0 + px;
        if (x >= w) break;
        const PixStage L = s_l[px];
        const int lo = s_lo[px], hi = s_hi[px];
        uint16_t* out = cost + ((size_t)y * w + x) * D;
        for (int di = lane; di < D; di += 64) {
            const int d = d0 + di;
            // right pixel x - d sits at s_r[(x - d) - xs] = s_r[px + (d1 - 1 - d)]
            const PixStage R = s_r[px + (d1 - 1 - d)];
            uint16_t v = kInv16;
            if (d >= lo && d < hi && !(R.ggg >> 24)) {
                const int pop = SYM ? __popc(L.clo ^ R.clo) : (__popc(L.clo ^ R.clo) + __popc(L.chi ^ R.chi));
                float c = s_lc[pop];
                if (use_ad) {
                    const int ad = abs((int)(L.ggg & 0xFF) - (int)(R.ggg & 0xFF));
                    c = c + s_la[ad];
                }
                if (use_grad) {
                    const int g = abs((int)((L.ggg >> 8) & 0xFF) - (int)((R.ggg >> 8) & 0xFF)) +
                                  abs((int)((L.ggg >> 16) & 0xFF) - (int)((R.ggg >> 16) & 0xFF));
                    c = c + s_lg[g];
                }
                const int q = (int)(c * scale + 0.5f);
                v = (uint16_t)clampi(q, 0, cost_max);
            }
            out[di] = v;
        }
    }
}

hipError_t launch_cost(hipStream_t s, const apg_hip_params& p, const CostLut* d_lut, int w, int h,
                       const uint8_t* gl, const uint8_t* gxl, const uint8_t* gyl,
                       const uint8_t* gr, const uint8_t* gxr, const uint8_t* gyr,
                       const uint32_t* c32l, const uint32_t* c32r, const uint64_t* c64l, const uint64_t* c64r,
                       const int16_t* dmin, const int16_t* dmax, uint16_t* cost) {
    dim3 grid((w + COST_TX - 1) / COST_TX, h);
    if (p.census_type == 1)
        hipLaunchKernelGGL(cost_kernel<true>, grid, dim3(256), 0, s, d_lut, w, h, p.min_disparity, p.max_disparity,
                           p.cost_max, p.use_ad, p.use_grad, gl, gxl, gyl, gr, gxr, gyr, c32l, c32r, c64l, c64r,
                           dmin, dmax, cost);
    else
        hipLaunchKernelGGL(cost_kernel<false>, grid, dim3(256), 0, s, d_lut, w, h, p.min_disparity, p.max_disparity,
                           p.cost_max, p.use_ad, p.use_grad, gl, gxl, gyl, gr, gxr, gyr, c32l, c32r, c64l, c64r,
                           dmin, dmax, cost);
    return hipGetLastError();
}

// ------------------------------------------------------------ right WTA ----
constexpr int RW_TX = 64, RW_DC = 32;

// wta_right_from_left_volume (reference src/pipeline.cpp:24-85): for right
// pixel xr, earliest argmin over d of C(xr + d, d) on the raw volume.
// Block = 4 rows x 64 right pixels; the diagonal band of each d-chunk is staged
// in LDS so the volume is read with row-contiguous loads.
__global__ __launch_bounds__(256) void right_wta_kernel(int w, int h, int d0, int D, int subpixel,
                                                        const uint16_t* __restrict__ cost, float* __restrict__ dr) {
    __shared__ uint16_t band[4][RW_TX + RW_DC][RW_DC + 1];
    const int lane = threadIdx.x, row = threadIdx.y;
    const int y = blockIdx.y * 4 + row;
    const int x0 = blockIdx.x * RW_TX;
    const int xr = x0 + lane;
    uint32_t best = 0xFFFFFFFFu;
    int best_d = -1;
    for (int dc = 0; dc < D; dc += RW_DC) {
        if (y < h) {
            // band[i][j] = C(x0 + d0 + dc + i - j + j ..) : pixel xl = x0 + d0 + dc + i, disparity index dc + j
            // 8 loads in flight per lane (clamped inside the row and the
            // volume, masked after), so a batch waits once
            constexpr int NK = (RW_TX + RW_DC) * RW_DC;
            for (int k0 = lane; k0 < NK; k0 += 64 * 8) {
                uint16_t v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int k = min(k0 + 64 * u, NK - 1);
                    const int i = k / RW_DC, j = k % RW_DC;
                    const int xl = min(max(x0 + d0 + dc + i, 0), w - 1);
                    v[u] = cost[((size_t)y * w + xl) * D + min(dc + j, D - 1)];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int k = k0 + 64 * u;
                    if (k >= NK) break;
                    const int i = k / RW_DC, j = k % RW_DC;
                    const int xl = x0 + d0 + dc + i;
                    band[row][i][j] = (xl >= 0 && xl < w && dc + j < D) ? v[u] : kInv16;
                }
            }
        }
        __syncthreads();
        if (y < h && xr < w) {
            const int jn = min(RW_DC, D - dc);
            for (int j = 0; j < jn; ++j) {
                // xl = xr + d = x0 + lane + d0 + dc + j -> i = lane + j
                const uint16_t c = band[row][lane + j][j];
                if (c != kInv16 && c < best) { best = c; best_d = d0 + dc + j; }
            }
        }
        __syncthreads();
    }
    if (y >= h || xr >= w) return;
    float dval = -1.f;
    if (best_d >= 0) {
        dval = (float)best_d;
        if (subpixel) {
            const int di = best_d - d0;
            const int xm = xr + best_d - 1, xc = xr + best_d, xp = xr + best_d + 1;
            if (xm >= 0 && xm < w && xc >= 0 && xc < w && xp >= 0 && xp < w && di > 0 && di + 1 < D) {
                const uint16_t cm = cost[((size_t)y * w + xm) * D + di - 1];
                const uint16_t c0 = cost[((size_t)y * w + xc) * D + di];
                const uint16_t cp = cost[((size_t)y * w + xp) * D + di + 1];
                if (cm != kInv16 && c0 != kInv16 && cp != kInv16) dval += legacy_delta((float)cm, (float)c0, (float)cp);
            }
        }
    }
    dr[(size_t)y * w + xr] = dval;
}

hipError_t launch_right_wta(hipStream_t s, const apg_hip_params& p, int w, int h, const uint16_t* cost, float* dr) {
    dim3 grid((w + RW_TX - 1) / RW_TX, (h + 3) / 4), block(64, 4);
    hipLaunchKernelGGL(right_wta_kernel, grid, block, 0, s, w, h, p.min_disparity, p.max_disparity - p.min_disparity,
                       p.post_subpixel, cost, dr);
    return hipGetLastError();
}

// ----------------------------------------------------------- cross arms ----
// CostAggregator::build_cross_arms (reference src/cost_aggregator.cpp:17-63).
__global__ void arms_kernel(const uint8_t* __restrict__ g, int w, int h, int Lmax, int tau, uint32_t* __restrict__ arms) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const int c0 = g[(size_t)y * w + x];
    int l = 0, r = 0, u = 0, v = 0;
    while (l < Lmax && x - l - 1 >= 0 && abs((int)g[(size_t)y * w + x - l - 1] - c0) <= tau) ++l;
    while (r < Lmax && x + r + 1 < w && abs((int)g[(size_t)y * w + x + r + 1] - c0) <= tau) ++r;
    while (u < Lmax && y - u - 1 >= 0 && abs((int)g[(size_t)(y - u - 1) * w + x] - c0) <= tau) ++u;
    while (v < Lmax && y + v + 1 < h && abs((int)g[(size_t)(y + v + 1) * w + x] - c0) <= tau) ++v;
    arms[(size_t)y * w + x] = (uint32_t)l | ((uint32_t)r << 8) | ((uint32_t)u << 16) | ((uint32_t)v << 24);
}

// The same arms with every neighbour loaded up front (LM >= Lmax values per
// direction, clamped inside the image, all in flight together) and the four
// scans run on registers: the first neighbour past the border, past Lmax or
// off by more than tau ends an arm.
template <int LM>
__global__ __launch_bounds__(128) void arms_batch_kernel(const uint8_t* __restrict__ g, int w, int h, int Lmax,
                                                         int tau, uint32_t* __restrict__ arms) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    const int xc = min(x, w - 1);
    const uint8_t* row = g + (size_t)y * w;
    const int c0 = row[xc];
    uint8_t vl[LM], vr[LM], vu[LM], vd[LM];
#pragma unroll
    for (int k = 0; k < LM; ++k) {
        vl[k] = row[max(xc - k - 1, 0)];
        vr[k] = row[min(xc + k + 1, w - 1)];
        vu[k] = g[(size_t)max(y - k - 1, 0) * w + xc];
        vd[k] = g[(size_t)min(y + k + 1, h - 1) * w + xc];
    }
    if (x >= w) return;
    int l = 0, r = 0, u = 0, v = 0;
    bool gl = true, gr = true, gu = true, gd = true;
#pragma unroll
    for (int k = 0; k < LM; ++k) {
        const bool in = k < Lmax;
        gl = gl && in && x - k - 1 >= 0 && abs((int)vl[k] - c0) <= tau;
        gr = gr && in && x + k + 1 < w && abs((int)vr[k] - c0) <= tau;
        gu = gu && in && y - k - 1 >= 0 && abs((int)vu[k] - c0) <= tau;
        gd = gd && in && y + k + 1 < h && abs((int)vd[k] - c0) <= tau;
        l += gl;
        r += gr;
        u += gu;
        v += gd;
    }
    arms[(size_t)y * w + x] = (uint32_t)l | ((uint32_t)r << 8) | ((uint32_t)u << 16) | ((uint32_t)v << 24);
}

hipError_t launch_cross_arms(hipStream_t s, const apg_hip_params& p, const uint8_t* gray, int w, int h, uint32_t* arms) {
    const int Lmax = p.max_arm_length > 1 ? p.max_arm_length : 1;
    const dim3 grid((w + 127) / 128, h);
    if (Lmax <= 17)
        hipLaunchKernelGGL(arms_batch_kernel<17>, grid, dim3(128), 0, s, gray, w, h, Lmax, p.color_threshold, arms);
    else if (Lmax <= 32)
        hipLaunchKernelGGL(arms_batch_kernel<32>, grid, dim3(128), 0, s, gray, w, h, Lmax, p.color_threshold, arms);
    else
        hipLaunchKernelGGL(arms_kernel, grid, dim3(128), 0, s, gray, w, h, Lmax, p.color_threshold, arms);
    return hipGetLastError();
}

}  // namespace apgk
