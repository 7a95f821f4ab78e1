// gfx950 kernels of the refiner (reference src/refiner.cpp:14-98).
//
// The reference sweeps the unreliable pixels in raster order (backward on odd
// iterations); each takes the best of up to four candidates -- its own value,
// the two neighbours the sweep has already updated, and its value plus a
// jitter drawn from std::uniform_int_distribution<int>(-r, r) over one
// std::mt19937(12345) stream -- by the raw LUT cost at that disparity.
// On the device:
//   * which pixels draw a jitter depends only on the values before the sweep
//     (unreliable, non-empty range, value >= 0), so a draw's index is an
//     exclusive prefix sum over the sweep order;
//   * libstdc++ 11 maps a 32-bit engine output e to (e * s) >> 32 - r with
//     s = 2r + 1 and accepts it iff (e * s) mod 2^32 >= 2^32 mod s, retrying
//     otherwise (Lemire, "Fast Random Integer Generation in an Interval",
//     ACM TOMACS 29(1), 2019, algorithm 5): whether an output is accepted does
//     not depend on the data, so the k-th draw of a sweep is the k-th accepted
//     output after the outputs the earlier sweeps consumed (a prefix sum over
//     the acceptance flags, then a compaction);
//   * the candidates that do not depend on the sweep (own value, jitter) are
//     evaluated for every pixel in parallel;
//   * a pixel depends on its left and upper neighbours (mirrored for backward
//     sweeps), so one workgroup walks the anti-diagonals x + y = s, a thread
//     per row, the upper neighbours' results passed through LDS.
#include <algorithm>

#include <hipcub/hipcub.hpp>

#include "device_common.hpp"
#include "kernels.hpp"

namespace apgk {

namespace {

// a neighbour with no candidate (the reference pushes only values >= 0, so
// any negative value means none for a neighbour; the own and jitter
// candidates, whose clamped disparity may be negative, carry flags instead)
constexpr float kNoCand = -1.f;
constexpr uint32_t kHasOwn = 1u, kHasJitter = 2u;

__device__ __forceinline__ int sweep_index(bool fwd, int w, int h, int x, int y) {
    return fwd ? y * w + x : (h - 1 - y) * w + (w - 1 - x);
}

// 1 where the pixel at sweep position j draws a jitter
__global__ void refine_mask_kernel(int w, int h, int fwd, const uint8_t* __restrict__ reliable,
                                   const int16_t* __restrict__ dmin, const int16_t* __restrict__ dmax,
                                   const float* __restrict__ disp, uint32_t* __restrict__ flag) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)w * h) return;
    const int x = (int)(q % w), y = (int)(q / w);
    flag[sweep_index(fwd, w, h, x, y)] = (!reliable[q] && dmax[q] > dmin[q] && disp[q] >= 0.f) ? 1u : 0u;
}

// acceptance of engine outputs base .. base + n - 1 for the range s (t = 2^32 mod s)
__global__ void jitter_accept_kernel(const uint32_t* __restrict__ eng, const uint32_t* __restrict__ base, size_t n,
                                     uint32_t s, uint32_t t, uint32_t* __restrict__ flag) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (uint32_t)((uint64_t)eng[*base + i] * s) >= t ? 1u : 0u;
}

// acc[i] -> slot pos[i] of the list of accepted offsets
__global__ void jitter_compact_kernel(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos, size_t n,
                                      uint32_t* __restrict__ acc) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) acc[pos[i]] = (uint32_t)i;
}

// after a sweep: the engine outputs it consumed end at its last accepted one
__global__ void jitter_advance_kernel(uint32_t* __restrict__ base, const uint32_t* __restrict__ dflag,
                                      const uint32_t* __restrict__ didx, size_t npx, const uint32_t* __restrict__ acc) {
    const uint32_t ndraw = didx[npx - 1] + dflag[npx - 1];
    if (ndraw) *base += acc[ndraw - 1] + 1u;
}

// Anti-diagonal-major order of the sweep coordinates (xs, ys): diagonal
// s = xs + ys, then ys.  Pixels before diagonal s: every row ys < h holds
// min(max(s - ys, 0), w) of them.
__device__ __forceinline__ uint32_t diag_start(int s, int w, int h) {
    const int64_t n1 = min(max(s - w + 1, 0), h);   // rows with all w pixels before s
    const int64_t n2 = min(max(s, 0), h);           // rows with some
    const int64_t a = s - n1, b = s - n2;           // (s - ys) for ys = n1 .. n2 - 1: a down to b + 1
    return (uint32_t)(n1 * w + (a * (a + 1) - b * (b + 1)) / 2);
}
__device__ __forceinline__ uint32_t diag_index(int xs, int ys, int w, int h) {
    const int s = xs + ys;
    return diag_start(s, w, h) + (uint32_t)(ys - max(0, s - w + 1));
}

// one view's terms of a pixel: census (two words), gray | gx << 8 | gy << 16
__global__ void refine_terms_kernel(int n, const uint8_t* __restrict__ g, const uint8_t* __restrict__ gx,
                                    const uint8_t* __restrict__ gy, const uint32_t* __restrict__ c32,
                                    const uint64_t* __restrict__ c64, uint4* __restrict__ rec) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    uint4 r;
    if (c32) { r.x = c32[q]; r.y = 0u; }
    else { const uint64_t c = c64[q]; r.x = (uint32_t)c; r.y = (uint32_t)(c >> 32); }
    r.z = (uint32_t)g[q] | ((uint32_t)gx[q] << 8) | ((uint32_t)gy[q] << 16);
    r.w = 0u;
    rec[q] = r;
}

// local_cost (reference src/refiner.cpp:14-38) from the two views' terms
__device__ __forceinline__ float rec_cost(const uint4& l, const uint4& r, bool sym, int use_ad, int use_grad,
                                          const float* lc, const float* la, const float* lg) {
    const int pop = sym ? __popc(l.x ^ r.x) : __popc(l.x ^ r.x) + __popc(l.y ^ r.y);
    float c = lc[pop];
    if (use_ad) c += la[abs((int)(l.z & 0xFFu) - (int)(r.z & 0xFFu))];
    if (use_grad)
        c += lg[abs((int)((l.z >> 8) & 0xFFu) - (int)((r.z >> 8) & 0xFFu)) +
                abs((int)((l.z >> 16) & 0xFFu) - (int)((r.z >> 16) & 0xFFu))];
    return c;
}

// Per pixel, stored in anti-diagonal order: value, range, whether the sweep
// changes it, the candidates that do not depend on the sweep (own value,
// jitter: cost and clamped disparity, present when `has` says so) and the
// left view's terms.
struct SweepRec {
    float cur;
    int lohi;           // lo | hi << 16
    uint32_t act;       // unreliable with a non-empty range
    uint32_t has;       // kHasOwn | kHasJitter: the reference pushes that candidate
    float4 st;          // cost, disparity of the own value; of the jitter
    uint4 lt;           // left-view terms
};

__global__ void refine_static_kernel(const CostLut* __restrict__ lut, int w, int h, int fwd, int r, int sym,
                                     int use_ad, int use_grad, const uint4* __restrict__ lrec,
                                     const uint4* __restrict__ rrec, const uint8_t* __restrict__ reliable,
                                     const int16_t* __restrict__ dmin, const int16_t* __restrict__ dmax,
                                     const float* __restrict__ disp, const uint32_t* __restrict__ eng,
                                     uint32_t eng_last, const uint32_t* __restrict__ base,
                                     const uint32_t* __restrict__ acc, const uint32_t* __restrict__ didx,
                                     SweepRec* __restrict__ out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= w * h) return;
    const int x = q % w, y = q / w;
    const int xs = fwd ? x : w - 1 - x, ys = fwd ? y : h - 1 - y;
    const int lo = dmin[q], hi = dmax[q];
    const float cur = disp[q];
    SweepRec o;
    o.cur = cur;
    o.lohi = (int)(uint16_t)lo | (hi << 16);
    o.act = (!reliable[q] && hi > lo) ? 1u : 0u;
    o.has = 0u;
    o.st = make_float4(0.f, 0.f, 0.f, 0.f);
    o.lt = lrec[q];
    auto cost = [&](int d) {
        const int xr = x - d;
        if (xr < 0 || xr >= w || d < lo || d >= hi) return 1e6f;
        return rec_cost(o.lt, rrec[(size_t)y * w + xr], sym, use_ad, use_grad, lut->census, lut->ad, lut->grad);
    };
    if (o.act && cur >= 0.f) {
        const int dc = clampi((int)roundf(cur), lo, hi - 1);
        o.st.x = cost(dc);
        o.st.y = (float)dc;
        o.has = kHasOwn;
        const uint32_t s = 2u * (uint32_t)r + 1u;
        const uint32_t e = eng[min(*base + acc[didx[sweep_index(fwd, w, h, x, y)]], eng_last)];
        const int jit = (int)(((uint64_t)e * s) >> 32) - r;
        const float dj = cur + (float)jit;
        if (!(dj < 0.f)) {
            const int d = clampi((int)roundf(dj), lo, hi - 1);
            o.st.z = cost(d);
            o.st.w = (float)d;
            o.has |= kHasJitter;
        }
    }
    out[diag_index(xs, ys, w, h)] = o;
}

// results of the sweep (anti-diagonal order) back to the image
__global__ void refine_unpack_kernel(int w, int h, int fwd, const float* __restrict__ res, float* __restrict__ disp) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= w * h) return;
    const int x = q % w, y = q / w;
    disp[q] = res[diag_index(fwd ? x : w - 1 - x, fwd ? y : h - 1 - y, w, h)];
}

// One workgroup, a thread per row (rows t, t + blockDim, ...), anti-diagonals
// of the sweep order in turn; the result of (x, y - 1) reaches row y through
// LDS (double-buffered by diagonal parity).  The threads of a diagonal read
// consecutive records, loaded one diagonal ahead, so a step waits only for
// the right-view terms of its two neighbour candidates.
constexpr int kSweepThreads = 1024;
constexpr int kSweepRows = 4;   // rows per thread: h <= 4096

__global__ __launch_bounds__(kSweepThreads) void refine_sweep_kernel(const CostLut* __restrict__ lut, int w, int h,
                                                                     int fwd, int sym, int use_ad, int use_grad,
                                                                     const SweepRec* __restrict__ rec,
                                                                     const uint4* __restrict__ rrec,
                                                                     float* __restrict__ res_out) {
    extern __shared__ float sh[];
    float* s_lc = sh;
    float* s_la = s_lc + 65;
    float* s_lg = s_la + 256;
    float* s_up = s_lg + 511;   // [2][h]
    const int tid = threadIdx.x;
    for (int i = tid; i < 65; i += kSweepThreads) s_lc[i] = lut->census[i];
    for (int i = tid; i < 256; i += kSweepThreads) s_la[i] = lut->ad[i];
    for (int i = tid; i < 511; i += kSweepThreads) s_lg[i] = lut->grad[i];
    __syncthreads();
    float left[kSweepRows];
    SweepRec nx[kSweepRows];
#pragma unroll
    for (int k = 0; k < kSweepRows; ++k) {
        left[k] = kNoCand;
        const int ys = tid + k * kSweepThreads;
        nx[k] = rec[ys < h ? diag_index(0, ys, w, h) : 0];   // the row's first pixel
    }
    const int ns = w + h - 1;
    for (int s = 0; s < ns; ++s) {
        float* up_r = s_up + ((s + 1) & 1) * h;   // written at diagonal s - 1
        float* up_w = s_up + (s & 1) * h;
        const uint32_t ds = diag_start(s, w, h), ds1 = diag_start(s + 1, w, h);
        const int y0 = max(0, s - w + 1), y1 = max(0, s + 1 - w + 1);
#pragma unroll
        for (int k = 0; k < kSweepRows; ++k) {
            const int ys = tid + k * kSweepThreads;
            const int xs = s - ys;
            if (ys >= h || xs < 0 || xs >= w) continue;
            const SweepRec p = nx[k];
            const int x = fwd ? xs : w - 1 - xs, y = fwd ? ys : h - 1 - ys;
            const int lo = (int)(int16_t)(p.lohi & 0xFFFF), hi = p.lohi >> 16;
            float res = p.cur;
            if (p.act) {
                float best = 1e9f;
                int bd = p.cur >= 0.f ? clampi((int)roundf(p.cur), lo, hi - 1) : lo;
                // candidates in the reference's order: own value, the neighbour
                // along the row, the neighbour across rows, the jitter
                if ((p.has & kHasOwn) && p.st.x < best) { best = p.st.x; bd = (int)p.st.y; }
                const float nb[2] = {xs > 0 ? left[k] : kNoCand, ys > 0 ? up_r[ys - 1] : kNoCand};
                float nc[2];
                int nd[2];
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    nd[i] = nb[i] < 0.f ? lo : clampi((int)roundf(nb[i]), lo, hi - 1);
                    const int xr = x - nd[i];
                    const bool in = xr >= 0 && xr < w;
                    const uint4 rt = rrec[(size_t)y * w + (in ? xr : 0)];
                    nc[i] = in ? rec_cost(p.lt, rt, sym, use_ad, use_grad, s_lc, s_la, s_lg) : 1e6f;
                }
#pragma unroll
                for (int i = 0; i < 2; ++i)
                    if (!(nb[i] < 0.f) && nc[i] < best) { best = nc[i]; bd = nd[i]; }
                if ((p.has & kHasJitter) && p.st.z < best) { best = p.st.z; bd = (int)p.st.w; }
                res = (float)bd;
            }
            left[k] = res;
            up_w[ys] = res;
            res_out[ds + (uint32_t)(ys - y0)] = res;
            // the row's pixel of the next diagonal, loaded after this step's
            // neighbour terms (vector memory completes in order)
            if (xs + 1 < w) nx[k] = rec[ds1 + (uint32_t)(ys - y1)];
        }
        __syncthreads();
    }
}

}  // namespace

size_t refine_scratch_bytes(int w, int h) {
    const size_t n = (size_t)w * h, nw = n + kRefineSlack;
    size_t temp = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, temp, (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)nw);
    // flags, draw indices, sweep results, acceptance flags, their positions,
    // accepted offsets, left and right terms, sweep records, engine base, scan temp
    return 3 * n * 4 + 3 * nw * 4 + 2 * n * 16 + n * sizeof(SweepRec) + 256 + ((temp + 255) & ~(size_t)255);
}

hipError_t launch_refine(hipStream_t st, const apg_hip_params& p, const CostLut* d_lut, int w, int h,
                         const RefineInputs& in, const int16_t* dmin, const int16_t* dmax, const uint8_t* reliable,
                         float* disp, const uint32_t* eng, size_t eng_len, void* scratch, size_t scratch_bytes) {
    const size_t n = (size_t)w * h, nw = n + kRefineSlack;
    if (h > kSweepThreads * kSweepRows || n >= (1u << 31)) return hipErrorInvalidValue;
    if (scratch_bytes < refine_scratch_bytes(w, h)) return hipErrorInvalidValue;
    const int iters = p.refine_iterations;
    if (!eng || !scratch || eng_len < (size_t)(iters + 1) * nw) return hipErrorInvalidValue;
    // engine offsets (base, acc, the clamp index) are 32-bit
    if (eng_len > 0xFFFFFFFFull) return hipErrorInvalidValue;
    uint8_t* sp = (uint8_t*)scratch;
    uint32_t* dflag = (uint32_t*)sp; sp += n * 4;
    uint32_t* didx = (uint32_t*)sp; sp += n * 4;
    float* sres = (float*)sp; sp += n * 4;
    uint32_t* aflag = (uint32_t*)sp; sp += nw * 4;
    uint32_t* apos = (uint32_t*)sp; sp += nw * 4;
    uint32_t* acc = (uint32_t*)sp; sp += nw * 4;
    uint4* lrec = (uint4*)sp; sp += n * 16;
    uint4* rrec = (uint4*)sp; sp += n * 16;
    SweepRec* srec = (SweepRec*)sp; sp += n * sizeof(SweepRec);
    uint32_t* base = (uint32_t*)sp; sp += 256;
    void* temp = sp;
    size_t temp_bytes = scratch_bytes - (size_t)(sp - (uint8_t*)scratch);
    const int sym = in.c32l ? 1 : 0;
    hipError_t e = launch_fill_u32(st, 1, 0u, base);   // engine outputs consumed so far
    if (e != hipSuccess) return e;
    const unsigned gb = (unsigned)((n + 255) / 256), gw = (unsigned)((nw + 255) / 256);
    hipLaunchKernelGGL(refine_terms_kernel, dim3(gb), dim3(256), 0, st, (int)n, in.gl, in.gxl, in.gyl, in.c32l,
                       in.c64l, lrec);
    hipLaunchKernelGGL(refine_terms_kernel, dim3(gb), dim3(256), 0, st, (int)n, in.gr, in.gxr, in.gyr, in.c32r,
                       in.c64r, rrec);
    const size_t shm = (65 + 256 + 511 + 2 * (size_t)h) * sizeof(float);
    for (int it = 0; it < iters; ++it) {
        const int fwd = (it % 2) == 0 ? 1 : 0;
        const int r = std::max(1, p.refine_random_radius >> it);
        const uint32_t s = 2u * (uint32_t)r + 1u;
        const uint32_t t = (uint32_t)(0u - s) % s;   // 2^32 mod s
        hipLaunchKernelGGL(refine_mask_kernel, dim3(gb), dim3(256), 0, st, w, h, fwd, reliable, dmin, dmax, disp, dflag);
        if ((e = hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, dflag, didx, (int)n, st)) != hipSuccess) return e;
        hipLaunchKernelGGL(jitter_accept_kernel, dim3(gw), dim3(256), 0, st, eng, base, nw, s, t, aflag);
        if ((e = hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, aflag, apos, (int)nw, st)) != hipSuccess) return e;
        hipLaunchKernelGGL(jitter_compact_kernel, dim3(gw), dim3(256), 0, st, aflag, apos, nw, acc);
        hipLaunchKernelGGL(refine_static_kernel, dim3(gb), dim3(256), 0, st, d_lut, w, h, fwd, r, sym, p.use_ad,
                           p.use_grad, lrec, rrec, reliable, dmin, dmax, disp, eng, (uint32_t)(eng_len - 1), base,
                           acc, didx, srec);
        hipLaunchKernelGGL(refine_sweep_kernel, dim3(1), dim3(kSweepThreads), shm, st, d_lut, w, h, fwd, sym,
                           p.use_ad, p.use_grad, srec, rrec, sres);
        hipLaunchKernelGGL(refine_unpack_kernel, dim3(gb), dim3(256), 0, st, w, h, fwd, sres, disp);
        hipLaunchKernelGGL(jitter_advance_kernel, dim3(1), dim3(1), 0, st, base, dflag, didx, n, acc);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace apgk
