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
#include <utility>

#include <hipcub/hipcub.hpp>

#include "device_common.hpp"
#include "kernels.hpp"

namespace apgk {

namespace {

// f(integral_constant<int, 0>) ... f(integral_constant<int, N-1>)
template <typename F, int... I>
__device__ __forceinline__ void static_for_b_impl(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for_b(F&& f) {
    static_for_b_impl(f, std::make_integer_sequence<int, N>{});
}

// a neighbour with no candidate (the reference pushes only values >= 0, so
// any negative value means none for a neighbour; the own and jitter
// candidates, whose clamped disparity may be negative, carry flags instead)
constexpr float kNoCand = -1.f;
constexpr uint32_t kHasOwn = 1u, kHasJitter = 2u;

__device__ __forceinline__ int sweep_index(bool fwd, int w, int h, int x, int y) {
    return fwd ? y * w + x : (h - 1 - y) * w + (w - 1 - x);
}

// Per sweep position i < nw, one pair of flags: x = 1 where the pixel at
// sweep position i (i < n) draws a jitter, y = 1 where engine output
// base + i is accepted for the range s (t = 2^32 mod s).  One exclusive scan
// of the pairs gives each pixel's draw index and each accepted output's slot.
__global__ void refine_flags_kernel(int w, int h, int fwd, const uint8_t* __restrict__ reliable,
                                    const int16_t* __restrict__ dmin, const int16_t* __restrict__ dmax,
                                    const float* __restrict__ disp, const uint32_t* __restrict__ eng,
                                    const uint32_t* __restrict__ base, size_t n, size_t nw, uint32_t s, uint32_t t,
                                    uint2* __restrict__ flag) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nw) return;
    uint32_t d = 0u;
    if (i < n) {
        const int ys = (int)(i / (size_t)w), xs = (int)(i - (size_t)ys * w);
        const int x = fwd ? xs : w - 1 - xs, y = fwd ? ys : h - 1 - ys;
        const size_t q = (size_t)y * w + x;
        d = (!reliable[q] && dmax[q] > dmin[q] && disp[q] >= 0.f) ? 1u : 0u;
    }
    const uint32_t a = (uint32_t)((uint64_t)eng[*base + i] * s) >= t ? 1u : 0u;
    flag[i] = make_uint2(d, a);
}

struct PairSum {
    __host__ __device__ uint2 operator()(const uint2& a, const uint2& b) const {
        return make_uint2(a.x + b.x, a.y + b.y);
    }
};

// acc[i] -> slot pos[i].y of the list of accepted offsets
__global__ void jitter_compact_kernel(const uint2* __restrict__ flag, const uint2* __restrict__ pos, size_t n,
                                      uint32_t* __restrict__ acc) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i].y) acc[pos[i].y] = (uint32_t)i;
}

// after a sweep: the engine outputs it consumed end at its last accepted one
__global__ void jitter_advance_kernel(uint32_t* __restrict__ base, const uint2* __restrict__ flag,
                                      const uint2* __restrict__ pos, size_t npx, const uint32_t* __restrict__ acc) {
    const uint32_t ndraw = pos[npx - 1].x + flag[npx - 1].x;
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
                                     const uint32_t* __restrict__ acc, const uint2* __restrict__ dpos,
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
        const uint32_t e = eng[min(*base + acc[dpos[sweep_index(fwd, w, h, x, y)].x], eng_last)];
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

// Banded form of the same sweep: the rows of the sweep order are split into
// bands of 64, one wavefront (one workgroup) per band, lane l holding row
// b * 64 + l.  At local step k lane l handles sweep column xs = k - l, so the
// band's anti-diagonals advance one per step; a pixel's upper neighbour is the
// lane below's result of the previous step (a DPP shift), its left neighbour
// the lane's own; no barrier is needed.
//  * Records are band-major, slot ((b * kb + k) << 6) + l, written by
//    refine_band_static_kernel for every slot (slots outside the image hold a
//    record whose own candidate always wins with value -1), so each step
//    loads a coalesced block at a wave-uniform base.  A record carries the
//    static candidates already resolved: cost sc, float value sd and its
//    rounded value sdi of the winner between the own value and the jitter
//    (a jitter winner's cost is stored as the next float up, so that a
//    neighbour with an equal cost, which the reference tries first, wins).
//  * The neighbours' values travel as rounded integers (negative: none), so
//    the sweep never rounds.
//  * A band's first row takes its upper neighbours from the band above's last
//    row through `edge`: lane 63 stores each result, tagged (bit 31 set), with
//    an agent-scope relaxed atomic store at the index of its step; the band
//    below loads the values of a block of PD columns with one relaxed,
//    agent-scope load a block ahead of its use, and re-polls only if a column
//    inside the row does not carry the tag yet.  Each value is its own flag,
//    so no fence is needed.  The wait is bounded (bs.timeout ticks of the wall clock from the
//    kernel's start and an iteration cap per wait); a band that gives up sets
//    *bs.err, takes no neighbour from above for the rest of the sweep and
//    runs to its end, so every wave exits.  A negative timeout makes every
//    band but the first give up at once.  Bands wait only on bands with
//    lower indices, which each XCD dispatches first, and the grid holds at
//    most one band per CU.
//  * The right-view terms a pixel's candidates need, columns x - d for d in
//    [d0, d0 + D), sit in an LDS ring per row (RS = max(64, 2^ceil(log2 D))
//    entries of 8 bytes): the window slides by one column per step, the
//    entering column loaded PD steps ahead from a copy of the right-view
//    terms padded by kBandPadCols entries at both ends, so every load is in
//    bounds without a clamp.  Columns outside the image hold another row's
//    terms; no candidate reads them (their cost is the out-of-range 1e6).
//    Lanes before the start of their row write the columns their window will
//    hold anyway (RS >= 64 keeps those writes off the other columns' slots).
//  * Records, results and the entering columns are read and written through
//    buffer resources, the per-step offsets scalar; the host keeps every
//    offset below 2^31 bytes.
constexpr int kBandRows = 64;
constexpr int kBandPD = 8;          // steps of prefetch (records, entering columns, edge values)
constexpr int kBandSlack = 96;      // local steps past w in a band's slots: the skew (63) and two prefetch spans
constexpr int kBandPadCols = 2048;  // right-view terms padding at both ends
constexpr int kBandLdsBytes = kBandRows * 256 * 8 + (65 + 256 + 511) * 4;   // rings of D <= 256, LUTs

__host__ __device__ constexpr int band_steps(int w) { return w + kBandSlack; }

// The record of band slot (xs, ys): A = {sc, sd, sdi, lo}, B = {hi - 1,
// census, gray, gx << 8 | gy << 16}.
__device__ __forceinline__ void band_static_record(const CostLut* __restrict__ lut, int w, int h, int fwd, int r,
                                                   int use_ad, int use_grad, const uint4* __restrict__ lrec,
                                                   const uint4* __restrict__ rrec,
                                                   const uint8_t* __restrict__ reliable,
                                                   const int16_t* __restrict__ dmin, const int16_t* __restrict__ dmax,
                                                   const float* __restrict__ disp, const uint32_t* __restrict__ eng,
                                                   uint32_t eng_last, const uint32_t* __restrict__ base,
                                                   const uint32_t* __restrict__ acc,
                                                   const uint2* __restrict__ dpos, int xs, int ys, float4& ra,
                                                   uint4& rb) {
    if (xs < 0 || xs >= w || ys >= h) {
        ra = make_float4(-1.f, -1.f, __int_as_float(-1), __int_as_float(0));
        rb = make_uint4(0u, 0u, 0u, 0u);
        return;
    }
    const int x = fwd ? xs : w - 1 - xs, y = fwd ? ys : h - 1 - ys;
    const size_t q = (size_t)y * w + x;
    const int lo = dmin[q], hi = dmax[q];
    const float cur = disp[q];
    const uint4 lt = lrec[q];
    auto cost = [&](int d) {
        const int xr = x - d;
        if (xr < 0 || xr >= w || d < lo || d >= hi) return 1e6f;
        return rec_cost(lt, rrec[(size_t)y * w + xr], true, use_ad, use_grad, lut->census, lut->ad, lut->grad);
    };
    float sc, sd;
    int sdi;
    if (reliable[q] || hi <= lo) {   // not swept: the value stays, nothing beats it
        sc = -1.f;
        sd = cur;
        sdi = cur >= 0.f ? (int)roundf(cur) : -1;
    } else {
        float best = 1e9f;
        int bd = cur >= 0.f ? clampi((int)roundf(cur), lo, hi - 1) : lo;
        bool jit_wins = false;
        if (cur >= 0.f) {
            const int dc = clampi((int)roundf(cur), lo, hi - 1);
            best = cost(dc);
            bd = dc;
            const uint32_t s = 2u * (uint32_t)r + 1u;
            const uint32_t e = eng[min(*base + acc[dpos[sweep_index(fwd, w, h, x, y)].x], eng_last)];
            const int jit = (int)(((uint64_t)e * s) >> 32) - r;
            const float dj = cur + (float)jit;
            if (!(dj < 0.f)) {
                const int d = clampi((int)roundf(dj), lo, hi - 1);
                const float cj = cost(d);
                if (cj < best) {
                    best = cj;
                    bd = d;
                    jit_wins = true;
                }
            }
        }
        // the jitter is tried after the neighbours, so they win its ties
        // (costs are finite and >= 0: the next float up is the next bit pattern)
        sc = jit_wins ? __int_as_float(__float_as_int(best + 0.f) + 1) : best;
        sd = (float)bd;
        sdi = bd;
    }
    ra = make_float4(sc, sd, __int_as_float(sdi), __int_as_float(lo));
    rb = make_uint4((uint32_t)(hi - 1), lt.x, lt.z & 0xFFu, lt.z & 0xFFFF00u);
}

// Slot records of one band (blockIdx.y), kStaticDiag consecutive steps
// (blockIdx.x) of 256 / kStaticDiag of its rows (blockIdx.z), one slot per
// thread: a row's pixels on consecutive lanes, so the per-pixel reads and the
// draw-index -> accepted-offset -> engine chain are coalesced; the records
// are written to their slots directly (staging them in LDS to write whole
// steps measured slower: the occupancy it costs hides less of the chain's
// latency).
constexpr int kStaticDiag = 32;
__global__ __launch_bounds__(256) void refine_band_static_kernel(const CostLut* __restrict__ lut, int w, int h,
                                                                  int fwd, int r, int kb, int use_ad, int use_grad,
                                                                  const uint4* __restrict__ lrec,
                                                                  const uint4* __restrict__ rrec,
                                                                  const uint8_t* __restrict__ reliable,
                                                                  const int16_t* __restrict__ dmin,
                                                                  const int16_t* __restrict__ dmax,
                                                                  const float* __restrict__ disp,
                                                                  const uint32_t* __restrict__ eng, uint32_t eng_last,
                                                                  const uint32_t* __restrict__ base,
                                                                  const uint32_t* __restrict__ acc,
                                                                  const uint2* __restrict__ dpos,
                                                                  float4* __restrict__ ra, uint4* __restrict__ rb) {
    const int b = blockIdx.y, k0 = blockIdx.x * kStaticDiag;
    const int t = threadIdx.x;
    const int j = t % kStaticDiag, l = t / kStaticDiag + (256 / kStaticDiag) * blockIdx.z;
    float4 A;
    uint4 B;
    band_static_record(lut, w, h, fwd, r, use_ad, use_grad, lrec, rrec, reliable, dmin, dmax, disp, eng,
                       eng_last, base, acc, dpos, k0 + j - l, b * kBandRows + l, A, B);
    if (k0 + j < kb) {
        const size_t sl = ((size_t)b * kb + k0 + j) * kBandRows + l;
        ra[sl] = A;
        rb[sl] = B;
    }
}

// the right view's terms (census, gray | gx << 8 | gy << 16) at [kBandPadCols + q]
__global__ void refine_band_terms_kernel(int n, const uint4* __restrict__ rrec, uint2* __restrict__ rt) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const uint4 r = rrec[q];
    rt[kBandPadCols + q] = make_uint2(r.x, r.z);
}

// sweep results, band-major, back to the image
__global__ void refine_band_unpack_kernel(int w, int h, int fwd, int kb, const float* __restrict__ res,
                                          float* __restrict__ disp) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= w * h) return;
    const int x = q % w, y = q / w;
    const int xs = fwd ? x : w - 1 - x, ys = fwd ? y : h - 1 - y;
    const int b = ys / kBandRows, l = ys % kBandRows;
    disp[q] = res[(((size_t)b * kb + (size_t)(xs + l)) << 6) + l];
}

struct BandSync {
    uint32_t* edge;     // [bands + 1][kb]: row b + 1 the tagged results of band b's last row, at its step index (zeroed before the sweep)
    uint32_t* err;      // set when a wait gave up (sticky: cleared by the host after reading it)
    long long timeout;  // wall-clock ticks the waits may take; < 0: give up at once
};

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// the median of three (v_med3_i32): min(max(v, lo), hi) when lo <= hi
__device__ __forceinline__ int med3_i32(int v, int lo, int hi) {
    int r;
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(v), "v"(lo), "v"(hi));
    return r;
}
// c - 8 a (v_mad_i32_i24) for |a| < 2^23
__device__ __forceinline__ int sub8x_i24(int c, int a) {
    int r;
    asm("v_mad_i32_i24 %0, %1, -8, %2" : "=v"(r) : "v"(a), "v"(c));
    return r;
}

template <int PD, bool AD, bool GRAD>
__global__ __launch_bounds__(kBandRows) void refine_band_kernel(const CostLut* __restrict__ lut, int w, int h,
                                                                 int fwd, int d0, int D, int rs_log2, int kb,
                                                                 int krun, const float4* __restrict__ ra,
                                                                 const uint4* __restrict__ rb,
                                                                 const uint2* __restrict__ rt,
                                                                 float* __restrict__ out, BandSync bs) {
    // LDS (static, so its base is a constant): row l's window ring at byte
    // l * RS * 8, so a column's slot is one AND-OR of its byte offset, then
    // the census, AD and gradient LUTs
    __shared__ __align__(16) unsigned char shb[kBandLdsBytes];
    const int lane = threadIdx.x, b = blockIdx.x;
    const uint32_t rowbase = (uint32_t)lane << (rs_log2 + 3);
    const uint32_t mask8 = (8u << rs_log2) - 1u;
    const uint32_t lut_c = (uint32_t)kBandRows << (rs_log2 + 3);
    const uint32_t lut_a = lut_c + 65 * 4, lut_g = lut_a + 256 * 4;
    auto ring = [&](int col8) { return (u32x2*)(shb + (((uint32_t)col8 & mask8) | rowbase)); };
    auto lutf = [&](uint32_t base, uint32_t i) { return *(const float*)(shb + base + 4 * i); };
    {
        constexpr int NLUT = 65 + 256 + 511, NL = (NLUT + 63) / 64;
        float lv[NL];
#pragma unroll
        for (int k = 0; k < NL; ++k) lv[k] = lut->census[min(lane + 64 * k, NLUT - 1)];
#pragma unroll
        for (int k = 0; k < NL; ++k)
            if (lane + 64 * k < NLUT) *(float*)(shb + lut_c + 4 * (lane + 64 * k)) = lv[k];
    }
    const int ys = b * kBandRows + lane;
    const int yc = min(ys, h - 1);
    const int y = fwd ? yc : h - 1 - yc;
    // the window of the row's first pixel: columns [x0 - d0 - D + 1, x0 - d0]
    {
        const uint2* rtp = rt + kBandPadCols + (size_t)y * w;
        const int x0 = fwd ? 0 : w - 1;
        const int c0 = x0 - d0 - D + 1;
        for (int j0 = 0; j0 < D; j0 += 32) {
            uint2 t[32];
#pragma unroll
            for (int j = 0; j < 32; ++j) t[j] = rtp[c0 + min(j0 + j, D - 1)];
#pragma unroll
            for (int j = 0; j < 32; ++j)
                if (j0 + j < D) *ring((c0 + j0 + j) * 8) = u32x2{t[j].x, t[j].y};
        }
    }
    // The band's records, results and the right-view terms through buffer
    // resources: per-step offsets are scalar.  The entering column at step k
    // is x - d0 (fwd, x = k - lane) or x - d0 - D + 1 (backward,
    // x = w - 1 - (k - lane)); its byte offset in rt is tvoff + 8k (fwd) or
    // tvoff + 8 (kend - k) (backward), kend past the last step loaded.
    const size_t slot0 = ((size_t)b * kb) << 6;
    const __amdgpu_buffer_rsrc_t rsa = raw_rsrc(ra + slot0), rsb = raw_rsrc(rb + slot0);
    const __amdgpu_buffer_rsrc_t rso = raw_rsrc(out + slot0), rst = raw_rsrc(rt);
    const int kend = krun + PD;
    const uint32_t col = (uint32_t)(kBandPadCols + y * w + (fwd ? -lane - d0 : w - d0 - D + lane));
    const uint32_t tvoff = fwd ? col * 8u : (col - (uint32_t)kend) * 8u;
    const uint32_t rvoff = (uint32_t)lane * 16u;
    u32x4 pa[PD], pb[PD];
    u32x2 pc[PD];
    auto issue = [&](int k, int q) {
        pa[q] = buf_ld128(rsa, rvoff, (uint32_t)k * 1024u);
        pb[q] = buf_ld128(rsb, rvoff, (uint32_t)k * 1024u);
        pc[q] = __builtin_amdgcn_raw_buffer_load_b64(rst, (int)tvoff, (fwd ? k : kend - k) * 8, 0);
    };
    // Upper neighbours of lane 0: the band above's lane 63 stores its result
    // of step k at edge row b, index k, i.e. column c at index c + 63.  Lane j
    // of one load holds column k0 + j of a block of PD steps, loaded a block
    // ahead.
    const uint32_t* edge_in = bs.edge + (size_t)b * kb + (kBandRows - 1);
    uint32_t* edge_out = bs.edge + (size_t)(b + 1) * kb;
    const int el = min(lane, PD - 1);
    // lane 63's store of step k's result: the same instruction as an
    // agent-scope relaxed atomic store (sc1); the other lanes' offsets lie
    // past the resource's records, so the hardware drops their stores and no
    // lane branches
    const __amdgpu_buffer_rsrc_t rse = __builtin_amdgcn_make_buffer_rsrc(edge_out, 0, kb * 4, 0x00020000);
    const uint32_t evoff = lane == kBandRows - 1 ? 0u : 0x80000000u;
    auto edge_load = [&](int k0) {
        return __hip_atomic_load(edge_in + k0 + el, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    uint32_t pev = edge_load(0);
#pragma unroll
    for (int j = 0; j < PD; ++j) issue(j, j);
    const long long t_start = (long long)wall_clock64();
    bool give_up = bs.timeout < 0 && b > 0;
    int ri_prev = -1;                                  // the lane's result of the previous step, rounded
    int x8 = (fwd ? -lane : w - 1 + lane) * 8;        // 8 x, x the lane's image column at the step
    const int xd8 = fwd ? 8 : -8;
    const int ent8 = (fwd ? -d0 : -d0 - D + 1) * 8;   // 8 (entering column - x)
    const uint32_t w8 = (uint32_t)w * 8u;
    for (int k0 = 0; k0 < krun; k0 += PD) {
        uint32_t ev = pev;
        pev = edge_load(k0 + PD);
        int dec = -1;
        if (b > 0 && !give_up) {
            // the block's columns inside the row must carry the tag (bit 31)
            const bool need = lane < PD && k0 + lane < w;
            if (__ballot(need && !(ev >> 31))) {
                for (int it = 0; it < (1 << 22); ++it) {
                    ev = __hip_atomic_load(edge_in + k0 + el, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (!__ballot(need && !(ev >> 31)) || (long long)wall_clock64() - t_start > bs.timeout) break;
                    __builtin_amdgcn_s_sleep(1);
                }
                if (__ballot(need && !(ev >> 31))) give_up = true;
            }
            if (!give_up) dec = (int)(ev & 0x7FFFFFFFu) - 1;
        }
        static_for_b<PD>([&](auto j_t) {
            constexpr int q = decltype(j_t)::value;
            const int k = k0 + q;
            const u32x4 A = pa[q], B = pb[q];
            *ring(x8 + ent8) = pc[q];   // the window takes the entering column
            const int up = __builtin_amdgcn_update_dpp(__builtin_amdgcn_readlane(dec, q), ri_prev, kDppWaveShr1,
                                                       0xF, 0xF, false);
            const float sc = __uint_as_float(A.x), sd = __uint_as_float(A.y);
            const int sdi = (int)A.z, lo = (int)A.w, hm1 = (int)B.x;
            // a neighbour's value clamped to the pixel's range (for swept
            // pixels lo <= hm1; the others keep their value whatever the
            // candidates cost)
            auto cand = [&](int nbv, float& c, int& nd) {
                nd = med3_i32(nbv, lo, hm1);
                const int xr8 = sub8x_i24(x8, nd);   // |nd| <= 2^15 (16-bit ranges)
                const u32x2 t = *ring(xr8);
                float v = lutf(lut_c, __popc(B.y ^ t.x));
                if (AD) v += lutf(lut_a, __builtin_amdgcn_sad_u8(B.z, t.y & 0xFFu, 0u));
                if (GRAD) v += lutf(lut_g, __builtin_amdgcn_sad_u8(B.w, t.y & 0xFFFF00u, 0u));
                v = (uint32_t)xr8 < w8 ? v : 1e6f;
                c = nbv >= 0 ? v : __builtin_inff();
            };
            float cl, cu;
            int nl, nu;
            cand(ri_prev, cl, nl);
            cand(up, cu, nu);
            // candidates in the reference's order: own value, the neighbour
            // along the row, the neighbour across rows, the jitter
            const bool wl = cl < sc;
            const float b1 = wl ? cl : sc;
            int ri = wl ? nl : sdi;
            const bool wu = cu < b1;
            ri = wu ? nu : ri;
            const float res = (wl || wu) ? (float)ri : sd;
            ri_prev = ri;
            buf_st32(rso, __float_as_uint(res), (uint32_t)lane * 4u, (uint32_t)k * 256u);
            __builtin_amdgcn_raw_buffer_store_b32((uint32_t)(ri + 1) | 0x80000000u, rse, (int)evoff, k * 4, 16);
            x8 += xd8;
            issue(k + PD, q);
        });
    }
    if (give_up && lane == 0) __hip_atomic_store(bs.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace

size_t refine_scratch_bytes(int w, int h) {
    const size_t n = (size_t)w * h, nw = n + kRefineSlack;
    size_t temp = 0;
    (void)hipcub::DeviceScan::ExclusiveScan(nullptr, temp, (const uint2*)nullptr, (uint2*)nullptr, PairSum(),
                                            make_uint2(0u, 0u), (int)nw);
    const size_t nbands = (size_t)(h + kBandRows - 1) / kBandRows;
    const size_t kb = (size_t)band_steps(w), slots = nbands * kb * kBandRows;
    auto r256 = [](size_t v) { return (v + 255) & ~(size_t)255; };
    // flag pairs and their exclusive sums (draw indices, accepted slots),
    // sweep results (image or band-major), accepted offsets, left and right
    // terms, the sweep records (diagonal or band-major), padded right terms,
    // engine base, edge values and the give-up flag, scan temp
    return r256(4 * nw * 4) + r256(std::max(n, slots) * 4) + r256(nw * 4) + r256(2 * n * 16) +
           r256(std::max(n * sizeof(SweepRec), slots * 32)) + r256((n + 2 * (size_t)kBandPadCols) * 8) + 256 +
           r256(((nbands + 1) * kb + 1) * 4) + r256(temp);
}

template <int PD>
using BandKernel = void (*)(const CostLut*, int, int, int, int, int, int, int, int, const float4*, const uint4*,
                            const uint2*, float*, BandSync);
template <int PD>
BandKernel<PD> band_kernel(int use_ad, int use_grad) {
    if (use_ad) return use_grad ? refine_band_kernel<PD, true, true> : refine_band_kernel<PD, true, false>;
    return use_grad ? refine_band_kernel<PD, false, true> : refine_band_kernel<PD, false, false>;
}

hipError_t launch_refine(hipStream_t st, const apg_hip_params& p, const CostLut* d_lut, int w, int h,
                         const RefineInputs& in, const int16_t* dmin, const int16_t* dmax, const uint8_t* reliable,
                         float* disp, const uint32_t* eng, size_t eng_len, void* scratch, size_t scratch_bytes,
                         const RefineBanding& bd, uint32_t* status, bool* banded_out) {
    const size_t n = (size_t)w * h, nw = n + kRefineSlack;
    if (h > kSweepThreads * kSweepRows || n >= (1u << 31)) return hipErrorInvalidValue;
    if (scratch_bytes < refine_scratch_bytes(w, h)) return hipErrorInvalidValue;
    const int iters = p.refine_iterations;
    if (!eng || !scratch || eng_len < (size_t)(iters + 1) * nw) return hipErrorInvalidValue;
    // engine offsets (base, acc, the clamp index) are 32-bit
    if (eng_len > 0xFFFFFFFFull) return hipErrorInvalidValue;
    const int nbands = (h + kBandRows - 1) / kBandRows;
    const int kb = band_steps(w);
    const size_t slots = (size_t)nbands * kb * kBandRows;
    auto r256 = [](size_t v) { return (v + 255) & ~(size_t)255; };
    uint8_t* sp = (uint8_t*)scratch;
    uint2* flag = (uint2*)sp;   // (draws, accepted) per sweep position
    uint2* pos = flag + nw;     // their exclusive sums: draw index, accepted slot
    sp += r256(4 * nw * 4);
    float* sres = (float*)sp; sp += r256(std::max(n, slots) * 4);
    uint32_t* acc = (uint32_t*)sp;
    sp += r256(nw * 4);
    uint4* lrec = (uint4*)sp;
    uint4* rrec = lrec + n;
    sp += r256(2 * n * 16);
    SweepRec* srec = (SweepRec*)sp;           // single-workgroup sweep: records in diagonal order
    float4* bra = (float4*)sp;                // banded sweep: band-major records
    uint4* brb = (uint4*)(bra + slots);
    sp += r256(std::max(n * sizeof(SweepRec), slots * 32));
    uint2* rt = (uint2*)sp; sp += r256((n + 2 * (size_t)kBandPadCols) * 8);
    uint32_t* base = (uint32_t*)sp; sp += 256;
    uint32_t* edge = (uint32_t*)sp;
    sp += r256(((size_t)(nbands + 1) * kb + 1) * 4);
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
    // the banded sweep: 32-bit census, D <= 256 (a row's window in LDS), the
    // entering columns within the padding, one band per CU at most (all bands
    // resident)
    const int D = p.max_disparity - p.min_disparity, d0 = p.min_disparity;
    int rs_log2 = 6;
    while ((1 << rs_log2) < D) ++rs_log2;
    const bool banded = bd.enable && status && sym && D > 0 && D <= 256 &&
                        std::abs(d0) + D + 2 * kBandSlack <= kBandPadCols &&
                        (n + 2 * (size_t)kBandPadCols) * 8 < ((size_t)1 << 31) &&
                        (size_t)(kb + kBandPD) * 1024 < ((size_t)1 << 31) && nbands <= bd.max_bands;
    const BandKernel<kBandPD> bk = band_kernel<kBandPD>(p.use_ad, p.use_grad);
    if (banded) {
        hipLaunchKernelGGL(refine_band_terms_kernel, dim3(gb), dim3(256), 0, st, (int)n, rrec, rt);
    }
    if (banded_out) *banded_out = banded;
    // a band runs whole prefetch spans: the last lane's last column is step w + 62
    const int krun = (w + kBandRows - 1 + kBandPD - 1) / kBandPD * kBandPD;
    for (int it = 0; it < iters; ++it) {
        const int fwd = (it % 2) == 0 ? 1 : 0;
        const int r = std::max(1, p.refine_random_radius >> it);
        const uint32_t s = 2u * (uint32_t)r + 1u;
        const uint32_t t = (uint32_t)(0u - s) % s;   // 2^32 mod s
        hipLaunchKernelGGL(refine_flags_kernel, dim3(gw), dim3(256), 0, st, w, h, fwd, reliable, dmin, dmax, disp,
                           eng, base, n, nw, s, t, flag);
        if ((e = hipcub::DeviceScan::ExclusiveScan(temp, temp_bytes, flag, pos, PairSum(), make_uint2(0u, 0u), (int)nw,
                                                   st)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(jitter_compact_kernel, dim3(gw), dim3(256), 0, st, flag, pos, nw, acc);
        if (banded) {
            hipLaunchKernelGGL(refine_band_static_kernel, dim3((kb + kStaticDiag - 1) / kStaticDiag, nbands, kBandRows * kStaticDiag / 256),
                               dim3(256), 0, st, d_lut, w, h, fwd, r, kb, p.use_ad, p.use_grad, lrec, rrec, reliable,
                               dmin, dmax, disp, eng, (uint32_t)(eng_len - 1), base, acc, pos, bra, brb);
            if ((e = launch_fill_u32(st, (size_t)(nbands + 1) * kb, 0u, edge)) != hipSuccess) return e;
            const BandSync bs{edge, status, bd.wait_ticks};
            hipLaunchKernelGGL(bk, dim3(nbands), dim3(kBandRows), 0, st, d_lut, w, h, fwd, d0, D, rs_log2, kb, krun,
                               bra, brb, rt, sres, bs);
            hipLaunchKernelGGL(refine_band_unpack_kernel, dim3(gb), dim3(256), 0, st, w, h, fwd, kb, sres, disp);
        } else {
            hipLaunchKernelGGL(refine_static_kernel, dim3(gb), dim3(256), 0, st, d_lut, w, h, fwd, r, sym, p.use_ad,
                               p.use_grad, lrec, rrec, reliable, dmin, dmax, disp, eng, (uint32_t)(eng_len - 1), base,
                               acc, pos, srec);
            hipLaunchKernelGGL(refine_sweep_kernel, dim3(1), dim3(kSweepThreads), shm, st, d_lut, w, h, fwd, sym,
                               p.use_ad, p.use_grad, srec, rrec, sres);
            hipLaunchKernelGGL(refine_unpack_kernel, dim3(gb), dim3(256), 0, st, w, h, fwd, sres, disp);
        }
        hipLaunchKernelGGL(jitter_advance_kernel, dim3(1), dim3(1), 0, st, base, flag, pos, n, acc);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace apgk
