// gfx950 kernels: Cross H and V aggregation (integer windowed means via packed
// (sum << 8 | count) prefix sums) and the SGM path scans (one wavefront per
// scanline, disparities across lanes as packed u16 pairs, DPP neighbour
// exchange and wave-min reduction).
#include <type_traits>

#include "device_common.hpp"
#include "kernels.hpp"

namespace apgk {

// Exact floor(n / c) for 1 <= c <= 2*kMaxArm+1 and n <= (2*kMaxArm+1)*65535:
// m = ceil(2^32 / c), q = umulhi(n, m) is exact while n*(m*c - 2^32) < 2^32.
__device__ __forceinline__ uint32_t div_small(uint32_t n, uint32_t c, uint32_t m) {
    return c == 1 ? n : __umulhi(n, m);
}
__device__ __forceinline__ uint32_t magic_for(uint32_t c) {
    return c <= 1 ? 0u : (uint32_t)((0x100000000ull + c - 1) / c);
}

constexpr int CH_TX = 64;

// Horizontal pass of CostAggregator::aggregate_hv (reference
// src/cost_aggregator.cpp:74-105): T(x,d) = floor(sum / count) over the valid
// C(xx,d), xx in [x-L, x+R], d inside xx's range; 65535 when d is outside
// x's range or count == 0.  Block = one row segment of CH_TX pixels and one
// 64-wide d chunk (lane = d); the row's packed prefix sums over the segment
// plus arm halo live in LDS.
__global__ __launch_bounds__(256) void cross_h_kernel(int w, int h, int d0, int D, int Lmax,
                                                      const int16_t* __restrict__ dmin, const int16_t* __restrict__ dmax,
                                                      const uint32_t* __restrict__ arms,
                                                      const uint16_t* __restrict__ cost, uint16_t* __restrict__ tmp) {
    extern __shared__ uint32_t sh[];
    uint32_t* magic = sh;                 // [2*kMaxArm + 2]
    uint32_t* P = sh + 2 * kMaxArm + 2;   // [(NCOL + 1) * 64]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int y = blockIdx.y;
    const int x0 = blockIdx.x * CH_TX;
    const int dc = blockIdx.z * 64;
    const int di = dc + lane;
    const int d = d0 + di;
    const int base = x0 - Lmax;
    const int NCOL = CH_TX + 2 * Lmax;
    for (int c = threadIdx.x; c < 2 * kMaxArm + 2; c += 256) magic[c] = magic_for((uint32_t)c);
    const int seg = (NCOL + 3) / 4;
    const int c_beg = wave * seg, c_end = min(NCOL, c_beg + seg);
    uint32_t acc = 0;
    if (wave == 0) P[lane] = 0;
    // columns in batches of CH_B: every load is issued (clamped inside the
    // image and the chunk), so a batch waits once
    constexpr int CH_B = 8;
    const int dic = min(di, D - 1);
    for (int col0 = c_beg; col0 < c_end; col0 += CH_B) {
        uint32_t cv[CH_B];
        int lo[CH_B], hi[CH_B];
#pragma unroll
        for (int i = 0; i < CH_B; ++i) {
            const int xc = min(max(base + col0 + i, 0), w - 1);
            const size_t q = (size_t)y * w + xc;
            cv[i] = cost[q * D + dic];
            lo[i] = dmin[q];
            hi[i] = dmax[q];
        }
#pragma unroll
        for (int i = 0; i < CH_B; ++i) {
            const int col = col0 + i;
            if (col >= c_end) break;
            const int xx = base + col;
            const bool ok = xx >= 0 && xx < w && di < D && d >= lo[i] && d < hi[i] && cv[i] != kInv16;
            acc += ok ? ((cv[i] << 8) | 1u) : 0u;
            P[(col + 1) * 64 + lane] = acc;
        }
    }
    __syncthreads();
    uint32_t off = 0;
    for (int ww = 0; ww < wave; ++ww) {
        const int e = min(NCOL, (ww + 1) * seg);
        if (e > ww * seg) off += P[e * 64 + lane];
    }
    __syncthreads();
    if (off)
        for (int col = c_beg; col < c_end; ++col) P[(col + 1) * 64 + lane] += off;
    __syncthreads();
    if (di >= D) return;
    for (int px = wave; px < CH_TX; px += 4) {
        const int x = x0 + px;
        if (x >= w) break;
        const size_t q = (size_t)y * w + x;
        const uint32_t a = arms[q];
        const int L = a & 0xFF, R = (a >> 8) & 0xFF;
        uint16_t out = kInv16;
        if (d >= dmin[q] && d < dmax[q]) {
            const uint32_t s = P[(x + R + 1 - base) * 64 + lane] - P[(x - L - base) * 64 + lane];
            const uint32_t cnt = s & 0xFF;
            if (cnt) out = (uint16_t)div_small(s >> 8, cnt, magic[cnt]);
        }
        tmp[q * D + di] = out;
    }
}

hipError_t launch_cross_h(hipStream_t s, const apg_hip_params& p, int w, int h, const int16_t* dmin,
                          const int16_t* dmax, const uint32_t* arms, const uint16_t* cost, uint16_t* tmp) {
    const int D = p.max_disparity - p.min_disparity;
    const int Lmax = p.max_arm_length > 1 ? p.max_arm_length : 1;
    const size_t shm = (2 * kMaxArm + 2 + (size_t)(CH_TX + 2 * Lmax + 1) * 64) * sizeof(uint32_t);
    dim3 grid((w + CH_TX - 1) / CH_TX, h, (D + 63) / 64);
    hipLaunchKernelGGL(cross_h_kernel, grid, dim3(256), shm, s, w, h, p.min_disparity, D, Lmax, dmin, dmax, arms,
                       cost, tmp);
    return hipGetLastError();
}

constexpr int CV_PF = 16;   // rows fetched per batch

// Vertical pass of CostAggregator::aggregate_hv (reference
// src/cost_aggregator.cpp:107-138): same windowed mean over y in [y-U, y+D].
// One wavefront per (column, 64-wide d chunk) slides down the column keeping
// the running packed prefix P(k) = sum over rows < k in an LDS ring of
// RS >= 2*Lmax + 2 entries; rows are fetched CV_PF at a time.
template <typename T>
__global__ __launch_bounds__(64) void cross_v_kernel(int w, int h, int d0, int D, int Lmax, int rs_mask,
                                                     const int16_t* __restrict__ dmin, const int16_t* __restrict__ dmax,
                                                     const uint32_t* __restrict__ arms,
                                                     const T* __restrict__ tmp, T* __restrict__ cost) {
    constexpr T INV = (T)~(T)0;
    extern __shared__ uint32_t sh[];
    uint32_t* magic = sh;                  // [2*kMaxArm + 2]
    uint32_t* ring = sh + 2 * kMaxArm + 2; // [RS * 64]
    const int lane = threadIdx.x;
    const int x = blockIdx.x;
    const int di = blockIdx.y * 64 + lane;
    const int d = d0 + di;
    const bool act = di < D;
    for (int c = lane; c < 2 * kMaxArm + 2; c += 64) magic[c] = magic_for((uint32_t)c);
    ring[lane] = 0;
    uint32_t acc = 0;
    const int total = h + Lmax;   // fill index k runs over [0, h), output row y = k - Lmax
    for (int kb = 0; kb < total; kb += CV_PF) {
        uint32_t v[CV_PF];
#pragma unroll
        for (int i = 0; i < CV_PF; ++i) {
            const int k = kb + i;
            v[i] = 0;
            if (k < h && act) {
                const size_t q = (size_t)k * w + x;
                const T c = tmp[q * D + di];
                if (c != INV && d >= dmin[q] && d < dmax[q]) v[i] = ((uint32_t)c << 8) | 1u;
            }
        }
#pragma unroll
        for (int i = 0; i < CV_PF; ++i) {
            const int k = kb + i;
            if (k < h) {
                acc += v[i];
                ring[((k + 1) & rs_mask) * 64 + lane] = acc;
            }
            const int y = k - Lmax;
            if (y >= 0 && y < h && act) {
                const size_t q = (size_t)y * w + x;
                const uint32_t a = arms[q];
                const int U = (a >> 16) & 0xFF, Dn = a >> 24;
                T out = INV;
                if (d >= dmin[q] && d < dmax[q]) {
                    const uint32_t s = ring[((y + Dn + 1) & rs_mask) * 64 + lane] - ring[((y - U) & rs_mask) * 64 + lane];
                    const uint32_t cnt = s & 0xFF;
                    if (cnt) out = (T)div_small(s >> 8, cnt, magic[cnt]);
                }
                cost[q * D + di] = out;
            }
        }
    }
}

constexpr int CV2_PF = 16;

// Vertical Cross pass, sliding form: one wavefront per (column, 128-wide d
// chunk), two disparities per lane, packed prefix sums of the last
// RS >= 2*Lmax+2 rows in an LDS ring, rows fetched CV2_PF at a time with the
// next batch in flight while the current one is consumed.  T entries outside
// the pixel's range are already invalid (written so by the H pass), so only
// the output side applies the range.  D must be even when T is u16.
template <typename T>
__global__ __launch_bounds__(64) void cross_v2_kernel(int w, int h, int d0, int D, int Lmax, int rs_mask,
                                                      const int16_t* __restrict__ dmin, const int16_t* __restrict__ dmax,
                                                      const uint32_t* __restrict__ arms, const T* __restrict__ tmp,
                                                      T* __restrict__ cost) {
    constexpr uint32_t TINV = (T)~(T)0;
    extern __shared__ uint32_t sh[];
    uint32_t* magic = sh;                   // [2*kMaxArm + 2]
    uint32_t* ring = sh + 2 * kMaxArm + 2;  // [RS][2][64]
    const int lane = threadIdx.x;
    const int x = blockIdx.x;
    const int di = blockIdx.y * 128 + 2 * lane;   // first of the lane's two disparities
    const int dA = d0 + di, dB = dA + 1;
    for (int c = lane; c < 2 * kMaxArm + 2; c += 64) magic[c] = magic_for((uint32_t)c);
    ring[lane] = 0;
    ring[64 + lane] = 0;
    uint32_t accA = 0, accB = 0;
    const int total = h + Lmax;
    // Raw words are prefetched unconditionally (row index clamped) and decoded
    // at use, so the next batch stays in flight while this one is consumed.
    // D is even (launcher): one aligned word per lane; lanes past D read pair 0.
    const int dl0 = di < D ? di : 0;
    // Per batch: the metadata (arms, range) of the batch's output rows, one
    // row per lane, then the T words of its input rows; issued in the order
    // they are consumed so a wait for the former never waits for the latter.
    uint32_t nxt[CV2_PF], nxt_arm, nxt_rng;
    auto fetch = [&](int kb) {
        const int yl = min(max(kb - Lmax + lane, 0), h - 1);
        nxt_arm = arms[(size_t)yl * w + x];
        nxt_rng = (uint32_t)(uint16_t)dmin[(size_t)yl * w + x] | ((uint32_t)(uint16_t)dmax[(size_t)yl * w + x] << 16);
#pragma unroll
        for (int i = 0; i < CV2_PF; ++i) {
            const int k = min(kb + i, h - 1);
            const T* row = tmp + ((size_t)k * w + x) * D;
            if (sizeof(T) == 2) nxt[i] = *(const uint32_t*)(row + dl0);
            else nxt[i] = *(const uint16_t*)((const uint8_t*)row + dl0);
        }
    };
    auto decode = [&](uint32_t raw) -> uint32_t {   // -> u16 halves, invalid 65535
        uint32_t a, b;
        if (sizeof(T) == 2) { a = raw & 0xFFFFu; b = raw >> 16; }
        else { a = raw & 0xFFu; b = (raw >> 8) & 0xFFu; a = a == 0xFFu ? 0xFFFFu : a; b = b == 0xFFu ? 0xFFFFu : b; }
        if (di >= D) a = 0xFFFFu;
        if (di + 1 >= D) b = 0xFFFFu;
        return a | (b << 16);
    };
    fetch(0);
    for (int kb = 0; kb < total; kb += CV2_PF) {
        uint32_t cur[CV2_PF];
#pragma unroll
        for (int i = 0; i < CV2_PF; ++i) cur[i] = nxt[i];
        const uint32_t meta_arm = nxt_arm, meta_rng = nxt_rng;
        fetch(kb + CV2_PF);
#pragma unroll
        for (int i = 0; i < CV2_PF; ++i) {
            const int k = kb + i;
            if (k < h) {
                const uint32_t c = decode(cur[i]);
                const uint32_t ca = c & 0xFFFFu, cb = c >> 16;
                accA += ca != 0xFFFFu ? (ca << 8) | 1u : 0u;
                accB += cb != 0xFFFFu ? (cb << 8) | 1u : 0u;
                const int slot = ((k + 1) & rs_mask) * 128;
                ring[slot + lane] = accA;
                ring[slot + 64 + lane] = accB;
            }
            const int y = k - Lmax;
            if (y >= 0 && y < h) {
                const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)meta_arm, i);
                const uint32_t rg = (uint32_t)__builtin_amdgcn_readlane((int)meta_rng, i);
                const int lo = (int16_t)(rg & 0xFFFF), hi = (int16_t)(rg >> 16);
                const int U = (a >> 16) & 0xFF, Dn = a >> 24;
                const int s1 = ((y + Dn + 1) & rs_mask) * 128, s0 = ((y - U) & rs_mask) * 128;
                const uint32_t sa = ring[s1 + lane] - ring[s0 + lane];
                const uint32_t sb = ring[s1 + 64 + lane] - ring[s0 + 64 + lane];
                T oa = (T)TINV, ob = (T)TINV;
                if (dA >= lo && dA < hi && (sa & 0xFF)) oa = (T)div_small(sa >> 8, sa & 0xFF, magic[sa & 0xFF]);
                if (dB >= lo && dB < hi && (sb & 0xFF)) ob = (T)div_small(sb >> 8, sb & 0xFF, magic[sb & 0xFF]);
                T* o = cost + ((size_t)y * w + x) * D + di;
                if (di < D) {
                    if (sizeof(T) == 1) *(uint16_t*)o = (uint16_t)((uint32_t)oa | ((uint32_t)ob << 8));
                    else *(uint32_t*)o = (uint32_t)oa | ((uint32_t)ob << 16);
                }
            }
        }
    }
}

// Vertical Cross pass over u8 T for ranges that depend on the column only
// (the dense range init: [min_disparity, min(max_disparity, x+1))) and
// min_disparity >= 0.  Then T(x, y, d) is valid iff d lies in column x's
// range (the pixel itself is in its own H window and x - d < W), so the
// count of valid entries in a vertical window is the window length
// U + Dn + 1 or 0, and only sums need prefixes: u16 pairs (window sums
// < 125 * 255 < 2^15, exact modulo 2^16).  Invalid entries (255) join the
// sums of d outside the range and are masked at the output.
// Grid: (column, row segment, 128-wide d chunk), one wavefront each; the
// prefix of input rows [y0 - Lmax, y1 + Lmax) lives in an LDS ring of
// 2*Lmax + 2 rows.  floor(s / n) = umulhi(2s, ceil(2^31 / n)), exact for
// s * n < 2^31.
// rows per prefetch batch; at 1280x720 D=128: 16: 77 us, 12: 84 us, 8: 83 us with 90-row
// segments (with 180-row segments 16: 83 us, 8: 87 us, 32: 98 us)
constexpr int CV3_PF = 16;

// TWO (D <= 64): two columns per wave, a half-wave (32 lanes, 2 d each) per
// column, so no lane idles; each column has its own window metadata.
template <bool TWO>
__device__ __forceinline__ void cross_v3_body(const uint3 bid, int w, int h, int d0, int D, int Lmax, int seg,
                                              const int16_t* __restrict__ dmin, const int16_t* __restrict__ dmax,
                                              const uint32_t* __restrict__ arms, const uint8_t* __restrict__ tmp,
                                              uint8_t* __restrict__ cost) {
    extern __shared__ uint32_t sh[];
    const int lane = threadIdx.x;
    constexpr int NC = TWO ? 2 : 1;              // columns per wave
    const int half = TWO ? lane >> 5 : 0;        // the lane's column of the wave
    const int xw = bid.x * NC;                   // the wave's first column
    const int x = xw + half;                     // the lane's column
    const int y0 = bid.y * seg, y1 = min(h, y0 + seg);
    if (y0 >= h) return;
    const int di = bid.z * 128 + 2 * (TWO ? lane & 31 : lane);
    const int rs = 2 * Lmax + 2;
    const int ya = max(0, y0 - Lmax), yb = min(h, y1 + Lmax), kend = y1 + Lmax, nk = kend - ya;
    uint32_t* ring = sh;                       // [rs][64] u16-pair prefix sums
    uint4* meta = (uint4*)(sh + rs * 64);      // [NC][nk] per step: write slot, window slots, magic
    // step k (input row k, output row y = k - Lmax): byte offsets of the ring
    // rows P(k + 1), P(y + Dn + 1), P(y - U) and ceil(2^31 / (U + Dn + 1))
    for (int rr = lane; rr < NC * nk; rr += 64) {
        const int c = TWO ? rr / nk : 0, r = rr - c * nk;
        const int k = ya + r, y = k - Lmax;
        uint4 m;
        m.x = (uint32_t)((r + 1) % rs) * 256u;
        m.y = m.z = m.w = 0u;
        if (y >= y0 && xw + c < w) {
            const uint32_t a = arms[(size_t)y * w + xw + c];
            const int U = (a >> 16) & 0xFF, Dn = a >> 24;
            m.y = (uint32_t)((y + Dn + 1 - ya) % rs) * 256u;
            m.z = (uint32_t)((y - U - ya) % rs) * 256u;
            const uint32_t n = (uint32_t)(U + Dn + 1);
            m.w = (0x80000000u + n - 1u) / n;
        }
        meta[rr] = m;
    }
    if (TWO) meta += half * nk;   // the lane's column's metadata
    ring[lane] = 0u;   // P(ya)
    __syncthreads();
    const int xc = min(x, w - 1);
    const int lo = dmin[(size_t)y0 * w + xc], hi = dmax[(size_t)y0 * w + xc];
    const int dA = d0 + di;
    const bool inA = di < D && x < w;   // D even: the pair is inside or outside as a whole
    const uint32_t omask = ((inA && dA >= lo && dA < hi) ? 0u : 0xFFu) |
                           ((inA && dA + 1 >= lo && dA + 1 < hi) ? 0u : 0xFF00u);
    // byte offset of the lane's pair from the wave's first column
    const uint32_t voff = inA ? (uint32_t)(di + half * D) : 0u;
    // stores of lanes outside the volume go past the buffer's records, where
    // the hardware drops them: no lane branches around a store
    const uint32_t soff = inA ? voff : 0x80000000u;
    const uint32_t lofs = 4u * (uint32_t)lane;
    const auto rt = raw_rsrc(tmp);
    const auto rc = raw_rsrc(cost);
    const uint32_t rowb = (uint32_t)w * (uint32_t)D, colb = (uint32_t)xw * (uint32_t)D;
    const int kout = y0 + Lmax;   // first step with an output row
    uint32_t acc = 0;
    auto step = [&](int k, uint32_t raw, uint4 m) {
        acc = as_u32(as_v2(acc) + as_v2(bytes_to_halves(raw)));
        *(uint32_t*)((uint8_t*)ring + m.x + lofs) = acc;
        if (k < kout) return;
        const uint32_t p1 = *(const uint32_t*)((const uint8_t*)ring + m.y + lofs);
        const uint32_t p0 = *(const uint32_t*)((const uint8_t*)ring + m.z + lofs);
        const uint32_t s = as_u32(as_v2(p1) - as_v2(p0));
        const uint32_t s2 = as_u32(as_v2(s) + as_v2(s));
        const uint32_t qa = __umulhi(s2 & 0xFFFFu, m.w), qb = __umulhi(s2 >> 16, m.w);
        const uint32_t out = (qa | (qb << 8)) | omask;
        buf_st16(rc, out, soff, (uint32_t)(k - Lmax) * rowb + colb);
    };
    uint32_t X[CV3_PF], Y[CV3_PF];
    // a batch's rows past the image bottom reload its last row (their prefix
    // entries are never read); whole batches inside it skip the clamps
    auto load = [&](uint32_t (&b)[CV3_PF], int kb) {
        if (kb + CV3_PF <= yb) {
#pragma unroll
            for (int i = 0; i < CV3_PF; ++i) b[i] = buf_ld16(rt, voff, (uint32_t)(kb + i) * rowb + colb);
        } else {
#pragma unroll
            for (int i = 0; i < CV3_PF; ++i) b[i] = buf_ld16(rt, voff, (uint32_t)min(kb + i, yb - 1) * rowb + colb);
        }
    };
    // OUT 0: no step of the batch has an output row; 1: every step has one and
    // lies before kend; 2: either may vary (checked per step)
    auto batch = [&](const uint32_t (&b)[CV3_PF], int kb, auto out_t) {
        constexpr int OUT = decltype(out_t)::value;
        uint4 m[CV3_PF];
        if (OUT == 2) {
#pragma unroll
            for (int i = 0; i < CV3_PF; ++i) m[i] = meta[min(kb + i - ya, nk - 1)];
        } else {
            const uint4* mb = meta + (kb - ya);
#pragma unroll
            for (int i = 0; i < CV3_PF; ++i) m[i] = mb[i];
        }
#pragma unroll
        for (int i = 0; i < CV3_PF; ++i) {
            if (OUT == 2) {
                if (kb + i < kend) step(kb + i, b[i], m[i]);
            } else {
                acc = as_u32(as_v2(acc) + as_v2(bytes_to_halves(b[i])));
                *(uint32_t*)((uint8_t*)ring + m[i].x + lofs) = acc;
                if (OUT == 1) {
                    const uint32_t p1 = *(const uint32_t*)((const uint8_t*)ring + m[i].y + lofs);
                    const uint32_t p0 = *(const uint32_t*)((const uint8_t*)ring + m[i].z + lofs);
                    const uint32_t sd = as_u32(as_v2(p1) - as_v2(p0));
                    const uint32_t s2 = as_u32(as_v2(sd) + as_v2(sd));
                    const uint32_t qa = __umulhi(s2 & 0xFFFFu, m[i].w), qb = __umulhi(s2 >> 16, m[i].w);
                    buf_st16(rc, __builtin_amdgcn_perm(qb, qa, 0x0c0c0400u) | omask, soff,
                                 (uint32_t)(kb + i - Lmax) * rowb + colb);
                }
            }
        }
    };
    auto run = [&](const uint32_t (&b)[CV3_PF], int kb) {
        if (kb + CV3_PF <= kout) batch(b, kb, std::integral_constant<int, 0>{});
        else if (kb >= kout && kb + CV3_PF <= kend) batch(b, kb, std::integral_constant<int, 1>{});
        else batch(b, kb, std::integral_constant<int, 2>{});
    };
    load(X, ya);
    load(Y, ya + CV3_PF);
    for (int kb = ya; kb < kend; kb += 2 * CV3_PF) {
        run(X, kb);
        load(X, kb + 2 * CV3_PF);
        run(Y, kb + CV3_PF);
        load(Y, kb + 3 * CV3_PF);
    }
}

template <bool TWO>
__global__ __launch_bounds__(64) void cross_v3_kernel(int w, int h, int d0, int D, int Lmax, int seg,
                                                      const int16_t* __restrict__ dmin,
                                                      const int16_t* __restrict__ dmax,
                                                      const uint32_t* __restrict__ arms,
                                                      const uint8_t* __restrict__ tmp, uint8_t* __restrict__ cost) {
    cross_v3_body<TWO>(block3(), w, h, d0, D, Lmax, seg, dmin, dmax, arms, tmp, cost);
}

template <bool TWO>
struct CrossV3Fn {
    static constexpr int kThreads = 64;
    template <typename... A>
    __device__ static void run(const uint3 b, const A&... a) { cross_v3_body<TWO>(b, a...); }
};

// Vertical Cross pass over u8 T for per-pixel ranges (the prior's): the grid
// of cross_v3_kernel (column, row segment, 128-wide d chunk; the prefix of
// input rows [y0 - Lmax, y1 + Lmax) in LDS rings of 2*Lmax + 2 rows).  Since
// the number of valid entries of a window varies with d, a second ring keeps
// u8-pair prefix counts beside the doubled u16-pair prefix sums (a segment
// has at most 255 input rows, so a count byte never carries; window sums stay
// below 125 * 254 < 2^15, so doubled prefixes are exact modulo 2^16).  Per
// step the ring slots it touches and the output pixel's range come from a
// table built once per segment.  Rows past the image bottom add to prefixes
// no output reads (a window ends at row h - 1 at most).
constexpr int CV4_PF = 16;

__global__ __launch_bounds__(64) void cross_v4_kernel(int w, int h, int d0, int D, int Lmax, int seg,
                                                      const int16_t* __restrict__ dmin,
                                                      const int16_t* __restrict__ dmax,
                                                      const uint32_t* __restrict__ arms,
                                                      const uint8_t* __restrict__ tmp, uint8_t* __restrict__ cost) {
    extern __shared__ uint32_t sh[];
    const int lane = threadIdx.x;
    const int x = blockIdx.x;
    const int y0 = blockIdx.y * seg, y1 = min(h, y0 + seg);
    if (y0 >= h) return;
    const int di = blockIdx.z * 128 + 2 * lane;
    const int rs = 2 * Lmax + 2;
    const int ya = max(0, y0 - Lmax), yb = min(h, y1 + Lmax), kend = y1 + Lmax, nk = kend - ya;
    uint32_t* mag = sh;                                // [2*kMaxArm + 2] ceil(2^31 / n)
    uint32_t* ring = sh + 2 * kMaxArm + 2;             // [rs][64]: doubled u16-pair prefix sums
    uint16_t* cring = (uint16_t*)(ring + rs * 64);     // [rs][64]: u8-pair prefix valid counts
    uint2* meta = (uint2*)(cring + ((rs * 64 + 3) & ~3));   // [nk]: slots P(k+1) | P(y+Dn+1) << 8 | P(y-U) << 16, range
    for (int c = lane; c < 2 * kMaxArm + 2; c += 64) mag[c] = c ? (0x80000000u + (uint32_t)c - 1u) / (uint32_t)c : 0u;
    for (int r = lane; r < nk; r += 64) {
        const int k = ya + r, y = k - Lmax;
        uint2 m;
        m.x = (uint32_t)((r + 1) % rs);
        m.y = 0u;
        if (y >= y0) {
            const size_t q = (size_t)y * w + x;
            const uint32_t a = arms[q];
            const int U = (a >> 16) & 0xFF, Dn = a >> 24;
            m.x |= (uint32_t)((y + Dn + 1 - ya) % rs) << 8 | (uint32_t)((y - U - ya) % rs) << 16;
            m.y = (uint32_t)(uint16_t)dmin[q] | ((uint32_t)(uint16_t)dmax[q] << 16);
        }
        meta[r] = m;
    }
    ring[lane] = 0u;   // P(ya)
    cring[lane] = 0u;
    __syncthreads();
    const int dA = d0 + di;
    const bool inA = di < D;   // D even: the pair is inside or outside as a whole
    const uint32_t voff = inA ? (uint32_t)di : 0u;
    const uint32_t soff = inA ? voff : 0x80000000u;   // past the records: the store is dropped
    const auto rt = raw_rsrc(tmp);
    const auto rc = raw_rsrc(cost);
    const uint32_t rowb = (uint32_t)w * (uint32_t)D, colb = (uint32_t)x * (uint32_t)D;
    const int kout = y0 + Lmax;   // first step with an output row
    uint32_t acc = 0, cacc = 0;
    // kt decides whether the step has an output row (kt >= kout); k is its input row
    auto step = [&](int kt, uint32_t raw, uint2 m, int k) {
        const bool va = inA && (raw & 0xFFu) != 0xFFu, vb = inA && (raw & 0xFF00u) != 0xFF00u;
        // the two T bytes as u16 halves, the invalid ones zeroed; acc += 2 * them
        const uint32_t pr = bytes_to_halves(raw) & ((va ? 0xFFFFu : 0u) | (vb ? 0xFFFF0000u : 0u));
        {
            uint32_t nacc;
            asm("v_pk_mad_u16 %0, %1, %2, %3" : "=v"(nacc) : "v"(pr), "s"(0x00020002u), "v"(acc));
            acc = nacc;
        }
        cacc += (va ? 1u : 0u) | (vb ? 0x100u : 0u);
        ring[(m.x & 0xFFu) * 64 + lane] = acc;
        cring[(m.x & 0xFFu) * 64 + lane] = (uint16_t)cacc;
        if (kt < kout) return;
        const uint32_t s1 = (m.x >> 8) & 0xFFu, s0 = m.x >> 16;
        const uint32_t sw = as_u32(as_v2(ring[s1 * 64 + lane]) - as_v2(ring[s0 * 64 + lane]));
        const uint32_t wc = (uint32_t)cring[s1 * 64 + lane] - (uint32_t)cring[s0 * 64 + lane];
        const int lo = (int16_t)(m.y & 0xFFFFu), hi = (int16_t)(m.y >> 16);
        const uint32_t ca = wc & 0xFFu, cb = (wc >> 8) & 0xFFu;
        // both magic numbers are read whatever the range test says (a count
        // is at most the window, 2 Lmax + 1, and mag[0] = 0), so the test is a
        // select and no lane branches
        const uint32_t qa = __umulhi(sw & 0xFFFFu, mag[ca]), qb = __umulhi(sw >> 16, mag[cb]);
        const uint32_t oa = (dA >= lo && dA < hi && ca) ? qa : 0xFFu;
        const uint32_t ob = (dA + 1 >= lo && dA + 1 < hi && cb) ? qb : 0xFFu;
        buf_st16(rc, __builtin_amdgcn_perm(ob, oa, 0x0c0c0400u), soff, (uint32_t)(k - Lmax) * rowb + colb);
    };
    uint32_t X[CV4_PF], Y[CV4_PF];
    // as in cross_v3_kernel: unclamped loads for whole batches inside the
    // image, and batches classified once by their output rows
    auto load = [&](uint32_t (&bf)[CV4_PF], int kb) {
        if (kb + CV4_PF <= yb) {
#pragma unroll
            for (int i = 0; i < CV4_PF; ++i) bf[i] = buf_ld16(rt, voff, (uint32_t)(kb + i) * rowb + colb);
        } else {
#pragma unroll
            for (int i = 0; i < CV4_PF; ++i) bf[i] = buf_ld16(rt, voff, (uint32_t)min(kb + i, yb - 1) * rowb + colb);
        }
    };
    auto batch = [&](const uint32_t (&bf)[CV4_PF], int kb, auto out_t) {
        constexpr int OUT = decltype(out_t)::value;   // 0 none, 1 every step, 2 checked
        uint2 m[CV4_PF];
        if (OUT == 2) {
#pragma unroll
            for (int i = 0; i < CV4_PF; ++i) m[i] = meta[min(kb + i - ya, nk - 1)];
        } else {
            const uint2* mb = meta + (kb - ya);
#pragma unroll
            for (int i = 0; i < CV4_PF; ++i) m[i] = mb[i];
        }
#pragma unroll
        for (int i = 0; i < CV4_PF; ++i) {
            if (OUT == 2) {
                if (kb + i < kend) step(kb + i, bf[i], m[i], kb + i);
            } else {
                step(OUT == 1 ? kout : kout - 1, bf[i], m[i], kb + i);
            }
        }
    };
    auto run = [&](const uint32_t (&bf)[CV4_PF], int kb) {
        if (kb + CV4_PF <= kout) batch(bf, kb, std::integral_constant<int, 0>{});
        else if (kb >= kout && kb + CV4_PF <= kend) batch(bf, kb, std::integral_constant<int, 1>{});
        else batch(bf, kb, std::integral_constant<int, 2>{});
    };
    load(X, ya);
    load(Y, ya + CV4_PF);
    for (int kb = ya; kb < kend; kb += 2 * CV4_PF) {
        run(X, kb);
        load(X, kb + 2 * CV4_PF);
        run(Y, kb + CV4_PF);
        load(Y, kb + 3 * CV4_PF);
    }
}

hipError_t launch_cross_v_t(hipStream_t s, const apg_hip_params& p, int w, int h, const int16_t* dmin,
                            const int16_t* dmax, const uint32_t* arms, const void* tmp, void* cost, bool u8,
                            bool column_ranges, const KTune& tn, LaunchSink* sink) {
    {
        const int D = p.max_disparity - p.min_disparity;
        const int Lmax = p.max_arm_length > 1 ? p.max_arm_length : 1;
        if (u8 && column_ranges && p.min_disparity >= 0 && D % 2 == 0 && (size_t)w * h * D < (1ull << 31)) {
            const int dch = (D + 127) / 128;
            // D <= 64: two columns per wave (a half-wave each)
            const bool two = D <= 64;
            const int nxw = two ? (w + 1) / 2 : w;   // waves across a row
            // 1280x720 D=128, arm 17: 8 segments of 90 rows take 77 us, 4 of 180 rows 83 us,
            // 2 of 360 rows 117 us, 12 of 60 rows 83 us
            const int target = tn.cross_v_waves > 0 ? tn.cross_v_waves : 10240;
            int nseg = (target + nxw * dch - 1) / (nxw * dch);
            nseg = max(1, min(nseg, (h + 63) / 64));
            const int seg = (h + nseg - 1) / nseg;
            nseg = (h + seg - 1) / seg;
            const size_t shm =
                (size_t)(2 * Lmax + 2) * 64 * sizeof(uint32_t) + (size_t)(two ? 2 : 1) * (seg + 2 * Lmax) * 16;
            const dim3 grid(nxw, nseg, dch);
            if (sink)
                return two ? sink->add<CrossV3Fn<true>>(grid, shm, w, h, p.min_disparity, D, Lmax, seg, dmin, dmax,
                                                         arms, (const uint8_t*)tmp, (uint8_t*)cost)
                           : sink->add<CrossV3Fn<false>>(grid, shm, w, h, p.min_disparity, D, Lmax, seg, dmin, dmax,
                                                          arms, (const uint8_t*)tmp, (uint8_t*)cost);
            if (two)
                hipLaunchKernelGGL(cross_v3_kernel<true>, grid, dim3(64), shm, s, w, h, p.min_disparity, D, Lmax, seg,
                                   dmin, dmax, arms, (const uint8_t*)tmp, (uint8_t*)cost);
            else
                hipLaunchKernelGGL(cross_v3_kernel<false>, grid, dim3(64), shm, s, w, h, p.min_disparity, D, Lmax,
                                   seg, dmin, dmax, arms, (const uint8_t*)tmp, (uint8_t*)cost);
            return hipGetLastError();
        }
    }
    if (sink) return hipErrorNotSupported;   // the other forms have no group launch
    {
        // per-pixel ranges: the segmented grid with counted prefixes
        const int D = p.max_disparity - p.min_disparity;
        const int Lmax = p.max_arm_length > 1 ? p.max_arm_length : 1;
        if (u8 && D % 2 == 0 && (size_t)w * h * D < (1ull << 31)) {
            const int dch = (D + 127) / 128;
            const int target = tn.cross_v_waves > 0 ? tn.cross_v_waves : 16384;
            int nseg = (target + w * dch - 1) / (w * dch);
            nseg = max(1, min(nseg, (h + 63) / 64));
            // a segment's input rows (its own plus the arm halo) fit a count byte
            nseg = max(nseg, (h + (255 - 2 * Lmax) - 1) / (255 - 2 * Lmax));
            const int seg = (h + nseg - 1) / nseg;
            nseg = (h + seg - 1) / seg;
            const size_t rs = (size_t)(2 * Lmax + 2);
            const size_t shm = (2 * kMaxArm + 2) * sizeof(uint32_t) + rs * 64 * 4 + ((rs * 64 + 3) & ~(size_t)3) * 2 +
                               (size_t)(seg + 2 * Lmax) * 8;
            hipLaunchKernelGGL(cross_v4_kernel, dim3(w, nseg, dch), dim3(64), shm, s, w, h, p.min_disparity, D, Lmax,
                               seg, dmin, dmax, arms, (const uint8_t*)tmp, (uint8_t*)cost);
            return hipGetLastError();
        }
    }
    {
        const int D = p.max_disparity - p.min_disparity;
        const int Lmax = p.max_arm_length > 1 ? p.max_arm_length : 1;
        if (D % 2 == 0) {
            int rs = 1;
            while (rs < 2 * Lmax + 2) rs <<= 1;
            const size_t shm = (2 * kMaxArm + 2 + (size_t)rs * 128) * sizeof(uint32_t);
            dim3 grid(w, (D + 127) / 128);
            if (u8)
                hipLaunchKernelGGL(cross_v2_kernel<uint8_t>, grid, dim3(64), shm, s, w, h, p.min_disparity, D, Lmax,
                                   rs - 1, dmin, dmax, arms, (const uint8_t*)tmp, (uint8_t*)cost);
            else
                hipLaunchKernelGGL(cross_v2_kernel<uint16_t>, grid, dim3(64), shm, s, w, h, p.min_disparity, D, Lmax,
                                   rs - 1, dmin, dmax, arms, (const uint16_t*)tmp, (uint16_t*)cost);
            return hipGetLastError();
        }
    }
    const int D = p.max_disparity - p.min_disparity;
    const int Lmax = p.max_arm_length > 1 ? p.max_arm_length : 1;
    int rs = 1;
    while (rs < 2 * Lmax + 2) rs <<= 1;
    const size_t shm = (2 * kMaxArm + 2 + (size_t)rs * 64) * sizeof(uint32_t);
    dim3 grid(w, (D + 63) / 64);
    if (u8)
        hipLaunchKernelGGL(cross_v_kernel<uint8_t>, grid, dim3(64), shm, s, w, h, p.min_disparity, D, Lmax, rs - 1, dmin,
                           dmax, arms, (const uint8_t*)tmp, (uint8_t*)cost);
    else
        hipLaunchKernelGGL(cross_v_kernel<uint16_t>, grid, dim3(64), shm, s, w, h, p.min_disparity, D, Lmax, rs - 1,
                           dmin, dmax, arms, (const uint16_t*)tmp, (uint16_t*)cost);
    return hipGetLastError();
}

hipError_t launch_cross_v(hipStream_t s, const apg_hip_params& p, int w, int h, const int16_t* dmin,
                          const int16_t* dmax, const uint32_t* arms, const uint16_t* tmp, uint16_t* cost) {
    return launch_cross_v_t(s, p, w, h, dmin, dmax, arms, tmp, cost, false, false);
}

// ------------------------------------------------------------------ SGM ----
// Scanline geometry of direction (dx, dy) (reference src/sgm_optimizer.cpp:186-259):
// rows for horizontal paths, columns for vertical ones and, for diagonals, the
// W + H - 1 rays that start on the border they enter from.
__device__ __forceinline__ void scan_start(int s, int w, int h, int dx, int dy, int& sx, int& sy, int& n) {
    if (dy == 0) {
        sy = s; sx = dx > 0 ? 0 : w - 1; n = w;
    } else if (dx == 0) {
        sx = s; sy = dy > 0 ? 0 : h - 1; n = h;
    } else {
        if (s < w) { sx = s; sy = dy > 0 ? 0 : h - 1; }
        else { const int k = s - w + 1; sx = dx > 0 ? 0 : w - 1; sy = dy > 0 ? k : h - 1 - k; }
        const int nx = dx > 0 ? w - sx : sx + 1;
        const int ny = dy > 0 ? h - sy : sy + 1;
        n = min(nx, ny);
    }
}

constexpr int SG_PF = 8;   // scanline steps fetched per batch

// Loads the packed pair (d = 2*m, 2*m+1) of a D-vector; out-of-range halves
// read as 65535 (invalid).
template <bool EVEN_D>
__device__ __forceinline__ uint32_t load_pair(const uint16_t* __restrict__ v, int m, int D) {
    const int d = 2 * m;
    if (EVEN_D) return d < D ? *(const uint32_t*)(v + d) : 0xFFFFFFFFu;
    const uint32_t lo = d < D ? v[d] : 0xFFFFu;
    const uint32_t hi = d + 1 < D ? v[d + 1] : 0xFFFFu;
    return lo | (hi << 16);
}
template <bool EVEN_D>
__device__ __forceinline__ void store_pair(uint16_t* __restrict__ v, int m, int D, uint32_t val) {
    const int d = 2 * m;
    if (EVEN_D) {
        if (d < D) *(uint32_t*)(v + d) = val;
    } else {
        if (d < D) v[d] = (uint16_t)val;
        if (d + 1 < D) v[d + 1] = (uint16_t)(val >> 16);
    }
}

// One wavefront runs the recurrence of process_pixel (reference
// src/sgm_optimizer.cpp:21-81) along one scanline.  Lane l holds K packed
// u16 pairs covering d = 2*(l*K + j) + {0,1}.  L and S use 65535 as the
// path-infinity / invalid value: saturating adds keep invalid states at 65535
// and (host-checked) npath*(cost_max + P2max) < 65535 keeps valid sums exact.
//   L(d) = C(d) + min(L'(d), L'(d-1)+P1, L'(d+1)+P1, minL'+P2) - minL'
//   L(d) = C(d) when the previous pixel has no valid state (or at a start).
// MODE 0 writes S = L, MODE 1 writes S = sat(S + L).
template <int K, bool EVEN_D, int MODE>
__global__ __launch_bounds__(64) void sgm_dir_kernel(apg_hip_params p, int w, int h, int D, int dx, int dy,
                                                     const uint8_t* __restrict__ gray, const uint16_t* __restrict__ cost,
                                                     uint16_t* __restrict__ sum) {
    const int lane = threadIdx.x;
    int sx, sy, n;
    scan_start(blockIdx.x, w, h, dx, dy, sx, sy, n);
    const long step = (long)dy * w + dx;
    const long p0 = (long)sy * w + sx;
    const uint32_t P1 = (uint32_t)p.P1;
    uint32_t prev[K];
#pragma unroll
    for (int j = 0; j < K; ++j) prev[j] = 0xFFFFFFFFu;
    uint32_t minprev = 0xFFFFu;   // wave-uniform; 0xFFFF = no valid state
    int gprev = 0;
    for (int tb = 0; tb < n; tb += 64) {
        // gray of the next 64 pixels of the scanline, one per lane
        const int tl = tb + lane;
        const uint32_t gchunk = tl < n ? gray[p0 + step * tl] : 0u;
        for (int t0 = tb; t0 < min(n, tb + 64); t0 += SG_PF) {
            uint32_t c[SG_PF][K];
            uint32_t so[SG_PF][K];
#pragma unroll
            for (int i = 0; i < SG_PF; ++i) {
                const int t = t0 + i;
                if (t < n) {
                    const uint16_t* cv = cost + (size_t)(p0 + step * t) * D;
#pragma unroll
                    for (int j = 0; j < K; ++j) c[i][j] = load_pair<EVEN_D>(cv, lane * K + j, D);
                    if (MODE == 1) {
                        const uint16_t* sv = sum + (size_t)(p0 + step * t) * D;
#pragma unroll
                        for (int j = 0; j < K; ++j) so[i][j] = load_pair<EVEN_D>(sv, lane * K + j, D);
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < SG_PF; ++i) {
                const int t = t0 + i;
                if (t >= n) break;
                const int g = __builtin_amdgcn_readlane((int)gchunk, t - tb);
                uint32_t L[K];
                if (t == 0 || minprev == 0xFFFFu) {
#pragma unroll
                    for (int j = 0; j < K; ++j) L[j] = c[i][j];
                } else {
                    const uint32_t P2 = (uint32_t)adaptive_p2(p, g - gprev);
                    const u16x2 mp = splat2(minprev);
                    const u16x2 mp2 = add_sat2(mp, splat2(P2));
                    const u16x2 p1 = splat2(P1);
                    const uint32_t left_hi = wave_shr1(prev[K - 1], 0xFFFFFFFFu);   // lane-1's last pair
                    const uint32_t right_lo = wave_shl1(prev[0], 0xFFFFFFFFu);      // lane+1's first pair
#pragma unroll
                    for (int j = 0; j < K; ++j) {
                        const uint32_t pj = prev[j];
                        const uint32_t before = j == 0 ? left_hi : prev[j - 1];
                        const uint32_t after = j == K - 1 ? right_lo : prev[j + 1];
                        const u16x2 dm1 = as_v2(__builtin_amdgcn_alignbit(pj, before, 16));
                        const u16x2 dp1 = as_v2(__builtin_amdgcn_alignbit(after, pj, 16));
                        u16x2 t2 = min2(as_v2(pj), min2(add_sat2(dm1, p1), add_sat2(dp1, p1)));
                        t2 = min2(t2, mp2);
                        L[j] = as_u32(add_sat2(as_v2(c[i][j]), t2 - mp));
                    }
                }
                uint32_t lm = 0xFFFFu;
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    prev[j] = L[j];
                    lm = min(lm, min(L[j] & 0xFFFFu, L[j] >> 16));
                }
                minprev = wave_min_u32(lm);
                gprev = g;
                uint16_t* sv = sum + (size_t)(p0 + step * t) * D;
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    const uint32_t out = MODE == 0 ? L[j] : as_u32(add_sat2(as_v2(so[i][j]), as_v2(L[j])));
                    store_pair<EVEN_D>(sv, lane * K + j, D, out);
                }
            }
        }
    }
}

template <int K, bool EVEN_D>
static hipError_t launch_sgm_k(hipStream_t s, const apg_hip_params& p, int w, int h, int D, const uint8_t* gray,
                               const uint16_t* cost, uint16_t* sum, int dx, int dy, int mode) {
    const int nscan = dy == 0 ? h : (dx == 0 ? w : w + h - 1);
    if (mode == 0)
        hipLaunchKernelGGL((sgm_dir_kernel<K, EVEN_D, 0>), dim3(nscan), dim3(64), 0, s, p, w, h, D, dx, dy, gray, cost, sum);
    else
        hipLaunchKernelGGL((sgm_dir_kernel<K, EVEN_D, 1>), dim3(nscan), dim3(64), 0, s, p, w, h, D, dx, dy, gray, cost, sum);
    return hipGetLastError();
}

hipError_t launch_sgm_dir(hipStream_t s, const apg_hip_params& p, int w, int h, int D, const uint8_t* gray,
                          const uint16_t* cost, uint16_t* sum, int dx, int dy, int mode) {
    const int K = (D + 127) / 128;
    const bool even = (D % 2) == 0;
#define APG_SGM_CASE(KK)                                                                           \
    case KK:                                                                                       \
        return even ? launch_sgm_k<KK, true>(s, p, w, h, D, gray, cost, sum, dx, dy, mode)         \
                    : launch_sgm_k<KK, false>(s, p, w, h, D, gray, cost, sum, dx, dy, mode);
    switch (K) {
        APG_SGM_CASE(1)
        APG_SGM_CASE(2)
        APG_SGM_CASE(3)
        APG_SGM_CASE(4)
        default: return hipErrorInvalidValue;
    }
#undef APG_SGM_CASE
}

// u32 accumulator variant of one path (SgmOptimizer::aggregate_path with a
// caller accumulator, reference src/sgm_optimizer.cpp:21-81): acc == INV32
// stays INV32, an invalid C sets INV32, otherwise acc += L.  The path L is
// recomputed by a MODE-0 scan into a scratch u16 volume first.
__global__ void add_path32_kernel(size_t n, const uint16_t* __restrict__ L, uint32_t* __restrict__ acc) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint16_t l = L[i];
    if (l == kInv16) acc[i] = kInv32;
    else if (acc[i] != kInv32) acc[i] += l;
}

__global__ void widen16_kernel(size_t n, const uint16_t* __restrict__ in, uint32_t* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint16_t v = in[i];
    out[i] = v == kInv16 ? kInv32 : (uint32_t)v;
}

hipError_t launch_widen16(hipStream_t s, size_t n, const uint16_t* in, uint32_t* out) {
    hipLaunchKernelGGL(widen16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, in, out);
    return hipGetLastError();
}

hipError_t launch_sgm_dir32_ws(hipStream_t s, const apg_hip_params& p, int w, int h, int D, const uint8_t* gray,
                               const uint16_t* cost, uint32_t* acc, int dx, int dy, uint16_t* L) {
    const size_t n = (size_t)w * h * D;
    hipError_t e = launch_sgm_dir(s, p, w, h, D, gray, cost, L, dx, dy, 0);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(add_path32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, L, acc);
    return hipGetLastError();
}

hipError_t launch_sgm_dir32(hipStream_t s, const apg_hip_params& p, int w, int h, int D, const uint8_t* gray,
                            const uint16_t* cost, uint32_t* acc, int dx, int dy) {
    // scratch L volume owned by the caller via acc's tail is not available;
    // allocate on the stream (stream-ordered allocator keeps it on device).
    const size_t n = (size_t)w * h * D;
    uint16_t* L = nullptr;
    hipError_t e = hipMallocAsync((void**)&L, n * sizeof(uint16_t), s);
    if (e != hipSuccess) return e;
    e = launch_sgm_dir(s, p, w, h, D, gray, cost, L, dx, dy, 0);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(add_path32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, L, acc);
        e = hipGetLastError();
    }
    hipError_t e2 = hipFreeAsync(L, s);
    return e != hipSuccess ? e : e2;
}

}  // namespace apgk
