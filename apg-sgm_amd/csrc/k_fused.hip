// gfx950 kernels of the fused (non-materialising) schedule:
//   cost_crossh_kernel : matching cost + right-view WTA candidates + Cross H
//                        pass in one pass over a row segment (the raw cost
//                        volume never reaches HBM);
//   rwta_finalize_kernel: right-view disparity with the legacy subpixel step
//                        from the per-pixel winning keys.
#include <type_traits>

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

// Sliding form of the same pass for u8 T and ranges that depend on the
// column only (the dense range init [min_disparity, min(max_disparity, x+1))).
// One wavefront walks a row segment of `seg` output columns plus the arm
// halo, 128 disparities at a time (lane l: d-indices 2l, 2l+1).  Per column it
// computes the two costs, adds them to u16-pair prefix sums kept in an LDS
// ring of 2*Lmax+2 columns, and Lmax columns behind emits T of one output
// column.  The prefix sums hold twice the cost sums (window sums < 125*255,
// doubled < 2^16, exact modulo 2^16), which is the numerator the divisor's
// magic takes.  With column ranges, C(xx, d) is valid iff d <= xx < W + d, so
// the valid count of the window [x-L, x+R] is |[max(x-L, d), min(x+R, W-1+d)]|
// and only sums need prefixes; once x - Lmax >= the chunk's largest d (and
// d >= 0) the window is whole for every lane, its count L+R+1 and the
// divisor uniform (the FAST phase).
// floor(s / n) = umulhi(2s, ceil(2^31 / n)) (exact for s * n < 2^31).
// Records are pre-formatted for the cost terms: g repeated in four bytes, so
// v_sad_u8 of two of them is 4|dg|, the AD LUT's byte offset; 4gx | 4gy << 16,
// so v_sad_u16 gives the gradient LUT's byte offset.  The right-pixel records
// are split by index parity: the lanes of one column read consecutive
// 16-byte records (no LDS bank conflicts).
// The loop is software-pipelined: the records of column c+1 are read while
// column c computes, and an output column's window sums are read right after
// the prefix write and consumed one column later.
constexpr int K2_SEG = 64;   // measured best at 1280x720 D=128 (more waves; the halo costs less)

struct K2Shm {
    size_t lut, mag, col, chi, rng, outr, rre, rro, key, ring, cring, total;
    size_t col_st, chi_st, rng_st, outr_st, rr_st, key_st;   // per-row copy strides
};

// gr: per-pixel ranges (the general form): each column's [dmin, dmax) and a
// second ring of u8-pair valid counts.  nr: image rows per wave (2 when
// D <= 64: each half-wave runs one row); the per-row arrays hold nr copies,
// copy r at offset r * (the array's stride member).
__host__ __device__ inline K2Shm k2_layout(int D, int Lmax, int seg, bool sym, bool gr = false, int nr = 1) {
    const int ncol = seg + 2 * Lmax;
    const int nrr = (ncol + D + 2) / 2 + 1;   // records per parity array
    K2Shm s;
    size_t o = 0;
    s.lut = o; o += (65 + 256 + 511) * 4;
    s.mag = o; o += (2 * kMaxArm + 2) * 4;
    o = (o + 15) & ~(size_t)15;
    s.col_st = (size_t)(ncol + 2) * 16;        // + the records read ahead of the last column
    s.col = o; o += nr * s.col_st;
    s.chi_st = sym ? 0 : (size_t)(ncol + 2) * 4;
    s.chi = o; o += nr * s.chi_st;
    s.rng_st = gr ? (size_t)(ncol + 2) * 4 : 0;
    s.rng = o; o += nr * s.rng_st;
    o = (o + 15) & ~(size_t)15;
    s.outr_st = (size_t)(seg + 1) * 16;
    s.outr = o; o += nr * s.outr_st;
    s.rr_st = (size_t)nrr * 16;
    s.rre = o; o += nr * s.rr_st;
    s.rro = o; o += nr * s.rr_st;
    s.key_st = ((size_t)(seg + D - 1) * 4 + 15) & ~(size_t)15;
    s.key = o; o += nr * s.key_st;
    o = (o + 15) & ~(size_t)15;
    s.ring = o; o += (size_t)(2 * Lmax + 2) * 256;
    s.cring = o; o += gr ? (size_t)(2 * Lmax + 2) * 128 : 0;
    s.total = o;
    return s;
}

constexpr int kFarLeft = -(1 << 24);   // x of a column outside the image: x - d is never in [0, W)

// (g, gx, gy) -> the two pre-formatted record words
__device__ __forceinline__ uint32_t rep4(uint32_t g) { return g * 0x01010101u; }
__device__ __forceinline__ uint32_t grad4(uint32_t gx, uint32_t gy) { return (gx << 2) | (gy << 18); }

// GR: ranges per pixel (dmin, dmax: the HighQuality prior's).  A state then
// also needs d in its own pixel's range, a Cross-H window counts its valid
// states from u8-pair prefix counts (a segment has fewer than 256 columns),
// and an output pixel's T is valid for d in its own range with a non-empty
// window (reference src/cost_aggregator.cpp:74-105).
// R2 (D <= 64): lanes 0-31 run row 2 * blockIdx.y and lanes 32-63 the row
// below, each half over the D disparities (FULL: D == 64), so no lane idles
// for want of disparities; each row has its own records, arms and keys.
template <bool SYM, bool FULL, bool NOCLAMP, bool GR, bool R2>
__device__ __forceinline__ void cost_crossh2_body(const uint3 bid, const CostLut* __restrict__ lut, int w, int h,
                                                  int d0, int D, int cost_max, int Lmax, int seg,
                                                  const uint8_t* __restrict__ gl, const uint8_t* __restrict__ gxl,
                                                  const uint8_t* __restrict__ gyl, const uint8_t* __restrict__ gr,
                                                  const uint8_t* __restrict__ gxr, const uint8_t* __restrict__ gyr,
                                                  const uint32_t* __restrict__ c32l,
                                                  const uint32_t* __restrict__ c32r,
                                                  const uint64_t* __restrict__ c64l,
                                                  const uint64_t* __restrict__ c64r, const int16_t* __restrict__ dmin,
                                                  const int16_t* __restrict__ dmax, const uint32_t* __restrict__ arms,
                                                  uint8_t* __restrict__ out, uint32_t* __restrict__ rkeys) {
    extern __shared__ __align__(16) unsigned char shm[];
    constexpr int NR = R2 ? 2 : 1;
    const K2Shm L = k2_layout(D, Lmax, seg, SYM, GR, NR);
    float* s_lc = (float*)(shm + L.lut);
    float* s_la = s_lc + 65;
    float* s_lg = s_la + 256;
    uint32_t* s_mag = (uint32_t*)(shm + L.mag);
    // Per-row arrays (NR copies, strides in L):
    //   col: left records: census lo, rep4(g), grad4(gx, gy), x (kFarLeft outside)
    //   chi: census hi (64-bit census); rng (GR): dmin | dmax << 16 (empty outside the image)
    //   outr: ring byte offsets of P(x+R+1) and of P(x-L), ceil(2^31/(L+R+1)),
    //         (x-L) | (x+R) << 16 (GR: the pixel's dmin | dmax << 16)
    //   rre / rro: right-pixel records by index parity; key: right-view WTA keys
    uint32_t* s_key = (uint32_t*)(shm + L.key);
    unsigned char* s_ring = shm + L.ring;    // [2*Lmax+2][64 lanes] u16-pair doubled prefix sums
    unsigned char* s_cring = shm + L.cring;  // GR: [2*Lmax+2][64 lanes] u8-pair prefix valid counts

    const int lane = threadIdx.x;
    const int rsub = R2 ? lane >> 5 : 0;    // the lane's row of the wave
    const int ln = R2 ? lane & 31 : lane;   // the lane within its row's group
    const int yw = bid.y * NR;         // the wave's first row
    const bool yin = yw + rsub < h;
    const int x0 = bid.x * seg;
    const int ncol = seg + 2 * Lmax, rs = 2 * Lmax + 2;
    const int xb = x0 - Lmax;              // column 0
    const int xs = xb - (d0 + D - 1);      // right pixel of record 0
    const int kbase = x0 - (d0 + D - 1);   // right pixel of s_key[0]
    const int nown = min(seg, w - x0);
    const size_t roww = (size_t)yw * w;   // the wave's first row

    // The staging below issues all of a batch's global loads before the first
    // of them is used, so a wave waits for memory once per batch rather than
    // once per loop trip.  The LUTs are contiguous floats in CostLut and in LDS.
    {
        constexpr int NLUT = 65 + 256 + 511, NL = (NLUT + 63) / 64;
        float lv[NL];
#pragma unroll
        for (int k = 0; k < NL; ++k) lv[k] = lut->census[min(lane + 64 * k, NLUT - 1)];
#pragma unroll
        for (int k = 0; k < NL; ++k)
            if (lane + 64 * k < NLUT) s_lc[lane + 64 * k] = lv[k];
    }
    for (int n = lane; n < 2 * kMaxArm + 2; n += 64) s_mag[n] = lut->mag31[n];
    __syncthreads();   // the arms' records below read s_mag
    for (int r = 0; r < NR && yw + r < h; ++r) {
    const size_t row = roww + (size_t)r * w;
    uint4* s_colw = (uint4*)(shm + L.col + r * L.col_st);
    uint32_t* s_chiw = (uint32_t*)(shm + L.chi + r * L.chi_st);
    uint32_t* s_rngw = (uint32_t*)(shm + L.rng + r * L.rng_st);
    uint4* s_outw = (uint4*)(shm + L.outr + r * L.outr_st);
    const size_t rro_w = L.rro + r * L.rr_st, rre_w = L.rre + r * L.rr_st;
    // left records: columns xb .. xb + ncol - 1 (kFarLeft outside the image)
    for (int c0 = 0; c0 <= ncol + 1; c0 += 256) {
        uint32_t lo[4], hi[4], g[4], gx[4], gy[4], rg[4];
        bool in[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int c = c0 + lane + 64 * j, x = xb + c;
            in[j] = c < ncol && x >= 0 && x < w;
            const size_t q = row + clampi(x, 0, w - 1);
            if (SYM) { lo[j] = c32l[q]; hi[j] = 0u; }
            else { const uint64_t v = c64l[q]; lo[j] = (uint32_t)v; hi[j] = (uint32_t)(v >> 32); }
            g[j] = gl[q];
            gx[j] = gxl[q];
            gy[j] = gyl[q];
            if (GR) rg[j] = (uint32_t)(uint16_t)dmin[q] | ((uint32_t)(uint16_t)dmax[q] << 16);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int c = c0 + lane + 64 * j;
            if (c <= ncol + 1) {
                s_colw[c] = in[j] ? make_uint4(lo[j], rep4(g[j]), grad4(gx[j], gy[j]), (uint32_t)(xb + c))
                                  : make_uint4(0u, 0u, 0u, (uint32_t)kFarLeft);
                if (!SYM) s_chiw[c] = in[j] ? hi[j] : 0u;
                if (GR) s_rngw[c] = in[j] ? rg[j] : 0u;
            }
        }
    }
    for (int o0 = 0; o0 <= seg; o0 += 256) {
        uint32_t a[4], rg[4];
        bool in[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int o = o0 + lane + 64 * j, x = x0 + o;
            in[j] = o < seg && x < w;
            const size_t q = row + min(x, w - 1);
            a[j] = arms ? arms[q] : 0u;   // no Cross pass: windows of one column
            if (GR) rg[j] = (uint32_t)(uint16_t)dmin[q] | ((uint32_t)(uint16_t)dmax[q] << 16);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int o = o0 + lane + 64 * j;
            if (o > seg) continue;
            uint4 r = make_uint4(0u, 0u, 0u, 0u);
            if (in[j]) {
                const int x = x0 + o;
                const int La = a[j] & 0xFF, Ra = (a[j] >> 8) & 0xFF;
                const int co = Lmax + o;   // column index of x; P(j) sums columns < j
                r.x = (uint32_t)(((co + Ra + 1) % rs) * 256);
                r.y = (uint32_t)(((co - La) % rs) * 256);
                r.z = s_mag[La + Ra + 1];
                r.w = GR ? rg[j] : (uint32_t)(x - La) | ((uint32_t)(x + Ra) << 16);
            }
            s_outw[o] = r;
        }
    }
    // right pixel xs + i -> parity array i & 1, entry i >> 1
    for (int i0 = 0; i0 < ncol + D + 1; i0 += 256) {
        uint32_t lo[4], hi[4], g[4], gx[4], gy[4];
        bool in[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int i = i0 + lane + 64 * j, xr = xs + i;
            in[j] = i < ncol + D - 1 && xr >= 0 && xr < w;
            const size_t q = row + clampi(xr, 0, w - 1);
            if (SYM) { lo[j] = c32r[q]; hi[j] = 0u; }
            else { const uint64_t v = c64r[q]; lo[j] = (uint32_t)v; hi[j] = (uint32_t)(v >> 32); }
            g[j] = gr[q];
            gx[j] = gxr[q];
            gy[j] = gyr[q];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int i = i0 + lane + 64 * j;
            if (i >= ncol + D + 1) continue;
            const uint4 r = in[j] ? make_uint4(lo[j], rep4(g[j]), grad4(gx[j], gy[j]), SYM ? 0u : hi[j])
                                  : make_uint4(0u, 0u, 0u, 0u);
            ((uint4*)(shm + ((i & 1) ? rro_w : rre_w)))[i >> 1] = r;
        }
    }
    }
    for (int i = lane; i < NR * (int)(L.key_st / 4); i += 64) s_key[i] = 0xFFFFFFFFu;
    __syncthreads();

    const float scale = lut->scale;
    const auto rout = raw_rsrc(out);
    const uint32_t lofs = 4u * (uint32_t)lane;
    const uint32_t uw = (uint32_t)w;
    const uint32_t la_base = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) float*)s_la;
    const uint32_t lg_base = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) float*)s_lg;
    // the lane's row's copies of the per-row arrays
    const uint4* s_colr = (const uint4*)(shm + L.col + rsub * L.col_st);
    const uint32_t* s_chir = (const uint32_t*)(shm + L.chi + rsub * L.chi_st);
    const uint32_t* s_rngr = (const uint32_t*)(shm + L.rng + rsub * L.rng_st);
    const uint4* s_outr = (const uint4*)(shm + L.outr + rsub * L.outr_st);
    uint32_t* s_keyr = (uint32_t*)(shm + L.key + rsub * L.key_st);
    const size_t rr_r = rsub * L.rr_st;
    // T of the lane's row: its byte offset from the wave's first row
    const uint32_t voff_row = (uint32_t)rsub * uw * (uint32_t)D;
    for (int dc = 0; dc < D; dc += 128) {
        const int di0 = dc + 2 * ln;   // the lane's pair (di0, di0 + 1); D even
        const int dA = d0 + di0;
        const bool act = (FULL || di0 < D) && (!R2 || yin);
        const int dtop = d0 + min(dc + 127, D - 1);   // largest d of the chunk
        const bool nonneg = d0 >= 0;
        // record of the right pixel x - d for (c, di) is c + D - 1 - di = c + 2m + 1
        // (m = D/2 - 1 - di0/2); lanes past D read records 1 + c (their states are masked)
        const int m = act ? D / 2 - 1 - di0 / 2 : 0;
        auto rr_at = [&](int j) -> uint4 {   // record 2m + 1 + j
            const int u = j + 1;
            return ((const uint4*)(shm + ((u & 1) ? L.rro : L.rre) + rr_r))[(u >> 1) + m];
        };
        // s_key entry of the right pixel xr = x - dA is xr - kbase
        uint32_t* key_lane = s_keyr - dA - kbase;
        *(uint32_t*)(s_ring + lofs) = 0u;   // P(0)
        if (GR) *(uint16_t*)(s_cring + lofs / 2) = 0u;
        uint32_t acc = 0u, cacc = 0u;
        uint32_t ringw = 256u;              // ring byte offset of P(c + 1)
        const uint32_t ringe = (uint32_t)rs * 256u;
        // pipeline registers: records of the next two columns, LUT terms of the next one
        uint4 cr_n = s_colr[0], r0_n = rr_at(0), r1_n = rr_at(-1), or_n = s_outr[0];
        uint32_t ch_n = SYM ? 0u : s_chir[0];
        uint32_t rg_n = GR ? s_rngr[0] : 0u, rg_c = 0u;
        float tc0, ta0, tg0, tc1, ta1, tg1;
        // LDS byte addresses of the AD and gradient LUT entries come out of the
        // SAD instructions themselves (the table base is their accumulator)
        auto lds_f32 = [](uint32_t a) -> float { return *(const __attribute__((address_space(3))) float*)(uintptr_t)a; };
        auto lut_terms = [&](const uint4& cr, uint32_t chl, const uint4& r0, const uint4& r1) {
            const int pop0 = SYM ? __popc(cr.x ^ r0.x) : __popc(cr.x ^ r0.x) + __popc(chl ^ r0.w);
            const int pop1 = SYM ? __popc(cr.x ^ r1.x) : __popc(cr.x ^ r1.x) + __popc(chl ^ r1.w);
            tc0 = s_lc[pop0];
            ta0 = lds_f32(__builtin_amdgcn_sad_u8(r0.y, cr.y, la_base));
            tg0 = lds_f32(__builtin_amdgcn_sad_u16(r0.z, cr.z, lg_base));
            tc1 = s_lc[pop1];
            ta1 = lds_f32(__builtin_amdgcn_sad_u8(r1.y, cr.y, la_base));
            tg1 = lds_f32(__builtin_amdgcn_sad_u16(r1.z, cr.z, lg_base));
        };
        lut_terms(cr_n, ch_n, r0_n, r1_n);   // column 0
        uint4 cr_c = cr_n;                   // column 0's record (its x)
        cr_n = s_colr[1];
        if (!SYM) ch_n = s_chir[1];
        if (GR) {
            rg_c = rg_n;
            rg_n = s_rngr[1];
        }
        r0_n = rr_at(1);
        r1_n = rr_at(0);
        uint32_t pp1 = 0u, pp0 = 0u, cp1 = 0u, cp0 = 0u;
        uint4 por = make_uint4(0u, 0u, 0u, 0u);
        int pxo = 0;
        bool pend = false;

        // T of the column whose window sums were read last step; FAST: whole
        // windows for every lane (count L+R+1, T valid)
        auto flush = [&](auto fast_t) {
            constexpr bool FAST = decltype(fast_t)::value;
            const uint32_t sw = as_u32(as_v2(pp1) - as_v2(pp0));   // doubled window sums
            uint32_t q0, q1;
            if (FAST) {
                q0 = __umulhi(sw & 0xFFFFu, por.z);
                q1 = __umulhi(sw >> 16, por.z);
            } else if (GR) {
                // own range of the output pixel, valid counts of the window
                const int lo = (int)(int16_t)(por.w & 0xFFFFu), hi = (int)(int16_t)(por.w >> 16);
                const uint32_t wc = cp1 - cp0;   // per byte: prefix counts never fall
                uint32_t qq[2];
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const int d = dA + e;
                    const uint32_t cnt = (wc >> (8 * e)) & 0xFFu;
                    const uint32_t se = e ? (sw >> 16) : (sw & 0xFFFFu);
                    const bool v = d >= lo && d < hi && cnt > 0;
                    // the magic number is read on every lane (the index kept
                    // in the table) and the test is a select: no lane branches
                    const uint32_t qv = __umulhi(se, s_mag[min(cnt, (uint32_t)(2 * kMaxArm + 1))]);
                    qq[e] = v ? qv : 0xFFu;
                }
                q0 = qq[0];
                q1 = qq[1];
            } else {
                const int xlo = (int)(por.w & 0xFFFFu), xhi = (int)(por.w >> 16);
                uint32_t qq[2];
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const int d = dA + e;
                    const int cnt = min(xhi, w - 1 + d) - max(xlo, d) + 1;
                    const uint32_t se = e ? (sw >> 16) : (sw & 0xFFFFu);
                    const bool v = d <= pxo && cnt > 0;   // d in [min, min(max, x+1)), window not empty
                    const uint32_t qv = __umulhi(se, s_mag[min(max(cnt, 0), 2 * kMaxArm + 1)]);
                    qq[e] = v ? qv : 0xFFu;
                }
                q0 = qq[0];
                q1 = qq[1];
            }
            // bytes q0, q1 (each < 256) -> one 16-bit store
            // an inactive lane's offset lies past the records: the store is dropped
            buf_st16(rout, __builtin_amdgcn_perm(q1, q0, 0x0c0c0400u), act ? (uint32_t)di0 + voff_row : 0x80000000u,
                     (uint32_t)((int)roww + pxo) * (uint32_t)D);
        };

        // keys: own column (right-view WTA candidates); outp: an output column's
        // window sums are read; allv: every lane's right pixels x - d, x - d - 1
        // lie in the image (no per-lane validity test); fast: the column flushed
        // in this step has whole windows
        auto step = [&](int c, auto keys_t, auto outp_t, auto allv_t, auto fast_t) {
            constexpr bool keys = decltype(keys_t)::value, outp = decltype(outp_t)::value;
            constexpr bool allv = decltype(allv_t)::value;
            const uint4 cr = cr_c;
            const uint32_t rg = rg_c;
            // same operation order as state_cost (reference src/cost_computer.cpp:176-199)
            const float f0 = (tc0 + ta0) + tg0;
            const float f1 = (tc1 + ta1) + tg1;
            // LUT terms of column c + 1 and records of column c + 2 (padded past the end)
            lut_terms(cr_n, ch_n, r0_n, r1_n);
            cr_c = cr_n;
            cr_n = s_colr[c + 2];
            if (!SYM) ch_n = s_chir[c + 2];
            if (GR) {
                rg_c = rg_n;
                rg_n = s_rngr[c + 2];
            }
            // d + 1 at column c + 2 pairs with the right pixel of d at column c + 1
            r1_n = r0_n;
            r0_n = rr_at(c + 2);
            if (pend) flush(fast_t);
            // NOCLAMP: the host showed every LUT term >= 0 and the largest
            // quantised sum <= cost_max, so the clamp of the reference is the identity
            const uint32_t cst0 = NOCLAMP ? (uint32_t)(int)(f0 * scale + 0.5f)
                                          : (uint32_t)clampi((int)(f0 * scale + 0.5f), 0, cost_max);
            const uint32_t cst1 = NOCLAMP ? (uint32_t)(int)(f1 * scale + 0.5f)
                                          : (uint32_t)clampi((int)(f1 * scale + 0.5f), 0, cost_max);
            const int xr0 = (int)cr.w - dA;   // right pixel of d; d + 1 has xr0 - 1
            bool v0 = allv ? act : act && (uint32_t)xr0 < uw;
            bool v1 = allv ? act : act && (uint32_t)(xr0 - 1) < uw;
            if (GR) {   // d in the column's own range [dmin, dmax)
                const int lo = (int)(int16_t)(rg & 0xFFFFu), hi = (int)(int16_t)(rg >> 16);
                v0 = v0 && dA >= lo && dA < hi;
                v1 = v1 && dA + 1 >= lo && dA + 1 < hi;
            }
            if (keys) {
                uint32_t* k0p = FULL ? key_lane + (int)cr.w : s_keyr + max(xr0 - kbase, 1);
                atomicMin(k0p, v0 ? (cst0 << 16) | (uint32_t)di0 : 0xFFFFFFFFu);
                atomicMin(k0p - 1, v1 ? (cst1 << 16) | (uint32_t)(di0 + 1) : 0xFFFFFFFFu);
            }
            const uint32_t pair = (v0 ? cst0 : 0u) | ((v1 ? cst1 : 0u) << 16);
            // acc += 2 * pair per u16 lane in one v_pk_mad_u16 (no carry between halves)
            {
                uint32_t nacc;
                asm("v_pk_mad_u16 %0, %1, %2, %3" : "=v"(nacc) : "v"(pair), "s"(0x00020002u), "v"(acc));
                acc = nacc;
            }
            *(uint32_t*)(s_ring + ringw + lofs) = acc;   // P(c + 1)
            if (GR) {
                cacc += (v0 ? 1u : 0u) | (v1 ? 0x100u : 0u);
                *(uint16_t*)(s_cring + ringw / 2 + lofs / 2) = (uint16_t)cacc;
            }
            ringw = ringw + 256u == ringe ? 0u : ringw + 256u;
            pend = outp;
            if (outp) {
                por = or_n;
                or_n = s_outr[c - 2 * Lmax + 1];
                pxo = x0 + c - 2 * Lmax;
                pp1 = *(const uint32_t*)(s_ring + por.x + lofs);
                pp0 = *(const uint32_t*)(s_ring + por.y + lofs);
                if (GR) {
                    cp1 = *(const uint16_t*)(s_cring + por.x / 2 + lofs / 2);
                    cp0 = *(const uint16_t*)(s_cring + por.y / 2 + lofs / 2);
                }
            }
        };

        // phases of c: keys for own columns [Lmax, Lmax + nown), output (Lmax
        // behind) for [2Lmax, 2Lmax + nown); columns past 2Lmax + nown lie beyond
        // the image and feed nothing.  The image columns where every d of the
        // chunk has its right pixel inside the image, [vx0, vx1], in c: [cvb,
        // cve).  Steps from cf on flush columns x >= dtop + Lmax (whole windows).
        // The loops over allv columns run two columns per iteration so the
        // pipeline registers alternate instead of being copied.
        const int kb = Lmax, ke = Lmax + nown, ob = 2 * Lmax, oe = 2 * Lmax + nown;
        const int vx0 = max(dtop, 0), vx1 = min(w - 1 + d0 + dc, w - 1);
        const int cvb = min(oe, max(0, vx0 - xb)), cve = max(cvb, min(oe, vx1 - xb + 1));
        const int cf = nonneg && !GR ? min(oe, max(ob + 1, dtop + Lmax - x0 + 2 * Lmax + 1)) : oe;
        using TT = std::true_type;
        using FF = std::false_type;
        int c = 0;
        auto run = [&](int nx, auto k_t, auto o_t, auto v_t, auto f_t) {
            if constexpr (decltype(v_t)::value)
                for (; c + 1 < nx; c += 2) {
                    step(c, k_t, o_t, v_t, f_t);
                    step(c + 1, k_t, o_t, v_t, f_t);
                }
            for (; c < nx; ++c) step(c, k_t, o_t, v_t, f_t);
        };
        auto run_o = [&](int nx, auto k_t, auto v_t, bool inF) {
            if (inF) run(nx, k_t, TT{}, v_t, TT{});
            else run(nx, k_t, TT{}, v_t, FF{});
        };
        while (c < oe) {
            const bool inK = c >= kb && c < ke, inO = c >= ob, inV = c >= cvb && c < cve, inF = c >= cf;
            int nx = oe;
            if (c < kb) nx = min(nx, kb);
            if (c < ke) nx = min(nx, ke);
            if (c < ob) nx = min(nx, ob);
            if (c < cvb) nx = min(nx, cvb);
            if (c < cve) nx = min(nx, cve);
            if (c < cf) nx = min(nx, cf);
            if (inV) {
                if (inO) {
                    if (inK) run_o(nx, TT{}, TT{}, inF);
                    else run_o(nx, FF{}, TT{}, inF);
                } else {
                    if (inK) run(nx, TT{}, FF{}, TT{}, FF{});
                    else run(nx, FF{}, FF{}, TT{}, FF{});
                }
            } else {
                if (inO) {
                    if (inK) run_o(nx, TT{}, FF{}, inF);
                    else run_o(nx, FF{}, FF{}, inF);
                } else {
                    if (inK) run(nx, TT{}, FF{}, FF{}, FF{});
                    else run(nx, FF{}, FF{}, FF{}, FF{});
                }
            }
        }
        if (pend) flush(FF{});
    }
    __syncthreads();
    for (int r = 0; r < NR && yw + r < h; ++r) {
        const uint32_t* s_keyw = (const uint32_t*)(shm + L.key + r * L.key_st);
        for (int i = lane; i < seg + D - 1; i += 64) {
            const int xr = kbase + i;
            const uint32_t k = s_keyw[i];
            if (xr >= 0 && xr < w && k != 0xFFFFFFFFu) atomicMin(&rkeys[roww + (size_t)r * w + xr], k);
        }
    }
}

template <bool SYM, bool FULL, bool NOCLAMP, bool GR, bool R2>
__global__ __launch_bounds__(64) void cost_crossh2_kernel(
    const CostLut* __restrict__ lut, int w, int h, int d0, int D, int cost_max, int Lmax, int seg,
    const uint8_t* __restrict__ gl, const uint8_t* __restrict__ gxl, const uint8_t* __restrict__ gyl,
    const uint8_t* __restrict__ gr, const uint8_t* __restrict__ gxr, const uint8_t* __restrict__ gyr,
    const uint32_t* __restrict__ c32l, const uint32_t* __restrict__ c32r, const uint64_t* __restrict__ c64l,
    const uint64_t* __restrict__ c64r, const int16_t* __restrict__ dmin, const int16_t* __restrict__ dmax,
    const uint32_t* __restrict__ arms, uint8_t* __restrict__ out, uint32_t* __restrict__ rkeys) {
    cost_crossh2_body<SYM, FULL, NOCLAMP, GR, R2>(block3(), lut, w, h, d0, D, cost_max, Lmax, seg, gl, gxl, gyl, gr,
                                                  gxr, gyr, c32l, c32r, c64l, c64r, dmin, dmax, arms, out, rkeys);
}

template <bool SYM, bool FULL, bool NOCLAMP, bool GR, bool R2>
struct CostCrossH2Fn {
    static constexpr int kThreads = 64;
    template <typename... A>
    __device__ static void run(const uint3 b, const A&... a) { cost_crossh2_body<SYM, FULL, NOCLAMP, GR, R2>(b, a...); }
};

hipError_t launch_cost_crossh(hipStream_t s, const apg_hip_params& p, const CostLut* d_lut, int w, int h,
                              const uint8_t* gl, const uint8_t* gxl, const uint8_t* gyl, const uint8_t* gr,
                              const uint8_t* gxr, const uint8_t* gyr, const uint32_t* c32l, const uint32_t* c32r,
                              const uint64_t* c64l, const uint64_t* c64r, const int16_t* dmin, const int16_t* dmax,
                              const uint32_t* arms, void* out, bool out_u8, uint32_t* rkeys, bool column_ranges,
                              bool clamp_free, const KTune& tn, LaunchSink* sink) {
    const int D = p.max_disparity - p.min_disparity;
    const bool cross = p.agg_enable != 0;
    const int Lmax = cross ? (p.max_arm_length > 1 ? p.max_arm_length : 1) : 0;
    // built-in segment: 64 columns up to D = 128; 128 columns beyond (1920x1080 D=256:
    // 771 against 816 us, and 1.63 against 1.85 ms over the HighQuality prior's ranges)
    int seg2 = tn.cost_segment > 0 ? tn.cost_segment : (D > 128 ? 2 * K2_SEG : K2_SEG);
    seg2 = w < seg2 ? w : seg2;
    // per-pixel ranges: the valid counts are u8 prefixes over a segment's columns
    const bool granges = !column_ranges;
    // without the Cross pass (Fast) the same kernel runs with one-column
    // windows (Lmax 0, no arms): T is then the cost itself, 255 where invalid
    if (out_u8 && D % 2 == 0 && (size_t)w * h * D < (1ull << 31) && (!granges || seg2 + 2 * Lmax <= 255)) {
        const int seg = seg2;
        const bool r2 = D <= 64;   // two rows per wave
        const K2Shm L2 = k2_layout(D, Lmax, seg, p.census_type == 1, granges, r2 ? 2 : 1);
        const bool full = r2 ? D == 64 : D % 128 == 0;
        const dim3 grid((w + seg - 1) / seg, r2 ? (h + 1) / 2 : h);
        if (sink && !granges) {
            // the column-range instances as group launches
            auto rec = [&](auto fn) {
                return sink->add<decltype(fn)>(grid, L2.total, d_lut, w, h, p.min_disparity, D, p.cost_max, Lmax, seg,
                                               gl, gxl, gyl, gr, gxr, gyr, c32l, c32r, c64l, c64r, dmin, dmax,
                                               cross ? arms : nullptr, (uint8_t*)out, rkeys);
            };
            const bool sym = p.census_type == 1;
#define APG_G2(S, F, N) (r2 ? rec(CostCrossH2Fn<S, F, N, false, true>{}) : rec(CostCrossH2Fn<S, F, N, false, false>{}))
            if (clamp_free)
                return sym ? (full ? APG_G2(true, true, true) : APG_G2(true, false, true))
                           : (full ? APG_G2(false, true, true) : APG_G2(false, false, true));
            return sym ? (full ? APG_G2(true, true, false) : APG_G2(true, false, false))
                       : (full ? APG_G2(false, true, false) : APG_G2(false, false, false));
#undef APG_G2
        }
        if (sink) return hipErrorNotSupported;   // per-pixel ranges: no group form
        decltype(&cost_crossh2_kernel<true, true, true, false, false>) k;
#define APG_K2R(S, F, N, G) (r2 ? cost_crossh2_kernel<S, F, N, G, true> : cost_crossh2_kernel<S, F, N, G, false>)
#define APG_K2(S, F, N) (granges ? APG_K2R(S, F, N, true) : APG_K2R(S, F, N, false))
        if (clamp_free)
            k = p.census_type == 1 ? (full ? APG_K2(true, true, true) : APG_K2(true, false, true))
                                   : (full ? APG_K2(false, true, true) : APG_K2(false, false, true));
        else
            k = p.census_type == 1 ? (full ? APG_K2(true, true, false) : APG_K2(true, false, false))
                                   : (full ? APG_K2(false, true, false) : APG_K2(false, false, false));
#undef APG_K2
#undef APG_K2R
        if (L2.total > 64 * 1024) {
            hipError_t e =
                hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)L2.total);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(k, grid, dim3(64), L2.total, s, d_lut, w, h,
                           p.min_disparity, D,
                           p.cost_max, Lmax, seg, gl, gxl, gyl, gr, gxr, gyr, c32l, c32r, c64l, c64r, dmin, dmax,
                           cross ? arms : nullptr, (uint8_t*)out, rkeys);
        return hipGetLastError();
    }
    if (sink) return hipErrorNotSupported;
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
__device__ __forceinline__ void rwta_finalize_body(const uint3 bid, const CostLut* __restrict__ lut, int w, int h,
                                                   int d0, int D, int cost_max, int use_ad, int use_grad,
                                                   int subpixel, const uint8_t* __restrict__ gl,
                                                   const uint8_t* __restrict__ gxl, const uint8_t* __restrict__ gyl,
                                                   const uint8_t* __restrict__ gr, const uint8_t* __restrict__ gxr,
                                                   const uint8_t* __restrict__ gyr, const uint32_t* __restrict__ c32l,
                                                   const uint32_t* __restrict__ c32r,
                                                   const uint64_t* __restrict__ c64l,
                                                   const uint64_t* __restrict__ c64r,
                                                   const int16_t* __restrict__ dmin, const int16_t* __restrict__ dmax,
                                                   const uint32_t* __restrict__ rkeys, float* __restrict__ dr) {
    const int xr = bid.x * blockDim.x + threadIdx.x, y = bid.y;
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

template <bool SYM>
__global__ __launch_bounds__(128) void rwta_finalize_kernel(const CostLut* __restrict__ lut, int w, int h, int d0,
                                                            int D, int cost_max, int use_ad, int use_grad,
                                                            int subpixel, const uint8_t* __restrict__ gl,
                                                            const uint8_t* __restrict__ gxl,
                                                            const uint8_t* __restrict__ gyl,
                                                            const uint8_t* __restrict__ gr,
                                                            const uint8_t* __restrict__ gxr,
                                                            const uint8_t* __restrict__ gyr,
                                                            const uint32_t* __restrict__ c32l,
                                                            const uint32_t* __restrict__ c32r,
                                                            const uint64_t* __restrict__ c64l,
                                                            const uint64_t* __restrict__ c64r,
                                                            const int16_t* __restrict__ dmin,
                                                            const int16_t* __restrict__ dmax,
                                                            const uint32_t* __restrict__ rkeys,
                                                            float* __restrict__ dr) {
    rwta_finalize_body<SYM>(block3(), lut, w, h, d0, D, cost_max, use_ad, use_grad, subpixel, gl, gxl, gyl, gr, gxr,
                            gyr, c32l, c32r, c64l, c64r, dmin, dmax, rkeys, dr);
}

template <bool SYM>
struct RwtaFinalizeFn {
    static constexpr int kThreads = 128;
    template <typename... A>
    __device__ static void run(const uint3 b, const A&... a) { rwta_finalize_body<SYM>(b, a...); }
};

hipError_t launch_rwta_finalize(hipStream_t s, const apg_hip_params& p, const CostLut* d_lut, int w, int h,
                                const uint8_t* gl, const uint8_t* gxl, const uint8_t* gyl, const uint8_t* gr,
                                const uint8_t* gxr, const uint8_t* gyr, const uint32_t* c32l, const uint32_t* c32r,
                                const uint64_t* c64l, const uint64_t* c64r, const int16_t* dmin, const int16_t* dmax,
                                const uint32_t* rkeys, float* dr, LaunchSink* sink) {
    const int D = p.max_disparity - p.min_disparity;
    dim3 grid((w + 127) / 128, h);
    if (sink)
        return p.census_type == 1
                   ? sink->add<RwtaFinalizeFn<true>>(grid, 0, d_lut, w, h, p.min_disparity, D, p.cost_max, p.use_ad,
                                                     p.use_grad, p.post_subpixel, gl, gxl, gyl, gr, gxr, gyr, c32l,
                                                     c32r, c64l, c64r, dmin, dmax, rkeys, dr)
                   : sink->add<RwtaFinalizeFn<false>>(grid, 0, d_lut, w, h, p.min_disparity, D, p.cost_max, p.use_ad,
                                                      p.use_grad, p.post_subpixel, gl, gxl, gyl, gr, gxr, gyr, c32l,
                                                      c32r, c64l, c64r, dmin, dmax, rkeys, dr);
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
