// gfx950 SGM kernel of the fused schedule.
//
// Every path value is L_r(p,d) = C'(p,d) + delta_r(p,d) with
// delta_r = min(L'(d), L'(d+-1)+P1, minL'+P2) - minL' in [0, P2] (0 at a scan
// start or after a pixel without valid state), so the path sum is
//   S(p,d) = npath * C'(p,d) + sum_r delta_r(p,d)
// and only the deltas (u8, P2max <= 254) are stored, one plane per direction.
// One wavefront owns one scanline (a row, a column or a diagonal ray) and runs
// its two opposite directions as independent chains: chain A walks forward,
// chain B backward.  Lane l holds disparity pairs (2(lK+j), 2(lK+j)+1),
// j < K, as packed u16 with 65535 as path infinity.  Loads and stores use
// buffer instructions with scalar per-step offsets; the C' words of the next
// SL_PF steps of both chains are in flight while the current ones are used.
#include <type_traits>
#include <utility>

#include "device_common.hpp"
#include "kernels.hpp"

namespace apgk {

constexpr int SL_PF = 8;   // steps per prefetch batch (two batches in flight)
// paired planes: positions whose two arrivals are fewer steps apart keep the
// first delta in LDS (a parked delta is stored at least four batches before
// the second arrival's prefetch reads it back)
constexpr int kPairGap = 48;

// Raw C' words of one lane: K pairs of TC (2, 4 or 8 bytes)
template <typename TC, int K>
struct CRaw {
    static constexpr int kBytes = 2 * K * (int)sizeof(TC);
    uint32_t w[kBytes > 4 ? 2 : 1];
};

// voff2: offset of the second word of an 8-byte lane (the first word's offset
// again when the lane's last pair lies past D, so nothing past the pixel is read)
template <typename TC, int K>
__device__ __forceinline__ CRaw<TC, K> load_c(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t voff2, uint32_t soff) {
    CRaw<TC, K> v;
    if constexpr (CRaw<TC, K>::kBytes == 2) {
        v.w[0] = buf_ld16(r, voff, soff);
    } else if constexpr (CRaw<TC, K>::kBytes == 4) {
        v.w[0] = buf_ld32(r, voff, soff);
    } else {
        v.w[0] = buf_ld32(r, voff, soff);
        v.w[1] = buf_ld32(r, voff2, soff);
    }
    return v;
}

// pair j of the lane -> u16 halves, invalid (u8 255 / u16 65535) -> 65535.
// ALLV: the caller knows every state of the pixel is valid (no u8 255).
template <typename TC, int K, bool ALLV = false>
__device__ __forceinline__ uint32_t decode_c(const CRaw<TC, K>& v, int j) {
    if constexpr (sizeof(TC) == 2) {
        return v.w[j];
    } else {
        const uint32_t x = __builtin_amdgcn_perm(0u, v.w[0], j == 0 ? 0x0c010c00u : 0x0c030c02u);
        if constexpr (ALLV) return x;
        const u16x2 t = (as_v2(x) + splat2(1)) >> (u16x2)8;   // 1 where the byte is 255
        return as_u32(t * splat2(0xFF00) + as_v2(x));
    }
}

// Scanline f (0 rows, 1 columns, 2 diagonal +x+y rays, 3 diagonal +x-y rays),
// member s: first pixel, pixel step and length of chain A.
__device__ __forceinline__ void family_line(int f, int s, int w, int h, int& p0, int& step, int& n) {
    int sx, sy, dx, dy;
    if (f == 0) { dx = 1; dy = 0; sx = 0; sy = s; n = w; }
    else if (f == 1) { dx = 0; dy = 1; sx = s; sy = 0; n = h; }
    else {
        dx = 1; dy = f == 2 ? 1 : -1;
        if (s < w) { sx = s; sy = dy > 0 ? 0 : h - 1; }
        else { const int k = s - w + 1; sx = 0; sy = dy > 0 ? k : h - 1 - k; }
        const int nx = w - sx, ny = dy > 0 ? h - sy : sy + 1;
        n = min(nx, ny);
    }
    p0 = sy * w + sx;
    step = dy * w + dx;
}

// One chain's state and step (reference src/sgm_optimizer.cpp:21-81).
template <int K>
struct Chain {
    uint32_t prev[K];
    uint32_t minprev;   // wave-uniform, 65535 when no valid state
};

// L = C + min(L'(d), L'(d-1)+P1, L'(d+1)+P1, minL'+P2) - minL'; with no valid
// previous state (all 65535, minL' = 65535) every term saturates and L = C,
// the delta 0.  p1s is P1 in both halves; p1l / p1r (pairs 0 and K-1) hold
// 65535 instead in the half whose neighbour lies outside [0, 128K), so the
// saturating add ignores the neighbour the DPP shift filled with 0.
// Returns the lane's minimum over its new L values (min of halves); the
// caller reduces it over the wave into ch.minprev.
template <int K>
__device__ __forceinline__ uint32_t chain_step(Chain<K>& ch, const uint32_t (&c)[K], uint32_t P2, uint32_t p1s,
                                               uint32_t p1l, uint32_t p1r, uint32_t (&dl)[K]) {
    const uint32_t mp = ch.minprev;
    const uint32_t mp2 = min(mp + P2, 0xFFFFu);
    const uint32_t left_hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)ch.prev[K - 1], kDppWaveShr1, 0xF, 0xF, true);
    const uint32_t right_lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)ch.prev[0], kDppWaveShl1, 0xF, 0xF, true);
    uint32_t lm = 0xFFFFFFFFu, Ln[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const uint32_t pj = ch.prev[j];
        const uint32_t before = j == 0 ? left_hi : ch.prev[j - 1];
        const uint32_t after = j == K - 1 ? right_lo : ch.prev[j + 1];
        // (L'(d-1), L'(d)) and (L'(d+1), L'(d+2)) of the pair's first d, plus P1
        const u16x2 dm1 = add_sat2(as_v2(__builtin_amdgcn_alignbit(pj, before, 16)), as_v2(j == 0 ? p1l : p1s));
        const u16x2 dp1 = add_sat2(as_v2(__builtin_amdgcn_alignbit(after, pj, 16)), as_v2(j == K - 1 ? p1r : p1s));
        u16x2 t = min2(min2(dm1, dp1), as_v2(pj));
        t = min2(t, splat2(mp2));
        const u16x2 dd = t - splat2(mp);
        dl[j] = as_u32(dd);
        Ln[j] = as_u32(add_sat2(as_v2(c[j]), dd));
        lm = as_u32(min2(as_v2(lm), as_v2(Ln[j])));
    }
#pragma unroll
    for (int j = 0; j < K; ++j) ch.prev[j] = Ln[j];
    return min(lm & 0xFFFFu, lm >> 16);
}

// delta pairs (u16 halves < 256) -> bytes of the lane's plane word
template <int K>
__device__ __forceinline__ uint32_t pack_deltas(const uint32_t (&dl)[K]) {
    if constexpr (K == 1) return __builtin_amdgcn_perm(0u, dl[0], 0x0c0c0200u);
    else return __builtin_amdgcn_perm(dl[1], dl[0], 0x06040200u);
}

template <typename TC, int K, bool FULL>
__device__ __forceinline__ void sgm_lines_body(const uint3 bid, const apg_hip_params& p, int w, int h, int D,
                                               int nfam, const uint8_t* __restrict__ gray,
                                               const TC* __restrict__ cost, const PlaneSet& planes) {
    const int lane = threadIdx.x;
    // block -> (family, member): rows, then columns, then the two diagonal ray sets
    int b = bid.x, f = 0;
    {
        const int cnt[4] = {h, w, w + h - 1, w + h - 1};
        while (f < nfam - 1 && b >= cnt[f]) { b -= cnt[f]; ++f; }
    }
    int p0, step, n;
    family_line(f, b, w, h, p0, step, n);
    const auto rc = raw_rsrc(cost);
    const auto rpa = raw_rsrc(planes.p[2 * f]);
    const auto rpb = raw_rsrc(planes.p[2 * f + 1]);
    constexpr uint32_t esz = sizeof(TC);
    // lanes whose pairs lie past D read the pixel's first word and store nothing
    const bool lin = FULL || 2 * K * lane < D;
    const uint32_t cvoff = lin ? (uint32_t)(2 * K * lane) * esz : 0u;
    const uint32_t pvoff = (uint32_t)(2 * K * lane);
    bool inr[K];
#pragma unroll
    for (int j = 0; j < K; ++j) inr[j] = FULL || 2 * (K * lane + j) < D;
    const uint32_t cvoff2 = (FULL || inr[K - 1]) ? cvoff + 4 : cvoff;
    const uint32_t P1 = (uint32_t)p.P1;
    const uint32_t p1s = P1 | (P1 << 16);
    // lane 0's first d-1 and lane 63's last d+1 lie outside [0, 128K)
    const uint32_t p1l = lane == 0 ? (0xFFFFu | (P1 << 16)) : p1s;
    const uint32_t p1r = lane == 63 ? (P1 | 0xFFFF0000u) : p1s;
    const int pA0 = p0, pB0 = p0 + (n - 1) * step;
    const uint32_t pstep = (uint32_t)step * (uint32_t)D;
    Chain<K> A, B;
#pragma unroll
    for (int j = 0; j < K; ++j) A.prev[j] = B.prev[j] = 0xFFFFFFFFu;
    A.minprev = B.minprev = 0xFFFFu;
    // the scanline's gray values, staged in LDS once; adaptive P2 of the next
    // 64 steps of each chain is computed one step per lane from them
    extern __shared__ uint8_t s_gray[];   // [n]
    for (int t = lane; t < n; t += 64) s_gray[t] = gray[p0 + t * step];
    __syncthreads();
    uint32_t p2A = 0, p2B = 0;
    auto refresh_p2 = [&](int t0) {
        // chain A at step t: pixel t after t-1; chain B: pixel n-1-t after n-t
        const int tl = min(t0 + lane, n - 1);
        const int ga = s_gray[tl], gpa = s_gray[max(tl - 1, 0)];
        const int gb = s_gray[n - 1 - tl], gpb = s_gray[min(n - tl, n - 1)];
        p2A = (uint32_t)adaptive_p2(p, ga - gpa);
        p2B = (uint32_t)adaptive_p2(p, gb - gpb);
    };
    CRaw<TC, K> XA[SL_PF], XB[SL_PF], YA[SL_PF], YB[SL_PF];
    auto load = [&](CRaw<TC, K> (&ra)[SL_PF], CRaw<TC, K> (&rb)[SL_PF], int t0) {
#pragma unroll
        for (int i = 0; i < SL_PF; ++i) {
            const int t = min(t0 + i, n - 1);
            ra[i] = load_c<TC, K>(rc, cvoff, cvoff2, (uint32_t)(pA0 + t * step) * D * esz);
            rb[i] = load_c<TC, K>(rc, cvoff, cvoff2, (uint32_t)(pB0 - t * step) * D * esz);
        }
    };
    uint32_t soA = (uint32_t)pA0 * (uint32_t)D, soB = (uint32_t)pB0 * (uint32_t)D;   // plane offsets
    auto run = [&](const CRaw<TC, K> (&ra)[SL_PF], const CRaw<TC, K> (&rb)[SL_PF], int t0) {
#pragma unroll
        for (int i = 0; i < SL_PF; ++i) {
            const int t = t0 + i;
            if (t >= n) break;
            if (i == 0 && (t & 63) == 0) refresh_p2(t);
            const uint32_t P2a = (uint32_t)__builtin_amdgcn_readlane((int)p2A, t & 63);
            const uint32_t P2b = (uint32_t)__builtin_amdgcn_readlane((int)p2B, t & 63);
            uint32_t ca[K], cb[K], da[K], db[K];
#pragma unroll
            for (int j = 0; j < K; ++j) {
                ca[j] = decode_c<TC, K>(ra[i], j);
                cb[j] = decode_c<TC, K>(rb[i], j);
                if (!FULL && !inr[j]) ca[j] = cb[j] = 0xFFFFFFFFu;
            }
            uint32_t ma = chain_step<K>(A, ca, P2a, p1s, p1l, p1r, da);
            uint32_t mb = chain_step<K>(B, cb, P2b, p1s, p1l, p1r, db);
            wave_min2_u32(ma, mb);
            A.minprev = ma;
            B.minprev = mb;
            const uint32_t wa = pack_deltas<K>(da), wb = pack_deltas<K>(db);
            if (lin) {
                if (K == 1 || (!FULL && !inr[K - 1])) {
                    // K == 2 and only pair 0 inside D: the lane's word is half a word
                    buf_st16(rpa, wa, pvoff, soA);
                    buf_st16(rpb, wb, pvoff, soB);
                } else {
                    buf_st32(rpa, wa, pvoff, soA);
                    buf_st32(rpb, wb, pvoff, soB);
                }
            }
            soA += pstep;
            soB -= pstep;
        }
    };
    load(XA, XB, 0);
    load(YA, YB, SL_PF);
    for (int t0 = 0; t0 < n; t0 += 2 * SL_PF) {
        run(XA, XB, t0);
        load(XA, XB, t0 + 2 * SL_PF);
        run(YA, YB, t0 + SL_PF);
        load(YA, YB, t0 + 3 * SL_PF);
    }
}

template <typename TC, int K, bool FULL>
__global__ __launch_bounds__(64) void sgm_lines_kernel(apg_hip_params p, int w, int h, int D, int nfam,
                                                       const uint8_t* __restrict__ gray, const TC* __restrict__ cost,
                                                       PlaneSet planes) {
    sgm_lines_body<TC, K, FULL>(block3(), p, w, h, D, nfam, gray, cost, planes);
}

template <typename TC, int K, bool FULL>
struct SgmLinesFn {
    static constexpr int kThreads = 64;
    template <typename... A>
    __device__ static void run(const uint3 b, const A&... a) { sgm_lines_body<TC, K, FULL>(b, a...); }
};

// Batched-transfer form for u8 C' with D % 16 == 0, D <= 256.  Each chain
// moves SB steps of C' (SB pixel rows) from HBM into LDS with one 16-byte
// load per lane, reads its own 2K bytes per step from LDS one step ahead, and
// stages its deltas in LDS, written back SB steps at a time with one 16-byte
// store per lane.  That is 4 vector-memory instructions per SB steps of the
// wave instead of 4 per step, so the loads of the next batch stay in flight
// however the wait counts are placed.
constexpr int SB = 8;   // steps per staged batch

// One scanline chain job (or both chains of the line when DUAL), all state
// local: waves that run several jobs reuse the LDS staging buffers.
template <int K, bool FULL, bool DUAL>
__device__ __forceinline__ void sgm_job(const apg_hip_params& p, int w, int h, int D, int f, int b, int chain,
                                        const uint8_t* __restrict__ gray, const uint8_t* __restrict__ cost,
                                        const PlaneSet& planes, uint8_t* sh) {
    constexpr int RB = 128 * K;   // staged row bytes (>= D)
    uint8_t* s_c = sh;                       // [chain][buffer][SB][RB] C' rows
    uint8_t* s_d = sh + 4 * SB * RB;         // [chain][buffer][SB][RB] delta rows
    uint8_t* s_gray = sh + 8 * SB * RB;      // [n]
    const int lane = threadIdx.x;
    int p0, step, n;
    family_line(f, b, w, h, p0, step, n);
    const auto rc = raw_rsrc(cost);
    // chain A walks from the line's first pixel; a single-chain wave of the
    // backward direction is chain A of the reversed line
    const int pF = p0, pL = p0 + (n - 1) * step;
    const int pA0 = chain ? pL : pF, pB0 = chain ? pF : pL;
    if (chain) step = -step;
    // columns whose every disparity is valid (0 <= x - d < w for all d): no
    // invalid C' byte there, so a batch inside them skips the sentinel decode
    const int xlo = p.min_disparity + D - 1, xhi = w - 1 + p.min_disparity;
    const int dxA = f == 1 ? 0 : (chain ? -1 : 1);   // x step of chain A: rows and rays move +x forward
    const int xA0 = pA0 % w, xB0 = pB0 % w;
    auto batch_valid = [&](int k) {
        if (planes.general_ranges) return false;
        const int t0 = k * SB, t1 = min(k * SB + SB, n) - 1;
        const int a0 = xA0 + dxA * t0, a1 = xA0 + dxA * t1;
        const int b0 = xB0 - dxA * t0, b1 = xB0 - dxA * t1;
        bool v = min(a0, a1) >= xlo && max(a0, a1) <= xhi;
        if (DUAL) v = v && min(b0, b1) >= xlo && max(b0, b1) <= xhi;
        return v;
    };
    const auto rpa = raw_rsrc(planes.p[2 * f + chain]);
    const auto rpb = raw_rsrc(planes.p[2 * f + 1]);
    // batch transfer: lane -> (row j, 16-byte chunk c); NLD instructions cover SB rows
    const int lpx = D / 16, ppl = 64 / lpx;
    constexpr int NLD = K == 1 ? 1 : 2;
    const int jl = lane / lpx, cl = lane % lpx;
    bool inr[K];
#pragma unroll
    for (int j = 0; j < K; ++j) inr[j] = FULL || 2 * (K * lane + j) < D;
    const uint32_t P1 = (uint32_t)p.P1;
    const uint32_t p1s = P1 | (P1 << 16);
    const uint32_t p1l = lane == 0 ? (0xFFFFu | (P1 << 16)) : p1s;
    const uint32_t p1r = lane == 63 ? (P1 | 0xFFFF0000u) : p1s;
    Chain<K> A, B;
#pragma unroll
    for (int j = 0; j < K; ++j) A.prev[j] = B.prev[j] = 0xFFFFFFFFu;
    A.minprev = B.minprev = 0xFFFFu;
    for (int t = lane; t < n; t += 64) s_gray[t] = gray[pA0 + t * step];
    uint32_t p2A = 0, p2B = 0;
    auto refresh_p2 = [&](int t0) {
        const int tl = min(t0 + lane, n - 1);
        const int ga = s_gray[tl], gpa = s_gray[max(tl - 1, 0)];
        const int gb = s_gray[n - 1 - tl], gpb = s_gray[min(n - tl, n - 1)];
        p2A = (uint32_t)adaptive_p2(p, ga - gpa);
        p2B = (uint32_t)adaptive_p2(p, gb - gpb);
    };
    // byte offset of (row j, chunk c) of batch k for chain A / B
    auto voffA = [&](int k, int j) { return (uint32_t)(pA0 + min(k * SB + j, n - 1) * step) * (uint32_t)D + cl * 16u; };
    auto voffB = [&](int k, int j) { return (uint32_t)(pB0 - min(k * SB + j, n - 1) * step) * (uint32_t)D + cl * 16u; };
    u32x4 gA[NLD], gB[NLD];
    auto gload = [&](int k) {
#pragma unroll
        for (int r = 0; r < NLD; ++r) {
            const int j = min(r * ppl + jl, SB - 1);
            gA[r] = buf_ld128(rc, voffA(k, j), 0u);
            if (DUAL) gB[r] = buf_ld128(rc, voffB(k, j), 0u);
        }
    };
    auto lstage = [&](int k) {   // batch k's rows -> LDS buffer k & 1
        const int bf = k & 1;
#pragma unroll
        for (int r = 0; r < NLD; ++r) {
            const int j = r * ppl + jl;
            if (j < SB) {
                *(u32x4*)(s_c + ((0 * 2 + bf) * SB + j) * RB + cl * 16) = gA[r];
                if (DUAL) *(u32x4*)(s_c + ((1 * 2 + bf) * SB + j) * RB + cl * 16) = gB[r];
            }
        }
    };
    auto flush = [&](int k) {    // deltas of batch k (LDS buffer k & 1) -> planes
        const int bf = k & 1;
#pragma unroll
        for (int r = 0; r < NLD; ++r) {
            const int j = r * ppl + jl;
            if (j < SB && k * SB + j < n) {
                const u32x4 va = *(const u32x4*)(s_d + ((0 * 2 + bf) * SB + j) * RB + cl * 16);
                __builtin_amdgcn_raw_buffer_store_b128(va, rpa, (int)voffA(k, j), 0, 0);
                if (DUAL) {
                    const u32x4 vb = *(const u32x4*)(s_d + ((1 * 2 + bf) * SB + j) * RB + cl * 16);
                    __builtin_amdgcn_raw_buffer_store_b128(vb, rpb, (int)voffB(k, j), 0, 0);
                }
            }
        }
    };
    auto cword = [&](int chain, int k, int i) -> uint32_t {   // the lane's C' bytes of step i of batch k
        const uint8_t* q = s_c + ((chain * 2 + (k & 1)) * SB + i) * RB + 2 * K * lane;
        return K == 1 ? (uint32_t)*(const uint16_t*)q : *(const uint32_t*)q;
    };
    const int nb = (n + SB - 1) / SB;
    gload(0);
    lstage(0);
    if (nb > 1) gload(1);
    uint32_t nA = cword(0, 0, 0), nB = DUAL ? cword(1, 0, 0) : 0u;
    for (int k = 0; k < nb; ++k) {
        // staged rows first: the stores and loads issued after them are newer,
        // so waiting for the rows never waits for a store
        if (k + 1 < nb) lstage(k + 1);
        if (k > 0) flush(k - 1);
        if (k + 2 < nb) gload(k + 2);
        const int bf = k & 1;
        auto run_batch = [&](auto allv_tag) {
        constexpr bool ALLV = decltype(allv_tag)::value;
#pragma unroll
        for (int i = 0; i < SB; ++i) {
            const int t = k * SB + i;
            if (t >= n) break;
            if (i == 0 && (t & 63) == 0) refresh_p2(t);
            const uint32_t wa = nA, wb = nB;
            // next step's words (the next batch's rows are already staged)
            if (i + 1 < SB) {
                nA = cword(0, k, i + 1);
                if (DUAL) nB = cword(1, k, i + 1);
            } else {
                nA = cword(0, k + 1, 0);
                if (DUAL) nB = cword(1, k + 1, 0);
            }
            const uint32_t P2a = (uint32_t)__builtin_amdgcn_readlane((int)p2A, t & 63);
            uint32_t ca[K], cb[K], da[K], db[K];
#pragma unroll
            for (int j = 0; j < K; ++j) {
                CRaw<uint8_t, K> xa, xb;
                xa.w[0] = wa;
                xb.w[0] = wb;
                ca[j] = decode_c<uint8_t, K, ALLV>(xa, j);
                cb[j] = DUAL ? decode_c<uint8_t, K, ALLV>(xb, j) : 0u;
                if (!FULL && !inr[j]) ca[j] = cb[j] = 0xFFFFFFFFu;
            }
            uint32_t ma = chain_step<K>(A, ca, P2a, p1s, p1l, p1r, da);
            if (DUAL) {
                const uint32_t P2b = (uint32_t)__builtin_amdgcn_readlane((int)p2B, t & 63);
                uint32_t mb = chain_step<K>(B, cb, P2b, p1s, p1l, p1r, db);
                wave_min2_u32(ma, mb);
                B.minprev = mb;
            } else {
                ma = wave_min_u32(ma);
            }
            A.minprev = ma;
            uint8_t* qa = s_d + ((0 * 2 + bf) * SB + i) * RB + 2 * K * lane;
            uint8_t* qb = s_d + ((1 * 2 + bf) * SB + i) * RB + 2 * K * lane;
            if (K == 1) {
                *(uint16_t*)qa = (uint16_t)pack_deltas<K>(da);
                if (DUAL) *(uint16_t*)qb = (uint16_t)pack_deltas<K>(db);
            } else {
                *(uint32_t*)qa = pack_deltas<K>(da);
                if (DUAL) *(uint32_t*)qb = pack_deltas<K>(db);
            }
        }
        };
        if (batch_valid(k)) run_batch(std::true_type{});
        else run_batch(std::false_type{});
    }
    flush(nb - 1);
}

// DUAL: one wave runs both chains of its scanline (two independent
// recurrences interleaved).  Otherwise a wave runs one chain (forward or
// backward) of one scanline: twice the waves, each with half the work, which
// measured faster at 1280x720 D=128 (more waves share each SIMD's issue
// slots while the longest scanlines bound the duration).
template <int K, bool FULL, bool DUAL>
__device__ __forceinline__ void sgm_lines4_body(const uint3 blk, const apg_hip_params& p, int w, int h, int D,
                                                int nfam, const uint8_t* __restrict__ gray,
                                                const uint8_t* __restrict__ cost, const PlaneSet& planes) {
    extern __shared__ __align__(16) uint8_t sh[];
    const int bid = blk.x;
    if (DUAL) {
        int b = bid, f = 0;
        const int cnt[4] = {h, w, w + h - 1, w + h - 1};
        while (f < nfam - 1 && b >= cnt[f]) { b -= cnt[f]; ++f; }
        sgm_job<K, FULL, true>(p, w, h, D, f, b, 0, gray, cost, planes, sh);
    } else {
        int b = bid >> 1, f = 0;
        const int cnt[4] = {h, w, w + h - 1, w + h - 1};
        while (f < nfam - 1 && b >= cnt[f]) { b -= cnt[f]; ++f; }
        sgm_job<K, FULL, false>(p, w, h, D, f, b, bid & 1, gray, cost, planes, sh);
    }
}

template <int K, bool FULL, bool DUAL>
__global__ __launch_bounds__(64) void sgm_lines4_kernel(apg_hip_params p, int w, int h, int D, int nfam,
                                                        const uint8_t* __restrict__ gray,
                                                        const uint8_t* __restrict__ cost, PlaneSet planes) {
    sgm_lines4_body<K, FULL, DUAL>(block3(), p, w, h, D, nfam, gray, cost, planes);
}

template <int K, bool FULL, bool DUAL>
struct SgmLines4Fn {
    static constexpr int kThreads = 64;
    template <typename... A>
    __device__ static void run(const uint3 b, const A&... a) { sgm_lines4_body<K, FULL, DUAL>(b, a...); }
};

// Half-wave form for u8 C' with D % 16 == 0, D <= 128: one wave runs both
// chains of its scanline, chain A (forward) in lanes 0-31 and chain B
// (backward) in lanes 32-63.  Half-lane l holds the disparities 4l..4l+3 as
// two packed u16 pairs, so the pair (4l+1, 4l+2) is the +1 neighbour of the
// first pair and the -1 neighbour of the second, and each half reduces its
// minimum with four in-row rotations and one row swap (v_permlane16_swap):
// about 31 VALU per step of two chains against about 25 per step of one
// chain in the 64-lane form.  minL' and the per-step P2 live in VGPRs as
// u16 pairs (one value per half), so the recurrence needs no readlane.
// The C' rows of PF batches are in flight per wave (register slot k % PF
// holds batch k); the loop runs PF batches per trip so every slot index is
// static.
template <int N, typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
// f(integral_constant<int, 0>) ... f(integral_constant<int, N-1>)
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl<N>(f, std::make_integer_sequence<int, N>{});
}

// plane of (family f, chain c): families whose bit is set in planes.paired
// hold one plane (both chains), the others two
__device__ __forceinline__ uint8_t* family_plane(const PlaneSet& planes, int f, int c) {
    int base = 0;
    for (int g = 0; g < f; ++g) base += (planes.paired >> g) & 1 ? 1 : 2;
    return planes.p[base + ((planes.paired >> f) & 1 ? 0 : c)];
}

template <bool FULL, int PF, bool PAIRED>
__device__ __forceinline__ void sgm_half_body(const apg_hip_params& p, int w, int h, int D, int f, int b,
                                              const uint8_t* __restrict__ gray, const uint8_t* __restrict__ cost,
                                              const PlaneSet& planes, int nmax, uint8_t* sh) {
    constexpr int RB = 128;                  // staged row bytes (>= D)
    uint8_t* s_c = sh;                       // [chain][buffer][SB][RB] C' rows
    uint8_t* s_d = sh + 4 * SB * RB;         // [chain][buffer][SB][RB] delta rows
    uint8_t* s_p2 = sh + 8 * SB * RB;        // [chain][nmax] P2 per step
    uint8_t* s_mid = s_p2 + 2 * nmax;        // PAIRED: [kPairGap + 1][RB] first deltas near the crossing,
                                             // then one spare row
    const int lane = threadIdx.x;
    const int hf = lane >> 5, hl = lane & 31;
    int p0, step, n;
    family_line(f, b, w, h, p0, step, n);
    const auto rc = raw_rsrc(cost);
    const int pA0 = p0, pB0 = p0 + (n - 1) * step;
    const int xlo = p.min_disparity + D - 1, xhi = w - 1 + p.min_disparity;
    const int dxA = f == 1 ? 0 : 1;
    const int xA0 = pA0 % w, xB0 = pB0 % w;
    auto batch_valid = [&](int k) {
        if (planes.general_ranges) return false;
        const int t0 = k * SB, t1 = min(k * SB + SB, n) - 1;
        const int a0 = xA0 + dxA * t0, a1 = xA0 + dxA * t1;
        const int b0 = xB0 - dxA * t0, b1 = xB0 - dxA * t1;
        return min(a0, a1) >= xlo && max(a0, a1) <= xhi && min(b0, b1) >= xlo && max(b0, b1) <= xhi;
    };
    // PAIRED: one plane per family holds delta_A + delta_B (both <= P2 <= 127).
    // The first chain to reach a position parks its delta there; the second
    // adds its own and stores the sum.  Positions within kPairGap / 2 of the
    // crossing keep the first delta in s_mid instead, so a parked delta is
    // always stored at least four batches before the second chain loads it.
    const auto rpa = raw_rsrc(family_plane(planes, f, 0));
    const auto rpb = raw_rsrc(family_plane(planes, f, 1));
    // P2 of every step of both chains: chain A at step t is pixel t after t-1,
    // chain B pixel n-1-t after n-t (the first step of each compares a pixel
    // with itself).  Each trip issues the gray loads of eight steps per lane
    // before using any of them.
    for (int t0 = 0; t0 < n; t0 += 512) {
        int ga[8], gpa[8], gb[8], gpb[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int t = min(t0 + lane + 64 * j, n - 1), u = n - 1 - t;
            ga[j] = gray[pA0 + t * step];
            gpa[j] = gray[pA0 + max(t - 1, 0) * step];
            gb[j] = gray[pA0 + u * step];
            gpb[j] = gray[pA0 + min(u + 1, n - 1) * step];
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int t = t0 + lane + 64 * j;
            if (t < n) {
                s_p2[t] = (uint8_t)adaptive_p2(p, ga[j] - gpa[j]);
                s_p2[nmax + t] = (uint8_t)adaptive_p2(p, gb[j] - gpb[j]);
            }
        }
    }
    // the P2 words past the last step are read with their batch (never used)
    for (int t = n + lane; t < ((n + SB - 1) / SB) * SB; t += 64) {
        s_p2[t] = 0;
        s_p2[nmax + t] = 0;
    }
    // batch transfer: lane -> (row j, 16-byte chunk c); one instruction per chain covers SB rows
    const int lpx = D / 16;
    const int jl = lane / lpx, cl = lane % lpx;
    const bool inr = FULL || 4 * hl < D;
    const uint32_t P1 = (uint32_t)p.P1;
    const uint32_t p1s = P1 | (P1 << 16);
    // half-lane 0's d-1 and half-lane 31's d+1 lie outside [0, 128)
    const uint32_t p1l = hl == 0 ? (0xFFFFu | (P1 << 16)) : p1s;
    const uint32_t p1r = hl == 31 ? (P1 | 0xFFFF0000u) : p1s;
    uint32_t prev0 = 0xFFFFFFFFu, prev1 = 0xFFFFFFFFu, mp = 0xFFFFu;   // minL' of the lane's chain
    auto voffA = [&](int k, int j) { return (uint32_t)(pA0 + min(k * SB + j, n - 1) * step) * (uint32_t)D + cl * 16u; };
    auto voffB = [&](int k, int j) { return (uint32_t)(pB0 - min(k * SB + j, n - 1) * step) * (uint32_t)D + cl * 16u; };
    // batch k's rows arrive in register slot k % PF, PF batches ahead of use
    u32x4 gA[PF], gB[PF];
    // PAIRED: parked deltas of batch k (register slot k % PF), loaded with the
    // C' rows of batch k + PF - 2 (two batches before the flush that needs them)
    u32x4 pkA[PF], pkB[PF];
    // kp = k - 2: PAIRED, the batch whose parked deltas are loaded into slot kp % PF.
    // Only a far second arrival's row is read; the other lanes read the plane's
    // first 16 bytes (inside every plane: D >= 16), one cache line per load.
    auto gload = [&](int k, int kp, auto slot_t) {
        constexpr int q = decltype(slot_t)::value;
        const int j = min(jl, SB - 1);
        gA[q] = buf_ld128(rc, voffA(k, j), 0u);
        gB[q] = buf_ld128(rc, voffB(k, j), 0u);
        if constexpr (PAIRED) {
            constexpr int qp = (q + PF - 2) % PF;
            const int t = kp * SB + jl;
            const bool far2 = jl < SB && t < n && 2 * t - (n - 1) >= kPairGap;
            pkA[qp] = buf_ld128_dev(rpa, far2 ? voffA(kp, jl) : 0u, 0u);
            pkB[qp] = buf_ld128_dev(rpb, far2 ? voffB(kp, jl) : 0u, 0u);
        }
    };
    auto lstage = [&](int k, auto slot_t) {   // batch k's rows -> LDS buffer k & 1
        constexpr int q = decltype(slot_t)::value;
        const int bf = k & 1;
        if (jl < SB) {
            *(u32x4*)(s_c + ((0 * 2 + bf) * SB + jl) * RB + cl * 16) = gA[q];
            *(u32x4*)(s_c + ((1 * 2 + bf) * SB + jl) * RB + cl * 16) = gB[q];
        }
    };
    const int nb = (n + SB - 1) / SB;
    auto flush = [&](int k, auto slot_t) {    // deltas of batch k (LDS buffer k & 1) -> planes
        constexpr int q = decltype(slot_t)::value;
        const int bf = k & 1;
        const int t = k * SB + jl;
        if (jl < SB && t < n) {
            u32x4 va = *(const u32x4*)(s_d + ((0 * 2 + bf) * SB + jl) * RB + cl * 16);
            u32x4 vb = *(const u32x4*)(s_d + ((1 * 2 + bf) * SB + jl) * RB + cl * 16);
            if constexpr (PAIRED) {
                // A reaches its row's position t first iff e <= 0, B reaches
                // n-1-t first iff e < 0.  A first arrival parks its delta in the
                // plane (far from the crossing) or copies it to s_mid (near; the
                // plane store goes to the spare bytes).  A second arrival stores
                // its delta plus the first's: the parked one (far), the other
                // chain's staged row when both arrivals fall in this batch, else
                // the s_mid copy made by an earlier flush.
                const int e = 2 * t - (n - 1);
                const bool near = e > -kPairGap && e < kPairGap;
                const int t1 = n - 1 - t;                    // the other chain's step at this row's positions
                const bool same = t1 >= k * SB && t1 < k * SB + SB;
                const int midA = t - ((n - 1) >> 1) + kPairGap / 2;        // s_mid row of position t
                const int midB = (n - 1 - t) - ((n - 1) >> 1) + kPairGap / 2;
                const int spare_row = kPairGap + 1;
                const uint8_t* srcA = same ? s_d + ((1 * 2 + bf) * SB + (t1 - k * SB)) * RB
                                           : s_mid + (near && e > 0 ? midA : spare_row) * RB;
                const uint8_t* srcB = same ? s_d + ((0 * 2 + bf) * SB + (t1 - k * SB)) * RB
                                           : s_mid + (near && e >= 0 ? midB : spare_row) * RB;
                const u32x4 oa = *(const u32x4*)(srcA + cl * 16), ob = *(const u32x4*)(srcB + cl * 16);
                *(u32x4*)(s_mid + (near && e <= 0 ? midA : spare_row) * RB + cl * 16) = va;
                *(u32x4*)(s_mid + (near && e < 0 ? midB : spare_row) * RB + cl * 16) = vb;
                const u32x4 z = {0u, 0u, 0u, 0u};
                va = va + (e <= 0 ? z : (near ? oa : pkA[q]));
                vb = vb + (e < 0 ? z : (near ? ob : pkB[q]));
                const uint32_t spare = (uint32_t)((size_t)w * h * D) + cl * 16u;
                __builtin_amdgcn_raw_buffer_store_b128(va, rpa, (int)(near && e <= 0 ? spare : voffA(k, jl)), 0, 0);
                __builtin_amdgcn_raw_buffer_store_b128(vb, rpb, (int)(near && e < 0 ? spare : voffB(k, jl)), 0, 0);
            } else {
                __builtin_amdgcn_raw_buffer_store_b128(va, rpa, (int)voffA(k, jl), 0, 0);
                __builtin_amdgcn_raw_buffer_store_b128(vb, rpb, (int)voffB(k, jl), 0, 0);
            }
        }
    };
    // the lane's C' words (4 d of its chain) and P2 pairs of all SB steps of a
    // batch are read into register set k & 1 one batch ahead, so no LDS
    // latency lies on the recurrence
    const uint32_t lofs = (uint32_t)(hf * 2 * SB * RB + 4 * hl);
    const uint8_t* p2h = s_p2 + hf * nmax;
    uint32_t W[2][SB], P[2][SB];
    auto wread = [&](int k, auto set_t) {
        constexpr int e = decltype(set_t)::value;
#pragma unroll
        for (int i = 0; i < SB; ++i) W[e][i] = *(const uint32_t*)(s_c + lofs + ((k & 1) * SB + i) * RB);
        const uint2 pb = *(const uint2*)(p2h + k * SB);   // P2 of the batch's steps, one byte each
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t sel = 0x0c000c00u | (uint32_t)i | ((uint32_t)i << 16);   // byte i -> both halves
            P[e][i] = __builtin_amdgcn_perm(0u, pb.x, sel);
            P[e][4 + i] = __builtin_amdgcn_perm(0u, pb.y, sel);
        }
    };
    // loads are issued unconditionally (past the end: the last batch again), so
    // the wait before staging a batch counts the loads issued after it and
    // leaves them in flight
    static_for<PF>([&](auto q_t) {
        constexpr int q = decltype(q_t)::value;
        gload(min(q, nb - 1), q - 2, q_t);   // parked deltas of batches -2 .. 1: never far second arrivals
    });
    lstage(0, std::integral_constant<int, 0>{});
    __syncthreads();
    wread(0, std::integral_constant<int, 0>{});
    // INNER: batch k is not the last one (all SB steps exist, batch k + 1 exists)
    auto batch = [&](int k, auto q_t, auto inner_t) {
        constexpr int q = decltype(q_t)::value;
        constexpr bool INNER = decltype(inner_t)::value;
        // staged rows first: the stores and loads issued after them are newer,
        // so waiting for the rows never waits for a store
        if (INNER || k + 1 < nb) {
            lstage(k + 1, std::integral_constant<int, (q + 1) % PF>{});
            wread(k + 1, std::integral_constant<int, (q + 1) & 1>{});
        }
        // the previous batch's deltas: in an inner batch after half its steps, so
        // the wait for the LDS reads of the flush finds the word reads done
        using PrevSlot = std::integral_constant<int, (q + PF - 1) % PF>;   // slot of batch k - 1
        if (!INNER && k > 0) flush(k - 1, PrevSlot{});
        gload(min(k + PF, nb - 1), k + PF - 2, q_t);
        const int bf = k & 1;
        auto run_batch = [&](auto allv_tag) {
            constexpr bool ALLV = decltype(allv_tag)::value;
#pragma unroll
            for (int i = 0; i < SB; ++i) {
                const int t = k * SB + i;
                if (!INNER && t >= n) break;
                if (INNER && i == SB / 2 && k > 0) flush(k - 1, PrevSlot{});
                const uint32_t wc = W[q & 1][i], p2p = P[q & 1][i];
                CRaw<uint8_t, 2> x;
                x.w[0] = wc;
                uint32_t c0 = decode_c<uint8_t, 2, ALLV>(x, 0), c1 = decode_c<uint8_t, 2, ALLV>(x, 1);
                if (!FULL && !inr) c0 = c1 = 0xFFFFFFFFu;
                // L = C + min(L'(d), L'(d-1)+P1, L'(d+1)+P1, minL'+P2) - minL'
                //   = C + min(u - minL', P2), u = min(L'(d), L'(d+-1)+P1) >= minL'
                // (path values stay far below 65535 - P2 unless all are 65535,
                // where u = minL' and the delta is 0), so after the reduction
                // only three operations remain on the recurrence
                const uint32_t left_hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)prev1, kDppWaveShr1, 0xF, 0xF, true);
                const uint32_t right_lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)prev0, kDppWaveShl1, 0xF, 0xF, true);
                const u16x2 sl = add_sat2(as_v2(__builtin_amdgcn_alignbit(prev0, left_hi, 16)), as_v2(p1l));
                const u16x2 sm = add_sat2(as_v2(__builtin_amdgcn_alignbit(prev1, prev0, 16)), as_v2(p1s));
                const u16x2 sr = add_sat2(as_v2(__builtin_amdgcn_alignbit(right_lo, prev1, 16)), as_v2(p1r));
                const u16x2 u0 = min2(min2(sl, sm), as_v2(prev0));
                const u16x2 u1 = min2(min2(sm, sr), as_v2(prev1));
                const u16x2 d0 = min2(u0 - splat2(mp), as_v2(p2p)), d1 = min2(u1 - splat2(mp), as_v2(p2p));
                prev0 = as_u32(add_sat2(as_v2(c0), d0));
                prev1 = as_u32(add_sat2(as_v2(c1), d1));
                const uint32_t lm2 = as_u32(min2(as_v2(prev0), as_v2(prev1)));
                uint32_t m = min(lm2 & 0xFFFFu, lm2 >> 16);
                // minimum over the half: rotations within each 16-lane row, then
                // rows 0/1 and 2/3 exchanged
                m = min(m, dpp<kDppRowRor8>(0xFFFFFFFFu, m));
                m = min(m, dpp<kDppRowRor4>(0xFFFFFFFFu, m));
                m = min(m, dpp<kDppRowRor2>(0xFFFFFFFFu, m));
                m = min(m, dpp<kDppRowRor1>(0xFFFFFFFFu, m));
                const auto sw = __builtin_amdgcn_permlane16_swap(m, m, false, false);
                mp = min((uint32_t)sw[0], (uint32_t)sw[1]);
                *(uint32_t*)(s_d + lofs + (bf * SB + i) * RB) =
                    __builtin_amdgcn_perm(as_u32(d1), as_u32(d0), 0x06040200u);
            }
        };
        if (!INNER || batch_valid(k)) run_batch(std::integral_constant<bool, INNER>{});
        else run_batch(std::false_type{});
    };
    // groups of PF inner batches (batch k in slot k % PF), then the rest
    int k0 = 0;
    for (; k0 + PF < nb; k0 += PF)
        static_for<PF>([&](auto q_t) { batch(k0 + decltype(q_t)::value, q_t, std::true_type{}); });
    static_for<PF>([&](auto q_t) {
        const int k = k0 + decltype(q_t)::value;
        if (k < nb) batch(k, q_t, std::false_type{});
    });
    static_for<PF>([&](auto q_t) {
        if ((nb - 1) % PF == decltype(q_t)::value) flush(nb - 1, q_t);
    });
}

// Scanline family f, member b of block blockIdx.x.  PAIRED: every family's
// bit is set in planes.paired (sgm_pair_mask pairs all or none).
template <bool FULL, int PF, bool PAIRED>
__device__ __forceinline__ void sgm_half_body(const uint3 bid, const apg_hip_params& p, int w, int h, int D, int nfam,
                                              const uint8_t* __restrict__ gray, const uint8_t* __restrict__ cost,
                                              const PlaneSet& planes, int nmax) {
    extern __shared__ __align__(16) uint8_t sh[];
    int b = bid.x, f = 0;
    {
        const int cnt[4] = {h, w, w + h - 1, w + h - 1};
        while (f < nfam - 1 && b >= cnt[f]) { b -= cnt[f]; ++f; }
    }
    sgm_half_body<FULL, PF, PAIRED>(p, w, h, D, f, b, gray, cost, planes, nmax, sh);
}

template <bool FULL, int PF, bool PAIRED>
__global__ __launch_bounds__(64) void sgm_half_kernel(apg_hip_params p, int w, int h, int D, int nfam,
                                                      const uint8_t* __restrict__ gray,
                                                      const uint8_t* __restrict__ cost, PlaneSet planes, int nmax) {
    sgm_half_body<FULL, PF, PAIRED>(block3(), p, w, h, D, nfam, gray, cost, planes, nmax);
}

template <bool FULL, int PF, bool PAIRED>
struct SgmHalfFn {
    static constexpr int kThreads = 64;
    template <typename... A>
    __device__ static void run(const uint3 b, const A&... a) { sgm_half_body<FULL, PF, PAIRED>(b, a...); }
};

// Quarter-wave form for u8 C', 4 paths, D % 16 == 0, D <= 128: one wave runs
// the forward and backward chains of two scanlines of one family (rows 2i,
// 2i+1 or columns 2j, 2j+1), 16 lanes per chain (chain q = lane >> 4: q & 1
// the direction, q >> 1 the scanline).  Half-lane l holds the disparities
// 8l..8l+7 as four packed u16 pairs; the three inner neighbour pairs are
// shared by adjacent pairs, and each chain's minimum is one 16-lane row
// reduced by four rotations.  About 48 VALU per step of four chains, and one
// wave per SIMD at 1280x720, so the recurrence's latency is hidden by the
// four pairs' independent work rather than by other waves.  Measured 225 us
// against 231 us for the half-wave form at 1280x720 D=128, but slower on
// most other shapes (one wave per SIMD exposes the recurrence's latency when
// the scanlines are long or few): a launch-shape option, not the default.
template <bool FULL, int PF>
__device__ __forceinline__ void sgm_quad_body(const uint3 bid, const apg_hip_params& p, int w, int h, int D,
                                              const uint8_t* __restrict__ gray, const uint8_t* __restrict__ cost,
                                              const PlaneSet& planes, int nmax) {
    extern __shared__ __align__(16) uint8_t sh[];
    constexpr int RB = 128;                  // staged row bytes (>= D)
    constexpr int NC = 4;                    // chains per wave
    uint8_t* s_c = sh;                       // [chain][buffer][SB][RB] C' rows
    uint8_t* s_d = sh + NC * 2 * SB * RB;    // [chain][buffer][SB][RB] delta rows
    uint32_t* s_p2 = (uint32_t*)(sh + 2 * NC * 2 * SB * RB);   // [chain][nmax] P2 | P2 << 16 per step
    const int lane = threadIdx.x;
    const int hq = lane >> 4, hl = lane & 15;
    // block -> family (rows, then columns) and its scanline pair
    const int nrow = (h + 1) / 2;
    const int f = (int)bid.x < nrow ? 0 : 1;
    const int sp = f == 0 ? (int)bid.x : (int)bid.x - nrow;
    const int nline = f == 0 ? h : w;
    // scanlines s0 = 2sp and s1 = 2sp + 1 (s0 again past the last one: the same
    // values are computed twice and stored twice)
    int p0a, step, n;
    family_line(f, 2 * sp, w, h, p0a, step, n);
    int p0b = p0a;
    if (2 * sp + 1 < nline) { int st2, n2; family_line(f, 2 * sp + 1, w, h, p0b, st2, n2); }
    const auto rc = raw_rsrc(cost);
    // chain q: first pixel and pixel step
    auto cfirst = [&](int q) { const int pl = (q >> 1) ? p0b : p0a; return (q & 1) ? pl + (n - 1) * step : pl; };
    auto cstep = [&](int q) { return (q & 1) ? -step : step; };
    const int xlo = p.min_disparity + D - 1, xhi = w - 1 + p.min_disparity;
    auto batch_valid = [&](int k) {
        if (planes.general_ranges) return false;
        const int t0 = k * SB, t1 = min(k * SB + SB, n) - 1;
        if (f == 1) {   // columns: x fixed per scanline
            const int xa = p0a % w, xb2 = p0b % w;
            return min(xa, xb2) >= xlo && max(xa, xb2) <= xhi;
        }
        // rows: forward chains at x = t, backward at x = w - 1 - t
        return t0 >= xlo && t1 <= xhi && w - 1 - t1 >= xlo && w - 1 - t0 <= xhi;
    };
    // P2 of every step of each chain (forward: pixel t after t-1; backward:
    // pixel n-1-t after n-t)
    for (int i = lane; i < 2 * n; i += 64) {
        const int s = i >= n, t = i - s * n;
        const int pl = s ? p0b : p0a;
        const int ga = gray[pl + t * step], gpa = gray[pl + max(t - 1, 0) * step];
        const int u = n - 1 - t;
        const int gb = gray[pl + u * step], gpb = gray[pl + min(u + 1, n - 1) * step];
        const uint32_t a = (uint32_t)adaptive_p2(p, ga - gpa), bb = (uint32_t)adaptive_p2(p, gb - gpb);
        s_p2[(2 * s) * nmax + t] = a | (a << 16);
        s_p2[(2 * s + 1) * nmax + t] = bb | (bb << 16);
    }
    for (int t = n + lane; t < ((n + SB - 1) / SB) * SB; t += 64)
#pragma unroll
        for (int q = 0; q < NC; ++q) s_p2[q * nmax + t] = 0u;
    // batch transfer: lane -> (row j, 16-byte chunk c); one instruction per chain covers SB rows
    const int lpx = D / 16;
    const int jl = lane / lpx, cl = lane % lpx;
    const bool inr = FULL || 8 * hl < D;
    const uint32_t P1 = (uint32_t)p.P1;
    const uint32_t p1s = P1 | (P1 << 16);
    // half-lane 0's d-1 and half-lane 15's d+1 lie outside [0, 128)
    const uint32_t p1l = hl == 0 ? (0xFFFFu | (P1 << 16)) : p1s;
    const uint32_t p1r = hl == 15 ? (P1 | 0xFFFF0000u) : p1s;
    uint32_t pv[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mp = 0xFFFFu;
    auto voff = [&](int q, int k, int j) {
        return (uint32_t)(cfirst(q) + min(k * SB + j, n - 1) * cstep(q)) * (uint32_t)D + cl * 16u;
    };
    // batch k's rows arrive in register slot k % PF, PF batches ahead of use
    u32x4 g[PF][NC];
    auto gload = [&](int k, auto slot_t) {
        constexpr int s = decltype(slot_t)::value;
        const int j = min(jl, SB - 1);
#pragma unroll
        for (int q = 0; q < NC; ++q) g[s][q] = buf_ld128(rc, voff(q, k, j), 0u);
    };
    auto lstage = [&](int k, auto slot_t) {   // batch k's rows -> LDS buffer k & 1
        constexpr int s = decltype(slot_t)::value;
        const int bf = k & 1;
        if (jl < SB) {
#pragma unroll
            for (int q = 0; q < NC; ++q) *(u32x4*)(s_c + ((q * 2 + bf) * SB + jl) * RB + cl * 16) = g[s][q];
        }
    };
    auto flush = [&](int k) {    // deltas of batch k (LDS buffer k & 1) -> planes
        const int bf = k & 1;
        if (jl < SB && k * SB + jl < n) {
#pragma unroll
            for (int q = 0; q < NC; ++q) {
                const u32x4 v = *(const u32x4*)(s_d + ((q * 2 + bf) * SB + jl) * RB + cl * 16);
                __builtin_amdgcn_raw_buffer_store_b128(v, raw_rsrc(planes.p[2 * f + (q & 1)]), (int)voff(q, k, jl), 0,
                                                       0);
            }
        }
    };
    // the lane's C' words (8 d of its chain) and P2 pairs of all SB steps of a
    // batch, read into register set k & 1 one batch ahead
    const uint32_t lofs = (uint32_t)(hq * 2 * SB * RB + 8 * hl);
    const uint32_t* p2h = s_p2 + hq * nmax;
    uint2 W[2][SB];
    uint32_t P[2][SB];
    auto wread = [&](int k, auto set_t) {
        constexpr int e = decltype(set_t)::value;
#pragma unroll
        for (int i = 0; i < SB; ++i) W[e][i] = *(const uint2*)(s_c + lofs + ((k & 1) * SB + i) * RB);
        const u32x4* pp = (const u32x4*)(p2h + k * SB);
        const u32x4 a = pp[0], b = pp[1];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            P[e][i] = a[i];
            P[e][4 + i] = b[i];
        }
    };
    const int nb = (n + SB - 1) / SB;
    // loads are issued unconditionally (past the end: the last batch again), so
    // the wait before staging a batch counts the loads issued after it
    static_for<PF>([&](auto q_t) { gload(min((int)decltype(q_t)::value, nb - 1), q_t); });
    lstage(0, std::integral_constant<int, 0>{});
    __syncthreads();
    wread(0, std::integral_constant<int, 0>{});
    // INNER: batch k is not the last one (all SB steps exist, batch k + 1 exists)
    auto batch = [&](int k, auto q_t, auto inner_t) {
        constexpr int q = decltype(q_t)::value;
        constexpr bool INNER = decltype(inner_t)::value;
        if (INNER || k + 1 < nb) {
            lstage(k + 1, std::integral_constant<int, (q + 1) % PF>{});
            wread(k + 1, std::integral_constant<int, (q + 1) & 1>{});
        }
        if (!INNER && k > 0) flush(k - 1);
        gload(min(k + PF, nb - 1), q_t);
        const int bf = k & 1;
        auto run_batch = [&](auto allv_tag) {
            constexpr bool ALLV = decltype(allv_tag)::value;
#pragma unroll
            for (int i = 0; i < SB; ++i) {
                const int t = k * SB + i;
                if (!INNER && t >= n) break;
                if (INNER && i == SB / 2 && k > 0) flush(k - 1);
                const uint2 wc = W[q & 1][i];
                const uint32_t p2p = P[q & 1][i];
                uint32_t c[4];
                {
                    CRaw<uint8_t, 2> x0, x1;
                    x0.w[0] = wc.x;
                    x1.w[0] = wc.y;
                    c[0] = decode_c<uint8_t, 2, ALLV>(x0, 0);
                    c[1] = decode_c<uint8_t, 2, ALLV>(x0, 1);
                    c[2] = decode_c<uint8_t, 2, ALLV>(x1, 0);
                    c[3] = decode_c<uint8_t, 2, ALLV>(x1, 1);
                }
                if (!FULL && !inr) c[0] = c[1] = c[2] = c[3] = 0xFFFFFFFFu;
                // L = C + min(u - minL', P2), u = min(L'(d), L'(d+-1)+P1) (as in the half-wave form)
                const uint32_t left_hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)pv[3], kDppRowShr1, 0xF, 0xF, true);
                const uint32_t right_lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)pv[0], kDppRowShl1, 0xF, 0xF, true);
                // s[j]: (L'(d-1), L'(d)) + P1 of pair j's first d, j = 0..4 (s[j+1] is pair j's d+1 term)
                u16x2 s[5];
                s[0] = add_sat2(as_v2(__builtin_amdgcn_alignbit(pv[0], left_hi, 16)), as_v2(p1l));
#pragma unroll
                for (int j = 1; j < 4; ++j) s[j] = add_sat2(as_v2(__builtin_amdgcn_alignbit(pv[j], pv[j - 1], 16)), as_v2(p1s));
                s[4] = add_sat2(as_v2(__builtin_amdgcn_alignbit(right_lo, pv[3], 16)), as_v2(p1r));
                uint32_t dl[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const u16x2 uu = min2(min2(s[j], s[j + 1]), as_v2(pv[j]));
                    const u16x2 dd = min2(uu - splat2(mp), as_v2(p2p));
                    dl[j] = as_u32(dd);
                    pv[j] = as_u32(add_sat2(as_v2(c[j]), dd));
                }
                const uint32_t lm2 = as_u32(min2(min2(as_v2(pv[0]), as_v2(pv[1])), min2(as_v2(pv[2]), as_v2(pv[3]))));
                uint32_t m = min(lm2 & 0xFFFFu, lm2 >> 16);
                // minimum over the chain's 16-lane row: every lane ends with it
                m = min(m, dpp<kDppRowRor8>(0xFFFFFFFFu, m));
                m = min(m, dpp<kDppRowRor4>(0xFFFFFFFFu, m));
                m = min(m, dpp<kDppRowRor2>(0xFFFFFFFFu, m));
                mp = min(m, dpp<kDppRowRor1>(0xFFFFFFFFu, m));
                *(uint2*)(s_d + lofs + (bf * SB + i) * RB) =
                    make_uint2(__builtin_amdgcn_perm(dl[1], dl[0], 0x06040200u),
                               __builtin_amdgcn_perm(dl[3], dl[2], 0x06040200u));
            }
        };
        if (!INNER || batch_valid(k)) run_batch(std::integral_constant<bool, INNER>{});
        else run_batch(std::false_type{});
    };
    // groups of PF inner batches (batch k in slot k % PF), then the rest
    int k0 = 0;
    for (; k0 + PF < nb; k0 += PF)
        static_for<PF>([&](auto q_t) { batch(k0 + decltype(q_t)::value, q_t, std::true_type{}); });
    static_for<PF>([&](auto q_t) {
        const int k = k0 + decltype(q_t)::value;
        if (k < nb) batch(k, q_t, std::false_type{});
    });
    flush(nb - 1);
}

template <bool FULL, int PF>
__global__ __launch_bounds__(64) void sgm_quad_kernel(apg_hip_params p, int w, int h, int D,
                                                      const uint8_t* __restrict__ gray,
                                                      const uint8_t* __restrict__ cost, PlaneSet planes, int nmax) {
    sgm_quad_body<FULL, PF>(block3(), p, w, h, D, gray, cost, planes, nmax);
}

template <bool FULL, int PF>
struct SgmQuadFn {
    static constexpr int kThreads = 64;
    template <typename... A>
    __device__ static void run(const uint3 b, const A&... a) { sgm_quad_body<FULL, PF>(b, a...); }
};

// batches of C' rows loaded ahead per wave: 3 or 4 measured 242 us against
// 262 us for 2 at 1280x720 D=128 (6 no better); even, so the register set of
// a batch's words (k & 1) is static within the PF-batch trip
constexpr int kHalfPF = 4;

int sgm_pair_mask(const apg_hip_params& p, int D, bool cost_u8, const KTune& tn) {
    const int p2max = p.P2_base > p.P1 + 1 ? p.P2_base : p.P1 + 1;
    if (!(cost_u8 && D % 16 == 0 && D > 0 && D <= 128 && (tn.sgm_dual == 0 || tn.sgm_dual == 3) && p2max <= 127))
        return 0;
    // every family (all or none: the half-wave kernel's PAIRED form assumes
    // it).  At 1280x720 D=128 the SGM and WTA took 227 + 138 us with no family
    // paired, 250 + 107 with the columns only, 280 + 73 with all.
    return p.paths == 8 ? 0xF : 0x3;
}

template <typename TC>
static hipError_t launch_lines_t(hipStream_t s, const apg_hip_params& p, int w, int h, int D, const uint8_t* gray,
                                 const TC* cost, const PlaneSet& planes, const KTune& tn, LaunchSink* sink) {
    const int K = (D + 127) / 128;
    const int nfam = p.paths == 8 ? 4 : 2;
    const int nblk = h + w + (nfam == 4 ? 2 * (w + h - 1) : 0);
    const bool full = D == 128 * K;
    const size_t shm = (size_t)(w > h ? w : h);   // one scanline of gray values
    if (sizeof(TC) == 1 && D % 16 == 0 && K == 1 && nfam == 2 && tn.sgm_dual == 4) {
        // both chains of two scanlines per wave, 16 lanes each; P2 of every step staged
        const int nmax = ((w > h ? w : h) + SB - 1) / SB * SB;
        const size_t shmq = (size_t)16 * SB * 128 + (size_t)4 * nmax * 4;
        const int nq = (h + 1) / 2 + (w + 1) / 2;
        if (sink)
            return full ? sink->add<SgmQuadFn<true, kHalfPF>>(dim3(nq), shmq, p, w, h, D, gray, (const uint8_t*)cost,
                                                             planes, nmax)
                        : sink->add<SgmQuadFn<false, kHalfPF>>(dim3(nq), shmq, p, w, h, D, gray, (const uint8_t*)cost,
                                                              planes, nmax);
        if (shmq > 64 * 1024) {
            hipError_t e = hipFuncSetAttribute(full ? (const void*)sgm_quad_kernel<true, kHalfPF>
                                                    : (const void*)sgm_quad_kernel<false, kHalfPF>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmq);
            if (e != hipSuccess) return e;
        }
        if (full)
            hipLaunchKernelGGL((sgm_quad_kernel<true, kHalfPF>), dim3(nq), dim3(64), shmq, s, p, w, h, D, gray,
                               (const uint8_t*)cost, planes, nmax);
        else
            hipLaunchKernelGGL((sgm_quad_kernel<false, kHalfPF>), dim3(nq), dim3(64), shmq, s, p, w, h, D, gray,
                               (const uint8_t*)cost, planes, nmax);
        return hipGetLastError();
    }
    if (sizeof(TC) == 1 && D % 16 == 0 && K == 1 && (tn.sgm_dual == 0 || tn.sgm_dual == 3)) {
        // both chains of a scanline per wave, 32 lanes each; P2 of every step staged
        const int nmax = ((w > h ? w : h) + SB - 1) / SB * SB;   // whole batches of P2 bytes
        const size_t shmh =
            (size_t)8 * SB * 128 + (size_t)2 * nmax + (planes.paired ? (size_t)(kPairGap + 2) * 128 : 0);
        const void* kf = planes.paired ? (full ? (const void*)sgm_half_kernel<true, kHalfPF, true>
                                               : (const void*)sgm_half_kernel<false, kHalfPF, true>)
                                       : (full ? (const void*)sgm_half_kernel<true, kHalfPF, false>
                                               : (const void*)sgm_half_kernel<false, kHalfPF, false>);
        if (shmh > 64 * 1024) {
            hipError_t e = hipFuncSetAttribute(kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmh);
            if (e != hipSuccess) return e;
        }
#define APG_HALF(FF, PP)                                                                                      \
    do {                                                                                                      \
        if (sink)                                                                                             \
            return sink->add<SgmHalfFn<FF, kHalfPF, PP>>(dim3(nblk), shmh, p, w, h, D, nfam, gray,            \
                                                         (const uint8_t*)cost, planes, nmax);                 \
        hipLaunchKernelGGL((sgm_half_kernel<FF, kHalfPF, PP>), dim3(nblk), dim3(64), shmh, s, p, w, h, D, nfam, \
                           gray, (const uint8_t*)cost, planes, nmax);                                         \
    } while (0)
        if (planes.paired) {
            if (full) APG_HALF(true, true);
            else APG_HALF(false, true);
        } else {
            if (full) APG_HALF(true, false);
            else APG_HALF(false, false);
        }
#undef APG_HALF
        return hipGetLastError();
    }
    if (planes.paired) return hipErrorInvalidValue;   // only the half-wave form pairs its planes
    if (sizeof(TC) == 1 && D % 16 == 0 && K <= 2) {
        const size_t shm4 = (size_t)8 * SB * 128 * K + shm;
        const int nblk4 = 2 * nblk;
#define APG_LINES4(KK, FF)                                                                                        \
    do {                                                                                                          \
        if (tn.sgm_dual == 1) {                                                                                   \
            if (sink)                                                                                             \
                return sink->add<SgmLines4Fn<KK, FF, true>>(dim3(nblk), shm4, p, w, h, D, nfam, gray,             \
                                                            (const uint8_t*)cost, planes);                        \
            hipLaunchKernelGGL((sgm_lines4_kernel<KK, FF, true>), dim3(nblk), dim3(64), shm4, s, p, w, h, D, nfam, \
                               gray, (const uint8_t*)cost, planes);                                               \
        } else {                                                                                                  \
            if (sink)                                                                                             \
                return sink->add<SgmLines4Fn<KK, FF, false>>(dim3(nblk4), shm4, p, w, h, D, nfam, gray,           \
                                                             (const uint8_t*)cost, planes);                       \
            hipLaunchKernelGGL((sgm_lines4_kernel<KK, FF, false>), dim3(nblk4), dim3(64), shm4, s, p, w, h, D,     \
                               nfam, gray, (const uint8_t*)cost, planes);                                         \
        }                                                                                                         \
    } while (0)
        if (K == 1) {
            if (full) APG_LINES4(1, true);
            else APG_LINES4(1, false);
        } else {
            if (full) APG_LINES4(2, true);
            else APG_LINES4(2, false);
        }
#undef APG_LINES4
        return hipGetLastError();
    }
#define APG_LINES(KK, FF)                                                                                      \
    do {                                                                                                       \
        if (sink) {                                                                                            \
            if constexpr (sizeof(TC) == 1)                                                                     \
                return sink->add<SgmLinesFn<TC, KK, FF>>(dim3(nblk), shm, p, w, h, D, nfam, gray, cost, planes); \
            return hipErrorNotSupported;                                                                       \
        }                                                                                                      \
        hipLaunchKernelGGL((sgm_lines_kernel<TC, KK, FF>), dim3(nblk), dim3(64), shm, s, p, w, h, D, nfam, gray,   \
                           cost, planes);                                                                      \
    } while (0)
    if (K == 1) {
        if (full) APG_LINES(1, true);
        else APG_LINES(1, false);
    } else if (K == 2) {
        if (full) APG_LINES(2, true);
        else APG_LINES(2, false);
    } else {
        return hipErrorInvalidValue;
    }
#undef APG_LINES
    return hipGetLastError();
}

hipError_t launch_sgm_lines(hipStream_t s, const apg_hip_params& p, int w, int h, int D, const uint8_t* gray,
                            const void* cost, bool cost_u8, const PlaneSet& planes, const KTune& tn,
                            LaunchSink* sink) {
    if (D % 2 || D > 256 || (size_t)w * h * D * (cost_u8 ? 1 : 2) >= (1ull << 31)) return hipErrorInvalidValue;
    return cost_u8 ? launch_lines_t<uint8_t>(s, p, w, h, D, gray, (const uint8_t*)cost, planes, tn, sink)
                   : launch_lines_t<uint16_t>(s, p, w, h, D, gray, (const uint16_t*)cost, planes, tn, sink);
}

}  // namespace apgk
