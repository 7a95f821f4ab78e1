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
#include "device_common.hpp"
#include "kernels.hpp"

namespace apgk {

constexpr int SL_PF = 8;   // steps per prefetch batch (two batches in flight)

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

// pair j of the lane -> u16 halves, invalid (u8 255 / u16 65535) -> 65535
template <typename TC, int K>
__device__ __forceinline__ uint32_t decode_c(const CRaw<TC, K>& v, int j) {
    if constexpr (sizeof(TC) == 2) {
        return v.w[j];
    } else {
        const uint32_t x = __builtin_amdgcn_perm(0u, v.w[0], j == 0 ? 0x0c010c00u : 0x0c030c02u);
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
__global__ __launch_bounds__(64) void sgm_lines_kernel(apg_hip_params p, int w, int h, int D, int nfam,
                                                       const uint8_t* __restrict__ gray, const TC* __restrict__ cost,
                                                       PlaneSet planes) {
    const int lane = threadIdx.x;
    // block -> (family, member): rows, then columns, then the two diagonal ray sets
    int b = blockIdx.x, f = 0;
    {
        const int cnt[4] = {h, w, w + h - 1, w + h - 1};
        while (f < nfam - 1 && b >= cnt[f]) { b -= cnt[f]; ++f; }
    }
    int p0, step, n;
    family_line(f, b, w, h, p0, step, n);
    // the longest scanlines bound the kernel's duration: their waves win
    // instruction arbitration against the shorter ones sharing a SIMD
    if (n >= w && n >= h) __builtin_amdgcn_s_setprio(2);
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
    // adaptive P2 of the next 64 steps of each chain, one step per lane; the
    // gray values of the 64 steps after those are loaded one chunk ahead
    uint32_t p2A = 0, p2B = 0, lastA = 0, lastB = 0, ngA, ngB;
    auto load_gray = [&](int t0) {
        const int tl = min(t0 + lane, n - 1);
        ngA = gray[pA0 + tl * step];
        ngB = gray[pB0 - tl * step];
    };
    load_gray(0);
    auto refresh_p2 = [&](int t0) {
        const uint32_t ga = ngA, gb = ngB;
        load_gray(t0 + 64);
        const uint32_t gpa = (uint32_t)__builtin_amdgcn_update_dpp((int)lastA, (int)ga, kDppWaveShr1, 0xF, 0xF, false);
        const uint32_t gpb = (uint32_t)__builtin_amdgcn_update_dpp((int)lastB, (int)gb, kDppWaveShr1, 0xF, 0xF, false);
        p2A = (uint32_t)adaptive_p2(p, (int)ga - (int)gpa);
        p2B = (uint32_t)adaptive_p2(p, (int)gb - (int)gpb);
        lastA = (uint32_t)__builtin_amdgcn_readlane((int)ga, 63);
        lastB = (uint32_t)__builtin_amdgcn_readlane((int)gb, 63);
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

template <typename TC>
static hipError_t launch_lines_t(hipStream_t s, const apg_hip_params& p, int w, int h, int D, const uint8_t* gray,
                                 const TC* cost, const PlaneSet& planes) {
    const int K = (D + 127) / 128;
    const int nfam = p.paths == 8 ? 4 : 2;
    const int nblk = h + w + (nfam == 4 ? 2 * (w + h - 1) : 0);
    const bool full = D == 128 * K;
#define APG_LINES(KK, FF)                                                                                      \
    hipLaunchKernelGGL((sgm_lines_kernel<TC, KK, FF>), dim3(nblk), dim3(64), 0, s, p, w, h, D, nfam, gray, cost, \
                       planes)
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
                            const void* cost, bool cost_u8, const PlaneSet& planes) {
    if (D % 2 || D > 256 || (size_t)w * h * D * (cost_u8 ? 1 : 2) >= (1ull << 31)) return hipErrorInvalidValue;
    return cost_u8 ? launch_lines_t<uint8_t>(s, p, w, h, D, gray, (const uint8_t*)cost, planes)
                   : launch_lines_t<uint16_t>(s, p, w, h, D, gray, (const uint16_t*)cost, planes);
}

}  // namespace apgk
