// gfx950 SGM kernels of the fused schedule.
//
// Every path value is L_r(p,d) = C'(p,d) + delta_r(p,d) with
// delta_r = min(L'(d), L'(d+-1)+P1, minL'+P2) - minL' in [0, P2] (0 at a scan
// start or after a pixel without valid state), so the path sum is
//   S(p,d) = npath * C'(p,d) + sum_r delta_r(p,d).
// One wavefront owns one scanline family member (a row, a column or a
// diagonal ray) and runs BOTH of its opposite directions at once: chain A
// walks forward, chain B backward.  Until they meet each chain stores its
// delta at its own pixel; after the meeting point each chain reads the other
// direction's delta at its pixel and stores the pair sum Delta (u8, <= 2*P2
// <= 254).  Traffic per state and direction pair: C' read twice, one delta
// store, one delta load, one Delta store.  All families of one launch run
// concurrently (H + W waves for 4 paths, + 2(W + H - 1) rays for 8 paths).
// The u16 recurrence keeps 65535 as path infinity exactly as the plain scan.
#include "device_common.hpp"
#include "kernels.hpp"

namespace apgk {

// C' element of the fused schedule: u8 with 255 = invalid when every
// reachable cost is <= 254, else u16 with 65535 = invalid.
// Loads are split from decoding so that prefetched words stay in flight
// until they are used (decoding right after a load would force the wait).
// raw_cpair: the lane's pair (d = 2m, 2m+1) as stored; lanes past D read the
// pixel's first pair instead (always in bounds) and are masked by decode.
template <typename TC>
__device__ __forceinline__ uint32_t raw_cpair(const TC* __restrict__ v, int m, int D) {
    const int d = 2 * m < D ? 2 * m : 0;   // D even on this path (checked by the launcher)
    if (sizeof(TC) == 1) return *(const uint16_t*)((const uint8_t*)v + d);
    return *(const uint32_t*)((const uint16_t*)v + d);
}

// raw pair -> u16 halves with the invalid value 65535 (u8: 255 -> 65535)
template <typename TC>
__device__ __forceinline__ uint32_t decode_cpair(uint32_t raw, bool in_range) {
    if (!in_range) return 0xFFFFFFFFu;
    if (sizeof(TC) == 2) return raw;
    const uint32_t x = (raw & 0xFFu) | ((raw & 0xFF00u) << 8);
    const uint32_t inv = ((x + 0x00010001u) >> 8) & 0x00010001u;   // 1 in halves equal to 255
    return x | (inv * 0xFF00u);
}

// delta pair (u16 halves, each < 256) -> 2 bytes
__device__ __forceinline__ uint32_t pack_bytes(uint32_t v) { return (v & 0xFFu) | ((v >> 8) & 0xFF00u); }
__device__ __forceinline__ uint32_t unpack_bytes(uint32_t b) { return (b & 0xFFu) | ((b & 0xFF00u) << 8); }

// One recurrence step for a chain (reference src/sgm_optimizer.cpp:44-80).
// Returns the delta pairs in dl[] and updates prev[]; returns the new minL.
template <int K>
__device__ __forceinline__ uint32_t sgm_step(const apg_hip_params& p, uint32_t (&prev)[K], const uint32_t (&c)[K],
                                             uint32_t minprev, int g, int gprev, bool start, uint32_t (&dl)[K]) {
    uint32_t L[K];
    if (start || minprev == 0xFFFFu) {
#pragma unroll
        for (int j = 0; j < K; ++j) {
            L[j] = c[j];
            dl[j] = 0;
        }
    } else {
        const uint32_t P2 = (uint32_t)adaptive_p2(p, g - gprev);
        const u16x2 mp = splat2(minprev);
        const u16x2 mp2 = add_sat2(mp, splat2(P2));
        const u16x2 p1 = splat2((uint32_t)p.P1);
        const uint32_t left_hi = wave_shr1(prev[K - 1], 0xFFFFFFFFu);
        const uint32_t right_lo = wave_shl1(prev[0], 0xFFFFFFFFu);
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const uint32_t pj = prev[j];
            const uint32_t before = j == 0 ? left_hi : prev[j - 1];
            const uint32_t after = j == K - 1 ? right_lo : prev[j + 1];
            const u16x2 dm1 = as_v2(__builtin_amdgcn_alignbit(pj, before, 16));
            const u16x2 dp1 = as_v2(__builtin_amdgcn_alignbit(after, pj, 16));
            u16x2 t2 = min2(as_v2(pj), min2(add_sat2(dm1, p1), add_sat2(dp1, p1)));
            t2 = min2(t2, mp2);
            const u16x2 dd = t2 - mp;
            dl[j] = as_u32(dd);
            L[j] = as_u32(add_sat2(as_v2(c[j]), dd));
        }
    }
    uint32_t lm = 0xFFFFu;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        prev[j] = L[j];
        lm = min(lm, min(L[j] & 0xFFFFu, L[j] >> 16));
    }
    return wave_min_u32(lm);
}

constexpr int SF_PF = 8;   // steps per prefetch batch (per chain)

// Scanline of family f, member s: start pixel, step and length of chain A.
__device__ __forceinline__ void family_line(int f, int s, int w, int h, long& p0, long& step, int& n) {
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
    p0 = (long)sy * w + sx;
    step = (long)dy * w + dx;
}

template <typename TC, int K>
__global__ __launch_bounds__(64) void sgm_pairs_kernel(apg_hip_params p, int w, int h, int D, int nfam,
                                                       const uint8_t* __restrict__ gray, const TC* __restrict__ cost,
                                                       uint8_t* dsum0, uint8_t* dsum1, uint8_t* dsum2, uint8_t* dsum3) {
    const int lane = threadIdx.x;
    // block -> (family, member): rows, then columns, then the two diagonal ray sets
    int b = blockIdx.x, f = 0;
    const int cnt[4] = {h, w, w + h - 1, w + h - 1};
    while (f < nfam - 1 && b >= cnt[f]) { b -= cnt[f]; ++f; }
    uint8_t* dsum = f == 0 ? dsum0 : f == 1 ? dsum1 : f == 2 ? dsum2 : dsum3;
    long p0, step;
    int n;
    family_line(f, b, w, h, p0, step, n);
    uint32_t prevA[K], prevB[K];
#pragma unroll
    for (int j = 0; j < K; ++j) prevA[j] = prevB[j] = 0xFFFFFFFFu;
    uint32_t minA = 0xFFFFu, minB = 0xFFFFu;
    int gA = 0, gB = 0;
    bool inr[K];   // lane's pair j lies inside D
#pragma unroll
    for (int j = 0; j < K; ++j) inr[j] = 2 * (lane * K + j) + 1 < D;
    // lane byte offset of pair j in a delta plane row: 2*(lane*K + j); lanes past D
    // write/read pair 0 of the pixel with an unchanged value (see below)
    auto doff = [&](int j) { return inr[j] ? 2 * (lane * K + j) : 0; };
    // Raw words of one batch of SF_PF steps for both chains: C' pairs and the
    // other direction's delta.  The delta word is fetched from the delta plane
    // when its producing step (n-1-t) finished before the fetch is issued
    // (before step `done_before`), else from the C' pair just fetched (a cache
    // hit, ignored).  The next batch is in flight while the current one runs.
    uint32_t rA[SF_PF][K], rB[SF_PF][K], qA[SF_PF][K], qB[SF_PF][K];
    auto fetch = [&](int t0, int done_before) {
#pragma unroll
        for (int i = 0; i < SF_PF; ++i) {
            const int t = min(t0 + i, n - 1);
            const long pa = p0 + step * t, pb = p0 + step * (n - 1 - t);
            const TC* va = cost + (size_t)pa * D;
            const TC* vb = cost + (size_t)pb * D;
            const bool od = 2 * t > n - 1 && n - 1 - t < done_before;
#pragma unroll
            for (int j = 0; j < K; ++j) {
                rA[i][j] = raw_cpair<TC>(va, lane * K + j, D);
                rB[i][j] = raw_cpair<TC>(vb, lane * K + j, D);
                const uint8_t* ea = od ? dsum + (size_t)pa * D + doff(j) : (const uint8_t*)va;
                const uint8_t* eb = od ? dsum + (size_t)pb * D + doff(j) : (const uint8_t*)vb;
                qA[i][j] = *(const uint16_t*)ea;
                qB[i][j] = *(const uint16_t*)eb;
            }
        }
    };
    fetch(0, -SF_PF);
    uint32_t gchA = 0, gchB = 0;
    for (int t0 = 0; t0 < n; t0 += SF_PF) {
        if ((t0 & 63) == 0) {
            // gray of the next 64 pixels of both chains, one per lane
            const int tl = min(t0 + lane, n - 1);
            gchA = gray[p0 + step * tl];
            gchB = gray[p0 + step * (n - 1 - tl)];
        }
        uint32_t cA[SF_PF][K], cB[SF_PF][K], xA[SF_PF][K], xB[SF_PF][K];
#pragma unroll
        for (int i = 0; i < SF_PF; ++i)
#pragma unroll
            for (int j = 0; j < K; ++j) {
                cA[i][j] = rA[i][j];
                cB[i][j] = rB[i][j];
                xA[i][j] = qA[i][j];
                xB[i][j] = qB[i][j];
            }
        // deltas of this batch were prefetched iff their producer preceded t0 - SF_PF
        const int pf_done = t0 - SF_PF;
        // a batch that needs some delta not prefetched re-reads all its deltas at use time
        const int tl_last = min(t0 + SF_PF - 1, n - 1);
        bool slow = false;
        for (int t = t0; t <= tl_last; ++t) slow |= (2 * t > n - 1) && !(n - 1 - t < pf_done);
        if (t0 + SF_PF < n) fetch(t0 + SF_PF, t0);
#pragma unroll
        for (int i = 0; i < SF_PF; ++i) {
            const int t = t0 + i;
            if (t >= n) break;
            const int ga = __builtin_amdgcn_readlane((int)gchA, t & 63);
            const int gb = __builtin_amdgcn_readlane((int)gchB, t & 63);
            uint32_t ca[K], cb[K], dA[K], dB[K];
#pragma unroll
            for (int j = 0; j < K; ++j) {
                ca[j] = decode_cpair<TC>(cA[i][j], inr[j]);
                cb[j] = decode_cpair<TC>(cB[i][j], inr[j]);
            }
            minA = sgm_step<K>(p, prevA, ca, minA, ga, gA, t == 0, dA);
            minB = sgm_step<K>(p, prevB, cb, minB, gb, gB, t == 0, dB);
            gA = ga;
            gB = gb;
            const long pa = p0 + step * t, pb = p0 + step * (n - 1 - t);
            uint8_t* da = dsum + (size_t)pa * D;
            uint8_t* db = dsum + (size_t)pb * D;
            // Before the chains meet each stores its own delta; at the meeting
            // pixel (odd n) both store the pair sum; afterwards each adds the
            // delta the other direction stored here at step n-1-t.
            const bool second = 2 * t > n - 1, middle = 2 * t == n - 1;
#pragma unroll
            for (int j = 0; j < K; ++j) {
                uint32_t oA = unpack_bytes(xA[i][j]), oB = unpack_bytes(xB[i][j]);
                if (slow && second) {
                    oA = unpack_bytes(*(const uint16_t*)(da + doff(j)));
                    oB = unpack_bytes(*(const uint16_t*)(db + doff(j)));
                }
                const uint32_t addA = second ? oA : (middle ? dB[j] : 0u);
                const uint32_t addB = second ? oB : (middle ? dA[j] : 0u);
                if (inr[j]) {
                    *(uint16_t*)(da + doff(j)) = (uint16_t)pack_bytes(dA[j] + addA);
                    *(uint16_t*)(db + doff(j)) = (uint16_t)pack_bytes(dB[j] + addB);
                }
            }
        }
    }
}

template <typename TC>
static hipError_t launch_pairs_t(hipStream_t s, const apg_hip_params& p, int w, int h, int D, const uint8_t* gray,
                                 const TC* cost, uint8_t* const dsum[4]) {
    const int K = (D + 127) / 128;
    const int nfam = p.paths == 8 ? 4 : 2;
    const int nblk = h + w + (nfam == 4 ? 2 * (w + h - 1) : 0);
#define APG_PAIRS(KK)                                                                                          \
    case KK:                                                                                                   \
        hipLaunchKernelGGL((sgm_pairs_kernel<TC, KK>), dim3(nblk), dim3(64), 0, s, p, w, h, D, nfam, gray, cost, \
                           dsum[0], dsum[1], dsum[2], dsum[3]);                                                \
        break;
    switch (K) {
        APG_PAIRS(1)
        APG_PAIRS(2)
        APG_PAIRS(3)
        APG_PAIRS(4)
        default: return hipErrorInvalidValue;
    }
#undef APG_PAIRS
    return hipGetLastError();
}

hipError_t launch_sgm_pairs(hipStream_t s, const apg_hip_params& p, int w, int h, int D, const uint8_t* gray,
                            const void* cost, bool cost_u8, uint8_t* const dsum[4]) {
    if (D % 2) return hipErrorInvalidValue;
    return cost_u8 ? launch_pairs_t<uint8_t>(s, p, w, h, D, gray, (const uint8_t*)cost, dsum)
                   : launch_pairs_t<uint16_t>(s, p, w, h, D, gray, (const uint16_t*)cost, dsum);
}

}  // namespace apgk
