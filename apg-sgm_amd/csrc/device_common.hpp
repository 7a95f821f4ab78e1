// Shared device helpers for the gfx950 kernels: packed u16 arithmetic
// (v_pk_*_u16), wave64 DPP reductions and the stage-parameter block that is
// passed by value to every kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/apg_sgm_hip.h"

namespace apgk {

constexpr uint16_t kInv16 = 0xFFFFu;          // kInvalidCost (reference include/apg_sgm/types.hpp:44)
constexpr uint32_t kInv32 = 0xFFFFFFFFu;      // kInvalidCost32 (reference include/apg_sgm/types.hpp:45)
constexpr int kWave = 64;

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t as_u32(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ u16x2 as_v2(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
__device__ __forceinline__ u16x2 splat2(uint32_t s) { return (u16x2)((unsigned short)s); }
__device__ __forceinline__ u16x2 add_sat2(u16x2 a, u16x2 b) { return __builtin_elementwise_add_sat(a, b); }
__device__ __forceinline__ u16x2 min2(u16x2 a, u16x2 b) { return __builtin_elementwise_min(a, b); }

// DPP control words (CDNA3/CDNA4 ISA, DPP_CTRL enumeration)
constexpr int kDppRowShl1 = 0x101, kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppWaveShl1 = 0x130, kDppWaveShr1 = 0x138;
constexpr int kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;
constexpr int kDppRowRor1 = 0x121, kDppRowRor2 = 0x122, kDppRowRor4 = 0x124, kDppRowRor8 = 0x128;

template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ uint32_t dpp(uint32_t old, uint32_t src) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)src, CTRL, ROW_MASK, 0xF, false);
}

// lane l receives lane l-1's value; lane 0 receives `fill`.
__device__ __forceinline__ uint32_t wave_shr1(uint32_t v, uint32_t fill) { return dpp<kDppWaveShr1>(fill, v); }
// lane l receives lane l+1's value; lane 63 receives `fill`.
__device__ __forceinline__ uint32_t wave_shl1(uint32_t v, uint32_t fill) { return dpp<kDppWaveShl1>(fill, v); }

// Minimum over the 64 lanes, returned wave-uniform. Requires all lanes active.
// Lanes without a DPP source read the min identity, so each level compiles
// to one v_min_u32_dpp (row_shr 1/2/4/8 within rows of 16, then the row
// broadcasts 15 and 31 leave the total in lane 63).
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    v = min(v, dpp<kDppRowShr1>(0xFFFFFFFFu, v));
    v = min(v, dpp<kDppRowShr2>(0xFFFFFFFFu, v));
    v = min(v, dpp<kDppRowShr4>(0xFFFFFFFFu, v));
    v = min(v, dpp<kDppRowShr8>(0xFFFFFFFFu, v));
    v = min(v, dpp<kDppRowBcast15, 0xA>(0xFFFFFFFFu, v));
    v = min(v, dpp<kDppRowBcast31, 0xC>(0xFFFFFFFFu, v));
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Minima over the 64 lanes of two independent values, levels interleaved so
// each DPP read of one value follows the other value's write (no hazard NOPs).
__device__ __forceinline__ void wave_min2_u32(uint32_t& a, uint32_t& b) {
    a = min(a, dpp<kDppRowShr1>(0xFFFFFFFFu, a));
    b = min(b, dpp<kDppRowShr1>(0xFFFFFFFFu, b));
    a = min(a, dpp<kDppRowShr2>(0xFFFFFFFFu, a));
    b = min(b, dpp<kDppRowShr2>(0xFFFFFFFFu, b));
    a = min(a, dpp<kDppRowShr4>(0xFFFFFFFFu, a));
    b = min(b, dpp<kDppRowShr4>(0xFFFFFFFFu, b));
    a = min(a, dpp<kDppRowShr8>(0xFFFFFFFFu, a));
    b = min(b, dpp<kDppRowShr8>(0xFFFFFFFFu, b));
    a = min(a, dpp<kDppRowBcast15, 0xA>(0xFFFFFFFFu, a));
    b = min(b, dpp<kDppRowBcast15, 0xA>(0xFFFFFFFFu, b));
    a = min(a, dpp<kDppRowBcast31, 0xC>(0xFFFFFFFFu, a));
    b = min(b, dpp<kDppRowBcast31, 0xC>(0xFFFFFFFFu, b));
    a = (uint32_t)__builtin_amdgcn_readlane((int)a, 63);
    b = (uint32_t)__builtin_amdgcn_readlane((int)b, 63);
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Raw (stride 0) buffer resource over a device allocation, so a load or store
// addresses base + lane offset (VGPR) + wave offset (SGPR): per-step address
// updates stay scalar and 32-bit.  Dword 3 is the gfx9 default format word.
// Callers keep every offset inside the allocation (< 2^31 bytes).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void* p) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, 0x7FFFFFFF, 0x00020000);
}
__device__ __forceinline__ uint32_t buf_ld16(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, (int)soff, 0);
}
__device__ __forceinline__ void buf_st16(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v, r, (int)voff, (int)soff, 0);
}
__device__ __forceinline__ uint32_t buf_ld32(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0);
}
__device__ __forceinline__ void buf_st32(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b32(v, r, (int)voff, (int)soff, 0);
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ u32x4 buf_ld128(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0);
}
// the same at device scope (sc1): served by L2, never by a CU's L1, for data
// that this kernel wrote earlier
__device__ __forceinline__ u32x4 buf_ld128_dev(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 16);
}

// u8 pair (bytes 0, 1 of `raw`) -> u16 halves
__device__ __forceinline__ uint32_t bytes_to_halves(uint32_t raw) { return __builtin_amdgcn_perm(0u, raw, 0x0c010c00u); }

// detail::adaptive_p2 (reference src/sgm_common.hpp:25-39)
__device__ __forceinline__ int adaptive_p2(const apg_hip_params& p, int dI) {
    dI = dI < 0 ? -dI : dI;
    int P2 = p.P2_base;
    if (!p.adaptive_penalties) return max(P2, p.P1 + 1);
    if (p.piecewise_p2) {
        if (dI >= p.grad_t2) P2 = p.P2_base / 4;
        else if (dI >= p.grad_t1) P2 = p.P2_base / 2;
        else P2 = p.P2_base;
    } else {
        P2 = (int)((float)p.P2_base / (1.f + p.p2_alpha * (float)dI) + 0.5f);
    }
    return max(P2, p.P1 + 1);
}

// SubpixelEstimator::compute_legacy_delta (reference src/subpixel_estimator.hpp:115-124).
// Built with -ffp-contract=off and IEEE division, so it rounds as the host does.
__device__ __forceinline__ float legacy_delta(float cm, float c0, float cp) {
    const float denom = cm - 2.0f * c0 + cp;
    if (fabsf(denom) > 1e-6f) {
        const float delta = 0.5f * (cm - cp) / denom;
        if (delta < -0.5f) return -0.5f;
        if (delta > 0.5f) return 0.5f;
        return delta;
    }
    return 0.0f;
}

// Per-pixel cost descriptor: census (hi word only for the 64-bit census),
// gray | gx << 8 | gy << 16.
struct PixDesc {
    uint32_t clo, chi;
    uint32_t ggg;
};

// Quantised matching cost of one state (reference src/cost_computer.cpp:176-199):
// c = census_term + ad_term + grad_term in that order, q = (int)(c*scale + 0.5f),
// clamped to [0, cost_max].  lut_* hold the host-built float terms.
// The AD / gradient LUTs hold +0.0f when the term is disabled, and c >= 0,
// so adding their entry unconditionally leaves c bit-identical to skipping
// the term; the three table reads are independent and issue together.
template <bool SYM>
__device__ __forceinline__ uint16_t state_cost(const PixDesc& L, const PixDesc& R, const float* lut_c,
                                               const float* lut_a, const float* lut_g, float scale, int cost_max,
                                               int use_ad, int use_grad) {
    (void)use_ad;
    (void)use_grad;
    const int pop = SYM ? __popc(L.clo ^ R.clo) : (__popc(L.clo ^ R.clo) + __popc(L.chi ^ R.chi));
    const int ad = abs((int)(L.ggg & 0xFF) - (int)(R.ggg & 0xFF));
    const int g = abs((int)((L.ggg >> 8) & 0xFF) - (int)((R.ggg >> 8) & 0xFF)) +
                  abs((int)((L.ggg >> 16) & 0xFF) - (int)((R.ggg >> 16) & 0xFF));
    const float tc = lut_c[pop], ta = lut_a[ad], tg = lut_g[g];
    const float c = (tc + ta) + tg;
    const int q = (int)(c * scale + 0.5f);
    return (uint16_t)clampi(q, 0, cost_max);
}

}  // namespace apgk
