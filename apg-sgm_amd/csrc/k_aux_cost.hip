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
    out->mag31[0] = 0u;
    for (uint32_t n = 1; n < 2 * kMaxArm + 2; ++n) out->mag31[n] = (0x80000000u + n - 1u) / n;
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
template <>
struct GroupByRef<AuxPair> : std::true_type {};

__device__ __forceinline__ void aux_pixel(const AuxImage& im, int w, int h, int stride, int ch, int census_type,
                                          int gmin, int gmax, uint8_t (&tile)[AUX_TY + 6][AUX_TX + 8]);

// ------------------------------------------------------- aux, pair form ----
// Both views in one launch (blockIdx.z = view), four horizontally adjacent
// pixels per thread, one u32 of the gray image per tile word.  A block covers
// 64 columns x 16 rows; its LDS tile holds the clamp-to-edge gray values of
// 16 + 2*RH rows (RH = 3 for the census, max(3, Lmax) where the left view also
// builds the Cross arms) and 104 columns (20 on each side: census 4, arms
// Lmax <= 17 rounded up to whole words).  Tile rows are 48 words apart, so the
// two or four rows that one LDS access of a wave touches fall in disjoint banks.
//
// Byte-parallel comparison: for bytes a, b of two words,
//   t = (b | 0x80) - (a & 0x7f) has top bit [b & 0x7f >= a & 0x7f] (no borrow
//   leaves a byte), and "b >= a" = top bit of ((a ^ b) & b) | (~(a ^ b) & t).
// Each census comparison yields that bit for the thread's four pixels at bit 7
// of each byte; acc = (acc >> 1) with the top bits replaced collects eight
// comparisons per byte (first at bit 0), and a 4x4 byte transpose makes the
// per-pixel census words.
constexpr int AX_PX = 64, AX_ROWS = 16, AX_HALO = 20, AX_TW = (AX_PX + 2 * AX_HALO) / 4, AX_S = 48;
constexpr int AX_LMAX = 17;   // longest arm the pair form builds

__device__ __forceinline__ uint32_t bsel(uint32_t m, uint32_t a, uint32_t b) { return (m & a) | (~m & b); }
// top bit of each byte: b >= a (unsigned bytes)
__device__ __forceinline__ uint32_t byte_ge(uint32_t b, uint32_t a) {
    const uint32_t t = (b | 0x80808080u) - (a & 0x7F7F7F7Fu);
    return bsel(a ^ b, b, t);
}
// per-byte |a - b|
__device__ __forceinline__ uint32_t byte_absdiff(uint32_t a, uint32_t b) {
    const uint32_t one = (~byte_ge(b, a) >> 7) & 0x01010101u;   // 1 where a > b
    const uint32_t m = (one << 8) - one;                         // 0xff where a > b
    return bsel(m, a, b) - bsel(m, b, a);
}
// bytes j of g0..g3 -> word j (byte k from g_k)
__device__ __forceinline__ void byte_transpose(const uint32_t (&g)[4], uint32_t (&c)[4]) {
    const uint32_t x0 = __builtin_amdgcn_perm(g[1], g[0], 0x05010400u), x1 = __builtin_amdgcn_perm(g[1], g[0], 0x07030602u);
    const uint32_t y0 = __builtin_amdgcn_perm(g[3], g[2], 0x05010400u), y1 = __builtin_amdgcn_perm(g[3], g[2], 0x07030602u);
    c[0] = __builtin_amdgcn_perm(y0, x0, 0x05040100u);
    c[1] = __builtin_amdgcn_perm(y0, x0, 0x07060302u);
    c[2] = __builtin_amdgcn_perm(y1, x1, 0x05040100u);
    c[3] = __builtin_amdgcn_perm(y1, x1, 0x07060302u);
}

constexpr int kAuxTile = (AX_ROWS + 2 * AX_LMAX) * AX_S;   // words

template <int CT>
__device__ __forceinline__ void aux4_pair_body(const uint3 bid, const AuxPair& ap, int w, int h, int stride, int ch,
                                               int gmin, int gmax, uint32_t* __restrict__ arms, int Lmax, int tau) {
    __shared__ uint32_t tile[kAuxTile];
    const int view = bid.z;
    // the pointers read as global ones (group launches read ap from memory)
    const AuxImage& imr = ap.img[view];
    const AuxImage im{global_ptr(imr.src), global_ptr(imr.gray), global_ptr(imr.gx), global_ptr(imr.gy),
                      global_ptr(imr.c32), global_ptr(imr.c64), global_ptr(imr.keys), global_ptr(imr.dmin),
                      global_ptr(imr.dmax)};
    const bool do_arms = view == 0 && arms != nullptr;
    const int rh = do_arms ? max(3, Lmax) : 3;
    const int th = AX_ROWS + 2 * rh;
    const int x0 = bid.x * AX_PX, y0 = bid.y * AX_ROWS;
    const int tid = threadIdx.x;
    // every word's load is issued before any is stored (a word wholly inside
    // the image of a 4-byte aligned gray row is one load; the edge words and
    // colour images are gathered byte by byte)
    constexpr int NT = ((AX_ROWS + 2 * AX_LMAX) * AX_TW + 255) / 256;
    const bool aligned = ch == 1 && (((uintptr_t)im.src) & 3u) == 0 && (stride & 3) == 0;
    uint32_t tv[NT];
#pragma unroll
    for (int k = 0; k < NT; ++k) {
        const int i = min(tid + 256 * k, th * AX_TW - 1);
        const int r = i / AX_TW, c = i - r * AX_TW;
        const int sy = clampi(y0 - rh + r, 0, h - 1);
        const int xa = x0 - AX_HALO + 4 * c;
        const uint8_t* rowp = im.src + (size_t)sy * stride;
        if (aligned && xa >= 0 && xa + 3 < w) {
            tv[k] = *(const uint32_t*)(rowp + xa);
        } else {
            uint32_t v = 0u;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const uint8_t* px = rowp + (size_t)clampi(xa + b, 0, w - 1) * ch;
                const uint32_t g = ch == 1 ? px[0] : ((77u * px[0] + 150u * px[1] + 29u * px[2]) >> 8);
                v |= g << (8 * b);
            }
            tv[k] = v;
        }
    }
#pragma unroll
    for (int k = 0; k < NT; ++k) {
        const int i = tid + 256 * k;
        if (i < th * AX_TW) {
            const int r = i / AX_TW, c = i - r * AX_TW;
            tile[r * AX_S + c] = tv[k];
        }
    }
    __syncthreads();
    const int tx = tid & 15, ty = tid >> 4;
    const int x = x0 + 4 * tx, y = y0 + ty;
    if (x >= w || y >= h) return;
    const uint32_t* tc = tile + (rh + ty) * AX_S + AX_HALO / 4 + tx;   // the thread's word of its row
    // rows -3..3, words -1..1 around the thread's word
    uint32_t W[7][3];
#pragma unroll
    for (int r = 0; r < 7; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) W[r][k] = tc[(r - 3) * AX_S + k - 1];
    // the four bytes at columns x + dx .. x + dx + 3 of row y + dy
    auto win = [&](int dy, int dx) -> uint32_t {
        const uint32_t* r = W[dy + 3];
        if (dx == -4) return r[0];
        if (dx == 0) return r[1];
        if (dx == 4) return r[2];
        return dx < 0 ? __builtin_amdgcn_alignbyte(r[1], r[0], (uint32_t)(4 + dx))
                      : __builtin_amdgcn_alignbyte(r[2], r[1], (uint32_t)dx);
    };
    const uint32_t center = win(0, 0);
    const uint32_t gxw = byte_absdiff(win(0, 1), win(0, -1));
    const uint32_t gyw = byte_absdiff(win(1, 0), win(-1, 0));
    const bool wide = (w & 3) == 0 && x + 3 < w;   // four pixels, word-aligned rows
    const int nq = min(4, w - x);
    const size_t pi = (size_t)y * w + x;
    if (wide) {
        *(uint32_t*)(im.gray + pi) = center;
        *(uint32_t*)(im.gx + pi) = gxw;
        *(uint32_t*)(im.gy + pi) = gyw;
    } else {
        for (int j = 0; j < nq; ++j) {
            im.gray[pi + j] = (uint8_t)(center >> (8 * j));
            im.gx[pi + j] = (uint8_t)(gxw >> (8 * j));
            im.gy[pi + j] = (uint8_t)(gyw >> (8 * j));
        }
    }
    if (im.keys) {
        if (wide) *(uint4*)(im.keys + pi) = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
        else for (int j = 0; j < nq; ++j) im.keys[pi + j] = 0xFFFFFFFFu;
    }
    if (im.dmin) {   // as range_kernel
        // [gmin, max(min(gmax, x + 1), gmin)) per pixel, computed where stored:
        // an array indexed by the edge loop below went to LDS (4 KB per block)
        auto hi = [&](int j) { return (uint32_t)(uint16_t)max(min(gmax, x + j + 1), gmin); };
        if (wide) {
            const uint32_t lo2 = (uint32_t)(uint16_t)gmin * 0x10001u;
            *(uint2*)(im.dmin + pi) = make_uint2(lo2, lo2);
            *(uint2*)(im.dmax + pi) = make_uint2(hi(0) | hi(1) << 16, hi(2) | hi(3) << 16);
        } else {
            for (int j = 0; j < nq; ++j) {
                im.dmin[pi + j] = (int16_t)gmin;
                im.dmax[pi + j] = (int16_t)hi(j);
            }
        }
    }
    // census: "b >= a" bits are collected and inverted at the end (a > b)
    if constexpr (CT == 1) {
        // symmetric pairs (dy, dx) > (-dy, -dx) over the first half of the 9x7 window
        uint32_t g[4], acc = 0u;
        int b = 0;
#pragma unroll
        for (int dy = -3; dy <= 0; ++dy)
#pragma unroll
            for (int dx = -4; dx <= 4; ++dx) {
                if (dy == 0 && dx >= 0) continue;
                acc = bsel(0x80808080u, byte_ge(win(-dy, -dx), win(dy, dx)), acc >> 1);
                ++b;
                if ((b & 7) == 0) g[(b >> 3) - 1] = acc;
            }
        g[3] = (acc >> 1) | 0x80808080u;   // 31 comparisons: bit 31 of every census word is 0
        uint32_t c[4];
        byte_transpose(g, c);
        const uint4 cw = make_uint4(~c[0], ~c[1], ~c[2], ~c[3]);
        if (wide) *(uint4*)(im.c32 + pi) = cw;
        else {
            const uint32_t cv[4] = {cw.x, cw.y, cw.z, cw.w};
            for (int j = 0; j < nq; ++j) im.c32[pi + j] = cv[j];
        }
    } else {
        // every neighbour of the 9x7 window against the centre
        uint32_t g[8], acc = 0u;
        int b = 0;
#pragma unroll
        for (int dy = -3; dy <= 3; ++dy)
#pragma unroll
            for (int dx = -4; dx <= 4; ++dx) {
                if (dy == 0 && dx == 0) continue;
                acc = bsel(0x80808080u, byte_ge(center, win(dy, dx)), acc >> 1);
                ++b;
                if ((b & 7) == 0) g[(b >> 3) - 1] = acc;
            }
        g[7] = (acc >> 2) | 0xC0C0C0C0u;   // 62 comparisons: bits 62, 63 are 0
        uint32_t lo[4], hi[4];
        byte_transpose({g[0], g[1], g[2], g[3]}, lo);
        byte_transpose({g[4], g[5], g[6], g[7]}, hi);
        uint64_t cv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) cv[j] = (uint64_t)~lo[j] | ((uint64_t)~hi[j] << 32);
        if (wide) {
            *(ulonglong2*)(im.c64 + pi) = make_ulonglong2(cv[0], cv[1]);
            *(ulonglong2*)(im.c64 + pi + 2) = make_ulonglong2(cv[2], cv[3]);
        } else {
            for (int j = 0; j < nq; ++j) im.c64[pi + j] = cv[j];
        }
    }
    if (!do_arms) return;
    // CostAggregator::build_cross_arms (reference src/cost_aggregator.cpp:17-63):
    // an arm grows while the next neighbour differs from the centre by at most
    // tau, up to Lmax.  Neighbours past the border repeat the edge pixel in the
    // tile; capping each arm at the distance to the border afterwards gives the
    // reference's arm, which stops there.  |a - c| <= tau <=> a - (c - tau) <= 2 tau
    // in unsigned arithmetic.
    const uint32_t t2 = 2u * (uint32_t)tau;
    uint32_t ct[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) ct[j] = ((center >> (8 * j)) & 0xFFu) - (uint32_t)tau;
    uint32_t H[11];   // words -5..5 of the thread's row
#pragma unroll
    for (int k = 0; k < 11; ++k) H[k] = tc[k - 5];
    int al[4], ar[4], au[4], ad[4];
    bool gl[4], gr[4], gu[4], gd[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        al[j] = ar[j] = au[j] = ad[j] = 0;
        gl[j] = gr[j] = gu[j] = gd[j] = true;
    }
    auto okb = [&](uint32_t word, int byte, int j) { return ((word >> (8 * byte)) & 0xFFu) - ct[j] <= t2; };
#pragma unroll
    for (int k = 1; k <= AX_LMAX; ++k) {
        const bool in = k <= Lmax;
        const int kr = in ? k : Lmax;   // rows past Lmax are not in the tile; their tests are masked
        const uint32_t up = tc[-kr * AX_S], dn = tc[kr * AX_S];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int ql = j - k + 20, qr = j + k + 20;   // byte offsets from word -5
            gl[j] = gl[j] & in & okb(H[ql >> 2], ql & 3, j);
            gr[j] = gr[j] & in & okb(H[qr >> 2], qr & 3, j);
            gu[j] = gu[j] & in & okb(up, j, j);
            gd[j] = gd[j] & in & okb(dn, j, j);
            al[j] += gl[j];
            ar[j] += gr[j];
            au[j] += gu[j];
            ad[j] += gd[j];
        }
    }
    uint32_t av[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int xj = x + j;
        av[j] = (uint32_t)min(al[j], xj) | ((uint32_t)min(ar[j], w - 1 - xj) << 8) |
                ((uint32_t)min(au[j], y) << 16) | ((uint32_t)min(ad[j], h - 1 - y) << 24);
    }
    if (wide) *(uint4*)(arms + pi) = make_uint4(av[0], av[1], av[2], av[3]);
    else for (int j = 0; j < nq; ++j) arms[pi + j] = av[j];
}

template <int CT>
__global__ __launch_bounds__(256) void aux4_pair_kernel(AuxPair ap, int w, int h, int stride, int ch, int gmin,
                                                        int gmax, uint32_t* __restrict__ arms, int Lmax, int tau) {
    aux4_pair_body<CT>(block3(), ap, w, h, stride, ch, gmin, gmax, arms, Lmax, tau);
}

template <int CT>
struct Aux4PairFn {
    static constexpr int kThreads = 256;
    template <typename... A>
    __device__ static void run(const uint3 b, const A&... a) { aux4_pair_body<CT>(b, a...); }
};

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
                           const AuxViews& L, const AuxViews& R, uint32_t* rkeys, int16_t* dmin, int16_t* dmax,
                           uint32_t* arms, LaunchSink* sink) {
    AuxPair ap;
    ap.img[0] = AuxImage{L.src, L.gray, L.gx, L.gy, L.c32, L.c64, nullptr, dmin, dmax};
    ap.img[1] = AuxImage{R.src, R.gray, R.gx, R.gy, R.c32, R.c64, rkeys, nullptr, nullptr};
    const int Lmax = p.max_arm_length > 1 ? p.max_arm_length : 1;
    if (arms && Lmax > AX_LMAX) return hipErrorInvalidValue;
    dim3 grid((w + AX_PX - 1) / AX_PX, (h + AX_ROWS - 1) / AX_ROWS, 2), block(256);
    if (sink)
        return p.census_type == 1 ? sink->add<Aux4PairFn<1>>(grid, 0, ap, w, h, stride, ch, p.min_disparity,
                                                             p.max_disparity, arms, Lmax, p.color_threshold)
                                  : sink->add<Aux4PairFn<0>>(grid, 0, ap, w, h, stride, ch, p.min_disparity,
                                                             p.max_disparity, arms, Lmax, p.color_threshold);
    if (p.census_type == 1)
        hipLaunchKernelGGL(aux4_pair_kernel<1>, grid, block, 0, s, ap, w, h, stride, ch, p.min_disparity,
                           p.max_disparity, arms, Lmax, p.color_threshold);
    else
        hipLaunchKernelGGL(aux4_pair_kernel<0>, grid, block, 0, s, ap, w, h, stride, ch, p.min_disparity,
                           p.max_disparity, arms, Lmax, p.color_threshold);
    return hipGetLastError();
}

bool aux_pair_builds_arms(const apg_hip_params& p) {
    return p.agg_enable && (p.max_arm_length > 1 ? p.max_arm_length : 1) <= AX_LMAX;
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
