// Seeded synthetic stereo pairs for benchmarks and tests (host code).
// Integer-only PRNG (splitmix64) so that every host reproduces the bytes;
// only kind 2 evaluates a float surface (sinf), identical on one image.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/apg_sgm_hip.h"

namespace {

struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    int uniform(int lo, int hi) {  // inclusive
        if (hi <= lo) return lo;
        return lo + (int)(next() % (uint64_t)(hi - lo + 1));
    }
    uint8_t byte() { return (uint8_t)(next() >> 56); }
    // N(0, sigma=2) rounded: Irwin-Hall sum of 12 16-bit uniforms has mean
    // 393210 and standard deviation ~65536; (s - mean) / 32768 has sigma 2.
    int gauss2() {
        int64_t s = 0;
        for (int i = 0; i < 12; ++i) s += (int64_t)(next() >> 48);
        return (int)((s - 393210 + 16384) >> 15);
    }
};

inline uint8_t clamp8(int v) { return (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v)); }

void texture(Rng& r, int w, int h, std::vector<uint8_t>& out) {
    std::vector<uint8_t> noise((size_t)w * h);
    for (auto& v : noise) v = r.byte();
    out.assign((size_t)w * h, 0);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            int s = 0;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int xx = std::min(std::max(x + dx, 0), w - 1), yy = std::min(std::max(y + dy, 0), h - 1);
                    s += noise[(size_t)yy * w + xx];
                }
            out[(size_t)y * w + x] = (uint8_t)((s + 4) / 9);
        }
}

}  // namespace

extern "C" int apg_synth_pair(int kind, int w, int h, int max_disp, uint64_t seed, uint8_t* left, uint8_t* right,
                              float* gt_out) {
    if (w <= 0 || h <= 0 || !left || !right || max_disp < 2) return 1;
    Rng rng{seed * 0x2545F4914F6CDD1Dull + 0x1234567ull};
    std::vector<uint8_t> tex;
    texture(rng, w, h, tex);
    const size_t n = (size_t)w * h;
    std::vector<float> gt(n, 0.f);
    const int dcap = std::max(1, max_disp - 1);
    if (kind == 0 || kind == 1) {
        std::vector<int> di(n, 0);
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                int bg = 16;
                if (kind == 1) bg = 4 + (int)(((int64_t)(x + y) * 24) / std::max(1, w + h));
                di[(size_t)y * w + x] = std::min(bg, dcap);
            }
        const int nrect = kind == 0 ? 4 : 8;
        const int dlo = kind == 0 ? 8 : 4;
        const int dhi = std::max(dlo, std::min(kind == 0 ? 56 : 120, max_disp - 8));
        const int smin = std::max(2, std::min(60, w / 4)), smax = std::max(smin, std::min(200, w / 2));
        const int tmin = std::max(2, std::min(60, h / 4)), tmax = std::max(tmin, std::min(200, h / 2));
        for (int k = 0; k < nrect; ++k) {
            const int rw = rng.uniform(smin, smax), rh = rng.uniform(tmin, tmax);
            const int rx = rng.uniform(0, std::max(0, w - rw)), ry = rng.uniform(0, std::max(0, h - rh));
            const int d = std::min(rng.uniform(dlo, dhi), dcap);
            for (int y = ry; y < std::min(h, ry + rh); ++y)
                for (int x = rx; x < std::min(w, rx + rw); ++x) di[(size_t)y * w + x] = d;
        }
        // forward warp with a z-buffer: right(x - d) = left(x), larger d wins
        std::vector<int> zb(n, -1);
        std::vector<uint8_t> rt(n, 0);
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                const int d = di[(size_t)y * w + x];
                const int xr = x - d;
                if (xr < 0) continue;
                const size_t q = (size_t)y * w + xr;
                if (d > zb[q]) { zb[q] = d; rt[q] = tex[(size_t)y * w + x]; }
            }
        for (size_t q = 0; q < n; ++q)
            if (zb[q] < 0) rt[q] = rng.byte();
        int gq = 256, bias = 0;
        if (kind == 1) { gq = rng.uniform(230, 282); bias = rng.uniform(-10, 10); }
        for (size_t q = 0; q < n; ++q) {
            left[q] = clamp8((int)tex[q] + rng.gauss2());
            int rv = rt[q];
            if (kind == 1) rv = ((rv * gq + 128) >> 8) + bias;
            right[q] = clamp8(rv + rng.gauss2());
            gt[q] = (float)di[q];
        }
    } else {
        const float lo = std::min(10.f, (float)dcap), hi = std::max(lo, std::min(240.f, (float)(max_disp - 10)));
        const float mid = 0.5f * (lo + hi);
        const float a = (float)rng.uniform(-600, 600) * 1e-4f;   // |dd/dx| <= 0.06
        const float b = (float)rng.uniform(-400, 400) * 1e-4f;   // |dd/dy| <= 0.04
        const float amp = std::min(20.f, 0.25f * (hi - lo));
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                float d = mid + a * (float)(x - w / 2) + b * (float)(y - h / 2) +
                          amp * sinf(6.2831853f * (float)x / 200.f);
                d = std::min(std::max(d, lo), hi);
                gt[(size_t)y * w + x] = d;
            }
        // backward warp with linear interpolation: right(xr) = left(xr + d(xr))
        for (int y = 0; y < h; ++y)
            for (int xr = 0; xr < w; ++xr) {
                const float xs = (float)xr + gt[(size_t)y * w + xr];
                const int x0 = (int)floorf(xs);
                const float f = xs - (float)x0;
                int v;
                if (x0 + 1 < w) {
                    const float l0 = tex[(size_t)y * w + x0], l1 = tex[(size_t)y * w + x0 + 1];
                    v = (int)floorf(l0 + f * (l1 - l0) + 0.5f);
                } else {
                    v = rng.byte();
                }
                right[(size_t)y * w + xr] = clamp8(v + rng.gauss2());
            }
        for (size_t q = 0; q < n; ++q) left[q] = clamp8((int)tex[q] + rng.gauss2());
    }
    if (gt_out) memcpy(gt_out, gt.data(), n * sizeof(float));
    return 0;
}
