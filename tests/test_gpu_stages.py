"""GPU stage kernels called one at a time through the C ABI, against the
oracle on the same inputs (bit-exact)."""
import ctypes as C

import numpy as np
import pytest

import apg_amd as A
import pyoracle as O


@pytest.mark.gpu
@pytest.mark.parametrize("lmax,tau", [(17, 12), (1, 12), (62, 40), (5, 0), (30, 300), (9, -1)])
def test_cross_arms_match_oracle(ctx, lmax, tau):
    # smooth regions (long arms, early-exit loop runs to Lmax), noise (short
    # arms), image borders on every side of odd-sized tiles
    print(f"cross arms Lmax={lmax} tau={tau}", flush=True)
    rng = np.random.default_rng(lmax * 1000 + tau)
    h, w = 77, 203
    yy, xx = np.mgrid[0:h, 0:w]
    gray = ((xx // 23 + yy // 19) * 37 % 256).astype(np.uint8)
    noisy = rng.random((h, w)) < 0.3
    gray[noisy] = rng.integers(0, 256, noisy.sum(), dtype=np.uint8)
    p = A.params(1, 16, max_arm_length=lmax, color_threshold=tau)
    c = O.cfg_from(p.as_dict())
    want = O.cross_arms(c, gray)
    dg = ctx.to_device(gray)
    darms = ctx.array((h, w), np.uint32)
    ctx._check(ctx.lib.apg_hip_cross_arms(ctx.h, C.byref(p), dg.ptr, w, h, darms.ptr), "cross_arms")
    got = darms.download()
    for k, name in enumerate(("left", "right", "up", "down")):
        g = ((got >> (8 * k)) & 0xFF).astype(np.int16)
        assert np.array_equal(g, want[k]), f"{name} arm differs at {np.argwhere(g != want[k])[:3].tolist()}"
    dg.free()
    darms.free()


@pytest.mark.gpu
@pytest.mark.parametrize("paths,cost_max", [(4, 16383), (8, 9000), (4, 255)])
def test_sgm_optimize_stage_u16_and_u32_sums(ctx, paths, cost_max):
    # apg_hip_sgm_optimize on the oracle's aggregated volume: u16 sums widened,
    # or past the u16 bound u32 accumulation (d_sum16 then holds one path)
    p = A.params(1, 24, paths=paths, cost_max=cost_max)
    c = O.cfg_from(p.as_dict())
    left, right, _ = A.synth_pair(1, 90, 60, 24, 8)
    o = O.pipeline(c, left, right, intermediates=True)
    g = O.aux(c, left)[0]
    h, w, D = o["cost_agg"].shape
    print(f"sgm_optimize paths={paths} cost_max={cost_max}", flush=True)
    dg, dc = ctx.to_device(g), ctx.to_device(o["cost_agg"])
    s16, s32 = ctx.array((h, w, D), np.uint16), ctx.array((h, w, D), np.uint32)
    ctx._check(ctx.lib.apg_hip_sgm_optimize(ctx.h, C.byref(p), w, h, dg.ptr, dc.ptr, s16.ptr, s32.ptr), "optimize")
    np.testing.assert_array_equal(s32.download(), o["sgm"])
    for a in (dg, dc, s16, s32):
        a.free()
