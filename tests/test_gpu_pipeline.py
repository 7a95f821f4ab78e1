"""GPU parity of the device pipeline (apg_hip_pipeline, C ABI) against the CPU
oracle on seeded synthetic pairs, for both schedules:
  stage-by-stage: every intermediate (C, C', S, right WTA, WTA output, best,
                  second) and every output must be bit-exact;
  fused:          S, right WTA, WTA output, best, second and every output
                  (disparity, confidence, invalid_reason, reliable mask)
                  must be bit-exact."""
import numpy as np
import pytest

import apg_amd as A
import pyoracle as O

# (kind, w, h, D, mode, overrides) -- odd sizes, odd D, D > 128, negative
# min disparity, 64-bit census, 8 paths, non-piecewise P2, u16 C' (cost_max
# above 254), large P2 (stage-by-stage fallback).
CASES = [
    (0, 64, 48, 16, 1, {}),
    (0, 97, 61, 24, 1, {}),
    (1, 160, 120, 48, 0, {}),
    (1, 160, 120, 48, 1, {}),
    (2, 131, 77, 33, 1, {}),
    (0, 96, 40, 12, 1, {"min_disparity": -4}),
    (1, 200, 64, 130, 1, {}),
    (2, 300, 50, 200, 1, {}),
    (0, 90, 60, 40, 1, {"paths": 8}),
    (1, 120, 80, 32, 1, {"census_type": 0}),
    (2, 100, 70, 28, 1, {"piecewise_p2": 0, "agg_iterations": 2}),
    (0, 33, 27, 16, 1, {"max_arm_length": 5, "color_threshold": 30}),
    (1, 150, 90, 64, 1, {"cost_max": 600}),
    (2, 140, 70, 40, 1, {"P2_base": 200}),
    (0, 1, 9, 4, 1, {}),
    (1, 70, 1, 8, 0, {}),
    # batched-transfer SGM (u8 C', D % 16 == 0) across the all-valid column
    # test: negative and positive min disparity, rays of 8 paths, K = 2
    (1, 200, 90, 64, 1, {"min_disparity": -16}),
    (2, 230, 70, 48, 1, {"min_disparity": 20}),
    (0, 180, 120, 64, 1, {"paths": 8}),
    (1, 420, 90, 192, 1, {"paths": 8}),
    # a negative AD weight: costs below 0 and above cost_max are clamped
    (2, 150, 90, 64, 1, {"eta_ad": -0.3, "cost_max": 200}),
    # path sums past the u16 bound: u32 accumulation (4 and 8 paths)
    (1, 120, 80, 40, 1, {"cost_max": 16383}),
    (2, 96, 64, 32, 1, {"paths": 8, "cost_max": 9000}),
    # median windows of radius 2, 5 and 9 (the general selection kernel)
    (0, 110, 70, 32, 1, {"post_median_radius": 2}),
    (1, 130, 90, 40, 1, {"post_median_radius": 5}),
    (2, 75, 50, 24, 0, {"post_median_radius": 9}),
]

FUSED_KEYS = ["sgm", "disp_right", "disp_wta", "best", "second", "after_refine",
              "disparity", "confidence", "invalid_reason", "reliable"]
PLAIN_KEYS = ["cost", "cost_agg"] + FUSED_KEYS


def _check(g, o, keys):
    for k in keys:
        a, b = getattr(g, k), o[k]
        assert a.shape == b.shape, k
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            i = tuple(bad[0])
            raise AssertionError(f"{k}: {len(bad)} mismatches, first at {i}: gpu={a[i]} oracle={b[i]}")


def _case(kind, w, h, D, mode, over):
    p = A.params(mode, D, **over)
    if "min_disparity" in over:
        p.max_disparity = over["min_disparity"] + D
    left, right, _ = A.synth_pair(kind, w, h, max(D, 2), 1234 + w)
    o = O.pipeline(O.cfg_from(p.as_dict()), left, right, intermediates=True)
    return p, left, right, o


@pytest.mark.gpu
@pytest.mark.parametrize("kind,w,h,D,mode,over", CASES)
def test_pipeline_stage_by_stage_bit_exact(ctx, kind, w, h, D, mode, over):
    print(f"stage-by-stage kind={kind} {w}x{h} D={D} mode={mode} {over}", flush=True)
    p, left, right, o = _case(kind, w, h, D, mode, over)
    A.set_schedule(ctx, 1)
    try:
        g = A.run_pipeline(ctx, p, left, right, intermediates=True)
    finally:
        A.set_schedule(ctx, 0)
    _check(g, o, PLAIN_KEYS)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,w,h,D,mode,over", CASES)
def test_pipeline_fused_bit_exact(ctx, kind, w, h, D, mode, over):
    print(f"fused kind={kind} {w}x{h} D={D} mode={mode} {over}", flush=True)
    p, left, right, o = _case(kind, w, h, D, mode, over)
    A.set_schedule(ctx, 0)
    g = A.run_pipeline(ctx, p, left, right, intermediates=A.FUSED_INTERMEDIATES)
    _check(g, o, FUSED_KEYS)
    # and without any intermediate requested (the production call)
    g2 = A.run_pipeline(ctx, p, left, right)
    _check(g2, o, ["disparity", "confidence", "invalid_reason", "reliable"])


# every SGM wave shape of the fused schedule (launch-shape option sgm_dual):
# 64-lane dual and single chains, 32-lane half waves, and (4 paths) 16-lane
# quarter waves over two scanlines -- odd scanline counts (the quarter form's
# last wave holds one scanline twice), D < 128, negative min disparity, 8 paths
SGM_SHAPE_CASES = [
    (1, 160, 121, 128, {}),
    (2, 97, 64, 48, {"min_disparity": -8}),
    (0, 131, 77, 64, {}),
    (1, 90, 60, 32, {"paths": 8}),
    (2, 200, 90, 112, {"min_disparity": 10}),
]


@pytest.mark.gpu
@pytest.mark.parametrize("sgm_dual", [1, 2, 3, 4])
@pytest.mark.parametrize("kind,w,h,D,over", SGM_SHAPE_CASES)
def test_pipeline_sgm_wave_shapes(ctx, kind, w, h, D, over, sgm_dual):
    print(f"sgm_dual={sgm_dual} kind={kind} {w}x{h} D={D} {over}", flush=True)
    p, left, right, o = _case(kind, w, h, D, 1, over)
    A.set_schedule(ctx, 0)
    A.set_tuning(ctx, (0, sgm_dual, 0))
    try:
        g = A.run_pipeline(ctx, p, left, right, intermediates=A.FUSED_INTERMEDIATES)
    finally:
        A.set_tuning(ctx, None)
    _check(g, o, FUSED_KEYS)


# the fused schedule without S requested (the WTA stage keeps its sums in
# LDS only): odd and even widths, narrow and wide images, negative and
# positive min disparity, Fast (no Cross)
SGM_WTA_CASES = [
    (1, 160, 121, 128, 1, {}),
    (2, 97, 64, 48, 1, {"min_disparity": -8}),
    (0, 131, 77, 64, 1, {}),
    (2, 200, 90, 112, 1, {"min_disparity": 10}),
    (1, 1, 9, 16, 1, {}),
    (0, 2, 5, 16, 1, {}),
    (1, 3, 7, 16, 1, {}),
    (2, 17, 11, 32, 1, {}),
    (0, 95, 33, 16, 1, {}),
    (1, 96, 20, 128, 1, {}),
    (2, 129, 40, 80, 1, {"min_disparity": -20}),
    (0, 400, 30, 128, 1, {}),
    (1, 301, 25, 96, 0, {}),
    (2, 777, 12, 64, 1, {"min_disparity": 5}),
]
WTA_KEYS = ["disp_right", "disp_wta", "best", "second", "after_refine", "disparity", "confidence",
            "invalid_reason", "reliable"]


@pytest.mark.gpu
@pytest.mark.parametrize("kind,w,h,D,mode,over", SGM_WTA_CASES)
def test_pipeline_fused_wta_outputs(ctx, kind, w, h, D, mode, over):
    print(f"fused wta outputs kind={kind} {w}x{h} D={D} mode={mode} {over}", flush=True)
    p, left, right, o = _case(kind, w, h, D, mode, over)
    A.set_schedule(ctx, 0)
    g = A.run_pipeline(ctx, p, left, right, intermediates=tuple(WTA_KEYS[:5]))
    _check(g, o, WTA_KEYS)
    # the same pair without intermediates (the WTA outputs go to the workspace)
    g2 = A.run_pipeline(ctx, p, left, right)
    _check(g2, o, WTA_KEYS[5:])


# the aux pass (four pixels per thread over a 64 x 16 tile, the left view's
# Cross arms from the same tile up to arm length 17, the separate arms kernel
# beyond): widths around the tile and word edges, heights around the row
# tiles, arm lengths 1..30, thresholds 0 and 255, the 64-bit census
AUX_CASES = [
    (0, 63, 15, 16, 1, {"max_arm_length": 1}),
    (1, 65, 17, 16, 1, {"max_arm_length": 2, "color_threshold": 0}),
    (2, 66, 33, 16, 1, {"max_arm_length": 17, "color_threshold": 40}),
    (0, 67, 16, 16, 1, {"max_arm_length": 18}),
    (1, 130, 47, 24, 1, {"max_arm_length": 4, "color_threshold": 255}),
    (2, 129, 50, 16, 1, {"census_type": 0, "max_arm_length": 16}),
    (0, 4, 3, 4, 1, {"max_arm_length": 17}),
    (1, 200, 31, 32, 1, {"max_arm_length": 30, "color_threshold": 12}),
    (2, 6, 40, 4, 1, {"max_arm_length": 9}),
]


@pytest.mark.gpu
@pytest.mark.parametrize("sched", [1, 0], ids=["stage_by_stage", "fused"])
@pytest.mark.parametrize("kind,w,h,D,mode,over", AUX_CASES)
def test_pipeline_aux_pair_edges(ctx, kind, w, h, D, mode, over, sched):
    print(f"aux pair sched={sched} kind={kind} {w}x{h} D={D} {over}", flush=True)
    p, left, right, o = _case(kind, w, h, D, mode, over)
    A.set_schedule(ctx, sched)
    try:
        g = A.run_pipeline(ctx, p, left, right, intermediates=True if sched == 1 else A.FUSED_INTERMEDIATES)
    finally:
        A.set_schedule(ctx, 0)
    _check(g, o, PLAIN_KEYS if sched == 1 else FUSED_KEYS)


# fill + median: one tiled pass for median radius 0 / 1 and max_fill_gap <=
# 62 (a 64-column halo), the row fill and the median kernels otherwise; gaps
# 0, 5, 40, 62, 63, 100, widths that leave partial tiles, a strict LR check
# (more holes)
POST_CASES = [
    (0, 150, 40, 32, 1, {"post_max_fill_gap": 0}),
    (1, 130, 35, 48, 1, {"post_max_fill_gap": 62, "post_lr_max_diff": 0}),
    (2, 140, 33, 48, 1, {"post_max_fill_gap": 63, "post_lr_max_diff": 0}),
    (0, 200, 20, 64, 1, {"post_max_fill_gap": 5, "post_median_radius": 0}),
    (1, 65, 17, 16, 1, {"post_max_fill_gap": 40}),
    (2, 300, 48, 96, 1, {"post_lr_max_diff": 0}),
    (0, 190, 70, 64, 0, {"post_max_fill_gap": 100, "post_lr_max_diff": 0}),
]


@pytest.mark.gpu
@pytest.mark.parametrize("kind,w,h,D,mode,over", POST_CASES)
def test_pipeline_fill_median_fused(ctx, kind, w, h, D, mode, over):
    print(f"fill+median kind={kind} {w}x{h} D={D} {over}", flush=True)
    p, left, right, o = _case(kind, w, h, D, mode, over)
    A.set_schedule(ctx, 0)
    g = A.run_pipeline(ctx, p, left, right)
    _check(g, o, WTA_KEYS[5:])
    g = A.run_pipeline(ctx, p, left, right, intermediates=A.FUSED_INTERMEDIATES)
    _check(g, o, FUSED_KEYS)


@pytest.mark.gpu
@pytest.mark.parametrize("h,w", [(19, 67), (33, 130)])
def test_pipeline_color_input_odd_widths(ctx, h, w):
    rng = np.random.default_rng(w)
    left = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    right = np.roll(left, -5, axis=1)
    p = A.params(1, 16, census_type=0 if w > 100 else 1)
    o = O.pipeline(O.cfg_from(p.as_dict()), left, right, intermediates=True)
    for sched in (1, 0):
        print(f"color {w}x{h} sched={sched}", flush=True)
        A.set_schedule(ctx, sched)
        g = A.run_pipeline(ctx, p, left, right, intermediates=True if sched == 1 else A.FUSED_INTERMEDIATES)
        _check(g, o, PLAIN_KEYS if sched == 1 else FUSED_KEYS)
    A.set_schedule(ctx, 0)


@pytest.mark.gpu
def test_pipeline_color_input(ctx):
    rng = np.random.default_rng(5)
    left = rng.integers(0, 256, (45, 70, 3), dtype=np.uint8)
    right = np.roll(left, -6, axis=1)
    p = A.params(1, 24)
    o = O.pipeline(O.cfg_from(p.as_dict()), left, right, intermediates=True)
    for sched in (1, 0):
        A.set_schedule(ctx, sched)
        g = A.run_pipeline(ctx, p, left, right, intermediates=A.FUSED_INTERMEDIATES)
        _check(g, o, FUSED_KEYS)
    A.set_schedule(ctx, 0)


@pytest.mark.gpu
def test_pipeline_full_size_fused_vs_oracle(ctx):
    """Balanced 1280x720 D=128 (the bench workload): fused outputs bit-exact."""
    p = A.params(1, 128)
    left, right, _ = A.synth_pair(1, 1280, 720, 128, 20260101)
    O.set_threads(16)
    o = O.pipeline(O.cfg_from(p.as_dict()), left, right, intermediates=True)
    g = A.run_pipeline(ctx, p, left, right, intermediates=("disp_right", "disp_wta", "best", "second"))
    _check(g, o, ["disp_right", "disp_wta", "best", "second", "disparity", "confidence", "invalid_reason", "reliable"])
