"""GPU parity of the device pipeline (apg_hip_pipeline, C ABI) against the CPU
oracle on seeded synthetic pairs: every intermediate (C, C', S, right WTA, WTA
output, best, second) and every output (disparity, confidence,
invalid_reason, reliable mask) must be bit-exact."""
import numpy as np
import pytest

import apg_amd as A
import pyoracle as O

# (kind, w, h, max_disp, mode, overrides) -- odd sizes, odd D, D > 128,
# negative min disparity, 64-bit census, 8 paths, non-piecewise P2.
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
]


def _check(g, o, keys):
    for k in keys:
        a, b = getattr(g, k), o[k]
        assert a.shape == b.shape, k
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            i = tuple(bad[0])
            raise AssertionError(f"{k}: {len(bad)} mismatches, first at {i}: gpu={a[i]} oracle={b[i]}")


@pytest.mark.gpu
@pytest.mark.parametrize("kind,w,h,D,mode,over", CASES)
def test_pipeline_bit_exact(ctx, kind, w, h, D, mode, over):
    print(f"case kind={kind} {w}x{h} D={D} mode={mode} {over}", flush=True)
    p = A.params(mode, D, **over)
    if "min_disparity" in over:
        p.max_disparity = over["min_disparity"] + D
    left, right, _ = A.synth_pair(kind, w, h, max(D, 2), 1234 + w)
    g = A.run_pipeline(ctx, p, left, right, intermediates=True)
    c = O.cfg_from(p.as_dict())
    o = O.pipeline(c, left, right, intermediates=True)
    _check(g, o, ["cost", "cost_agg", "sgm", "disp_right", "disp_wta", "best", "second",
                  "disparity", "confidence", "invalid_reason", "reliable"])


@pytest.mark.gpu
def test_pipeline_color_input(ctx):
    rng = np.random.default_rng(5)
    left = rng.integers(0, 256, (45, 70, 3), dtype=np.uint8)
    right = np.roll(left, -6, axis=1)
    p = A.params(1, 24)
    g = A.run_pipeline(ctx, p, left, right, intermediates=True)
    o = O.pipeline(O.cfg_from(p.as_dict()), left, right, intermediates=True)
    _check(g, o, ["cost", "cost_agg", "sgm", "disparity", "confidence", "invalid_reason", "reliable"])
