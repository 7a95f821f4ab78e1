"""HighQuality pipeline on the GPU (prior-narrowed ranges, 8 paths, refiner)
against the oracle and against the golden vectors the reference build wrote:
ranges, prior, every volume, d_before_refine and every output bit-exact."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import apg_amd as A
import pyoracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# (kind, w, h, D, overrides): both census types, the case-B budget, a prior
# that finds too few supports, D > 128 (K = 2 lanes), odd sizes
CASES = [
    (2, 160, 120, 48, {}),
    (0, 200, 150, 64, {}),
    (1, 256, 128, 96, {"census_type": 0}),
    (2, 120, 90, 32, {"prior_max_supports": 10, "prior_min_supports": 5}),
    (0, 96, 64, 24, {"prior_min_supports": 10 ** 6}),
    (1, 333, 101, 160, {}),
    # the refiner alone: a tall image (the sweep's diagonals clamp on the width),
    # one and four sweeps (last sweep backward / forward), none, radius 1 and 37
    (2, 64, 200, 32, {}),
    (0, 128, 96, 32, {"refine_iterations": 1}),
    (1, 128, 96, 32, {"refine_iterations": 4, "refine_random_radius": 37}),
    (2, 112, 80, 32, {"refine_iterations": 0}),
    (0, 100, 70, 32, {"refine_random_radius": 1}),
    # ranges reaching below 0
    (2, 128, 96, 32, {"min_disparity": -8}),
    # heights around one wavefront of rows
    (1, 90, 64, 24, {}),
    (0, 90, 65, 24, {}),
    # the banded sweep (64-row bands handing their last row over column by
    # column): narrow, many bands, D = 256 windows, each cost-term combination
    (0, 20, 300, 16, {}),
    (2, 97, 1000, 32, {}),
    (1, 200, 300, 256, {}),
    (2, 160, 200, 64, {"use_ad": 0}),
    (0, 150, 130, 48, {"use_grad": 0}),
    (1, 140, 70, 32, {"use_ad": 0, "use_grad": 0, "min_disparity": 5}),
]

KEYS = ["cost", "cost_agg", "disp_right", "disp_wta", "best", "second", "dmin", "dmax", "d_prior", "after_refine",
        "disparity", "confidence", "invalid_reason", "reliable"]


def _cmp(got, want, what):
    assert got.shape == want.shape, what
    if not np.array_equal(got, want, equal_nan=True):
        bad = np.argwhere(got != want)
        i = tuple(bad[0])
        raise AssertionError(f"{what}: {len(bad)} mismatches, first at {i}: gpu={got[i]} oracle={want[i]}")


@pytest.mark.gpu
@pytest.mark.parametrize("kind,w,h,D,over", CASES)
def test_hq_pipeline_matches_oracle(ctx, kind, w, h, D, over):
    print(f"HQ kind={kind} {w}x{h} D={D} {over}", flush=True)
    p = A.params(2, D, **over)
    left, right, _ = A.synth_pair(kind, w, h, D, 4321 + w)
    o = O.pipeline(O.cfg_from(p.as_dict()), left, right, intermediates=True)
    g = A.run_pipeline(ctx, p, left, right, intermediates=True)
    for k in KEYS:
        _cmp(getattr(g, k), o[k], k)
    # the u32 sums of the oracle and the GPU's u16 sums widened on the way out
    _cmp(g.sgm, o["sgm"], "sgm")


FUSED_KEYS = ["dmin", "dmax", "d_prior", "disp_wta", "after_refine", "disparity", "confidence", "invalid_reason",
              "reliable"]


@pytest.mark.gpu
@pytest.mark.parametrize("kind,w,h,D,over", CASES)
def test_hq_fused_schedule_matches_oracle(ctx, kind, w, h, D, over):
    """The same cases without raw volumes requested, so the fused general-range
    Cost + Cross kernels and the delta-plane SGM run where they apply."""
    print(f"HQ fused kind={kind} {w}x{h} D={D} {over}", flush=True)
    p = A.params(2, D, **over)
    left, right, _ = A.synth_pair(kind, w, h, D, 4321 + w)
    o = O.pipeline(O.cfg_from(p.as_dict()), left, right, intermediates=True)
    g = A.run_pipeline(ctx, p, left, right, intermediates=tuple(k for k in FUSED_KEYS[:5]))
    for k in FUSED_KEYS:
        _cmp(getattr(g, k), o[k], k)


def _golden_small():
    z = np.load(os.path.join(GOLD, "golden_small.npz"))
    return z, sorted({k.split("__")[0] for k in z.files if k.startswith("hq_")})


@pytest.mark.gpu
@pytest.mark.parametrize("name", _golden_small()[1])
def test_hq_golden_vectors(ctx, name):
    # straight against the reference build's outputs (packed-volume pipeline)
    z, _ = _golden_small()
    p = A.Params()
    for k, v in json.loads(bytes(z[f"{name}__cfg"]).decode()).items():
        setattr(p, k, v)
    print(f"HQ golden {name}", flush=True)
    g = A.run_pipeline(ctx, p, z[f"{name}__left"], z[f"{name}__right"], intermediates=True)
    for k in ("cost", "cost_agg", "best", "second", "disparity", "confidence", "invalid_reason", "reliable",
              "disp_right", "dmin", "dmax"):
        _cmp(getattr(g, k), z[f"{name}__{k}"], k)
    _cmp(g.disp_wta, z[f"{name}__d_before_refine"], "d_before_refine")


@pytest.mark.gpu
def test_prior_entry_point_matches_oracle(ctx):
    p = A.params(2, 64)
    left, right, _ = A.synth_pair(2, 240, 160, 64, 77)
    c = O.cfg_from(p.as_dict())
    _, gxl, gyl, cl = O.aux(c, left)
    _, _, _, cr = O.aux(c, right)
    want_min, want_max, want_dp, want_n = O.prior_estimate(c, gxl, gyl, cl, cr)
    h, w = left.shape
    dgx, dgy, dcl, dcr = (ctx.to_device(a) for a in (gxl, gyl, cl, cr))
    dmin, dmax, dp = ctx.array((h, w), np.int16), ctx.array((h, w), np.int16), ctx.array((h, w), np.float32)
    n = C.c_int(-1)
    print("prior entry point 240x160 D=64", flush=True)
    ctx._check(ctx.lib.apg_hip_prior_estimate(ctx.h, C.byref(p), w, h, dgx.ptr, dgy.ptr, dcl.ptr, dcr.ptr, None,
                                              None, dmin.ptr, dmax.ptr, dp.ptr, C.byref(n)), "prior")
    assert n.value == want_n
    _cmp(dmin.download(), want_min, "dmin")
    _cmp(dmax.download(), want_max, "dmax")
    _cmp(dp.download(), want_dp, "d_prior")


@pytest.mark.gpu
@pytest.mark.parametrize("shift,maxd,over", [
    (-2, 16, {"min_disparity": -8, "prior_search_radius": 1}),
    (4, 0, {"min_disparity": -16, "refine_random_radius": 3}),
])
def test_hq_negative_ranges_match_oracle(ctx, shift, maxd, over):
    # a scene at negative disparity, and a range that ends at 0
    rng = np.random.default_rng(17 + shift)
    left = rng.integers(0, 256, (48, 80), dtype=np.uint8)
    right = np.roll(left, shift, axis=1)
    p = A.params(2, maxd, **over)
    print(f"HQ negative range shift={shift} max={maxd} {over}", flush=True)
    o = O.pipeline(O.cfg_from(p.as_dict()), left, right, intermediates=True)
    g = A.run_pipeline(ctx, p, left, right, intermediates=True)
    for k in KEYS:
        _cmp(getattr(g, k), o[k], k)



@pytest.mark.gpu
def test_refiner_band_wait_gives_up_and_recovers(ctx):
    """The banded sweep's hand-off wait, made to give up by the host: the call
    fails with status 8 and every wave exits (the next call, with the default
    wait, runs bit-exact against the oracle)."""
    p = A.params(2, 64)
    left, right, _ = A.synth_pair(0, 200, 150, 64, 3)
    print("refiner band wait: forced give-up", flush=True)
    ctx._check(ctx.lib.apg_hip_set_refine_wait(ctx.h, -1), "set_refine_wait")
    try:
        with pytest.raises(A.ApgError) as ei:
            A.run_pipeline(ctx, p, left, right)
        assert ei.value.code == 8
    finally:
        ctx._check(ctx.lib.apg_hip_set_refine_wait(ctx.h, 0), "set_refine_wait")
    print("refiner band wait: default", flush=True)
    g = A.run_pipeline(ctx, p, left, right)
    o = O.pipeline(O.cfg_from(p.as_dict()), left, right)
    for k in ("disparity", "invalid_reason", "reliable"):
        assert np.array_equal(getattr(g, k), o[k], equal_nan=True), k
