"""Pins the CPU oracle (oracle/liboracle.so) before it is trusted as the GPU
checker: the reference's own frozen fingerprints and KATs, golden vectors
written from a build of the reference sources (tests/golden/), and
closed-form / brute-force properties chosen so that a dropped term, a wrong
sign or index or a transposed operand fails at least one check."""
import json
import os

import numpy as np
import pytest

import apg_amd as A
import pyoracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# --------------------------------------------- reference fingerprints --------
def pattern(w, h, ch, kind):
    """generate_pattern of tests/test_api_contract.cpp:43-74 (Checkerboard,
    DeterministicNoise)."""
    x = np.arange(w)[None, :]
    y = np.arange(h)[:, None]
    if kind == "checker":
        v = np.where(((x // 4 + y // 4) % 2) == 0, 40, 210)
    else:
        v = (x * 37 + y * 73 + x * y * 13) % 256
    v = np.broadcast_to(v, (h, w)).astype(np.int64)
    if ch == 1:
        return np.ascontiguousarray(v.astype(np.uint8))
    return np.ascontiguousarray(np.stack([v, (v + 30) % 256, (v + 60) % 256], -1).astype(np.uint8))


# tests/test_api_contract.cpp:243-257 (frozen; identical for Dense and Packed)
FINGERPRINTS = {
    "Odd_Checkerboard": (33, 27, 1, "checker", 0, 612, 0x1710451f7eab215, 0xe61822cc0fbe3a38,
                         0x22081e533cd30004, 0x9ade9d63e4fcae7b),
    "Color_Noise": (32, 24, 3, "noise", 0, 720, 0x7e8e0fa784351325, 0x5860e8f479cbe0fd,
                    0x82f55cc1fd68a695, 0x9db48cb5d76242b5),
}


def fingerprint(o):
    return ((o["disparity"] < 0).sum(), (o["reliable"] != 0).sum(), O.fnv1a64(o["disparity"]),
            O.fnv1a64(o["confidence"]), O.fnv1a64(o["invalid_reason"]), O.fnv1a64(o["reliable"]))


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("name", list(FINGERPRINTS))
def test_frozen_fingerprints(mode, name):
    w, h, ch, kind, *exp = FINGERPRINTS[name]
    img = pattern(w, h, ch, kind)
    o = O.pipeline(O.cfg(mode, 16), img, img)
    assert fingerprint(o) == tuple(exp)


def test_frozen_fingerprint_negative_range():
    img = pattern(48, 36, 1, "noise")
    o = O.pipeline(O.cfg(1, 8, min_disparity=-4), img, img)
    assert fingerprint(o) == (0, 178, 0x1d8e1d22713fcee9, 0x77f3c5cdbc782a05, 0x8ec5011cb7fb425, 0x3563e2187fb4a965)


# ------------------------------------------------------------ KATs -----------
def test_kat_sgm_1d_noise_suppression():
    """tests/test_synthetic.cpp:26-71: with P1=25, P2=80 (fixed) the isolated
    perturbation at pixel 2 must not win; disparity stays 0."""
    c = O.cfg(1, 3, P1=25, P2_base=80, adaptive_penalties=0)
    cost = np.zeros((1, 5, 3), np.uint16)
    cost[0, :, :] = [10, 60, 70]
    cost[0, 2, :] = [30, 20, 70]
    gray = np.full((1, 5), 128, np.uint8)
    acc = O.sgm_path(c, gray, cost, np.zeros((1, 5, 3), np.uint32), 1, 0)
    dmin = np.zeros((1, 5), np.int16)
    dmax = np.full((1, 5), 3, np.int16)
    disp, _, _ = O.wta(c, acc, dmin, dmax)
    assert disp[0, 2] == pytest.approx(0.0, abs=0.01)
    # closed form of the recurrence at pixel 2, d=1: L = 20 + min(L1(1)=60, L1(0)+P1=35, L1(2)+P1=95, 10+80) - 10
    assert acc[0, 2, 1] == 20 + 35 - 10


def test_kat_wta_high_cost_and_tie():
    """tests/test_sanity.cpp:1858-1938: costs above 32767, invalid mixtures and
    the earliest-index tie break."""
    c = O.cfg(1, 64, post_subpixel=0)
    tw, th = 16, 8
    vol = np.empty((th, tw, 64), np.uint16)
    for y in range(th):
        for x in range(tw):
            s = 45000 + np.arange(64) * 10
            s[10] = s[20] = 40000
            vol[y, x] = s
    dmin = np.zeros((th, tw), np.int16)
    dmax = np.full((th, tw), 64, np.int16)
    disp, best, second = O.wta(c, vol, dmin, dmax)
    assert (disp == 10).all() and (best == 40000).all() and (second == 40000).all()
    vol32 = vol.astype(np.uint32)
    vol32[:, :, 0] = 0xFFFFFFFF
    d2, b2, s2 = O.wta(c, vol32, dmin, dmax)
    assert (d2 == 10).all()


@pytest.mark.parametrize("cm,c0,cp,exp", [
    (20, 10, 20, 0.0),                      # symmetric
    (12, 10, 20, 0.5 * (12 - 20) / (12 - 20 + 20)),   # left-biased, closed form
    (20, 10, 12, 0.5 * (20 - 12) / (20 - 20 + 12)),   # right-biased
    (10, 10, 10, 0.0),                      # flat: denominator 0
    (10, 20, 10, 0.0),                      # peak: 0.5*0/-20
    (100, 10, 11, 0.5 * 89 / 91),           # unclamped, |delta| < 0.5
    (40, 12, 10, 0.5),                      # 0.5*30/26 > 0.5 -> clamp +0.5
    (10, 12, 40, -0.5),                     # clamp -0.5
])
def test_subpixel_legacy_cases(cm, c0, cp, exp):
    """tests/test_subpixel.cpp:10-99 (Legacy) and the closed form."""
    d = O.subpixel_legacy(cm, c0, cp)
    assert d == pytest.approx(exp, abs=1e-7)
    assert -0.5 <= d <= 0.5


# -------------------------------------------------- reference golden vectors -
def _small_cases():
    z = np.load(os.path.join(GOLD, "golden_small.npz"))
    names = sorted({k.split("__")[0] for k in z.files})
    return z, names


@pytest.mark.parametrize("name", _small_cases()[1])
def test_golden_small(name):
    z, _ = _small_cases()
    cfg = O.cfg_from(json.loads(bytes(z[f"{name}__cfg"]).decode()))
    o = O.pipeline(cfg, z[f"{name}__left"], z[f"{name}__right"], intermediates=True)
    for k in ("cost", "cost_agg", "sgm", "disp_wta", "best", "second", "disparity", "confidence",
              "invalid_reason", "reliable", "disp_right", "dmin", "dmax"):
        np.testing.assert_array_equal(o[k], z[f"{name}__{k}"], err_msg=k)
    np.testing.assert_array_equal(o["disp_wta"], z[f"{name}__d_before_refine"], err_msg="d_before_refine")


def _hashed():
    with open(os.path.join(GOLD, "golden_hashes.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("name", sorted(_hashed()))
def test_golden_hashes(name):
    g = _hashed()[name]
    gen = g["generator"]
    left, right, _ = A.synth_pair(gen["kind"], gen["w"], gen["h"], gen["max_disp"], gen["seed"])
    assert hex(O.fnv1a64(left)) == g["input_fnv"]["left"] and hex(O.fnv1a64(right)) == g["input_fnv"]["right"]
    o = O.pipeline(O.cfg_from(g["cfg"]), left, right, intermediates=True)
    o["d_before_refine"] = o["disp_wta"]
    got = {k: hex(O.fnv1a64(o[k])) for k in g["fnv"]}
    assert got == g["fnv"]


# ----------------------------------------- closed forms and brute force -----
def test_census_definition_bruteforce():
    """Symmetric census bit b: I(p+o_b) > I(p-o_b) over the half window
    dy<0 || (dy==0 && dx<0) in raster order, clamp-to-edge sampling."""
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (13, 17), dtype=np.uint8)
    g, gx, gy, cen = O.aux(O.cfg(1, 8), img)
    h, w = img.shape
    offs = [(dx, dy) for dy in range(-3, 4) for dx in range(-4, 5) if dy < 0 or (dy == 0 and dx < 0)]
    assert len(offs) == 31

    def s(x, y):
        return int(img[min(max(y, 0), h - 1), min(max(x, 0), w - 1)])
    for y in range(h):
        for x in range(w):
            bits = 0
            for b, (dx, dy) in enumerate(offs):
                if s(x + dx, y + dy) > s(x - dx, y - dy):
                    bits |= 1 << b
            assert cen[y, x] == bits
            assert gx[y, x] == abs(s(x + 1, y) - s(x - 1, y))
            assert gy[y, x] == abs(s(x, y + 1) - s(x, y - 1))


def test_cost_is_monotone_and_zero_on_identity():
    """Identical left/right at d=0 gives cost 0; invalid outside [lo,hi) and x-d<0."""
    rng = np.random.default_rng(4)
    img = rng.integers(0, 256, (10, 20), dtype=np.uint8)
    c = O.cfg(1, 6)
    L = O.aux(c, img)
    dmin, dmax = O.range_init(c, 20, 10)
    cost = O.cost_volume(c, L, L, dmin, dmax)
    assert (cost[:, :, 0] == 0).all()
    for x in range(20):
        for d in range(6):
            assert (cost[:, x, d] == 65535).all() == (d > x)


def test_cross_windowed_mean_bruteforce():
    """Cross H: integer mean of valid values over [x-L, x+R]; V likewise."""
    rng = np.random.default_rng(5)
    h, w, D = 9, 23, 5
    c = O.cfg(1, D, max_arm_length=4)
    gray = rng.integers(0, 40, (h, w), dtype=np.uint8)
    cost = rng.integers(0, 256, (h, w, D)).astype(np.uint16)
    cost[rng.random((h, w, D)) < 0.2] = 65535
    dmin, dmax = O.range_init(c, w, h)
    La, Ra, Ua, Da = O.cross_arms(c, gray)
    tmp = O.cross_h(c, dmin, dmax, La, Ra, cost)
    for y in range(h):
        for x in range(w):
            for d in range(D):
                if not (dmin[y, x] <= d < dmax[y, x]):
                    assert tmp[y, x, d] == 65535
                    continue
                vals = [int(cost[y, xx, d]) for xx in range(x - La[y, x], x + Ra[y, x] + 1)
                        if dmin[y, xx] <= d < dmax[y, xx] and cost[y, xx, d] != 65535]
                assert tmp[y, x, d] == (sum(vals) // len(vals) if vals else 65535)
    out = O.cross_v(c, dmin, dmax, Ua, Da, tmp)
    for y in range(h):
        for x in range(w):
            for d in range(D):
                if not (dmin[y, x] <= d < dmax[y, x]):
                    continue
                vals = [int(tmp[yy, x, d]) for yy in range(y - Ua[y, x], y + Da[y, x] + 1)
                        if dmin[yy, x] <= d < dmax[yy, x] and tmp[yy, x, d] != 65535]
                assert out[y, x, d] == (sum(vals) // len(vals) if vals else 65535)
    # arms: |I - I0| <= tau run, capped
    for y in range(h):
        for x in range(w):
            l = 0
            while l < 4 and x - l - 1 >= 0 and abs(int(gray[y, x - l - 1]) - int(gray[y, x])) <= 12:
                l += 1
            assert La[y, x] == l


def test_sgm_path_order_independence_and_sentinels():
    """Path sums are integer adds: any path order gives the same S; an invalid
    C' is invalid in S; L_r - C is within [0, P2max] for every valid state."""
    rng = np.random.default_rng(6)
    h, w, D = 11, 14, 9
    c = O.cfg(1, D, paths=8)
    gray = rng.integers(0, 256, (h, w), dtype=np.uint8)
    cost = rng.integers(0, 200, (h, w, D)).astype(np.uint16)
    cost[rng.random((h, w, D)) < 0.15] = 65535
    full = O.sgm_optimize(c, gray, cost)
    dirs = [(1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (1, -1), (-1, 1), (-1, -1)]
    acc = np.zeros((h, w, D), np.uint32)
    for dx, dy in reversed(dirs):
        acc = O.sgm_path(c, gray, cost, acc, dx, dy)
    np.testing.assert_array_equal(acc, full)
    assert ((cost == 65535) == (full == 0xFFFFFFFF)).all()
    one = O.sgm_path(c, gray, cost, np.zeros((h, w, D), np.uint32), 1, 0)
    v = cost != 65535
    diff = one[v].astype(np.int64) - cost[v]
    assert diff.min() >= 0 and diff.max() <= 120


def test_sgm_first_pixel_and_restart():
    """A scan start (and a pixel after an all-invalid pixel) takes L = C."""
    c = O.cfg(1, 4)
    gray = np.zeros((1, 3), np.uint8)
    cost = np.array([[[5, 1, 7, 3], [65535] * 4, [9, 2, 4, 8]]], np.uint16)
    acc = O.sgm_path(c, gray, cost, np.zeros((1, 3, 4), np.uint32), 1, 0)
    assert list(acc[0, 0]) == [5, 1, 7, 3]
    assert list(acc[0, 2]) == [9, 2, 4, 8]


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("w,h,D,mode,over", [
    (1, 1, 4, 1, {}), (2, 3, 4, 1, {}), (9, 7, 16, 1, {}), (5, 40, 8, 0, {}), (40, 5, 64, 1, {}),
    (30, 20, 8, 1, {"min_disparity": 3}), (25, 18, 10, 1, {"min_disparity": -6}),
    (50, 30, 20, 1, {"paths": 8, "agg_iterations": 3}), (33, 21, 15, 1, {"census_type": 0, "paths": 8}),
    # HighQuality: prior ranges (enough supports, too few, the case-B budget), refiner
    (64, 48, 16, 2, {}), (96, 64, 24, 2, {"census_type": 0}),
    (120, 80, 32, 2, {"prior_max_supports": 10, "prior_min_supports": 5}),
    (40, 30, 8, 2, {"prior_min_supports": 100000}),
    # path sums past the u16 bound, wide median windows
    (36, 24, 12, 1, {"cost_max": 16383}), (30, 22, 10, 1, {"paths": 8, "cost_max": 9000}),
    (34, 26, 10, 1, {"post_median_radius": 5}), (29, 31, 8, 0, {"post_median_radius": 9}),
])
def test_oracle_matches_reference_build(w, h, D, mode, over):
    """Edge sizes (1x1, thin, D > width, shifted / negative ranges) against
    the reference classes compiled from their own sources."""
    rng = np.random.default_rng(w * 100 + h)
    left = rng.integers(0, 256, (h, w), dtype=np.uint8)
    right = np.roll(left, -2, axis=1)
    c = O.cfg(mode, over.get("min_disparity", 0) + D, **over)
    rs = O.ref_stages(c, left, right)
    rp = O.ref_pipeline(c, left, right)
    o = O.pipeline(c, left, right, intermediates=True)
    for k in ("cost", "cost_agg", "sgm", "disp_wta", "best", "second"):
        np.testing.assert_array_equal(o[k], rs[k], err_msg=k)
    for k in ("disparity", "confidence", "invalid_reason", "reliable", "disp_right", "dmin", "dmax"):
        np.testing.assert_array_equal(o[k], rp[k], err_msg=k)
    np.testing.assert_array_equal(o["disp_wta"], rp["d_before_refine"], err_msg="d_before_refine")


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("shift,maxd,over", [
    (-2, 16, {"min_disparity": -8, "prior_search_radius": 1}),
    (2, 16, {"min_disparity": -8}),
    (4, 0, {"min_disparity": -16, "refine_random_radius": 3}),
])
def test_oracle_hq_negative_ranges_match_reference_build(shift, maxd, over):
    """HighQuality with ranges reaching below 0: a scene at positive and at
    negative disparity, and a range that ends at 0 (every pixel's range lies
    below 0, so the refiner's clamped candidates would be negative)."""
    rng = np.random.default_rng(17 + shift)
    left = rng.integers(0, 256, (48, 80), dtype=np.uint8)
    right = np.roll(left, shift, axis=1)
    c = O.cfg(2, maxd, **over)
    rp = O.ref_pipeline(c, left, right)
    o = O.pipeline(c, left, right, intermediates=True)
    for k in ("disparity", "confidence", "invalid_reason", "reliable", "disp_right", "dmin", "dmax"):
        np.testing.assert_array_equal(o[k], rp[k], err_msg=k)
    np.testing.assert_array_equal(o["disp_wta"], rp["d_before_refine"], err_msg="d_before_refine")
