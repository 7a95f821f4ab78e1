"""Closed-form and invariant checks of the oracle's HighQuality stages (the
prior and the refiner), independent of the reference build; the golden
vectors in tests/golden/ pin their exact outputs."""
import numpy as np

import pyoracle as O


def _aux_pair(left, right, c):
    g, gx, gy, cl = O.aux(c, left)
    _, _, _, cr = O.aux(c, right)
    return gx, gy, cl, cr


def test_prior_supports_recover_a_pure_shift():
    # right(x) = left(x + 8): every right-view-consistent support at x >= 16
    # must report d = 8 (census WTA over the whole range)
    rng = np.random.default_rng(21)
    left = rng.integers(0, 256, (96, 160), dtype=np.uint8)
    right = np.roll(left, -8, axis=1)
    c = O.cfg(2, 32)
    gx, gy, cl, cr = _aux_pair(left, right, c)
    sup = O.prior_supports(c, gx, gy, cl, cr)
    assert len(sup) > 100
    inside = [s for s in sup if 16 <= s[0] < 160 - 16]
    assert inside and all(s[2] == 8.0 for s in inside)
    # samples sit on even coordinates at least 2 px from the border, one cell grid of 16
    assert all(s[0] % 2 == 0 and s[1] % 2 == 0 and 2 <= s[0] < 158 and 2 <= s[1] < 94 for s in sup)
    assert all(0.0 <= s[3] <= 1.0 for s in sup)


def test_prior_ranges_contain_the_shift_and_stay_in_the_dense_range():
    rng = np.random.default_rng(22)
    left = rng.integers(0, 256, (80, 128), dtype=np.uint8)
    right = np.roll(left, -5, axis=1)
    c = O.cfg(2, 48)
    gx, gy, cl, cr = _aux_pair(left, right, c)
    dmin, dmax, dp, n = O.prior_estimate(c, gx, gy, cl, cr)
    assert n >= c.prior_min_supports
    xs = np.arange(128)[None, :].repeat(80, 0)
    # within [min_disparity, min(max_disparity, x + 1)) of the dense init
    assert (dmin >= 0).all() and (dmax <= np.minimum(48, xs + 1)).all() and (dmax >= dmin).all()
    # a confident, zero-spread prior of 5 keeps d = 5 inside and narrows the range
    ok = xs >= 5
    assert ((dmin[ok] <= 5) & (5 < dmax[ok])).all()
    assert (dmax - dmin)[:, 40:].mean() < 48
    assert np.allclose(dp[dp >= 0], 5.0)


def test_prior_off_and_too_few_supports_keep_the_dense_range():
    rng = np.random.default_rng(23)
    left = rng.integers(0, 256, (40, 64), dtype=np.uint8)
    right = np.roll(left, -3, axis=1)
    for over in ({"prior_enable": 0}, {"prior_min_supports": 10 ** 6}):
        c = O.cfg(2, 16, **over)
        gx, gy, cl, cr = _aux_pair(left, right, c)
        dmin, dmax, dp, _ = O.prior_estimate(c, gx, gy, cl, cr)
        d0, d1 = O.range_init(c, 64, 40)
        assert np.array_equal(dmin, d0) and np.array_equal(dmax, d1)
        assert (dp == -1).all()


def test_refiner_invariants():
    # reliable pixels and pixels with an empty range are untouched; results
    # are integers inside the pixel's range; zero iterations is the identity
    rng = np.random.default_rng(24)
    left = rng.integers(0, 256, (30, 50), dtype=np.uint8)
    right = np.roll(left, -4, axis=1)
    c = O.cfg(2, 16)
    o = O.pipeline(c, left, right, intermediates=True)
    c0 = O.cfg(2, 16, refine_iterations=0)
    o0 = O.pipeline(c0, left, right, intermediates=True)
    coff = O.cfg(2, 16, refine_enable=0)
    ooff = O.pipeline(coff, left, right, intermediates=True)
    assert np.array_equal(o0["disparity"], ooff["disparity"])
    assert np.array_equal(o["disp_wta"], o0["disp_wta"])   # d_before_refine does not depend on the refiner
