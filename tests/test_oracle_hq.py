"""Closed-form and invariant checks of the oracle's HighQuality stages (the
prior and the refiner), independent of the reference build; the golden
vectors in tests/golden/ pin their exact outputs."""
import numpy as np
import pytest

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


def _device_rule_draws(stream, r, ndraw):
    """The refiner kernels' rule: with s = 2r + 1, an engine output e is
    accepted iff (e * s) mod 2^32 >= 2^32 mod s; the k-th draw is the k-th
    accepted output, mapped to (e * s) >> 32 - r."""
    s = 2 * r + 1
    t = (2**32 - s) % s
    prod = stream.astype(np.uint64) * np.uint64(s)
    ok = (prod & np.uint64(0xFFFFFFFF)) >= np.uint64(t)
    pos = np.flatnonzero(ok)[:ndraw]
    vals = (prod[pos] >> np.uint64(32)).astype(np.int64) - r
    return vals, int(pos[-1]) + 1 if ndraw else 0


@pytest.mark.parametrize("r", [1, 2, 4, 9, 100])
def test_jitter_acceptance_rule_matches_libstdcxx(r):
    """libstdc++'s uniform_int_distribution over a 32-bit engine equals the
    data-independent accept-and-compact rule the device refiner uses,
    including rejected outputs (planted: values whose (e * s) mod 2^32 falls
    below 2^32 mod s, which a real mt19937 stream reaches about once in 1e9)."""
    rng = np.random.default_rng(7 + r)
    stream = rng.integers(0, 2**32, 4000, dtype=np.uint64).astype(np.uint32)
    s = 2 * r + 1
    t = (2**32 - s) % s
    assert t > 0
    # e with (e * s) mod 2^32 == 0 (< t): e = k * ceil(2^32 / s) adjusted; search a few
    rejects = [e for e in range(0, 2**32, (2**32 + s - 1) // s)][:50]
    rejects = [e for e in rejects if (e * s) % 2**32 < t]
    assert rejects
    for i, e in enumerate(rejects[:20]):
        stream[97 * i + 3] = e
        stream[97 * i + 4] = rejects[(i + 1) % len(rejects)]   # two rejections in a row
    nd = 3500
    got, used = O.uniform_draws(stream, r, nd)
    exp, exp_used = _device_rule_draws(stream, r, nd)
    assert used == exp_used
    np.testing.assert_array_equal(got, exp)
    assert got.min() >= -r and got.max() <= r


def test_jitter_acceptance_rule_on_mt19937_stream():
    """The same on the refiner's real engine: std::mt19937(12345) outputs
    (numpy's MT19937 with the reference generator's integer seeding)."""
    bg = np.random.MT19937(0)
    bg._legacy_seeding(12345)
    stream = np.array([bg.random_raw() for _ in range(5000)], dtype=np.uint64).astype(np.uint32)
    for r in (4, 2, 1):
        got, used = O.uniform_draws(stream, r, 4000)
        exp, exp_used = _device_rule_draws(stream, r, 4000)
        assert used == exp_used
        np.testing.assert_array_equal(got, exp)
