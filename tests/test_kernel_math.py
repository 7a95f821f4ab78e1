"""Integer identities the gfx950 kernels rely on, checked exhaustively on the CPU.

These pin the arithmetic shortcuts of the device code (magic-number division,
u8 <-> u16 sentinel decoding, packed prefix differences) independently of the
oracle: each test states the closed form and compares it with the shortcut
over the whole input domain the kernels can reach.
"""
import numpy as np

K_MAX_ARM = 62                    # apg-sgm_amd/csrc/kernels.hpp kMaxArm
WIN_MAX = 2 * K_MAX_ARM + 1       # longest Cross window (rows or columns)


def umulhi(a, b):
    return ((a.astype(np.uint64) * np.uint64(b)) >> np.uint64(32)).astype(np.uint64)


def test_cross_v_doubled_magic_division_is_exact():
    # cross_v3_kernel: floor(s / n) = umulhi(2s, ceil(2^31 / n)) for window
    # sums s < WIN_MAX * 255 (u8 entries, invalid 255 included) and n <= WIN_MAX
    s = np.arange(WIN_MAX * 255 + 1, dtype=np.uint64)
    for n in range(1, WIN_MAX + 1):
        m = (0x80000000 + n - 1) // n
        assert m < 2**32
        got = umulhi(2 * s, m)
        np.testing.assert_array_equal(got, s // np.uint64(n), err_msg=f"n={n}")


def test_cross_div_small_magic_is_exact_at_the_extremes():
    # div_small: floor(v / c) = umulhi(v, ceil(2^32 / c)) for c >= 2 and
    # v <= WIN_MAX * 65535 (u16 windows); the error term grows with v, so the
    # top of the range and the values just below each multiple are the worst
    vmax = WIN_MAX * 65535
    for c in range(2, WIN_MAX + 1):
        m = (2**32 + c - 1) // c
        top = np.arange(vmax - 4 * c, vmax + 1, dtype=np.uint64)
        near = np.arange(c - 1, vmax + 1, c, dtype=np.uint64)   # v = k*c - 1
        for v in (top, near, near[:-1] + np.uint64(1)):
            np.testing.assert_array_equal(umulhi(v, m), v // np.uint64(c), err_msg=f"c={c}")


def test_u8_sentinel_decode():
    # sgm_lines_kernel decode_c: per u16 half, x + ((x + 1) >> 8) * 0xFF00
    # maps 0..254 to themselves and 255 (u8 invalid) to 65535
    b = np.arange(256, dtype=np.uint32)
    got = (b + ((b + 1) >> 8) * 0xFF00) & 0xFFFF
    want = np.where(b == 255, 0xFFFF, b)
    np.testing.assert_array_equal(got, want)


def test_packed_prefix_difference_is_exact_modulo_2_16():
    # cross_v3_kernel keeps u16 prefix sums that wrap; a window difference of
    # at most WIN_MAX entries of <= 255 is < 2^16 and survives the wrap
    rng = np.random.default_rng(7)
    vals = rng.integers(0, 256, size=20000).astype(np.uint64)
    pref = np.concatenate([np.zeros(1, np.uint64), np.cumsum(vals)])
    mask = np.uint64(0xFFFF)
    wrapped = pref & mask
    for win in (1, 2, 35, WIN_MAX):
        exact = pref[win:] - pref[:-win]
        diff = (wrapped[win:] - wrapped[:-win]) & mask
        np.testing.assert_array_equal(diff, exact)
    assert WIN_MAX * 255 * 2 < 2**16   # the doubled sum also fits a u16 half


def test_prior_sort_leaves_ties_where_std_sort_does(tmp_path):
    # k_prior.hip sums support weights in std::sort order (key-only
    # comparator), so its device sort must reproduce libstdc++'s placement of
    # equal keys; tests/cpp/prior_sort_main.cpp compares the two over random,
    # structured and adversarial (heapsort-branch) inputs.
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / "prior_sort_main")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(root, "apg-sgm_amd", "csrc"),
                    os.path.join(root, "tests", "cpp", "prior_sort_main.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout + out.stderr
