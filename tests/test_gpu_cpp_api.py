"""GPU tests of the C++ stage classes and of the session C ABI (the drop-in
boundaries), against the CPU oracle."""
import ctypes as C
import json
import os
import subprocess

import numpy as np
import pytest

import apg_amd as A
import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "apg-sgm_amd", "_build")


def _oracle_hashes(left, right, mode, dmax, dmin=0):
    c = O.cfg(mode, dmax, min_disparity=dmin)
    o = O.pipeline(c, left, right, intermediates=True)
    return {k: hex(O.fnv1a64(o[k])) for k in o}


@pytest.mark.gpu
@pytest.mark.parametrize("w,h,mode,dmax,dmin", [(96, 64, 1, 32, 0), (75, 41, 0, 20, 0), (60, 44, 1, 8, -4)])
def test_stage_classes_bit_exact(tmp_path, w, h, mode, dmax, dmin):
    left, right, _ = A.synth_pair(1, w, h, dmax - dmin, 99)
    lp, rp = tmp_path / "l.raw", tmp_path / "r.raw"
    left.tofile(lp)
    right.tofile(rp)
    print(f"stage classes {w}x{h} mode={mode} [{dmin},{dmax})", flush=True)
    out = subprocess.run([os.path.join(BUILD, "stage_classes_main"), str(lp), str(rp), str(w), str(h), str(mode),
                          str(dmax), str(dmin)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = json.loads(out.stdout)
    exp = _oracle_hashes(left, right, mode, dmax, dmin)
    for k in ("cost", "cost_agg", "sgm", "disp_wta", "best", "second", "disparity", "confidence",
              "invalid_reason", "reliable", "disp_right"):
        assert got[k] == exp[k], k
    # stage-by-stage post chain == pipeline output
    assert got["disparity_stages"] == exp["disparity"]
    # one LR path over the raw volume
    c = O.cfg(mode, dmax, min_disparity=dmin)
    g, gx, gy, cen = O.aux(c, left)
    R = O.aux(c, right)
    dmn, dmx = O.range_init(c, w, h)
    cost = O.cost_volume(c, (g, gx, gy, cen), R, dmn, dmx)
    acc = O.sgm_path(c, g, cost, np.zeros(cost.shape, np.uint32), 1, 0)
    assert got["path_lr_raw"] == hex(O.fnv1a64(acc))


@pytest.mark.gpu
def test_c_api_smoke_program(tmp_path):
    out_path = tmp_path / "disp.f32"
    r = subprocess.run([os.path.join(BUILD, "c_api_smoke"), str(out_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    w, h, ch = 64, 40, 3
    x = np.arange(w)[None, :, None]
    y = np.arange(h)[:, None, None]
    c = np.arange(ch)[None, None, :]
    left = ((x * 37 + y * 73 + x * y * 13 + c * 30) % 256).astype(np.uint8)
    right = (((x + 4) * 37 + y * 73 + (x + 4) * y * 13 + c * 30) % 256).astype(np.uint8)
    o = O.pipeline(O.cfg(1, 24), left, right)
    np.testing.assert_array_equal(np.fromfile(out_path, np.float32).reshape(h, w), o["disparity"])


class ImageView(C.Structure):
    _fields_ = [("data", C.c_void_p), ("width", C.c_int), ("height", C.c_int), ("stride", C.c_int),
                ("channels", C.c_int)]


@pytest.mark.gpu
def test_session_c_abi_matches_oracle_and_fingerprints():
    lib = A.load()
    lib.apg_session_create.restype = C.c_void_p
    lib.apg_result_create.restype = C.c_void_p
    lib.apg_result_get_disparity.restype = C.POINTER(C.c_float)
    lib.apg_result_get_confidence.restype = C.POINTER(C.c_float)
    lib.apg_result_get_invalid_reason.restype = C.POINTER(C.c_uint8)
    lib.apg_result_get_reliable_mask.restype = C.POINTER(C.c_uint8)
    lib.apg_session_get_last_error.restype = C.c_char_p
    for f in ("apg_session_process", "apg_result_get_disparity", "apg_result_get_confidence",
              "apg_result_get_invalid_reason", "apg_result_get_reliable_mask", "apg_session_destroy",
              "apg_result_destroy", "apg_session_get_last_error"):
        getattr(lib, f).argtypes = None
    s = C.c_void_p(lib.apg_session_create(1, 1))
    r = C.c_void_p(lib.apg_result_create())
    # reference fingerprint case: 33x27 checkerboard, max_disparity 16 (session default is 128)
    from test_oracle import pattern
    img = pattern(33, 27, 1, "checker")
    v = ImageView(img.ctypes.data, 33, 27, 33, 1)
    opts_s = C.c_void_p(lib.apg_session_create(1, 0))
    rc = lib.apg_session_process(opts_s, C.byref(v), C.byref(v), r)
    assert rc == 0, lib.apg_session_get_last_error(opts_s)
    ww, hh = C.c_int(), C.c_int()
    d = np.ctypeslib.as_array(lib.apg_result_get_disparity(r, C.byref(ww), C.byref(hh)), (27, 33)).copy()
    o = O.pipeline(O.cfg(1, 128), img, img)
    np.testing.assert_array_equal(d, o["disparity"])
    # a realistic pair through BoundedReuse, twice
    left, right, _ = A.synth_pair(0, 160, 96, 128, 5)
    lv = ImageView(left.ctypes.data, 160, 96, 160, 1)
    rv = ImageView(right.ctypes.data, 160, 96, 160, 1)
    o = O.pipeline(O.cfg(1, 128), left, right)
    for _ in range(2):
        assert lib.apg_session_process(s, C.byref(lv), C.byref(rv), r) == 0
        d = np.ctypeslib.as_array(lib.apg_result_get_disparity(r, C.byref(ww), C.byref(hh)), (96, 160)).copy()
        cf = np.ctypeslib.as_array(lib.apg_result_get_confidence(r, C.byref(ww), C.byref(hh)), (96, 160)).copy()
        rs = np.ctypeslib.as_array(lib.apg_result_get_invalid_reason(r, C.byref(ww), C.byref(hh)), (96, 160)).copy()
        rl = np.ctypeslib.as_array(lib.apg_result_get_reliable_mask(r, C.byref(ww), C.byref(hh)), (96, 160)).copy()
        np.testing.assert_array_equal(d, o["disparity"])
        np.testing.assert_array_equal(cf, o["confidence"])
        np.testing.assert_array_equal(rs, o["invalid_reason"])
        np.testing.assert_array_equal(rl, o["reliable"])
    lib.apg_session_destroy(s)
    lib.apg_session_destroy(opts_s)
    lib.apg_result_destroy(r)


def _pack(dense, dmin, dmax, d0):
    """Dense [h][w][D] over [d0, d0 + D) -> ragged states in raster order."""
    h, w, D = dense.shape
    out = []
    for q in range(h * w):
        y, x = divmod(q, w)
        lo, hi = int(dmin[y, x]), int(dmax[y, x])
        if hi > lo:
            out.append(dense[y, x, lo - d0:hi - d0])
    return np.concatenate(out) if out else np.empty(0, dense.dtype)


@pytest.mark.gpu
@pytest.mark.parametrize("w,h,dmax", [(96, 64, 24), (130, 70, 40)])
def test_packed_stage_classes_match_oracle(tmp_path, w, h, dmax):
    # HighQuality ranges from the prior; the oracle's dense volumes over the
    # same ranges, packed, are what the reference's packed methods compute
    left, right, _ = A.synth_pair(2, w, h, dmax, 55)
    lp, rp = tmp_path / "l.raw", tmp_path / "r.raw"
    left.tofile(lp)
    right.tofile(rp)
    print(f"packed stage classes {w}x{h} D={dmax}", flush=True)
    out = subprocess.run([os.path.join(BUILD, "packed_classes_main"), str(lp), str(rp), str(w), str(h), str(dmax),
                          str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    c = O.cfg(2, dmax)
    o = O.pipeline(c, left, right, intermediates=True)
    rd = lambda name, dt: np.fromfile(tmp_path / name, dt)  # noqa: E731
    dmn, dmx = rd("dmin.bin", np.int16).reshape(h, w), rd("dmax.bin", np.int16).reshape(h, w)
    np.testing.assert_array_equal(dmn, o["dmin"])
    np.testing.assert_array_equal(dmx, o["dmax"])
    np.testing.assert_array_equal(rd("cost.bin", np.uint16), _pack(o["cost"], dmn, dmx, 0))
    np.testing.assert_array_equal(rd("cost_agg.bin", np.uint16), _pack(o["cost_agg"], dmn, dmx, 0))
    s32 = _pack(o["sgm"], dmn, dmx, 0)
    np.testing.assert_array_equal(rd("sgm32.bin", np.uint32), s32)
    np.testing.assert_array_equal(rd("sgm16.bin", np.uint16), np.where(s32 == 0xFFFFFFFF, 0xFFFF, s32).astype(np.uint16))
    for k, f in (("disp_wta", "disp.bin"), ("disp_wta", "disp16.bin"), ("best", "best.bin"), ("second", "second.bin")):
        np.testing.assert_array_equal(rd(f, np.float32).reshape(h, w), o[k], err_msg=f)
    # one left-to-right path over the raw ragged cost into a zeroed u32 accumulator
    g, gx, gy, cen = O.aux(c, left)
    acc = O.sgm_path(c, g, o["cost"], np.zeros(o["cost"].shape, np.uint32), 1, 0)
    np.testing.assert_array_equal(rd("path.bin", np.uint32), _pack(acc, dmn, dmx, 0))


@pytest.mark.gpu
def test_client_on_reference_header_paths(tmp_path):
    # a C++ client including every header by the reference's path: session,
    # matcher with a retained workspace, HighQuality stage chain with the refiner
    w, h, D = 120, 80, 32
    left, right, _ = A.synth_pair(2, w, h, D, 31)
    lp, rp = tmp_path / "l.raw", tmp_path / "r.raw"
    left.tofile(lp)
    right.tofile(rp)
    print(f"reference-header client {w}x{h} D={D}", flush=True)
    out = subprocess.run([os.path.join(BUILD, "reference_headers_main"), str(lp), str(rp), str(w), str(h), str(D)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = json.loads(out.stdout)
    bal = O.pipeline(O.cfg(1, D), left, right)
    for k in ("disparity", "confidence", "invalid_reason", "reliable"):
        assert got[k] == hex(O.fnv1a64(bal[k])), k
    assert got["matcher_disparity"] == hex(O.fnv1a64(bal["disparity"]))
    hq = O.pipeline(O.cfg(2, D), left, right, intermediates=True)
    assert got["hq_dmin"] == hex(O.fnv1a64(hq["dmin"]))
    assert got["hq_dmax"] == hex(O.fnv1a64(hq["dmax"]))
    assert got["hq_disp_wta"] == hex(O.fnv1a64(hq["disp_wta"]))
    assert got["hq_after_refine"] == hex(O.fnv1a64(hq["after_refine"]))
    assert got["hq_supports"] > 0
