"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) and, when it was
built, of the reference-source harness (oracle/_ref/libapgref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker, never as the product.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libapgref.so")


class Cfg(C.Structure):
    """orc_cfg (oracle/oracle.h); same field order as the product's params."""
    _fields_ = [
        ("mode", C.c_int), ("min_disparity", C.c_int), ("max_disparity", C.c_int),
        ("census_type", C.c_int),
        ("lambda_census", C.c_float), ("lambda_ad", C.c_float), ("lambda_grad", C.c_float),
        ("eta_ad", C.c_float), ("mu_grad", C.c_float),
        ("use_ad", C.c_int), ("use_grad", C.c_int), ("cost_max", C.c_int),
        ("agg_enable", C.c_int), ("max_arm_length", C.c_int), ("color_threshold", C.c_int),
        ("agg_iterations", C.c_int),
        ("paths", C.c_int), ("P1", C.c_int), ("P2_base", C.c_int), ("adaptive_penalties", C.c_int),
        ("grad_t1", C.c_int), ("grad_t2", C.c_int), ("p2_alpha", C.c_float), ("piecewise_p2", C.c_int),
        ("conf_uniqueness_ratio", C.c_float), ("conf_lr_max_diff", C.c_int),
        ("conf_min_texture", C.c_float), ("conf_reliable_threshold", C.c_float),
        ("post_subpixel", C.c_int), ("post_lr_check", C.c_int), ("post_lr_max_diff", C.c_int),
        ("post_median_radius", C.c_int), ("post_max_fill_gap", C.c_int),
        ("prior_enable", C.c_int), ("prior_uniqueness_ratio", C.c_float), ("prior_texture_threshold", C.c_int),
        ("prior_lr_max_diff", C.c_int), ("prior_search_radius", C.c_int), ("prior_min_supports", C.c_int),
        ("prior_max_supports", C.c_int),
        ("refine_enable", C.c_int), ("refine_iterations", C.c_int), ("refine_random_radius", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Inter(C.Structure):
    _fields_ = [("cost", C.c_void_p), ("cost_agg", C.c_void_p), ("sgm", C.c_void_p), ("disp_right", C.c_void_p),
                ("disp_wta", C.c_void_p), ("best", C.c_void_p), ("second", C.c_void_p),
                ("dmin", C.c_void_p), ("dmax", C.c_void_p), ("d_prior", C.c_void_p),
                ("after_refine", C.c_void_p)]


class Support(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("disparity", C.c_float), ("confidence", C.c_float)]


_orc = None
_ref = None


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def oracle():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            raise FileNotFoundError(f"{ORACLE_SO} missing: run `make -C oracle`")
        _orc = C.CDLL(ORACLE_SO)
        _orc.orc_fnv1a64.restype = C.c_uint64
        _orc.orc_fnv1a64.argtypes = [C.c_void_p, C.c_size_t]
        _orc.orc_subpixel_legacy.restype = C.c_float
        _orc.orc_subpixel_legacy.argtypes = [C.c_float] * 3
        _orc.orc_adaptive_p2.restype = C.c_int
        _orc.orc_adaptive_p2.argtypes = [C.POINTER(Cfg), C.c_int]
    return _orc


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise FileNotFoundError(f"{REF_SO} missing: run `make -C oracle ref` where the reference sources exist")
        _ref = C.CDLL(REF_SO)
        _ref.ref_time_pipeline.restype = C.c_double
        _ref.ref_max_threads.restype = C.c_int
    return _ref


def cfg(mode=1, max_disparity=128, **over) -> Cfg:
    c = Cfg()
    oracle().orc_config_from_mode(mode, max_disparity, C.byref(c))
    for k, v in over.items():
        setattr(c, k, v)
    return c


def cfg_from(d: dict) -> Cfg:
    """Cfg from a dict; fields it lacks keep the preset of its mode."""
    c = cfg(d.get("mode", 1), d.get("max_disparity", 128))
    for k, _ in Cfg._fields_:
        if k in d:
            setattr(c, k, d[k])
    return c


def set_threads(n):
    oracle().orc_set_threads(int(n))


def fnv1a64(arr: np.ndarray) -> int:
    a = np.ascontiguousarray(arr)
    return int(oracle().orc_fnv1a64(_p(a), a.nbytes))


def aux(c: Cfg, img: np.ndarray):
    """gray, gx, gy, census (u32 or u64) of one image (H x W [x 3] u8)."""
    img = np.ascontiguousarray(img)
    h, w = img.shape[:2]
    ch = 1 if img.ndim == 2 else img.shape[2]
    g = np.empty((h, w), np.uint8); gx = np.empty_like(g); gy = np.empty_like(g)
    o = oracle()
    o.orc_gray_grad(_p(img), w, h, w * ch, ch, _p(g), _p(gx), _p(gy))
    if c.census_type == 1:
        cen = np.empty((h, w), np.uint32)
        o.orc_census_sym(_p(g), w, h, _p(cen))
    else:
        cen = np.empty((h, w), np.uint64)
        o.orc_census64(_p(g), w, h, _p(cen))
    return g, gx, gy, cen


def range_init(c: Cfg, w, h):
    dmin = np.empty((h, w), np.int16); dmax = np.empty((h, w), np.int16)
    oracle().orc_range_init(C.byref(c), w, h, _p(dmin), _p(dmax))
    return dmin, dmax


def cost_volume(c: Cfg, L, R, dmin, dmax):
    """L, R: aux() tuples."""
    gl, gxl, gyl, cl = L
    gr, gxr, gyr, cr = R
    h, w = gl.shape
    D = c.max_disparity - c.min_disparity
    out = np.empty((h, w, D), np.uint16)
    sym = c.census_type == 1
    oracle().orc_cost_volume(C.byref(c), w, h, _p(gl), _p(gxl), _p(gyl), _p(gr), _p(gxr), _p(gyr),
                             _p(cl) if sym else None, _p(cr) if sym else None,
                             None if sym else _p(cl), None if sym else _p(cr), _p(dmin), _p(dmax), _p(out))
    return out


def right_wta(c: Cfg, cost):
    h, w, _ = cost.shape
    out = np.empty((h, w), np.float32)
    oracle().orc_right_wta(C.byref(c), w, h, _p(np.ascontiguousarray(cost)), _p(out))
    return out


def cross_arms(c: Cfg, gray):
    h, w = gray.shape
    arms = [np.empty((h, w), np.int16) for _ in range(4)]
    oracle().orc_cross_arms(C.byref(c), _p(gray), w, h, *[_p(a) for a in arms])
    return arms


def cross_aggregate(c: Cfg, gray, dmin, dmax, cost):
    h, w, _ = cost.shape
    out = np.ascontiguousarray(cost).copy()
    oracle().orc_cross_aggregate(C.byref(c), _p(gray), w, h, _p(dmin), _p(dmax), _p(out))
    return out


def cross_h(c: Cfg, dmin, dmax, La, Ra, cost):
    h, w, D = cost.shape
    out = np.empty_like(cost)
    oracle().orc_cross_h(C.byref(c), w, h, _p(dmin), _p(dmax), _p(La), _p(Ra), _p(np.ascontiguousarray(cost)), _p(out))
    return out


def cross_v(c: Cfg, dmin, dmax, Ua, Da, tmp):
    h, w, D = tmp.shape
    out = np.empty_like(tmp)
    oracle().orc_cross_v(C.byref(c), w, h, _p(dmin), _p(dmax), _p(Ua), _p(Da), _p(np.ascontiguousarray(tmp)), _p(out))
    return out


def sgm_optimize(c: Cfg, gray, cost):
    h, w, D = cost.shape
    acc = np.empty((h, w, D), np.uint32)
    oracle().orc_sgm_optimize(C.byref(c), _p(gray), w, h, _p(np.ascontiguousarray(cost)), _p(acc))
    return acc


def sgm_path(c: Cfg, gray, cost, acc, dx, dy):
    h, w, D = cost.shape
    acc = np.ascontiguousarray(acc).copy()
    oracle().orc_sgm_path(C.byref(c), _p(gray), w, h, D, _p(np.ascontiguousarray(cost)), _p(acc), dx, dy)
    return acc


def wta(c: Cfg, vol, dmin, dmax):
    h, w, _ = vol.shape
    disp = np.empty((h, w), np.float32); best = np.empty_like(disp); second = np.empty_like(disp)
    vol = np.ascontiguousarray(vol)
    fn = oracle().orc_wta_u32 if vol.dtype == np.uint32 else oracle().orc_wta_u16
    fn(C.byref(c), w, h, _p(vol), _p(dmin), _p(dmax), _p(disp), _p(best), _p(second))
    return disp, best, second


def subpixel_legacy(cm, c0, cp):
    return float(oracle().orc_subpixel_legacy(cm, c0, cp))


def pipeline(c: Cfg, left, right, intermediates=False):
    """Full dense pipeline -> dict of outputs (and intermediates)."""
    left = np.ascontiguousarray(left); right = np.ascontiguousarray(right)
    h, w = left.shape[:2]
    ch = 1 if left.ndim == 2 else left.shape[2]
    D = c.max_disparity - c.min_disparity
    out = {"disparity": np.empty((h, w), np.float32), "confidence": np.empty((h, w), np.float32),
           "invalid_reason": np.empty((h, w), np.uint8), "reliable": np.empty((h, w), np.uint8)}
    inter = None
    if intermediates:
        out.update({"cost": np.empty((h, w, D), np.uint16), "cost_agg": np.empty((h, w, D), np.uint16),
                    "sgm": np.empty((h, w, D), np.uint32), "disp_right": np.empty((h, w), np.float32),
                    "disp_wta": np.empty((h, w), np.float32), "best": np.empty((h, w), np.float32),
                    "second": np.empty((h, w), np.float32), "dmin": np.empty((h, w), np.int16),
                    "dmax": np.empty((h, w), np.int16), "d_prior": np.empty((h, w), np.float32),
                    "after_refine": np.empty((h, w), np.float32)})
        inter = Inter(*[out[k].ctypes.data for k in ("cost", "cost_agg", "sgm", "disp_right", "disp_wta", "best",
                                                     "second", "dmin", "dmax", "d_prior", "after_refine")])
    rc = oracle().orc_pipeline(C.byref(c), _p(left), _p(right), w, h, w * ch, ch, _p(out["disparity"]),
                               _p(out["confidence"]), _p(out["invalid_reason"]), _p(out["reliable"]),
                               C.byref(inter) if inter is not None else None)
    if rc:
        raise RuntimeError(f"oracle pipeline status {rc}")
    return out


# ---------------------------------------------------------------- reference --

def ref_stages(c: Cfg, left, right):
    """Stage outputs of the reference classes (dense): C, C', S, WTA disp/best/second, census, gray, range."""
    left = np.ascontiguousarray(left); right = np.ascontiguousarray(right)
    h, w = left.shape[:2]
    ch = 1 if left.ndim == 2 else left.shape[2]
    D = c.max_disparity - c.min_disparity
    o = {"cost": np.empty((h, w, D), np.uint16), "cost_agg": np.empty((h, w, D), np.uint16),
         "sgm": np.empty((h, w, D), np.uint32), "disp_wta": np.empty((h, w), np.float32),
         "best": np.empty((h, w), np.float32), "second": np.empty((h, w), np.float32),
         "census": np.empty((h, w), np.uint32), "gray": np.empty((h, w), np.uint8),
         "dmin": np.empty((h, w), np.int16), "dmax": np.empty((h, w), np.int16)}
    ref().ref_stages(C.byref(c), _p(left), _p(right), w, h, w * ch, ch, *[_p(o[k]) for k in (
        "cost", "cost_agg", "sgm", "disp_wta", "best", "second")],
        _p(o["census"]) if c.census_type == 1 else None, _p(o["gray"]), _p(o["dmin"]), _p(o["dmax"]))
    if c.census_type != 1:
        del o["census"]
    return o


def ref_pipeline(c: Cfg, left, right, backend=0):
    left = np.ascontiguousarray(left); right = np.ascontiguousarray(right)
    h, w = left.shape[:2]
    ch = 1 if left.ndim == 2 else left.shape[2]
    o = {"disparity": np.empty((h, w), np.float32), "confidence": np.empty((h, w), np.float32),
         "invalid_reason": np.empty((h, w), np.uint8), "reliable": np.empty((h, w), np.uint8),
         "disp_right": np.empty((h, w), np.float32), "d_before_refine": np.empty((h, w), np.float32),
         "dmin": np.empty((h, w), np.int16), "dmax": np.empty((h, w), np.int16)}
    rc = ref().ref_pipeline(C.byref(c), backend, _p(left), _p(right), w, h, w * ch, ch,
                            *[_p(o[k]) for k in ("disparity", "confidence", "invalid_reason", "reliable",
                                                 "disp_right", "d_before_refine", "dmin", "dmax")])
    if rc:
        raise RuntimeError(f"reference pipeline status {rc}")
    return o


def ref_sgm_path(c: Cfg, gray, cost, acc, dx, dy):
    h, w, D = cost.shape
    acc = np.ascontiguousarray(acc).copy()
    ref().ref_sgm_path(C.byref(c), _p(gray), w, h, D, _p(np.ascontiguousarray(cost)), _p(acc), dx, dy)
    return acc


def ref_time_pipeline(c: Cfg, left, right, threads=0, warmup=1, reps=1, backend=1):
    left = np.ascontiguousarray(left); right = np.ascontiguousarray(right)
    h, w = left.shape
    r = ref()
    r.ref_time_pipeline.argtypes = [C.POINTER(Cfg), C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                    C.c_int, C.c_int]
    return float(r.ref_time_pipeline(C.byref(c), backend, _p(left), _p(right), w, h, threads, warmup, reps))


def ref_max_threads():
    return int(ref().ref_max_threads())


def uniform_draws(stream: np.ndarray, r: int, ndraw: int):
    """std::uniform_int_distribution<int>(-r, r) draws over a replayed engine
    stream (libstdc++'s own algorithm): (draws, outputs consumed)."""
    o = oracle()
    s = np.ascontiguousarray(stream, dtype=np.uint32)
    out = np.empty(ndraw, np.int32)
    used = C.c_size_t()
    o.orc_uniform_draws.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_void_p, C.POINTER(C.c_size_t)]
    o.orc_uniform_draws(_p(s), s.size, r, ndraw, _p(out), C.byref(used))
    return out, used.value


def prior_supports(c: Cfg, gx, gy, census_l, census_r):
    """extract_supports -> list of (x, y, disparity, confidence)."""
    h, w = gx.shape
    cap = max(1, c.prior_max_supports)
    buf = (Support * cap)()
    sym = census_l.dtype == np.uint32
    n = oracle().orc_prior_supports(C.byref(c), w, h, _p(gx), _p(gy), _p(census_l) if sym else None,
                                    _p(census_r) if sym else None, None if sym else _p(census_l),
                                    None if sym else _p(census_r), buf, cap)
    return [(buf[i].x, buf[i].y, buf[i].disparity, buf[i].confidence) for i in range(min(n, cap))]


def prior_estimate(c: Cfg, gx, gy, census_l, census_r):
    """estimate -> (dmin, dmax, d_prior, support count)."""
    h, w = gx.shape
    dmin = np.empty((h, w), np.int16); dmax = np.empty((h, w), np.int16); dp = np.empty((h, w), np.float32)
    sym = census_l.dtype == np.uint32
    n = oracle().orc_prior_estimate(C.byref(c), w, h, _p(gx), _p(gy), _p(census_l) if sym else None,
                                    _p(census_r) if sym else None, None if sym else _p(census_l),
                                    None if sym else _p(census_r), _p(dmin), _p(dmax), _p(dp))
    return dmin, dmax, dp, int(n)
