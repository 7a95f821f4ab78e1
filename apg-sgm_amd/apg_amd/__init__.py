"""ctypes bindings of libapg_sgm_amd.so (the gfx950 build of the apg-sgm
dense-stereo hot path).

Only the in-tree shared library is loaded; there is no fallback.  When the
library or a GPU is missing every compute call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libapg_sgm_amd.so")

STATUS = {
    0: "Ok", 1: "EmptyInput", 2: "SizeMismatch", 3: "InvalidDisparityRange", 4: "InvalidConfig",
    5: "UnsupportedBackend", 6: "ResourceLimitExceeded", 7: "AllocationFailure", 8: "InternalError",
}


class Params(C.Structure):
    """apg_hip_params (include/apg_sgm_hip.h)."""
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


class Timing(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "aux_ms", "prior_ms", "cost_ms", "right_wta_ms", "cross_ms", "sgm_ms", "wta_ms",
        "confidence_ms", "refine_ms", "post_ms", "total_ms")]


class Intermediates(C.Structure):
    _fields_ = [("d_cost", C.c_void_p), ("d_cost_agg", C.c_void_p), ("d_sum32", C.c_void_p),
                ("d_disp_right", C.c_void_p), ("d_disp_wta", C.c_void_p), ("d_best", C.c_void_p),
                ("d_second", C.c_void_p), ("d_dmin", C.c_void_p), ("d_dmax", C.c_void_p), ("d_prior", C.c_void_p),
                ("d_after_refine", C.c_void_p)]


class Tuning(C.Structure):
    """apg_hip_tuning: launch shapes of the fused schedule (0 = built-in)."""
    _fields_ = [("cost_segment", C.c_int), ("sgm_dual", C.c_int), ("cross_v_waves", C.c_int)]

    def as_tuple(self):
        return (self.cost_segment, self.sgm_dual, self.cross_v_waves)


class Pair(C.Structure):
    """apg_hip_pair: one stereo pair of a batch (device pointers)."""
    _fields_ = [("params", Params), ("d_left", C.c_void_p), ("d_right", C.c_void_p),
                ("width", C.c_int), ("height", C.c_int), ("stride", C.c_int), ("channels", C.c_int),
                ("d_disparity", C.c_void_p), ("d_confidence", C.c_void_p),
                ("d_invalid_reason", C.c_void_p), ("d_reliable", C.c_void_p)]


class Request(C.Structure):
    """apg_hip_request: one serving request (host pointers)."""
    _fields_ = [("params", Params), ("left", C.c_void_p), ("right", C.c_void_p),
                ("width", C.c_int), ("height", C.c_int), ("stride", C.c_int), ("channels", C.c_int),
                ("disparity", C.c_void_p), ("confidence", C.c_void_p),
                ("invalid_reason", C.c_void_p), ("reliable", C.c_void_p)]


BATCH_GRAPH = 1
BATCH_PER_PAIR = 2


class ApgError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"{STATUS.get(code, code)} ({code}): {msg}")
        self.code = code


_lib = None


def load():
    """Load the in-tree library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    lib = C.CDLL(LIB_PATH)
    vp, i, sz = C.c_void_p, C.c_int, C.c_size_t
    P = C.POINTER(Params)
    sigs = {
        "apg_hip_params_from_mode": (None, [i, i, P]),
        "apg_hip_validate_params": (i, [P]),
        "apg_hip_ctx_create": (i, [i, C.POINTER(vp)]),
        "apg_hip_ctx_destroy": (None, [vp]),
        "apg_hip_synchronize": (i, [vp]),
        "apg_hip_stream": (vp, [vp]),
        "apg_hip_workspace_bytes": (sz, [vp]),
        "apg_hip_release_workspace": (None, [vp]),
        "apg_hip_enable_stage_timing": (None, [vp, i]),
        "apg_hip_last_timing": (i, [vp, C.POINTER(Timing)]),
        "apg_hip_last_error": (C.c_char_p, [vp]),
        "apg_hip_malloc": (i, [vp, sz, C.POINTER(vp)]),
        "apg_hip_free": (i, [vp, vp]),
        "apg_hip_upload": (i, [vp, vp, vp, sz]),
        "apg_hip_download": (i, [vp, vp, vp, sz]),
        "apg_hip_memset": (i, [vp, vp, i, sz]),
        "apg_hip_aux": (i, [vp, P, vp, i, i, i, i, vp, vp, vp, vp, vp]),
        "apg_hip_range_init": (i, [vp, P, i, i, vp, vp]),
        "apg_hip_cost_volume": (i, [vp, P, i, i] + [vp] * 13),
        "apg_hip_right_wta": (i, [vp, P, i, i, vp, vp]),
        "apg_hip_cross_arms": (i, [vp, P, vp, i, i, vp]),
        "apg_hip_cross_aggregate": (i, [vp, P, i, i, vp, vp, vp, vp, vp]),
        "apg_hip_sgm_optimize": (i, [vp, P, i, i, vp, vp, vp, vp]),
        "apg_hip_sgm_path": (i, [vp, P, i, i, i, vp, vp, vp, i, i]),
        "apg_hip_wta": (i, [vp, P, i, i, vp, i, vp, vp, vp, vp, vp]),
        "apg_hip_confidence": (i, [vp, P, i, i] + [vp] * 9),
        "apg_hip_lr_check": (i, [vp, P, i, i, vp, vp, vp]),
        "apg_hip_fill_holes": (i, [vp, P, i, i, vp, vp, vp, vp, vp]),
        "apg_hip_median": (i, [vp, P, i, i, vp, vp]),
        "apg_hip_pipeline": (i, [vp, P, vp, vp, i, i, i, i, vp, vp, vp, vp, C.POINTER(Intermediates)]),
        "apg_hip_set_schedule": (None, [vp, i]),
        "apg_hip_bench_pipeline": (i, [vp, P, vp, vp, i, i, i, i, vp, vp, vp, vp, i, i, C.POINTER(C.c_double),
                                       C.POINTER(Timing)]),
        "apg_hip_pipeline_workspace_bytes": (sz, [P, i, i]),
        "apg_hip_bench_pipeline_steps": (i, [vp, P, vp, vp, i, i, i, i, vp, vp, vp, vp, i, i,
                                             C.POINTER(C.c_float)]),
        "apg_hip_pipeline_static": (i, [P, vp, vp, i, i, i, i, vp, vp, vp, vp, vp]),
        "apg_hip_pipeline_static_bytes": (sz, [P, i, i]),
        "apg_hip_static_scratch_bytes": (sz, []),
        "apg_hip_static_last_error": (C.c_char_p, []),
        "apg_hip_prior_estimate": (i, [vp, P, i, i] + [vp] * 9 + [C.POINTER(i)]),
        "apg_hip_refine": (i, [vp, P, i, i] + [vp] * 14),
        "apg_hip_set_tuning": (i, [vp, C.POINTER(Tuning)]),
        "apg_hip_set_refine_wait": (i, [vp, C.c_longlong]),
        "apg_hip_get_tuning": (i, [vp, C.POINTER(Tuning)]),
        "apg_hip_batch_create": (i, [i, i, C.POINTER(vp)]),
        "apg_hip_batch_destroy": (None, [vp]),
        "apg_hip_batch_stream": (vp, [vp]),
        "apg_hip_batch_synchronize": (i, [vp]),
        "apg_hip_batch_workspace_bytes": (sz, [vp]),
        "apg_hip_batch_last_error": (C.c_char_p, [vp]),
        "apg_hip_batch_run": (i, [vp, C.POINTER(Pair), i, i]),
        "apg_hip_batch_group_stats": (i, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "apg_hip_batch_set_grouping": (i, [vp, C.c_size_t, C.c_size_t]),
        "apg_hip_batch_autotune": (i, [vp, C.POINTER(Pair), i, C.c_char_p]),
        "apg_hip_batch_tuning_for": (i, [vp, P, i, i, C.POINTER(Tuning)]),
        "apg_hip_server_create": (i, [i, i, sz, C.POINTER(vp)]),
        "apg_hip_server_destroy": (None, [vp]),
        "apg_hip_server_last_error": (C.c_char_p, [vp]),
        "apg_hip_server_submit": (i, [vp, C.POINTER(Request), C.c_uint64]),
        "apg_hip_server_poll": (i, [vp, C.POINTER(C.c_uint64), C.POINTER(i), i]),
        "apg_hip_server_pending": (i, [vp]),
        "apg_hip_server_resident_bytes": (sz, [vp]),
        "apg_hip_server_replay": (i, [vp, C.POINTER(Request), C.POINTER(C.c_double), i,
                                      C.POINTER(C.c_double)]),
        "apg_synth_pair": (i, [i, i, i, i, C.c_uint64, vp, vp, vp]),
        "apg_hip_version_string": (C.c_char_p, []),
    }
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def params(mode=1, max_disparity=128, **over) -> Params:
    p = Params()
    load().apg_hip_params_from_mode(mode, max_disparity, C.byref(p))
    for k, v in over.items():
        setattr(p, k, v)
    return p


def synth_pair(kind, w, h, max_disp, seed):
    """Seeded synthetic pair -> (left u8 HxW, right u8 HxW, gt f32 HxW)."""
    left = np.empty((h, w), np.uint8)
    right = np.empty((h, w), np.uint8)
    gt = np.empty((h, w), np.float32)
    rc = load().apg_synth_pair(kind, w, h, max_disp, seed, left.ctypes.data, right.ctypes.data, gt.ctypes.data)
    if rc:
        raise ApgError(rc, "apg_synth_pair")
    return left, right, gt


class DeviceArray:
    """A device allocation owned by a Context (freed with it or on del)."""

    def __init__(self, ctx, shape, dtype):
        self.ctx, self.shape, self.dtype = ctx, tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        ptr = C.c_void_p()
        ctx._check(ctx.lib.apg_hip_malloc(ctx.h, max(self.nbytes, 1), C.byref(ptr)), "malloc")
        self.ptr = ptr.value

    def upload(self, arr):
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        self.ctx._check(self.ctx.lib.apg_hip_upload(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes), "upload")
        return self

    def download(self):
        out = np.empty(self.shape, self.dtype)
        self.ctx._check(self.ctx.lib.apg_hip_download(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes), "download")
        return out

    def fill(self, byte):
        self.ctx._check(self.ctx.lib.apg_hip_memset(self.ctx.h, self.ptr, byte, self.nbytes), "memset")
        return self

    def free(self):
        if self.ptr and self.ctx.h:
            self.ctx.lib.apg_hip_free(self.ctx.h, self.ptr)
        self.ptr = None


class Context:
    """apg_hip_ctx: a HIP stream + device workspace on one GPU."""

    def __init__(self, device=0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.apg_hip_ctx_create(device, C.byref(h))
        if rc:
            raise ApgError(rc, "apg_hip_ctx_create (is a GPU visible?)")
        self.h = h.value
        self._arrays = []

    def _check(self, rc, what):
        if rc:
            msg = self.lib.apg_hip_last_error(self.h)
            raise ApgError(rc, f"{what}: {msg.decode() if msg else ''}")

    def array(self, shape, dtype):
        a = DeviceArray(self, shape, dtype)
        self._arrays.append(a)
        return a

    def to_device(self, arr):
        arr = np.ascontiguousarray(arr)
        return self.array(arr.shape, arr.dtype).upload(arr)

    def sync(self):
        self._check(self.lib.apg_hip_synchronize(self.h), "sync")

    def close(self):
        if self.h:
            for a in self._arrays:
                a.free()
            self._arrays = []
            self.lib.apg_hip_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class PipelineOutputs:
    disparity: np.ndarray
    confidence: np.ndarray
    invalid_reason: np.ndarray
    reliable: np.ndarray
    cost: np.ndarray | None = None
    cost_agg: np.ndarray | None = None
    sgm: np.ndarray | None = None
    disp_right: np.ndarray | None = None
    disp_wta: np.ndarray | None = None
    best: np.ndarray | None = None
    second: np.ndarray | None = None
    dmin: np.ndarray | None = None
    dmax: np.ndarray | None = None
    d_prior: np.ndarray | None = None
    after_refine: np.ndarray | None = None


ALL_INTERMEDIATES = ("cost", "cost_agg", "sgm", "disp_right", "disp_wta", "best", "second", "dmin", "dmax",
                     "d_prior", "after_refine")
# requesting the raw (cost) or post-Cross (cost_agg) volume selects the stage-by-stage schedule
FUSED_INTERMEDIATES = ("sgm", "disp_right", "disp_wta", "best", "second", "dmin", "dmax", "d_prior",
                       "after_refine")


def set_schedule(ctx: "Context", schedule: int):
    """0: fused schedule when eligible (default); 1: always stage by stage."""
    ctx.lib.apg_hip_set_schedule(ctx.h, schedule)


def run_pipeline(ctx: Context, p: Params, left: np.ndarray, right: np.ndarray, intermediates=False) -> PipelineOutputs:
    """Upload a pair, run apg_hip_pipeline, download results and, optionally,
    materialised intermediates: True = all of ALL_INTERMEDIATES, or an
    iterable of names."""
    h, w = left.shape[:2]
    ch = 1 if left.ndim == 2 else left.shape[2]
    D = p.max_disparity - p.min_disparity
    dl, dr = ctx.to_device(left), ctx.to_device(right)
    out_d = ctx.array((h, w), np.float32)
    out_c = ctx.array((h, w), np.float32)
    out_r = ctx.array((h, w), np.uint8)
    out_m = ctx.array((h, w), np.uint8)
    inter = None
    arrs = {}
    want = ALL_INTERMEDIATES if intermediates is True else tuple(intermediates or ())
    if want:
        shapes = {"cost": ((h, w, D), np.uint16), "cost_agg": ((h, w, D), np.uint16), "sgm": ((h, w, D), np.uint32),
                  "disp_right": ((h, w), np.float32), "disp_wta": ((h, w), np.float32), "best": ((h, w), np.float32),
                  "second": ((h, w), np.float32), "dmin": ((h, w), np.int16), "dmax": ((h, w), np.int16),
                  "d_prior": ((h, w), np.float32), "after_refine": ((h, w), np.float32)}
        arrs = {k: ctx.array(*shapes[k]) for k in want}
        inter = Intermediates(*[arrs[k].ptr if k in arrs else None for k in ALL_INTERMEDIATES])
    rc = ctx.lib.apg_hip_pipeline(ctx.h, C.byref(p), dl.ptr, dr.ptr, w, h, w * ch, ch, out_d.ptr, out_c.ptr,
                                  out_r.ptr, out_m.ptr, C.byref(inter) if inter is not None else None)
    ctx._check(rc, "apg_hip_pipeline")
    ctx.sync()
    res = PipelineOutputs(out_d.download(), out_c.download(), out_r.download(), out_m.download(),
                          **{k: v.download() for k, v in arrs.items()})
    for a in [dl, dr, out_d, out_c, out_r, out_m] + list(arrs.values()):
        a.free()
    return res


def set_tuning(ctx: Context, tuning=None):
    """Launch shapes for later pipeline calls on ctx: a Tuning, a
    (cost_segment, sgm_dual, cross_v_waves) tuple, or None for the built-in."""
    if tuning is None:
        ctx._check(ctx.lib.apg_hip_set_tuning(ctx.h, None), "set_tuning")
        return
    t = tuning if isinstance(tuning, Tuning) else Tuning(*tuning)
    ctx._check(ctx.lib.apg_hip_set_tuning(ctx.h, C.byref(t)), "set_tuning")


def workspace_bytes(p: Params, w: int, h: int) -> int:
    """Device bytes of the pipeline workspace for a (w, h) pair (host-only)."""
    return int(load().apg_hip_pipeline_workspace_bytes(C.byref(p), w, h))


class DevicePair:
    """Device inputs and outputs of one pair (allocated on ctx)."""

    def __init__(self, ctx: Context, p: Params, left: np.ndarray, right: np.ndarray):
        self.params = p
        self.h, self.w = left.shape[:2]
        self.ch = 1 if left.ndim == 2 else left.shape[2]
        self.left, self.right = ctx.to_device(left), ctx.to_device(right)
        self.disp = ctx.array((self.h, self.w), np.float32)
        self.conf = ctx.array((self.h, self.w), np.float32)
        self.reason = ctx.array((self.h, self.w), np.uint8)
        self.rel = ctx.array((self.h, self.w), np.uint8)

    def as_pair(self) -> Pair:
        return Pair(self.params, self.left.ptr, self.right.ptr, self.w, self.h, self.w * self.ch, self.ch,
                    self.disp.ptr, self.conf.ptr, self.reason.ptr, self.rel.ptr)

    def clear_outputs(self, byte=0xFF):
        for a in (self.disp, self.conf, self.reason, self.rel):
            a.fill(byte)

    def outputs(self):
        return (self.disp.download(), self.conf.download(), self.reason.download(), self.rel.download())


class Batch:
    """apg_hip_batch: many pairs over `slots` concurrent pipelines, optionally
    replayed as one hipGraph."""

    def __init__(self, device=0, slots=4):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.apg_hip_batch_create(device, slots, C.byref(h))
        if rc:
            raise ApgError(rc, "apg_hip_batch_create")
        self.h = h.value

    def _check(self, rc, what):
        if rc:
            msg = self.lib.apg_hip_batch_last_error(self.h)
            raise ApgError(rc, f"{what}: {msg.decode() if msg else ''}")

    @staticmethod
    def pair_array(pairs):
        arr = (Pair * len(pairs))()
        for i, q in enumerate(pairs):
            arr[i] = q.as_pair() if isinstance(q, DevicePair) else q
        return arr

    def run(self, pairs, graph=False, grouped=True):
        arr = pairs if isinstance(pairs, C.Array) else self.pair_array(pairs)
        flags = (BATCH_GRAPH if graph else 0) | (0 if grouped else BATCH_PER_PAIR)
        self._check(self.lib.apg_hip_batch_run(self.h, arr, len(arr), flags), "batch_run")

    def set_grouping(self, chunk_bytes=0, max_pair_states=0):
        """Group launches: workspace bytes per chunk and the largest pair
        (W*H*D) that joins them (0: the defaults)."""
        self._check(self.lib.apg_hip_batch_set_grouping(self.h, int(chunk_bytes), int(max_pair_states)),
                    "set_grouping")

    def group_stats(self):
        """(pairs in group launches, group launches, pairs run one at a time)
        of the last pair list run with group launches."""
        g, n, s = C.c_int(), C.c_int(), C.c_int()
        self._check(self.lib.apg_hip_batch_group_stats(self.h, C.byref(g), C.byref(n), C.byref(s)), "group_stats")
        return g.value, n.value, s.value

    def autotune(self, pairs, cache_path=None):
        arr = pairs if isinstance(pairs, C.Array) else self.pair_array(pairs)
        path = cache_path.encode() if cache_path else None
        self._check(self.lib.apg_hip_batch_autotune(self.h, arr, len(arr), path), "batch_autotune")

    def tuning_for(self, p: Params, w: int, h: int) -> Tuning:
        t = Tuning()
        self._check(self.lib.apg_hip_batch_tuning_for(self.h, C.byref(p), w, h, C.byref(t)), "tuning_for")
        return t

    def sync(self):
        self._check(self.lib.apg_hip_batch_synchronize(self.h), "batch_synchronize")

    def workspace_bytes(self) -> int:
        return int(self.lib.apg_hip_batch_workspace_bytes(self.h))

    def close(self):
        if self.h:
            self.lib.apg_hip_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HostRequest:
    """Host buffers of one serving request; keeps them alive with the Request."""

    def __init__(self, p: Params, left: np.ndarray, right: np.ndarray):
        self.left = np.ascontiguousarray(left)
        self.right = np.ascontiguousarray(right)
        h, w = self.left.shape[:2]
        ch = 1 if self.left.ndim == 2 else self.left.shape[2]
        self.disparity = np.full((h, w), np.nan, np.float32)
        self.confidence = np.full((h, w), np.nan, np.float32)
        self.invalid_reason = np.full((h, w), 0xFF, np.uint8)
        self.reliable = np.full((h, w), 0xFF, np.uint8)
        self.req = Request(p, self.left.ctypes.data, self.right.ctypes.data, w, h, w * ch, ch,
                           self.disparity.ctypes.data, self.confidence.ctypes.data,
                           self.invalid_reason.ctypes.data, self.reliable.ctypes.data)

    def outputs(self):
        return (self.disparity, self.confidence, self.invalid_reason, self.reliable)


class Server:
    """apg_hip_server: continuous batching of host requests under an HBM budget."""

    def __init__(self, device=0, slots=4, budget_bytes=0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.apg_hip_server_create(device, slots, budget_bytes, C.byref(h))
        if rc:
            raise ApgError(rc, "apg_hip_server_create")
        self.h = h.value

    def _check(self, rc, what):
        if rc:
            msg = self.lib.apg_hip_server_last_error(self.h)
            raise ApgError(rc, f"{what}: {msg.decode() if msg else ''}")

    def submit(self, r: HostRequest, tag: int):
        self._check(self.lib.apg_hip_server_submit(self.h, C.byref(r.req), tag), "server_submit")

    def poll(self, max_done=64):
        tags = (C.c_uint64 * max_done)()
        sts = (C.c_int * max_done)()
        k = self.lib.apg_hip_server_poll(self.h, tags, sts, max_done)
        if k < 0:
            self._check(8, "server_poll")
        return [(int(tags[i]), int(sts[i])) for i in range(k)]

    def pending(self) -> int:
        return int(self.lib.apg_hip_server_pending(self.h))

    def resident_bytes(self) -> int:
        return int(self.lib.apg_hip_server_resident_bytes(self.h))

    def replay(self, reqs, arrival_s):
        """Open-loop replay of (HostRequest, arrival second) -> latencies (s)."""
        n = len(reqs)
        arr = (Request * n)(*[r.req for r in reqs])
        arrival = (C.c_double * n)(*[float(a) for a in arrival_s])
        lat = (C.c_double * n)()
        self._check(self.lib.apg_hip_server_replay(self.h, arr, arrival, n, lat), "server_replay")
        return np.ctypeslib.as_array(lat).copy()

    def close(self):
        if self.h:
            self.lib.apg_hip_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
