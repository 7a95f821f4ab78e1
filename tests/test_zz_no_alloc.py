"""apg_hip_pipeline_static (the allocation-free entry point) against
apg_hip_pipeline through the C ABI: two calls in a row on the same device
array, the second on a smaller pair with other parameters, each bit-exact
against apg_hip_pipeline on the same inputs; pairs beyond the array and
configurations off the fused schedule are refused before any launch."""
import ctypes as C

import numpy as np
import pytest

import apg_amd as A

OUT4 = ("disparity", "confidence", "invalid_reason", "reliable")


def _static(ctx, p, left, right):
    h, w = left.shape
    dl, dr = ctx.to_device(left), ctx.to_device(right)
    o = [ctx.array((h, w), t) for t in (np.float32, np.float32, np.uint8, np.uint8)]
    rc = ctx.lib.apg_hip_pipeline_static(C.byref(p), dl.ptr, dr.ptr, w, h, w, 1, o[0].ptr, o[1].ptr, o[2].ptr,
                                         o[3].ptr, ctx.lib.apg_hip_stream(ctx.h))
    if rc:
        raise A.ApgError(rc, ctx.lib.apg_hip_static_last_error().decode())
    ctx.sync()
    out = [a.download() for a in o]
    for a in [dl, dr] + o:
        a.free()
    return out


@pytest.mark.gpu
def test_static_entry_matches_pipeline_on_two_calls(ctx):
    lib = A.load()
    cases = [(1, 1280, 720, 128, 20260101), (0, 640, 480, 64, 7)]   # second call smaller, Fast
    for k, (mode, w, h, D, seed) in enumerate(cases):
        p = A.params(mode, D)
        need = lib.apg_hip_pipeline_static_bytes(C.byref(p), w, h)
        assert 0 < need <= lib.apg_hip_static_scratch_bytes()
        left, right, _ = A.synth_pair(1 if k == 0 else 0, w, h, D, seed)
        print(f"static call {k}: mode {mode} {w}x{h} D={D} ({need >> 20} MiB)", flush=True)
        got = _static(ctx, p, left, right)
        want = A.run_pipeline(ctx, p, left, right)
        for n, g in zip(OUT4, got):
            np.testing.assert_array_equal(g, getattr(want, n), err_msg=f"call {k} {n}")


@pytest.mark.gpu
def test_static_entry_refuses_what_does_not_fit(ctx):
    left, right, _ = A.synth_pair(0, 64, 48, 16, 3)
    big = A.params(1, 256)
    w, h = 1920, 1080
    bufs = [ctx.array((h, w), t) for t in (np.uint8, np.uint8, np.float32, np.float32, np.uint8, np.uint8)]
    rc = ctx.lib.apg_hip_pipeline_static(C.byref(big), *[b.ptr for b in bufs[:2]], w, h, w, 1,
                                         *[b.ptr for b in bufs[2:]], ctx.lib.apg_hip_stream(ctx.h))
    assert rc == 6 and b"static scratch" in ctx.lib.apg_hip_static_last_error()
    for b in bufs:
        b.free()
    hq = A.params(2, 16)
    with pytest.raises(A.ApgError) as e:
        _static(ctx, hq, left, right)
    assert e.value.code == 5
