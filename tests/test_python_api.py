"""Python API (apg_sgm: StereoSession / StereoResult / PipelineStats and the
enums), mirroring the reference's python/tests/test_bindings.py: option
validation on the CPU; on the GPU, the frozen C++ contract fingerprints of
the reference (test_bindings.py:305-347), strided and read-only inputs,
result lifetime, serialised use from threads and the memory policies."""
import concurrent.futures
import gc

import numpy as np
import pytest

import apg_amd as A
import apg_sgm as apg
import pyoracle as O


def fnv1a64(b: bytes) -> int:
    h = 0xCBF29CE484222325
    for x in b:
        h ^= x
        h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_module_surface_and_enums():
    assert set(apg.__all__) >= {"StereoSession", "StereoResult", "PipelineStats", "StageTiming", "QualityMode",
                                "VolumeBackend", "MemoryPolicy", "StatusCode", "HierarchicalFallback"}
    assert int(apg.StatusCode.Ok) == 0 and int(apg.StatusCode.InternalError) == 8
    assert int(apg.QualityMode.HighQuality) == 2 and int(apg.VolumeBackend.Packed) == 2
    assert int(apg.MemoryPolicy.BoundedReuse) == 1


@pytest.mark.parametrize("kw", [{"mode": "ultra"}, {"backend": "cuda"}, {"memory_policy": "keep"},
                                {"hierarchical_fallback": "closest"}])
def test_bad_option_strings_raise_value_error(kw):
    with pytest.raises(ValueError):
        apg.StereoSession(**kw)


def test_invalid_config_raises_runtime_error():
    with pytest.raises(RuntimeError):
        apg.StereoSession(max_disparity=0, min_disparity=0)


def _checker(w, h, ch=1):
    yy, xx = np.mgrid[0:h, 0:w]
    if ch == 1:
        return np.where(((xx // 4 + yy // 4) % 2) == 0, 40, 210).astype(np.uint8)
    v = (xx * 37 + yy * 73 + xx * yy * 13) % 256
    return np.stack([v, (v + 30) % 256, (v + 60) % 256], axis=-1).astype(np.uint8)


@pytest.mark.gpu
def test_frozen_contract_fingerprints_through_python():
    # the reference's frozen outputs for left == right inputs (Fast, packed backend)
    left1 = _checker(33, 27)
    r1 = apg.StereoSession(mode="fast", max_disparity=16, backend="packed", threads=1).process(left1, left1.copy())
    assert fnv1a64(r1.disparity.tobytes()) == 0x1710451F7EAB215
    assert fnv1a64(r1.confidence.tobytes()) == 0xE61822CC0FBE3A38
    assert fnv1a64(r1.invalid_reason.tobytes()) == 0x22081E533CD30004
    assert fnv1a64(r1.reliable_mask.tobytes()) == 0x9ADE9D63E4FCAE7B
    left2 = _checker(32, 24, 3)
    r2 = apg.StereoSession(mode="fast", max_disparity=16, backend="packed", threads=1).process(left2, left2.copy())
    assert fnv1a64(r2.disparity.tobytes()) == 0x7E8E0FA784351325
    assert fnv1a64(r2.confidence.tobytes()) == 0x5860E8F479CBE0FD
    assert fnv1a64(r2.invalid_reason.tobytes()) == 0x82F55CC1FD68A695
    assert fnv1a64(r2.reliable_mask.tobytes()) == 0x9DB48CB5D76242B5


@pytest.mark.gpu
@pytest.mark.parametrize("mode,name", [(0, "fast"), (1, "balanced"), (2, "high_quality")])
def test_session_matches_oracle(mode, name):
    left, right, _ = A.synth_pair(1, 160, 96, 48, 321 + mode)
    s = apg.StereoSession(mode=name, max_disparity=48)
    r = s.process(left, right)
    o = O.pipeline(O.cfg(mode, 48), left, right)
    np.testing.assert_array_equal(r.disparity, o["disparity"])
    np.testing.assert_array_equal(r.confidence, o["confidence"])
    np.testing.assert_array_equal(r.invalid_reason, o["invalid_reason"])
    np.testing.assert_array_equal(r.reliable_mask, o["reliable"])
    assert (r.width, r.height, r.empty) == (160, 96, False)
    assert s.frame_count == 1 and s.total_time_ms > 0
    d = r.stats.to_dict()
    assert 0.0 <= d["valid_disparity_ratio"] <= 1.0 and d["timing"]["total_ms"] > 0


@pytest.mark.gpu
def test_strided_readonly_inputs_and_result_lifetime():
    left, right, _ = A.synth_pair(0, 120, 80, 32, 9)
    s = apg.StereoSession(mode="balanced", max_disparity=32)
    ref = s.process(left, right).disparity.copy()
    # padded rows, a column-sliced view and read-only buffers give the same result
    pl = np.zeros((80, 200), np.uint8)
    pr = np.zeros((80, 200), np.uint8)
    pl[:, :120], pr[:, :120] = left, right
    np.testing.assert_array_equal(s.process(pl[:, :120], pr[:, :120]).disparity, ref)
    wide_l = np.repeat(left, 2, axis=1)[:, ::2]
    wide_r = np.repeat(right, 2, axis=1)[:, ::2]
    np.testing.assert_array_equal(s.process(wide_l, wide_r).disparity, ref)
    lro, rro = left.copy(), right.copy()
    lro.flags.writeable = False
    rro.flags.writeable = False
    np.testing.assert_array_equal(s.process(lro, rro).disparity, ref)
    # views outlive the session and the result object
    view = s.process(left, right).disparity
    del s
    gc.collect()
    np.testing.assert_array_equal(view, ref)


@pytest.mark.gpu
def test_memory_policies_and_reset():
    left, right, _ = A.synth_pair(0, 96, 64, 24, 5)
    s = apg.StereoSession(mode="balanced", max_disparity=24)
    s.process(left, right)
    assert s.allocated_bytes == 0   # release_per_frame
    b = apg.StereoSession(mode="balanced", max_disparity=24, memory_policy="bounded_reuse",
                          max_retained_bytes=1 << 30)
    b.process(left, right)
    assert b.allocated_bytes > 0
    b.reset()
    assert b.allocated_bytes == 0 and b.frame_count == 0
    with pytest.raises(ValueError):
        b.process(left, right[:, :-1])


@pytest.mark.gpu
def test_threads_on_one_session_are_serialised():
    left, right, _ = A.synth_pair(0, 128, 96, 16, 6)
    s = apg.StereoSession(mode="fast", max_disparity=16)
    want = s.process(left, right).disparity.copy()
    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as ex:
        outs = list(ex.map(lambda _: s.process(left, right).disparity.copy(), range(8)))
    assert all(np.array_equal(o, want) for o in outs)
    assert s.frame_count == 9


@pytest.mark.gpu
def test_bounded_reuse_across_sizes_and_threads():
    # a retaining workspace stages images and results through pinned host
    # memory that grows with the frame and is then reused; result planes of
    # the right size are reused as well.  Frames of changing sizes, then one
    # session per thread at once (the copy threads are shared), all against
    # the oracle.
    sizes = [(160, 96), (200, 80), (150, 60), (200, 80), (96, 64)]
    frames = []
    for i, (w, h) in enumerate(sizes):
        left, right, _ = A.synth_pair(i % 3, w, h, 48, 300 + i)
        o = O.pipeline(O.cfg(1, 48), left, right)
        frames.append((left, right, o))
    s = apg.StereoSession(mode="balanced", max_disparity=48, memory_policy="bounded_reuse",
                          max_retained_bytes=1 << 30)
    for i, (left, right, o) in enumerate(frames):
        print(f"bounded reuse frame {i}: {left.shape[1]}x{left.shape[0]}", flush=True)
        r = s.process(left, right)
        np.testing.assert_array_equal(r.disparity, o["disparity"])
        np.testing.assert_array_equal(r.confidence, o["confidence"])
        np.testing.assert_array_equal(r.invalid_reason, o["invalid_reason"])
        np.testing.assert_array_equal(r.reliable_mask, o["reliable"])

    def run(i):
        left, right, o = frames[i % len(frames)]
        t = apg.StereoSession(mode="balanced", max_disparity=48, memory_policy="bounded_reuse",
                              max_retained_bytes=1 << 30)
        ok = True
        for _ in range(3):
            r = t.process(left, right)
            ok = ok and np.array_equal(r.disparity, o["disparity"]) and np.array_equal(r.confidence, o["confidence"])
        return ok
    print("bounded reuse: 4 sessions on 4 threads", flush=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as ex:
        assert all(ex.map(run, range(4)))
