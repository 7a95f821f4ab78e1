"""CPU checks of the shared library: it loads, exports every symbol that
include/*.h declares, its presets equal the reference's (via the oracle's
restatement), parameter validation and the synthetic generator."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import apg_amd as A
import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in ("apg_sgm_hip.h", "apg_sgm_c_api.h"):
        with open(os.path.join(ROOT, "include", fn)) as f:
            txt = f.read()
        for m in re.finditer(r"APG_(?:HIP_)?API\s+[^;(]*?\b(apg_\w+)\s*\(", txt, re.S):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    names = declared_symbols()
    assert len(names) >= 40
    out = subprocess.run(["nm", "-D", "--defined-only", A.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = A.load()
    for n in names:
        assert getattr(lib, n) is not None


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("dmax", [16, 128, 256])
def test_presets_match_reference_from_mode(mode, dmax):
    p = A.params(mode, dmax)
    c = O.cfg(mode, dmax)
    assert p.as_dict() == c.as_dict()


def test_validate_params_codes():
    lib = A.load()

    def v(**over):
        p = A.params(1, 128)
        for k, val in over.items():
            setattr(p, k, val)
        return lib.apg_hip_validate_params(C.byref(p))
    assert v() == 0
    assert v(max_disparity=0) == 3
    assert v(P2_base=5, P1=10) == 4
    assert v(P1=-1) == 4
    assert v(grad_t1=30, grad_t2=10) == 4
    assert v(cost_max=70000) == 4
    # past the reference's u16 criterion npath*(cost_max + P2max) < 65535 the sums
    # are u32; one path's values (cost_max + P2max) must still fit u16
    assert v(cost_max=20000) == 0
    assert v(cost_max=65415) == 5
    assert v(max_disparity=1024) == 5
    assert v(max_arm_length=100) == 5
    assert v(refine_enable=1, refine_iterations=64) == 0
    assert v(refine_enable=1, refine_iterations=65) == 5
    assert v(refine_enable=0, refine_iterations=65) == 0
    assert v(refine_enable=1, refine_random_radius=1024) == 0
    assert v(refine_enable=1, refine_random_radius=1025) == 5


def test_synth_deterministic_and_shaped():
    a = A.synth_pair(1, 320, 240, 96, 7)
    b = A.synth_pair(1, 320, 240, 96, 7)
    c = A.synth_pair(1, 320, 240, 96, 8)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert not np.array_equal(a[0], c[0])
    left, right, gt = a
    assert left.shape == (240, 320) and gt.min() >= 4 and gt.max() <= 96 - 8
    # the right image is the left shifted by the ground truth where visible
    l2, r2, g2 = A.synth_pair(0, 200, 100, 64, 3)
    y, x = 50, 150
    d = int(g2[y, x])
    assert abs(int(l2[y, x]) - int(r2[y, x - d])) <= 16


def test_no_gpu_fails_loudly():
    """Without a visible GPU the device entry points return an error (no CPU fallback)."""
    try:
        ctx = A.Context(0)
    except A.ApgError:
        return
    ctx.close()
    pytest.skip("a GPU is visible here")


def test_pipeline_workspace_bytes_closed_form():
    # host-only sizing used by the serving form's HBM admission:
    # per pixel 6 (gray/gx/gy x2) + 8 (symmetric census x2) + 4 (range) +
    # 12 (arms, right disparity, post scratch) + 4 (right-WTA keys) = 34 B;
    # per state C' (u8) + T (u8, Cross only) + one u8 delta plane per path, or,
    # where the half-wave SGM pairs a scanline's two directions (D % 16 == 0,
    # D <= 128, P2 <= 127), one plane per scanline family
    def ws(mode, D, w, h, **over):
        return A.workspace_bytes(A.params(mode, D, **over), w, h)
    for (w, h, D) in [(640, 480, 64), (1280, 720, 128), (1920, 1080, 256), (333, 101, 48), (200, 90, 40)]:
        n, nv = w * h, w * h * D
        pairs = D % 16 == 0 and D <= 128
        planes4, planes8 = (2, 4) if pairs else (4, 8)
        spare = 256 if pairs else 0   # bytes past each paired plane's states
        lut = ws(1, D, w, h) - (34 * n + (2 + planes4) * nv + planes4 * spare)
        assert 0 < lut < 64 * 1024
        assert ws(0, D, w, h) == 34 * n + (1 + planes4) * nv + planes4 * spare + lut          # Fast: no Cross, no T
        assert ws(1, D, w, h, paths=8) == 34 * n + (2 + planes8) * nv + planes8 * spare + lut
        # P2 above 127: one plane per direction
        assert ws(1, D, w, h, P2_base=200) == 34 * n + 6 * nv + lut
    # stage-by-stage schedule (max(P2, P1+1) > 254): three u16 volumes, best / second
    n, nv = 200 * 100, 200 * 100 * 32
    lut = ws(1, 32, 200, 100) - (34 * n + 4 * nv + 2 * 256)   # paired planes at D = 32
    assert ws(1, 32, 200, 100, P2_base=300) == (34 - 4 + 8) * n + 6 * nv + lut
    assert A.workspace_bytes(A.params(1, 32, min_disparity=40), 10, 10) == 0   # max <= min


def test_workspace_bytes_count_the_refiner():
    # HighQuality admission must see the refiner's random stream
    # (4 B x (iterations + 1) x (w * h + 4096)) and its sweep scratch
    w, h = 640, 360
    hq = A.workspace_bytes(A.params(2, 64), w, h)
    no_ref = A.workspace_bytes(A.params(2, 64, refine_enable=0), w, h)
    assert hq - no_ref >= 4 * 4 * (w * h + 4096) + 100 * w * h


REF_HEADER_PATHS = ["cost_computer", "cost_aggregator", "sgm_optimizer", "postprocess", "prior_estimator",
                    "confidence_estimator", "refiner", "session", "workspace", "execution", "result",
                    "pipeline_stats", "pipeline", "buffers", "config", "image", "image_view", "packed_volume",
                    "search_range", "status", "types", "metrics", "ablation"]


@pytest.mark.parametrize("name", REF_HEADER_PATHS)
def test_reference_header_path_compiles_alone(tmp_path, name):
    # each header the reference's clients include, by its path, in a TU of its own
    src = tmp_path / "tu.cpp"
    src.write_text(f'#include "apg_sgm/{name}.hpp"\nint main() {{ return 0; }}\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


def test_reference_headers_client_is_built():
    assert os.path.exists(os.path.join(ROOT, "apg-sgm_amd", "_build", "reference_headers_main"))


@pytest.mark.parametrize("over,ok", [
    ({"cost_max": 16383}, 0), ({"paths": 8, "cost_max": 9000}, 0), ({"post_median_radius": 9}, 0),
    ({"post_median_radius": 64}, 0), ({"post_median_radius": 65}, 5), ({"cost_max": 65500}, 5),
])
def test_validate_u32_sums_and_median_radius(over, ok):
    # u32 path sums past the u16 bound; one path's values must still fit u16
    assert A.load().apg_hip_validate_params(C.byref(A.params(1, 64, **over))) == ok
