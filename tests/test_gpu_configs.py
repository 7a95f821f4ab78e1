"""The five benchmark configurations on the GPU, each at its own workload.

* Plain Balanced / Fast 640x480 D=64 and the other hashed cases: every output
  and intermediate's FNV-1a-64 against the reference build's hashes
  (tests/golden/golden_hashes.json, written by make_golden.py), on the fused
  schedule and on the stage-by-stage schedule.  This is the method of the
  reference's own frozen fingerprints (tests/test_api_contract.cpp:243-257).
* HighQuality 1920x1080 D=256 with the bench seed, and a small D=256 case:
  ranges, d_before_refine and every output bit-exact against the oracle.
* The uneven 64-pair batch: every pair of a graph replay bit-identical to
  apg_hip_pipeline on that pair alone, two pairs against the oracle.
* The serving mix: each request's results the same alone and in the mix.
* HighQuality twice through one default (ReleasePerFrame) StereoSession,
  each frame against the oracle: the refiner's random stream must be
  regenerated after the workspace is released.
"""
import json
import os

import numpy as np
import pytest

import apg_amd as A
import pyoracle as O
from apg_amd.workloads import SEED_SERVE, SERVE_CLASSES, batch_specs, serve_specs

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
OUT4 = ("disparity", "confidence", "invalid_reason", "reliable")


def _hashed():
    with open(os.path.join(GOLD, "golden_hashes.json")) as f:
        return json.load(f)


def _cmp(got, want, what):
    assert got.shape == want.shape, what
    if not np.array_equal(got, want, equal_nan=True):
        bad = np.argwhere(~((got == want) | (np.isnan(got) & np.isnan(want))) if got.dtype.kind == "f"
                          else got != want)
        i = tuple(bad[0])
        raise AssertionError(f"{what}: {len(bad)} mismatches, first at {i}: gpu={got[i]} want={want[i]}")


# ---------------------------------------------------- golden hashes (C0) ----

# the fused schedule keeps C and C' on chip (u8 T / C'); everything else is
# materialised on request.  HighQuality runs its own per-range schedule.
FUSED_KEYS = ("sgm", "disp_right", "disp_wta", "best", "second", "dmin", "dmax")


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", [0, 1], ids=["fused", "stage_by_stage"])
@pytest.mark.parametrize("name", sorted(_hashed()))
def test_golden_hashes_on_gpu(ctx, name, schedule):
    g = _hashed()[name]
    gen = g["generator"]
    left, right, _ = A.synth_pair(gen["kind"], gen["w"], gen["h"], gen["max_disp"], gen["seed"])
    assert hex(O.fnv1a64(left)) == g["input_fnv"]["left"] and hex(O.fnv1a64(right)) == g["input_fnv"]["right"]
    p = A.Params()
    for k, v in g["cfg"].items():
        setattr(p, k, v)
    print(f"golden {name} schedule={schedule}", flush=True)
    A.set_schedule(ctx, schedule)
    try:
        inter = True if schedule == 1 else FUSED_KEYS
        r = A.run_pipeline(ctx, p, left, right, intermediates=inter)
    finally:
        A.set_schedule(ctx, 0)
    got = {k: getattr(r, k) for k in OUT4}
    keys = list(OUT4) + (["cost", "cost_agg"] if schedule == 1 else []) + list(FUSED_KEYS)
    for k in keys:
        if k not in got:
            got[k] = getattr(r, k)
    got["d_before_refine"] = r.disp_wta
    bad = [k for k in g["fnv"] if k in got and hex(O.fnv1a64(got[k])) != g["fnv"][k]]
    missing = [k for k in g["fnv"] if k not in got]
    assert not bad, f"{name}: FNV differs from the reference build for {bad}"
    assert missing == ([] if schedule == 1 else ["cost", "cost_agg"]), missing


# ---------------------------------------------- HighQuality full size (C2) --

SEED_HQ = 20260404


@pytest.mark.gpu
@pytest.mark.parametrize("w,h,D,seed", [(600, 200, 256, 99), (1920, 1080, 256, SEED_HQ)],
                         ids=["hq_600x200_d256", "hq_1920x1080_d256"])
def test_hq_d256_matches_oracle(ctx, w, h, D, seed):
    p = A.params(2, D)
    left, right, _ = A.synth_pair(2, w, h, D, seed)
    print(f"HQ {w}x{h} D={D} seed={seed}: device", flush=True)
    g = A.run_pipeline(ctx, p, left, right, intermediates=("disp_wta", "dmin", "dmax", "disp_right", "after_refine"))
    print("oracle", flush=True)
    o = O.pipeline(O.cfg_from(p.as_dict()), left, right, intermediates=True)
    for k in ("dmin", "dmax", "disp_right", "after_refine") + OUT4:
        _cmp(getattr(g, k), o[k], k)
    _cmp(g.disp_wta, o["disp_wta"], "d_before_refine")


# ------------------------------------------------------ the batch (C3) ----

@pytest.mark.gpu
def test_batch_of_64_uneven_pairs_graph_replay(ctx):
    specs = batch_specs(64)
    print(f"C3 batch: {len(specs)} pairs, {sum(w * h * d for w, h, d, _ in specs) / 1e9:.2f} G states", flush=True)
    A.set_schedule(ctx, 0)
    pairs, inputs = [], []
    for w, h, d, sd in specs:
        left, right, _ = A.synth_pair(0, w, h, d, sd)
        inputs.append((left, right))
        pairs.append(A.DevicePair(ctx, A.params(1, d), left, right))
    alone = []
    for q in pairs:
        ctx._check(ctx.lib.apg_hip_pipeline(ctx.h, A.C.byref(q.params), q.left.ptr, q.right.ptr, q.w, q.h, q.w, 1,
                                            q.disp.ptr, q.conf.ptr, q.reason.ptr, q.rel.ptr, None), "pipeline")
        ctx.sync()
        alone.append(q.outputs())
    ctx.lib.apg_hip_release_workspace(ctx.h)
    print("alone done; graph capture + 2 replays", flush=True)
    with A.Batch(0, 4) as b:
        arr = b.pair_array(pairs)
        for rep in range(3):
            for q in pairs:
                q.clear_outputs()
            b.run(arr, graph=True)
            b.sync()
            for i, q in enumerate(pairs):
                for n, x, y in zip(OUT4, q.outputs(), alone[i]):
                    _cmp(x, y, f"graph call {rep} pair {i} {specs[i][:3]} {n}")
    small = sorted(range(len(specs)), key=lambda i: specs[i][0] * specs[i][1] * specs[i][2])[:2]
    for i in small:
        w, h, d, _ = specs[i]
        print(f"oracle check pair {i}: {w}x{h} D={d}", flush=True)
        o = O.pipeline(O.cfg_from(A.params(1, d).as_dict()), *inputs[i])
        for n, x in zip(OUT4, alone[i]):
            _cmp(x, o[n], f"pair {i} {n}")
    for q in pairs:
        for a in (q.left, q.right, q.disp, q.conf, q.reason, q.rel):
            a.free()


# ---------------------------------------------------- serving mix (C4) ----

@pytest.mark.gpu
@pytest.mark.parametrize("rate", [500.0, 1e6], ids=["500_per_s", "saturated"])
def test_serving_mix_same_alone_and_in_mix(ctx, rate):
    trace = serve_specs(40, rate, SEED_SERVE + 7)
    pool = {}
    for ci, (w, h, d) in enumerate(SERVE_CLASSES):
        for v in range(4):
            pool[(ci, v)] = A.synth_pair(0, w, h, d, SEED_SERVE + 10 * ci + v)[:2]
    A.set_schedule(ctx, 0)
    alone = {}
    for _, ci, mode, v in trace:
        if (ci, mode, v) not in alone:
            r = A.run_pipeline(ctx, A.params(mode, SERVE_CLASSES[ci][2]), *pool[(ci, v)])
            alone[(ci, mode, v)] = tuple(getattr(r, k) for k in OUT4)
    ctx.lib.apg_hip_release_workspace(ctx.h)
    # one request object (own output buffers) per arrival
    reqs = [A.HostRequest(A.params(mode, SERVE_CLASSES[ci][2]), *pool[(ci, v)]) for _, ci, mode, v in trace]
    print(f"C4 serving: {len(reqs)} requests at {rate}/s, classes "
          f"{[sum(1 for t in trace if t[1] == c) for c in range(3)]}", flush=True)
    with A.Server(0, 4) as s:
        lat = s.replay(reqs, [t for t, _, _, _ in trace])
        assert s.pending() == 0
    assert np.all(lat > 0)
    for k, ((_, ci, mode, v), r) in enumerate(zip(trace, reqs)):
        for n, x, y in zip(OUT4, r.outputs(), alone[(ci, mode, v)]):
            _cmp(x, y, f"request {k} class {ci} mode {mode} {n}")


# ------------------------------------- HighQuality frames on one session --

@pytest.mark.gpu
def test_hq_two_frames_through_default_session():
    import apg_sgm as apg
    s = apg.StereoSession(mode="high_quality", max_disparity=48)
    p = A.params(2, 48)
    for f, seed in enumerate((5, 6)):
        left, right, _ = A.synth_pair(2, 160, 120, 48, seed)
        print(f"HQ session frame {f}", flush=True)
        r = s.process(left, right)
        o = O.pipeline(O.cfg_from(p.as_dict()), left, right)
        _cmp(np.asarray(r.disparity), o["disparity"], f"frame {f} disparity")
        _cmp(np.asarray(r.invalid_reason), o["invalid_reason"], f"frame {f} invalid_reason")
        assert s.allocated_bytes == 0
