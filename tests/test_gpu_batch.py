"""Batched and serving forms (apg_hip_batch_*, apg_hip_server_*, C ABI).

Every pair's outputs (disparity, confidence, invalid_reason, reliable mask)
must be bit-identical to apg_hip_pipeline on that pair alone, whatever slot
it lands on, whatever runs beside it, whatever launch shapes it uses, and on
every replay of a captured graph.  The alone results are themselves pinned
to the oracle by test_gpu_pipeline.py; two small pairs of the mix are
checked against the oracle here as well.
"""
import json

import numpy as np
import pytest

import apg_amd as A
import pyoracle as O

# (kind, w, h, D, mode): uneven sizes, D from 16 to 256 (K = 1 and 2 SGM
# lanes, full and partial), Balanced and Fast, 8 paths, 64-bit census
MIX = [
    (1, 320, 240, 64, 1, {}),
    (0, 160, 120, 32, 0, {}),
    (2, 400, 200, 128, 1, {}),
    (1, 256, 160, 48, 1, {}),
    (0, 96, 64, 16, 0, {}),
    (1, 560, 96, 256, 1, {}),
    (2, 420, 150, 192, 1, {}),
    (0, 128, 100, 40, 1, {"paths": 8}),
    (1, 208, 136, 64, 1, {"census_type": 0}),
    (2, 300, 180, 96, 0, {}),
    (0, 640, 48, 128, 1, {}),
    (1, 72, 300, 24, 1, {}),
]


def _params(mode, D, over):
    return A.params(mode, D, **over)


def _inputs(i, kind, w, h, D):
    left, right, _ = A.synth_pair(kind, w, h, D, 777 + 31 * i)
    return left, right


@pytest.fixture(scope="module")
def alone(ctx):
    """apg_hip_pipeline on each pair alone (default launch shapes)."""
    A.set_schedule(ctx, 0)
    A.set_tuning(ctx, None)
    res = []
    for i, (kind, w, h, D, mode, over) in enumerate(MIX):
        left, right = _inputs(i, kind, w, h, D)
        g = A.run_pipeline(ctx, _params(mode, D, over), left, right)
        res.append((g.disparity, g.confidence, g.invalid_reason, g.reliable))
    return res


@pytest.fixture(scope="module")
def device_pairs(ctx):
    pairs = []
    for i, (kind, w, h, D, mode, over) in enumerate(MIX):
        left, right = _inputs(i, kind, w, h, D)
        pairs.append(A.DevicePair(ctx, _params(mode, D, over), left, right))
    return pairs


def _same(got, want, what):
    names = ("disparity", "confidence", "invalid_reason", "reliable")
    for n, a, b in zip(names, got, want):
        assert a.shape == b.shape, (what, n)
        if not np.array_equal(a, b, equal_nan=True):
            bad = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b))) if a.dtype.kind == "f" else a != b)
            raise AssertionError(f"{what}: {n} differs at {len(bad)} pixels, first {tuple(bad[0])}")


@pytest.mark.gpu
def test_alone_results_of_small_pairs_match_oracle(alone):
    for i in (1, 4):
        kind, w, h, D, mode, over = MIX[i]
        print(f"oracle check pair {i}: {w}x{h} D={D} mode={mode}", flush=True)
        left, right = _inputs(i, kind, w, h, D)
        o = O.pipeline(O.cfg_from(_params(mode, D, over).as_dict()), left, right)
        _same(alone[i], (o["disparity"], o["confidence"], o["invalid_reason"], o["reliable"]), f"pair {i}")


@pytest.mark.gpu
@pytest.mark.parametrize("slots,grouped", [(1, True), (3, True), (4, True), (1, False), (4, False)])
def test_batch_direct_matches_alone(alone, device_pairs, slots, grouped):
    print(f"batch direct slots={slots} grouped={grouped}", flush=True)
    with A.Batch(0, slots) as b:
        for rep in range(2):
            for q in device_pairs:
                q.clear_outputs()
            b.run(device_pairs, grouped=grouped)
            b.sync()
            for i, q in enumerate(device_pairs):
                _same(q.outputs(), alone[i], f"slots={slots} grouped={grouped} call {rep} pair {i}")


# pairs of the mix whose schedule has a group form for every step: the
# fused column-range schedule (Fast, Balanced; u8 C', arms from the aux pass)
GROUPABLE = [i for i, (kind, w, h, D, mode, over) in enumerate(MIX) if mode in (0, 1)]


@pytest.mark.gpu
def test_group_launches_span_pairs(alone, device_pairs):
    # one slot: every Fast / Balanced pair of the mix runs in group launches,
    # seven steps each, so fewer launches than pairs x steps means launches
    # that cover several pairs
    print("group launch plan, 1 slot", flush=True)
    with A.Batch(0, 1) as b:
        for q in device_pairs:
            q.clear_outputs()
        b.run(device_pairs)
        b.sync()
        grouped, launches, single = b.group_stats()
        print(f"grouped={grouped} launches={launches} single={single}", flush=True)
        assert grouped + single == len(MIX)
        assert grouped == len(GROUPABLE)
        assert 0 < launches < 7 * grouped
        for i, q in enumerate(device_pairs):
            _same(q.outputs(), alone[i], f"grouped pair {i}")


@pytest.mark.gpu
def test_grouping_size_threshold(alone, device_pairs):
    # pairs above the state threshold run with their own kernels, before the
    # group launches of the others, on the same slot
    small = [i for i in GROUPABLE if MIX[i][1] * MIX[i][2] * MIX[i][3] <= 1_000_000]
    print(f"grouping threshold 1 M states: {len(small)} small pairs", flush=True)
    with A.Batch(0, 1) as b:
        b.set_grouping(0, 1_000_000)
        for rep in range(2):
            for q in device_pairs:
                q.clear_outputs()
            b.run(device_pairs, graph=rep == 1)
            b.sync()
            grouped, launches, single = b.group_stats()
            assert (grouped, single) == (len(small), len(MIX) - len(small))
            assert 0 < launches < 7 * grouped
            for i, q in enumerate(device_pairs):
                _same(q.outputs(), alone[i], f"threshold call {rep} pair {i}")


@pytest.mark.gpu
def test_group_launches_of_one_shape(ctx):
    # eight pairs of one size and mode (one kernel instance per step): seven
    # launches in all, each pair's outputs as alone
    print("group launches, 8 pairs of 200x120 D=64 Balanced", flush=True)
    pars = A.params(1, 64)
    pairs, want = [], []
    for i in range(8):
        left, right = _inputs(200 + i, i % 3, 200, 120, 64)
        g = A.run_pipeline(ctx, pars, left, right)
        want.append((g.disparity, g.confidence, g.invalid_reason, g.reliable))
        pairs.append(A.DevicePair(ctx, pars, left, right))
    with A.Batch(0, 1) as b:
        arr = b.pair_array(pairs)
        for rep in range(3):
            for q in pairs:
                q.clear_outputs()
            b.run(arr, graph=True)
            b.sync()
            assert b.group_stats() == (8, 7, 0)
            for i, q in enumerate(pairs):
                _same(q.outputs(), want[i], f"call {rep} pair {i}")


@pytest.mark.gpu
@pytest.mark.parametrize("grouped", [True, False])
def test_batch_graph_replays_match_alone(ctx, alone, device_pairs, grouped):
    # the first graph call runs directly and captures; each later call is one
    # graph launch.  Outputs are cleared and unrelated eager work runs on
    # another stream between replays, and every replay is checked.
    print(f"batch graph capture + 3 replays, grouped={grouped}", flush=True)
    other = A.DevicePair(ctx, A.params(1, 32), *_inputs(99, 0, 96, 64, 32))
    with A.Batch(0, 4) as b:
        arr = b.pair_array(device_pairs)
        for rep in range(4):
            for q in device_pairs:
                q.clear_outputs()
            b.run(arr, graph=True, grouped=grouped)
            b.sync()
            for i, q in enumerate(device_pairs):
                _same(q.outputs(), alone[i], f"graph call {rep} pair {i}")
            b2 = b.pair_array([other])
            with A.Batch(0, 1) as eager:
                eager.run(b2)
                eager.sync()


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(64, 0, 1024), (512, 1, 16384), (32, 1, 256), (1024, 0, 65536), (128, 1, 2048)])
def test_launch_shapes_do_not_change_results(ctx, alone, shape):
    print(f"launch shapes {shape}", flush=True)
    try:
        A.set_tuning(ctx, shape)
        for i in (0, 2, 5, 6, 7, 8):
            kind, w, h, D, mode, over = MIX[i]
            left, right = _inputs(i, kind, w, h, D)
            g = A.run_pipeline(ctx, _params(mode, D, over), left, right)
            _same((g.disparity, g.confidence, g.invalid_reason, g.reliable), alone[i], f"shape {shape} pair {i}")
    finally:
        A.set_tuning(ctx, None)


@pytest.mark.gpu
def test_autotune_cache_round_trip(tmp_path, alone, device_pairs):
    path = str(tmp_path / "shapes.json")
    sub = [device_pairs[i] for i in (0, 2, 5)]
    print("autotune 3 buckets", flush=True)
    with A.Batch(0, 2) as b:
        b.autotune(sub, path)
        tuned = [b.tuning_for(q.params, q.w, q.h).as_tuple() for q in sub]
        doc = json.load(open(path))
        assert doc["arch"].startswith("gfx")
        assert len(doc["entries"]) == 3
        for q in device_pairs:
            q.clear_outputs()
        b.run(device_pairs)
        b.sync()
        for i, q in enumerate(device_pairs):
            _same(q.outputs(), alone[i], f"tuned batch pair {i}")
    with A.Batch(0, 2) as b2:
        # every bucket is in the cache: nothing is re-timed, the shapes are read back
        b2.autotune(sub, path)
        assert [b2.tuning_for(q.params, q.w, q.h).as_tuple() for q in sub] == tuned


def _requests():
    reqs = []
    for i, (kind, w, h, D, mode, over) in enumerate(MIX):
        left, right = _inputs(i, kind, w, h, D)
        reqs.append(A.HostRequest(_params(mode, D, over), left, right))
    return reqs


@pytest.mark.gpu
@pytest.mark.parametrize("spacing_ms", [0.0, 0.3])
def test_server_results_same_alone_and_in_mix(alone, spacing_ms):
    # each request twice, interleaved, so equal pairs run beside different neighbours
    reqs = _requests() + _requests()[::-1]
    n = len(MIX)
    arrival = [k * spacing_ms * 1e-3 for k in range(len(reqs))]
    print(f"server replay {len(reqs)} requests spacing {spacing_ms} ms", flush=True)
    with A.Server(0, 4) as s:
        lat = s.replay(reqs, arrival)
        assert s.pending() == 0
        assert 0 < s.resident_bytes() <= 64 << 30
    assert np.all(lat > 0)
    for k, r in enumerate(reqs):
        i = k if k < n else 2 * n - 1 - k
        _same(r.outputs(), alone[i], f"request {k} (pair {i})")


@pytest.mark.gpu
def test_server_budget_admission(alone):
    reqs = _requests()
    need = max(A.workspace_bytes(r.req.params, r.req.width, r.req.height) for r in reqs)
    # room for about one large workspace: admission must wait and release idle
    # slots, and every result must still be the same
    budget = int(need * 1.6)
    print(f"server budget {budget} bytes", flush=True)
    with A.Server(0, 4, budget) as s:
        lat = s.replay(reqs, [0.0] * len(reqs))
        assert s.resident_bytes() <= budget
        big = A.HostRequest(A.params(1, 256), *_inputs(5, 1, 1600, 900, 256))
        with pytest.raises(A.ApgError) as e:
            s.submit(big, 1)
        assert e.value.code == 6
    assert np.all(lat > 0)
    for i, r in enumerate(reqs):
        _same(r.outputs(), alone[i], f"budget request {i}")


@pytest.mark.gpu
def test_server_budget_admission_with_highquality(ctx):
    # the HighQuality workspace includes the refiner's random stream and sweep
    # scratch: a budget of one HQ request's bytes admits it alone, and a
    # Balanced request queued beside it waits for the HQ slot to be released
    hq = A.HostRequest(A.params(2, 64), *_inputs(50, 2, 320, 240, 64))
    bal = A.HostRequest(A.params(1, 32), *_inputs(51, 1, 160, 120, 32))
    ws_hq = A.workspace_bytes(hq.req.params, 320, 240)
    no_ref = A.params(2, 64, refine_enable=0)
    assert ws_hq - A.workspace_bytes(no_ref, 320, 240) >= 4 * 4 * (320 * 240 + 4096)
    budget = ws_hq + 320 * 240 * 16
    print(f"server HQ budget {budget} bytes", flush=True)
    with A.Server(0, 2, budget) as s:
        s.replay([hq, bal], [0.0, 0.0])
        assert s.resident_bytes() <= budget
    for r, kind, w, h, D, mode, i in ((hq, 2, 320, 240, 64, 2, 50), (bal, 1, 160, 120, 32, 1, 51)):
        left, right = _inputs(i, kind, w, h, D)
        o = O.pipeline(O.cfg_from(A.params(mode, D).as_dict()), left, right)
        _same(r.outputs(), (o["disparity"], o["confidence"], o["invalid_reason"], o["reliable"]), f"mode {mode}")


@pytest.mark.gpu
def test_batch_graph_with_highquality(ctx):
    # HighQuality pairs (device prior, refiner with its banded sweep) in a
    # captured batch beside a Balanced pair: every replay, with unrelated
    # eager work between replays, equals each pair run alone
    specs = [(2, 200, 150, 64, 0, 41), (2, 160, 120, 48, 1, 42), (1, 200, 150, 64, 2, 43)]
    pairs, want = [], []
    for mode, w, h, D, kind, seed in specs:
        left, right, _ = A.synth_pair(kind, w, h, D, seed)
        g = A.run_pipeline(ctx, A.params(mode, D), left, right)
        want.append((g.disparity, g.confidence, g.invalid_reason, g.reliable))
        pairs.append(A.DevicePair(ctx, A.params(mode, D), left, right))
    other = A.DevicePair(ctx, A.params(1, 32), *_inputs(98, 0, 96, 64, 32))
    with A.Batch(0, 2) as b:
        arr = b.pair_array(pairs)
        for rep in range(4):
            print(f"HighQuality batch graph call {rep}", flush=True)
            for q in pairs:
                q.clear_outputs()
            b.run(arr, graph=True)
            b.sync()
            for i, q in enumerate(pairs):
                _same(q.outputs(), want[i], f"HighQuality graph call {rep} pair {i}")
            with A.Batch(0, 1) as eager:
                eager.run(eager.pair_array([other]))
                eager.sync()
