"""The benchmark app (apg-sgm_amd/apg-benchmark) against the reference's own
(reference apps/benchmark_stereo.cpp): same manifest in, same CSV out.

Fixture: tests/golden/bench_data (a synthetic three-case dataset) and
tests/golden/bench_reference.csv, the CSV the reference app wrote for it
(tests/golden/make_benchmark_csv.py).  Run-to-run columns (wall-clock times)
and the memory columns that count this build's device workspace instead of the
reference's host volumes (32-bit sums are never materialised here, and the
ragged ranges of the packed configurations index a dense device volume) are
left out of the comparison; every other column, accuracy metrics, search
widths, support statistics, the dense cost-volume size and the telemetry block
included, must match character for character.
"""
import csv
import io
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
DATA = os.path.join(GOLD, "bench_data")
APP = os.path.join(ROOT, "apg-sgm_amd", "apg-benchmark")
METRICS = os.path.join(ROOT, "apg-sgm_amd", "_build", "metrics_main")


def _rows(text):
    r = list(csv.reader(io.StringIO(text)))
    return r[0], r[1:]


def _reference():
    with open(os.path.join(GOLD, "bench_reference.csv")) as f:
        return _rows(f.read())


def _needs(path):
    if not os.path.exists(path):
        pytest.skip("not built: %s (make -C apg-sgm_amd)" % path)


@pytest.mark.parametrize("case,aid", [("slanted_vis", "A"), ("slanted_vis", "D"), ("slanted", "A"), ("slanted", "D")])
def test_metrics_match_reference_csv(case, aid):
    """apg::evaluate_stereo on the reference's own disparity map gives the
    reference app's metric columns (global, KITTI D1, edge split, and the
    all/noc/occ/boundary/texture regions)."""
    _needs(METRICS)
    args = [METRICS, os.path.join(DATA, "ref_%s_%s_disp.pfm" % (case, aid)), os.path.join(DATA, case + "_gt.pfm"),
            os.path.join(DATA, case + "_left.pgm")]
    if case == "slanted_vis":
        args.append(os.path.join(DATA, case + "_vis.pgm"))
    out = subprocess.run(args, capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    got = json.loads(out.stdout)
    head, rows = _reference()
    row = next(r for r in rows if r[head.index("case")] == case and r[head.index("ablation_id")] == aid)
    assert len(got) == 74
    for k, v in got.items():
        assert row[head.index(k)] == v, (k, row[head.index(k)], v)


def test_metrics_plant_sanity():
    """The comparison above can fail: the ground truth against itself gives
    zero error, which differs from the reference row."""
    _needs(METRICS)
    gt = os.path.join(DATA, "slanted_gt.pfm")
    out = subprocess.run([METRICS, gt, gt, os.path.join(DATA, "slanted_left.pgm")], capture_output=True, text=True,
                         timeout=60)
    got = json.loads(out.stdout)
    assert got["epe"] == "0.0000" and got["bad_0_5"] == "0.0000"
    head, rows = _reference()
    row = next(r for r in rows if r[head.index("case")] == "slanted" and r[head.index("ablation_id")] == "A")
    assert row[head.index("epe")] != got["epe"]


@pytest.mark.parametrize("args,msg", [
    (["--manifest", "x", "--hierarchical"], "not part of this device build"),
    (["--manifest", "x", "--subpixel-method", "parabola"], "legacy only"),
    (["--manifest", "x", "--backend", "cuda"], "unrecognized --backend"),
    ([], "--manifest is required"),
    (["--manifest", "/nonexistent/manifest.txt"], "no valid test cases"),
])
def test_app_rejects_before_any_device_work(args, msg):
    _needs(APP)
    out = subprocess.run([APP] + args, capture_output=True, text=True, timeout=60)
    assert out.returncode == 1
    assert msg in out.stderr


TIMING = {"time_total_ms", "time_aux_ms", "time_prior_ms", "time_cost_ms", "time_right_wta_ms", "time_cross_ms",
          "time_sgm_ms", "time_wta_ms", "time_confidence_ms", "time_refine_ms", "time_post_ms"}
DEVICE_MEMORY = {"cost32_bytes", "peak_bytes", "peak_workspace_bytes", "retained_workspace_bytes"}


@pytest.mark.gpu
def test_benchmark_csv_matches_reference(tmp_path):
    _needs(APP)
    out_csv = tmp_path / "ablation.csv"
    print("apg-benchmark on tests/golden/bench_data, 3 cases x 7 configurations", flush=True)
    env = dict(os.environ, OMP_NUM_THREADS="4")
    r = subprocess.run([APP, "--manifest", os.path.join(DATA, "manifest.txt"), "--output", str(out_csv),
                        "--warmup", "0", "--repeat", "1"], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    head, rows = _rows(out_csv.read_text())
    rhead, rrows = _reference()
    assert head == rhead
    assert len(rows) == len(rrows) == 21
    bad = []
    for row, ref in zip(rows, rrows):
        assert len(row) == len(head)
        for k, a, b in zip(head, row, ref):
            skip = k in TIMING or k in DEVICE_MEMORY or (k == "cost_bytes" and row[0] == "packed")
            if not skip and a != b:
                bad.append((row[head.index("case")], row[head.index("ablation_id")], k, a, b))
    print("mismatching columns:", sorted({b[2] for b in bad}))
    assert not bad, bad[:40]
