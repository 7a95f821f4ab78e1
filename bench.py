#!/usr/bin/env python3
"""Benchmark of the MI355X dense-stereo hot path.

Workload (N=1): Balanced mode, one 1280x720 pair, D=128 (the synthetic
ramp + 8 rectangles + gain/bias generator), inputs resident in HBM.  One step =
one full device pipeline (aux, range, cost, right WTA, Cross H+V, 4-path SGM,
WTA + subpixel, confidence, LR check, fill, median).  Timed with HIP events on
the library's stream.  Prints ONE JSON line.

The CPU baseline (reported, not the target) is the reference implementation
compiled from its own sources (oracle/_ref) when present, else the repo's C
oracle, on the same pair with all host cores: one parity run (bit-exact
against the GPU disparity), then about 10 s of repeated runs, median.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "apg-sgm_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
SEED = 20260101
W, H, DMAX = 1280, 720, 128
# Algorithmic HBM bytes per disparity state (compulsory stage-boundary volume
# traffic of u16 volumes): cost write 2 + Cross read/write 4 + SGM read C' /
# write S 4 + WTA read S 2 = 12 for the pipeline.  Per timed stage of the
# fused schedule: "cost" = cost + right-WTA candidates + Cross H (writes T: 2),
# "cross" = Cross V (reads T, writes C': 4), "sgm" = all path scans (reads C',
# writes S: 4), "wta" = WTA + confidence + LR (reads S: 2).
BYTES_PER_STATE = {"pipeline": 12.0, "sgm": 4.0, "cross": 4.0, "cost": 2.0, "wta": 2.0}


def cpu_baseline(left, right, p, gpu_disp):
    """The reference pipeline (oracle/_ref, compiled from the reference's
    sources) on the bench pair, OpenMP over all host threads: one run for
    parity with the GPU disparity, then repeated runs for about 10 s, median
    per-pair time.  The repo's C oracle stands in when the reference build is
    absent."""
    import pyoracle as O
    c = O.cfg_from(p.as_dict())
    states = left.shape[0] * left.shape[1] * (p.max_disparity - p.min_disparity)
    t0 = time.perf_counter()
    if O.ref_available():
        out = O.ref_pipeline(c, left, right)
        kind, threads = "reference", O.ref_max_threads()
    else:
        threads = os.cpu_count() or 1
        O.set_threads(threads)
        out = O.pipeline(c, left, right)
        kind = "port"
    t1 = time.perf_counter() - t0
    parity = bool(np.array_equal(out["disparity"], gpu_disp))
    reps = int(min(60, max(3, round(10.0 / max(t1, 1e-3)))))
    if kind == "reference":
        ms = O.ref_time_pipeline(c, left, right, threads=0, warmup=0, reps=reps)
    else:
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            O.pipeline(c, left, right)
            ts.append((time.perf_counter() - t) * 1e3)
        ms = float(sorted(ts)[len(ts) // 2])
    return {"value": round(states / (ms * 1e-3) / 1e6, 1), "unit": "Mdisp/s", "cores": threads, "kind": kind,
            "sample": f"Balanced pipeline on the bench pair {left.shape[1]}x{left.shape[0]} D={DMAX}: median of "
                      f"{reps} runs ({ms:.1f} ms per pair) after one parity run, OpenMP over {threads} host threads"
            }, parity


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_balanced_1280x720_d128.md")
STAGE_KERNEL = {"cost": "cost_crossh2_kernel", "cross": "cross_v3_kernel", "sgm": "sgm_lines4_kernel",
                "wta": "wta_fused_kernel"}


def measured_traffic(stage):
    """HBM bytes per launch of the stage's kernel from the committed counter
    summary (FETCH_SIZE + WRITE_SIZE runs, scaled by the copy-kernel
    calibration), or None when the summary has no line for it."""
    try:
        lines = open(TRAFFIC_FILE).read().splitlines()
    except OSError:
        return None
    for ln in lines:
        f = ln.split()
        if len(f) == 2 and f[0] == STAGE_KERNEL.get(stage) and f[1].isdigit():
            return int(f[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--schedule", type=int, default=0, help="0 fused (default), 1 stage by stage")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MB",
                    help="first run one device-to-device copy of MB MiB (a kernel of known traffic, "
                         "to scale the FETCH_SIZE / WRITE_SIZE counters of a profiled run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs one GPU per process; --gpus {args.gpus} is not supported "
              "(pairs are independent; multi-GPU would be host-side dispatch only)", file=sys.stderr)
        sys.exit(2)

    if args.copy_calib > 0:
        import torch
        n = args.copy_calib << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty(n, dtype=torch.uint8, device="cuda")
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    import apg_amd as A
    p = A.params(1, DMAX)
    left, right, _ = A.synth_pair(1, W, H, DMAX, SEED)
    states = W * H * DMAX
    ctx = A.Context(0)
    A.set_schedule(ctx, args.schedule)
    dl, dr = ctx.to_device(left), ctx.to_device(right)
    od = ctx.array((H, W), np.float32)
    oc = ctx.array((H, W), np.float32)
    orr = ctx.array((H, W), np.uint8)
    om = ctx.array((H, W), np.uint8)
    ms = C.c_double()
    tim = A.Timing()
    rc = ctx.lib.apg_hip_bench_pipeline(ctx.h, C.byref(p), dl.ptr, dr.ptr, W, H, W, 1, od.ptr, oc.ptr, orr.ptr,
                                        om.ptr, args.warmup, args.steps, C.byref(ms), C.byref(tim))
    ctx._check(rc, "bench")
    ms_step = ms.value / args.steps
    gpu_disp = od.download()
    value = states / (ms_step * 1e-3) / 1e6

    stages = {k: getattr(tim, k) for k in ("aux_ms", "prior_ms", "cost_ms", "right_wta_ms", "cross_ms",
                                            "sgm_ms", "wta_ms", "confidence_ms", "post_ms", "total_ms")}
    dom = max(("cost", "cross", "sgm", "wta"), key=lambda k: stages[k + "_ms"])
    dom_ms = stages[dom + "_ms"]
    achieved = states * BYTES_PER_STATE[dom] / (dom_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": STAGE_KERNEL[dom], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom),
                "pipeline_frac": round(states * BYTES_PER_STATE["pipeline"] / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    out = {"metric": "Mdisp/s (W*H*D per sec), Balanced 1280x720 D=128", "value": round(value, 1),
           "unit": "Mdisp/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "u16", "data": "synthetic",
           "config": {"workload": "Balanced single pair 1280x720 D=128 (ramp + 8 rects, gain/bias)",
                      "width": W, "height": H, "disparities": DMAX, "mode": "Balanced", "seed": SEED},
           "stage_ms": {k: round(v, 4) for k, v in stages.items()}, "roofline": roofline}
    if not args.no_cpu:
        cb, parity = cpu_baseline(left, right, p, gpu_disp)
        out["cpu_baseline"] = cb
        out["parity_vs_cpu"] = parity
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
