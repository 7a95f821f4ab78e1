#!/usr/bin/env python3
"""Benchmark of the MI355X dense-stereo hot path.

Workload (N=1): Balanced mode, one 1280x720 pair, D=128 (the synthetic
ramp + 8 rectangles + gain/bias generator), inputs resident in HBM.  One step =
one full device pipeline (aux, range, cost, right WTA, Cross H+V, 4-path SGM,
WTA + subpixel, confidence, LR check, fill, median).  Timed with HIP events on
the library's stream.  Prints ONE JSON line.

The CPU baseline (reported, not the target) is the reference implementation
compiled from its own sources (oracle/_ref) when present, else the repo's C
oracle, run once on the same pair with all host cores.  The GPU disparity of
the bench pair is compared bit-exactly with that CPU run.
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
    import pyoracle as O
    c = O.cfg_from(p.as_dict())
    cores = os.cpu_count() or 1
    t0 = time.perf_counter()
    if O.ref_available():
        out = O.ref_pipeline(c, left, right)
        kind = "reference"
    else:
        O.set_threads(cores)
        out = O.pipeline(c, left, right)
        kind = "port"
    dt = time.perf_counter() - t0
    states = left.shape[0] * left.shape[1] * (p.max_disparity - p.min_disparity)
    parity = bool(np.array_equal(out["disparity"], gpu_disp))
    return {"value": states / dt / 1e6, "unit": "Mdisp/s", "cores": cores, "kind": kind,
            "sample": f"1 Balanced pipeline on the bench pair {left.shape[1]}x{left.shape[0]} D={DMAX}, "
                      f"{dt:.2f} s wall, OpenMP over {cores} host threads"}, parity


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
    roofline = {"bound": "hbm", "kernel": f"{dom} stage", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
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
