#!/usr/bin/env python3
"""Writes the golden vectors of tests/golden/ from a build of the reference's
own sources (oracle/_ref/libapgref.so, built by `make -C oracle ref` from the
reference checkout).  Run where that build exists:

    python tests/golden/make_golden.py

Outputs
  golden_small.npz   full stage arrays for tiny pairs (inputs included)
  golden_hashes.json FNV-1a-64 of every stage output for larger pairs whose
                     inputs come from the seeded generator (input hashes are
                     stored too, so a generator change is detected)
Each case runs the reference classes stage by stage (C, C', S, WTA disp /
best / second) and the full StereoMatcher pipeline (disparity, confidence,
invalid_reason, reliable mask, right disparity, disparity before the
refiner, search range).  HighQuality cases run the prior (ranges narrowed
per pixel), 8 paths and the refiner; the pipeline goes through the
reference's packed volume, the stage classes through its dense one.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "apg-sgm_amd")]

import apg_amd as A  # noqa: E402  (generator only)
import pyoracle as O  # noqa: E402

STAGE_KEYS = ("cost", "cost_agg", "sgm", "disp_wta", "best", "second")
PIPE_KEYS = ("disparity", "confidence", "invalid_reason", "reliable", "disp_right", "d_before_refine", "dmin",
             "dmax")

# name: (generator kind, w, h, max_disp, seed, mode, overrides)
SMALL = {
    "bal_48x32_d16": (0, 48, 32, 16, 11, 1, {}),
    "fast_40x30_d12": (1, 40, 30, 12, 12, 0, {}),
    "bal_neg_44x28_d12": (2, 44, 28, 12, 13, 1, {"min_disparity": -4}),
    "p8_c64_36x30_d16": (1, 36, 30, 16, 14, 1, {"paths": 8, "census_type": 0}),
    # HighQuality: prior-narrowed ranges, 8 paths, refiner
    "hq_96x64_d24": (2, 96, 64, 24, 15, 2, {}),
    "hq_case_b_80x60_d16": (0, 80, 60, 16, 16, 2, {"prior_max_supports": 8, "prior_min_supports": 4}),
}
HASHED = {
    "c0_balanced_640x480_d64": (0, 640, 480, 64, 1, 1, {}),
    "c0_fast_640x480_d64": (0, 640, 480, 64, 1, 0, {}),
    "bal_320x240_d96": (1, 320, 240, 96, 2, 1, {}),
    "p8_256x200_d64": (2, 256, 200, 64, 3, 1, {"paths": 8}),
    "bal_c64_200x150_d40": (1, 200, 150, 40, 4, 1, {"census_type": 0}),
    "bal_odd_251x137_d57": (2, 251, 137, 57, 5, 1, {"piecewise_p2": 0, "agg_iterations": 2}),
    "hq_320x240_d64": (2, 320, 240, 64, 6, 2, {}),
    "hq_c64_256x180_d48": (1, 256, 180, 48, 7, 2, {"census_type": 0}),
    "hq_480x270_d128": (2, 480, 270, 128, 8, 2, {}),
}


def run_case(kind, w, h, D, seed, mode, over):
    left, right, _ = A.synth_pair(kind, w, h, D, seed)
    # D is the number of disparities; the range is [min_disparity, min_disparity + D)
    c = O.cfg(mode, over.get("min_disparity", 0) + D, **over)
    st = O.ref_stages(c, left, right)
    pp = O.ref_pipeline(c, left, right, backend=0)   # Auto: packed for HighQuality
    out = {k: st[k] for k in STAGE_KEYS}
    out.update({k: pp[k] for k in PIPE_KEYS})
    return left, right, c, out


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libapgref.so is not built (needs the reference sources)")
    O.ref().ref_max_threads()
    small = {}
    for name, (kind, w, h, D, seed, mode, over) in SMALL.items():
        left, right, c, out = run_case(kind, w, h, D, seed, mode, over)
        small[f"{name}__left"] = left
        small[f"{name}__right"] = right
        small[f"{name}__cfg"] = np.frombuffer(json.dumps(c.as_dict()).encode(), np.uint8)
        for k, v in out.items():
            small[f"{name}__{k}"] = v
    np.savez_compressed(os.path.join(HERE, "golden_small.npz"), **small)
    hashed = {}
    for name, (kind, w, h, D, seed, mode, over) in HASHED.items():
        left, right, c, out = run_case(kind, w, h, D, seed, mode, over)
        hashed[name] = {
            "generator": {"kind": kind, "w": w, "h": h, "max_disp": D, "seed": seed},
            "cfg": c.as_dict(),
            "input_fnv": {"left": hex(O.fnv1a64(left)), "right": hex(O.fnv1a64(right))},
            "fnv": {k: hex(O.fnv1a64(v)) for k, v in out.items()},
        }
        print(name, "done", flush=True)
    with open(os.path.join(HERE, "golden_hashes.json"), "w") as f:
        json.dump(hashed, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
