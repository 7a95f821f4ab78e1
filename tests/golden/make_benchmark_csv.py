"""Writes the benchmark-app fixture: a small synthetic stereo dataset
(tests/golden/bench_data: PGM pairs, PFM ground truth, a visibility mask and a
manifest) and the CSV that the reference's own benchmark app writes for it
(reference apps/benchmark_stereo.cpp; binary built from the reference sources
by `make -C oracle ref`), tests/golden/bench_reference.csv, plus the
reference's disparity maps of configurations A and D for the two cases with
ground truth (bench_data/ref_<case>_<id>_disp.pfm), on which the accuracy
metrics are checked without a GPU.

    python tests/golden/make_benchmark_csv.py

The scene: a textured right image; the left image samples it at x - d(x, y)
(linear interpolation) with d a slanted background plane, a fronto-parallel
box and a disc, plus a flat low-texture patch, so that every metric region of
the app's CSV (occluded, boundary, low texture ...) is populated.
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DATA = os.path.join(HERE, "bench_data")
REF_APP = os.path.join(ROOT, "oracle", "_ref", "apg-benchmark")


def write_pgm(path, img):
    h, w = img.shape
    with open(path, "wb") as f:
        f.write(b"P5\n%d %d\n255\n" % (w, h))
        f.write(np.ascontiguousarray(img, dtype=np.uint8).tobytes())


def write_pfm(path, img):
    """Grayscale little-endian PFM, rows stored bottom-up."""
    h, w = img.shape
    with open(path, "wb") as f:
        f.write(b"Pf\n%d %d\n-1.0\n" % (w, h))
        f.write(np.ascontiguousarray(img[::-1], dtype="<f4").tobytes())


def scene(w, h, seed):
    rng = np.random.default_rng(seed)
    # right view: band-limited noise (2x2 blocks upsampled) plus fine noise
    coarse = rng.integers(0, 256, size=(h // 2 + 1, w // 2 + 1)).astype(np.float32)
    right = np.repeat(np.repeat(coarse, 2, 0), 2, 1)[:h, :w] * 0.7 + rng.integers(0, 77, size=(h, w))
    right[h // 8 : h // 8 + 12, w // 2 : w // 2 + 20] = 128.0   # low-texture patch
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    d = 6.0 + 0.04 * xx + 0.02 * yy                              # slanted background
    box = (xx >= w * 0.55) & (xx < w * 0.8) & (yy >= h * 0.3) & (yy < h * 0.7)
    d[box] = 18.0
    disc = (xx - w * 0.3) ** 2 + (yy - h * 0.6) ** 2 < (h * 0.18) ** 2
    d[disc] = 12.5
    xs = xx - d
    x0 = np.floor(xs).astype(np.int64)
    fr = xs - x0
    ok = (x0 >= 0) & (x0 + 1 < w)
    x0c = np.clip(x0, 0, w - 1)
    x1c = np.clip(x0 + 1, 0, w - 1)
    rows = np.arange(h)[:, None]
    left = right[rows, x0c] * (1 - fr) + right[rows, x1c] * fr
    left = np.where(ok, left, rng.integers(0, 256, size=(h, w)))
    left = np.clip(np.rint(left + rng.normal(0, 1.5, size=(h, w))), 0, 255).astype(np.uint8)
    right = np.clip(np.rint(right), 0, 255).astype(np.uint8)
    gt = d.astype(np.float32)
    gt[~ok] = np.inf                      # not matchable: no ground truth
    gt[: 2, :] = -1.0                     # a couple of rows without ground truth
    vis = np.where(ok, 255, 0).astype(np.uint8)
    vis[:, : int(w * 0.1)] = 100          # a band the mask calls not visible
    return left, right, gt, vis


def main():
    os.makedirs(DATA, exist_ok=True)
    lines = ["# name left right gt [dmax [dmin [vis]]]"]
    for name, (w, h, seed, dmax, dmin, with_vis, with_gt) in {
        "slanted_vis": (128, 80, 7, 32, 0, True, True),
        "slanted": (112, 72, 11, 48, 0, False, True),
        "nogt": (96, 64, 13, 32, 0, False, False),
    }.items():
        left, right, gt, vis = scene(w, h, seed)
        write_pgm(os.path.join(DATA, name + "_left.pgm"), left)
        write_pgm(os.path.join(DATA, name + "_right.pgm"), right)
        gt_name = name + "_gt.pfm" if with_gt else "absent_gt.pfm"
        if with_gt:
            write_pfm(os.path.join(DATA, gt_name), gt)
        line = "%s %s_left.pgm %s_right.pgm %s %d %d" % (name, name, name, gt_name, dmax, dmin)
        if with_vis:
            write_pgm(os.path.join(DATA, name + "_vis.pgm"), vis)
            line += " %s_vis.pgm" % name
        lines.append(line)
    with open(os.path.join(DATA, "manifest.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    if not os.path.exists(REF_APP):
        sys.exit("build the reference app first: make -C oracle ref")
    out = os.path.join(HERE, "bench_reference.csv")
    env = dict(os.environ, OMP_NUM_THREADS="4")
    subprocess.run([REF_APP, "--manifest", os.path.join(DATA, "manifest.txt"), "--output", out,
                    "--warmup", "0", "--repeat", "1"], check=True, env=env)
    with tempfile.TemporaryDirectory() as tmp:
        subprocess.run([REF_APP, "--manifest", os.path.join(DATA, "manifest.txt"), "--output",
                        os.path.join(tmp, "ad.csv"), "--ablation", "A,D", "--save_disp", tmp, "--warmup", "0",
                        "--repeat", "1"], check=True, env=env, stdout=subprocess.DEVNULL)
        for case in ("slanted_vis", "slanted"):
            for aid in ("A", "D"):
                shutil.copy(os.path.join(tmp, "%s_%s_disp.pfm" % (case, aid)),
                            os.path.join(DATA, "ref_%s_%s_disp.pfm" % (case, aid)))
    print("wrote", out)


if __name__ == "__main__":
    main()
