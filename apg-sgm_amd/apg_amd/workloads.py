"""Seeded synthetic workloads of the mixed-size forms.

* batch_specs: the uneven batch (64 pairs, W uniform in [320, 1920] step 16,
  H in [240, 1080] step 8, D from {64, 128, 192, 256} restricted to D < W/2,
  Balanced; pair i generated with seed + i by the noise + 4 rectangles
  generator).
* serve_specs: Poisson arrivals at a fixed rate, size mix 60% 640x480 D64 /
  30% 1280x720 D128 / 10% 1920x1080 D256, modes 70% Balanced / 30% Fast.

Both draw from an integer-only splitmix64 stream, so every host reproduces
the same lists.
"""
from __future__ import annotations

import math

SEED_BATCH = 20260202
SEED_SERVE = 20260303

SERVE_CLASSES = [(640, 480, 64), (1280, 720, 128), (1920, 1080, 256)]


def splitmix64(state):
    """Integer-only seeded stream (splitmix64): yields 64-bit values."""
    m = (1 << 64) - 1
    while True:
        state = (state + 0x9E3779B97F4A7C15) & m
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
        yield z ^ (z >> 31)


def batch_specs(n=64, seed=SEED_BATCH):
    """[(w, h, D, pair_seed)] of the uneven batch."""
    rng = splitmix64(seed)
    specs = []
    for i in range(n):
        w = 320 + 16 * (next(rng) % 101)
        h = 240 + 8 * (next(rng) % 106)
        ds = [d for d in (64, 128, 192, 256) if d < w / 2]
        specs.append((w, h, ds[next(rng) % len(ds)], seed + i))
    return specs


def serve_specs(n, rate, seed):
    """[(arrival_s, class, mode, variant)]: exponential gaps at `rate` per
    second, class index into SERVE_CLASSES, mode 1 Balanced / 0 Fast, one of
    four input variants per class."""
    rng = splitmix64(seed)
    t, out = 0.0, []
    for _ in range(n):
        u = (next(rng) >> 11) * (1.0 / (1 << 53))
        t += -math.log1p(-u) / rate
        c = next(rng) % 10
        cls = 0 if c < 6 else (1 if c < 9 else 2)
        mode = 1 if next(rng) % 10 < 7 else 0
        out.append((t, cls, mode, next(rng) % 4))
    return out
