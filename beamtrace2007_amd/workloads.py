"""Job draws of the batched and served soft-shadow workloads.

Batched (64 jobs in one persistent launch): each job picks scene A (the
noisy icosphere blob, ~345K triangles) or B (the atrium, ~66K) by a seeded
coin, a square image side uniform in 128..1024, and a light scale
log-uniform in [0.25, 4] that multiplies the area light's U and V vectors as
bt_render's --light-scale does (apps/bt_render.cpp:289-290).

Served (a Poisson stream at a fixed rate): 70% soft-shadow and 30%
point-shadow requests on the conference hall, image sides uniform in
256..1024.

Draws use numpy's PCG64 generator with the given seed, so they are the same
on every machine.
"""
from __future__ import annotations

import math

import numpy as np

BATCH_SCENES = ("blob", "atrium")


def scaled_light(light9, scale: float) -> np.ndarray:
    """The area light with U and V scaled in float32, as the reference does
    (L.U = cfg.lightU * Real(lightScale))."""
    L = np.array(light9, np.float32).copy()
    L[3:] = L[3:] * np.float32(scale)
    return L


def batch_draw(n: int = 64, seed: int = 64, side_lo: int = 128, side_hi: int = 1024) -> list[dict]:
    """[{scene, side, light_scale}] for the batched workload."""
    g = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(n):
        scene = BATCH_SCENES[0] if g.random() < 0.5 else BATCH_SCENES[1]
        side = int(g.integers(side_lo, side_hi + 1))
        scale = float(math.exp(g.uniform(math.log(0.25), math.log(4.0))))
        out.append({"scene": scene, "side": side, "light_scale": scale})
    return out


def served_draw(n: int, qps: float, seed: int = 5, side_lo: int = 256, side_hi: int = 1024) -> list[dict]:
    """[{t, kind, side}] arrival times (s) of a Poisson stream at `qps`."""
    g = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for _ in range(n):
        t += float(g.exponential(1.0 / qps))
        kind = "soft" if g.random() < 0.7 else "point"
        side = int(g.integers(side_lo, side_hi + 1))
        out.append({"t": t, "kind": kind, "side": side})
    return out
