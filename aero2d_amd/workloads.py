"""Synthetic workloads of the batched implicit Euler path (seeded, deterministic).

    config1()   NACA 0012, 128 x 64, M 0.5, alpha 2 deg, wall-normal ratio 1.1 to 20 chords
    config2()   64 random NACA 4-digit sections on 256 x 128 (FP64 batch)
    config3()   32 NACA 0012/2412 transonic cases on 512 x 256 (mixed vs FP64)
    config4()   256 random sections on {128x64, 256x128, 384x192, 512x256}
    config5()   request generator for the serving loop (config-3 physics, sizes
                up to 768 x 384)

Grid recipe. ogrid marches nc layers over 20 chords with geometric ratio q;
the first spacing follows ds0 = 20 (q - 1)/(q^nc - 1). With q = q64^(64/nc) all
sizes are geometrically alike (q64 = 1.1 is the 128 x 64 grid of config 1).
The marching generator's outer-circle snap inverts the last layer when that
layer gets thin against the deviation of the marched layer from a circle, and
highly cambered sections can cross normals at high resolution; grid_for()
therefore tries q64 = 1.1, 1.12, 1.15 in that order and keeps the first grid
whose cells all have positive area (the generator's own check), recording q64.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .batch import Case
from .geometry import Airfoil, naca4, naca4_params
from .grid import OGrid, ogrid, self_similar_spacing

__all__ = ["grid_for", "config1", "config2", "config3", "config4", "config5", "Request",
           "SEED_BASE", "Q64_LADDER"]

SEED_BASE = 20261203
Q64_LADDER = (1.1, 1.12, 1.15)
RADIUS = 20.0


def grid_for(af: Airfoil, nc: int, q64_ladder=Q64_LADDER) -> tuple[OGrid, float]:
    """O-grid with nc wall-normal cells to 20 chords; returns (grid, q64 used).
    Raises RuntimeError when every ratio in the ladder gives inverted cells."""
    last = None
    for q64 in q64_ladder:
        try:
            return ogrid(af, nj=nc + 1, ds0=self_similar_spacing(nc, RADIUS, q64), radius=RADIUS), q64
        except RuntimeError as e:
            last = e
    raise last


def config1() -> list[Case]:
    """Plain single case: NACA 0012 on 128 x 64, M 0.5, alpha 2 deg, ratio 1.1 to
    20 chords (ds0 = 4.4965e-3)."""
    g = ogrid(naca4("0012", 65), nj=65, ds0=self_similar_spacing(64, RADIUS, 1.1), radius=RADIUS)
    return [Case(g, 0.5, 2.0, "NACA 0012")]


def config2(n: int = 64, ni: int = 256, nc: int = 128, seed: int = SEED_BASE + 2) -> list[Case]:
    """Uniform batch: m ~ U[0, 4]%, p ~ U[0.3, 0.5], t ~ U[9, 15]%,
    M ~ U[0.3, 0.7], alpha ~ U[-4, 6] deg, all on ni x nc."""
    rng = np.random.default_rng(seed)
    cases = []
    for k in range(n):
        m, p, t = rng.uniform(0, 0.04), rng.uniform(0.3, 0.5), rng.uniform(0.09, 0.15)
        M, a = rng.uniform(0.3, 0.7), rng.uniform(-4.0, 6.0)
        g, _ = grid_for(naca4_params(m, p, t, ni // 2 + 1), nc)
        cases.append(Case(g, M, a, f"m{m:.3f} p{p:.2f} t{t:.3f}"))
    return cases


def config3(n: int = 32, ni: int = 512, nc: int = 256, seed: int = SEED_BASE + 3) -> list[Case]:
    """Transonic batch: NACA 0012 or 2412 (seeded coin), M ~ U[0.72, 0.82],
    alpha ~ U[0, 2.5] deg, on ni x nc."""
    rng = np.random.default_rng(seed)
    cases = []
    grids = {}
    for k in range(n):
        code = "0012" if rng.uniform() < 0.5 else "2412"
        M, a = rng.uniform(0.72, 0.82), rng.uniform(0.0, 2.5)
        if code not in grids:
            grids[code] = grid_for(naca4(code, ni // 2 + 1), nc)[0]
        cases.append(Case(grids[code], M, a, f"NACA {code}"))
    return cases


CONFIG4_SIZES = ((128, 64), (256, 128), (384, 192), (512, 256))


def config4(n: int = 256, seed: int = SEED_BASE + 4, sizes=CONFIG4_SIZES):
    """Uneven batch: size uniform over `sizes`, m ~ U[0, 6]%, p ~ U[0.2, 0.6],
    t ~ U[8, 18]%, M ~ U[0.3, 0.8], alpha ~ U[-4, 8] deg. A draw whose grid
    inverts at every ratio of the ladder is redrawn; returns (cases, redraws)."""
    rng = np.random.default_rng(seed)
    cases, redraws = [], 0
    while len(cases) < n:
        ni, nc = sizes[rng.integers(len(sizes))]
        m, p, t = rng.uniform(0, 0.06), rng.uniform(0.2, 0.6), rng.uniform(0.08, 0.18)
        M, a = rng.uniform(0.3, 0.8), rng.uniform(-4.0, 8.0)
        try:
            g, _ = grid_for(naca4_params(m, p, t, ni // 2 + 1), nc)
        except RuntimeError:
            redraws += 1
            continue
        cases.append(Case(g, M, a, f"{ni}x{nc} m{m:.3f} p{p:.2f} t{t:.3f}"))
    return cases, redraws


CONFIG5_SIZES = ((128, 64), (256, 128), (384, 192), (512, 256), (768, 384))


@dataclass
class Request:
    """One serving request: arrival time (s from the stream start) and its case."""
    t_arrive: float
    case: Case


def config5(rate: float, n: int, seed: int = SEED_BASE + 5, sizes=CONFIG5_SIZES) -> list[Request]:
    """Open-loop Poisson stream of n single-case requests at `rate` per second:
    exponential inter-arrival times, config-3 physics (NACA 0012/2412,
    M ~ U[0.72, 0.82], alpha ~ U[0, 2.5]), size uniform over `sizes`."""
    rng = np.random.default_rng(seed + int(rate * 1000))
    t = 0.0
    grids = {}
    out = []
    for k in range(n):
        t += rng.exponential(1.0 / rate)
        ni, nc = sizes[rng.integers(len(sizes))]
        code = "0012" if rng.uniform() < 0.5 else "2412"
        M, a = rng.uniform(0.72, 0.82), rng.uniform(0.0, 2.5)
        key = (code, ni, nc)
        if key not in grids:
            grids[key] = grid_for(naca4(code, ni // 2 + 1), nc)[0]
        out.append(Request(t, Case(grids[key], M, a, f"NACA {code} {ni}x{nc}")))
    return out
