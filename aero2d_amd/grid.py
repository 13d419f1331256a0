"""Structured O-grid around an airfoil (host side, NumPy FP64).

Hyperbolic marching: layers grow along smoothed normals with a geometric
wall-normal spacing and along-layer smoothing; the outer layer is snapped to a
circle about (0.5, 0). Index i runs counter-clockwise along the contour
(periodic), j runs from the wall outward. Behaviour follows aero2d/grid.py:15-81.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .geometry import Airfoil

__all__ = ["OGrid", "ogrid", "growth_ratio", "self_similar_spacing"]


@dataclass
class OGrid:
    """Node coordinates of shape (ni, nj) (aero2d/grid.py:15-31)."""
    x: np.ndarray
    y: np.ndarray

    @property
    def shape(self):
        return self.x.shape

    def cell_areas(self) -> np.ndarray:
        """Cell areas (ni, nj-1) from the cross product of the two diagonals.
        The node loop (i,j)->(i+1,j)->(i+1,j+1)->(i,j+1) is clockwise, hence the
        minus sign (aero2d/grid.py:24-31)."""
        x, y = self.x, self.y
        x1, y1 = np.roll(x, -1, 0), np.roll(y, -1, 0)
        d1x, d1y = x1[:, 1:] - x[:, :-1], y1[:, 1:] - y[:, :-1]
        d2x, d2y = x[:, 1:] - x1[:, :-1], y[:, 1:] - y1[:, :-1]
        return -0.5 * (d1x * d2y - d1y * d2x)


def growth_ratio(ds0: float, total: float, n: int) -> float:
    """Ratio r with ds0 (r^n - 1)/(r - 1) = total, by 200 bisections on
    [1 + 1e-9, 2] (aero2d/grid.py:34-41)."""
    lo, hi = 1.0 + 1e-9, 2.0
    for _ in range(200):
        r = 0.5 * (lo + hi)
        s = ds0 * (r**n - 1) / (r - 1)
        lo, hi = (r, hi) if s < total else (lo, r)
    return 0.5 * (lo + hi)


def _smooth_periodic(a: np.ndarray, eps: float, passes: int) -> np.ndarray:
    """Explicit Laplacian smoothing along the periodic i axis (aero2d/grid.py:44-47)."""
    for _ in range(passes):
        a = a + eps * (np.roll(a, 1, 0) + np.roll(a, -1, 0) - 2 * a)
    return a


def ogrid(af: Airfoil, nj: int = 64, ds0: float = 2e-3, radius: float = 40.0,
          normal_smoothing: int = 20) -> OGrid:
    """March an O-grid with nj node layers from the airfoil contour (aero2d/grid.py:50-81).

    `radius` is the total marched wall-normal distance. Raises RuntimeError when
    any cell has non-positive area (aero2d/grid.py:77-80)."""
    pts = af.coords()[:-1].astype(float)          # drop the repeated trailing-edge point
    ni = len(pts)
    center = np.array([0.5, 0.0])
    r = growth_ratio(ds0, radius, nj - 1)
    d = ds0 * r ** np.arange(nj - 1)
    X = np.zeros((ni, nj, 2))
    X[:, 0] = pts
    for j in range(1, nj):
        cur = X[:, j - 1]
        t = np.roll(cur, -1, 0) - np.roll(cur, 1, 0)
        t /= np.linalg.norm(t, axis=1, keepdims=True)
        n = np.column_stack([t[:, 1], -t[:, 0]])
        # normals are smoothed harder away from the first layer
        n = _smooth_periodic(n, 0.25, normal_smoothing if j > 1 else 2)
        n /= np.linalg.norm(n, axis=1, keepdims=True)
        new = cur + d[j - 1] * n
        frac = j / (nj - 1)
        new = _smooth_periodic(new, 0.1 + 0.3 * frac, 2 + int(8 * frac))
        X[:, j] = new
    # outer layer: an exact circle, polar angles kept
    rel = X[:, -1] - center
    ang = np.arctan2(rel[:, 1], rel[:, 0])
    rmean = np.mean(np.linalg.norm(rel, axis=1))
    X[:, -1] = center + rmean * np.column_stack([np.cos(ang), np.sin(ang)])
    g = OGrid(X[..., 0], X[..., 1])
    a = g.cell_areas()
    if a.min() <= 0:
        raise RuntimeError(f"O-grid has inverted cells: min area {a.min():.3e}")
    return g


def self_similar_spacing(nc: int, radius: float = 20.0, q64: float = 1.1) -> float:
    """First wall spacing ds0 for nc cells over `radius` with ratio q = q64^(64/nc).

    A fixed ratio of 1.1 over 20 chords is only usable near nc = 64 (at nc = 128
    it gives ds0 ~ 1e-5 and the marching grid inverts). Scaling the ratio as
    q64^(64/nc) keeps the grids of every size geometrically alike:
    ds0 = radius (q - 1)/(q^nc - 1)."""
    q = q64 ** (64.0 / nc)
    return radius * (q - 1.0) / (q**nc - 1.0)
