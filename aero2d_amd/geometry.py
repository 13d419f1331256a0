"""Airfoil geometry for the implicit Euler path: the NACA 4-digit family.

Contour convention (XFOIL's, kept from aero2d/geometry.py:1-6): start at the
trailing edge, run over the upper surface to the leading edge and back over the
lower surface to the trailing edge. First and last points are the trailing edge.
Chord 1, leading edge at (0, 0), trailing edge at (1, 0).

Host-side input generation (NumPy, FP64): nothing here touches the GPU.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

__all__ = ["Airfoil", "naca4", "naca4_params", "cosine_spacing", "NACA_TMR_X0"]


@dataclass
class Airfoil:
    """Closed contour (aero2d/geometry.py:15-35)."""
    x: np.ndarray
    y: np.ndarray
    name: str = "airfoil"

    @property
    def n(self) -> int:
        return len(self.x)

    def coords(self) -> np.ndarray:
        return np.column_stack([self.x, self.y])

    @staticmethod
    def from_file(path: str) -> "Airfoil":
        """Selig format: a name line, then `x y` pairs (aero2d/geometry.py:28-35)."""
        with open(path) as f:
            lines = [ln.strip() for ln in f if ln.strip()]
        pts = np.array([[float(v) for v in ln.split()[:2]] for ln in lines[1:]])
        return Airfoil(pts[:, 0], pts[:, 1], lines[0])


def cosine_spacing(n: int) -> np.ndarray:
    """x = (1 - cos(beta)) / 2 for beta evenly spaced on [0, pi] (aero2d/geometry.py:41-43)."""
    beta = np.linspace(0.0, np.pi, n)
    return 0.5 * (1.0 - np.cos(beta))


# Zero of the original NACA half-thickness polynomial (a4 = -0.1015).
NACA_TMR_X0 = 1.008930411365


def naca4_params(m: float, p: float, t: float, n_side: int = 100, closed_te: bool = True,
                 scaled_te: bool = False, name: str | None = None) -> Airfoil:
    """NACA 4-digit section from real-valued camber m, camber position p and
    thickness t (fractions of chord), with `n_side` points per side.

    The same equations as aero2d/geometry.py:49-87, taking (m, p, t) directly so
    that random draws need not be rounded to digits; `naca4(code)` calls this with
    m = d0/100, p = d1/10, t = d23/100 and gives identical points.
    closed_te replaces a4 with -0.1036 (sharp trailing edge); scaled_te instead
    extends the a4 = -0.1015 polynomial to its zero at x = 1.00893 and rescales.
    """
    xc = cosine_spacing(n_side)
    if scaled_te:
        xs = xc * NACA_TMR_X0
        a4 = -0.1015
        yt = 5 * t * (0.2969 * np.sqrt(xs) - 0.1260 * xs - 0.3516 * xs**2
                      + 0.2843 * xs**3 + a4 * xs**4) / NACA_TMR_X0
        yt[-1] = 0.0
        closed_te = True
    else:
        a4 = -0.1036 if closed_te else -0.1015
        yt = 5 * t * (0.2969 * np.sqrt(xc) - 0.1260 * xc - 0.3516 * xc**2
                      + 0.2843 * xc**3 + a4 * xc**4)
    yc = np.zeros_like(xc)
    dyc = np.zeros_like(xc)
    if m > 0 and p > 0:
        fore = xc < p
        yc[fore] = m / p**2 * (2 * p * xc[fore] - xc[fore] ** 2)
        dyc[fore] = 2 * m / p**2 * (p - xc[fore])
        aft = ~fore
        yc[aft] = m / (1 - p) ** 2 * ((1 - 2 * p) + 2 * p * xc[aft] - xc[aft] ** 2)
        dyc[aft] = 2 * m / (1 - p) ** 2 * (p - xc[aft])
    th = np.arctan(dyc)
    xu, yu = xc - yt * np.sin(th), yc + yt * np.cos(th)
    xl, yl = xc + yt * np.sin(th), yc - yt * np.cos(th)
    x = np.concatenate([xu[::-1], xl[1:]])
    y = np.concatenate([yu[::-1], yl[1:]])
    if closed_te:
        y[0] = y[-1] = 0.5 * (y[0] + y[-1])
    if name is None:
        name = f"NACA m={m:.4g} p={p:.4g} t={t:.4g}"
    return Airfoil(x, y, name)


def naca4(code: str, n_side: int = 100, closed_te: bool = True, scaled_te: bool = False) -> Airfoil:
    """NACA 4-digit airfoil from its code, e.g. "0012" (aero2d/geometry.py:49)."""
    m = int(code[0]) / 100.0
    p = int(code[1]) / 10.0
    t = int(code[2:]) / 100.0
    return naca4_params(m, p, t, n_side, closed_te, scaled_te, name=f"NACA {code}")
