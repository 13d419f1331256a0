"""Airfoil geometry for the implicit Euler path: the NACA 4-digit family and
contours from coordinate files (Selig or Lednicer), normalised, smoothed and
re-distributed for the grid generator.

Contour convention (XFOIL's, kept from aero2d/geometry.py:1-6): start at the
trailing edge, run over the upper surface to the leading edge and back over the
lower surface to the trailing edge. First and last points are the trailing edge.
Chord 1, leading edge at (0, 0), trailing edge at (1, 0).

Host-side input generation (NumPy, FP64): nothing here touches the GPU.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

__all__ = ["Airfoil", "naca4", "naca4_params", "cosine_spacing", "NACA_TMR_X0", "normalize", "repanel",
           "smooth_airfoil", "load_airfoil"]


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


# ---------------------------------------------------------------------------
# Contours from coordinate files (aero2d/geometry.py:244-448)
# ---------------------------------------------------------------------------
def _unit_chord(x, y):
    """Scale and rotate so that the trailing edge (midpoint of the contour's
    ends) lands on (1, 0) and the point farthest from it, the leading edge, on
    (0, 0) (aero2d/geometry.py:272-284)."""
    te = np.array([0.5 * (x[0] + x[-1]), 0.5 * (y[0] + y[-1])])
    le_i = int(np.argmax(np.hypot(x - te[0], y - te[1])))
    le = np.array([x[le_i], y[le_i]])
    chord = te - le
    length = float(np.hypot(chord[0], chord[1]))
    ang = float(np.arctan2(chord[1], chord[0]))
    c, s = np.cos(-ang), np.sin(-ang)
    u, v = (x - le[0]) / length, (y - le[1]) / length
    return u * c - v * s, u * s + v * c


def _signed_area(x, y):
    return 0.5 * float(np.sum(x[:-1] * y[1:] - x[1:] * y[:-1]))


def normalize(af: Airfoil) -> Airfoil:
    """The package's contour convention for any input (aero2d/geometry.py:244-269):
    a Lednicer layout (two halves, each from the leading to the trailing edge)
    is joined into one contour; coordinates that are not already at unit chord
    (leading edge at x = 0 and trailing edge at x = 1, within 2e-3) are scaled
    and rotated to it, and the contour is made counter-clockwise (positive
    signed area: trailing edge, upper surface, leading edge, lower surface)."""
    x, y = np.asarray(af.x, float), np.asarray(af.y, float)
    jumps = np.nonzero(np.diff(x) < -0.5)[0]
    if x[0] < 0.5 and len(jumps):
        k = int(jumps[0]) + 1                      # start of the second half
        xu, yu = x[:k][::-1], y[:k][::-1]          # upper half, trailing edge first
        lo = slice(k + 1, None) if np.isclose(x[k], x[0]) else slice(k, None)
        x, y = np.concatenate([xu, x[lo]]), np.concatenate([yu, y[lo]])
    at_unit = abs(x.min()) < 2e-3 and abs(x.max() - 1.0) < 2e-3 and abs(0.5 * (y[0] + y[-1])) < 2e-3
    if not at_unit:
        x, y = _unit_chord(x, y)
    if _signed_area(x, y) < 0:
        x, y = x[::-1], y[::-1]
    return Airfoil(x, y, af.name)


def _turning_points(x, y, angle_deg):
    """Interior nodes where the contour's direction turns by more than angle_deg."""
    t = np.column_stack([np.diff(x), np.diff(y)])
    t /= np.maximum(np.linalg.norm(t, axis=1), 1e-300)[:, None]
    cosang = np.clip(np.einsum("ij,ij->i", t[:-1], t[1:]), -1.0, 1.0)
    return np.nonzero(np.degrees(np.arccos(cosang)) > angle_deg)[0] + 1


class _ArcSpline:
    """x(s), y(s) over arc length, one interpolating spline per smooth piece:
    the contour is cut at the given node indices (cubic on pieces of four or
    more nodes, the interpolating polynomial of lower degree otherwise), so a
    corner is kept instead of rounded off into an overshoot."""

    def __init__(self, s, x, y, cuts):
        from scipy.interpolate import CubicSpline, make_interp_spline
        edges = np.concatenate([[0], np.asarray(cuts, int), [len(s) - 1]])
        self.s_edges = s[edges]
        self.pieces = []
        for a, b in zip(edges[:-1], edges[1:]):
            ss, xx, yy = s[a:b + 1], x[a:b + 1], y[a:b + 1]
            if len(ss) >= 4:
                self.pieces.append((CubicSpline(ss, xx), CubicSpline(ss, yy)))
            else:
                k = len(ss) - 1
                self.pieces.append((make_interp_spline(ss, xx, k=k), make_interp_spline(ss, yy, k=k)))

    def __call__(self, st):
        st = np.asarray(st, float)
        piece = np.clip(np.searchsorted(self.s_edges, st, side="right") - 1, 0, len(self.pieces) - 1)
        x, y = np.empty_like(st), np.empty_like(st)
        for k, (fx, fy) in enumerate(self.pieces):
            sel = piece == k
            if sel.any():
                x[sel], y[sel] = fx(st[sel]), fy(st[sel])
        return x, y


def repanel(af: Airfoil, n_side: int = 100, spacing: str = "cosine", keep_frame: bool = False,
            corner_angle: float = 30.0, te_merge: float = 2e-3) -> Airfoil:
    """Re-distribute a contour to n_side points per side along an arc-length
    spline (aero2d/geometry.py:320-391).

    spacing "cosine" clusters points at both edges; "viscous" strongly at the
    leading edge and moderately at the trailing edge. Corners (direction
    turning by more than corner_angle) cut the spline and each keeps a node,
    its neighbours shifted with it by a linearly decaying amount over six
    nodes; the two corners of a blunt trailing-edge base (within 0.5 % of the
    chord of the contour's ends) are not kept. The leading edge is the spline's
    minimum of x. The result is normalised (or, keep_frame, only made
    counter-clockwise). Ends closer than te_merge of the chord are merged;
    otherwise a closing segment is appended."""
    x, y = np.asarray(af.x, float), np.asarray(af.y, float)
    keep = np.concatenate([[True], np.hypot(np.diff(x), np.diff(y)) > 1e-10])
    x, y = x[keep], y[keep]
    s = np.concatenate([[0.0], np.cumsum(np.hypot(np.diff(x), np.diff(y)))])
    cuts = _turning_points(x, y, corner_angle) if corner_angle else np.zeros(0, int)
    near = 0.005 * max(float(np.ptp(x)), 1e-12)
    at_start = (np.hypot(x[cuts] - x[0], y[cuts] - y[0]) <= near) & (s[cuts] < 0.5 * s[-1])
    at_end = (np.hypot(x[cuts] - x[-1], y[cuts] - y[-1]) <= near) & (s[cuts] > 0.5 * s[-1])
    if at_start.any() and at_end.any():            # a blunt base: spline across it
        cuts = cuts[~(at_start | at_end)]
    spl = _ArcSpline(s, x, y, cuts)
    probe = np.linspace(0.0, s[-1], 20001)
    s_le = probe[np.argmin(spl(probe)[0])]
    if spacing == "viscous":
        u = np.linspace(0.0, 1.0, n_side)          # 0 at the leading edge, 1 at the trailing edge
        g = 0.55 * (1 - np.cos(0.5 * np.pi * u)) + 0.45 * 0.5 * (1 - np.cos(np.pi * u))
        s_up = s_le * (1 - g[::-1])
        s_lo = s_le + (s[-1] - s_le) * g
    else:
        b = np.linspace(0.0, np.pi, n_side)
        s_up = s_le * 0.5 * (1 - np.cos(b))
        s_lo = s_le + (s[-1] - s_le) * 0.5 * (1 - np.cos(b))
    st = np.concatenate([s_up, s_lo[1:]])
    m = 6
    pinned = set()
    for sc in s[cuts]:
        k = int(np.argmin(np.abs(st[1:-1] - sc))) + 1
        shift = sc - st[k]
        for j in range(-m, m + 1):
            i = k + j
            if 0 < i < len(st) - 1 and i not in pinned:
                st[i] += shift * (1.0 - abs(j) / (m + 1))
        st[k] = sc
        pinned.add(k)
    st = np.sort(st)
    out = Airfoil(*spl(st), af.name)
    if keep_frame:
        if _signed_area(out.x, out.y) < 0:
            out = Airfoil(out.x[::-1].copy(), out.y[::-1].copy(), af.name)
    else:
        out = normalize(out)
    gap = np.hypot(out.x[0] - out.x[-1], out.y[0] - out.y[-1])
    if gap <= te_merge * max(float(np.ptp(out.x)), 1e-12):
        out.x[0] = out.x[-1] = 0.5 * (out.x[0] + out.x[-1])
        out.y[0] = out.y[-1] = 0.5 * (out.y[0] + out.y[-1])
    else:
        out = Airfoil(np.r_[out.x, out.x[0]], np.r_[out.y, out.y[0]], out.name)
    return out


def smooth_airfoil(af: Airfoil, strength: float = 1.0) -> Airfoil:
    """Smooth measured coordinates (aero2d/geometry.py:394-421): each surface
    becomes a smoothing spline y(u) with u = sqrt(x - x_le) (smooth through a
    round leading edge), its smoothing parameter chosen by generalised
    cross-validation and multiplied by `strength`. The contour is normalised
    first; the leading-edge point is kept."""
    from scipy.interpolate import make_smoothing_spline
    af = normalize(af)
    x, y = np.asarray(af.x, float), np.asarray(af.y, float)
    ile = int(np.argmin(x))
    xle, yle = x[ile], y[ile]
    sides = []
    for xs, ys in ((x[:ile + 1][::-1], y[:ile + 1][::-1]), (x[ile:], y[ile:])):
        u = np.sqrt(np.maximum(xs - xle, 0.0))
        keep = np.concatenate([[True], np.diff(u) > 1e-9])
        u, ys = u[keep], ys[keep]
        spl = make_smoothing_spline(u, ys)
        if strength != 1.0 and getattr(spl, "lam", None) is not None:
            spl = make_smoothing_spline(u, ys, lam=spl.lam * strength)
        ysm = spl(u)
        ysm[0] = yle
        sides.append((xle + u ** 2, ysm))
    (xu, yu), (xl, yl) = sides
    return Airfoil(np.concatenate([xu[::-1], xl[1:]]), np.concatenate([yu[::-1], yl[1:]]), af.name)


def load_airfoil(path: str, n_side: int = 100, smooth: bool = False, te_merge: float = 2e-3) -> Airfoil:
    """Read a coordinate file, Selig or Lednicer, with or without a header
    (aero2d/geometry.py:424-448): rows of two numbers (comma or blank separated)
    are points, the first other non-blank line is the name, a Lednicer
    point-count row (both values above 1.5) is skipped. Fewer than ten points
    raise ValueError. The contour is normalised (or smoothed) and
    re-distributed to n_side points per side."""
    rows, name = [], None
    with open(path, encoding="utf-8", errors="replace") as f:
        for line in f:
            parts = line.replace(",", " ").split()
            try:
                vals = [float(v) for v in parts[:2]]
            except ValueError:
                vals = []
            if len(vals) == 2:
                rows.append(vals)
            elif name is None and line.strip():
                name = line.strip()
    pts = np.array(rows, float).reshape(-1, 2)
    if len(pts) and pts[0, 0] > 1.5 and pts[0, 1] > 1.5:
        pts = pts[1:]
    if len(pts) < 10:
        raise ValueError(f"{path}: too few contour points ({len(pts)})")
    af = Airfoil(pts[:, 0], pts[:, 1], name or "airfoil")
    af = smooth_airfoil(af) if smooth else normalize(af)
    return repanel(af, n_side, te_merge=te_merge)
