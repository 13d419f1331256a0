"""Application entry of the Euler path: a polar over angles of attack on one
airfoil, the reference's `core.run_euler` (aero2d/core.py:220-260) on the HIP
solver.

    from aero2d_amd.core import run_euler
    from aero2d_amd import naca4
    p = run_euler(naca4("2412", 100), [0, 2, 4], 0.6)      # p.cl, p.cd, p.cm, p.ok, p.cp

All angles run as one batch of the implicit solver in the mixed form (FP32
Jacobian-vector products and preconditioner, FP64 residual, update and
forces; the reference runs this path in FP32).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .euler import EulerSolver
from .geometry import Airfoil, repanel
from .grid import ogrid
from .implicit import ImplicitEuler

__all__ = ["GRIDS", "Polar", "run_euler", "grid_key"]

# Euler grids: (points per airfoil side, wall-normal node layers, first
# wall-normal spacing), keyed as in the reference's interface
# (aero2d/core.py:14-18); English aliases map to the same presets.
GRIDS = {
    "грубая (быстро)": (65, 49, 2e-3),
    "средняя": (97, 65, 1.5e-3),
    "мелкая (медленно)": (129, 97, 1e-3),
}
GRID_ALIASES = {"coarse": "грубая (быстро)", "medium": "средняя", "fine": "мелкая (медленно)"}


def grid_key(grid: str) -> str:
    """The GRIDS key for a preset name or its English alias."""
    key = GRID_ALIASES.get(grid, grid)
    if key not in GRIDS:
        raise KeyError(f"unknown grid preset {grid!r}; one of {list(GRIDS) + list(GRID_ALIASES)}")
    return key


@dataclass
class Polar:
    """A polar over angles of attack (aero2d/core.py:21-44): coefficients per
    angle, `ok` where the solve converged, and the wall Cp per angle."""
    label: str
    method: str
    mach: float
    alpha: np.ndarray
    cl: np.ndarray
    cd: np.ndarray
    cm: np.ndarray
    ok: np.ndarray
    cp_x: list = field(default_factory=list)
    cp: list = field(default_factory=list)
    notes: list = field(default_factory=list)
    re: float = float("nan")
    cdf: np.ndarray | None = None
    xtr_u: np.ndarray | None = None
    xtr_l: np.ndarray | None = None
    bl: list = field(default_factory=list)
    cp_inv: list = field(default_factory=list)
    field_ctx: dict | None = None
    ch: np.ndarray | None = None
    chord_ref: float = 1.0
    xref: tuple = (0.25, 0.0)
    elem: list | None = None


def run_euler(af: Airfoil, alphas, mach: float, grid: str = "грубая (быстро)", n_iter: int = 100,
              callback=None) -> Polar:
    """Euler polar of `af` at Mach `mach` (aero2d/core.py:220-260).

    The contour is re-distributed to the preset's points per side, an O-grid
    marched to 40 chords, and every angle solved in one batch to a residual
    drop of 1e-6 (at most n_iter Newton steps). callback(text, fraction) is
    called with progress; returning True stops the solve and returns the
    current state. Raises ValueError below M 0.2, where the solver has no
    low-Mach preconditioning."""
    if mach < 0.2:
        raise ValueError("the Euler solver does not run below M = 0.2 (no low-Mach preconditioning); "
                         "use a panel method there")
    nside, nj, ds0 = GRIDS[grid_key(grid)]
    if callback:
        callback("Building the grid...", 0.0)
    g = ogrid(repanel(af, nside), nj=nj, ds0=ds0, radius=40.0)
    solver = ImplicitEuler(EulerSolver(g, dtype=np.float32))
    alphas = np.atleast_1d(np.asarray(alphas, float))

    def progress(it, n, logres, cl):
        if callback is None:
            return False
        worst = np.nanmax(logres) if np.any(np.isfinite(logres)) else 0.0
        return callback(f"Iteration {it}/{n}: residual down {-worst:.1f} orders (worst case)", it / n)

    if callback:
        callback("Solving (first iteration includes set-up)...", 0.0)
    try:
        r = solver.solve(np.full_like(alphas, mach), alphas, n_iter=n_iter, tol=1e-6, verbose=False,
                         callback=progress)
    finally:
        solver.b.close()
    xw = np.asarray(solver.b.m_np.xw)
    ok = np.asarray(r.converged, bool)
    notes = []
    if not ok.all():
        notes.append("Not converged at alpha = " + ", ".join(f"{a:g}" for a in alphas[~ok]) +
                     " deg; their values are unreliable (usually transonic flow with strong shocks or "
                     "large angles). Try a finer grid or more iterations.")
    notes.append("Euler is inviscid: Cd is wave drag only (shocks), with no friction or separation.")
    name = grid_key(grid).split(" ")[0]
    return Polar(f"{af.name}, Euler ({name}), M={mach:g}", "euler", mach, alphas, r.cl, r.cd, r.cm, ok,
                 [xw] * len(alphas), [c for c in r.cp], notes)
