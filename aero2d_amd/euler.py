"""Finite-volume Euler discretisation on an O-grid — the API of aero2d/euler.py
with its device work done by the HIP library.

Cell-centred finite volumes, HLLC flux, MUSCL on primitives with the smooth
van Albada limiter, mirror-ghost wall, characteristic far field with the
compressible point-vortex correction. Non-dimensionalisation:
rho_inf = 1, a_inf = 1, p_inf = 1/gamma, |V_inf| = M (aero2d/euler.py:1-15).

Metrics are computed in FP64 NumPy on the host, as the reference does
(aero2d/euler.py:44-60), and again by the library from the same nodes. Every
field operation (residual, forces, local time step, HLLC) runs on the GPU;
there is no CPU fallback.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import _abi
from .grid import OGrid

GAMMA = 1.4

__all__ = ["GAMMA", "Metrics", "compute_metrics", "cons_to_prim", "prim_to_cons", "hllc_flux",
           "EulerSolver", "EulerResult"]


@dataclass
class Metrics:
    """aero2d/euler.py:33-41."""
    vol: np.ndarray      # (ni, nc)
    Si: np.ndarray       # (2, ni, nc)
    Sj: np.ndarray       # (2, ni, nc+1)
    xc: np.ndarray
    yc: np.ndarray
    xw: np.ndarray
    yw: np.ndarray


def compute_metrics(g: OGrid) -> Metrics:
    """Face vectors, areas, centres and wall midpoints (aero2d/euler.py:44-60)."""
    x, y = g.x.astype(np.float64), g.y.astype(np.float64)
    vol = g.cell_areas()
    dx = x[:, 1:] - x[:, :-1]
    dy = y[:, 1:] - y[:, :-1]
    Si = np.stack([-dy, dx])
    x1, y1 = np.roll(x, -1, 0), np.roll(y, -1, 0)
    Sj = np.stack([y1 - y, -(x1 - x)])
    xc = 0.25 * (x[:, :-1] + x[:, 1:] + x1[:, :-1] + x1[:, 1:])
    yc = 0.25 * (y[:, :-1] + y[:, 1:] + y1[:, :-1] + y1[:, 1:])
    xw = 0.5 * (x[:, 0] + x1[:, 0])
    yw = 0.5 * (y[:, 0] + y1[:, 0])
    return Metrics(vol, Si, Sj, xc, yc, xw, yw)


def cons_to_prim(U):
    """(rho, rho u, rho v, E) -> (rho, u, v, p) (aero2d/euler.py:66-71)."""
    U = np.asarray(U, np.float64)
    rho = U[0]
    u = U[1] / rho
    v = U[2] / rho
    p = (GAMMA - 1) * (U[3] - 0.5 * rho * (u * u + v * v))
    return np.stack([rho, u, v, p])


def prim_to_cons(W):
    """aero2d/euler.py:74-76."""
    rho, u, v, p = np.asarray(W, np.float64)
    return np.stack([rho, rho * u, rho * v, p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v)])


def hllc_flux(WL, WR, S):
    """HLLC flux through faces with area vectors S (aero2d/euler.py:79-121),
    evaluated by the library's device flux. WL, WR (4, ...), S (2, ...)."""
    WL = np.asarray(WL, np.float64)
    shape = WL.shape[1:]
    n = int(np.prod(shape)) if shape else 1
    a = np.ascontiguousarray(WL.reshape(4, n))
    b = np.ascontiguousarray(np.asarray(WR, np.float64).reshape(4, n))
    s = np.ascontiguousarray(np.asarray(S, np.float64).reshape(2, n))
    F = np.empty((4, n))
    _abi.lib()
    _abi.call("a2d_hllc_flux", n, _abi.dptr(a), _abi.dptr(b), _abi.dptr(s), _abi.dptr(F))
    return F.reshape((4,) + shape)


def _precision_of(dtype, precision):
    if precision is not None:
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        return precision
    return "fp64" if np.dtype(dtype) == np.float64 else "mixed"


class EulerSolver:
    """Discretisation of one O-grid (aero2d/euler.py:218-250).

    dtype float64 selects the FP64 form; float32 (the reference default)
    selects the mixed form, whose residual, update and forces stay FP64 while
    the Jacobian-vector product, preconditioner and Krylov vectors are FP32.
    `precision` ("fp64" / "mixed") overrides dtype. irs_eps and mg_levels
    belong to the reference's explicit scheme and are stored only.
    stencil_march: the residual/tangent kernel (a2d_knobs.stencil_march: 0 LDS
    tiles, 2 column march, 1 the march for batches that fill the device).
    """

    def __init__(self, grid: OGrid, dtype=np.float32, cfl: float = 2.5, irs_eps: float = 0.6,
                 order: int = 2, vortex_correction: bool = True, mg_levels: int = 1,
                 lim_k: float = 0.02, precision: str | None = None, stencil_march: int = 1):
        if order not in (1, 2):
            raise ValueError("order must be 1 or 2")
        self.dtype = np.dtype(dtype)
        self.precision = _precision_of(dtype, precision)
        self.grid = grid
        m = compute_metrics(grid)
        self.m_np = m
        self.vol, self.Si, self.Sj = m.vol, m.Si, m.Sj
        self.xf = 0.5 * (grid.x[:, -1] + np.roll(grid.x[:, -1], -1))
        self.yf = 0.5 * (grid.y[:, -1] + np.roll(grid.y[:, -1], -1))
        self.xw, self.yw = m.xw, m.yw
        self.cfl, self.irs_eps, self.order, self.lim_k = cfl, irs_eps, order, lim_k
        self.vortex = vortex_correction
        self.stencil_march = stencil_march
        self.mg_levels = mg_levels
        self.ni, self.nc = m.vol.shape
        self._pool = None
        self._case = None

    # ---- knobs shared with ImplicitEuler ---------------------------------------
    def knobs(self, **over) -> _abi.Knobs:
        k = _abi.default_knobs()
        k.cfl_explicit = self.cfl
        k.lim_k = self.lim_k
        k.vortex = int(bool(self.vortex))
        k.order = int(self.order)
        k.stencil_march = int(self.stencil_march)
        for name, v in over.items():
            setattr(k, name, v)
        return k

    @property
    def prec_code(self) -> int:
        return _abi.PREC_FP64 if self.precision == "fp64" else _abi.PREC_MIXED

    # ---- single-case device scratch ----------------------------------------------
    def _scratch(self, winf, U=None):
        if self._pool is None:
            self._pool = _abi.Pool(0, self.knobs())
            self._case = self._pool.add_case(self.grid.x, self.grid.y, 0.5, 0.0, self.prec_code)
        winf = np.asarray(winf, np.float64)
        mach = math.hypot(winf[1], winf[2])
        alpha = math.degrees(math.atan2(winf[2], winf[1]))
        _abi.call("a2d_case_set_freestream", self._pool.handle, self._case, mach, alpha)
        if U is not None:
            Uc = np.ascontiguousarray(U, np.float64)
            _abi.call("a2d_case_set_state", self._pool.handle, self._case, _abi.dptr(Uc))
        return self._pool.handle, self._case

    # ---- reference API -------------------------------------------------------------
    def freestream(self, mach, alpha_deg):
        """winf = (1, M cos a, M sin a, 1/gamma) (aero2d/euler.py:253-256)."""
        a = np.deg2rad(alpha_deg)
        mach = np.asarray(mach, np.float64)
        return np.stack([np.ones_like(mach), mach * np.cos(a), mach * np.sin(a),
                         np.ones_like(mach) / GAMMA])

    def init_state(self, winf):
        """Uniform freestream state (4, ni, nc) (aero2d/euler.py:383-384)."""
        winf = np.asarray(winf, np.float64)
        return np.broadcast_to(prim_to_cons(winf[:, None, None]), (4, self.ni, self.nc)).copy()

    def residual(self, U, winf, circ):
        """R(U) with circulation circ, on the GPU (aero2d/euler.py:259-287).
        Returns (R, W) with W = cons_to_prim(U)."""
        h, c = self._scratch(winf, U)
        R = np.empty((4, self.ni, self.nc))
        rn = np.zeros(1)
        _abi.call("a2d_residual", h, c, float(circ), _abi.dptr(R), _abi.dptr(rn))
        return R, cons_to_prim(U)

    def local_dt(self, W):
        """Local pseudo-time step (aero2d/euler.py:289-296), from the GPU."""
        U = prim_to_cons(W)
        h, c = self._scratch(self._winf_guess(), U)
        dt = np.empty((self.ni, self.nc))
        _abi.call("a2d_local_dt", h, c, _abi.dptr(dt))
        return dt

    def _winf_guess(self):
        return self.freestream(0.5, 0.0)

    def wall_pressure(self, W):
        """aero2d/euler.py:315-316."""
        return 1.5 * W[3, :, 0] - 0.5 * W[3, :, 1]

    def forces(self, U, winf):
        """Cl, Cd, Cm and wall Cp, integrated in FP64 on the GPU (aero2d/euler.py:318-331)."""
        h, c = self._scratch(winf, U)
        out = np.empty(3)
        cp = np.empty(self.ni)
        circ = np.empty(1)
        _abi.call("a2d_forces", h, c, _abi.dptr(out), _abi.dptr(cp), _abi.dptr(circ))
        return out[0], out[1], out[2], cp

    def close(self):
        if self._pool is not None:
            self._pool.close()
            self._pool = None


@dataclass
class EulerResult:
    """aero2d/euler.py:438-448, plus converged_tol: rn/r0 met the solve's tol."""
    mach: np.ndarray
    alpha: np.ndarray
    cl: np.ndarray
    cd: np.ndarray
    cm: np.ndarray
    cp: np.ndarray
    res_hist: np.ndarray
    U: np.ndarray
    converged: np.ndarray | None = None
    converged_tol: np.ndarray | None = None
    iterations_per_case: np.ndarray | None = None
