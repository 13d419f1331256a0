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
    `precision` ("fp64" / "mixed") overrides dtype. cfl, irs_eps and
    mg_levels are the explicit scheme's (`solve`), which runs in FP64 whatever
    the dtype; with mg_levels > 1 every level must have even cell counts
    along i and j (ValueError otherwise, as the reference raises).
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
        self.iterations = 0
        self.cl_history = np.zeros((0,))
        self._levels = [(np.asarray(grid.x, np.float64), np.asarray(grid.y, np.float64))]
        for _ in range(mg_levels - 1):
            x, y = self._levels[-1]
            if (x.shape[0] % 2) or ((x.shape[1] - 1) % 2):
                raise ValueError("multigrid needs an even number of cells along i and j on every level")
            self._levels.append((x[::2, ::2], y[::2, ::2]))
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

    # ---- the explicit scheme ------------------------------------------------------
    def explicit_batch(self, machs, alphas, U0=None):
        """Device cases of the explicit scheme for each (M, alpha): a new pool
        holding, per case, one FP64 case per grid level, each linked to the
        next coarser one. U0 (B, 4, ni, nc) or (4, ni, nc): starting state
        (freestream otherwise). Returns (pool, ids of the finest level); the
        caller closes the pool."""
        pool = _abi.Pool(0, self.knobs(irs_eps=float(self.irs_eps)))
        try:
            ids = []
            for b in range(len(machs)):
                prev = None
                for x, y in self._levels:
                    cid = pool.add_case(x, y, machs[b], alphas[b], _abi.PREC_FP64)
                    if prev is None:
                        ids.append(cid)
                    else:
                        pool.link_coarse(prev, cid)
                    prev = cid
            if U0 is not None:
                U0 = np.asarray(U0, np.float64)
                for b, cid in enumerate(ids):
                    Ub = np.ascontiguousarray(U0[b] if U0.ndim == 4 else U0)
                    _abi.call("a2d_case_set_state", pool.handle, cid, _abi.dptr(Ub))
        except Exception:
            pool.close()
            raise
        return pool, ids

    def solve(self, machs, alphas, n_iter=4000, chunk=200, tol=1e-7, verbose=True, U0=None,
              tol_cl=2e-5, tol_cd=2e-6):
        """Batched explicit solve (aero2d/euler.py:386-435): chunks of `chunk`
        cycles of four-stage Runge-Kutta with local time steps, residual
        smoothing and (mg_levels > 1) a FAS V-cycle, Gamma = Cl |V_inf| / 2
        after every cycle (0 at the start); stop when every case's res fell
        below tol of its first chunk's, or Cl and Cd of every case moved less
        than tol_cl, tol_cd over the last chunk. Raises FloatingPointError when
        the state is not finite at the end. All cases run as one batch on the
        GPU in FP64. Returns EulerResult with res_hist (n_chunks, B); sets
        self.iterations and self.cl_history."""
        machs = np.atleast_1d(np.asarray(machs, np.float64))
        alphas = np.atleast_1d(np.asarray(alphas, np.float64))
        if machs.shape != alphas.shape or machs.ndim != 1:
            raise ValueError("machs and alphas must be 1-D arrays of one length")
        nb = len(machs)
        pool, ids = self.explicit_batch(machs, alphas, U0)
        try:
            hist, clh, cdh = [], [], []
            res0 = None
            self.iterations = 0
            for it in range(0, n_iter, chunk):
                out = pool.explicit_cycles(ids, chunk, fresh=(it == 0))
                self.iterations = it + chunk
                res, cl, cd = out[:, 0].copy(), out[:, 1].copy(), out[:, 2].copy()
                if res0 is None:
                    res0 = res
                hist.append(res)
                clh.append(cl)
                cdh.append(cd)
                if verbose:
                    with np.errstate(divide="ignore", invalid="ignore"):
                        worst = np.max(np.log10(res / res0))
                    print(f"  iter {it + chunk:6d}  log10(res/res0) max {worst:+.2f}  Cl {np.array2string(cl, precision=4)}")
                if not np.all(np.isfinite(res)):
                    break
                with np.errstate(divide="ignore", invalid="ignore"):
                    if np.all(res / res0 < tol):
                        break
                if len(clh) > 2 and np.all(np.abs(clh[-1] - clh[-2]) < tol_cl) \
                        and np.all(np.abs(cdh[-1] - cdh[-2]) < tol_cd):
                    break
            self.cl_history = np.array(clh)
            U = np.empty((nb, 4, self.ni, self.nc))
            for b, cid in enumerate(ids):
                _abi.call("a2d_case_get_state", pool.handle, cid, _abi.dptr(U[b]))
            if not np.all(np.isfinite(U)):
                raise FloatingPointError("the solution diverged (NaN)")
            cl, cd, cm = np.empty(nb), np.empty(nb), np.empty(nb)
            cp = np.empty((nb, self.ni))
            f3, circ = np.empty(3), np.empty(1)
            for b, cid in enumerate(ids):
                _abi.call("a2d_forces", pool.handle, cid, _abi.dptr(f3), _abi.dptr(cp[b]), _abi.dptr(circ))
                cl[b], cd[b], cm[b] = f3
        finally:
            pool.close()
        return EulerResult(machs, alphas, cl, cd, cm, cp, np.array(hist), U)

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
