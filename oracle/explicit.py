"""Oracle: the reference's explicit solver (PyTorch CPU, FP64) — test
infrastructure only (see oracle/__init__.py).

Four-stage Runge-Kutta pseudo-time iteration with local time steps, implicit
residual smoothing (periodic constant-coefficient tridiagonal along i solved
by FFT, tridiagonal along j), an optional FAS V-cycle over coarser grids made
of every other node (first order there), and the chunked host loop with the
circulation updated after every cycle (aero2d/euler.py:219-250, 298-312,
334-436). One case at a time (no batch axes): the explicit form is checked
case by case.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import flow

DT = torch.float64
RK_STAGES = (0.25, 1.0 / 3.0, 0.5, 1.0)        # aero2d/euler.py:340


@dataclass
class Level:
    """One grid level of the explicit solver (aero2d/euler.py:219-250)."""
    g: dict
    ni: int
    nc: int
    order: int
    irs_den_i: torch.Tensor
    coarse: "Level | None"


def make_levels(x, y, irs_eps=0.6, order=2, mg_levels=1) -> Level:
    """Level chain: each coarser level takes every other node in both
    directions and runs first order (aero2d/euler.py:238-250)."""
    g = flow.case_geometry(x, y)
    ni, nc = g["vol"].shape
    k = np.arange(ni)
    den = torch.tensor(1 + 2 * irs_eps * (1 - np.cos(2 * np.pi * k / ni)), dtype=DT)
    coarse = None
    if mg_levels > 1:
        if ni % 2 or nc % 2:
            raise ValueError("multigrid needs even cell counts in i and j")
        coarse = make_levels(x[::2, ::2], y[::2, ::2], irs_eps, order=1, mg_levels=mg_levels - 1)
    return Level(g, ni, nc, order, den, coarse)


def smooth(R, lev: Level, eps):
    """Implicit residual smoothing (aero2d/euler.py:298-312): along i the
    periodic system (-eps, 1 + 2 eps, -eps) by FFT, along j the tridiagonal
    (-eps, 1 + 2 eps, -eps) with 1 + eps on both end rows, by a Thomas solve."""
    if eps <= 0:
        return R
    Rh = torch.fft.fft(R, dim=1)
    R = torch.real(torch.fft.ifft(Rh / lev.irs_den_i[None, :, None], dim=1)).to(DT)
    n = lev.nc
    d = torch.full((n,), 1 + 2 * eps, dtype=DT)
    d[0] = d[-1] = 1 + eps
    lo = torch.full((n,), -eps, dtype=DT)        # sub-diagonal (row j couples j - 1)
    lo[0] = 0.0
    up = torch.full((n,), -eps, dtype=DT)        # super-diagonal
    up[-1] = 0.0
    # Thomas along j for every (component, i)
    cp = torch.empty(n, dtype=DT)
    x = torch.empty_like(R)
    cp[0] = up[0] / d[0]
    y = torch.empty_like(R)
    y[:, :, 0] = R[:, :, 0] / d[0]
    for j in range(1, n):
        den = d[j] - lo[j] * cp[j - 1]
        cp[j] = up[j] / den
        y[:, :, j] = (R[:, :, j] - lo[j] * y[:, :, j - 1]) / den
    x[:, :, n - 1] = y[:, :, n - 1]
    for j in range(n - 2, -1, -1):
        x[:, :, j] = y[:, :, j] - cp[j] * x[:, :, j + 1]
    return x


def iterate(U, winf, circ, lev: Level, cfl, eps, lim_k, vortex, P=None):
    """One four-stage RK iteration (aero2d/euler.py:334-345). Returns U and the
    density change rate sqrt(mean(((U - U0)_rho / dt)^2))."""
    U0 = U
    dt = flow.local_dt(flow.cons_to_prim(U), lev.g, cfl)
    for a in RK_STAGES:
        R, _ = flow.residual(U, winf, circ, lev.g, lim_k=lim_k, vortex=vortex, order=lev.order)
        if P is not None:
            R = R + P
        R = smooth(R, lev, eps)
        U = U0 - a * (dt / lev.g["vol"])[None] * R
    res = torch.sqrt(torch.mean(((U[0] - U0[0]) / dt) ** 2))
    return U, res


def sum4(a):
    """Sum over the 2 x 2 fine cells of each coarse cell (aero2d/euler.py:348-350)."""
    return a[..., 0::2, 0::2] + a[..., 1::2, 0::2] + a[..., 0::2, 1::2] + a[..., 1::2, 1::2]


def prolong(d):
    """Bilinear interpolation of a coarse correction (weights 3/4, 1/4): periodic
    along i, the end values held along j (aero2d/euler.py:352-364)."""
    e = 0.75 * d + 0.25 * torch.roll(d, 1, dims=1)
    o = 0.75 * d + 0.25 * torch.roll(d, -1, dims=1)
    d = torch.stack([e, o], dim=2).reshape(d.shape[0], -1, d.shape[2])
    dm = torch.cat([d[:, :, :1], d[:, :, :-1]], dim=2)
    dp = torch.cat([d[:, :, 1:], d[:, :, -1:]], dim=2)
    e = 0.75 * d + 0.25 * dm
    o = 0.75 * d + 0.25 * dp
    return torch.stack([e, o], dim=3).reshape(d.shape[0], d.shape[1], -1)


def cycle(U, winf, circ, lev: Level, cfl, eps, lim_k, vortex, P=None):
    """FAS V-cycle (aero2d/euler.py:367-381)."""
    U, res = iterate(U, winf, circ, lev, cfl, eps, lim_k, vortex, P)
    if lev.coarse is not None:
        cs = lev.coarse
        R, _ = flow.residual(U, winf, circ, lev.g, lim_k=lim_k, vortex=vortex, order=lev.order)
        if P is not None:
            R = R + P
        vol = lev.g["vol"]
        Uc0 = sum4(U * vol[None]) / sum4(vol)[None]
        Rc, _ = flow.residual(Uc0, winf, circ, cs.g, lim_k=lim_k, vortex=vortex, order=cs.order)
        Pc = sum4(R) - Rc
        Uc, _ = cycle(Uc0, winf, circ, cs, cfl, eps, lim_k, vortex, Pc)
        U = U + prolong(Uc - Uc0)
        U, _ = iterate(U, winf, circ, lev, cfl, eps, lim_k, vortex, P)
    return U, res


@dataclass
class ExplicitResult:
    cl: float
    cd: float
    cm: float
    cp: np.ndarray
    res_hist: np.ndarray
    cl_hist: np.ndarray
    iterations: int
    U: torch.Tensor


def solve(x, y, mach, alpha, n_iter=4000, chunk=200, tol=1e-7, U0=None, tol_cl=2e-5, tol_cd=2e-6,
          cfl=2.5, irs_eps=0.6, order=2, vortex=True, mg_levels=1, lim_k=0.02) -> ExplicitResult:
    """The reference's batched explicit solve for one case
    (aero2d/euler.py:386-435): chunks of `chunk` cycles, Gamma = Cl |V|/2
    after every cycle (0 at the start), stop when the chunk's residual fell
    below tol of the first chunk's or Cl, Cd moved less than tol_cl, tol_cd
    over the last chunk; FloatingPointError when the state is not finite."""
    lev = make_levels(x, y, irs_eps, order, mg_levels)
    winf = flow.freestream(torch.tensor(mach, dtype=DT), torch.tensor(alpha, dtype=DT))
    U = flow.init_state(winf, lev.ni, lev.nc) if U0 is None else torch.as_tensor(U0, dtype=DT)
    circ = torch.zeros((), dtype=DT)
    hist, clh, cdh = [], [], []
    res0 = None
    it = 0
    for it in range(0, n_iter, chunk):
        res = None
        for _ in range(chunk):
            U, res = cycle(U, winf, circ, lev, cfl, irs_eps, lim_k, vortex)
            cl = flow.forces(U, winf, lev.g)[0]
            circ = 0.5 * cl * torch.sqrt(winf[1] ** 2 + winf[2] ** 2)
        res = float(res)
        if res0 is None:
            res0 = res
        cl, cd, _, _ = flow.forces(U, winf, lev.g)
        hist.append(res)
        clh.append(float(cl))
        cdh.append(float(cd))
        if not np.isfinite(res):
            break
        if res / res0 < tol:
            break
        if len(clh) > 2 and abs(clh[-1] - clh[-2]) < tol_cl and abs(cdh[-1] - cdh[-2]) < tol_cd:
            break
    if not bool(torch.all(torch.isfinite(U))):
        raise FloatingPointError("the solution diverged (NaN)")
    cl, cd, cm, cp = flow.forces(U, winf, lev.g)
    return ExplicitResult(float(cl), float(cd), float(cm), cp.numpy(), np.array(hist), np.array(clh),
                          it + chunk, U)
