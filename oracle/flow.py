"""Oracle: finite-volume Euler residual on an O-grid (PyTorch CPU, FP64).

Test infrastructure only (see oracle/__init__.py). Every function keeps the
reference's variable-first indexing of aero2d/euler.py (W[0] is density,
W[:, :, 0] the first wall-normal layer) and accepts optional TRAILING case axes:
U (4, ni, nc, *B), winf (4, *B), circ (*B), geometry arrays (..., *B). A batch
therefore runs as one set of tensor ops, the way jax.vmap batches the reference.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

GAMMA = 1.4           # aero2d/euler.py:27
DT = torch.float64


@dataclass
class Metrics:
    """aero2d/euler.py:33-41 (FP64 NumPy)."""
    vol: np.ndarray      # (ni, nc)
    Si: np.ndarray       # (2, ni, nc)    normal of i-face between cells i-1 and i
    Sj: np.ndarray       # (2, ni, nc+1)  normal of j-face j, j = 0 is the wall
    xc: np.ndarray
    yc: np.ndarray
    xw: np.ndarray       # (ni,) wall-face midpoints
    yw: np.ndarray


def cell_areas(x, y):
    """aero2d/grid.py:24-31."""
    x1, y1 = np.roll(x, -1, 0), np.roll(y, -1, 0)
    d1x, d1y = x1[:, 1:] - x[:, :-1], y1[:, 1:] - y[:, :-1]
    d2x, d2y = x[:, 1:] - x1[:, :-1], y[:, 1:] - y1[:, :-1]
    return -0.5 * (d1x * d2y - d1y * d2x)


def compute_metrics(x: np.ndarray, y: np.ndarray) -> Metrics:
    """aero2d/euler.py:44-60."""
    x, y = x.astype(np.float64), y.astype(np.float64)
    vol = cell_areas(x, y)
    dx = x[:, 1:] - x[:, :-1]
    dy = y[:, 1:] - y[:, :-1]
    Si = np.stack([-dy, dx])
    x1, y1 = np.roll(x, -1, 0), np.roll(y, -1, 0)
    dx = x1 - x
    dy = y1 - y
    Sj = np.stack([dy, -dx])
    xc = 0.25 * (x[:, :-1] + x[:, 1:] + x1[:, :-1] + x1[:, 1:])
    yc = 0.25 * (y[:, :-1] + y[:, 1:] + y1[:, :-1] + y1[:, 1:])
    xw = 0.5 * (x[:, 0] + x1[:, 0])
    yw = 0.5 * (y[:, 0] + y1[:, 0])
    return Metrics(vol, Si, Sj, xc, yc, xw, yw)


def case_geometry(x: np.ndarray, y: np.ndarray) -> dict:
    """Per-case tensors the residual needs (aero2d/euler.py:226-237): metrics,
    outer-face midpoints xf, yf and wall midpoints xw, yw, all FP64."""
    m = compute_metrics(x, y)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=DT)
    xf = 0.5 * (x[:, -1] + np.roll(x[:, -1], -1))
    yf = 0.5 * (y[:, -1] + np.roll(y[:, -1], -1))
    return dict(vol=t(m.vol), Si=t(m.Si), Sj=t(m.Sj), xf=t(xf), yf=t(yf), xw=t(m.xw), yw=t(m.yw))


def stack_geometry(geoms: list[dict]) -> dict:
    """Stack same-size per-case geometry dicts along a new trailing case axis."""
    return {k: torch.stack([g[k] for g in geoms], dim=-1) for k in geoms[0]}


# ---------------------------------------------------------------------------
# thermodynamics and HLLC
# ---------------------------------------------------------------------------
def cons_to_prim(U):
    """(rho, rho u, rho v, E) -> (rho, u, v, p) (aero2d/euler.py:66-71)."""
    rho = U[0]
    u = U[1] / rho
    v = U[2] / rho
    p = (GAMMA - 1) * (U[3] - 0.5 * rho * (u * u + v * v))
    return torch.stack([rho, u, v, p])


def prim_to_cons(W):
    """aero2d/euler.py:74-76."""
    rho, u, v, p = W
    return torch.stack([rho, rho * u, rho * v, p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v)])


def hllc_flux(WL, WR, S):
    """HLLC flux through a face with area vector S (already times length), Einfeldt
    wave speeds from Roe averages (aero2d/euler.py:79-121)."""
    area = torch.sqrt(S[0] ** 2 + S[1] ** 2)
    nx, ny = S[0] / area, S[1] / area
    rL, uL, vL, pL = WL
    rR, uR, vR, pR = WR
    qL = uL * nx + vL * ny
    qR = uR * nx + vR * ny
    cL = torch.sqrt(GAMMA * pL / rL)
    cR = torch.sqrt(GAMMA * pR / rR)
    EL = pL / (GAMMA - 1) + 0.5 * rL * (uL**2 + vL**2)
    ER = pR / (GAMMA - 1) + 0.5 * rR * (uR**2 + vR**2)
    sl, sr = torch.sqrt(rL), torch.sqrt(rR)
    qt = (sl * qL + sr * qR) / (sl + sr)
    HL = (EL + pL) / rL
    HR = (ER + pR) / rR
    Ht = (sl * HL + sr * HR) / (sl + sr)
    ut = (sl * uL + sr * uR) / (sl + sr)
    vt = (sl * vL + sr * vR) / (sl + sr)
    ct = torch.sqrt(torch.clamp_min((GAMMA - 1) * (Ht - 0.5 * (ut**2 + vt**2)), 1e-12))
    SL = torch.minimum(qL - cL, qt - ct)
    SR = torch.maximum(qR + cR, qt + ct)
    SM = (pR - pL + rL * qL * (SL - qL) - rR * qR * (SR - qR)) / (rL * (SL - qL) - rR * (SR - qR))

    def F(r, u, v, p, E, q):
        return torch.stack([r * q, r * u * q + p * nx, r * v * q + p * ny, (E + p) * q])

    def Ustar(r, u, v, p, E, q, Sk):
        f = r * (Sk - q) / (Sk - SM)
        return torch.stack([f,
                            f * (u + (SM - q) * nx),
                            f * (v + (SM - q) * ny),
                            f * (E / r + (SM - q) * (SM + p / (r * (Sk - q))))])

    FL = F(rL, uL, vL, pL, EL, qL)
    FR = F(rR, uR, vR, pR, ER, qR)
    UL = torch.stack([rL, rL * uL, rL * vL, EL])
    UR = torch.stack([rR, rR * uR, rR * vR, ER])
    FsL = FL + SL * (Ustar(rL, uL, vL, pL, EL, qL, SL) - UL)
    FsR = FR + SR * (Ustar(rR, uR, vR, pR, ER, qR, SR) - UR)
    flux = torch.where(SL >= 0, FL, torch.where(SM >= 0, FsL, torch.where(SR > 0, FsR, FR)))
    return flux * area


def van_albada(a, b, eps):
    """Smooth van Albada slope limiter, the reference default LIMITER="smooth"
    (aero2d/euler.py:124-137)."""
    eps = eps + 1e-20
    return ((b * b + eps) * a + (a * a + eps) * b) / (a * a + b * b + 2 * eps)


# ---------------------------------------------------------------------------
# boundary conditions
# ---------------------------------------------------------------------------
def mirror(Wc, nx, ny):
    r, u, v, p = Wc
    vn = u * nx + v * ny
    return torch.stack([r, u - 2 * vn * nx, v - 2 * vn * ny, p])


def wall_ghost(W, Sj0):
    """Mirror ghosts of cells j=1, j=0 -> j=-2, j=-1 (aero2d/euler.py:143-153)."""
    area = torch.sqrt(Sj0[0] ** 2 + Sj0[1] ** 2)
    nx, ny = Sj0[0] / area, Sj0[1] / area
    return mirror(W[:, :, 1], nx, ny), mirror(W[:, :, 0], nx, ny)


def farfield_ghost(Wi, Sjf, winf, xf, yf, circ):
    """Characteristic far field (Riemann invariants) with the compressible
    point-vortex correction about (0.25, 0) (aero2d/euler.py:156-205)."""
    area = torch.sqrt(Sjf[0] ** 2 + Sjf[1] ** 2)
    nx, ny = Sjf[0] / area, Sjf[1] / area
    rinf, uinf, vinf, pinf = winf
    minf2 = (uinf**2 + vinf**2) / (GAMMA * pinf / rinf)
    dxv, dyv = xf - 0.25, yf
    alpha = torch.atan2(vinf, uinf)
    phi = torch.atan2(dyv, dxv)
    rr = torch.sqrt(dxv**2 + dyv**2)
    beta = torch.sqrt(torch.clamp_min(1 - minf2, 1e-6))
    fac = torch.where(minf2 < 1,
                      circ * beta / (2 * math.pi * rr * (1 - minf2 * torch.sin(phi - alpha) ** 2)),
                      torch.zeros_like(rr))
    ub = uinf + fac * torch.sin(phi)
    vb = vinf - fac * torch.cos(phi)
    q2inf = uinf**2 + vinf**2
    q2b = ub**2 + vb**2
    cinf2 = GAMMA * pinf / rinf
    cb2 = torch.clamp_min(cinf2 + 0.5 * (GAMMA - 1) * (q2inf - q2b), 1e-6)
    pb = pinf * (cb2 / cinf2) ** (GAMMA / (GAMMA - 1))
    rb = rinf * (cb2 / cinf2) ** (1 / (GAMMA - 1))
    r, u, v, p = Wi
    ci = torch.sqrt(GAMMA * p / r)
    cb = torch.sqrt(GAMMA * pb / rb)
    qni = u * nx + v * ny
    qnb = ub * nx + vb * ny
    Rp = qni + 2 * ci / (GAMMA - 1)
    Rm = qnb - 2 * cb / (GAMMA - 1)
    qn = 0.5 * (Rp + Rm)
    c = 0.25 * (GAMMA - 1) * (Rp - Rm)
    outflow = qn > 0
    ut = torch.where(outflow, u - qni * nx, ub - qnb * nx)
    vt = torch.where(outflow, v - qni * ny, vb - qnb * ny)
    s = torch.where(outflow, p / r**GAMMA, pb / rb**GAMMA)
    rho = (c * c / (GAMMA * s)) ** (1 / (GAMMA - 1))
    pres = rho * c * c / GAMMA
    Wb = torch.stack([rho, ut + qn * nx, vt + qn * ny, pres])
    qnorm_inf = uinf * nx + vinf * ny
    sup_in = (qnorm_inf < 0) & (torch.abs(qnorm_inf) > torch.sqrt(cinf2))
    sup_out = (qni > 0) & (qni > ci)
    Wb = torch.where(sup_in, torch.stack([rb, ub, vb, pb]), Wb)
    Wb = torch.where(sup_out, Wi, Wb)
    return Wb


# ---------------------------------------------------------------------------
# residual, local time step, forces
# ---------------------------------------------------------------------------
def _recon(Wm, W0, Wp, Wpp, eps):
    """MUSCL left/right states on the face between W0 and Wp (aero2d/euler.py:211-215)."""
    sL = van_albada(W0 - Wm, Wp - W0, eps)
    sR = van_albada(Wp - W0, Wpp - Wp, eps)
    return W0 + 0.5 * sL, Wp - 0.5 * sR


def _clamp_rho_p(W):
    return torch.stack([torch.clamp_min(W[0], 1e-6), W[1], W[2], torch.clamp_min(W[3], 1e-6)])


def freestream(mach, alpha_deg):
    """winf = (1, M cos a, M sin a, 1/gamma), shape (4, *B) (aero2d/euler.py:253-256)."""
    a = torch.deg2rad(torch.as_tensor(alpha_deg, dtype=DT))
    mach = torch.as_tensor(mach, dtype=DT)
    return torch.stack([torch.ones_like(mach), mach * torch.cos(a), mach * torch.sin(a),
                        torch.ones_like(mach) / GAMMA])


def init_state(winf, ni, nc):
    """Uniform conserved freestream (4, ni, nc, *B) (aero2d/euler.py:383-384)."""
    Uc = prim_to_cons(winf[:, None, None])
    return Uc.expand((4, ni, nc) + tuple(winf.shape[1:])).clone()


def residual(U, winf, circ, g, lim_k=0.02, vortex=True, order=2):
    """Second-order FV residual R = d_i F^i + d_j F^j of one case or a batch
    (aero2d/euler.py:259-287). Returns (R, W).

    The i-faces and j-faces go through ONE reconstruction and ONE HLLC call on
    their concatenation along axis 2: the same elementwise arithmetic as the
    reference's two calls, with half the tensor dispatches."""
    W = cons_to_prim(U)
    nc = W.shape[2]
    mach = torch.sqrt(winf[1] ** 2 + winf[2] ** 2)
    scale = torch.stack([winf[0], mach, mach, winf[3]])
    eps = ((lim_k * scale) ** 2)[:, None, None]
    Sj = g["Sj"]
    g2, g1 = wall_ghost(W, Sj[:, :, 0])
    Wf = farfield_ghost(W[:, :, -1], Sj[:, :, -1], winf, g["xf"], g["yf"],
                        circ * (1.0 if vortex else 0.0))
    # j-direction extended array [g2, g1, W..., Wf, Wf]
    We = torch.cat([g2[:, :, None], g1[:, :, None], W, Wf[:, :, None], Wf[:, :, None]], dim=2)
    # i-direction (periodic; face i between cells i-1 and i) stencil
    Wm2, Wm1, Wp1 = torch.roll(W, 2, 1), torch.roll(W, 1, 1), torch.roll(W, -1, 1)
    cat = lambda a, b: torch.cat([a, b], dim=2)
    if order == 2:
        WL, WR = _recon(cat(Wm2, We[:, :, 0:-3]), cat(Wm1, We[:, :, 1:-2]),
                        cat(W, We[:, :, 2:-1]), cat(Wp1, We[:, :, 3:]), eps)
    else:
        WL, WR = cat(Wm1, We[:, :, 1:-2]), cat(W, We[:, :, 2:-1])
    F = hllc_flux(_clamp_rho_p(WL), _clamp_rho_p(WR), cat(g["Si"], Sj))
    Fi, Fj = F[:, :, :nc], F[:, :, nc:]
    R = (torch.roll(Fi, -1, 1) - Fi) + (Fj[:, :, 1:] - Fj[:, :, :-1])
    return R, W


def local_dt(W, g, cfl=2.5):
    """Local pseudo-time step (aero2d/euler.py:289-296)."""
    r, u, v, p = W
    c = torch.sqrt(GAMMA * p / r)
    Si, Sj = g["Si"], g["Sj"]
    Si_c = 0.5 * (Si + torch.roll(Si, -1, 1))
    Sj_c = 0.5 * (Sj[:, :, 1:] + Sj[:, :, :-1])
    li = torch.abs(u * Si_c[0] + v * Si_c[1]) + c * torch.sqrt(Si_c[0] ** 2 + Si_c[1] ** 2)
    lj = torch.abs(u * Sj_c[0] + v * Sj_c[1]) + c * torch.sqrt(Sj_c[0] ** 2 + Sj_c[1] ** 2)
    return cfl * g["vol"] / (li + lj)


def wall_pressure(W):
    """Linear extrapolation to the wall (aero2d/euler.py:315-316)."""
    return 1.5 * W[3, :, 0] - 0.5 * W[3, :, 1]


def forces(U, winf, g):
    """Pressure-integrated Cl, Cd, Cm (about x = 0.25) and wall Cp
    (aero2d/euler.py:318-331)."""
    W = cons_to_prim(U)
    pw = wall_pressure(W)
    S = g["Sj"][:, :, 0]
    q = 0.5 * (winf[1] ** 2 + winf[2] ** 2)
    cp = (pw - winf[3]) / q
    fx = -torch.sum(cp * S[0], dim=0)
    fy = -torch.sum(cp * S[1], dim=0)
    mach = torch.sqrt(2 * q)
    ca, sa = winf[1] / mach, winf[2] / mach
    cl = fy * ca - fx * sa
    cd = fx * ca + fy * sa
    cm = torch.sum(cp * ((g["xw"] - 0.25) * S[1] - g["yw"] * S[0]), dim=0)
    return cl, cd, cm, cp


def circulation(U, winf, g):
    """Gamma = Cl |V_inf| / 2 from the current state (aero2d/implicit.py:160-164)."""
    cl = forces(U, winf, g)[0]
    return 0.5 * cl * torch.sqrt(winf[1] ** 2 + winf[2] ** 2)


def residual_with_circ(U, winf, g, vortex=True):
    """Residual with Gamma recomputed from U, the function the implicit solver
    linearises (aero2d/implicit.py:166-167)."""
    return residual(U, winf, circulation(U, winf, g), g, vortex=vortex)
