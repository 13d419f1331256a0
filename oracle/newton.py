"""Oracle: pseudo-transient Newton-Krylov solve (PyTorch CPU, FP64).

Test infrastructure only (see oracle/__init__.py). Restates aero2d/implicit.py
and the "batched" GMRES of jax.scipy.sparse.linalg (JAX >= 0.4.30, which the
reference calls at aero2d/implicit.py:184-185 and does not vendor).

Arrays carry the case axis LAST (U is (4, ni, nc, B)), so the reference's
indexing reads unchanged and a batch runs as one set of tensor ops, as
jax.vmap batches the reference. Block arrays are (ni, nc, B, 4, 4). GMRES uses
per-case masks where JAX's batched while_loops stop per case.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch
from torch.func import jvp

from . import flow
from .flow import DT, cons_to_prim, farfield_ghost, hllc_flux, mirror


# ---------------------------------------------------------------------------
# first-order face Jacobians (aero2d/implicit.py:32-59)
# ---------------------------------------------------------------------------
def _columns(fn, x, *others):
    """Forward-mode Jacobian of fn(x, *others) with respect to x, all four unit
    tangents at once along an extra trailing axis (what jax.jacfwd computes).
    x (4, *F) -> (*F, 4, 4) indexed [out, in]; `others` broadcast against x."""
    X = x.unsqueeze(-1).expand(x.shape + (4,)).clone()
    T = torch.zeros_like(X)
    for b in range(4):
        T[b, ..., b] = 1.0
    out = jvp(lambda u: fn(u, *[o.unsqueeze(-1) for o in others]), (X,), (T,))[1]
    return out.movedim(0, -2)


def _flux_c(UL, UR, S):
    return hllc_flux(cons_to_prim(UL), cons_to_prim(UR), S)


def face_jac(UL, UR, S):
    """dF/dUL, dF/dUR for a set of faces; UL, UR (4, *F), S (2, *F) -> 2 x (*F, 4, 4)
    (aero2d/implicit.py:36-41)."""
    AL = _columns(lambda u, ur, s: _flux_c(u, ur, s), UL, UR, S)
    AR = _columns(lambda u, ul, s: _flux_c(ul, u, s), UR, UL, S)
    return AL, AR


def _wall_flux(u, s):
    area = torch.sqrt(s[0] ** 2 + s[1] ** 2)
    w = cons_to_prim(u)
    return hllc_flux(mirror(w, s[0] / area, s[1] / area), w, s)


def wall_jac(U0, S):
    """Wall-face flux Jacobian, mirror ghost as the left state (aero2d/implicit.py:44-51)."""
    return _columns(_wall_flux, U0, S)


def _far_flux(u, s, x, y, winf, circ):
    w = cons_to_prim(u)
    return hllc_flux(w, farfield_ghost(w, s, winf, x, y, circ), s)


def far_jac(Ui, S, winf, xf, yf, circ):
    """Far-field face flux Jacobian with Gamma frozen (aero2d/implicit.py:54-59)."""
    return _columns(_far_flux, Ui, S, xf, yf, winf, circ)


def assemble(U, winf, circ, dtau_inv, g, vortex=True):
    """First-order block Jacobian plus vol/dtau on the diagonal; returns D, Bim,
    Bip, Bjm, Bjp, each (ni, nc, *B, 4, 4) (aero2d/implicit.py:81-110)."""
    AL, AR = face_jac(torch.roll(U, 1, 1), U, g["Si"])
    D = torch.roll(AL, -1, 0) - AR
    Bip = torch.roll(AR, -1, 0)
    Bim = -AL
    AL, AR = face_jac(U[:, :, :-1], U[:, :, 1:], g["Sj"][:, :, 1:-1])
    z = torch.zeros_like(AL[:, :1])
    D = D + torch.cat([AL, z], 1) - torch.cat([z, AR], 1)
    Bjp = torch.cat([AR, z], 1)
    Bjm = torch.cat([z, -AL], 1)
    Aw = wall_jac(U[:, :, 0], g["Sj"][:, :, 0])
    Af = far_jac(U[:, :, -1], g["Sj"][:, :, -1], winf, g["xf"], g["yf"],
                 circ * (1.0 if vortex else 0.0))
    D = torch.cat([D[:, :1] - Aw[:, None], D[:, 1:-1], D[:, -1:] + Af[:, None]], 1)
    eye = torch.eye(4, dtype=U.dtype)
    D = D + (g["vol"] * dtau_inv)[..., None, None] * eye
    return D, Bim, Bip, Bjm, Bjp


# ---------------------------------------------------------------------------
# block line solves along j and the line-Jacobi preconditioner
# ---------------------------------------------------------------------------
def factor_lines(D, Bjm, Bjp):
    """Block Thomas forward elimination along j for every i-line:
    den_j = D_j - A_j c_{j-1}, inv_j = den_j^-1, c_j = inv_j C_j
    (aero2d/implicit.py:112-128). Returns inv, c, A as (nc, ni, *B, 4, 4)."""
    Dt, Am, Cp = D.movedim(1, 0), Bjm.movedim(1, 0), Bjp.movedim(1, 0)
    cprev = torch.zeros_like(Dt[0])
    invs, cs = [], []
    for j in range(Dt.shape[0]):
        den = Dt[j] - Am[j] @ cprev
        inv = torch.linalg.inv(den)
        cprev = inv @ Cp[j]
        invs.append(inv)
        cs.append(cprev)
    return torch.stack(invs), torch.stack(cs), Am


def _bmv(A, x):
    return (A @ x[..., None])[..., 0]


def solve_lines(fac, r):
    """Forward d_j = inv_j (r_j - A_j d_{j-1}); backward x_j = d_j - c_j x_{j+1}
    (aero2d/implicit.py:130-148). r (4, ni, nc, *B) -> x of the same shape."""
    inv, c, Am = fac
    rt = r.movedim(0, -1).movedim(1, 0)          # (nc, ni, *B, 4)
    nc = rt.shape[0]
    d = []
    dprev = torch.zeros_like(rt[0])
    for j in range(nc):
        dprev = _bmv(inv[j], rt[j] - _bmv(Am[j], dprev))
        d.append(dprev)
    xs = [None] * nc
    xnext = torch.zeros_like(rt[0])
    for j in range(nc - 1, -1, -1):
        xnext = d[j] - _bmv(c[j], xnext)
        xs[j] = xnext
    return torch.stack(xs).movedim(0, 1).movedim(-1, 0)


def precond(fac, Bim, Bip, r, sweeps=4):
    """Line Jacobi: one line solve, then (sweeps-1) solves with the i-coupling
    moved to the right-hand side (aero2d/implicit.py:150-157)."""
    x = solve_lines(fac, r)
    for _ in range(sweeps - 1):
        xb = x.movedim(0, -1)                         # (ni, nc, *B, 4)
        off = _bmv(Bim, torch.roll(xb, 1, 0)) + _bmv(Bip, torch.roll(xb, -1, 0))
        x = solve_lines(fac, r - off.movedim(-1, 0))
    return x


def precond_rb(fac, Bim, Bip, r, sweeps=4):
    """Red-black line ordering of the same line solves (no counterpart in the
    reference, which uses line Jacobi): each sweep solves the odd i-lines with
    the current even neighbours, then the even i-lines with the just-updated
    odd ones. Starts from x = 0. Needs an even number of lines."""
    ni = r.shape[1]
    assert ni % 2 == 0
    x = torch.zeros_like(r)
    for _ in range(sweeps):
        for color in (1, 0):
            xb = x.movedim(0, -1)
            off = _bmv(Bim, torch.roll(xb, 1, 0)) + _bmv(Bip, torch.roll(xb, -1, 0))
            xn = solve_lines(fac, r - off.movedim(-1, 0))
            x[:, color::2] = xn[:, color::2]
    return x


# ---------------------------------------------------------------------------
# GMRES, JAX "batched" semantics (jax.scipy.sparse.linalg.gmres)
# ---------------------------------------------------------------------------
def _norm(x):
    """Per-case 2-norm of x (n..., B) over all but the last axis."""
    return torch.sqrt(torch.sum(x.reshape(-1, x.shape[-1]) ** 2, dim=0))


def _safe_normalize(x, thresh=None):
    """x/|x| where |x| > thresh (default: machine eps), else 0 and norm 0."""
    norm = _norm(x)
    if thresh is None:
        thresh = torch.full_like(norm, torch.finfo(x.dtype).eps)
    use = norm > thresh
    xn = torch.where(use, x / torch.where(use, norm, torch.ones_like(norm)), torch.zeros_like(x))
    return xn, torch.where(use, norm, torch.zeros_like(norm))


def _icgs(V, q, max_iterations=2):
    """Iterative classical Gram-Schmidt of q against the basis V (K, n..., B),
    restated from JAX's _iterative_classical_gram_schmidt: after the first
    projection pass a further pass runs only while k < max_iterations - 1,
    which with max_iterations=2 leaves a single pass. Unfilled basis slots are
    zero vectors and give zero overlaps, as in JAX's zero-padded basis."""
    K, B = V.shape[0], V.shape[-1]
    Vf = V.reshape(K, -1, B)
    r = torch.zeros((K, B), dtype=q.dtype)
    k = 0
    while True:
        h = torch.einsum("knb,nb->kb", Vf, q.reshape(-1, B))
        q = q - torch.einsum("kb,knb->nb", h, Vf).reshape(q.shape)
        r = r + h
        k += 1
        if not (k < max_iterations - 1):
            break
    return q, r


def gmres_batched(A, b, x0, M, tol=0.01, atol=0.0, restart=50):
    """One restart cycle (maxiter=1) of left-preconditioned GMRES with JAX's
    "batched" semantics: `restart` Arnoldi steps unless an exact breakdown,
    least squares by normal equations + Cholesky at the end of the cycle; the
    cycle is skipped when |M(b - A x0)| <= ptol = |M b| min(1, atol'/|b|)
    with atol' = max(tol |b|, atol). Vectors are (n..., B), one case per
    trailing index; A and M map such vectors to such vectors."""
    B = b.shape[-1]
    dtype = b.dtype
    eps = torch.finfo(dtype).eps
    b_norm = _norm(b)
    atol_ = torch.clamp_min(tol * b_norm, atol)
    ptol = _norm(M(b)) * torch.clamp_max(atol_ / b_norm, 1.0)
    unit_res, res_norm = _safe_normalize(M(b - A(x0)))
    run = res_norm > ptol
    V = torch.zeros((restart + 1,) + tuple(b.shape), dtype=dtype)
    V[0] = unit_res
    R = torch.eye(restart, restart + 1, dtype=dtype).repeat(B, 1, 1)
    breakdown = torch.zeros(B, dtype=torch.bool)
    for k in range(restart):
        active = ~breakdown
        w = M(A(V[k]))
        _, vn0 = _safe_normalize(w)
        w, h = _icgs(V, w)
        unit_v, vn1 = _safe_normalize(w, thresh=eps * vn0)
        h = h.clone()
        h[k + 1] = vn1
        V[k + 1] = torch.where(active, unit_v, V[k + 1])
        R[:, k] = torch.where(active[:, None], h.T, R[:, k])
        breakdown = breakdown | (vn1 == 0)
    beta = torch.zeros((B, restart + 1), dtype=dtype)
    beta[:, 0] = res_norm
    a = R.transpose(1, 2)                                 # (B, restart+1, restart)
    a2 = a.transpose(1, 2) @ a
    b2 = a.transpose(1, 2) @ beta[..., None]
    L, info = torch.linalg.cholesky_ex(a2)
    y = torch.cholesky_solve(b2, L)[..., 0]              # (B, restart)
    y = torch.where((info == 0)[:, None], y, torch.full_like(y, float("nan")))
    dx = torch.einsum("bk,knb->nb", y, V[:-1].reshape(restart, -1, B)).reshape(b.shape)
    x = x0 + dx
    return torch.where(run, x, x0)


# ---------------------------------------------------------------------------
# Newton step and host loop
# ---------------------------------------------------------------------------
@dataclass
class Params:
    """ImplicitEuler / EulerSolver defaults (aero2d/implicit.py:68-78,
    aero2d/euler.py:219-221)."""
    cfl0: float = 10.0
    cfl_max: float = 300.0
    restart: int = 50
    gtol: float = 0.01
    sweeps: int = 4
    max_update: float = 0.2
    cfl_explicit: float = 2.5
    vortex: bool = True


def step(U, winf, g, cfl, p: Params = Params()):
    """One pseudo-transient Newton step for a batch (aero2d/implicit.py:169-192).
    U (4, ni, nc, B), winf (4, B), g: geometry with trailing case axis, cfl (B,).
    Returns (U_new, rn, omega), the last two of shape (B,)."""
    vortex = p.vortex
    circ = flow.circulation(U, winf, g)
    R, W = flow.residual(U, winf, circ, g, vortex=vortex)
    dtau_inv = 1.0 / (flow.local_dt(W, g) / p.cfl_explicit * cfl)
    D, Bim, Bip, Bjm, Bjp = assemble(U, winf, circ, dtau_inv, g, vortex)
    fac = factor_lines(D, Bjm, Bjp)
    diag = (g["vol"] * dtau_inv)[None]
    res = lambda u: flow.residual_with_circ(u, winf, g, vortex)[0]
    matvec = lambda v: diag * v + jvp(res, (U,), (v,))[1]
    Mop = lambda r: precond(fac, Bim, Bip, r, p.sweeps)
    dU = gmres_batched(matvec, -R, Mop(-R), Mop, tol=p.gtol, atol=0.0, restart=p.restart)
    Wn = cons_to_prim(U + dU)
    rel = torch.maximum((torch.abs(Wn[0] - W[0]) / W[0]).amax(dim=(0, 1)),
                        (torch.abs(Wn[3] - W[3]) / W[3]).amax(dim=(0, 1)))
    omega = torch.clamp_max(p.max_update / (rel + 1e-30), 1.0)
    omega = torch.where(torch.isnan(rel), rel, omega)
    Unew = U + omega * dU
    rn = torch.sqrt(torch.mean((R[0] / g["vol"]) ** 2, dim=(0, 1)))
    return Unew, rn, omega


@dataclass
class SolveResult:
    """EulerResult fields (aero2d/euler.py:438-448) plus the loop's bookkeeping
    (aero2d/implicit.py:234-252)."""
    mach: np.ndarray
    alpha: np.ndarray
    cl: np.ndarray
    cd: np.ndarray
    cm: np.ndarray
    cp: np.ndarray
    res_hist: np.ndarray
    U: torch.Tensor
    converged: np.ndarray
    iterations: int
    cl_history: np.ndarray
    frozen: np.ndarray


def batched_forces(U, winf, g):
    return flow.forces(U, winf, g)


def solve(g, machs, alphas, n_iter=200, tol=1e-8, p: Params = Params(), U0=None,
          callback=None, per_case_stop=False):
    """Batched solve with the reference host loop (aero2d/implicit.py:194-253):
    SER CFL, rejection of non-finite steps with CFL 1, freezing after 5 failures,
    stop when every case has rn/r0 < tol or is frozen. g is a batched geometry
    dict (one entry per case). With per_case_stop, a case stops updating once it
    meets tol (the HIP library's per-case semantics); otherwise the whole batch
    iterates until all cases are done, as the reference does."""
    machs = np.asarray(machs, np.float64)
    alphas = np.asarray(alphas, np.float64)
    nb = len(machs)
    winf = flow.freestream(torch.as_tensor(machs), torch.as_tensor(alphas))
    ni, nc = g["vol"].shape[0], g["vol"].shape[1]
    U = flow.init_state(winf, ni, nc) if U0 is None else U0.clone()
    cfl = np.full(nb, p.cfl0)
    hist, clh = [], []
    r0 = None
    fails = np.zeros(nb, int)
    frozen = np.zeros(nb, bool)
    stopped = np.zeros(nb, bool)
    iterations = 0
    for it in range(n_iter):
        with torch.no_grad():
            Unew, rn, om = step(U, winf, g, torch.as_tensor(cfl, dtype=DT), p)
        rn, om = rn.numpy(), om.numpy()
        finite = np.isfinite(rn) & torch.isfinite(Unew).reshape(-1, nb).all(0).numpy()
        bad = ~finite | frozen | stopped
        fails += ~finite & ~stopped
        frozen |= fails >= 5
        U = torch.where(torch.as_tensor(bad), U, Unew)
        rn = np.where(finite, rn, np.nan)
        if r0 is None:
            r0 = np.where(np.isfinite(rn), np.maximum(rn, 1e-30), 1.0)
        cl = batched_forces(U, winf, g)[0].numpy()
        ser = p.cfl0 * (r0 / np.where(np.isfinite(rn), np.maximum(rn, 1e-30), r0))
        cfl = np.clip(ser, p.cfl0, p.cfl_max)
        cfl = np.where(om < 1.0, np.maximum(cfl * 0.5, 1.0), cfl)
        cfl = np.where(~finite, 1.0, cfl)
        hist.append(rn)
        clh.append(cl)
        iterations = it + 1
        done = (rn / r0 < tol) | frozen
        if per_case_stop:
            stopped |= done
        if callback is not None and callback(it + 1, n_iter, np.log10(rn / r0), cl):
            break
        if np.all(done | stopped):
            break
    hist = np.array(hist)
    cl, cd, cm, cp = (t.numpy() for t in batched_forces(U, winf, g))
    last = np.array([h[np.isfinite(h)][-1] if np.any(np.isfinite(h)) else np.nan for h in hist.T])
    converged = (~frozen) & (np.log10(last / r0) < -3)
    return SolveResult(machs, alphas, cl, cd, cm, cp, hist, U, converged, iterations,
                       np.array(clh), frozen)
