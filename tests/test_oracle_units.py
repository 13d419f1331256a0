"""Closed-form, invariant and brute-force checks of the oracle's functions.

Each check is chosen so that a plausible restatement mistake (a dropped term, a
wrong sign or index, a transposed block) fails it; none retypes the formula
under test."""
import math

import numpy as np
import pytest
import torch
from torch.func import jvp

from aero2d_amd import naca4, ogrid
from oracle import flow, newton
from tests.helpers import oracle_geo, perturbed_state, small_grid, winf_t

T = lambda a: torch.tensor(a, dtype=torch.float64)


def phys_flux(W, nx, ny):
    r, u, v, p = W
    q = u * nx + v * ny
    E = p / 0.4 + 0.5 * r * (u * u + v * v)
    return np.array([r * q, r * u * q + p * nx, r * v * q + p * ny, (E + p) * q])


def test_hllc_consistency():
    """F(W, W) = physical flux (reference tests/test_euler.py:15-24), FP64 tolerance."""
    W = np.array([1.2, 0.3, -0.1, 0.9])
    F = flow.hllc_flux(T(W[:, None]), T(W[:, None]), T([[0.6], [0.8]]))[:, 0].numpy()
    assert np.allclose(F, phys_flux(W, 0.6, 0.8), rtol=1e-13, atol=1e-15)


def test_hllc_antisymmetry_and_upwinding():
    rng = np.random.default_rng(0)
    n = 500
    WL = np.stack([rng.uniform(0.5, 2, n), rng.normal(0, .5, n), rng.normal(0, .5, n), rng.uniform(.3, 1.2, n)])
    WR = np.stack([rng.uniform(0.5, 2, n), rng.normal(0, .5, n), rng.normal(0, .5, n), rng.uniform(.3, 1.2, n)])
    S = rng.normal(size=(2, n))
    F = flow.hllc_flux(T(WL), T(WR), T(S)).numpy()
    Fr = flow.hllc_flux(T(WR), T(WL), T(-S)).numpy()
    assert np.allclose(F, -Fr, rtol=1e-12, atol=1e-13)
    # supersonic to the right: the left physical flux
    WLs = np.array([1.0, 3.0, 0.2, 0.7])
    WRs = np.array([0.8, 2.5, -0.1, 0.6])
    Fs = flow.hllc_flux(T(WLs[:, None]), T(WRs[:, None]), T([[2.0], [0.0]]))[:, 0].numpy()
    assert np.allclose(Fs, 2.0 * phys_flux(WLs, 1.0, 0.0), rtol=1e-13)


def test_hllc_stationary_contact():
    """q = 0 and equal pressure on both sides: only the pressure flux remains."""
    WL = np.array([1.0, 0.0, 0.3, 0.7])
    WR = np.array([0.4, 0.0, -0.2, 0.7])
    S = np.array([[1.5], [0.0]])
    F = flow.hllc_flux(T(WL[:, None]), T(WR[:, None]), T(S))[:, 0].numpy()
    assert np.allclose(F, [0.0, 0.7 * 1.5, 0.0, 0.0], atol=1e-14)


def test_prim_cons_roundtrip():
    W = T([1.1, 0.4, -0.2, 0.7])
    assert torch.allclose(flow.cons_to_prim(flow.prim_to_cons(W)), W, rtol=1e-14)


def test_van_albada_properties():
    a = T(np.linspace(-2, 2, 41))
    b = T(np.linspace(1.5, -1.0, 41))
    eps = 1e-4
    assert torch.allclose(flow.van_albada(a, b, eps), flow.van_albada(b, a, eps), rtol=1e-15)
    assert torch.allclose(flow.van_albada(a, a, eps), a, rtol=1e-14)
    # opposite slopes much larger than sqrt(eps) are limited to nearly zero
    assert abs(flow.van_albada(T(1.0), T(-1.0), 1e-8).item()) < 1e-7


def test_wall_ghost_reverses_normal_velocity():
    W = T(np.random.default_rng(1).normal(size=(4, 6, 3)) + np.array([2, 0, 0, 2])[:, None, None])
    Sj0 = T(np.random.default_rng(2).normal(size=(2, 6)))
    g2, g1 = flow.wall_ghost(W, Sj0)
    n = Sj0 / torch.sqrt(Sj0[0] ** 2 + Sj0[1] ** 2)
    qn = lambda w: w[1] * n[0] + w[2] * n[1]
    qt = lambda w: -w[1] * n[1] + w[2] * n[0]
    assert torch.allclose(qn(g1), -qn(W[:, :, 0]), atol=1e-14)
    assert torch.allclose(qt(g1), qt(W[:, :, 0]), atol=1e-14)
    assert torch.allclose(g2[0], W[0, :, 1]) and torch.allclose(g2[3], W[3, :, 1])


def test_farfield_freestream_and_vortex():
    w = winf_t(0.5, 3.0)
    ang = torch.linspace(0, 2 * math.pi, 17, dtype=torch.float64)[:-1]
    xf, yf = 0.5 + 20 * torch.cos(ang), 20 * torch.sin(ang)
    S = torch.stack([torch.cos(ang), torch.sin(ang)]) * 0.3
    Wi = w[:, None].expand(4, 16)
    Wb = flow.farfield_ghost(Wi, S, w, xf, yf, T(0.0))
    assert torch.allclose(Wb, Wi, atol=1e-13)
    # with circulation the boundary velocity is the freestream plus the
    # compressible point vortex  G beta / (2 pi r (1 - M^2 sin^2(phi - a)))
    G = 0.1
    Wb = flow.farfield_ghost(Wi, S, w, xf, yf, T(G))
    ub = flow.farfield_ghost(Wi, S, w, xf, yf, T(G))  # same call, deterministic
    assert torch.equal(Wb, ub)
    # along the outer normal the interior Riemann invariant is kept on outflow faces
    phi = torch.atan2(yf, xf - 0.25)
    r = torch.sqrt((xf - 0.25) ** 2 + yf ** 2)
    beta = math.sqrt(1 - 0.25)
    fac = G * beta / (2 * math.pi * r * (1 - 0.25 * torch.sin(phi - math.radians(3.0)) ** 2))
    # tangential velocity on inflow faces equals that of the vortex-corrected state
    nx, ny = torch.cos(ang), torch.sin(ang)
    inflow = (w[1] * nx + w[2] * ny) < 0
    ut_exp = -(w[1] + fac * torch.sin(phi)) * ny + (w[2] - fac * torch.cos(phi)) * nx
    ut_got = -Wb[1] * ny + Wb[2] * nx
    assert torch.allclose(ut_got[inflow], ut_exp[inflow], atol=1e-12)


def test_metrics_closed_cells_and_total_area():
    g = small_grid(33, 13)
    m = flow.compute_metrics(g.x, g.y)
    # outward face vectors of every cell sum to zero
    Si, Sj = m.Si, m.Sj
    sx = np.roll(Si[0], -1, 0) - Si[0] + Sj[0][:, 1:] - Sj[0][:, :-1]
    sy = np.roll(Si[1], -1, 0) - Si[1] + Sj[1][:, 1:] - Sj[1][:, :-1]
    assert np.max(np.abs(sx)) < 1e-13 and np.max(np.abs(sy)) < 1e-13
    # total area = area inside the outer layer - area of the airfoil polygon (shoelace)
    shoe = lambda x, y: 0.5 * abs(np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y))
    area = shoe(g.x[:, -1], g.y[:, -1]) - shoe(g.x[:, 0], g.y[:, 0])
    assert abs(m.vol.sum() - area) < 1e-10 * area


def test_freestream_preserved():
    """Reference tests/test_euler.py:32-40, at FP64 tolerance."""
    g = ogrid(naca4("0012", 33), nj=25)
    geo = oracle_geo(g)
    w = winf_t(0.5, 3.0)
    U = flow.init_state(w, *geo["vol"].shape)
    R, _ = flow.residual(U, w, T(0.0), geo, vortex=False)
    assert float(torch.max(torch.abs(R[:, :, 3:]) / geo["vol"][None, :, 3:])) < 1e-10


def test_residual_mirror_symmetry():
    """NACA 0012 at alpha = 0: the residual of a mirrored state is the mirrored residual."""
    g = small_grid(25, 11)
    geo = oracle_geo(g)
    ni, nc = geo["vol"].shape
    w = winf_t(0.6, 0.0)
    U = torch.tensor(perturbed_state(ni, nc, 0.6, 0.0, 4))
    mirror_i = lambda A: torch.flip(A, dims=[1])      # cell i <-> cell ni-1-i
    Um = mirror_i(U) * T([1, 1, -1, 1])[:, None, None]
    R, _ = flow.residual(U, w, T(0.0), geo)
    Rm, _ = flow.residual(Um, w, T(0.0), geo)
    assert torch.allclose(Rm, mirror_i(R) * T([1, 1, -1, 1])[:, None, None], atol=1e-10 * R.abs().max())


def test_residual_telescopes_to_boundary_fluxes():
    """Sum of R over all cells = far-field flux - wall flux (first order, where
    both boundary fluxes are a single HLLC of explicit states)."""
    g = small_grid(17, 9)
    geo = oracle_geo(g)
    ni, nc = geo["vol"].shape
    w = winf_t(0.5, 2.0)
    U = torch.tensor(perturbed_state(ni, nc, 0.5, 2.0, 8))
    circ = T(0.07)
    R, W = flow.residual(U, w, circ, geo, order=1)
    Sj = geo["Sj"]
    a0 = torch.sqrt(Sj[0, :, 0] ** 2 + Sj[1, :, 0] ** 2)
    ghost = flow.mirror(W[:, :, 0], Sj[0, :, 0] / a0, Sj[1, :, 0] / a0)
    Fw = flow.hllc_flux(ghost, W[:, :, 0], Sj[:, :, 0])
    Wb = flow.farfield_ghost(W[:, :, -1], Sj[:, :, -1], w, geo["xf"], geo["yf"], circ)
    Ff = flow.hllc_flux(W[:, :, -1], Wb, Sj[:, :, -1])
    assert torch.allclose(R.sum(dim=(1, 2)), Ff.sum(1) - Fw.sum(1), atol=1e-12)


def test_forces_uniform_pressure_and_rotation():
    g = small_grid(33, 9)
    geo = oracle_geo(g)
    ni, nc = geo["vol"].shape
    w = winf_t(0.5, 4.0)
    W = torch.stack([torch.ones(ni, nc), torch.zeros(ni, nc), torch.zeros(ni, nc),
                     torch.full((ni, nc), 0.9)]).double()
    cl, cd, cm, cp = flow.forces(flow.prim_to_cons(W), w, geo)
    assert abs(cl) < 1e-13 and abs(cd) < 1e-13 and abs(cm) < 1e-13
    U = torch.tensor(perturbed_state(ni, nc, 0.5, 4.0, 3))
    cl, cd, _, cp = flow.forces(U, w, geo)
    S = geo["Sj"][:, :, 0]
    fx, fy = -(cp * S[0]).sum(), -(cp * S[1]).sum()
    assert abs(cl ** 2 + cd ** 2 - (fx ** 2 + fy ** 2)) < 1e-13
    cl0, cd0, _, _ = flow.forces(U, winf_t(0.5, 0.0), geo)
    # at alpha = 0 lift is the y force and drag the x force
    assert abs(cl0 - (-(flow.forces(U, winf_t(0.5, 0.0), geo)[3] * S[1]).sum())) < 1e-14


def _fd_jvp(f, U, v, h=1e-6):
    return (f(U + h * v) - f(U - h * v)) / (2 * h)


def test_jvp_matches_central_differences_and_circulation_coupling():
    g = small_grid(17, 9)
    geo = oracle_geo(g)
    ni, nc = geo["vol"].shape
    w = winf_t(0.5, 2.0)
    U = torch.tensor(perturbed_state(ni, nc, 0.5, 2.0, 11))
    v = torch.tensor(np.random.default_rng(12).normal(size=U.shape))
    res = lambda u: flow.residual_with_circ(u, w, geo)[0]
    j = jvp(res, (U,), (v,))[1]
    fd = _fd_jvp(res, U, v)
    assert torch.max(torch.abs(j - fd)) < 1e-6 * torch.max(torch.abs(j))
    # a perturbation of the two wall layers only reaches the far-field rows
    # through the circulation Gamma(U)
    vw = torch.zeros_like(U)
    vw[:, :, :2] = v[:, :, :2]
    jw = jvp(res, (U,), (vw,))[1]
    assert torch.max(torch.abs(jw[:, :, -1])) > 1e-6
    jw_frozen = jvp(lambda u: flow.residual(u, w, flow.circulation(U, w, geo), geo)[0], (U,), (vw,))[1]
    assert torch.max(torch.abs(jw_frozen[:, :, -1])) == 0.0


def test_face_jacobians_match_differences():
    g = small_grid(17, 9)
    geo = oracle_geo(g)
    ni, nc = geo["vol"].shape
    U = torch.tensor(perturbed_state(ni, nc, 0.5, 2.0, 13))
    UL, UR, S = torch.roll(U, 1, 1), U, geo["Si"]
    AL, AR = newton.face_jac(UL, UR, S)
    h = 1e-6
    for b in range(4):
        e = torch.zeros_like(U)
        e[b] = h
        fdL = (newton._flux_c(UL + e, UR, S) - newton._flux_c(UL - e, UR, S)) / (2 * h)
        fdR = (newton._flux_c(UL, UR + e, S) - newton._flux_c(UL, UR - e, S)) / (2 * h)
        assert torch.allclose(AL[..., :, b], fdL.movedim(0, -1), atol=1e-6)
        assert torch.allclose(AR[..., :, b], fdR.movedim(0, -1), atol=1e-6)


def _dense_line_operator(D, Bjm, Bjp, i):
    nc = D.shape[1]
    A = np.zeros((4 * nc, 4 * nc))
    for j in range(nc):
        A[4 * j:4 * j + 4, 4 * j:4 * j + 4] = D[i, j]
        if j > 0:
            A[4 * j:4 * j + 4, 4 * (j - 1):4 * j] = Bjm[i, j]
        if j < nc - 1:
            A[4 * j:4 * j + 4, 4 * (j + 1):4 * j + 8] = Bjp[i, j]
    return A


def test_line_solve_and_preconditioner_match_dense_algebra():
    g = small_grid(9, 6, ds0=0.1, radius=3.0)          # 16 x 5 cells
    geo = oracle_geo(g)
    ni, nc = geo["vol"].shape
    w = winf_t(0.5, 2.0)
    U = torch.tensor(perturbed_state(ni, nc, 0.5, 2.0, 21))
    circ = flow.circulation(U, w, geo)
    dti = 1.0 / (flow.local_dt(flow.cons_to_prim(U), geo) / 2.5 * 20.0)
    D, Bim, Bip, Bjm, Bjp = (a.numpy() for a in newton.assemble(U, w, circ, dti, geo))
    fac = newton.factor_lines(*(torch.tensor(a) for a in (D, Bjm, Bjp)))
    r = np.random.default_rng(3).normal(size=(4, ni, nc))
    x = newton.solve_lines(fac, torch.tensor(r)).numpy()
    for i in range(ni):
        A = _dense_line_operator(D, Bjm, Bjp, i)
        xi = np.linalg.solve(A, r[:, i, :].T.reshape(-1)).reshape(nc, 4).T
        assert np.allclose(x[:, i, :], xi, rtol=1e-10, atol=1e-12)
    # 4 line-Jacobi sweeps == the dense iteration x <- L^-1 (r - B x), with
    # L the block-tridiagonal j-lines and B the periodic i-coupling
    N = 4 * ni * nc
    idx = lambda i, j: 4 * (i * nc + j)
    L = np.zeros((N, N))
    Bm = np.zeros((N, N))
    for i in range(ni):
        L[idx(i, 0):idx(i, 0) + 4 * nc, idx(i, 0):idx(i, 0) + 4 * nc] = _dense_line_operator(D, Bjm, Bjp, i)
        for j in range(nc):
            Bm[idx(i, j):idx(i, j) + 4, idx((i - 1) % ni, j):idx((i - 1) % ni, j) + 4] = Bim[i, j]
            Bm[idx(i, j):idx(i, j) + 4, idx((i + 1) % ni, j):idx((i + 1) % ni, j) + 4] = Bip[i, j]
    rv = r.transpose(1, 2, 0).reshape(-1)
    xv = np.linalg.solve(L, rv)
    for _ in range(3):
        xv = np.linalg.solve(L, rv - Bm @ xv)
    xp = newton.precond(fac, torch.tensor(Bim), torch.tensor(Bip), torch.tensor(r), 4).numpy()
    assert np.allclose(xp.transpose(1, 2, 0).reshape(-1), xv, rtol=1e-9, atol=1e-11)
    # red-black: block Gauss-Seidel over lines, odd lines first, from x = 0
    xv = np.zeros(N)
    for _ in range(3):
        for color in (1, 0):
            rows = np.concatenate([np.arange(idx(i, 0), idx(i, 0) + 4 * nc) for i in range(color, ni, 2)])
            rhs = rv - Bm @ xv
            xv[rows] = np.linalg.solve(L[np.ix_(rows, rows)], rhs[rows])
    xr = newton.precond_rb(fac, torch.tensor(Bim), torch.tensor(Bip), torch.tensor(r), 3).numpy()
    assert np.allclose(xr.transpose(1, 2, 0).reshape(-1), xv, rtol=1e-9, atol=1e-11)
    # and both orderings converge to the solution of the full system (L + B) x = r
    xf = np.linalg.solve(L + Bm, rv)
    xr = newton.precond_rb(fac, torch.tensor(Bim), torch.tensor(Bip), torch.tensor(r), 60).numpy()
    assert np.allclose(xr.transpose(1, 2, 0).reshape(-1), xf, rtol=1e-8, atol=1e-10)


def test_gmres_batched_is_the_krylov_least_squares_minimiser():
    """One cycle of m steps minimises |M(b - A x)| over x0 + K_m(MA, M r0):
    compare with a dense least-squares solve on an explicit Krylov basis; with
    m = n the cycle solves the system."""
    rng = np.random.default_rng(5)
    n, B = 24, 3
    As = [np.eye(n) * 3 + rng.normal(size=(n, n)) * 0.4 for _ in range(B)]
    Ms = [np.diag(1.0 / np.diag(A)) for A in As]
    b = rng.normal(size=(n, B))
    x0 = rng.normal(size=(n, B)) * 0.1
    At = torch.tensor(np.stack(As)); Mt = torch.tensor(np.stack(Ms))
    Aop = lambda v: torch.einsum("bij,jb->ib", At, v)
    Mop = lambda v: torch.einsum("bij,jb->ib", Mt, v)
    for m in (6, n):
        x = newton.gmres_batched(Aop, torch.tensor(b), torch.tensor(x0), Mop, tol=1e-12, restart=m).numpy()
        for k in range(B):
            MA = Ms[k] @ As[k]
            r0 = Ms[k] @ (b[:, k] - As[k] @ x0[:, k])
            K = np.stack([np.linalg.matrix_power(MA, p) @ r0 for p in range(m)], axis=1)
            Q, _ = np.linalg.qr(K)
            c, *_ = np.linalg.lstsq(MA @ Q, r0, rcond=None)
            xe = x0[:, k] + Q @ c
            assert np.allclose(x[:, k], xe, rtol=1e-8, atol=1e-10)
            if m == n:
                assert np.allclose(x[:, k], np.linalg.solve(As[k], b[:, k]), rtol=1e-8)
