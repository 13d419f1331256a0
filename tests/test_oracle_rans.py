"""Pins of the RANS oracle (oracle/rans.py) by closed forms, invariants and
dense algebra (CPU only). Each check is chosen so that a plausible slip (a
dropped term, a sign, a transposed operand) fails it."""
import math

import numpy as np
import pytest
import torch

from aero2d_amd import naca4, ogrid
from aero2d_amd.rans import ilu_structure
from aero2d_amd.rans_mesh import connectivity, ogrid_mesh
from oracle.rans import CV1, CW3, GAM, BlockILU, RansOracle, gmres

DT = torch.float64


@pytest.fixture(scope="module")
def mesh():
    g = ogrid(naca4("0012", 17), nj=13, ds0=1e-2, radius=6.0)
    return ogrid_mesh(g, quad_layers=6)


def _state(o, alpha, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    winf = o.freestream(alpha)
    xc = o.m.centroid
    W = np.tile(winf, (o.nc, 1))
    for k in range(5):
        c = rng.normal(size=3)
        f = c[0] * np.sin(2.1 * xc[:, 0] + c[1]) * np.cos(1.7 * xc[:, 1] + c[2])
        W[:, k] += amp * (abs(winf[k]) if k != 2 else o.mach) * f + (0.3 * f if k == 4 else 0.0)
    return o.cons(torch.tensor(W))


def _phys_flux(w, n):
    r, u, v, p, c = w
    un = u * n[0] + v * n[1]
    H = GAM / (GAM - 1) * p / r + 0.5 * (u * u + v * v)
    return np.array([r * un, r * u * un + p * n[0], r * v * un + p * n[1], r * H * un, r * un * c])


def test_mesh_cells_closed_and_oriented(mesh):
    """Every cell's outward face normals sum to zero (aero2d/rans_mesh.py tests:
    tests/test_rans.py:35-46); areas positive; wall and far faces labelled."""
    acc = np.zeros((mesh.n_cells, 2))
    L, R = mesh.face_cells[:, 0], mesh.face_cells[:, 1]
    s = mesh.face_normal * mesh.face_area[:, None]
    np.add.at(acc, L, s)
    np.add.at(acc, R[R >= 0], -s[R >= 0])
    assert np.abs(acc).max() < 1e-12
    assert mesh.volume.min() > 0 and (mesh.face_tag == 1).sum() == 32 and (mesh.face_tag == 2).sum() == 32
    assert abs(mesh.volume.sum() - (math.pi * np.max(np.hypot(mesh.nodes[:, 0] - 0.5, mesh.nodes[:, 1])) ** 2
                                     - 0.0817)) < 0.05 * mesh.volume.sum()   # annulus minus a NACA 0012 section
    with pytest.raises(RuntimeError):
        connectivity(np.array([[0, 0], [1, 0], [0, 1.0]]), [[0, 1, 2]], {}, set())


def test_roe_consistency_antisymmetry_and_upwinding(mesh):
    o = RansOracle(mesh, 0.5, 1e5)
    rng = np.random.default_rng(0)
    w = np.stack([rng.uniform(0.5, 1.5, 50), rng.normal(0, 0.5, 50), rng.normal(0, 0.5, 50),
                  rng.uniform(0.5, 1.0, 50), rng.uniform(0, 5, 50)], 1)
    w2 = np.stack([rng.uniform(0.5, 1.5, 50), rng.normal(0, 0.5, 50), rng.normal(0, 0.5, 50),
                   rng.uniform(0.5, 1.0, 50), rng.uniform(0, 5, 50)], 1)
    th = rng.uniform(0, 2 * np.pi, 50)
    n = np.stack([np.cos(th), np.sin(th)], 1)
    F = o.roe(torch.tensor(w), torch.tensor(w), torch.tensor(n)).numpy()
    Fp = np.array([_phys_flux(w[i], n[i]) for i in range(50)])
    assert np.abs(F - Fp).max() < 1e-13
    Fa = o.roe(torch.tensor(w), torch.tensor(w2), torch.tensor(n)).numpy()
    Fb = o.roe(torch.tensor(w2), torch.tensor(w), torch.tensor(-n)).numpy()
    assert np.abs(Fa + Fb).max() < 1e-13
    # supersonic along n on both sides: the upwind state's flux
    ws = w.copy()
    ws2 = w2.copy()
    a1 = np.sqrt(GAM * ws[:, 3] / ws[:, 0])
    a2 = np.sqrt(GAM * ws2[:, 3] / ws2[:, 0])
    ws[:, 1:3] = 2.5 * a1[:, None] * n
    ws2[:, 1:3] = 2.5 * a2[:, None] * n
    Fs = o.roe(torch.tensor(ws), torch.tensor(ws2), torch.tensor(n)).numpy()
    Fl = np.array([_phys_flux(ws[i], n[i]) for i in range(50)])
    assert np.abs(Fs - Fl).max() < 1e-12 * np.abs(Fl).max()


def test_viscous_flux_of_a_shear_layer(mesh):
    """u = a y, v = 0: tau_xy = (mu + mu_t) a, tau_xx = tau_yy = 0; the energy
    flux u tau.n plus conduction k dT/dn."""
    o = RansOracle(mesh, 0.5, 1e4)
    a, gT = 0.7, np.array([0.2, -0.1])
    wf = torch.tensor([[1.1, 0.3, 0.0, 0.8, 2.0]], dtype=DT)
    n = np.array([0.6, 0.8])
    t = lambda x: torch.tensor(np.asarray(x, float)[None], dtype=DT)   # noqa: E731
    F = o.visc_flux(wf, t([0.0, a]), t([0.0, 0.0]), t(gT), t([0.0, 0.0]), t(n)).numpy()[0]
    T = GAM * 0.8 / 1.1
    mu = o.mu_inf * T**1.5 * (1 + 110.4 / 288.15) / (T + 110.4 / 288.15)
    X = 1.1 * 2.0 * o.nu_inf / mu
    mut = 1.1 * 2.0 * o.nu_inf * X**3 / (X**3 + CV1**3)
    tau = (mu + mut) * a
    k = (mu / 0.72 + mut / 0.9) / (GAM - 1)
    assert np.allclose(F[:3], [0.0, tau * n[1], tau * n[0]], rtol=1e-13, atol=1e-18)
    assert abs(F[3] - (0.3 * tau * n[1] + k * gT @ n)) < 1e-13 * abs(F[3])


def test_sa_function_identities(mesh):
    """fv1(chi = cv1) = 1/2 and fw(r = 1) = 1: with nu~ = kappa^2 d^2 S~ (r = 1)."""
    o = RansOracle(mesh, 0.5, 1e5)
    rho = torch.tensor([1.0], dtype=DT)
    T = torch.tensor([1.0], dtype=DT)
    mu = o.mu_lam(T)
    chi = mu * CV1 / (rho * o.nu_inf)
    _, _, X, fv1, fn = o.sa_terms(rho, chi, T)
    assert abs(float(fv1) - 0.5) < 1e-14 and float(fn) == 1.0
    r = torch.tensor([1.0], dtype=DT)
    gg = r + 0.3 * (r**6 - r)
    fw = gg * ((1 + CW3**6) / (gg**6 + CW3**6)) ** (1.0 / 6.0)
    assert abs(float(fw) - 1.0) < 1e-15
    # the source's destruction with r = 1 equals cw1 rho nu~ chi / d^2
    d = 0.01
    nut = 40.0 * o.nu_inf
    St = nut / (0.41**2 * d**2)
    chi = torch.tensor([nut / o.nu_inf], dtype=DT)
    X = float(chi) * o.nu_inf / float(o.mu_lam(T))
    fv1 = X**3 / (X**3 + CV1**3)
    fv2 = 1 - X / (1 + X * fv1)
    Om = St - nut * fv2 / (0.41**2 * d**2)       # vorticity making S~ = St
    g = torch.zeros((1, 5, 2), dtype=DT)
    g[0, 2, 0] = Om
    w = torch.tensor([[1.0, 0.0, 0.0, 1 / GAM, float(chi)]], dtype=DT)
    S = float(o.sa_source(w, g, torch.tensor([d], dtype=DT)))
    cw1 = 0.1355 / 0.41**2 + (1 + 0.622) / (2 / 3)
    expect = 0.1355 * St * float(chi) - cw1 * nut * float(chi) / d**2
    assert abs(S - expect) < 1e-10 * abs(expect)


def test_gradients_exact_for_linear_fields(mesh):
    o = RansOracle(mesh, 0.5, 1e5)
    xc = mesh.centroid
    w = torch.zeros((o.nc, 5), dtype=DT)
    w[:, 0] = torch.tensor(1 + 0.3 * xc[:, 0] - 0.2 * xc[:, 1])
    w[:, 3] = torch.tensor(0.7 - 0.1 * xc[:, 0] + 0.4 * xc[:, 1])
    g = o.gradients(w, o.wall_values(w)).numpy()
    off = np.ones(o.nc, bool)
    off[o.cw] = False
    assert np.abs(g[off, 0] - [0.3, -0.2]).max() < 1e-12 and np.abs(g[off, 3] - [-0.1, 0.4]).max() < 1e-12


def test_residual_preserves_freestream_and_mirror_symmetry(mesh):
    o = RansOracle(mesh, 0.5, 1e5)
    winf = o.freestream(0.0)
    Q = o.cons(torch.tensor(np.tile(winf, (o.nc, 1))))
    R = o.residual(Q, winf, 0.0).numpy()
    far = mesh.wall_dist > 0.5
    assert np.abs(R[far, :4]).max() < 1e-14
    # NACA 0012 at alpha 0: a state mirrored about y = 0 gives a mirrored residual
    xc = mesh.centroid
    key = np.round(xc[:, 0], 9) * 1e3 + np.round(np.abs(xc[:, 1]), 9)
    order = np.lexsort((np.sign(xc[:, 1]), key))
    pair = np.empty(o.nc, int)
    a, b = order[0::2], order[1::2]
    ok = np.allclose(xc[a, 0], xc[b, 0], atol=1e-9) and np.allclose(xc[a, 1], -xc[b, 1], atol=1e-9)
    if not ok:
        pytest.skip("mesh cells are not mirror pairs")
    pair[a], pair[b] = b, a
    W = o.prim(_state(o, 0.0, 3)).numpy()
    Wm = W.copy()
    sym = 0.5 * (W + W[pair] * np.array([1, 1, -1, 1, 1]))
    Wm = sym
    Qs = o.cons(torch.tensor(Wm))
    Rs = o.residual(Qs, winf, 0.0).numpy()
    Rm = Rs[pair] * np.array([1, 1, -1, 1, 1])
    assert np.abs(Rs - Rm).max() < 1e-12 * np.abs(Rs).max()


def test_residual_telescopes_to_boundary_fluxes(mesh):
    """Interior face fluxes cancel in the sum over cells: with the source off
    (laminar, rho chi row replaced), sum R = wall + far-field fluxes."""
    o = RansOracle(mesh, 0.5, 1e4, turbulent=False, limiter_zone="off")
    Q = _state(o, 2.0, 5)
    winf = o.freestream(2.0)
    R = o.residual(Q, winf, 0.0).numpy()
    o2 = RansOracle(mesh, 0.5, 1e4, turbulent=False, limiter_zone="off")
    o2.J = dict(o.J)
    o2.J["A_i"] = torch.zeros_like(o.J["A_i"])       # interior fluxes off
    R2 = o2.residual(Q, winf, 0.0).numpy()
    assert np.abs(R[:, :4].sum(0) - R2[:, :4].sum(0)).max() < 1e-12


def test_jvp_matches_central_differences_and_is_linear(mesh):
    o = RansOracle(mesh, 0.5, 2e4, limiter_zone="off")
    Q = _state(o, 2.0, 6)
    winf = o.freestream(2.0)
    V = torch.tensor(np.random.default_rng(1).normal(size=(o.nc, 5)))
    J = o.jvp(Q, V, winf, 0.1).numpy()
    h = 1e-7
    fd = ((o.residual(Q + h * V, winf, 0.1) - o.residual(Q - h * V, winf, 0.1)) / (2 * h)).numpy()
    err = np.abs(J - fd).max(0) / np.abs(J).max(0)
    assert err[:4].max() < 1e-6 and err[4] < 1e-3
    J2 = o.jvp(Q, 2 * V, winf, 0.1).numpy()
    assert np.abs(J2 - 2 * J).max() < 1e-12 * np.abs(J).max()


def test_face_jacobians_match_differences(mesh):
    o = RansOracle(mesh, 0.5, 2e4)
    Q = _state(o, 2.0, 7)
    winf = o.freestream(2.0)
    jl, jr, jw, jf, js = o.jac1_parts(Q, winf, 0.05)
    J = o.J
    f = 17
    qL, qR = Q[J["iL"][f]].clone(), Q[J["iR"][f]].clone()
    h = 1e-7
    for k in range(5):
        e = torch.zeros(5, dtype=DT)
        e[k] = h
        dl = (o._face_int(qL + e, qR, J["n_i"][f], J["d_i"][f], J["dist_i"][f])
              - o._face_int(qL - e, qR, J["n_i"][f], J["d_i"][f], J["dist_i"][f])) / (2 * h)
        dr = (o._face_int(qL, qR + e, J["n_i"][f], J["d_i"][f], J["dist_i"][f])
              - o._face_int(qL, qR - e, J["n_i"][f], J["d_i"][f], J["dist_i"][f])) / (2 * h)
        assert torch.allclose(jl[f, :, k], dl, rtol=1e-6, atol=1e-9)
        assert torch.allclose(jr[f, :, k], dr, rtol=1e-6, atol=1e-9)


def test_block_ilu_exact_for_block_tridiagonal():
    """A block-tridiagonal matrix has no fill: ILU(0) is its exact LU (the
    reference's tests/test_rans.py:13-31 pattern)."""
    rng = np.random.default_rng(0)
    n, k = 40, 5
    rows, cols, blocks = [], [], []
    for i in range(n):
        for j in (i - 1, i, i + 1):
            if 0 <= j < n:
                rows.append(i)
                cols.append(j)
                b = rng.normal(size=(k, k)) * 0.3
                if i == j:
                    b += 4 * np.eye(k)
                blocks.append(b)
    A = np.zeros((n * k, n * k))
    for r, c, b in zip(rows, cols, blocks):
        A[r * k:(r + 1) * k, c * k:(c + 1) * k] = b
    ilu = BlockILU(n, rows, cols, k=k, perm=np.arange(n)).factor(blocks)
    x = rng.normal(size=n * k)
    assert np.allclose(ilu.solve(A @ x), x, atol=1e-10)
    # the product's structure helper gives the same pattern and order
    perm, r, c, pos = ilu_structure(n, rows, cols, fill=0, perm=np.arange(n))
    assert np.array_equal(r, ilu.r) and np.array_equal(c, ilu.c) and np.array_equal(pos, ilu.src_pos)


def test_block_ilu_with_full_fill_is_exact_and_fill_grows_the_pattern(mesh):
    o = RansOracle(mesh, 0.5, 2e4)
    Q = _state(o, 2.0, 8)
    winf = o.freestream(2.0)
    D, off = o.jac1_blocks(Q, winf, 0.0)
    D = D.numpy().copy()
    D[:, np.arange(5), np.arange(5)] += (o.vol / o.local_dt(Q, 5.0))[:, None]
    r, c = o.pattern()
    blocks = np.concatenate([D, off.numpy()])
    n, k = o.nc, 5
    A = np.zeros((n * k, n * k))
    for rr, cc, b in zip(r, c, blocks):
        A[rr * k:(rr + 1) * k, cc * k:(cc + 1) * k] += b
    nnz = [len(BlockILU(n, r, c, fill=f).r) for f in (0, 1, 3)]
    assert nnz[0] < nnz[1] < nnz[2]
    full = BlockILU(n, r, c, fill=n).factor(blocks)
    x = np.random.default_rng(2).normal(size=n * k)
    assert np.allclose(full.solve(A @ x), x, atol=1e-8 * np.abs(x).max())


def test_gmres_left_and_right_converge():
    """The reference's tests/test_rans.py:201-209 pattern: a well-conditioned
    system with a rough preconditioner converges to rtol both ways, and the
    residual estimate decreases monotonically."""
    rng = np.random.default_rng(3)
    n = 60
    A = np.eye(n) * 4 + rng.normal(size=(n, n)) * 0.3
    b = rng.normal(size=n)
    Pm = np.diag(1 / np.diag(A))
    for right in (False, True):
        x, info, h = gmres(lambda v: A @ v, b, prec=lambda v: Pm @ v, rtol=1e-10, maxiter=60, right=right)
        assert info == 0 and np.linalg.norm(A @ x - b) < 1e-8 * np.linalg.norm(b)
        assert all(h[i + 1] <= h[i] * (1 + 1e-12) for i in range(len(h) - 1))


def test_local_dt_scales_with_cfl_and_forces_vanish_for_uniform_pressure(mesh):
    o = RansOracle(mesh, 0.5, 1e5)
    winf = o.freestream(2.0)
    w = np.tile(winf, (o.nc, 1))
    w[:, 1:3] = 0.0
    Q = o.cons(torch.tensor(w)).numpy()
    dt1, dt3 = o.local_dt(Q, 1.0), o.local_dt(Q, 3.0)
    assert np.all(dt1 > 0) and np.allclose(dt3, 3 * dt1, rtol=1e-14)
    f = o.forces(Q, 2.0)
    assert abs(f["cl"]) < 1e-12 and abs(f["cd"]) < 1e-12 and abs(f["cm"]) < 1e-12


def test_raw_gradient_options_touch_only_their_terms(mesh):
    """sa_raw_grad / visc_raw_grad (aero2d/rans.py:309-310, 868, 897, 961)
    read the gradients before limiting and far-field zeroing. At first order
    (no limiting, no zeroing) they change nothing. With the limiter in the
    trailing-edge zone: visc_raw_grad leaves continuity (no viscous flux) alone and changes the
    other equations only in cells whose gradient is limited or zeroed and in
    their face neighbours; sa_raw_grad changes only the rho chi equation and
    only in those cells. A raw gradient leaking into the reconstruction would
    move continuity; a source read at the neighbours would spread the change."""
    base = dict(limiter_zone="te")
    o0 = RansOracle(mesh, 0.5, 2e4, **base)
    Q = _state(o0, 2.0, 21, amp=0.2)
    winf = o0.freestream(2.0)
    ov = RansOracle(mesh, 0.5, 2e4, visc_raw_grad=True, **base)
    osr = RansOracle(mesh, 0.5, 2e4, sa_raw_grad=True, **base)
    for order in (1,):
        R0 = o0.residual(Q, winf, 0.1, order=order)
        assert torch.equal(ov.residual(Q, winf, 0.1, order=order), R0)
        assert torch.equal(osr.residual(Q, winf, 0.1, order=order), R0)
    R0, Rv, Rs = (o.residual(Q, winf, 0.1) for o in (o0, ov, osr))
    phi = o0.phi(Q).numpy()
    diff = (phi < 1).any(1) & o0.J["lim_te"].numpy()
    diff[o0.cf] = True                         # far-field cells: zeroed gradients
    assert 0 < diff.sum() < o0.nc
    near = diff.copy()
    L, R = o0.iL, o0.iR
    near[L[diff[R]]] = True
    near[R[diff[L]]] = True
    dv = (Rv - R0).abs().numpy()
    assert torch.equal(Rv[:, 0], R0[:, 0])
    assert dv[~near].max() == 0.0 and dv[near, 1:].max() > 1e-8
    ds = (Rs - R0).abs().numpy()
    assert torch.equal(Rs[:, :4], R0[:, :4])
    assert ds[~diff, 4].max() == 0.0 and ds[diff, 4].max() > 1e-8
