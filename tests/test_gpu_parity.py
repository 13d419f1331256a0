"""HIP kernels vs the CPU oracle, FP64, on small grids (partial tiles included).

Tolerances: the FP64 kernels evaluate the same formulas as the oracle in a
different order (fused multiply-adds, reciprocal reuse), so agreement is at
round-off: 1e-11 relative to the field's magnitude for residual, Jacobian-vector
product and blocks, 1e-9 for the line solves (they amplify round-off by the
conditioning of the 4x4 blocks)."""
import math

import numpy as np
import pytest
import torch
from torch.func import jvp

from aero2d_amd import EulerSolver, _abi
from oracle import flow, newton
from tests.helpers import oracle_geo, perturbed_state, rel_err, small_grid, winf_t

pytestmark = pytest.mark.gpu

CASES = [  # (n_side, nj, mach, alpha, seed)
    (17, 9, 0.5, 2.0, 1),
    (25, 13, 0.8, 1.25, 2),
    (33, 21, 0.3, -3.0, 3),
]


def _setup(n_side, nj, mach, alpha, seed, precision="fp64", shape=None):
    g = small_grid(n_side, nj)
    s = EulerSolver(g, dtype=np.float64, precision=precision)
    U = perturbed_state(s.ni, s.nc, mach, alpha, seed)
    h, c = s._scratch(s.freestream(mach, alpha), U)
    if shape == "march":                 # the column-march kernels instead of LDS tiles
        k = s._pool.knobs()
        k.stencil_march = 2
        s._pool.set_knobs(k)
    elif shape is not None:
        _abi.call("a2d_case_set_shape", h, c, shape)
    return g, s, U, h, c


def test_hllc_device_matches_oracle():
    rng = np.random.default_rng(7)
    n = 4096
    WL = np.stack([rng.uniform(0.3, 2.0, n), rng.normal(0, 0.8, n), rng.normal(0, 0.8, n),
                   rng.uniform(0.2, 1.5, n)])
    WR = np.stack([rng.uniform(0.3, 2.0, n), rng.normal(0, 0.8, n), rng.normal(0, 0.8, n),
                   rng.uniform(0.2, 1.5, n)])
    S = rng.normal(size=(2, n))
    from aero2d_amd.euler import hllc_flux
    F = hllc_flux(WL, WR, S)
    Fo = flow.hllc_flux(torch.tensor(WL), torch.tensor(WR), torch.tensor(S)).numpy()
    assert rel_err(F, Fo) < 1e-13


@pytest.mark.parametrize("shape", [0, 1, 2, 3, "march"])
@pytest.mark.parametrize("case", CASES)
def test_residual_matches_oracle(case, shape):
    g, s, U, h, c = _setup(*case, shape=shape)
    mach, alpha = case[2], case[3]
    geo = oracle_geo(g)
    w = winf_t(mach, alpha)
    for circ in (0.0, 0.137, float("nan")):
        R = np.empty((4, s.ni, s.nc))
        rn = np.zeros(1)
        _abi.call("a2d_residual", h, c, circ, _abi.dptr(R), _abi.dptr(rn))
        Ut = torch.tensor(U)
        ci = flow.circulation(Ut, w, geo) if math.isnan(circ) else torch.tensor(circ, dtype=torch.float64)
        Ro, _ = flow.residual(Ut, w, ci, geo)
        assert rel_err(R, Ro.numpy()) < 1e-11, (case, shape, circ)
        rno = torch.sqrt(torch.mean((Ro[0] / geo["vol"]) ** 2)).item()
        assert abs(rn[0] - rno) <= 1e-11 * rno


@pytest.mark.parametrize("case", CASES)
def test_forces_and_local_dt_match_oracle(case):
    g, s, U, h, c = _setup(*case)
    mach, alpha = case[2], case[3]
    geo = oracle_geo(g)
    w = winf_t(mach, alpha)
    cl, cd, cm, cp = s.forces(U, s.freestream(mach, alpha))
    clo, cdo, cmo, cpo = flow.forces(torch.tensor(U), w, geo)
    assert abs(cl - clo.item()) < 1e-12 and abs(cd - cdo.item()) < 1e-12 and abs(cm - cmo.item()) < 1e-12
    assert rel_err(cp, cpo.numpy()) < 1e-12
    from aero2d_amd.euler import cons_to_prim
    dt = s.local_dt(cons_to_prim(U))
    dto = flow.local_dt(flow.cons_to_prim(torch.tensor(U)), geo).numpy()
    assert rel_err(dt, dto) < 1e-12


@pytest.mark.parametrize("shape", [0, 2, "march"])
@pytest.mark.parametrize("case", CASES)
def test_jvp_matches_oracle(case, shape):
    g, s, U, h, c = _setup(*case, shape=shape)
    mach, alpha = case[2], case[3]
    geo = oracle_geo(g)
    w = winf_t(mach, alpha)
    rng = np.random.default_rng(case[4] + 100)
    v = rng.normal(size=U.shape) * np.array([1, 0.5, 0.5, 2.0])[:, None, None]
    out = np.empty_like(U)
    _abi.call("a2d_jvp", h, c, _abi.dptr(np.ascontiguousarray(v)), 10.0, 0, _abi.dptr(out))
    res = lambda u: flow.residual_with_circ(u, w, geo)[0]
    jo = jvp(res, (torch.tensor(U),), (torch.tensor(v),))[1].numpy()
    assert rel_err(out, jo) < 1e-11
    # with the pseudo-time diagonal at CFL 10
    _abi.call("a2d_jvp", h, c, _abi.dptr(np.ascontiguousarray(v)), 10.0, 1, _abi.dptr(out))
    Wt = flow.cons_to_prim(torch.tensor(U))
    diag = (geo["vol"] / (flow.local_dt(Wt, geo) / 2.5 * 10.0)).numpy()
    assert rel_err(out, diag[None] * v + jo) < 1e-11


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("case", CASES)
def test_assembly_matches_oracle(case, order):
    """order 1 keeps the blocks colour-split on the device; a2d_assemble
    returns them in the natural layout either way."""
    g, s, U, h, c = _setup(*case)
    k = s._pool.knobs()
    k.precond_order = order
    s._pool.set_knobs(k)
    mach, alpha = case[2], case[3]
    geo = oracle_geo(g)
    w = winf_t(mach, alpha)
    ni, nc = s.ni, s.nc
    blocks = [np.empty((ni, nc, 4, 4)) for _ in range(5)]
    _abi.call("a2d_assemble", h, c, 10.0, *[_abi.dptr(b) for b in blocks])
    Ut = torch.tensor(U)
    circ = flow.circulation(Ut, w, geo)
    Wt = flow.cons_to_prim(Ut)
    dti = 1.0 / (flow.local_dt(Wt, geo) / 2.5 * 10.0)
    ref = newton.assemble(Ut, w, circ, dti, geo)
    for got, exp in zip(blocks, ref):
        assert rel_err(got, exp.numpy()) < 1e-11


@pytest.mark.parametrize("sweep_lds", [2, 0])
@pytest.mark.parametrize("case", CASES)
def test_preconditioner_matches_oracle(case, sweep_lds):
    """Line-Jacobi preconditioner, its eliminations' intermediate values in
    LDS (sweep_lds = 2) or in HBM (0)."""
    g, s, U, h, c = _setup(*case)
    mach, alpha = case[2], case[3]
    geo = oracle_geo(g)
    w = winf_t(mach, alpha)
    k = s._pool.knobs()
    k.sweep_lds = sweep_lds
    s._pool.set_knobs(k)
    _abi.call("a2d_precond_setup", h, c, 25.0)
    rng = np.random.default_rng(5)
    r = np.ascontiguousarray(rng.normal(size=U.shape))
    x = np.empty_like(r)
    _abi.call("a2d_precond_apply", h, c, _abi.dptr(r), _abi.dptr(x))
    Ut = torch.tensor(U)
    circ = flow.circulation(Ut, w, geo)
    dti = 1.0 / (flow.local_dt(flow.cons_to_prim(Ut), geo) / 2.5 * 25.0)
    D, Bim, Bip, Bjm, Bjp = newton.assemble(Ut, w, circ, dti, geo)
    fac = newton.factor_lines(D, Bjm, Bjp)
    xo = newton.precond(fac, Bim, Bip, torch.tensor(r), 4).numpy()
    assert rel_err(x, xo) < 1e-9


@pytest.mark.parametrize("sweep_lds", [2, 0])
@pytest.mark.parametrize("precision, half", [("fp64", 0), ("mixed", 0), ("mixed", 1)])
@pytest.mark.parametrize("case", CASES)
def test_red_black_preconditioner_matches_oracle(case, precision, half, sweep_lds):
    """half = 1: the sweeps read the couplings from their FP16 copy (11-bit
    significand, relative rounding 2^-11 per entry): agreement to 5e-3 of the
    FP64 oracle; a wrong slot, row or index shows as O(1)."""
    g, s, U, h, c = _setup(*case, precision=precision)
    mach, alpha = case[2], case[3]
    geo = oracle_geo(g)
    w = winf_t(mach, alpha)
    k = s._pool.knobs()
    k.precond_order = 1
    k.precond_half = half
    k.sweep_lds = sweep_lds
    s._pool.set_knobs(k)
    _abi.call("a2d_precond_setup", h, c, 25.0)
    rng = np.random.default_rng(6)
    r = np.ascontiguousarray(rng.normal(size=U.shape))
    x = np.empty_like(r)
    _abi.call("a2d_precond_apply", h, c, _abi.dptr(r), _abi.dptr(x))
    Ut = torch.tensor(U)
    circ = flow.circulation(Ut, w, geo)
    dti = 1.0 / (flow.local_dt(flow.cons_to_prim(Ut), geo) / 2.5 * 25.0)
    D, Bim, Bip, Bjm, Bjp = newton.assemble(Ut, w, circ, dti, geo)
    fac = newton.factor_lines(D, Bjm, Bjp)
    xo = newton.precond_rb(fac, Bim, Bip, torch.tensor(r), 4).numpy()
    # FP32 blocks and sweeps in the mixed form
    err = rel_err(x, xo)
    print(precision, half, "rb precond rel err", err)
    assert err < (1e-9 if precision == "fp64" else (5e-3 if half else 2e-4))
    xj = newton.precond(fac, Bim, Bip, torch.tensor(r), 4).numpy()
    assert rel_err(x, xj) > 1e-6          # a different ordering really ran


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("nj, ds0, radius", [(3, 0.2, 1.0), (4, 0.2, 1.0), (10, 2e-2, 5.0), (11, 2e-2, 5.0)])
@pytest.mark.parametrize("precision, half", [("fp64", 0), ("mixed", 1)])
@pytest.mark.parametrize("sweep_lds", [2, 0])
def test_preconditioner_line_lengths(nj, ds0, radius, order, precision, half, sweep_lds):
    """Lines of 2, 3, 9 and 10 cells: the two-sided line elimination meets in
    the middle row whatever the parity of the length (and with an empty far
    half for 2 cells). Mixed form with FP16 couplings: 5e-3."""
    g = small_grid(17, nj, ds0=ds0, radius=radius)
    s = EulerSolver(g, dtype=np.float64, precision=precision)
    U = perturbed_state(s.ni, s.nc, 0.6, 1.0, 9)
    h, c = s._scratch(s.freestream(0.6, 1.0), U)
    k = s._pool.knobs()
    k.precond_order = order
    k.precond_half = half
    k.sweep_lds = sweep_lds
    s._pool.set_knobs(k)
    _abi.call("a2d_precond_setup", h, c, 40.0)
    r = np.ascontiguousarray(np.random.default_rng(11).normal(size=U.shape))
    x = np.empty_like(r)
    _abi.call("a2d_precond_apply", h, c, _abi.dptr(r), _abi.dptr(x))
    geo = oracle_geo(g)
    w = winf_t(0.6, 1.0)
    Ut = torch.tensor(U)
    circ = flow.circulation(Ut, w, geo)
    dti = 1.0 / (flow.local_dt(flow.cons_to_prim(Ut), geo) / 2.5 * 40.0)
    D, Bim, Bip, Bjm, Bjp = newton.assemble(Ut, w, circ, dti, geo)
    fac = newton.factor_lines(D, Bjm, Bjp)
    fn = newton.precond_rb if order else newton.precond
    xo = fn(fac, Bim, Bip, torch.tensor(r), 4).numpy()
    assert rel_err(x, xo) < (1e-9 if precision == "fp64" else 5e-3)


def _largest_served(precision):
    """768 x 384 cells (294,912), the largest size the serving workload draws
    (NACA 2412 on its stretched far-field grid), at a perturbed state."""
    from aero2d_amd import naca4
    from aero2d_amd.workloads import grid_for
    g = grid_for(naca4("2412", 385), 384)[0]
    mach, alpha = 0.7, 1.5
    s = EulerSolver(g, dtype=np.float64, precision=precision)
    assert (s.ni, s.nc) == (768, 384)
    U = perturbed_state(s.ni, s.nc, mach, alpha, 4, amp=0.02)
    h, c = s._scratch(s.freestream(mach, alpha), U)
    geo = oracle_geo(g)
    w = winf_t(mach, alpha)
    Ut = torch.tensor(U)
    return g, s, U, h, c, geo, w, Ut, flow.circulation(Ut, w, geo), mach, alpha


def test_largest_served_grid_matches_oracle():
    """At 768 x 384: residual, Jacobian-vector product and both preconditioner
    orderings vs the oracle at the small-grid tolerances, then one full FP64
    Newton step with GMRES(10) from the perturbed state, where the oracle's own
    response to a 1e-15 perturbation is ~1e-15, to 1e-10."""
    from aero2d_amd import ImplicitEuler
    g, s, U, h, c, geo, w, Ut, circ, mach, alpha = _largest_served("fp64")
    R = np.empty((4, s.ni, s.nc))
    rn = np.zeros(1)
    _abi.call("a2d_residual", h, c, float("nan"), _abi.dptr(R), _abi.dptr(rn))
    Ro, _ = flow.residual(Ut, w, circ, geo)
    assert rel_err(R, Ro.numpy()) < 1e-11
    v = np.ascontiguousarray(np.random.default_rng(3).normal(size=U.shape))
    out = np.empty_like(U)
    _abi.call("a2d_jvp", h, c, _abi.dptr(v), 10.0, 0, _abi.dptr(out))
    jo = jvp(lambda u: flow.residual_with_circ(u, w, geo)[0], (Ut,), (torch.tensor(v),))[1].numpy()
    assert rel_err(out, jo) < 1e-11
    dti = 1.0 / (flow.local_dt(flow.cons_to_prim(Ut), geo) / 2.5 * 25.0)
    D, Bim, Bip, Bjm, Bjp = newton.assemble(Ut, w, circ, dti, geo)
    fac = newton.factor_lines(D, Bjm, Bjp)
    r = np.ascontiguousarray(np.random.default_rng(5).normal(size=U.shape))
    for order, fn in ((0, newton.precond), (1, newton.precond_rb)):
        k = s._pool.knobs()
        k.precond_order = order
        s._pool.set_knobs(k)
        _abi.call("a2d_precond_setup", h, c, 25.0)
        x = np.empty_like(r)
        _abi.call("a2d_precond_apply", h, c, _abi.dptr(r), _abi.dptr(x))
        assert rel_err(x, fn(fac, Bim, Bip, torch.tensor(r), 4).numpy()) < 1e-9, order
    r1 = ImplicitEuler(s, gmres_restart=10).solve([mach], [alpha], n_iter=1, verbose=False, tol=0.0,
                                                 U0=U[None])
    ro = newton.solve(flow.stack_geometry([geo]), [mach], [alpha], n_iter=1, tol=0.0,
                      p=newton.Params(restart=10), U0=Ut[..., None])
    assert rel_err(r1.U[0], ro.U[..., 0].numpy()) < 1e-10
    assert abs(r1.cl[0] - ro.cl[0]) < 1e-10


def test_largest_served_grid_mixed_form_matches_oracle():
    """The mixed form's FP32 kernels at 768 x 384 against the FP64 oracle, at
    the small-grid bars: Jacobian-vector product 1e-4, red-black preconditioner
    2e-4 with FP32 couplings and 5e-3 reading their FP16 copy (a wrong index
    or dropped term shows as O(1))."""
    g, s, U, h, c, geo, w, Ut, circ, mach, alpha = _largest_served("mixed")
    v = np.ascontiguousarray(np.random.default_rng(3).normal(size=U.shape))
    out = np.empty_like(U)
    _abi.call("a2d_jvp", h, c, _abi.dptr(v), 10.0, 0, _abi.dptr(out))
    jo = jvp(lambda u: flow.residual_with_circ(u, w, geo)[0], (Ut,), (torch.tensor(v),))[1].numpy()
    assert rel_err(out, jo) < 1e-4
    dti = 1.0 / (flow.local_dt(flow.cons_to_prim(Ut), geo) / 2.5 * 25.0)
    D, Bim, Bip, Bjm, Bjp = newton.assemble(Ut, w, circ, dti, geo)
    fac = newton.factor_lines(D, Bjm, Bjp)
    r = np.ascontiguousarray(np.random.default_rng(5).normal(size=U.shape))
    xo = newton.precond_rb(fac, Bim, Bip, torch.tensor(r), 4).numpy()
    for half, tol in ((0, 2e-4), (1, 5e-3)):
        k = s._pool.knobs()
        k.precond_order = 1
        k.precond_half = half
        s._pool.set_knobs(k)
        _abi.call("a2d_precond_setup", h, c, 25.0)
        x = np.empty_like(r)
        _abi.call("a2d_precond_apply", h, c, _abi.dptr(r), _abi.dptr(x))
        assert rel_err(x, xo) < tol, half


@pytest.mark.parametrize("march", [1, 2])
@pytest.mark.parametrize("case", CASES[:2])
def test_newton_steps_match_oracle(case, march):
    """One and three full Newton steps (50 Arnoldi steps each) vs the oracle,
    from a non-uniform state: the oracle's own response to a 1e-15 perturbation
    of such a state is 1e-15, so the GPU must agree to 1e-10. march = 2 runs
    the residual (its norm partials included) and the tangent as the column
    march."""
    n_side, nj, mach, alpha, seed = case
    g = small_grid(n_side, nj)
    from aero2d_amd import ImplicitEuler
    s = EulerSolver(g, dtype=np.float64, stencil_march=march)
    geo = flow.stack_geometry([oracle_geo(g)])
    U0 = perturbed_state(s.ni, s.nc, mach, alpha, 7, amp=0.02)
    for steps in (1, 3):
        r = ImplicitEuler(s).solve([mach], [alpha], n_iter=steps, verbose=False, tol=0.0, U0=U0[None])
        ro = newton.solve(geo, [mach], [alpha], n_iter=steps, tol=0.0, U0=torch.tensor(U0)[..., None])
        assert rel_err(r.U[0], ro.U[..., 0].numpy()) < 1e-10, steps
        assert np.allclose(r.res_hist[:, 0], ro.res_hist[:, 0], rtol=1e-10, atol=0)
        assert abs(r.cl[0] - ro.cl[0]) < 1e-10


@pytest.mark.parametrize("case", CASES[:2])
def test_newton_steps_from_freestream_match_oracle(case):
    """From the exact freestream, every interior face starts at a tie of the
    Einfeldt min/max, where round-off picks the derivative branch; the oracle
    itself then moves by ~1.5e-6 (relative U) under a 1e-15 perturbation. The
    GPU must agree within that (5e-6) after one and three steps."""
    n_side, nj, mach, alpha, seed = case
    g = small_grid(n_side, nj)
    from aero2d_amd import ImplicitEuler
    s = EulerSolver(g, dtype=np.float64)
    geo = flow.stack_geometry([oracle_geo(g)])
    for steps in (1, 3):
        r = ImplicitEuler(s).solve([mach], [alpha], n_iter=steps, verbose=False, tol=0.0)
        ro = newton.solve(geo, [mach], [alpha], n_iter=steps, tol=0.0)
        assert rel_err(r.U[0], ro.U[..., 0].numpy()) < 5e-6, steps
        assert abs(r.cl[0] - ro.cl[0]) < 5e-5


@pytest.mark.parametrize("kernel", [None, "march"])
@pytest.mark.parametrize("case", CASES)
def test_mixed_jvp_and_blocks_match_oracle(case, kernel):
    """The mixed form's inner precision: the FP32 Jacobian-vector product
    (hardware 1-ulp reciprocal and square root; LDS tiles or the column march)
    and the FP32 preconditioner blocks against the FP64 oracle. FP32 rounding
    of O(1) fluxes and their differences bounds the agreement near 1e-6 of the
    field's magnitude; the test allows 1e-4 (a dropped term or wrong index
    shows as O(1))."""
    g, s, U, h, c = _setup(*case, precision="mixed", shape=kernel)
    mach, alpha = case[2], case[3]
    geo = oracle_geo(g)
    w = winf_t(mach, alpha)
    rng = np.random.default_rng(case[4] + 200)
    v = np.ascontiguousarray(rng.normal(size=U.shape) * np.array([1, 0.5, 0.5, 2.0])[:, None, None])
    out = np.empty_like(U)
    res = lambda u: flow.residual_with_circ(u, w, geo)[0]
    jo = jvp(res, (torch.tensor(U),), (torch.tensor(v),))[1].numpy()
    _abi.call("a2d_jvp", h, c, _abi.dptr(v), 10.0, 0, _abi.dptr(out))
    err = rel_err(out, jo)
    print("fp32 jvp rel err", err)
    assert err < 1e-4
    _abi.call("a2d_jvp", h, c, _abi.dptr(v), 10.0, 1, _abi.dptr(out))
    Wt = flow.cons_to_prim(torch.tensor(U))
    diag = (geo["vol"] / (flow.local_dt(Wt, geo) / 2.5 * 10.0)).numpy()
    assert rel_err(out, diag[None] * v + jo) < 1e-4
    ni, nc = s.ni, s.nc
    blocks = [np.empty((ni, nc, 4, 4)) for _ in range(5)]
    _abi.call("a2d_assemble", h, c, 10.0, *[_abi.dptr(b) for b in blocks])
    Ut = torch.tensor(U)
    circ = flow.circulation(Ut, w, geo)
    dti = 1.0 / (flow.local_dt(flow.cons_to_prim(Ut), geo) / 2.5 * 10.0)
    ref = newton.assemble(Ut, w, circ, dti, geo)
    for got, exp in zip(blocks, ref):
        e = rel_err(got, exp.numpy())
        print("fp32 block rel err", e)
        assert e < 1e-4
