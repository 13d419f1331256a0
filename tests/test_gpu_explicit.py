"""The explicit scheme on the GPU (a2d_explicit_cycles, EulerSolver.solve) vs
the CPU oracle's restatement of the reference's explicit solver
(oracle/explicit.py, aero2d/euler.py:298-435), FP64.

Tolerances: the kernels evaluate the oracle's formulas in another order (the
periodic smoothing by factored recursions instead of an FFT, the tridiagonal
solve by an elimination with precomputed pivots), so one cycle agrees at
round-off, 1e-11 relative to the state; a run of cycles keeps that up to the
growth of round-off through the iteration, 1e-9."""
import numpy as np
import pytest
import torch

from aero2d_amd import EulerSolver, ImplicitEuler, _abi
from oracle import explicit, flow
from tests.helpers import perturbed_state, rel_err, small_grid, winf_t

pytestmark = pytest.mark.gpu


def _oracle_cycle(g, U, mach, alpha, eps, order, mg, ncycles=1, fresh=True):
    """ncycles of the oracle's cycle with Gamma updated after each (0 at the start)."""
    lev = explicit.make_levels(g.x, g.y, eps, order, mg)
    winf = winf_t(mach, alpha)
    Ut = torch.tensor(U)
    circ = torch.zeros((), dtype=torch.float64)
    res = None
    for _ in range(ncycles):
        Ut, res = explicit.cycle(Ut, winf, circ, lev, 2.5, eps, 0.02, True)
        cl = flow.forces(Ut, winf, lev.g)[0]
        circ = 0.5 * cl * torch.sqrt(winf[1] ** 2 + winf[2] ** 2)
    cl, cd, cm, _ = flow.forces(Ut, winf, lev.g)
    return Ut.numpy(), float(res), float(cl), float(cd), float(circ)


def annulus(ni, nc, R=40.0):
    """O-grid between a circle of radius 0.5 and one of radius R (tall cases
    the airfoil grid generator does not make)."""
    from aero2d_amd.grid import OGrid
    th = 2 * np.pi * np.arange(ni) / ni
    r = np.geomspace(0.5, R, nc + 1)
    return OGrid(r[None, :] * np.cos(th)[:, None] + 0.5, r[None, :] * np.sin(th)[:, None])


# ni = 32, 200, 512, 1024 (every lane-chunk width of the i smoothing); nc = 8
# .. 300 (segment lengths 1 .. 64 of the j smoothing, partial segments), and
# nc = 520 (longer columns: the unsegmented j kernel)
GRIDS = [(17, 9), (101, 9), (257, 17), (513, 9), ("annulus", 64, 256), ("annulus", 32, 300),
         ("annulus", 16, 520)]


@pytest.mark.parametrize("eps", [0.6, 0.0])
@pytest.mark.parametrize("grid", GRIDS)
def test_one_iteration_matches_oracle(grid, eps):
    g = annulus(*grid[1:]) if grid[0] == "annulus" else small_grid(*grid)
    s = EulerSolver(g, dtype=np.float64, irs_eps=eps)
    mach, alpha = 0.6, 1.5
    U = perturbed_state(s.ni, s.nc, mach, alpha, 11)
    pool, ids = s.explicit_batch([mach], [alpha], U)
    try:
        out = pool.explicit_cycles(ids, 1, fresh=True)
        Ug = np.empty_like(U)
        _abi.call("a2d_case_get_state", pool.handle, ids[0], _abi.dptr(Ug))
    finally:
        pool.close()
    Uo, res, cl, cd, circ = _oracle_cycle(g, U, mach, alpha, eps, 2, 1)
    print(grid, eps, rel_err(Ug, Uo), out[0], res, cl)
    assert rel_err(Ug, Uo) < 1e-11
    assert abs(out[0, 0] - res) <= 1e-10 * res
    assert abs(out[0, 1] - cl) < 1e-11 and abs(out[0, 2] - cd) < 1e-11
    assert abs(out[0, 4] - circ) < 1e-11


@pytest.mark.parametrize("order", [2, 1])
@pytest.mark.parametrize("mg", [2, 3])
def test_fas_cycles_match_oracle(mg, order):
    """V-cycles over two and three levels (the coarsest 8 x 4 cells on 32 x 16)."""
    g = small_grid(17, 17)
    s = EulerSolver(g, dtype=np.float64, mg_levels=mg, order=order)
    mach, alpha = 0.75, 1.0
    U = perturbed_state(s.ni, s.nc, mach, alpha, 5)
    pool, ids = s.explicit_batch([mach], [alpha], U)
    try:
        out1 = pool.explicit_cycles(ids, 1, fresh=True)
        U1 = np.empty_like(U)
        _abi.call("a2d_case_get_state", pool.handle, ids[0], _abi.dptr(U1))
        out = pool.explicit_cycles(ids, 4, fresh=False)
        U5 = np.empty_like(U)
        _abi.call("a2d_case_get_state", pool.handle, ids[0], _abi.dptr(U5))
    finally:
        pool.close()
    Uo, res, cl, _, _ = _oracle_cycle(g, U, mach, alpha, 0.6, order, mg, 1)
    assert rel_err(U1, Uo) < 1e-11
    assert abs(out1[0, 0] - res) <= 1e-10 * res and abs(out1[0, 1] - cl) < 1e-11
    Uo, res, cl, _, _ = _oracle_cycle(g, U, mach, alpha, 0.6, order, mg, 5)
    assert rel_err(U5, Uo) < 1e-9
    assert abs(out[0, 0] - res) <= 1e-8 * res and abs(out[0, 1] - cl) < 1e-9


def test_solve_matches_oracle_chunks():
    """EulerSolver.solve against the oracle's solve, the same cycles (no stop)."""
    g = small_grid(13, 7)
    s = EulerSolver(g, dtype=np.float64, mg_levels=2)
    r = s.solve([0.75], [1.0], n_iter=60, chunk=20, tol=0.0, tol_cl=0.0, tol_cd=0.0, verbose=False)
    o = explicit.solve(g.x, g.y, 0.75, 1.0, n_iter=60, chunk=20, tol=0.0, tol_cl=0.0, tol_cd=0.0, mg_levels=2)
    assert s.iterations == o.iterations == 60
    assert r.res_hist.shape == (3, 1) and s.cl_history.shape == (3, 1)
    assert np.allclose(r.res_hist[:, 0], o.res_hist, rtol=1e-8, atol=0)
    assert np.allclose(s.cl_history[:, 0], o.cl_hist, rtol=0, atol=1e-10)
    assert abs(r.cl[0] - o.cl) < 1e-10 and abs(r.cd[0] - o.cd) < 1e-10 and abs(r.cm[0] - o.cm) < 1e-10
    assert rel_err(r.cp[0], o.cp) < 1e-9
    assert rel_err(r.U[0], o.U.numpy()) < 1e-9


def test_batch_entries_are_independent():
    """A case gives the same bits alone and in a batch of three."""
    g = small_grid(17, 9)
    s = EulerSolver(g, dtype=np.float64, mg_levels=2, stencil_march=0)
    machs, alphas = [0.5, 0.7, 0.8], [2.0, -1.0, 1.25]
    pool, ids = s.explicit_batch(machs, alphas)
    try:
        outb = pool.explicit_cycles(ids, 7, fresh=True)
        Ub = [np.empty((4, s.ni, s.nc)) for _ in ids]
        for U, c in zip(Ub, ids):
            _abi.call("a2d_case_get_state", pool.handle, c, _abi.dptr(U))
    finally:
        pool.close()
    for b in range(3):
        pool, ids = s.explicit_batch(machs[b:b + 1], alphas[b:b + 1])
        try:
            out = pool.explicit_cycles(ids, 7, fresh=True)
            U = np.empty((4, s.ni, s.nc))
            _abi.call("a2d_case_get_state", pool.handle, ids[0], _abi.dptr(U))
        finally:
            pool.close()
        assert np.array_equal(out[0], outb[b]) and np.array_equal(U, Ub[b])


def test_explicit_reaches_the_implicit_steady_state():
    """The explicit solve (two-level FAS, M 0.75) and the implicit FP64 solve
    find the same zero of the residual."""
    g = small_grid(13, 7)
    r = EulerSolver(g, dtype=np.float64, mg_levels=2).solve([0.75], [1.0], n_iter=1000, chunk=100, tol=1e-10,
                                                              tol_cl=0.0, tol_cd=0.0, verbose=False)
    assert r.res_hist[-1, 0] / r.res_hist[0, 0] < 1e-10
    ri = ImplicitEuler(EulerSolver(g, dtype=np.float64)).solve([0.75], [1.0], n_iter=60, tol=1e-12, verbose=False)
    assert abs(r.cl[0] - ri.cl[0]) < 1e-8 and abs(r.cd[0] - ri.cd[0]) < 1e-8


def test_explicit_refusals():
    g = small_grid(17, 9)
    s = EulerSolver(g, dtype=np.float64)
    pool = _abi.Pool(0, s.knobs())
    try:
        a = pool.add_case(g.x, g.y, 0.5, 1.0, _abi.PREC_FP64)
        m = pool.add_case(g.x, g.y, 0.5, 1.0, _abi.PREC_MIXED)
        with pytest.raises(_abi.HipError) as e:
            pool.explicit_cycles([m], 1, True)
        assert e.value.code == -5
        b = pool.add_case(g.x, g.y, 0.5, 1.0, _abi.PREC_FP64)
        with pytest.raises(_abi.HipError) as e:     # not half the size
            pool.link_coarse(a, b)
        assert e.value.code == -6
        c = pool.add_case(g.x[::2, ::2], g.y[::2, ::2], 0.5, 1.0, _abi.PREC_FP64)
        pool.link_coarse(a, c)
        with pytest.raises(_abi.HipError) as e:     # b has no coarse level
            pool.explicit_cycles([a, b], 1, True)
        assert e.value.code == -6
        pool.park(a, True)
        with pytest.raises(_abi.HipError) as e:
            pool.explicit_cycles([a], 1, True)
        assert e.value.code == -6
        pool.park(a, False)
        pool.remove_case(c)                          # unlinks
        out = pool.explicit_cycles([a, b], 1, True)
        assert np.all(np.isfinite(out))
    finally:
        pool.close()
    with pytest.raises(ValueError):
        EulerSolver(small_grid(17, 8), dtype=np.float64, mg_levels=2)   # 32 x 7 cells
    long_rows = EulerSolver(annulus(2050, 2), dtype=np.float64)          # rows beyond 2048 cells
    pool, ids = long_rows.explicit_batch([0.5], [1.0])
    try:
        with pytest.raises(_abi.HipError) as e:
            pool.explicit_cycles(ids, 1, True)
        assert e.value.code == -6
    finally:
        pool.close()
