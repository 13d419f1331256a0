"""Pins of the explicit-solver oracle (oracle/explicit.py, restating
aero2d/euler.py:298-436), on the CPU:

- residual smoothing solves (-eps, 1 + 2 eps, -eps) periodic along i and the
  same with 1 + eps end rows along j: applying both operators to the smoothed
  field gives the input back;
- restriction is the volume-weighted mean of the four fine cells, prolongation reproduces constant
  and linear corrections;
- the explicit solve (two-level FAS) drives the residual to zero, its state is
  a fixed point of the RK iteration and of the FAS cycle, and Newton steps of
  the implicit oracle from it leave it where it is.
"""
import numpy as np
import pytest
import torch

from oracle import explicit, flow, newton
from tests.helpers import perturbed_state, small_grid


def _apply_irs(x, eps):
    """(C_i C_j) x with the smoothing operators written out."""
    y = (1 + 2 * eps) * x - eps * (torch.roll(x, 1, 1) + torch.roll(x, -1, 1))
    z = (1 + 2 * eps) * y
    z[:, :, 0] = (1 + eps) * y[:, :, 0]
    z[:, :, -1] = (1 + eps) * y[:, :, -1]
    z[:, :, 1:] -= eps * y[:, :, :-1]
    z[:, :, :-1] -= eps * y[:, :, 1:]
    return z


@pytest.mark.parametrize("eps", [0.6, 0.25])
def test_smoothing_inverts_its_operators(eps):
    g = small_grid(17, 9)
    lev = explicit.make_levels(g.x, g.y, irs_eps=eps)
    R = torch.tensor(np.random.default_rng(1).normal(size=(4, lev.ni, lev.nc)))
    S = explicit.smooth(R, lev, eps)
    assert torch.max(torch.abs(_apply_irs(S, eps) - R)) < 1e-12
    assert explicit.smooth(R, lev, 0.0) is R


def test_restriction_and_prolongation():
    g = small_grid(17, 9)
    lev = explicit.make_levels(g.x, g.y, mg_levels=2)
    vol = lev.g["vol"]
    U = torch.tensor(perturbed_state(lev.ni, lev.nc, 0.5, 1.0, 3))
    Uc = explicit.sum4(U * vol[None]) / explicit.sum4(vol)[None]
    # the coarse value is the volume-weighted mean of its four fine cells
    assert torch.allclose((Uc * explicit.sum4(vol)[None]).sum((1, 2)), (U * vol[None]).sum((1, 2)),
                          rtol=1e-13, atol=0)
    assert torch.allclose(Uc[:, 1, 2], (U[:, 2:4, 4:6] * vol[2:4, 4:6]).sum((1, 2)) / vol[2:4, 4:6].sum(),
                          rtol=1e-14, atol=0)
    c = torch.full((4, 8, 4), 0.7, dtype=torch.float64)
    assert torch.allclose(explicit.prolong(c), torch.full((4, 16, 8), 0.7, dtype=torch.float64), atol=1e-15)
    I = torch.arange(8, dtype=torch.float64)[None, :, None]
    J = torch.arange(4, dtype=torch.float64)[None, None, :]
    d = (2.0 + 0.5 * I + 0.25 * J).expand(4, 8, 4)
    f = explicit.prolong(d)
    fi = (torch.arange(16, dtype=torch.float64)[None, :, None] - 0.5) / 2
    fj = (torch.arange(8, dtype=torch.float64)[None, None, :] - 0.5) / 2
    want = 2.0 + 0.5 * fi + 0.25 * fj
    assert torch.allclose(f[:, 2:-2, 1:-1], want.expand(4, 16, 8)[:, 2:-2, 1:-1], atol=1e-14)


@pytest.fixture(scope="module")
def steady():
    """Explicit solve with a two-level FAS cycle on a 24 x 6 grid, M 0.75,
    alpha 1, to a residual drop of 1e-10."""
    g = small_grid(13, 7)
    r = explicit.solve(g.x, g.y, 0.75, 1.0, n_iter=800, chunk=100, tol=1e-10, tol_cl=0.0, tol_cd=0.0,
                       mg_levels=2)
    return g, r


def _rnorm(U, g, mach, alpha, circ=None):
    winf = flow.freestream(torch.tensor(mach, dtype=torch.float64), torch.tensor(alpha, dtype=torch.float64))
    geo = flow.case_geometry(g.x, g.y)
    if circ is None:
        circ = 0.5 * flow.forces(U, winf, geo)[0] * mach
    R, _ = flow.residual(U, winf, circ, geo)
    return float(torch.sqrt(torch.mean(R ** 2))), winf, geo, circ


def test_explicit_solve_reaches_a_steady_state(steady):
    g, r = steady
    assert r.res_hist[-1] / r.res_hist[0] < 1e-10 and r.iterations <= 800
    n, winf, geo, _ = _rnorm(r.U, g, 0.75, 1.0)
    n0, *_ = _rnorm(flow.init_state(winf, *geo["vol"].shape), g, 0.75, 1.0, torch.zeros((), dtype=torch.float64))
    assert n < 1e-10 * n0                          # R(U) = 0 with the circulation of its own Cl


def test_steady_state_is_a_fixed_point_of_the_iteration_and_the_cycle(steady):
    g, r = steady
    _, winf, _, circ = _rnorm(r.U, g, 0.75, 1.0)
    for mg in (1, 2):
        lev = explicit.make_levels(g.x, g.y, mg_levels=mg)
        U1, res = explicit.iterate(r.U, winf, circ, lev, 2.5, 0.6, 0.02, True)
        assert float(torch.max(torch.abs(U1 - r.U))) < 1e-10 and float(res) < 1e-9
        U2, _ = explicit.cycle(r.U, winf, circ, lev, 2.5, 0.6, 0.02, True)
        assert float(torch.max(torch.abs(U2 - r.U))) < 1e-10


def test_implicit_oracle_agrees_with_the_explicit_steady_state(steady):
    """Newton steps from the explicit steady state leave Cl where it is: both
    solvers find the same zero of the residual."""
    g, r = steady
    geo = flow.stack_geometry([flow.case_geometry(g.x, g.y)])
    n = newton.solve(geo, [0.75], [1.0], n_iter=3, tol=0.0, U0=r.U[..., None])
    assert abs(float(n.cl[0]) - r.cl) < 1e-9
    assert float(torch.max(torch.abs(n.U[..., 0] - r.U))) < 1e-8
