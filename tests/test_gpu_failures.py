"""Failure handling of the batched solve (aero2d/implicit.py:213-233): a case
whose step is not finite keeps its state, retries at CFL 1 and is frozen after
five failures; the other cases of the batch are untouched (bit-identical to a
clean run). The failure is forced from the host through a NaN initial state."""
import numpy as np
import pytest

from aero2d_amd import EulerSolver, ImplicitEuler, naca4, ogrid

pytestmark = pytest.mark.gpu


def test_poisoned_case_freezes_and_others_are_unaffected():
    g = ogrid(naca4("0012", 33), nj=17, ds0=1e-2, radius=10.0)
    s = EulerSolver(g, dtype=np.float64)
    machs, alphas = [0.5, 0.6, 0.4], [2.0, 1.0, 0.0]
    clean = ImplicitEuler(s, per_case_stop=True).solve(machs, alphas, n_iter=60, tol=1e-9, verbose=False)
    U0 = np.stack([s.init_state(s.freestream(m, a)) for m, a in zip(machs, alphas)])
    U0[1, 0, 5, 3] = np.nan
    bad = ImplicitEuler(s, per_case_stop=True).solve(machs, alphas, n_iter=60, tol=1e-9, verbose=False, U0=U0)
    solver_frozen = bad.converged is not None and not bad.converged[1]
    assert solver_frozen
    assert np.isnan(bad.res_hist[0, 1])
    for k in (0, 2):
        assert bad.cl[k] == clean.cl[k] and bad.cd[k] == clean.cd[k]
        assert bad.converged_tol[k]


def test_callback_stops_the_solve():
    g = ogrid(naca4("0012", 33), nj=17, ds0=1e-2, radius=10.0)
    s = EulerSolver(g, dtype=np.float64)
    seen = []
    solver = ImplicitEuler(s)
    solver.solve([0.5], [2.0], n_iter=50, verbose=False,
                 callback=lambda it, n, lr, cl: seen.append(it) or it >= 4)
    assert seen == [1, 2, 3, 4] and solver.iterations == 4
