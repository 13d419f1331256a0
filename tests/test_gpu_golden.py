"""End-to-end FP64 GPU solves against the reference's published goldens.

benchmarks/results.json:77-131 of the reference: NACA 0012, naca4("0012", 65),
ogrid(nj=49, ds0=2e-3, radius=40) (128 x 48 cells), FP64:
  M 0.5, alpha 2:     Cl 0.28162520444340794, Cd -1.982688129501518e-4
  M 0.8, alpha 1.25:  Cl 0.3486395236312213,  Cd 0.02198245225446649
both converged (-13.15 and -12.01 orders after 80 Newton iterations). The GPU
FP64 form must reach rn/r0 <= 1e-10 and match Cl, Cd to 1e-8 (absolute)."""
import numpy as np
import pytest

from aero2d_amd import EulerSolver, ImplicitEuler, naca4, ogrid

pytestmark = pytest.mark.gpu

GOLD = {(0.5, 2.0): (0.28162520444340794, -1.982688129501518e-4),
        (0.8, 1.25): (0.3486395236312213, 0.02198245225446649)}


def test_fp64_solve_matches_reference_goldens():
    g = ogrid(naca4("0012", 65), nj=49, ds0=2e-3, radius=40.0)
    s = EulerSolver(g, dtype=np.float64)
    r = ImplicitEuler(s, per_case_stop=True).solve([0.5, 0.8], [2.0, 1.25], n_iter=120, tol=1e-11,
                                                   verbose=False)
    print("iterations", r.iterations_per_case, "cl", repr(r.cl), "cd", repr(r.cd), "cm", repr(r.cm))
    assert np.all(r.converged_tol)
    for k, key in enumerate(GOLD):
        cl, cd = GOLD[key]
        assert abs(r.cl[k] - cl) < 1e-8, (key, r.cl[k] - cl)
        assert abs(r.cd[k] - cd) < 1e-8, (key, r.cd[k] - cd)


def test_fp64_reference_trajectory():
    """80 Newton iterations of the reference batch of three reproduce the
    published FP64 residual drops (results.json:125-129): -13.15, -12.01, -4.63."""
    g = ogrid(naca4("0012", 65), nj=49, ds0=2e-3, radius=40.0)
    s = EulerSolver(g, dtype=np.float64)
    r = ImplicitEuler(s).solve([0.5, 0.8, 0.85], [2.0, 1.25, 1.0], n_iter=80, tol=1e-6,
                               verbose=False)
    drop = np.log10(r.res_hist[-1] / r.res_hist[0])
    print("drops", drop, "cl", repr(r.cl), "cd", repr(r.cd))
    pub = np.array([-13.152765674118024, -12.005115574724915, -4.628039205423194])
    assert np.all(np.abs(drop - pub) < 0.1), drop
    assert abs(r.cl[0] - 0.28162520444340794) < 1e-8
    assert abs(r.cl[1] - 0.3486395236312213) < 1e-8
    assert abs(r.cl[2] - 0.3399757887379958) < 1e-4
    assert abs(r.cd[2] - 0.05637206522325736) < 1e-4
