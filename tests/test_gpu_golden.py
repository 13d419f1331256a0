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


def test_config1_single_case_matches_oracle():
    """NACA 0012, 128 x 64, M 0.5, alpha 2, FP64 to rn/r0 <= 1e-10: Cl, Cd, Cm
    within 1e-8 of the oracle's solve (tests/golden/oracle_config1.json)."""
    import json
    import os
    from aero2d_amd.workloads import config1
    with open(os.path.join(os.path.dirname(__file__), "golden", "oracle_config1.json")) as f:
        ref = json.load(f)
    case = config1()[0]
    s = EulerSolver(case.grid, dtype=np.float64)
    r = ImplicitEuler(s).solve([case.mach], [case.alpha], n_iter=300, tol=1e-10, verbose=False)
    print("iterations", r.iterations_per_case, "cl", repr(r.cl[0]), "oracle", ref["cl"], ref["iterations"])
    assert r.converged_tol[0]
    assert abs(r.cl[0] - ref["cl"]) < 1e-8
    assert abs(r.cd[0] - ref["cd"]) < 1e-8
    assert abs(r.cm[0] - ref["cm"]) < 1e-8


def test_reference_integration_assertions():
    """The reference's tests/test_euler.py:43-65 on the GPU path: a 96 x 40 grid,
    M 0.5 at alpha 0 and 2 deg, 40 Newton iterations, tol 1e-5: |Cl(0)| < 2e-3,
    Cl(2) within 4% of the Prandtl-Glauert-corrected panel value (0.279,
    README.md:309-311 of the reference), |Cd| < 3e-3, residual drop < -3 orders."""
    g = ogrid(naca4("0012", 49), nj=41, ds0=3e-3)
    for dtype in (np.float32, np.float64):
        r = ImplicitEuler(EulerSolver(g, dtype=dtype)).solve([0.5, 0.5], [0.0, 2.0], n_iter=40,
                                                            verbose=False, tol=1e-5)
        assert abs(r.cl[0]) < 2e-3
        assert abs(r.cl[1] / 0.279 - 1) < 0.04
        assert abs(r.cd[1]) < 3e-3
        assert np.all(np.log10(r.res_hist[-1] / r.res_hist[0]) < -3)


def test_mixed_form_matches_fp64_on_the_same_solver_path():
    """Transonic batch (NACA 0012/2412, M 0.72-0.82, 128 x 64): the mixed form
    against the FP64 form run on the same solver path (the mixed form's
    Krylov, preconditioner and CFL settings). Both converge to rn/r0 <= 1e-10
    and agree per case to |dCl|, |dCm| <= 1e-7 and |dCd| <= 1e-8 (FP32 inner
    iterations, FP64 residual and update). Both run with the reference's step
    limit 0.2: with the mixed form's larger default limit, FP32 rounding alone
    can send a transonic case to another of its steady states, which is not a
    precision error. The sixth draw of this seed (NACA 2412, M 0.777, alpha
    0.2) is left out: its density residual stalls near 1e-3.7 in a limit cycle
    under every solver setting, the reference's included."""
    from aero2d_amd.batch import default_knobs, path_of, solve_batch
    from aero2d_amd.workloads import config3
    cases = config3(6, 128, 64)[:5]
    p64 = path_of("mixed")
    p64["knobs"].max_update = 0.2
    r64 = solve_batch(cases, "fp64", tol=1e-10, n_iter=600, **p64)
    rmx = solve_batch(cases, "mixed", tol=1e-10, n_iter=600, knobs=default_knobs("mixed", max_update=0.2))
    print("iterations", r64.iterations, rmx.iterations)
    print("dcl", np.abs(rmx.cl - r64.cl), "dcd", np.abs(rmx.cd - r64.cd))
    assert np.all(r64.converged) and np.all(rmx.converged)
    assert np.max(np.abs(rmx.cl - r64.cl)) < 1e-7
    assert np.max(np.abs(rmx.cd - r64.cd)) < 1e-8
    assert np.max(np.abs(rmx.cm - r64.cm)) < 1e-7


def test_fp16_storage_options_reach_the_same_steady_states():
    """Subsonic cases (one steady state each): the mixed form with the FP16
    coupling copy and the FP16 Krylov basis against FP32 for both. The
    storage only changes the inner iterations, so both converge to
    rn/r0 <= 1e-10 and agree to |dCl|, |dCm| <= 1e-8, |dCd| <= 1e-9, in about
    the same number of Newton steps."""
    from aero2d_amd.batch import default_knobs, solve_batch
    from aero2d_amd.workloads import config2
    cases = [c for c in config2(16, 128, 64) if c.mach < 0.6][:6]
    r16 = solve_batch(cases, "mixed", tol=1e-10, n_iter=300,
                      knobs=default_knobs("mixed", precond_half=1, basis_half=1))
    r32 = solve_batch(cases, "mixed", tol=1e-10, n_iter=300,
                      knobs=default_knobs("mixed", precond_half=0, basis_half=0))
    print("iterations", r16.iterations, r32.iterations, "arnoldi", r16.arnoldi.mean(), r32.arnoldi.mean())
    print("dcl", np.abs(r16.cl - r32.cl), "dcd", np.abs(r16.cd - r32.cd))
    assert len(cases) >= 4
    assert np.all(r16.converged) and np.all(r32.converged)
    assert np.max(np.abs(r16.cl - r32.cl)) < 1e-8
    assert np.max(np.abs(r16.cd - r32.cd)) < 1e-9
    assert np.max(np.abs(r16.cm - r32.cm)) < 1e-8
    assert r16.iterations.sum() <= 1.15 * r32.iterations.sum()
