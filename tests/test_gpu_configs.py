"""GPU tests at the workloads' own sizes and on reference branches that the
small-grid parity cases do not reach.

- 256 x 128 (the uniform FP64 batch): one Newton step of two cases against the
  oracle at 1e-10, and a case's bits alone vs in a six-case batch.
- The uneven batch's four sizes (128 x 64 .. 512 x 256) with autotuned tile
  shapes: the tuner's cache round trip, a case's bits alone vs in the mix,
  and graph replay vs eager launches.
- order = 1 (aero2d/euler.py:270-281), the supersonic far-field overrides and
  the M >= 1 branch without the vortex (aero2d/euler.py:171, 201-204): residual
  and Jacobian-vector product against the oracle at 1e-11.
- The mixed form with its shipped defaults on transonic draws against the FP64
  reference-path solve: |dCl| <= 1e-5 and |dCd| <= 1e-6, or a cross-restart
  shows both final states are steady states of the FP64 equations.
"""
import numpy as np
import pytest
import torch
from torch.func import jvp

from aero2d_amd import EulerSolver, _abi
from aero2d_amd.batch import cross_restart, default_knobs, solve_batch
from oracle import flow, newton
from tests.helpers import oracle_geo, perturbed_state, rel_err, small_grid, winf_t

pytestmark = pytest.mark.gpu


def test_full_size_uniform_batch_newton_step_matches_oracle():
    """Two cases of the 64-case batch (256 x 128, their own NACA sections),
    one FP64 Newton step with the reference's settings from a perturbed state:
    state after the step within 1e-10 (relative) of the oracle's, Cl within
    1e-10 (the oracle's response to a 1e-15 perturbation there is at
    round-off)."""
    from aero2d_amd.workloads import config2
    cases = config2(2, 256, 128)
    U0 = [perturbed_state(256, 128, c.mach, c.alpha, 30 + k, amp=0.02) for k, c in enumerate(cases)]
    r = solve_batch(cases, "fp64", tol=0.0, n_iter=1, states=U0, keep_state=True, stall=())
    for k, c in enumerate(cases):
        geo = flow.stack_geometry([oracle_geo(c.grid)])
        ro = newton.solve(geo, [c.mach], [c.alpha], n_iter=1, tol=0.0, U0=torch.tensor(U0[k])[..., None])
        err = rel_err(r.U[k], ro.U[..., 0].numpy())
        print(k, c.name, "rel err", err, "cl", r.cl[k], ro.cl[0])
        assert err < 1e-10
        assert abs(r.cl[k] - ro.cl[0]) < 1e-10


def test_full_size_case_is_the_same_alone_and_in_the_batch():
    """Six cases of 256 x 128, eight FP64 Newton steps: case 2's state has the
    same bits solved alone and inside the batch (every reduction is per case,
    in a fixed order)."""
    from aero2d_amd.workloads import config2
    cases = config2(6, 256, 128)
    full = solve_batch(cases, "fp64", tol=0.0, n_iter=8, keep_state=True, stall=())
    one = solve_batch([cases[2]], "fp64", tol=0.0, n_iter=8, keep_state=True, stall=())
    assert np.array_equal(full.U[2], one.U[0])
    assert full.cl[2] == one.cl[0] and full.cd[2] == one.cd[0]


def _uneven_cases():
    """One case of each size of the uneven batch, plus a second 128 x 64."""
    from aero2d_amd.workloads import CONFIG4_SIZES, config4
    cases, _ = config4(24)
    pick = []
    for ni, nc in CONFIG4_SIZES:
        pick.append(next(c for c in cases if c.grid.x.shape == (ni, nc + 1)))
    pick.append([c for c in cases if c.grid.x.shape == (128, 65)][1])
    return pick


def test_autotuner_cache_round_trip(tmp_path, monkeypatch):
    """shapes_for tunes each size it has no entry for, stores the fastest
    compiled shape in the JSON cache, and a second call answers from the cache
    alone (tune_missing=False) with the same shapes."""
    import json
    from aero2d_amd import autotune
    monkeypatch.setenv("AERO2D_TUNE_CACHE", str(tmp_path / "tiles.json"))
    cases = _uneven_cases()
    shapes = autotune.shapes_for(cases, "mixed")
    nshape = len(_abi.stencil_shapes())
    assert set(shapes) == {(128, 64), (256, 128), (384, 192), (512, 256)}
    assert all(0 <= s < nshape for s in shapes.values())
    data = json.loads((tmp_path / "tiles.json").read_text())
    assert len(data) == 4
    for (ni, nc), s in shapes.items():
        e = data[f"{ni}x{nc}/mixed"]
        assert e["shape"] == s and len(e["ms"]) == nshape
        assert e["ms"][str(s)] == min(e["ms"].values())
    again = autotune.shapes_for(cases, "mixed", tune_missing=False)
    assert again == shapes
    monkeypatch.setenv("AERO2D_TUNE_CACHE", str(tmp_path / "none.json"))
    assert autotune.shapes_for(cases, "mixed", tune_missing=False) == {}


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_uneven_batch_alone_equals_in_the_mix(precision):
    """The uneven batch's four sizes in one batch, each size on its own tile
    shape (not the default 16 x 16 for every size): after six Newton steps
    every case's state has the same bits as the case solved alone."""
    cases = _uneven_cases()
    shapes = {(128, 64): 1, (256, 128): 2, (384, 192): 3, (512, 256): 0}
    mix = solve_batch(cases, precision, tol=0.0, n_iter=6, keep_state=True, shapes=shapes, stall=())
    for k in (0, 2, 3):
        one = solve_batch([cases[k]], precision, tol=0.0, n_iter=6, keep_state=True, shapes=shapes, stall=())
        assert np.array_equal(mix.U[k], one.U[0]), k
        assert mix.cl[k] == one.cl[0]


def test_uneven_batch_graph_replay_equals_eager():
    """Graph-replayed Newton steps over the uneven mix (steps 2.. replayed)
    give the same bits as eager launches, step after step."""
    cases = _uneven_cases()
    shapes = {(128, 64): 1, (256, 128): 2, (384, 192): 3, (512, 256): 0}
    outs = {}
    for use_graph in (1, 0):
        kn = default_knobs("mixed")
        kn.use_graph = use_graph
        pool = _abi.Pool(0, kn)
        try:
            ids = np.array([pool.add_case(c.grid.x, c.grid.y, c.mach, c.alpha, _abi.PREC_MIXED) for c in cases],
                           np.int32)
            for k, c in enumerate(cases):
                _abi.call("a2d_case_set_shape", pool.handle, int(ids[k]), shapes[(c.grid.x.shape[0],
                                                                                   c.grid.x.shape[1] - 1)])
            hist = []
            for step in range(4):
                out = np.empty((len(ids), 10))
                cfl = np.full(len(ids), 10.0 * 1.5 ** step)
                _abi.call("a2d_newton_step", pool.handle, _abi.iptr(ids), len(ids), _abi.dptr(cfl), _abi.dptr(out))
                hist.append(out)
            c = cases[3]
            U = np.empty((4, c.grid.x.shape[0], c.grid.x.shape[1] - 1))
            _abi.call("a2d_case_get_state", pool.handle, int(ids[3]), _abi.dptr(U))
            hist.append(U)
            outs[use_graph] = hist
        finally:
            pool.close()
    for a, b in zip(outs[1], outs[0]):
        assert np.array_equal(a, b, equal_nan=True)


def _residual_and_jvp_vs_oracle(g, s, mach, alpha, U, order):
    geo = oracle_geo(g)
    w = winf_t(mach, alpha)
    h, c = s._scratch(s.freestream(mach, alpha), U)
    R = np.empty_like(U)
    rn = np.zeros(1)
    _abi.call("a2d_residual", h, c, float("nan"), _abi.dptr(R), _abi.dptr(rn))

    def res(u):
        return flow.residual(u, w, flow.circulation(u, w, geo), geo, order=order)[0]
    Ut = torch.tensor(U)
    assert rel_err(R, res(Ut).numpy()) < 1e-11
    rng = np.random.default_rng(11)
    v = np.ascontiguousarray(rng.normal(size=U.shape) * np.array([1, 0.5, 0.5, 2.0])[:, None, None])
    out = np.empty_like(U)
    _abi.call("a2d_jvp", h, c, _abi.dptr(v), 10.0, 0, _abi.dptr(out))
    jo = jvp(res, (Ut,), (torch.tensor(v),))[1].numpy()
    assert rel_err(out, jo) < 1e-11


@pytest.mark.parametrize("mach, alpha", [(0.5, 2.0), (0.8, 1.25)])
def test_first_order_residual_and_jvp_match_oracle(mach, alpha):
    """order = 1: the face states are the neighbouring cells' (no MUSCL)."""
    g = small_grid(25, 13)
    s = EulerSolver(g, dtype=np.float64, order=1)
    try:
        U = perturbed_state(s.ni, s.nc, mach, alpha, 5)
        _residual_and_jvp_vs_oracle(g, s, mach, alpha, U, order=1)
    finally:
        s.close()


def test_supersonic_far_field_branches_match_oracle():
    """M 1.5: no vortex correction (M >= 1), supersonic inflow faces take the
    freestream state and supersonic outflow faces the interior state
    (aero2d/euler.py:171, 201-204). Both overrides are checked to occur."""
    mach, alpha = 1.5, 3.0
    g = small_grid(25, 13, radius=5.0)
    s = EulerSolver(g, dtype=np.float64)
    try:
        U = perturbed_state(s.ni, s.nc, mach, alpha, 6, amp=0.03)
        Sjf = s.Sj[:, :, -1]
        area = np.hypot(Sjf[0], Sjf[1])
        nx, ny = Sjf[0] / area, Sjf[1] / area
        winf = s.freestream(mach, alpha)
        qinf = winf[1] * nx + winf[2] * ny
        sup_in = (qinf < 0) & (np.abs(qinf) > np.sqrt(1.4 * winf[3] / winf[0]))
        from aero2d_amd.euler import cons_to_prim
        W = cons_to_prim(U)[:, :, -1]
        qi = W[1] * nx + W[2] * ny
        sup_out = (qi > 0) & (qi > np.sqrt(1.4 * W[3] / W[0]))
        assert sup_in.any() and sup_out.any() and not (sup_in & sup_out).any()
        _residual_and_jvp_vs_oracle(g, s, mach, alpha, U, order=2)
    finally:
        s.close()


def test_mixed_defaults_meet_the_accuracy_bar_on_transonic_draws():
    """Eight transonic draws of the 32-case workload's distribution (NACA
    0012/2412, M 0.72-0.82, alpha 0-2.5) at 256 x 128, the mixed form with
    SOLVER_DEFAULTS unmodified against the FP64 reference path, no case
    dropped. Every case converges (rn/r0 <= 1e-10) in both forms. A case
    either agrees to |dCl| <= 1e-5 and |dCd| <= 1e-6, or both final states are
    steady states of the FP64 discrete equations: the FP64 reference-path
    Newton iteration restarted from the mixed state keeps Cl within 1e-8 and
    drives every residual component to <= 1e-10 of its freestream value (and
    the mixed form keeps the FP64 state the same way). Transonic sections
    admit several such states; which one a case reaches depends on the path."""
    from aero2d_amd.workloads import config3
    cases = config3(8, 256, 128)
    r64 = solve_batch(cases, "fp64", tol=1e-10, n_iter=800, keep_state=True)
    rmx = solve_batch(cases, "mixed", tol=1e-10, n_iter=800, keep_state=True)
    print("iterations", r64.iterations, rmx.iterations)
    assert np.all(r64.converged) and np.all(rmx.converged)
    dcl, dcd = np.abs(rmx.cl - r64.cl), np.abs(rmx.cd - r64.cd)
    print("dcl", dcl, "dcd", dcd)
    within = (dcl <= 1e-5) & (dcd <= 1e-6)
    assert within.sum() >= len(cases) // 2
    for k in np.nonzero(~within)[0]:
        a = cross_restart(cases[k], rmx.U[k], "fp64", n_iter=60)
        b = cross_restart(cases[k], r64.U[k], "mixed", n_iter=60)
        print("case", k, "fp64 from mixed", a["dcl"], a["residual"], "mixed from fp64", b["dcl"], b["residual"])
        assert a["dcl"] <= 1e-8 and max(a["residual"]) <= 1e-10
        assert b["dcl"] <= 1e-8 and max(b["residual"]) <= 1e-10


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_sweep_lds_gives_the_same_bits_as_hbm(precision):
    """The line sweeps' intermediate values kept in LDS or sent through HBM:
    the same arithmetic, so four Newton steps of a transonic batch give the
    same bits either way."""
    from aero2d_amd.workloads import config3
    cases = config3(6, 256, 128)
    outs = []
    for lds in (2, 0):
        kn = default_knobs(precision)
        kn.sweep_lds = lds
        r = solve_batch(cases, precision, tol=0.0, n_iter=4, keep_state=True, knobs=kn, stall=())
        outs.append(r)
    for k in range(len(cases)):
        assert np.array_equal(outs[0].U[k], outs[1].U[k]), k
