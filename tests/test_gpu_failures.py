"""Failure handling of the batched solve (aero2d/implicit.py:213-233): a case
whose step is not finite keeps its state, retries at CFL 1 and is frozen after
five failures; the other cases of the batch are untouched (bit-identical to a
clean run). The failure is forced from the host through a NaN initial state."""
import numpy as np
import pytest

from aero2d_amd import EulerSolver, ImplicitEuler, _abi, naca4, ogrid
from aero2d_amd.batch import default_knobs

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


def test_per_case_step_limit():
    """a2d_case_set_step_limit bounds one case's step without touching the
    others: two identical cases, one limited to 0.01, take the same Newton
    direction with omega = min(1, limit / change) each."""
    g = ogrid(naca4("0012", 33), nj=17, ds0=1e-2, radius=10.0)
    pool = _abi.Pool(0, default_knobs("mixed"))
    try:
        ids = np.array([pool.add_case(g.x, g.y, 0.5, 2.0, _abi.PREC_MIXED) for _ in range(2)], np.int32)
        _abi.call("a2d_case_set_step_limit", pool.handle, int(ids[1]), 0.01)
        out = np.empty((2, 10))
        cfl = np.full(2, 50.0)
        _abi.call("a2d_newton_step", pool.handle, _abi.iptr(ids), 2, _abi.dptr(cfl), _abi.dptr(out))
        rel, om = out[:, 7], out[:, 1]
        assert rel[0] == rel[1] and rel[0] > 0.01
        assert np.isclose(om[0], min(1.0, 0.5 / (rel[0] + 1e-30)), rtol=1e-15, atol=0)
        assert np.isclose(om[1], 0.01 / (rel[1] + 1e-30), rtol=1e-15, atol=0)
        _abi.call("a2d_case_set_step_limit", pool.handle, int(ids[1]), 0.0)   # back to the knob
        with pytest.raises(_abi.HipError):
            _abi.call("a2d_case_set_step_limit", pool.handle, int(ids[1]), -1.0)
    finally:
        pool.close()


def test_positivity_check_keeps_states_physical():
    """With the positivity knob, no accepted step leaves a cell with rho <= 0
    or p <= 0, however large the steps: four transonic cases (NACA 0012/2412,
    128 x 64) with a step limit of 2 (ten times the reference's), every state
    checked after every Newton step; the check must have rejected steps."""
    from aero2d_amd.batch import BatchSolve
    from aero2d_amd.euler import cons_to_prim
    from aero2d_amd.implicit import NewtonLoop
    from aero2d_amd.workloads import config3
    cases = config3(4, 128, 64)
    bs = BatchSolve(cases, "mixed", knobs=default_knobs("mixed", max_update=2.0, positivity=1)).upload()
    pool, ids = bs.pool, bs.ids
    loop = NewtonLoop(len(cases), 10.0, 300.0, 1000.0, 1e-3)
    rejected = 0
    try:
        for it in range(60):
            act = np.nonzero(~loop.frozen)[0]
            if len(act) == 0:
                break
            out = np.empty((len(act), 10))
            cfl = np.ascontiguousarray(loop.cfl[act])
            _abi.call("a2d_newton_step", pool.handle, _abi.iptr(np.ascontiguousarray(ids[act])), len(act),
                      _abi.dptr(cfl), _abi.dptr(out))
            loop.update(act, out)
            rejected += int(np.sum(out[:, 2] < 0.5))
            for k in act:
                ni, nj = cases[k].grid.x.shape
                U = np.empty((4, ni, nj - 1))
                _abi.call("a2d_case_get_state", pool.handle, int(ids[k]), _abi.dptr(U))
                W = cons_to_prim(U)
                assert np.all(W[0] > 0) and np.all(W[3] > 0), (it, k)
    finally:
        pool.close()
    print("rejected steps", rejected)
    assert rejected >= 1


@pytest.mark.parametrize("name", ["precond_order", "positivity", "precond_half", "basis_half"])
def test_invalid_switch_knob_is_refused(name):
    """Switch knobs take 0 or 1; any other value is refused with an error and
    the pool keeps its knobs."""
    pool = _abi.Pool(0, default_knobs("mixed"))
    try:
        k = pool.knobs()
        before = getattr(k, name)
        setattr(k, name, 2)
        with pytest.raises(_abi.HipError):
            pool.set_knobs(k)
        assert getattr(pool.knobs(), name) == before
    finally:
        pool.close()


def test_oversized_case_is_refused_before_any_device_work():
    """Kernels address inside a case with 32-bit byte offsets; the largest
    array is the block storage, 80 * cells values of the inner type. A case
    whose block storage reaches 4 GiB is refused with -6 by a2d_case_add (and
    a2d_case_bytes reports -6 for it); nothing is allocated."""
    pool = _abi.Pool(0, default_knobs("fp64"))
    try:
        ni, nc = 4096, 1640            # 6.72 M cells: 80 * 8 B * n = 4.30 GB >= 4 GiB in FP64
        assert 80 * 8 * ni * nc >= 1 << 32
        assert _abi.load().a2d_case_bytes(ni, nc, _abi.PREC_FP64, 50) == -6
        assert _abi.load().a2d_case_bytes(ni, nc, _abi.PREC_MIXED, 20) > 0      # half the bytes in FP32
        x = np.zeros((ni, nc + 1))
        y = np.zeros((ni, nc + 1))
        with pytest.raises(_abi.HipError) as e:
            pool.add_case(x, y, 0.5, 2.0, _abi.PREC_FP64)
        assert e.value.code == -6 and "4 GiB" in str(e.value)
        assert pool.used_bytes() == 0
    finally:
        pool.close()


def test_repeated_case_id_is_refused():
    """A batch that names a case twice would run that case's in-place kernels
    twice per launch: a2d_newton_step and a2d_time_kernel refuse it with -6
    and leave the state untouched."""
    g = ogrid(naca4("0012", 17), nj=9, ds0=2e-2, radius=5.0)
    pool = _abi.Pool(0, default_knobs("fp64"))
    try:
        a = pool.add_case(g.x, g.y, 0.5, 2.0, _abi.PREC_FP64)
        b = pool.add_case(g.x, g.y, 0.6, 1.0, _abi.PREC_FP64)
        U0 = np.empty((4, 32, 8))
        _abi.call("a2d_case_get_state", pool.handle, a, _abi.dptr(U0))
        ids = np.array([a, b, a], np.int32)
        out = np.empty((3, 10))
        cfl = np.full(3, 10.0)
        with pytest.raises(_abi.HipError) as e:
            _abi.call("a2d_newton_step", pool.handle, _abi.iptr(ids), 3, _abi.dptr(cfl), _abi.dptr(out))
        assert e.value.code == -6
        ms = np.zeros(1)
        with pytest.raises(_abi.HipError) as e:
            _abi.call("a2d_time_kernel", pool.handle, _abi.iptr(ids), 3, 0, 1, _abi.dptr(ms))
        assert e.value.code == -6
        U1 = np.empty((4, 32, 8))
        _abi.call("a2d_case_get_state", pool.handle, a, _abi.dptr(U1))
        assert np.array_equal(U0, U1)
    finally:
        pool.close()
