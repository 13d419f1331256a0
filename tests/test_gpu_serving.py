"""Serving loop and batch behaviour on the GPU.

The results of a case must not depend on which other cases share its Newton
steps: all reductions are per case and in a fixed order, and the host keeps
per-case CFL and failure state. Tolerance: 1e-12 absolute on Cl, Cd, Cm (the
kernels give bit-identical per-case arithmetic; the margin only absorbs
printing)."""
import numpy as np
import pytest

from aero2d_amd import _abi, naca4, naca4_params, ogrid
from aero2d_amd.batch import Case, default_knobs, solve_batch
from aero2d_amd.serve import SolveServer, serve_stream
from aero2d_amd.workloads import Request

pytestmark = pytest.mark.gpu


def _cases():
    g1 = ogrid(naca4("0012", 33), nj=17, ds0=1e-2, radius=10.0)            # 64 x 16
    g2 = ogrid(naca4("2412", 49), nj=25, ds0=6e-3, radius=10.0)            # 96 x 24
    g3 = ogrid(naca4_params(0.03, 0.35, 0.10, 41), nj=21, ds0=8e-3, radius=10.0)  # 80 x 20
    return [Case(g1, 0.5, 2.0, "a"), Case(g2, 0.6, 1.0, "b"), Case(g3, 0.45, -1.5, "c"),
            Case(g1, 0.55, 0.5, "d"), Case(g2, 0.4, 3.0, "e")]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_result_is_the_same_alone_and_in_the_mix(precision):
    cases = _cases()
    alone = [serve_stream([Request(0.0, c)], precision, tol=1e-9)[0][0] for c in cases[:2]]
    reqs = [Request(0.02 * k, c) for k, c in enumerate(cases)]
    mix, srv, wall = serve_stream(reqs, precision, tol=1e-9)
    by = {r.name: r for r in mix}
    assert len(mix) == len(cases) and all(r.converged for r in mix)
    for a in alone:
        m = by[a.name]
        assert a.iterations == m.iterations
        for f in ("cl", "cd", "cm"):
            assert abs(getattr(a, f) - getattr(m, f)) <= 1e-12, (a.name, f)


def test_batch_result_is_independent_of_batch_composition():
    cases = _cases()
    full = solve_batch(cases, "fp64", tol=1e-9, n_iter=300)
    sub = solve_batch([cases[3], cases[1]], "fp64", tol=1e-9, n_iter=300)
    assert np.array_equal(full.cl[[3, 1]], sub.cl) and np.array_equal(full.cd[[3, 1]], sub.cd)
    assert np.array_equal(full.iterations[[3, 1]], sub.iterations)


def test_large_batch_reduction_path_gives_the_same_bits():
    """A 512 x 256 case alone (128 update blocks) has its dot-product sums
    formed inside the CGS update; in a batch of six (768 update blocks, more
    than two per compute unit of a 256-CU device) they come from a separate
    per-case kernel. Both paths sum the same partials in the same order, so the
    states after three Newton steps must agree bit for bit."""
    from aero2d_amd.workloads import config3
    cases = config3(6, 512, 256)
    full = solve_batch(cases, "mixed", tol=1e-14, n_iter=3, keep_state=True)
    one = solve_batch([cases[4]], "mixed", tol=1e-14, n_iter=3, keep_state=True)
    assert full.iterations[4] == one.iterations[0] == 3
    assert np.array_equal(np.asarray(full.U[4]), np.asarray(one.U[0]))
    assert full.cl[4] == one.cl[0] and full.cd[4] == one.cd[0]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replays_match_eager_launches(precision):
    """From a batch's second Newton step on, the step is replayed from a
    captured HIP graph. Four steps with graphs (one eager, one capture and
    launch, two replays), each at a new CFL uploaded by the host and followed
    by state downloads and a synchronisation, give the same bits after every
    step as the same four steps launched eagerly."""
    cases = _cases()[:3]
    prec = _abi.PREC_FP64 if precision == "fp64" else _abi.PREC_MIXED
    hist = {}
    for use_graph in (1, 0):
        kn = default_knobs(precision)
        kn.use_graph = use_graph
        pool = _abi.Pool(0, kn)
        try:
            ids = np.array([pool.add_case(c.grid.x, c.grid.y, c.mach, c.alpha, prec) for c in cases], np.int32)
            out = np.empty((len(ids), 10))
            h = []
            for step in range(4):
                cfl = np.full(len(ids), 10.0 * 2.0 ** step)
                _abi.call("a2d_newton_step", pool.handle, _abi.iptr(ids), len(ids), _abi.dptr(cfl), _abi.dptr(out))
                h.append(out.copy())
                for k, c in enumerate(cases):
                    ni, nj = c.grid.x.shape
                    U = np.empty((4, ni, nj - 1))
                    _abi.call("a2d_case_get_state", pool.handle, int(ids[k]), _abi.dptr(U))
                    h.append(U)
            hist[use_graph] = h
        finally:
            pool.close()
    assert len(hist[1]) == len(hist[0]) == 4 * (1 + len(cases))
    for i, (a, b) in enumerate(zip(hist[1], hist[0])):
        assert np.array_equal(a, b, equal_nan=True), i


def test_admission_waits_for_hbm_and_keeps_arrival_order():
    g = _cases()[1].grid                                    # all requests 96 x 24
    cases = [Case(g, 0.4 + 0.05 * k, 0.5 * k, f"r{k}") for k in range(5)]
    one = _abi.case_bytes(96, 24, _abi.PREC_MIXED, 30)
    srv = SolveServer("mixed", hbm_budget=int(2.2 * one), tol=1e-8)
    try:
        futs = [srv.enqueue(c, 0.0) for c in cases]
        peak = 0
        while any(not f.done() for f in futs):
            peak = max(peak, srv.step())
            assert srv.pool.used_bytes() <= srv.budget
        assert 1 <= peak <= 2
        res = [f.result() for f in futs]
        assert all(r.converged for r in res)
        admits = [r.t_admit for r in res]
        assert admits == sorted(admits)
    finally:
        srv.close()


def test_admission_caps_cells_in_flight():
    """With ample HBM, the in-flight cell cap alone bounds the batch; requests
    wait in arrival order and each result equals its solve alone."""
    g = _cases()[1].grid                                    # all requests 96 x 24
    cases = [Case(g, 0.4 + 0.05 * k, 0.5 * k, f"r{k}") for k in range(5)]
    srv = SolveServer("mixed", tol=1e-8, max_cells=int(2.5 * 96 * 24))
    try:
        futs = [srv.enqueue(c, 0.0) for c in cases]
        peak = 0
        while any(not f.done() for f in futs):
            peak = max(peak, srv.step())
            assert srv.active_cells <= srv.max_cells
        assert peak == 2
        res = [f.result() for f in futs]
        assert all(r.converged for r in res)
        admits = [r.t_admit for r in res]
        assert admits == sorted(admits)
    finally:
        srv.close()
    alone = solve_batch([cases[4]], "mixed", tol=1e-8, n_iter=800)
    assert abs(res[4].cl - alone.cl[0]) <= 1e-12 and res[4].iterations == alone.iterations[0]


def test_request_larger_than_budget_is_refused():
    c = _cases()[1]
    srv = SolveServer("fp64", hbm_budget=1 << 20)
    try:
        f = srv.enqueue(c)
        with pytest.raises(_abi.BudgetExceeded):
            f.result(timeout=1)
    finally:
        srv.close()


def test_threaded_submit():
    cases = _cases()[:3]
    srv = SolveServer("mixed", tol=1e-8).start()
    try:
        futs = [srv.submit(c) for c in cases]
        res = [f.result(timeout=120) for f in futs]
        assert all(r.converged for r in res)
    finally:
        srv.close()


def test_bad_request_fails_on_its_own_future():
    """A request the server can never solve (an odd cell count around the
    contour under the mixed form's red-black line ordering) fails on its own
    future; a valid request submitted beside it completes, and the background
    loop keeps serving."""
    good = ogrid(naca4("0012", 33), nj=17, ds0=1e-2, radius=10.0)          # 64 x 16
    x, y = good.x[:-1].copy(), good.y[:-1].copy()                          # 63 cells around
    from aero2d_amd.grid import OGrid
    odd = OGrid(x, y)
    srv = SolveServer("mixed", hbm_budget=8 << 30, tol=1e-9).start()
    try:
        f_bad = srv.submit(Case(odd, 0.5, 2.0, "odd"))
        f_good = srv.submit(Case(good, 0.5, 2.0, "good"))
        with pytest.raises(ValueError):
            f_bad.result(timeout=60)
        r = f_good.result(timeout=120)
        assert r.converged
        f_again = srv.submit(Case(good, 0.55, 1.0, "again"))
        assert f_again.result(timeout=120).converged
        assert srv.error is None
    finally:
        srv.close()


def test_server_failure_fails_every_pending_request():
    """When the background loop stops on an unexpected error, every request
    waiting, in flight or submitted later fails instead of hanging."""
    g = ogrid(naca4("0012", 33), nj=17, ds0=1e-2, radius=10.0)
    srv = SolveServer("mixed", hbm_budget=8 << 30, tol=1e-9)

    def broken_step():
        srv._admit()
        raise RuntimeError("forced from the host")
    srv.step = broken_step
    srv.start()
    try:
        futs = [srv.submit(Case(g, 0.5, 2.0, f"r{k}")) for k in range(3)]
        for f in futs:
            with pytest.raises(RuntimeError):
                f.result(timeout=60)
        late = srv.submit(Case(g, 0.5, 2.0, "late"))
        with pytest.raises(RuntimeError):
            late.result(timeout=10)
        assert isinstance(srv.error, RuntimeError)
    finally:
        srv.close()


def test_parked_case_keeps_the_batch_and_its_graph():
    """A case parked in a batch takes no step: its state keeps its bits while
    the others advance, and the others' results equal a batch without it. The
    batch keeps its composition, so from its second step on every step
    replays the captured graph (a2d_pool_stats)."""
    cases = _cases()[:3]
    pool = _abi.Pool(0, default_knobs("mixed"))
    try:
        ids = np.array([pool.add_case(c.grid.x, c.grid.y, c.mach, c.alpha, _abi.PREC_MIXED) for c in cases],
                       np.int32)
        out = np.empty((3, 10))
        cfl = np.full(3, 10.0)
        _abi.call("a2d_newton_step", pool.handle, _abi.iptr(ids), 3, _abi.dptr(cfl), _abi.dptr(out))
        pool.park(int(ids[1]))
        ni, nj = cases[1].grid.x.shape
        U0 = np.empty((4, ni, nj - 1))
        _abi.call("a2d_case_get_state", pool.handle, int(ids[1]), _abi.dptr(U0))
        for _ in range(4):
            _abi.call("a2d_newton_step", pool.handle, _abi.iptr(ids), 3, _abi.dptr(cfl), _abi.dptr(out))
        U1 = np.empty_like(U0)
        _abi.call("a2d_case_get_state", pool.handle, int(ids[1]), _abi.dptr(U1))
        assert np.array_equal(U0, U1)
        st = pool.stats()
        assert st["steps"] == 5 and st["graph_replays"] == 4 and st["plans"] == 1
        ni, nj = cases[2].grid.x.shape
        Ua = np.empty((4, ni, nj - 1))
        _abi.call("a2d_case_get_state", pool.handle, int(ids[2]), _abi.dptr(Ua))
        with pytest.raises(_abi.HipError):
            _abi.call("a2d_case_park", pool.handle, int(ids[1]), 2)
    finally:
        pool.close()
    # the same five steps of cases 0 and 2 alone, with case 1 stepping once
    pool = _abi.Pool(0, default_knobs("mixed"))
    try:
        jd = np.array([pool.add_case(c.grid.x, c.grid.y, c.mach, c.alpha, _abi.PREC_MIXED) for c in cases],
                      np.int32)
        out = np.empty((3, 10))
        _abi.call("a2d_newton_step", pool.handle, _abi.iptr(jd), 3, _abi.dptr(np.full(3, 10.0)), _abi.dptr(out))
        sub = np.ascontiguousarray(jd[[0, 2]])
        for _ in range(4):
            _abi.call("a2d_newton_step", pool.handle, _abi.iptr(sub), 2, _abi.dptr(np.full(2, 10.0)),
                      _abi.dptr(out[:2].copy()))
        Ub = np.empty_like(Ua)
        _abi.call("a2d_case_get_state", pool.handle, int(jd[2]), _abi.dptr(Ub))
    finally:
        pool.close()
    assert np.array_equal(Ua, Ub)


def test_removed_case_slabs_are_reused():
    """Cases removed from a pool leave their slabs for the next cases of their
    size: admitting, removing and re-admitting allocates no new slab, and
    a2d_forces allocates nothing per call."""
    g = _cases()[0].grid
    pool = _abi.Pool(0, default_knobs("mixed"))
    try:
        a = [pool.add_case(g.x, g.y, 0.5, 1.0, _abi.PREC_MIXED) for _ in range(3)]
        for c in a:
            pool.remove_case(c)
        st = pool.stats()
        assert st["slabs_allocated"] == 3 and st["bytes_held"] > 0
        b = [pool.add_case(g.x, g.y, 0.6, 2.0, _abi.PREC_MIXED) for _ in range(3)]
        st = pool.stats()
        assert st["slabs_allocated"] == 3 and st["slabs_reused"] == 3 and st["bytes_held"] == 0
        f3, cp, circ = np.empty(3), np.empty(g.x.shape[0]), np.empty(1)
        for c in b:
            _abi.call("a2d_forces", pool.handle, c, _abi.dptr(f3), _abi.dptr(cp), _abi.dptr(circ))
    finally:
        pool.close()


def test_budget_frees_held_slabs():
    """With an HBM budget, slabs held for reuse are released when a case of
    another size needs the room: live plus held bytes stay within the budget."""
    g1, g2 = _cases()[0].grid, _cases()[1].grid
    one = _abi.case_bytes(g2.x.shape[0], g2.x.shape[1] - 1, _abi.PREC_MIXED, 20)
    pool = _abi.Pool(int(1.5 * one), default_knobs("mixed"))
    try:
        c = pool.add_case(g2.x, g2.y, 0.5, 1.0, _abi.PREC_MIXED)
        pool.remove_case(c)
        assert pool.stats()["bytes_held"] > 0
        d = pool.add_case(g1.x, g1.y, 0.5, 1.0, _abi.PREC_MIXED)      # a smaller size: no reuse
        st = pool.stats()
        assert pool.used_bytes() + st["bytes_held"] <= int(1.5 * one)
        pool.remove_case(d)
    finally:
        pool.close()


def test_serving_parks_finished_cases_and_replays_graphs():
    """At low load (no request waiting when a case finishes) the server parks
    finished cases, so most steps replay a captured graph; results are the
    same as without parking."""
    cases = _cases()
    reqs = [Request(0.0, c) for c in cases[:3]]
    res_p, srv_p, _ = serve_stream(reqs, "mixed", tol=1e-9)
    res_n, srv_n, _ = serve_stream(reqs, "mixed", tol=1e-9, park=False)
    by = {r.name: r for r in res_n}
    for r in res_p:
        assert r.iterations == by[r.name].iterations and abs(r.cl - by[r.name].cl) <= 1e-12
    sp, sn = srv_p.final_stats, srv_n.final_stats
    print("parking", sp, "no parking", sn)
    assert sp["graph_replay_fraction"] > sn["graph_replay_fraction"]
    assert sp["graph_replay_fraction"] > 0.9
