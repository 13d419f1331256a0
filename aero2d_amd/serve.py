"""Serving loop: continuous batching of single-case solves on one GPU.

Requests (an O-grid and a freestream) arrive at any time. A request is
admitted, in arrival order, when the device pool's HBM budget has room for
its footprint (a2d_case_bytes: state, metrics, Krylov basis and
preconditioner blocks) and the in-flight cell count stays under the cap that
fills the GPU; its state then stays on the device until it finishes. Every Newton step runs
over all admitted cases at once (one a2d_newton_step call), so requests that
arrive while others are in flight join the next step and their work overlaps
on the device; a case leaves as soon as it converges (rn/r0 <= tol), freezes
or reaches its step limit, and its memory admits the next request.

Each case's result depends only on that case: kernels reduce per case in a
fixed order and the host keeps per-case CFL/failure state, so a request gives
the same answer alone and in any mix (tests/test_gpu_serving.py).
"""
from __future__ import annotations

import queue
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass

import numpy as np

from . import _abi
from .batch import SOLVER_DEFAULTS, Case, cfl_policy, default_knobs
from .implicit import NewtonLoop, StallRule

__all__ = ["ServeResult", "SolveServer", "serve_stream", "censored_latency", "MAX_CELLS"]

# In-flight cell cap per form: the cell count at which one Newton step over
# the in-flight set takes about 250 ms on one MI355X (measured cost once the
# GPU is full: mixed ~8 ns, FP64 ~50 ns per cell and Newton step).
MAX_CELLS = {"mixed": 32_000_000, "fp64": 5_000_000}


@dataclass
class ServeResult:
    name: str
    cl: float
    cd: float
    cm: float
    converged: bool
    frozen: bool
    stalled: bool
    iterations: int
    rn_over_r0: float
    t_arrive: float
    t_admit: float
    t_done: float

    @property
    def latency(self) -> float:
        return self.t_done - self.t_arrive


class _Slot:
    __slots__ = ("case", "t_arrive", "t_admit", "future", "cid")

    def __init__(self, case, t_arrive, t_admit, future, cid):
        self.case, self.t_arrive, self.t_admit, self.future, self.cid = case, t_arrive, t_admit, future, cid


class SolveServer:
    """Continuous-batching solver over one device pool.

    precision "mixed" (default) or "fp64"; hbm_budget in bytes bounds the
    admitted cases' total footprint; max_batch bounds their count; max_cells
    bounds their total cell count, i.e. the time of one Newton step over the
    in-flight set (default MAX_CELLS[precision], about 250 ms per step on one
    MI355X: ~8 ns per cell and Newton step for the mixed form, ~50 ns for
    FP64, once the batch fills the GPU). Past that bound, admitting more only
    makes every admitted case progress more slowly, so later requests wait
    in arrival order instead. tol and n_iter are per request. Use submit()
    from any thread with start(), or drive it synchronously with step()
    (serve_stream does)."""

    def __init__(self, precision: str = "mixed", hbm_budget: int = 64 << 30, max_batch: int = 1024,
                 tol: float = 1e-10, n_iter: int = 800, cfl0: float = 10.0,
                 restart: int | None = None, early_exit: int | None = None,
                 cfl_max: float | None = None, max_cells: int | None = None,
                 stall: tuple | None = None, park: bool = True):
        if precision not in SOLVER_DEFAULTS:
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = precision
        self.prec = _abi.PREC_MIXED if precision == "mixed" else _abi.PREC_FP64
        kn = default_knobs(precision, restart=restart, early_exit=early_exit)
        self.restart = kn.gmres_restart
        self.order_rb = kn.precond_order == 1
        self.pool = _abi.Pool(hbm_budget, kn)
        self.budget = hbm_budget
        self.max_batch = max_batch
        self.max_cells = MAX_CELLS[precision] if max_cells is None else int(max_cells)
        self.active_cells = 0
        self.tol, self.n_iter = tol, n_iter
        cmax, clate, below = cfl_policy(precision, cfl_max)
        self.fallback_max_update = SOLVER_DEFAULTS[precision]["fallback_max_update"]
        fb = self.fallback_max_update > 0 and self.fallback_max_update != kn.max_update
        self.loop = NewtonLoop(4096, cfl0, cmax, clate, below, fb)
        stall = SOLVER_DEFAULTS[precision]["stall"] if stall is None else stall
        self.stall = StallRule(4096, *stall) if stall else None
        self.park = park
        self.parked: set[int] = set()      # finished cases kept in the batch, doing no work
        self.active: dict[int, _Slot] = {}
        self.waiting: list[tuple[Case, float, Future]] = []
        self.t0 = time.perf_counter()
        self._inbox: queue.Queue = queue.Queue()
        self._thread = None
        self._stop = threading.Event()
        self.error: BaseException | None = None     # set when the background loop stopped on one
        self.steps = 0

    # ---- clock -------------------------------------------------------------------
    def now(self) -> float:
        return time.perf_counter() - self.t0

    # ---- admission -------------------------------------------------------------------
    def footprint(self, case: Case) -> int:
        ni, nj = case.grid.x.shape
        return _abi.case_bytes(ni, nj - 1, self.prec, self.restart)

    def check(self, case: Case):
        """Raise when the request can never be admitted: a size the library
        refuses, a footprint above the whole HBM budget, or an odd cell count
        around the contour under the red-black line ordering (the mixed form's
        default), which pairs lines of two colours."""
        ni, nj = case.grid.x.shape
        if self.order_rb and ni % 2:
            raise ValueError(f"{ni} cells around: the red-black line ordering of the {self.precision} "
                             "form needs an even count")
        if self.footprint(case) > self.budget:
            raise _abi.BudgetExceeded("a2d_case_add", -2, "request larger than the HBM budget")

    def enqueue(self, case: Case, t_arrive: float | None = None, future: Future | None = None) -> Future:
        """Queue a request (not thread-safe; use submit() across threads). A
        request that can never be admitted fails at once on its own future."""
        f = future if future is not None else Future()
        try:
            self.check(case)
        except Exception as e:          # noqa: BLE001 - reported on the request's future
            f.set_exception(e)
            return f
        self.waiting.append((case, self.now() if t_arrive is None else t_arrive, f))
        return f

    @staticmethod
    def cells(case: Case) -> int:
        ni, nj = case.grid.x.shape
        return ni * (nj - 1)

    def _admit(self):
        while self.waiting and len(self.active) < self.max_batch:
            case, t_arr, fut = self.waiting[0]
            if self.pool.used_bytes() + self.footprint(case) > self.budget:
                break                                   # wait for memory, keep arrival order
            if self.active and self.active_cells + self.cells(case) > self.max_cells:
                break                                   # the GPU is full: wait, keep arrival order
            self.waiting.pop(0)
            try:
                cid = self.pool.add_case(case.grid.x, case.grid.y, case.mach, case.alpha, self.prec)
            except _abi.HipError as e:
                fut.set_exception(e)                    # this request fails; admission goes on
                continue
            self.loop.reset([cid])
            if self.stall is not None:
                self.stall.reset([cid])
            self.active[cid] = _Slot(case, t_arr, self.now(), fut, cid)
            self.active_cells += self.cells(case)

    def _finish(self, cid: int, stalled: bool = False):
        """Report a finished case; it leaves the batch at once when a request
        is waiting for its room, else it is parked (kept in the batch doing no
        work), so that the batch's composition, and its captured graph, stay."""
        s = self.active.pop(cid)
        self.active_cells -= self.cells(s.case)
        f3 = np.empty(3)
        cp = np.empty(s.case.grid.x.shape[0])
        circ = np.empty(1)
        _abi.call("a2d_forces", self.pool.handle, cid, _abi.dptr(f3), _abi.dptr(cp), _abi.dptr(circ))
        if self.park and not self.waiting:
            self.pool.park(cid)
            self.parked.add(cid)
        else:
            self.pool.remove_case(cid)
        rel = self.loop.last_rn[cid] / self.loop.r0[cid]
        res = ServeResult(s.case.name, float(f3[0]), float(f3[1]), float(f3[2]),
                          bool(rel <= self.tol and not self.loop.frozen[cid]), bool(self.loop.frozen[cid]),
                          bool(stalled), int(self.loop.iters[cid]), float(rel), s.t_arrive, s.t_admit,
                          self.now())
        s.future.set_result(res)

    def _unpark_all(self):
        for cid in self.parked:
            self.pool.remove_case(cid)
        self.parked.clear()

    # ---- one Newton step over everything admitted --------------------------------------------
    def step(self) -> int:
        """Admit what fits, take one Newton step over all active cases and
        retire the finished ones. Returns the number of active cases stepped.
        Parked cases leave the batch when a request waits (their memory is
        needed) or when no active case is left."""
        if self.parked and (self.waiting or not self.active):
            self._unpark_all()
        self._admit()
        if not self.active:
            return 0
        ids = np.array(sorted(self.active), np.int32)
        batch = np.array(sorted(set(self.active) | self.parked), np.int32)
        cfl = np.ascontiguousarray(self.loop.cfl[batch])
        out = np.empty((len(batch), 10))
        _abi.call("a2d_newton_step", self.pool.handle, _abi.iptr(batch), len(batch), _abi.dptr(cfl),
                  _abi.dptr(out))
        out = out[np.isin(batch, ids)]                  # parked cases' rows are stale
        rn, _ = self.loop.update(ids, out)
        for cid in self.loop.newly_failed:
            _abi.call("a2d_case_set_step_limit", self.pool.handle, int(cid), float(self.fallback_max_update))
        self.steps += 1
        rel = rn / self.loop.r0[ids]
        done = (rel <= self.tol) | self.loop.frozen[ids] | (self.loop.iters[ids] >= self.n_iter)
        stalled = np.zeros(len(ids), bool)
        if self.stall is not None:
            stalled = self.stall.update(ids, rel) & ~done
        for k in np.nonzero(done | stalled)[0]:
            self._finish(int(ids[k]), bool(stalled[k]))
        return len(ids)

    def stats(self) -> dict:
        """The pool's counters (a2d_pool_stats): steps, graph replays, plans built, slabs reused."""
        return self.pool.stats()

    # ---- threaded API ----------------------------------------------------------------------------
    def submit(self, case: Case) -> Future:
        """Thread-safe: queue a request for the background loop (start() first).
        After the loop has stopped on an error, the request fails at once."""
        f = Future()
        self._inbox.put((case, self.now(), f))
        if self.error is not None:      # the loop has stopped: nothing will take it
            self._drain_inbox(RuntimeError(f"solve server stopped: {self.error!r}"))
        return f

    def _drain_inbox(self, exc: BaseException):
        while True:
            try:
                f = self._inbox.get_nowait()[2]
            except queue.Empty:
                break
            if not f.done():
                f.set_exception(exc)

    def fail_all(self, exc: BaseException):
        """Fail every request still waiting, in flight or in the inbox."""
        pending = [f for _, _, f in self.waiting] + [s.future for s in self.active.values()]
        self.waiting.clear()
        self.active.clear()
        self.parked.clear()
        self.active_cells = 0
        for f in pending:
            if not f.done():
                f.set_exception(exc)
        self._drain_inbox(exc)

    def start(self):
        def run():
            try:
                while not self._stop.is_set():
                    while True:
                        try:
                            case, t, f = self._inbox.get_nowait()
                        except queue.Empty:
                            break
                        self.enqueue(case, t, f)
                    if self.step() == 0:
                        time.sleep(0.001)
            except BaseException as e:          # noqa: BLE001 - no request may hang
                self.error = e
                self.fail_all(RuntimeError(f"solve server stopped: {e!r}"))
        self._thread = threading.Thread(target=run, daemon=True)
        self._thread.start()
        return self

    def close(self):
        """Stop the background loop and release the pool; requests not
        finished by then fail with a RuntimeError ("closed")."""
        if self._thread is not None:
            self._stop.set()
            self._thread.join()
            self._thread = None
        self.fail_all(RuntimeError("solve server closed before the request finished"))
        self.pool.close()


def serve_stream(requests, precision: str = "mixed", hbm_budget: int = 64 << 30, tol: float = 1e-10,
                 horizon: float | None = None, **kw):
    """Replay an open-loop request stream (objects with .t_arrive and .case, t
    in seconds from the start) against a SolveServer in real time. Requests are
    released when the wall clock reaches their arrival time; the run ends when
    all released requests finish or, with `horizon`, at that many seconds.
    Returns (results, server, wall): the finished requests' results, and the
    server, whose `unfinished` lists the arrival times of the requests released
    but not finished by the end (waiting or in flight; their latency is at
    least wall - t_arrive, see censored_latency) and `final_stats` its pool's
    counters at the end."""
    srv = SolveServer(precision, hbm_budget, tol=tol, **kw)
    futures = []
    k = 0
    reqs = sorted(requests, key=lambda r: r.t_arrive)
    try:
        srv.t0 = time.perf_counter()
        while True:
            now = srv.now()
            while k < len(reqs) and reqs[k].t_arrive <= now:
                futures.append(srv.enqueue(reqs[k].case, reqs[k].t_arrive))
                k += 1
            if horizon is not None and now >= horizon:
                break
            if srv.step() == 0:
                if k >= len(reqs) and not srv.waiting:
                    break
                if k < len(reqs):
                    time.sleep(max(0.0, min(0.01, reqs[k].t_arrive - srv.now())))
        wall = srv.now()
        srv.final_stats = srv.stats()
    finally:
        srv.close()
    results = [f.result() for f in futures if f.done() and f.exception() is None]
    srv.unfinished = [reqs[i].t_arrive for i, f in enumerate(futures)
                      if not (f.done() and f.exception() is None)]
    return results, srv, wall


def censored_latency(results, unfinished, wall, q=(50, 99)) -> dict:
    """Latency percentiles over EVERY released request: a finished request
    counts its latency, an unfinished one wall - t_arrive (a lower bound: it is
    censored at the end of the run). Also the completed fraction. Raising
    censored values can only raise a percentile, so each percentile is a lower
    bound; it is exact unless a censored value is among the order statistics
    it interpolates (then it is listed in `censored_at`)."""
    lat = [r.latency for r in results] + [wall - t for t in unfinished]
    flag = [False] * len(results) + [True] * len(unfinished)
    out = {"released": len(lat), "completed": len(results),
           "completed_fraction": len(results) / len(lat) if lat else float("nan")}
    if not lat:
        return out
    order = np.argsort(lat, kind="stable")
    lat_s = np.asarray(lat)[order]
    flag_s = np.asarray(flag)[order]
    cens = []
    for p in q:
        out[f"p{p}_ms"] = float(np.percentile(lat_s, p)) * 1e3
        k_hi = int(np.ceil(p / 100 * (len(lat_s) - 1)))     # numpy's linear rule reads ranks <= k_hi
        if flag_s[:k_hi + 1].any():
            cens.append(p)
    out["censored_at"] = cens
    return out
