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
from .implicit import NewtonLoop

__all__ = ["ServeResult", "SolveServer", "serve_stream", "MAX_CELLS"]

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
                 cfl_max: float | None = None, max_cells: int | None = None):
        if precision not in SOLVER_DEFAULTS:
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = precision
        self.prec = _abi.PREC_MIXED if precision == "mixed" else _abi.PREC_FP64
        kn = default_knobs(precision, restart=restart, early_exit=early_exit)
        self.restart = kn.gmres_restart
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
        self.active: dict[int, _Slot] = {}
        self.waiting: list[tuple[Case, float, Future]] = []
        self.t0 = time.perf_counter()
        self._inbox: queue.Queue = queue.Queue()
        self._thread = None
        self._stop = threading.Event()
        self.steps = 0

    # ---- clock -------------------------------------------------------------------
    def now(self) -> float:
        return time.perf_counter() - self.t0

    # ---- admission -------------------------------------------------------------------
    def footprint(self, case: Case) -> int:
        ni, nj = case.grid.x.shape
        return _abi.case_bytes(ni, nj - 1, self.prec, self.restart)

    def enqueue(self, case: Case, t_arrive: float | None = None, future: Future | None = None) -> Future:
        """Queue a request (not thread-safe; use submit() across threads)."""
        f = future if future is not None else Future()
        if self.footprint(case) > self.budget:
            f.set_exception(_abi.BudgetExceeded("a2d_case_add", -2, "request larger than the HBM budget"))
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
            cid = self.pool.add_case(case.grid.x, case.grid.y, case.mach, case.alpha, self.prec)
            self.loop.reset([cid])
            self.active[cid] = _Slot(case, t_arr, self.now(), fut, cid)
            self.active_cells += self.cells(case)

    def _finish(self, cid: int):
        s = self.active.pop(cid)
        self.active_cells -= self.cells(s.case)
        f3 = np.empty(3)
        cp = np.empty(s.case.grid.x.shape[0])
        circ = np.empty(1)
        _abi.call("a2d_forces", self.pool.handle, cid, _abi.dptr(f3), _abi.dptr(cp), _abi.dptr(circ))
        self.pool.remove_case(cid)
        rel = self.loop.last_rn[cid] / self.loop.r0[cid]
        res = ServeResult(s.case.name, float(f3[0]), float(f3[1]), float(f3[2]),
                          bool(rel <= self.tol and not self.loop.frozen[cid]), bool(self.loop.frozen[cid]),
                          int(self.loop.iters[cid]), float(rel), s.t_arrive, s.t_admit, self.now())
        s.future.set_result(res)

    # ---- one Newton step over everything admitted --------------------------------------------
    def step(self) -> int:
        """Admit what fits, take one Newton step over all active cases and
        retire the finished ones. Returns the number of active cases stepped."""
        self._admit()
        if not self.active:
            return 0
        ids = np.array(sorted(self.active), np.int32)
        cfl = np.ascontiguousarray(self.loop.cfl[ids])
        out = np.empty((len(ids), 10))
        _abi.call("a2d_newton_step", self.pool.handle, _abi.iptr(ids), len(ids), _abi.dptr(cfl),
                  _abi.dptr(out))
        rn, _ = self.loop.update(ids, out)
        for cid in self.loop.newly_failed:
            _abi.call("a2d_case_set_step_limit", self.pool.handle, int(cid), float(self.fallback_max_update))
        self.steps += 1
        done = (rn / self.loop.r0[ids] <= self.tol) | self.loop.frozen[ids] | (self.loop.iters[ids] >= self.n_iter)
        for cid in ids[done]:
            self._finish(int(cid))
        return len(ids)

    # ---- threaded API ----------------------------------------------------------------------------
    def submit(self, case: Case) -> Future:
        """Thread-safe: queue a request for the background loop (start() first)."""
        f = Future()
        self._inbox.put((case, self.now(), f))
        return f

    def start(self):
        def run():
            while not self._stop.is_set():
                while True:
                    try:
                        case, t, f = self._inbox.get_nowait()
                    except queue.Empty:
                        break
                    self.enqueue(case, t, f)
                if self.step() == 0:
                    time.sleep(0.001)
        self._thread = threading.Thread(target=run, daemon=True)
        self._thread.start()
        return self

    def close(self):
        if self._thread is not None:
            self._stop.set()
            self._thread.join()
            self._thread = None
        self.pool.close()


def serve_stream(requests, precision: str = "mixed", hbm_budget: int = 64 << 30, tol: float = 1e-10,
                 horizon: float | None = None, **kw):
    """Replay an open-loop request stream (objects with .t_arrive and .case, t
    in seconds from the start) against a SolveServer in real time. Requests are
    released when the wall clock reaches their arrival time; the run ends when
    all released requests finish or, with `horizon`, at that many seconds
    (unfinished requests are reported). Returns (results, server, wall)."""
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
    finally:
        srv.close()
    results = [f.result() for f in futures if f.done() and f.exception() is None]
    return results, srv, wall
