"""Batched solves over cases with their own geometry and grid size.

The reference batches (M, alpha) over ONE grid (jax.vmap in
aero2d/implicit.py:204-210; one EulerSolver holds one grid,
aero2d/euler.py:226-237). Here every case carries its own O-grid, so a batch
may mix airfoils and sizes: all cases live in one device pool and each
Newton step is one library call over the active ones (kernels map their
blocks to cases through per-batch prefix tables). A case leaves the batch as
soon as it converges or freezes; results never depend on the other cases in
the batch (all reductions are case-local).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from .grid import OGrid
from .implicit import NewtonLoop

__all__ = ["Case", "BatchResult", "solve_batch", "default_knobs"]


@dataclass
class Case:
    """One solve request: an O-grid and a freestream."""
    grid: OGrid
    mach: float
    alpha: float
    name: str = ""


@dataclass
class BatchResult:
    cl: np.ndarray
    cd: np.ndarray
    cm: np.ndarray
    converged: np.ndarray        # rn/r0 <= tol and not frozen
    frozen: np.ndarray
    iterations: np.ndarray       # Newton steps taken per case
    rn_over_r0: np.ndarray       # final relative density residual
    rn: np.ndarray               # final absolute residual
    t_done: np.ndarray           # seconds from the start of the solve to each case's end
    wall: float                  # seconds for the whole batch
    arnoldi: np.ndarray          # mean Arnoldi steps per Newton step, per case
    U: list = field(default_factory=list)


def default_knobs(precision: str, restart: int = 50, early_exit: bool | None = None) -> _abi.Knobs:
    k = _abi.default_knobs()
    k.gmres_restart = restart
    k.early_exit = int(precision == "mixed") if early_exit is None else int(early_exit)
    return k


def solve_batch(cases: list[Case], precision: str = "fp64", tol: float = 1e-10, n_iter: int = 200,
                cfl0: float = 10.0, cfl_max: float = 300.0, knobs: _abi.Knobs | None = None,
                shapes: dict | None = None, keep_state: bool = False, pool_budget: int = 0,
                on_step=None) -> BatchResult:
    """Solve every case to rn/r0 <= tol (or n_iter Newton steps).

    precision: "fp64" or "mixed". shapes: optional {(ni, nc): stencil shape
    index} from the autotuner. on_step(it, n_active) is called after each step."""
    prec = _abi.PREC_FP64 if precision == "fp64" else _abi.PREC_MIXED
    kn = knobs if knobs is not None else default_knobs(precision)
    t0 = time.perf_counter()
    pool = _abi.Pool(pool_budget, kn)
    nb = len(cases)
    try:
        ids = np.empty(nb, np.int32)
        for k, c in enumerate(cases):
            ids[k] = pool.add_case(c.grid.x, c.grid.y, c.mach, c.alpha, prec)
            if shapes:
                sh = shapes.get(tuple(c.grid.x.shape[:1]) + (c.grid.x.shape[1] - 1,))
                if sh is not None:
                    _abi.call("a2d_case_set_shape", pool.handle, int(ids[k]), int(sh))
        loop = NewtonLoop(nb, cfl0, cfl_max)
        done = np.zeros(nb, bool)
        t_done = np.full(nb, np.nan)
        arn = np.zeros(nb)
        for it in range(n_iter):
            act = np.nonzero(~done)[0]
            if len(act) == 0:
                break
            sub = np.ascontiguousarray(ids[act])
            cfl = np.ascontiguousarray(loop.cfl[act])
            out = np.empty((len(act), 10))
            _abi.call("a2d_newton_step", pool.handle, _abi.iptr(sub), len(act), _abi.dptr(cfl),
                      _abi.dptr(out))
            rn, _ = loop.update(act, out)
            arn[act] += out[:, 6]
            fin = (rn / loop.r0[act] <= tol) | loop.frozen[act]
            now = time.perf_counter() - t0
            t_done[act[fin]] = now
            done[act[fin]] = True
            if on_step is not None:
                on_step(it, len(act))
        t_done[np.isnan(t_done)] = time.perf_counter() - t0
        cl = np.empty(nb)
        cd = np.empty(nb)
        cm = np.empty(nb)
        states = []
        for k in range(nb):
            f3 = np.empty(3)
            c = cases[k]
            cp = np.empty(c.grid.x.shape[0])
            circ = np.empty(1)
            _abi.call("a2d_forces", pool.handle, int(ids[k]), _abi.dptr(f3), _abi.dptr(cp), _abi.dptr(circ))
            cl[k], cd[k], cm[k] = f3
            if keep_state:
                ni, nj = c.grid.x.shape
                Uk = np.empty((4, ni, nj - 1))
                _abi.call("a2d_case_get_state", pool.handle, int(ids[k]), _abi.dptr(Uk))
                states.append(Uk)
    finally:
        pool.close()
    wall = time.perf_counter() - t0
    rel = loop.last_rn / loop.r0
    conv = (~loop.frozen) & (rel <= tol)
    return BatchResult(cl, cd, cm, conv, loop.frozen.copy(), loop.iters.copy(), rel, loop.last_rn.copy(),
                       t_done, wall, arn / np.maximum(loop.iters, 1), states)
