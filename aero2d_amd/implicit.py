"""Pseudo-transient Newton-Krylov solve — the API of aero2d/implicit.py.

Each Newton step solves (vol/dtau + dR/dU) dU = -R(U) with GMRES, left
preconditioned by the first-order block Jacobian inverted by line Jacobi
(block-tridiagonal solves along the wall-normal j lines, i-coupling by
sweeps). The whole step — residual, exact Jacobian-vector products, block
assembly and factorisation, Arnoldi, least squares, limiter and update — is
one call into the HIP library for the batch (a2d_newton_step), replayed from a
captured HIP graph after its first use. The host keeps the reference's
per-case control: SER CFL, rejection of non-finite steps, freezing after five
failures, the convergence test and the callback (aero2d/implicit.py:194-253).
"""
from __future__ import annotations

import numpy as np

from . import _abi
from .euler import EulerResult, EulerSolver

__all__ = ["ImplicitEuler", "NewtonLoop", "StallRule"]


class StallRule:
    """Retire cases whose convergence has stalled (limit cycles): a case that
    has taken `after` Newton steps and whose best rn/r0 over its last `window`
    steps is not `drop` times below its best before them. Per-case state lives
    in slots (reset() for a new case in a slot), so a batch and a serving loop
    apply the same rule case by case."""

    def __init__(self, nb, after, window, drop):
        self.after, self.window, self.drop = int(after), int(window), float(drop)
        self.ring = np.full((nb, self.window + 1), np.inf)   # best rn/r0 after each step
        self.count = np.zeros(nb, int)

    def reset(self, idx):
        self.ring[idx] = np.inf
        self.count[idx] = 0

    def update(self, idx, rel):
        """Record one step's rn/r0 (NaN for a rejected step) of the cases idx;
        returns which of them have stalled."""
        idx = np.asarray(idx)
        c = self.count[idx]
        prev = self.ring[idx, (c - 1) % (self.window + 1)]
        best = np.where(c > 0, np.fmin(prev, rel), np.fmin(np.inf, rel))
        best = np.where(np.isnan(best), np.inf, best)
        self.ring[idx, c % (self.window + 1)] = best
        self.count[idx] = c + 1
        old = self.ring[idx, (c + 1) % (self.window + 1)]   # best `window` steps ago
        return (c + 1 > self.window) & (c + 1 >= self.after) & ~(best * self.drop <= old)


class NewtonLoop:
    """Host bookkeeping of the reference loop for a set of cases
    (aero2d/implicit.py:211-244), one object per solve. `step_out` rows come
    from a2d_newton_step."""

    def __init__(self, nb, cfl0, cfl_max, cfl_late=None, late_below=0.0, fallback=False):
        """cfl_late / late_below: once a case's rn/r0 falls below late_below its
        CFL cap rises from cfl_max to cfl_late (None keeps the reference's
        single cap). fallback: record each case's first rejected step in
        `newly_failed` (the caller then lowers that case's step limit)."""
        self.cfl0, self.cfl_max = cfl0, cfl_max
        self.cfl_late, self.late_below = cfl_late, late_below
        self.fallback = fallback
        self.failed_once = np.zeros(nb, bool)
        self.newly_failed = np.zeros(0, int)
        self.late = np.zeros(nb, bool)          # rn/r0 has fallen below late_below
        self.newly_late = np.zeros(0, int)
        self.cfl = np.full(nb, float(cfl0))
        self.r0 = np.full(nb, np.nan)
        self.fails = np.zeros(nb, int)
        self.frozen = np.zeros(nb, bool)
        self.iters = np.zeros(nb, int)
        self.last_rn = np.full(nb, np.nan)

    def reset(self, idx):
        """Fresh per-case state for entries idx (a new case in a reused slot)."""
        self.cfl[idx] = float(self.cfl0)
        self.r0[idx] = np.nan
        self.fails[idx] = 0
        self.frozen[idx] = False
        self.iters[idx] = 0
        self.last_rn[idx] = np.nan
        self.failed_once[idx] = False
        self.late[idx] = False

    def update(self, idx, out):
        """Apply one step's results for batch entries idx; returns rn (NaN when
        the step was rejected) and cl."""
        rn, om, acc, cl = out[:, 0], out[:, 1], out[:, 2] > 0.5, out[:, 3]
        finite = np.isfinite(rn) & acc
        self.fails[idx] += ~finite
        first = ~finite & ~self.failed_once[idx]
        self.newly_failed = np.asarray(idx)[first] if self.fallback else np.zeros(0, int)
        self.failed_once[idx] |= ~finite
        self.frozen[idx] |= self.fails[idx] >= 5
        rn = np.where(finite, rn, np.nan)
        first = np.isnan(self.r0[idx])
        r0 = np.where(np.isfinite(rn), np.maximum(rn, 1e-30), 1.0)
        self.r0[idx] = np.where(first, r0, self.r0[idx])
        r0 = self.r0[idx]
        ser = self.cfl0 * (r0 / np.where(np.isfinite(rn), np.maximum(rn, 1e-30), r0))
        cap = np.full(len(idx), float(self.cfl_max))
        below = np.isfinite(rn) & (rn / r0 < self.late_below)
        if self.cfl_late is not None:
            cap = np.where(below, float(self.cfl_late), cap)
        cross = below & ~self.late[idx]
        self.newly_late = np.asarray(idx)[cross]
        self.late[idx] |= below
        cfl = np.clip(ser, self.cfl0, cap)
        cfl = np.where(om < 1.0, np.maximum(cfl * 0.5, 1.0), cfl)
        cfl = np.where(~finite, 1.0, cfl)
        self.cfl[idx] = cfl
        self.iters[idx] += 1
        self.last_rn[idx] = np.where(np.isfinite(rn), rn, self.last_rn[idx])
        return rn, cl


class ImplicitEuler:
    """aero2d/implicit.py:67-78. Extra knobs: early_exit (stop a case's Arnoldi
    loop once its preconditioned residual estimate meets gmres_tol; the
    default None means on for the mixed form, off — the reference's fixed
    `gmres_restart` steps — for FP64), per_case_stop (a converged case leaves
    the batch at once instead of iterating until all are done), use_graph."""

    def __init__(self, base: EulerSolver, cfl0: float = 10.0, cfl_max: float = 300.0,
                 gmres_restart: int = 50, gmres_tol: float = 0.01, jacobi_sweeps: int = 4,
                 max_update: float = 0.2, early_exit: bool | None = None,
                 per_case_stop: bool = False, use_graph: bool = True):
        self.b = base
        self.cfl0, self.cfl_max = cfl0, cfl_max
        self.restart, self.gtol = gmres_restart, gmres_tol
        self.sweeps = jacobi_sweeps
        self.max_update = max_update
        self.early_exit = (base.precision == "mixed") if early_exit is None else bool(early_exit)
        self.per_case_stop = per_case_stop
        self.use_graph = use_graph

    def knobs(self) -> _abi.Knobs:
        return self.b.knobs(gmres_restart=int(self.restart), gmres_tol=float(self.gtol),
                            jacobi_sweeps=int(self.sweeps), max_update=float(self.max_update),
                            early_exit=int(self.early_exit), use_graph=int(self.use_graph))

    def solve(self, machs, alphas, n_iter=200, tol=1e-8, verbose=True, U0=None, callback=None):
        """Batched solve over (M, alpha) cases on this grid (aero2d/implicit.py:194-253).

        callback(it, n_iter, log10_res, cl) after each iteration; True stops.
        A case whose step is not finite keeps its state and retries at CFL 1;
        after five failures it is frozen and reported not converged."""
        b = self.b
        machs = np.atleast_1d(np.asarray(machs, np.float64))
        alphas = np.atleast_1d(np.asarray(alphas, np.float64))
        nb = len(machs)
        pool = _abi.Pool(0, self.knobs())
        try:
            ids = np.array([pool.add_case(b.grid.x, b.grid.y, machs[k], alphas[k], b.prec_code)
                            for k in range(nb)], np.int32)
            if U0 is not None:
                U0 = np.asarray(U0, np.float64)
                for k in range(nb):
                    Uk = np.ascontiguousarray(U0[k])
                    _abi.call("a2d_case_set_state", pool.handle, int(ids[k]), _abi.dptr(Uk))
            loop = NewtonLoop(nb, self.cfl0, self.cfl_max)
            hist, clh = [], []
            stopped = np.zeros(nb, bool)
            cl_now = np.zeros(nb)
            self.iterations = 0
            for it in range(n_iter):
                act = np.nonzero(~loop.frozen & ~stopped)[0]
                rn_all = np.full(nb, np.nan)
                if len(act):
                    sub = np.ascontiguousarray(ids[act])
                    cfl = np.ascontiguousarray(loop.cfl[act])
                    out = np.empty((len(act), 10))
                    _abi.call("a2d_newton_step", pool.handle, _abi.iptr(sub), len(act),
                              _abi.dptr(cfl), _abi.dptr(out))
                    rn, cl = loop.update(act, out)
                    rn_all[act] = rn
                    cl_now[act] = cl
                # frozen / stopped cases keep their last residual in the history
                keep = ~np.isin(np.arange(nb), act)
                rn_all[keep] = np.nan if it == 0 else hist[-1][keep]
                hist.append(rn_all)
                clh.append(cl_now.copy())
                self.iterations = it + 1
                r0 = loop.r0
                if verbose and (it % 5 == 0 or it == n_iter - 1):
                    print(f"  it {it + 1:4d}  log10(res/res0) "
                          f"{np.array2string(np.log10(rn_all / r0), precision=2)}"
                          f"  CFL {np.array2string(loop.cfl, precision=0)}"
                          f"  Cl {np.array2string(cl_now, precision=4)}")
                done = (rn_all / r0 < tol) | loop.frozen
                if self.per_case_stop:
                    stopped |= done
                if callback is not None and callback(it + 1, n_iter, np.log10(rn_all / r0), cl_now.copy()):
                    break
                if np.all(done | stopped):
                    break
            self.cl_history = np.array(clh)
            self.frozen = loop.frozen.copy()
            hist = np.array(hist)
            cl = np.empty(nb)
            cd = np.empty(nb)
            cm = np.empty(nb)
            cp = np.empty((nb, b.ni))
            U = np.empty((nb, 4, b.ni, b.nc))
            for k in range(nb):
                f3 = np.empty(3)
                cpk = np.empty(b.ni)
                circ = np.empty(1)
                _abi.call("a2d_forces", pool.handle, int(ids[k]), _abi.dptr(f3), _abi.dptr(cpk),
                          _abi.dptr(circ))
                cl[k], cd[k], cm[k] = f3
                cp[k] = cpk
                Uk = np.empty((4, b.ni, b.nc))
                _abi.call("a2d_case_get_state", pool.handle, int(ids[k]), _abi.dptr(Uk))
                U[k] = Uk
        finally:
            pool.close()
        res = EulerResult(machs, alphas, cl, cd, cm, cp, hist, U)
        last = np.array([h[np.isfinite(h)][-1] if np.any(np.isfinite(h)) else np.nan for h in hist.T])
        res.converged = (~self.frozen) & (np.log10(last / loop.r0) < -3)
        res.converged_tol = (~self.frozen) & (last / loop.r0 < tol)
        res.iterations_per_case = loop.iters.copy()
        return res
