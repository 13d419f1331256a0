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
from .implicit import NewtonLoop, StallRule

__all__ = ["Case", "BatchResult", "BatchSolve", "solve_batch", "default_knobs", "cfl_policy",
           "path_of", "SOLVER_DEFAULTS"]


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


# Solver settings per form. "fp64" is the reference's algorithm (fixed 50
# Arnoldi steps, 4 line-Jacobi sweeps, CFL cap 300, step limit 0.2:
# aero2d/implicit.py:68-78). "mixed" is tuned for throughput: Arnoldi stops
# once the preconditioned residual estimate meets gmres_tol (at most 20
# steps), the preconditioner runs two red-black sweeps (as strong as four
# Jacobi sweeps here at half the cost), the CFL cap is the reference's 300
# until rn/r0 < 1e-3, then 1000 (a cap of 1000 from the start stalls transonic
# cases while their shocks form; raised late it shortens the tail), and the
# step limit (largest linearised relative change of rho and p per Newton
# step) is 0.5 instead of 0.2: shocks cross the grid in fewer steps (config
# 3: 235 instead of 302 Newton steps per case). Measured trade-off on the
# 256-case uneven batch: 233 cases converge at 0.5, 237 at 0.2 (243 at 0.3
# with the positivity check), at +17 % converged cases/s. Two optional
# safeguards are off by default because at 0.5 they changed no outcome:
# `fallback_max_update` drops a case to a smaller limit after its first
# rejected step, and the `positivity` knob rejects a step whose new state is
# not positive everywhere (p is quadratic in the step, so a large step can
# leave p <= 0; the reference accepts it and fails on every later step until
# the case freezes). The mixed form's line sweeps read the factored couplings
# from their FP16 copy (`precond_half`): the same Newton steps per case as
# with FP32 couplings, at three quarters of the preconditioner's cost. The
# Krylov basis is stored in FP16 (`basis_half`): a Newton step of 20 Arnoldi
# steps is 8-10 % cheaper; the 64-case batch converges in the same Newton
# steps, and on the transonic batch four runs each with the initial CFL
# perturbed by 1e-4 converge 28.5 cases on average with either basis (which
# limit-cycle-prone cases settle is chance there), at +2.3 % cases/s. Both
# forms solve the same discrete equations; where
# those have more than one steady state (transonic cases with a shock), the
# solver path decides which one a case reaches.
#
# Stall rule (both forms alike): a case that has taken 250 Newton steps and
# whose best rn/r0 over the last 200 steps is not 10x below its best before
# them is retired as not converged. Transonic cases on a limit cycle sit at
# rn/r0 ~ 1e-3 from step ~50 on; the slowest case that does converge (32
# transonic cases of 512 x 256) sat on a plateau for ~200 steps before its
# shock settled, which the window keeps.
STALL_RULE = (250, 200, 10.0)

SOLVER_DEFAULTS = {
    "fp64": dict(restart=50, early_exit=0, precond_order=0, jacobi_sweeps=4, cfl_max=300.0,
                 cfl_late=None, late_below=0.0, max_update=0.2, max_update_late=None,
                 fallback_max_update=0.0, positivity=0, precond_half=0, basis_half=0,
                 stall=STALL_RULE),
    "mixed": dict(restart=20, early_exit=1, precond_order=1, jacobi_sweeps=2, cfl_max=300.0,
                  cfl_late=1000.0, late_below=1e-3, max_update=0.5, max_update_late=None,
                  fallback_max_update=0.0, positivity=0, precond_half=1, basis_half=1,
                  stall=STALL_RULE),
}


def cfl_policy(precision: str, cfl_max: float | None = None, cfl_late: float | None = None,
               late_below: float | None = None) -> tuple:
    """(cfl_max, cfl_late, late_below) of the form, with optional overrides."""
    d = SOLVER_DEFAULTS[precision]
    return (d["cfl_max"] if cfl_max is None else float(cfl_max),
            d["cfl_late"] if cfl_late is None else float(cfl_late),
            d["late_below"] if late_below is None else float(late_below))


def default_knobs(precision: str, restart: int | None = None, early_exit: bool | None = None,
                  precond_order: int | None = None, jacobi_sweeps: int | None = None,
                  max_update: float | None = None, positivity: int | None = None,
                  precond_half: int | None = None, basis_half: int | None = None) -> _abi.Knobs:
    """Library knobs for the form's defaults, with optional overrides."""
    d = SOLVER_DEFAULTS[precision]
    k = _abi.default_knobs()
    k.gmres_restart = d["restart"] if restart is None else int(restart)
    k.early_exit = d["early_exit"] if early_exit is None else int(early_exit)
    k.precond_order = d["precond_order"] if precond_order is None else int(precond_order)
    k.jacobi_sweeps = d["jacobi_sweeps"] if jacobi_sweeps is None else int(jacobi_sweeps)
    k.max_update = d["max_update"] if max_update is None else float(max_update)
    k.positivity = d["positivity"] if positivity is None else int(positivity)
    k.precond_half = d["precond_half"] if precond_half is None else int(precond_half)
    k.basis_half = d["basis_half"] if basis_half is None else int(basis_half)
    return k


def path_of(form: str) -> dict:
    """solve_batch keyword arguments that put either precision on the solver
    path of `form` ("fp64" or "mixed"): its knobs, CFL policy and step-limit
    fallback. Running FP64 on the mixed form's path isolates the effect of
    precision."""
    d = SOLVER_DEFAULTS[form]
    kn = default_knobs("fp64", restart=d["restart"], early_exit=d["early_exit"],
                       precond_order=d["precond_order"], jacobi_sweeps=d["jacobi_sweeps"],
                       max_update=d["max_update"], positivity=d["positivity"])
    return dict(knobs=kn, cfl_max=d["cfl_max"], cfl_late=d["cfl_late"], late_below=d["late_below"],
                fallback_max_update=d["fallback_max_update"], max_update_late=d["max_update_late"] or 0.0,
                stall=d["stall"])


class BatchSolve:
    """A batch resident on the device: upload() places every case (grid
    metrics, freestream state, buffers) in one pool; run() then iterates the
    Newton loop, timing only device-resident work plus the host control."""

    def __init__(self, cases: list[Case], precision: str = "fp64", knobs: _abi.Knobs | None = None,
                 shapes: dict | None = None, pool_budget: int = 0, states: list | None = None):
        self.cases = cases
        self.states = states
        self.precision = precision
        self.prec = _abi.PREC_FP64 if precision == "fp64" else _abi.PREC_MIXED
        self.knobs = knobs if knobs is not None else default_knobs(precision)
        self.shapes = shapes
        self.pool_budget = pool_budget
        self.pool = None
        self.ids = None
        self.upload_s = 0.0

    def upload(self):
        t0 = time.perf_counter()
        self.pool = _abi.Pool(self.pool_budget, self.knobs)
        self.ids = np.empty(len(self.cases), np.int32)
        for k, c in enumerate(self.cases):
            self.ids[k] = self.pool.add_case(c.grid.x, c.grid.y, c.mach, c.alpha, self.prec)
            if self.shapes:
                sh = self.shapes.get((c.grid.x.shape[0], c.grid.x.shape[1] - 1))
                if sh is not None:
                    _abi.call("a2d_case_set_shape", self.pool.handle, int(self.ids[k]), int(sh))
            if self.states is not None and self.states[k] is not None:
                Uk = np.ascontiguousarray(self.states[k], np.float64)
                _abi.call("a2d_case_set_state", self.pool.handle, int(self.ids[k]), _abi.dptr(Uk))
        self.upload_s = time.perf_counter() - t0
        return self

    def run(self, tol: float = 1e-10, n_iter: int = 200, cfl0: float = 10.0, cfl_max: float | None = None,
            keep_state: bool = False, on_step=None, cfl_late: float | None = None,
            late_below: float | None = None, record_history: bool = False,
            fallback_max_update: float | None = None, max_update_late: float | None = None,
            stall: tuple | None = None) -> BatchResult:
        """Newton loop to rn/r0 <= tol per case (or n_iter steps); the pool is
        released at the end. fallback_max_update > 0: a case's step limit
        drops to it after the case's first rejected step (default: the form's).
        max_update_late > 0: a case's step limit rises to it once its rn/r0
        falls below late_below (default: the form's; 0 = never). stall =
        (after, window, drop): a case that has taken `after` Newton steps and
        whose best rn/r0 over the last `window` steps is not `drop` times below
        its best before them is retired as stalled (not converged); None
        takes the form's rule (STALL_RULE), () keeps every case until tol or
        n_iter."""
        if self.pool is None:
            self.upload()
        cases, ids, pool = self.cases, self.ids, self.pool
        nb = len(cases)
        cfl_max, cfl_late, late_below = cfl_policy(self.precision, cfl_max, cfl_late, late_below)
        if fallback_max_update is None:
            fallback_max_update = SOLVER_DEFAULTS[self.precision]["fallback_max_update"]
        if max_update_late is None:
            max_update_late = SOLVER_DEFAULTS[self.precision]["max_update_late"] or 0.0
        if stall is None:
            stall = SOLVER_DEFAULTS[self.precision]["stall"]
        stall = tuple(stall) if stall else None
        fallback = fallback_max_update > 0 and fallback_max_update != self.knobs.max_update
        staged = max_update_late > 0 and max_update_late != self.knobs.max_update
        t0 = time.perf_counter()
        try:
            loop = NewtonLoop(nb, cfl0, cfl_max, cfl_late, late_below, fallback)
            hist = []
            rule = StallRule(nb, *stall) if stall is not None else None
            stalled = np.zeros(nb, bool)
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
                for k in loop.newly_failed:
                    _abi.call("a2d_case_set_step_limit", pool.handle, int(ids[k]), float(fallback_max_update))
                if staged:
                    for k in loop.newly_late:
                        if not (fallback and loop.failed_once[k]):
                            _abi.call("a2d_case_set_step_limit", pool.handle, int(ids[k]),
                                      float(max_update_late))
                arn[act] += out[:, 6]
                rel = rn / loop.r0[act]
                if record_history:
                    h = np.full(nb, np.nan)
                    h[act] = rel
                    hist.append(h)
                fin = (rel <= tol) | loop.frozen[act]
                if rule is not None:
                    st = rule.update(act, rel) & ~fin
                    stalled[act[st]] = True
                    fin = fin | st
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
                _abi.call("a2d_forces", pool.handle, int(ids[k]), _abi.dptr(f3), _abi.dptr(cp),
                          _abi.dptr(circ))
                cl[k], cd[k], cm[k] = f3
            wall = time.perf_counter() - t0
            if keep_state:      # copied back outside the timed region
                for k in range(nb):
                    ni, nj = cases[k].grid.x.shape
                    Uk = np.empty((4, ni, nj - 1))
                    _abi.call("a2d_case_get_state", pool.handle, int(ids[k]), _abi.dptr(Uk))
                    states.append(Uk)
        finally:
            pool.close()
            self.pool = None
        rel = loop.last_rn / loop.r0
        conv = (~loop.frozen) & (rel <= tol)
        res = BatchResult(cl, cd, cm, conv, loop.frozen.copy(), loop.iters.copy(), rel, loop.last_rn.copy(),
                          t_done, wall, arn / np.maximum(loop.iters, 1), states)
        res.history = np.array(hist) if record_history else None
        res.upload_s = self.upload_s
        res.stalled = stalled
        return res


def solve_batch(cases: list[Case], precision: str = "fp64", tol: float = 1e-10, n_iter: int = 200,
                cfl0: float = 10.0, cfl_max: float | None = None, knobs: _abi.Knobs | None = None,
                shapes: dict | None = None, keep_state: bool = False, pool_budget: int = 0,
                on_step=None, cfl_late: float | None = None, late_below: float | None = None,
                record_history: bool = False, fallback_max_update: float | None = None,
                max_update_late: float | None = None, stall: tuple | None = None,
                states: list | None = None) -> BatchResult:
    """Upload every case, then solve each to rn/r0 <= tol (or n_iter Newton
    steps). precision: "fp64" or "mixed"; knobs and the CFL policy
    (cfl_max, cfl_late, late_below) default to SOLVER_DEFAULTS[precision].
    shapes: optional {(ni, nc): stencil shape index} from the autotuner.
    on_step(it, n_active) after each step. cfl_late / late_below: the CFL cap
    rises to cfl_late once rn/r0 < late_below (late_below = 0 never).
    fallback_max_update: a case's step limit after its first rejected step
    (0: keep the knobs' max_update). max_update_late: a case's step limit
    once its rn/r0 < late_below (0: keep it). stall: the stall rule of
    BatchSolve.run (None: the form's, (): off). states: optional initial states (4, ni, nc)
    per case (None entries start from the freestream).
    BatchResult.wall covers the Newton loop (inputs resident); .upload_s the
    host-to-device set-up before it."""
    bs = BatchSolve(cases, precision, knobs, shapes, pool_budget, states).upload()
    return bs.run(tol, n_iter, cfl0, cfl_max, keep_state, on_step, cfl_late, late_below, record_history,
                  fallback_max_update, max_update_late, stall)


def residual_components(case: Case, U: np.ndarray) -> np.ndarray:
    """Full FP64 residual of state U for `case`, per component: rms over cells
    of R_k / vol (the reference's rn for k = rho, aero2d/implicit.py:191),
    divided by the same quantity at the freestream state (the first Newton
    step's residual). Returns four ratios (rho, rho u, rho v, E)."""
    from .euler import EulerSolver
    s = EulerSolver(case.grid, dtype=np.float64)
    try:
        winf = s.freestream(case.mach, case.alpha)

        def rms(Uk):
            R, _ = s.residual(Uk, winf, float("nan"))
            return np.sqrt(np.mean((R / s.vol[None]) ** 2, axis=(1, 2)))
        return rms(np.asarray(U, np.float64)) / rms(s.init_state(winf))
    finally:
        s.close()


def cross_restart(case: Case, U: np.ndarray, precision: str = "fp64", n_iter: int = 40, **path) -> dict:
    """Restart a solve from state U (another solve's final state) and report
    whether U is a steady state of this form's discrete equations: the
    restarted solve (precision and solver path `path`, default the form's own)
    takes n_iter Newton steps; a steady U stays put. The restart begins at the
    path's largest CFL (where the solve that produced U ended), not at cfl0.
    Returns Cl/Cd/Cm of U and of the restarted state, their differences, and
    the residual components (residual_components) of U and of the restarted
    state."""
    if "cfl0" not in path:
        cmax, clate, _ = cfl_policy(precision, path.get("cfl_max"), path.get("cfl_late"), path.get("late_below"))
        path["cfl0"] = max(cmax, clate or 0.0)
    r0 = solve_batch([case], precision, tol=0.0, n_iter=0, states=[U])
    r1 = solve_batch([case], precision, tol=0.0, n_iter=n_iter, states=[U], keep_state=True, **path)
    comp = residual_components(case, r1.U[0])
    return {"cl": [float(r0.cl[0]), float(r1.cl[0])], "cd": [float(r0.cd[0]), float(r1.cd[0])],
            "cm": [float(r0.cm[0]), float(r1.cm[0])],
            "dcl": abs(float(r1.cl[0] - r0.cl[0])), "dcd": abs(float(r1.cd[0] - r0.cd[0])),
            "dcm": abs(float(r1.cm[0] - r0.cm[0])), "residual": comp.tolist(),
            "residual_start": residual_components(case, U).tolist(),
            "newton_steps": int(r1.iterations[0])}
