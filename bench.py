"""Benchmark: converged cases/s of the batched implicit Euler solve on one MI355X.

Workload (default, `--workload config3`): 32 transonic cases on 512 x 256
O-grids (NACA 0012 or 2412 by a seeded coin, M ~ U[0.72, 0.82], alpha ~
U[0, 2.5] deg; aero2d_amd.workloads.config3), each solved to a relative
density residual rn/r0 <= 1e-10 by the mixed-precision form (FP32
Jacobian-vector products, line preconditioner and Krylov arithmetic, FP64
residual, update and forces). A step is one complete batched solve of all
cases; W untimed solves, then K timed ones give `value`.

Beside it, once each: the plain FP64 form with the reference's algorithm
(speedup_mixed_vs_fp64), FP64 on the mixed form's solver path (splits the
speedup into precision and solver path), the mixed form with FP32 storage,
and the mixed form without the stall rule; per-case |dCl|, |dCd|, |dCm|
against the FP64 solve with cross-restart checks of the cases outside the
bar; live HIP-event rooflines (the mixed form's dominant kernel, the FP32
Jacobian-vector product; the FP64 residual; the FP64 coupled line sweep); a
CPU baseline (the FP64 oracle on the host cores, bounded sample).

Other workloads: config1 (one case), config2 (64 cases of 256 x 128, FP64
vs mixed), config4 (256 cases of uneven sizes), config5 (serving), explicit
(the reference's explicit scheme), kernels / kernels3 (kernel timings alone,
for counter runs).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "converged cases/s at resid 1e-10; p50/p99 latency; mixed vs FP64 on GPU"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per cell and launch (FP64): residual reads U(4), S_i(2),
# S_j(2), writes R(4) = 12 words
RES_BYTES_PER_CELL = 96
# one coupled line sweep of the preconditioner (sweeps 2-4 of each
# application; the kernel with the largest share of device time in the FP64
# solve): read g0 = S0 src (4), the neighbour lines' x (4), the blocks S0 Bim,
# S0 Bip, P and c (64), write x (4) = 76 words; the d values of the way in
# make an extra 4-word round trip
SWEEP_BYTES_PER_CELL = 608
# plain vector copy used to calibrate the HBM counters: 4 words read, 4 written
COPY_BYTES_PER_CELL = 64


def make_cases(n_cases, ni, nc, seed):
    from aero2d_amd.workloads import config2
    return config2(n_cases, ni, nc, seed)


def time_kernels(cases, precision, reps=20, knobs=None):
    """HIP-event durations (ms) of the main kernels on the pool stream, after
    one Newton step of the batch so every buffer holds live data."""
    from aero2d_amd import _abi
    from aero2d_amd.batch import default_knobs
    prec = _abi.PREC_FP64 if precision == "fp64" else _abi.PREC_MIXED
    pool = _abi.Pool(0, knobs if knobs is not None else default_knobs(precision))
    try:
        ids = np.array([pool.add_case(c.grid.x, c.grid.y, c.mach, c.alpha, prec) for c in cases], np.int32)
        out = np.empty((len(ids), 10))
        cfl = np.full(len(ids), 10.0)
        _abi.call("a2d_newton_step", pool.handle, _abi.iptr(ids), len(ids), _abi.dptr(cfl), _abi.dptr(out))
        names = ["residual", "jvp", "line_sweep", "cgs_dots", "assemble", "factor", "cgs_update",
                 "copy", "sweep_first", "sweep_rb_half", "precond"]
        res = {}
        for w, name in enumerate(names):
            ms = np.zeros(1)
            _abi.call("a2d_time_kernel", pool.handle, _abi.iptr(ids), len(ids), w, reps, _abi.dptr(ms))
            res[name] = float(ms[0])
    finally:
        pool.close()
    return res


# HBM bytes per launch of the 64-case batch (64 cases, 256 x 128, FP64) from
# two rocprofv3 --pmc runs of `bench.py --workload kernels` (FETCH_SIZE, then
# WRITE_SIZE), fetch scaled by 1.9936 (the copy kernel's known bytes over what
# the counter reads for it) and write by 1.0 (profiles/traffic_config2_fp64_pmc.txt).
TRAFFIC_CELLS = 64 * 256 * 128
MEASURED_TRAFFIC = {"line_sweep": 1483.1e6, "residual": 253.1e6}
# HBM bytes per launch at the transonic batch (32 x 512 x 256 cells): FETCH_SIZE
# x 1.9967 + WRITE_SIZE of `bench.py --workload kernels3`
# (profiles/kernels_config3_counters.txt)
MEASURED_TRAFFIC_C3 = {"sweep_rb_half": 402.6e6, "jvp": 442.9e6, "residual": 674.6e6}


def rooflines(kt, cells):
    """HBM rooflines from HIP-event kernel times (ms): the coupled line sweep
    (largest share of device time) and the residual kernel."""
    def one(name, bpc):
        ach = bpc * cells / (kt[name] * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS,
                "traffic": MEASURED_TRAFFIC.get(name) if cells == TRAFFIC_CELLS else None,
                "bytes_per_launch": bpc * cells, "ms_per_launch": kt[name]}
    sw = one("line_sweep", SWEEP_BYTES_PER_CELL)
    sw["kernel"] = "k_sweep<double, coupled, g0> (one coupled line-Jacobi sweep, every line)"
    res = one("residual", RES_BYTES_PER_CELL)
    res["kernel"] = "FP64 residual (LDS tiles or column march, as chosen for the batch)"
    return {"roofline": sw, "roofline_residual": res}


def run_kernels(a):
    """Kernel timings alone (no solve) on the default batch: used under
    rocprofv3 for the per-dispatch counters."""
    a.cases = a.cases or 64
    cases = make_cases(a.cases, a.ni, a.nc, a.seed)
    kt = time_kernels(cases, "fp64", reps=5)
    out = {"kernel_ms": kt, "cells": a.cases * a.ni * a.nc}
    out.update(rooflines(kt, out["cells"]))
    return out


def run_kernels3(a):
    """Kernel timings alone on the transonic batch (32 x 512 x 256 cells): the
    mixed form's kernels, then the FP64 residual and Jacobian-vector product on
    16 of the cases (used under rocprofv3 for the per-dispatch counters)."""
    from aero2d_amd.workloads import config3
    cases = config3(a.cases or 32)
    kt = time_kernels(cases, "mixed", reps=5)
    k64 = time_kernels(cases[:16], "fp64", reps=5)
    return {"kernel_ms_mixed": kt, "kernel_ms_fp64_16cases": k64,
            "cells": sum(c.grid.x.shape[0] * (c.grid.x.shape[1] - 1) for c in cases)}


def cpu_baseline(case, gpu_iters, steps=2, scale_by="the GPU run's mean"):
    """FP64 oracle Newton steps of one case on the host cores, scaled to a
    whole solve by `gpu_iters` Newton steps per case."""
    import torch
    from oracle import flow, newton
    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    geo = flow.stack_geometry([flow.case_geometry(case.grid.x, case.grid.y)])
    t = time.perf_counter()
    newton.solve(geo, [case.mach], [case.alpha], n_iter=steps, tol=0.0)
    dt = (time.perf_counter() - t) / steps
    return {"value": 1.0 / (dt * gpu_iters), "unit": "cases/s", "cores": cores, "kind": "port",
            "sample": (f"{steps} FP64 Newton steps of one {case.grid.x.shape[0]}x{case.grid.x.shape[1] - 1} "
                       f"case by the oracle ({dt:.2f} s/step), scaled by {scale_by} "
                       f"{gpu_iters:.1f} Newton steps per case")}


def per_case_errors(a, b):
    """|Cl|, |Cd|, |Cm| differences per case between two batch results."""
    return {"dcl": np.abs(a.cl - b.cl).tolist(), "dcd": np.abs(a.cd - b.cd).tolist(),
            "dcm": np.abs(a.cm - b.cm).tolist()}


def max_errors(a, b):
    """Largest |dCl|, |dCd|, |dCm| over the cases both results converged
    (an unconverged case has no answer to compare), and that case count."""
    both = a.converged & b.converged
    out = {"cases_compared": int(both.sum())}
    for k in ("cl", "cd", "cm"):
        d = np.abs(getattr(a, k) - getattr(b, k))[both]
        out["d" + k] = float(d.max()) if d.size else None
    return out


def mixed_path_fp64():
    """solve_batch arguments that run the FP64 form with the mixed form's
    solver settings (same preconditioner, Krylov size, early exit, CFL policy
    and step limit): differences from the mixed form are then due to precision
    alone."""
    from aero2d_amd.batch import path_of
    return path_of("mixed")


def steady_state_check(cases, ra, rb, k):
    """Full FP64 residual (all four components, rms over cells) of case k's
    final states in two results, relative to its freestream residual. Both
    below 1e-6 in every component (converged states sit at 1e-9..1e-8; the
    two states differ by ~1e-2 in Cl, so a state that was not steady would
    show a residual of that order) means both are steady states of the same
    discrete equations."""
    from aero2d_amd import EulerSolver
    c = cases[k]
    s = EulerSolver(c.grid, dtype=np.float64)
    try:
        winf = s.freestream(c.mach, c.alpha)

        def rms(U):
            R, _ = s.residual(U, winf, float("nan"))
            return np.sqrt(np.mean(R.reshape(4, -1) ** 2, axis=1))
        r0 = rms(s.init_state(winf))
        qa, qb = rms(ra.U[k]) / r0, rms(rb.U[k]) / r0
    finally:
        s.close()
    return {"case": int(k), "cl": [float(ra.cl[k]), float(rb.cl[k])],
            "max_rel_residual": [float(qa.max()), float(qb.max())],
            "both_steady": bool(qa.max() < 1e-6 and qb.max() < 1e-6)}


def compare_forms(r64, rmx, r64m, cases=None):
    """Per-case accuracy of the mixed form. precision_error_same_path: against
    FP64 on the same solver path. vs_reference_path: against FP64 with the
    reference's settings. A case above 1e-5 against the reference path is
    checked (when the final states are kept): if the full residual of both
    final states is at round-off, the two are distinct steady states of the
    same discrete equations (these transonic cases admit more than one; the
    solver path, and at large steps even the rounding, picks which)."""
    same = max_errors(rmx, r64m)
    ref = max_errors(rmx, r64)
    both = rmx.converged & r64.converged
    dref = np.abs(rmx.cl - r64.cl)
    dsame = np.abs(rmx.cl - r64m.cl)
    flagged = [int(i) for i in np.nonzero(both & (dref > 1e-5))[0]]
    out = {"precision_error_same_path": same, "vs_reference_path": ref,
           "cases_within_1e-5_of_reference_path": int((both & (dref <= 1e-5)).sum()),
           "other_steady_state_cases": flagged,
           "per_case_dcl_same_path": dsame.tolist(), "per_case_dcl_reference_path": dref.tolist()}
    if cases is not None and r64.U and rmx.U:
        checks = [steady_state_check(cases, r64, rmx, k) for k in flagged]
        out["steady_state_checks"] = checks
        out["unexplained_cases"] = [c["case"] for c in checks if not c["both_steady"]]
    return out


def run_config1(a):
    """Single case: NACA 0012 on 128 x 64 (ratio 1.1 to 20 chords), M 0.5,
    alpha 2 deg, to rn/r0 <= 1e-10, FP64 and mixed, against the committed
    oracle solve of the same case (tests/golden/oracle_config1.json)."""
    from aero2d_amd.batch import solve_batch
    from aero2d_amd.workloads import config1
    cases = config1()
    with open(os.path.join(ROOT, "tests", "golden", "oracle_config1.json")) as f:
        ref = json.load(f)
    res = {}
    for prec in ("fp64", "mixed"):
        solve_batch(cases, prec, tol=a.tol, n_iter=300)               # warm-up: plans, graphs
        r = solve_batch(cases, prec, tol=a.tol, n_iter=300)
        res[prec] = {"value": float(r.converged.sum() / r.wall), "wall_s": r.wall,
                     "converged": bool(r.converged[0]), "newton_steps": int(r.iterations[0]),
                     "cl": float(r.cl[0]), "cd": float(r.cd[0]), "cm": float(r.cm[0]),
                     "vs_oracle": {"dcl": abs(float(r.cl[0]) - ref["cl"]), "dcd": abs(float(r.cd[0]) - ref["cd"]),
                                   "dcm": abs(float(r.cm[0]) - ref["cm"])}}
        print(f"# {prec}: {r.wall:.3f} s, {int(r.iterations[0])} Newton steps, |dCl| vs oracle "
              f"{res[prec]['vs_oracle']['dcl']:.2e}", file=sys.stderr, flush=True)
    out = {"metric": METRIC, "value": res["fp64"]["value"], "unit": "cases/s", "n_gpus": 1, "steps": 1,
           "warmup": 1, "ms_per_step": 1e3 * res["fp64"]["wall_s"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "NACA 0012, 128x64 O-grid (ratio 1.1 to 20 chords), M 0.5, alpha 2 deg, "
                                  f"one case, rn/r0 <= {a.tol:g}"},
           "fp64": res["fp64"], "mixed": res["mixed"],
           "oracle": {"iterations": ref["iterations"], "cl": ref["cl"], "cd": ref["cd"], "cm": ref["cm"],
                      "full_solve_s_8_cores": ref["seconds"]}}
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cases[0], float(ref["iterations"]), steps=3,
                                           scale_by="the oracle's full solve of this case,")
    return out


def cross_restart_checks(cases, ra, rb, flagged, n_iter=80):
    """For each flagged case (converged in both results, outside the bar):
    is each end state a steady state of the same FP64 discrete equations?
    First directly: every FP64 residual component of the state <= 1e-9 of its
    freestream value (R(U) = 0 to the solves' tolerance; their stop rule bounds
    the density component by 1e-10), for both end states, which lie a Cl
    difference apart that a nonzero residual of that size could not hide.
    Then by restarts from each end state under the other form's iteration:
    FP64 on the mixed form's solver path from the mixed state (precision alone
    changed), the mixed form from the FP64 state, and, beside them, the FP64
    reference path (CFL cap 300) and a near-Newton FP64 iteration (reference
    knobs at CFL 3000) from the mixed state. A state is kept when Cl moves <=
    1e-8 and every residual component is <= 1e-9 after n_iter Newton steps.
    Both zeros of R and both kept: distinct steady states. Both zeros of R
    but one drifting under a pseudo-time iteration (a slow mode at that CFL):
    steady but drifting. An end state that is not a zero of R: unexplained."""
    from aero2d_amd.batch import cross_restart, path_of
    out = []
    kept = lambda r: r["dcl"] <= 1e-8 and max(r["residual"]) <= 1e-9
    for k in flagged:
        a = cross_restart(cases[k], ra.U[k], "fp64", n_iter=n_iter, **path_of("mixed"))
        aref = cross_restart(cases[k], ra.U[k], "fp64", n_iter=n_iter)
        anew = cross_restart(cases[k], ra.U[k], "fp64", n_iter=n_iter, cfl0=3000.0, cfl_max=3000.0)
        b = cross_restart(cases[k], rb.U[k], "mixed", n_iter=n_iter)
        steady = max(a["residual_start"]) <= 1e-9 and max(b["residual_start"]) <= 1e-9
        both_kept = kept(a) and kept(b)
        verdict = ("distinct steady states" if steady and both_kept else
                   "steady but drifting" if steady else "unexplained")
        out.append({"case": int(k), "cl_mixed": float(ra.cl[k]), "cl_fp64": float(rb.cl[k]),
                    "residual_mixed_state": a["residual_start"], "residual_fp64_state": b["residual_start"],
                    "fp64_from_mixed_state": a, "fp64_reference_path_from_mixed_state": aref,
                    "fp64_near_newton_from_mixed_state": anew, "mixed_from_fp64_state": b,
                    "kept": {"fp64_mixed_path": kept(a), "fp64_reference_path": kept(aref),
                             "fp64_near_newton": kept(anew), "mixed": kept(b)},
                    "verdict": verdict, "distinct_steady_states": verdict == "distinct steady states"})
        print(f"# cross restart case {k}: R(mixed end) {max(a['residual_start']):.1e}, R(fp64 end) "
              f"{max(b['residual_start']):.1e}; from mixed: fp64 mixed path dCl {a['dcl']:.1e} res "
              f"{max(a['residual']):.1e}, reference path dCl {aref['dcl']:.1e}, near-Newton dCl {anew['dcl']:.1e} "
              f"res {max(anew['residual']):.1e}; mixed from fp64 dCl {b['dcl']:.1e} res {max(b['residual']):.1e}: "
              f"{verdict}", file=sys.stderr, flush=True)
    return out


def accuracy(rmx, r64, cases, restart_checks=True):
    """Per-case |dCl|, |dCd|, |dCm| of the mixed form against the FP64
    reference-path solve, the count within the bar (|dCl| <= 1e-5 and
    |dCd| <= 1e-6) over the cases both converge, and cross-restart checks of
    the others."""
    both = rmx.converged & r64.converged
    dcl, dcd, dcm = (np.abs(getattr(rmx, k) - getattr(r64, k)) for k in ("cl", "cd", "cm"))
    within = both & (dcl <= 1e-5) & (dcd <= 1e-6)
    flagged = [int(k) for k in np.nonzero(both & ~within)[0]]
    out = {"cases_both_converged": int(both.sum()), "cases_within_bar": int(within.sum()),
           "bar": {"dcl": 1e-5, "dcd": 1e-6},
           "max_within_bar": {"dcl": float(dcl[within].max()) if within.any() else None,
                              "dcd": float(dcd[within].max()) if within.any() else None,
                              "dcm": float(dcm[within].max()) if within.any() else None},
           "per_case": {"dcl": dcl.tolist(), "dcd": dcd.tolist(), "dcm": dcm.tolist()},
           "outside_bar": flagged}
    if restart_checks and flagged:
        chk = cross_restart_checks(cases, rmx, r64, flagged)
        out["cross_restart"] = chk
        out["distinct_steady_state_cases"] = [c["case"] for c in chk if c["verdict"] == "distinct steady states"]
        out["steady_but_drifting_cases"] = [c["case"] for c in chk if c["verdict"] == "steady but drifting"]
        out["unexplained_cases"] = [c["case"] for c in chk if c["verdict"] == "unexplained"]
    else:
        out["distinct_steady_state_cases"] = []
        out["steady_but_drifting_cases"] = []
        out["unexplained_cases"] = []
    return out


def form_entry(r, walls=None, convs=None):
    walls = [r.wall] if walls is None else walls
    convs = [int(r.converged.sum())] if convs is None else convs
    lat = np.percentile(r.t_done, [50, 99]) * 1e3
    stalled = getattr(r, "stalled", np.zeros(len(r.cl), bool))
    return {"value": float(np.sum(convs) / np.sum(walls)), "unit": "cases/s",
            "converged": int(r.converged.sum()), "cases": len(r.cl),
            "not_converged": [int(k) for k in np.nonzero(~r.converged)[0]],
            "retired_by_stall_rule": [int(k) for k in np.nonzero(stalled)[0]],
            "wall_s": float(np.mean(walls)), "newton_steps_mean": float(r.iterations.mean()),
            "arnoldi_per_newton_mean": float(r.arnoldi.mean()),
            "latency_ms": {"p50": float(lat[0]), "p99": float(lat[1])},
            "upload_s": float(r.upload_s),
            "value_pcie_inclusive": float(np.sum(convs) / (np.sum(walls) + len(walls) * r.upload_s))}


# Dominant kernel of the mixed transonic solve (largest single share of device
# time, profiles/solve_config3_mixed_trace.txt): the coupled red-black half
# sweep with FP16 couplings, per cell of its colour: read S0 src (g0, 4 FP32),
# the two neighbour lines' x (8 FP32), the four coupling blocks (64 FP16);
# write x (4 FP32) = 192 B. Next: the FP32 Jacobian-vector product, 17 words
# per cell: read U 4, v 4, face metrics S_i 2 and S_j 2, diag 1; write 4.
SWEEP_RB_BYTES_PER_COLOUR_CELL = 192
JVP32_BYTES_PER_CELL = 68


def run_config3(a):
    """Transonic mixed vs FP64: 32 NACA 0012/2412 cases on 512 x 256, the
    workload the >= 10x mixed-over-FP64 target is stated on. K timed mixed
    solves after W untimed ones give `value`; the FP64 reference-path solve,
    FP64 on the mixed form's path, the mixed form with FP32 storage and the
    mixed form without the stall rule run once each."""
    from aero2d_amd.batch import SOLVER_DEFAULTS, default_knobs, path_of, solve_batch
    from aero2d_amd.workloads import config3
    cases = config3(a.cases or 32)
    n_iter = a.n_iter or 800
    if a.mixed_only:
        # the mixed solve alone (used under the profiler)
        rmx = solve_batch(cases, "mixed", tol=a.tol, n_iter=n_iter)
        return {"workload": f"{len(cases)} transonic cases on 512x256, mixed only", "wall_s": rmx.wall,
                "converged": int(rmx.converged.sum()), "newton_steps_mean": float(rmx.iterations.mean())}
    for w in range(a.warmup):
        r = solve_batch(cases, "mixed", tol=a.tol, n_iter=n_iter)
        print(f"# warmup mixed: {r.wall:.2f} s, converged {int(r.converged.sum())}", file=sys.stderr, flush=True)
    walls, convs, rmx = [], [], None
    for s in range(a.steps):
        # the last timed solve keeps its final states (copied after the timed region)
        rmx = solve_batch(cases, "mixed", tol=a.tol, n_iter=n_iter, keep_state=s == a.steps - 1)
        walls.append(rmx.wall)
        convs.append(int(rmx.converged.sum()))
        print(f"# mixed step {s}: {rmx.wall:.2f} s, converged {convs[-1]}/{len(cases)}, Newton steps mean "
              f"{rmx.iterations.mean():.1f}", file=sys.stderr, flush=True)
    r64 = solve_batch(cases, "fp64", tol=a.tol, n_iter=n_iter, keep_state=True)
    print(f"# fp64 (reference path): {r64.wall:.1f} s, converged {int(r64.converged.sum())}",
          file=sys.stderr, flush=True)
    r64m = solve_batch(cases, "fp64", tol=a.tol, n_iter=n_iter, **path_of("mixed"))
    print(f"# fp64 on the mixed form's path: {r64m.wall:.1f} s, converged {int(r64m.converged.sum())}",
          file=sys.stderr, flush=True)
    r32 = solve_batch(cases, "mixed", tol=a.tol, n_iter=n_iter,
                      knobs=default_knobs("mixed", precond_half=0, basis_half=0))
    print(f"# mixed, FP32 storage: {r32.wall:.1f} s, converged {int(r32.converged.sum())}",
          file=sys.stderr, flush=True)
    rns = solve_batch(cases, "mixed", tol=a.tol, n_iter=n_iter, stall=())
    print(f"# mixed, no stall rule: {rns.wall:.1f} s, converged {int(rns.converged.sum())}",
          file=sys.stderr, flush=True)
    vmx = float(np.sum(convs) / np.sum(walls))
    v64 = float(r64.converged.sum() / r64.wall)
    v64m = float(r64m.converged.sum() / r64m.wall)
    v32 = float(r32.converged.sum() / r32.wall)
    acc = accuracy(rmx, r64, cases, restart_checks=not a.no_restart_checks)
    acc32 = accuracy(r32, r64, cases, restart_checks=False)
    # precision alone: the mixed form against FP64 on the mixed form's solver path
    both_m = rmx.converged & r64m.converged
    dm = {k: np.abs(getattr(rmx, k) - getattr(r64m, k)) for k in ("cl", "cd", "cm")}
    within_m = both_m & (dm["cl"] <= 1e-5) & (dm["cd"] <= 1e-6)
    acc_same = {"cases_both_converged": int(both_m.sum()), "cases_within_bar": int(within_m.sum()),
                "outside_bar": [int(k) for k in np.nonzero(both_m & ~within_m)[0]],
                "max_within_bar": {k: float(v[within_m].max()) if within_m.any() else None for k, v in dm.items()},
                "per_case_dcl": dm["cl"].tolist()}
    out = {"metric": METRIC, "value": vmx, "unit": "cases/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": 1e3 * float(np.mean(walls)), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32 inner / f64 outer", "data": "synthetic",
           "config": {"workload": f"{len(cases)} transonic NACA 0012/2412 cases on 512x256 O-grids "
                                  f"(M 0.72-0.82, alpha 0-2.5 deg), mixed precision, rn/r0 <= {a.tol:g}; "
                                  "a step is one batched solve of all cases",
                      "cases": len(cases), "ni": 512, "nc": 256, "tol": a.tol, "n_iter": n_iter},
           "converged": int(rmx.converged.sum()),
           "speedup_mixed_vs_fp64": vmx / v64 if v64 > 0 else None,
           "speedup_split": {"precision": v64m and vmx / v64m, "solver_path": v64 and v64m / v64,
                             "note": "mixed / FP64-on-the-mixed-path is the precision part; "
                                     "FP64-on-the-mixed-path / FP64 reference path the solver-path part"},
           "claim": "the >= 10x target is stated for the mixed form's default (FP16 storage of the Krylov "
                    "basis and the factored couplings) against the FP64 reference algorithm; "
                    "mixed_fp32_storage is the form with FP32 storage",
           "mixed": form_entry(rmx, walls, convs) | {"knobs": SOLVER_DEFAULTS["mixed"],
                                                     "accuracy_vs_fp64": acc,
                                                     "accuracy_vs_fp64_same_path": acc_same},
           "fp64": form_entry(r64) | {"knobs": SOLVER_DEFAULTS["fp64"]},
           "fp64_mixed_path": form_entry(r64m),
           "mixed_fp32_storage": form_entry(r32) | {"speedup_vs_fp64": v32 / v64 if v64 > 0 else None,
                                                    "accuracy_vs_fp64": {k: acc32[k] for k in (
                                                        "cases_both_converged", "cases_within_bar",
                                                        "outside_bar", "max_within_bar")}},
           "mixed_no_stall_rule": form_entry(rns),
           "cases": [(c.name, c.mach, c.alpha) for c in cases]}
    if not a.no_kernel_timing:
        kt = time_kernels(cases, "mixed", reps=20)
        cells = sum(c.grid.x.shape[0] * (c.grid.x.shape[1] - 1) for c in cases)
        rl = {}
        for name, nbytes, kname in (
                ("sweep_rb_half", SWEEP_RB_BYTES_PER_COLOUR_CELL * cells // 2,
                 "k_sweep<float, coupled, g0, FP16 couplings, 32 lines, d in LDS> (one red-black half sweep)"),
                ("jvp", JVP32_BYTES_PER_CELL * cells,
                 "k_march<float, tangent> (FP32 Jacobian-vector product, column march)"),
                ("residual", RES_BYTES_PER_CELL * cells, "k_march<double, residual> (FP64 residual, column march)")):
            ach = nbytes / (kt[name] * 1e-3) / 1e9
            rl[name] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": ach / HBM_PEAK_GBS, "traffic": MEASURED_TRAFFIC_C3.get(name),
                        "bytes_per_launch": nbytes, "ms_per_launch": kt[name], "kernel": kname}
        out["roofline"] = rl["sweep_rb_half"]
        out["roofline_jvp"] = rl["jvp"]
        out["roofline_jvp"]["note"] = "instruction bound: VALU busy 0.93 (profiles/kernels_config3_counters.txt)"
        out["roofline_residual"] = rl["residual"]
        out["kernel_ms_mixed"] = kt
        k64 = time_kernels(cases[:16], "fp64", reps=10)
        c64 = sum(c.grid.x.shape[0] * (c.grid.x.shape[1] - 1) for c in cases[:16])
        sw = rooflines(k64, c64)["roofline"]
        sw["cells"] = c64
        out["roofline_fp64_sweep"] = sw
        out["kernel_ms_fp64_16cases"] = k64
    if not a.no_cpu_baseline:
        conv = r64.converged
        steps = float(r64.iterations[conv].mean()) if conv.any() else float(r64.iterations.mean())
        cb = cpu_baseline(cases[0], steps, scale_by="the GPU FP64 run's mean over converged cases,")
        # the oracle runs the FP64 reference algorithm: it converges the same cases
        cb["value"] *= float(conv.mean())
        cb["sample"] += f"; times the FP64 converged fraction {conv.mean():.3f}"
        out["cpu_baseline"] = cb
    return out


def run_config4(a):
    """Uneven persistent batch: 256 random sections, four sizes, mixed form,
    autotuned tile shapes."""
    from aero2d_amd.autotune import shapes_for
    from aero2d_amd.batch import solve_batch
    from aero2d_amd.workloads import config4
    cases, redraws = config4(a.cases or 256)
    shapes = shapes_for(cases, "mixed")
    r = solve_batch(cases, "mixed", tol=a.tol, n_iter=a.n_iter or 800, shapes=shapes)
    sizes = {}
    for c, conv, it in zip(cases, r.converged, r.iterations):
        key = f"{c.grid.x.shape[0]}x{c.grid.x.shape[1] - 1}"
        d = sizes.setdefault(key, {"cases": 0, "converged": 0, "newton_steps": 0})
        d["cases"] += 1
        d["converged"] += int(conv)
        d["newton_steps"] += int(it)
    lat = np.percentile(r.t_done, [50, 99]) * 1e3
    return {"metric": METRIC, "value": float(r.converged.sum() / r.wall), "unit": "cases/s", "n_gpus": 1,
            "steps": 1, "warmup": 0, "ms_per_step": 1e3 * r.wall, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32 inner / f64 outer", "data": "synthetic",
            "config": {"workload": f"{len(cases)} random NACA 4-digit cases on 128x64..512x256, mixed, "
                                   f"rn/r0 <= {a.tol:g}", "redraws": redraws},
            "converged": int(r.converged.sum()), "latency_ms": {"p50": float(lat[0]), "p99": float(lat[1])},
            "tile_shapes": {f"{k[0]}x{k[1]}": v for k, v in shapes.items()}, "by_size": sizes}


# explicit scheme, algorithmic bytes per cell and launch (FP64): smoothing
# along j reads R (4 words) and writes the smoothed residual (4); a later
# Runge-Kutta stage reads it (4), U0 (4) and vol/dt (1) and writes U (4)
IRS_J_BYTES_PER_CELL = 64
RK_ROWS_BYTES_PER_CELL = 104


def explicit_pool(cases, mg, knobs):
    """FP64 cases of the explicit scheme, each with mg grid levels (every
    other node of the level above), linked finest to coarsest."""
    from aero2d_amd import _abi
    pool = _abi.Pool(0, knobs)
    ids = []
    for c in cases:
        x, y, prev = c.grid.x, c.grid.y, None
        for _ in range(mg):
            cid = pool.add_case(x, y, c.mach, c.alpha, _abi.PREC_FP64)
            if prev is None:
                ids.append(cid)
            else:
                pool.link_coarse(prev, cid)
            prev, x, y = cid, x[::2, ::2], y[::2, ::2]
    return pool, np.array(ids, np.int32)


def run_explicit(a):
    """The explicit scheme (the reference's EulerSolver.solve: RK4, residual
    smoothing, FAS V-cycle), FP64: cycles/s of the transonic batch (32 x 512
    x 256) with one and two grid levels, HIP-event rooflines of its kernels, a
    converged solve with the reference's defaults on eight subsonic cases of
    the reference's golden grid (NACA 0012, 128 x 48), and the oracle's cycle
    time on the host as the CPU baseline."""
    import torch
    from aero2d_amd import EulerSolver, ImplicitEuler, _abi
    from aero2d_amd.workloads import config1, config3
    from oracle import explicit as oexplicit
    cases = config3(a.cases or 32)
    cells = sum(c.grid.x.shape[0] * (c.grid.x.shape[1] - 1) for c in cases)
    k = _abi.default_knobs()
    out = {"metric": "explicit case-cycles/s (RK4 + residual smoothing [+ FAS])", "unit": "case-cycles/s",
           "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic", "config": {"workload": f"{len(cases)} transonic cases on 512x256, explicit "
                                                       "cycles, CFL 2.5, irs_eps 0.6"}}
    cyc = max(2, a.steps)
    for mg in (1, 2):
        pool, ids = explicit_pool(cases, mg, k)
        try:
            pool.explicit_cycles(ids, max(1, a.warmup), True)
            t = time.perf_counter()
            r = pool.explicit_cycles(ids, cyc, False)
            dt = time.perf_counter() - t
            entry = {"ms_per_cycle": 1e3 * dt / cyc, "case_cycles_per_s": len(ids) * cyc / dt,
                     "finite": bool(np.all(np.isfinite(r)))}
            if mg == 1:
                kt = {}
                for name, w in (("residual", 0), ("irs_j", 11), ("rk_rows", 12)):
                    ms = np.zeros(1)
                    _abi.call("a2d_time_kernel", pool.handle, _abi.iptr(ids), len(ids), w, 5, _abi.dptr(ms))
                    kt[name] = float(ms[0])
                entry["kernel_ms"] = kt
                for name, bpc in (("irs_j", IRS_J_BYTES_PER_CELL), ("rk_rows", RK_ROWS_BYTES_PER_CELL),
                                  ("residual", RES_BYTES_PER_CELL)):
                    ach = bpc * cells / (kt[name] * 1e-3) / 1e9
                    entry["roofline_" + name] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                                                 "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": None,
                                                 "bytes_per_launch": bpc * cells, "ms_per_launch": kt[name]}
                entry["stage_kernels_ms"] = kt["residual"] + kt["irs_j"] + kt["rk_rows"]
            out[f"mg{mg}"] = entry
            print(f"# explicit mg{mg}: {json.dumps(entry)}", file=sys.stderr, flush=True)
        finally:
            pool.close()
    out["value"] = out["mg1"]["case_cycles_per_s"]
    out["ms_per_step"] = out["mg1"]["ms_per_cycle"]
    out["steps"], out["warmup"] = cyc, max(1, a.warmup)
    # converged solves: two grid levels to res/res0 < 1e-9 (chunks of 200
    # cycles), and the reference's defaults (one level, 4000 cycles, res/res0
    # < 1e-7 or Cl, Cd settled), against the implicit FP64 steady state
    g = config1()[0].grid
    machs = np.array([0.3, 0.4, 0.5, 0.5, 0.55, 0.6, 0.6, 0.65])
    alphas = np.array([0.0, 1.0, 2.0, 3.0, 1.5, 0.0, 2.0, 1.0])
    solves = [] if a.no_solves else [("solve_mg2", 2, dict(n_iter=20000, tol=1e-9, tol_cl=0.0, tol_cd=0.0)),
                                      ("solve_reference_defaults", 1, {})]
    if solves:
        ri = ImplicitEuler(EulerSolver(g, dtype=np.float64)).solve(machs, alphas, n_iter=200, tol=1e-12,
                                                                   verbose=False)
    for key, mg, kw in solves:
        s = EulerSolver(g, dtype=np.float64, mg_levels=mg)
        t = time.perf_counter()
        r = s.solve(machs, alphas, verbose=False, **kw)
        dt = time.perf_counter() - t
        drop = r.res_hist[-1] / r.res_hist[0]
        out[key] = {"workload": f"8 subsonic NACA 0012 cases (M 0.3-0.65) on {g.x.shape[0]}x{g.x.shape[1] - 1}, "
                                f"{mg} grid level(s), {kw or 'reference defaults'}",
                    "wall_s": dt, "cycles": s.iterations, "worst_log10_res_drop": float(np.max(np.log10(drop))),
                    "max_dcl_vs_implicit_fp64": float(np.max(np.abs(r.cl - ri.cl))),
                    "max_dcd_vs_implicit_fp64": float(np.max(np.abs(r.cd - ri.cd)))}
        print(f"# explicit {key}: {json.dumps(out[key])}", file=sys.stderr, flush=True)
    if not a.no_cpu_baseline:
        cores = len(os.sched_getaffinity(0))
        torch.set_num_threads(cores)
        c = cases[0]
        lev = oexplicit.make_levels(c.grid.x, c.grid.y)
        from oracle import flow
        winf = flow.freestream(torch.tensor(c.mach, dtype=torch.float64), torch.tensor(c.alpha, dtype=torch.float64))
        U = flow.init_state(winf, lev.ni, lev.nc)
        t = time.perf_counter()
        n = 2
        for _ in range(n):
            U, _ = oexplicit.cycle(U, winf, torch.zeros((), dtype=torch.float64), lev, 2.5, 0.6, 0.02, True)
        dt = (time.perf_counter() - t) / n
        out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "case-cycles/s", "cores": cores, "kind": "port",
                               "sample": f"{n} explicit cycles (one grid level) of one 512x256 case by the "
                                         f"oracle, {dt:.2f} s/cycle"}
    return out


def run_explicit_kernels(a):
    """The explicit scheme's kernels alone on the transonic batch (32 x 512 x
    256 cells, FP64, one grid level), after one cycle, with the plain copy for
    counter calibration (used under rocprofv3 for the per-dispatch counters)."""
    from aero2d_amd import _abi
    from aero2d_amd.workloads import config3
    cases = config3(a.cases or 32)
    pool, ids = explicit_pool(cases, 1, _abi.default_knobs())
    try:
        pool.explicit_cycles(ids, 1, True)
        kt = {}
        for name, w in (("residual", 0), ("irs_j", 11), ("rk_rows", 12), ("copy", 7)):
            ms = np.zeros(1)
            _abi.call("a2d_time_kernel", pool.handle, _abi.iptr(ids), len(ids), w, 5, _abi.dptr(ms))
            kt[name] = float(ms[0])
    finally:
        pool.close()
    return {"kernel_ms": kt, "cells": sum(c.grid.x.shape[0] * (c.grid.x.shape[1] - 1) for c in cases)}


def run_config5(a):
    """Serving: open-loop Poisson streams at several rates, mixed and FP64.
    Every released request counts in the latency percentiles (an unfinished
    one at its time so far, a lower bound: serve.censored_latency); each rate
    runs for the horizon (default 60 s) and reports the completed fraction,
    converged requests per second, the stall-rule retirements and the pool's
    graph-replay fraction."""
    from aero2d_amd.serve import censored_latency, serve_stream
    from aero2d_amd.workloads import config5
    out = {"metric": METRIC, "unit": "converged requests/s", "n_gpus": 1, "steps": 1, "warmup": 0,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "data": "synthetic",
           "config": {"workload": "open-loop Poisson single-case solves, NACA 0012/2412 transonic "
                                  f"(M 0.72-0.82), 128x64..768x384, {a.horizon:g} s per rate, rn/r0 <= {a.tol:g}"},
           "rates": {}}
    plan = {"mixed": [float(x) for x in a.rates.split(",")],
            "fp64": [float(x) for x in a.rates64.split(",")]}
    for prec in a.precisions.split(","):
        for rate in plan[prec]:
            reqs = config5(rate, max(8, int(np.ceil(rate * a.horizon))))
            res, srv, wall = serve_stream(reqs, prec, tol=a.tol, horizon=a.horizon)
            lat = censored_latency(res, srv.unfinished, wall)
            conv = sum(x.converged for x in res)
            st = srv.final_stats
            entry = {"rate": rate, "wall_s": wall, **lat,
                     "converged": int(conv), "converged_per_s": conv / wall,
                     "completed_per_s": len(res) / wall,
                     "stalled": int(sum(x.stalled for x in res)),
                     "newton_steps": srv.steps,
                     "graph_replay_fraction": st["graph_replay_fraction"], "plans_built": st["plans"],
                     "slabs_allocated": st["slabs_allocated"], "slabs_reused": st["slabs_reused"]}
            out["rates"].setdefault(prec, {})[str(rate)] = entry
            print(f"# {prec} rate {rate}: {json.dumps(entry)}", file=sys.stderr, flush=True)
    first = a.precisions.split(",")[0]
    out["value"] = max(e["converged_per_s"] for e in out["rates"][first].values())
    out["dtype"] = "f32 inner / f64 outer" if first == "mixed" else "f64"
    return out


def rans_mesh(n_side=97, nj=65, ds_wall=4e-6):
    """NACA 0012 (closed trailing edge) C-grid for the RANS workload: n_side
    intervals per side (steps 2e-3 at both edges), a wake cut to the 20-chord
    far field, first layer 1e-3 re-sampled to a wall cell of ds_wall (y+ ~ 1
    at Re 6e6, growing by 1.15 per column along the cut), all quadrilaterals."""
    from aero2d_amd import naca4
    from aero2d_amd.rans_mesh import cgrid, cgrid_mesh, refine_cwall
    g = refine_cwall(cgrid(naca4("0012", 201), n_side=n_side, nj=nj, ds0=1e-3, radius=20.0), ds_wall,
                     layers=1, ratio=1.2)
    return cgrid_mesh(g)


# NASA Turbulence Modeling Resource, 2-D NACA 0012 validation, SA, CFL3D on
# its finest grid (M 0.15, Re 6e6); and the reference's own figures on its
# 14k-cell mesh (README, "verify_rans.py quick", limiter at trailing edges)
RANS_TMR = {0.0: {"cl": 0.0, "cd": 0.00819}, 10.0: {"cl": 1.0909, "cd": 0.01231}}
RANS_REFERENCE = {0.0: {"iterations": 75, "wall_s": 209.0, "cores": 4},
                  10.0: {"iterations": 83, "wall_s": 256.0, "cores": 4, "cl": 1.0894, "cd": 0.01095}}


def rans_cpu_baseline(m, mach, re, alpha, gmres_per_iter, fill=3):
    """The oracle (PyTorch CPU FP64, oracle/rans.py) timed on the host cores:
    residual, Jacobian-vector product, first-order Jacobian, BlockILU factor and
    solve, each once on the same mesh and state, composed into one Newton
    iteration with the GPU run's mean GMRES iterations per Newton iteration."""
    import torch
    from oracle.rans import BlockILU, RansOracle
    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    o = RansOracle(m, mach, re)
    winf = o.freestream(alpha)
    f = np.clip(o.wall_d / 0.02, 0, 1)[:, None]
    w0 = np.tile(winf, (o.nc, 1))
    w0[:, 1:3] *= f
    Q = o.cons(torch.tensor(w0))
    v = torch.tensor(np.random.default_rng(0).normal(size=(o.nc, 5)))
    t = {}
    t0 = time.perf_counter()
    R = o.residual(Q, winf, 0.0)
    t["residual"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.jvp(Q, v, winf, 0.0)
    t["jvp"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    D, off = o.jac1_blocks(Q, winf, 0.0)
    D = D.numpy().copy()
    D[:, np.arange(5), np.arange(5)] += (o.vol / o.local_dt(Q.numpy(), 10.0))[:, None]
    t["jacobian"] = time.perf_counter() - t0
    r, c = o.pattern()
    t0 = time.perf_counter()
    ilu = BlockILU(o.nc, r, c, fill=fill).factor(np.concatenate([D, off.numpy()]))
    t["ilu_factor"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    ilu.solve(R.numpy().ravel())
    t["ilu_solve"] = time.perf_counter() - t0
    per_iter = (t["residual"] + t["jacobian"] + t["ilu_factor"]
                + (gmres_per_iter + 2) * (t["jvp"] + t["ilu_solve"]))
    return {"value": 1.0 / per_iter, "unit": "Newton iterations/s", "cores": cores, "kind": "port",
            "sample": "one oracle residual, JVP, first-order Jacobian, ILU(3) factor and solve on the same "
                      f"{o.nc}-cell mesh ({', '.join(f'{k} {v:.2f} s' for k, v in t.items())}), composed into "
                      f"a Newton iteration with {gmres_per_iter:.1f} GMRES iterations"}


def run_rans(a):
    """RANS (SA-neg) Newton-Krylov on the device, on the reference's own
    validation case (NACA 0012, M 0.15, Re 6e6; its solver defaults): a cold
    start at alpha 0 and alpha 10 continued from it, each to the reference's
    convergence rule, with forces against the NASA TMR values; Newton
    iterations per second; the device time of the kernels and the roofline of
    the dominant one (the ILU triangular solve); a polar stepped in case slots."""
    from aero2d_amd.rans import RansSolver
    m = rans_mesh()
    mach, re, alpha = 0.15, 6e6, 0.0
    s = RansSolver(m, mach, re)
    info = s.ilu_info()
    iters = a.n_iter or 300
    for _ in range(a.warmup):
        s.solve(alpha, max_iter=5, verbose=False)
    t0 = time.perf_counter()
    r = s.solve(alpha, max_iter=iters, verbose=False)
    wall = time.perf_counter() - t0
    its = r.iterations
    t0 = time.perf_counter()
    r10 = s.solve(10.0, Q0=r.state, alpha0=alpha, max_iter=iters, verbose=False)
    wall10 = time.perf_counter() - t0

    def solve_rec(res, w, al):
        tmr = RANS_TMR[al]
        ref = RANS_REFERENCE[al]
        return {"alpha": al, "status": res.status, "converged": bool(res.converged), "wall_s": w,
                "iterations": res.iterations, "cl": res.cl, "cd": res.cd, "cdp": res.cdp, "cdf": res.cdf,
                "tmr": tmr, "cd_vs_tmr": res.cd / tmr["cd"] - 1.0,
                "cl_vs_tmr": (res.cl / tmr["cl"] - 1.0) if tmr["cl"] else res.cl,
                "reference_published": ref, "wall_vs_reference": ref["wall_s"] / w,
                "gmres_iterations_mean": float(np.mean(res.history["gmres"])) if "gmres" in res.history else None}
    line = {"metric": "RANS Newton iterations/s (SA-neg, exact JVP, ILU(3)-GMRES)", "value": its / wall,
            "unit": "Newton iterations/s", "n_gpus": 1, "steps": its, "warmup": a.warmup,
            "ms_per_step": 1e3 * wall / its, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"RANS SA NACA 0012 M {mach} Re {re:g} alpha {alpha}, {m.n_cells}-cell C-grid "
                                   "(y+ ~ 1, 20-chord far field), cold start to the reference's convergence rule "
                                   "with its solver defaults",
                       "cells": m.n_cells, "ilu_nnz_blocks": info["nnz"], "ilu_levels": info["levels"],
                       "max_iter": iters},
            "solve": solve_rec(r, wall, 0.0), "solve_alpha10": solve_rec(r10, wall10, 10.0)}
    if a.prec_lag > 1:
        # the reference's prec_lag: the last factorisation kept up to prec_lag
        # iterations while the CFL stays within a factor 2 (its default, kept
        # above, factors every iteration)
        sl = RansSolver(m, mach, re)
        sl.prec_lag = a.prec_lag
        t0 = time.perf_counter()
        rl = sl.solve(alpha, max_iter=iters, verbose=False)
        wl = time.perf_counter() - t0
        t0 = time.perf_counter()
        rl10 = sl.solve(10.0, Q0=rl.state, alpha0=alpha, max_iter=iters, verbose=False)
        wl10 = time.perf_counter() - t0
        sl.close()
        line["prec_lag"] = {"prec_lag": a.prec_lag, "solve": solve_rec(rl, wl, 0.0),
                            "solve_alpha10": solve_rec(rl10, wl10, 10.0),
                            "factorisations": [int(np.sum(x.history["factored"])) for x in (rl, rl10)],
                            "newton_iterations_per_s": (rl.iterations + rl10.iterations) / (wl + wl10)}
    kt = {k: s.time_kernel(k, alpha, reps=10 if k == "ilu_factor" else 30) for k in s.KERNELS}
    line["kernel_ms"] = kt
    n, nnz = m.n_cells, info["nnz"]
    nfi = int((m.face_tag == 0).sum())
    solve_bytes = 204 * (nnz + n) + 32 * n + 160 * n   # staged blocks with columns, row records, y
    line["roofline"] = {"bound": "hbm", "achieved": solve_bytes / (kt["ilu_solve"] * 1e-3) / 1e9, "peak": 8000.0,
                        "unit": "GB/s", "traffic": None, "bytes_per_launch": solve_bytes,
                        "ms_per_launch": kt["ilu_solve"],
                        "frac": solve_bytes / (kt["ilu_solve"] * 1e-3) / 8e12,
                        "kernel": "k_ilu_solve (block ILU(3) forward and backward sweeps, one workgroup)",
                        "note": f"latency bound: {info['levels'] + info['levels_back']} dependent levels, "
                                f"{1e3 * kt['ilu_solve'] / (info['levels'] + info['levels_back']):.2f} us each"}
    mv_bytes = 8 * (22 * n + 12 * nfi)
    line["roofline_matvec"] = {"bound": "hbm", "achieved": mv_bytes / (kt["matvec"] * 1e-3) / 1e9, "peak": 8000.0,
                               "unit": "GB/s", "bytes_per_launch": mv_bytes, "ms_per_launch": kt["matvec"],
                               "frac": mv_bytes / (kt["matvec"] * 1e-3) / 8e12,
                               "kernel": "Jacobian-vector product (k_grad, k_limit, k_face, k_cell in dual numbers)"}
    if a.cases != 0:
        # a polar: every angle in its own case slot, all stepped together,
        # warm-started from the converged alpha-0 state turned to each angle
        # (the continuation the reference's polars use, aero2d/rans.py:2375-2436)
        nang = a.cases or 16
        alphas = list(np.linspace(-4.0, 11.0, nang))
        sp = RansSolver(m, mach, re)
        t0 = time.perf_counter()
        rp = sp.solve_polar(alphas, Q0=[r.state] * nang, alpha0=alpha, max_iter=iters)
        wp = time.perf_counter() - t0
        conv = int(sum(x.converged for x in rp))
        line["polar"] = {"angles": nang, "alphas": [float(x) for x in alphas], "wall_s": wp,
                         "start": "the converged alpha-0 state, turned to each angle",
                         "newton_iterations_per_s": float(sum(x.iterations for x in rp)) / wp,
                         "converged": conv, "converged_angles_per_s": conv / wp,
                         "single_angle_converged_per_s": (1.0 / wall10) if r10.converged else None,
                         "speedup_vs_one_at_a_time": (conv / wp) * wall10 if r10.converged else None,
                         "newton_iterations": [int(x.iterations) for x in rp],
                         "cl": [float(x.cl) for x in rp], "cd": [float(x.cd) for x in rp],
                         "status": [x.status for x in rp]}
        if a.polar_workers > 1:
            t0 = time.perf_counter()
            rw = sp.solve_polar(alphas, Q0=[r.state] * nang, alpha0=alpha, max_iter=iters, workers=a.polar_workers)
            ww = time.perf_counter() - t0
            cw = int(sum(x.converged for x in rw))
            line["polar"]["workers"] = {
                "workers": a.polar_workers, "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
                "wall_s": ww, "converged": cw, "converged_angles_per_s": cw / ww,
                "speedup_vs_one_solver": wp / ww,
                "same_states_as_one_solver": all(np.array_equal(x.state, y.state) for x, y in zip(rp, rw))}
        sp.close()
    if not a.no_cpu_baseline:
        gpi = line["solve"]["gmres_iterations_mean"] or 10.0
        line["cpu_baseline"] = rans_cpu_baseline(m, mach, re, alpha, gpi)
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="config3",
                    choices=["config1", "config2", "config3", "config4", "config5", "kernels", "kernels3",
                             "explicit", "explicit_kernels", "rans"])
    ap.add_argument("--rates", default="1,2,4,10,25,100,400", help="config5: mixed-form request rates")
    ap.add_argument("--rates64", default="0.25,0.5,1,10,100", help="config5: FP64 request rates")
    ap.add_argument("--horizon", type=float, default=60.0)
    ap.add_argument("--precisions", default="mixed,fp64")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cases", type=int, default=None)
    ap.add_argument("--ni", type=int, default=256)
    ap.add_argument("--nc", type=int, default=128)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--n-iter", type=int, default=None)
    ap.add_argument("--seed", type=int, default=20261203 + 2)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--mixed-only", action="store_true", help="config3: the mixed solve alone")
    ap.add_argument("--no-solves", action="store_true", help="explicit: cycle timing only")
    ap.add_argument("--polar-workers", type=int, default=4,
                    help="rans: solvers (each with its own stream and host thread) the polar's angles are dealt to")
    ap.add_argument("--prec-lag", type=int, default=3,
                    help="rans: also solve with the ILU factorisation kept up to this many iterations (1: skip)")
    ap.add_argument("--no-restart-checks", action="store_true",
                    help="config3: skip the cross-restart checks of cases outside the accuracy bar")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {a.gpus} is not supported (cases are independent; "
              "a multi-GPU run would shard the batch across GPUs with no communication)", file=sys.stderr)
        sys.exit(2)
    if a.workload == "rans" and a.polar_workers > 1:
        # the polar's worker solvers each drive their own stream; they run side
        # by side only on hardware queues of their own (HIP's default is 4),
        # which the runtime fixes when it starts, before any HIP call here
        q = max(int(os.environ.get("GPU_MAX_HW_QUEUES", "4")), 2 * a.polar_workers)
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(q, 32))

    if a.workload != "config2":
        fn = {"config1": run_config1, "config3": run_config3, "config4": run_config4, "config5": run_config5,
              "kernels": run_kernels, "kernels3": run_kernels3, "explicit": run_explicit,
              "explicit_kernels": run_explicit_kernels, "rans": run_rans}[a.workload]
        print(json.dumps(fn(a)), flush=True)
        return

    from aero2d_amd.batch import solve_batch

    a.cases = a.cases or 64
    a.n_iter = a.n_iter or 150
    cases = make_cases(a.cases, a.ni, a.nc, a.seed)
    print(f"# {len(cases)} cases on {a.ni}x{a.nc} grids", file=sys.stderr, flush=True)

    uploads = []

    def run(precision):
        for w in range(a.warmup):
            r = solve_batch(cases, precision, tol=a.tol, n_iter=a.n_iter)
            print(f"# warmup {precision}: {r.wall:.2f} s, converged {int(r.converged.sum())}",
                  file=sys.stderr, flush=True)
        walls, convs, res = [], [], None
        for s in range(a.steps):
            # the last step keeps its final states (copied after the timed region)
            res = solve_batch(cases, precision, tol=a.tol, n_iter=a.n_iter, keep_state=s == a.steps - 1)
            uploads.append(res.upload_s)
            walls.append(res.wall)
            convs.append(int(res.converged.sum()))
            print(f"# {precision} step {s}: {res.wall:.2f} s, converged {convs[-1]}/{len(cases)}, "
                  f"Newton steps mean {res.iterations.mean():.1f}, Arnoldi/step {res.arnoldi.mean():.1f}",
                  file=sys.stderr, flush=True)
        return res, float(np.sum(walls)), int(np.sum(convs))

    r64, wall64, conv64 = run("fp64")
    value = conv64 / wall64
    lat = np.percentile(r64.t_done, [50, 99]) * 1e3
    line = {
        "metric": METRIC, "value": value, "unit": "cases/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * wall64 / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{a.cases} NACA 4-digit cases on {a.ni}x{a.nc} O-grids, FP64, "
                               f"solved to rn/r0 <= {a.tol:g}",
                   "cases": a.cases, "ni": a.ni, "nc": a.nc, "tol": a.tol, "seed": a.seed,
                   "gmres": "fixed 50 Arnoldi steps (reference semantics)"},
        "converged": conv64 // a.steps, "latency_ms": {"p50": float(lat[0]), "p99": float(lat[1])},
        "newton_steps_mean": float(r64.iterations.mean()),
        "upload_s_per_step": float(np.mean(uploads)),
        "value_pcie_inclusive": conv64 / (wall64 + float(np.sum(uploads[:a.steps]))),
    }
    if not a.no_mixed:
        rmx, wallmx, convmx = run("mixed")
        r64m = solve_batch(cases, "fp64", tol=a.tol, n_iter=a.n_iter, **mixed_path_fp64())
        lm = np.percentile(rmx.t_done, [50, 99]) * 1e3
        line["mixed"] = {
            "value": convmx / wallmx, "unit": "cases/s", "speedup_vs_fp64": (convmx / wallmx) / value,
            "converged": convmx // a.steps, "latency_ms": {"p50": float(lm[0]), "p99": float(lm[1])},
            "newton_steps_mean": float(rmx.iterations.mean()),
            "arnoldi_per_newton_mean": float(rmx.arnoldi.mean()),
            "gmres": "early exit at the preconditioned residual estimate <= 0.01 |M b|"}
        cf = compare_forms(r64, rmx, r64m, cases)
        for k in ("per_case_dcl_same_path", "per_case_dcl_reference_path"):
            cf.pop(k)
        line["mixed"].update(cf)
    if not a.no_kernel_timing:
        kt = time_kernels(cases, "fp64")
        line.update(rooflines(kt, a.cases * a.ni * a.nc))
        line["kernel_ms"] = kt
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(cases[0], float(r64.iterations.mean()))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
