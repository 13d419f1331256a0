"""Benchmark: converged cases/s of the batched implicit Euler solve on one MI355X.

Workload (default): 64 NACA 4-digit cases on 256 x 128 O-grids, FP64, each
solved to a relative density residual rn/r0 <= 1e-10 — the uniform FP64 batch
whose throughput is the headline metric. Airfoils, Mach numbers and angles
are drawn from a fixed seed: m ~ U[0, 4]%, p ~ U[0.3, 0.5], t ~ U[9, 15]%,
M ~ U[0.3, 0.7], alpha ~ U[-4, 6] deg; grids ogrid(nj = 129, radius 20,
ds0 = 20 (q - 1)/(q^128 - 1), q = 1.1^(64/128)).

A step is one complete batched solve of all cases (warmup steps untimed).
The line also carries the mixed-precision form on the same batch, per-case
completion latency (p50/p99), a live HIP-event roofline of the residual
kernel and a CPU baseline (the FP64 oracle on the host cores, bounded sample).
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


def make_cases(n_cases, ni, nc, seed):
    from aero2d_amd import naca4_params, ogrid, self_similar_spacing
    from aero2d_amd.batch import Case
    rng = np.random.default_rng(seed)
    cases = []
    ds0 = self_similar_spacing(nc, 20.0)
    for k in range(n_cases):
        m = rng.uniform(0.0, 0.04)
        p = rng.uniform(0.3, 0.5)
        t = rng.uniform(0.09, 0.15)
        M = rng.uniform(0.3, 0.7)
        a = rng.uniform(-4.0, 6.0)
        g = ogrid(naca4_params(m, p, t, ni // 2 + 1), nj=nc + 1, ds0=ds0, radius=20.0)
        cases.append(Case(g, M, a, f"case{k}"))
    return cases


def time_kernels(cases, precision, reps=20):
    """HIP-event durations (ms) of the main kernels on the pool stream, after
    one Newton step of the batch so every buffer holds live data."""
    from aero2d_amd import _abi
    from aero2d_amd.batch import default_knobs
    prec = _abi.PREC_FP64 if precision == "fp64" else _abi.PREC_MIXED
    pool = _abi.Pool(0, default_knobs(precision))
    try:
        ids = np.array([pool.add_case(c.grid.x, c.grid.y, c.mach, c.alpha, prec) for c in cases], np.int32)
        out = np.empty((len(ids), 10))
        cfl = np.full(len(ids), 10.0)
        _abi.call("a2d_newton_step", pool.handle, _abi.iptr(ids), len(ids), _abi.dptr(cfl), _abi.dptr(out))
        names = ["residual", "jvp", "line_sweep", "cgs_dots", "assemble", "factor", "cgs_update"]
        res = {}
        for w, name in enumerate(names):
            ms = np.zeros(1)
            _abi.call("a2d_time_kernel", pool.handle, _abi.iptr(ids), len(ids), w, reps, _abi.dptr(ms))
            res[name] = float(ms[0])
    finally:
        pool.close()
    return res


def cpu_baseline(case, gpu_iters, steps=2):
    """FP64 oracle Newton steps of one case on the host cores."""
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
                       f"case by the oracle ({dt:.2f} s/step), scaled by the GPU run's mean "
                       f"{gpu_iters:.1f} Newton steps per case")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cases", type=int, default=64)
    ap.add_argument("--ni", type=int, default=256)
    ap.add_argument("--nc", type=int, default=128)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--n-iter", type=int, default=150)
    ap.add_argument("--seed", type=int, default=20261203 + 2)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kernel-timing", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {a.gpus} is not supported (cases are independent; "
              "a multi-GPU run would shard the batch across GPUs with no communication)", file=sys.stderr)
        sys.exit(2)

    from aero2d_amd.batch import solve_batch

    cases = make_cases(a.cases, a.ni, a.nc, a.seed)
    print(f"# {len(cases)} cases on {a.ni}x{a.nc} grids", file=sys.stderr, flush=True)

    def run(precision):
        for w in range(a.warmup):
            r = solve_batch(cases, precision, tol=a.tol, n_iter=a.n_iter)
            print(f"# warmup {precision}: {r.wall:.2f} s, converged {int(r.converged.sum())}",
                  file=sys.stderr, flush=True)
        walls, convs, res = [], [], None
        for s in range(a.steps):
            res = solve_batch(cases, precision, tol=a.tol, n_iter=a.n_iter)
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
    }
    if not a.no_mixed:
        rmx, wallmx, convmx = run("mixed")
        lm = np.percentile(rmx.t_done, [50, 99]) * 1e3
        line["mixed"] = {
            "value": convmx / wallmx, "unit": "cases/s", "speedup_vs_fp64": (convmx / wallmx) / value,
            "converged": convmx // a.steps, "latency_ms": {"p50": float(lm[0]), "p99": float(lm[1])},
            "max_abs_dcl": float(np.max(np.abs(rmx.cl - r64.cl))),
            "max_abs_dcd": float(np.max(np.abs(rmx.cd - r64.cd))),
            "max_abs_dcm": float(np.max(np.abs(rmx.cm - r64.cm))),
            "newton_steps_mean": float(rmx.iterations.mean()),
            "arnoldi_per_newton_mean": float(rmx.arnoldi.mean()),
            "gmres": "early exit at the preconditioned residual estimate <= 0.01 |M b|"}
    if not a.no_kernel_timing:
        kt = time_kernels(cases, "fp64")
        cells = a.cases * a.ni * a.nc
        ach = RES_BYTES_PER_CELL * cells / (kt["residual"] * 1e-3) / 1e9
        line["roofline"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": ach / HBM_PEAK_GBS, "traffic": None, "kernel": "residual (FP64)",
                            "bytes_per_launch": RES_BYTES_PER_CELL * cells, "ms_per_launch": kt["residual"]}
        line["kernel_ms"] = kt
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(cases[0], float(r64.iterations.mean()))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
