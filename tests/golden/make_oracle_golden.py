"""Regenerate tests/golden/oracle_results_grid.npz and oracle_results_grid.json.

Runs the oracle on the reference's own Euler benchmark setup
(benchmarks/run_benchmarks.py:49-87 of the reference): naca4("0012", 65),
ogrid(nj=49, ds0=2e-3, radius=40), FP64, cases (M, alpha) = (0.5, 2), (0.8, 1.25),
(0.85, 1), one batch, 80 Newton iterations (the third case never meets tol=1e-6,
so all 80 run, as in the reference). Records per-case Cl/Cd/Cm, the residual
history and the final states of the two converged cases, which the CPU tests
check against the reference's published numbers (benchmarks/results.json:77-131).

Takes ~8 minutes on 8 cores:  python tests/golden/make_oracle_golden.py
"""
import json
import os
import sys
import time

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

from aero2d_amd.geometry import naca4  # noqa: E402
from aero2d_amd.grid import ogrid  # noqa: E402
from oracle import flow, newton  # noqa: E402


def main():
    torch.set_num_threads(len(os.sched_getaffinity(0)))
    g = ogrid(naca4("0012", 65), nj=49, ds0=2e-3, radius=40.0)
    geo = flow.case_geometry(g.x, g.y)
    machs = [0.5, 0.8, 0.85]
    alphas = [2.0, 1.25, 1.0]
    gb = flow.stack_geometry([geo] * len(machs))
    t = time.time()
    r = newton.solve(gb, machs, alphas, n_iter=80, tol=1e-6,
                     callback=lambda it, n, lr, cl: print(it, lr, time.time() - t, flush=True) and False)
    drop = np.log10(r.res_hist[-1] / r.res_hist[0])
    out = dict(machs=machs, alphas=alphas, iterations=r.iterations,
               cl=r.cl.tolist(), cd=r.cd.tolist(), cm=r.cm.tolist(),
               log10_res_drop=drop.tolist(), res_hist=r.res_hist.tolist(),
               grid=list(g.shape), seconds=time.time() - t)
    with open(os.path.join(HERE, "oracle_results_grid.json"), "w") as f:
        json.dump(out, f, indent=1)
    U = r.U.numpy()                     # (4, ni, nc, B)
    np.savez_compressed(os.path.join(HERE, "oracle_results_grid.npz"),
                        U_m05=U[..., 0], U_m08=U[..., 1])
    print(json.dumps({k: out[k] for k in ("cl", "cd", "cm", "log10_res_drop")}, indent=1))


if __name__ == "__main__":
    main()
