"""Regenerate tests/golden/oracle_config1.json: the FP64 oracle solve of the
single-case workload (NACA 0012, 128 x 64 cells, ratio 1.1 to 20 chords,
M 0.5, alpha 2 deg) to rn/r0 <= 1e-10 with the reference's host loop.
Takes ~10 minutes on 8 cores:  python tests/golden/make_config1_golden.py"""
import json
import os
import sys
import time

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

from aero2d_amd.workloads import config1  # noqa: E402
from oracle import flow, newton  # noqa: E402


def main():
    torch.set_num_threads(len(os.sched_getaffinity(0)))
    case = config1()[0]
    geo = flow.stack_geometry([flow.case_geometry(case.grid.x, case.grid.y)])
    t = time.time()
    r = newton.solve(geo, [case.mach], [case.alpha], n_iter=300, tol=1e-10,
                     callback=lambda it, n, lr, cl: print(it, lr, flush=True) and False)
    out = dict(mach=case.mach, alpha=case.alpha, grid=list(case.grid.x.shape), iterations=r.iterations,
               cl=float(r.cl[0]), cd=float(r.cd[0]), cm=float(r.cm[0]),
               rn_over_r0=float(r.res_hist[-1][0] / r.res_hist[0][0]), seconds=time.time() - t)
    with open(os.path.join(HERE, "oracle_config1.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out)


if __name__ == "__main__":
    main()
