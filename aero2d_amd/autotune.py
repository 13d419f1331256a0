"""Tile-shape autotuner for the stencil kernels.

The residual and tangent-linear kernels are compiled for several tile shapes
TI x TJ (a2d_stencil_shapes). For a grid size and precision the tuner builds a
few cases of that size, takes one Newton step so every buffer is live, then
times the tangent-linear kernel (the one launched once per Arnoldi step) for
each shape with HIP events on the pool's stream (a2d_time_kernel) and keeps
the fastest. Results are cached in JSON keyed by "ni x nc / precision" at
$AERO2D_TUNE_CACHE (default ~/.cache/aero2d_amd/tiles.json).
"""
from __future__ import annotations

import json
import os

import numpy as np

from . import _abi
from .batch import Case, default_knobs

__all__ = ["tune", "shapes_for", "cache_path"]


def cache_path() -> str:
    return os.environ.get("AERO2D_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "aero2d_amd", "tiles.json"))


def _load() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _key(ni, nc, precision):
    return f"{ni}x{nc}/{precision}"


def tune(case: Case, precision: str, copies: int = 8, reps: int = 20, save: bool = True) -> dict:
    """Time every compiled shape on `copies` replicas of `case`; returns
    {"shape": best index, "ms": {shape: ms}} and stores it in the cache."""
    prec = _abi.PREC_FP64 if precision == "fp64" else _abi.PREC_MIXED
    shapes = _abi.stencil_shapes()
    pool = _abi.Pool(0, default_knobs(precision))
    ni, nj = case.grid.x.shape
    try:
        ids = np.array([pool.add_case(case.grid.x, case.grid.y, case.mach, case.alpha, prec)
                        for _ in range(copies)], np.int32)
        out = np.empty((copies, 10))
        cfl = np.full(copies, 10.0)
        _abi.call("a2d_newton_step", pool.handle, _abi.iptr(ids), copies, _abi.dptr(cfl), _abi.dptr(out))
        timing = {}
        for s in range(len(shapes)):
            for cid in ids:
                _abi.call("a2d_case_set_shape", pool.handle, int(cid), s)
            ms = np.zeros(1)
            _abi.call("a2d_time_kernel", pool.handle, _abi.iptr(ids), copies, 1, reps, _abi.dptr(ms))
            timing[s] = float(ms[0])
    finally:
        pool.close()
    best = min(timing, key=timing.get)
    entry = {"shape": best, "tile": list(shapes[best]), "ms": {str(k): v for k, v in timing.items()}}
    if save:
        data = _load()
        data[_key(ni, nj - 1, precision)] = entry
        os.makedirs(os.path.dirname(cache_path()), exist_ok=True)
        with open(cache_path(), "w") as f:
            json.dump(data, f, indent=1)
    return entry


def shapes_for(cases, precision: str, tune_missing: bool = True) -> dict:
    """{(ni, nc): shape index} for the sizes present in `cases`, from the cache,
    tuning sizes that are missing when tune_missing."""
    data = _load()
    out = {}
    for c in cases:
        ni, nj = c.grid.x.shape
        key = (ni, nj - 1)
        if key in out:
            continue
        e = data.get(_key(ni, nj - 1, precision))
        if e is None and tune_missing:
            e = tune(c, precision)
        if e is not None:
            out[key] = int(e["shape"])
    return out
