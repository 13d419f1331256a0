"""ctypes binding of libaero2d_hip.so (the C ABI in include/aero2d_hip.h).

The library is built in-tree by __graft_entry__.build() (make in
aero2d_amd/csrc). There is no fallback: when the library is missing or no HIP
device is present, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libaero2d_hip.so")

PREC_FP64 = 0
PREC_MIXED = 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i64 = ctypes.c_int64


class RansParams(ctypes.Structure):
    """a2d_rans_params (include/aero2d_hip.h)."""
    _fields_ = [("mach", ctypes.c_double), ("re", ctypes.c_double), ("chord", ctypes.c_double),
                ("xref_x", ctypes.c_double), ("xref_y", ctypes.c_double), ("chi_inf", ctypes.c_double),
                ("limiter_k", ctypes.c_double), ("zmin", ctypes.c_double), ("order", ctypes.c_int),
                ("turbulent", ctypes.c_int), ("vortex_bc", ctypes.c_int), ("limiter_zone", ctypes.c_int),
                ("sa_raw_grad", ctypes.c_int), ("visc_raw_grad", ctypes.c_int)]


class Knobs(ctypes.Structure):
    """a2d_knobs (include/aero2d_hip.h)."""
    _fields_ = [("cfl_explicit", ctypes.c_double), ("lim_k", ctypes.c_double),
                ("max_update", ctypes.c_double), ("gmres_tol", ctypes.c_double),
                ("vortex", ctypes.c_int), ("order", ctypes.c_int),
                ("gmres_restart", ctypes.c_int), ("jacobi_sweeps", ctypes.c_int),
                ("early_exit", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("precond_order", ctypes.c_int), ("positivity", ctypes.c_int),
                ("precond_half", ctypes.c_int), ("basis_half", ctypes.c_int),
                ("stencil_march", ctypes.c_int), ("sweep_lds", ctypes.c_int),
                ("irs_eps", ctypes.c_double)]


# name -> (restype, argtypes); the exported surface of include/aero2d_hip.h
SIGNATURES = {
    "a2d_init": (ctypes.c_int, [ctypes.c_int]),
    "a2d_last_error": (ctypes.c_char_p, []),
    "a2d_version": (ctypes.c_int, []),
    "a2d_default_knobs": (None, [ctypes.POINTER(Knobs)]),
    "a2d_pool_create": (ctypes.c_int, [_i64, ctypes.POINTER(_i64)]),
    "a2d_pool_destroy": (ctypes.c_int, [_i64]),
    "a2d_pool_set_knobs": (ctypes.c_int, [_i64, ctypes.POINTER(Knobs)]),
    "a2d_pool_get_knobs": (ctypes.c_int, [_i64, ctypes.POINTER(Knobs)]),
    "a2d_pool_used_bytes": (_i64, [_i64]),
    "a2d_case_bytes": (_i64, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "a2d_case_add": (ctypes.c_int, [_i64, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_int, _ip]),
    "a2d_case_remove": (ctypes.c_int, [_i64, ctypes.c_int]),
    "a2d_case_set_freestream": (ctypes.c_int, [_i64, ctypes.c_int, ctypes.c_double, ctypes.c_double]),
    "a2d_case_set_step_limit": (ctypes.c_int, [_i64, ctypes.c_int, ctypes.c_double]),
    "a2d_case_set_state": (ctypes.c_int, [_i64, ctypes.c_int, _dp]),
    "a2d_case_get_state": (ctypes.c_int, [_i64, ctypes.c_int, _dp]),
    "a2d_case_set_shape": (ctypes.c_int, [_i64, ctypes.c_int, ctypes.c_int]),
    "a2d_case_park": (ctypes.c_int, [_i64, ctypes.c_int, ctypes.c_int]),
    "a2d_pool_stats": (ctypes.c_int, [_i64, ctypes.POINTER(_i64), ctypes.c_int]),
    "a2d_stencil_shapes": (ctypes.c_int, [_ip, _ip, ctypes.c_int]),
    "a2d_residual": (ctypes.c_int, [_i64, ctypes.c_int, ctypes.c_double, _dp, _dp]),
    "a2d_forces": (ctypes.c_int, [_i64, ctypes.c_int, _dp, _dp, _dp]),
    "a2d_local_dt": (ctypes.c_int, [_i64, ctypes.c_int, _dp]),
    "a2d_jvp": (ctypes.c_int, [_i64, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_int, _dp]),
    "a2d_assemble": (ctypes.c_int, [_i64, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, _dp, _dp]),
    "a2d_precond_setup": (ctypes.c_int, [_i64, ctypes.c_int, ctypes.c_double]),
    "a2d_precond_apply": (ctypes.c_int, [_i64, ctypes.c_int, _dp, _dp]),
    "a2d_newton_step": (ctypes.c_int, [_i64, _ip, ctypes.c_int, _dp, _dp]),
    "a2d_case_link_coarse": (ctypes.c_int, [_i64, ctypes.c_int, ctypes.c_int]),
    "a2d_explicit_cycles": (ctypes.c_int, [_i64, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
    "a2d_hllc_flux": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, _dp]),
    "a2d_hllc_flux_noalloc": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, _dp]),
    "a2d_time_kernel": (ctypes.c_int, [_i64, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
    "a2d_rans_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _ip, _dp,
                                       _ip, _dp, _ip, _dp, _dp, ctypes.POINTER(RansParams),
                                       ctypes.POINTER(_i64)]),
    "a2d_rans_destroy": (ctypes.c_int, [_i64]),
    "a2d_rans_set_state": (ctypes.c_int, [_i64, _dp]),
    "a2d_rans_get_state": (ctypes.c_int, [_i64, _dp]),
    "a2d_rans_freeze_limiter": (ctypes.c_int, [_i64, _dp]),
    "a2d_rans_limiter": (ctypes.c_int, [_i64, _dp]),
    "a2d_rans_residual": (ctypes.c_int, [_i64, ctypes.c_int, _dp, ctypes.c_double, _dp, _dp]),
    "a2d_rans_jvp": (ctypes.c_int, [_i64, ctypes.c_int, _dp, ctypes.c_double, _dp, _dp]),
    "a2d_rans_local_dt": (ctypes.c_int, [_i64, ctypes.c_double, _dp]),
    "a2d_rans_forces": (ctypes.c_int, [_i64, ctypes.c_double, _dp, _dp, _dp]),
    "a2d_rans_jac1": (ctypes.c_int, [_i64, _dp, ctypes.c_double, _dp, _dp]),
    "a2d_rans_ilu_setup": (ctypes.c_int, [_i64, ctypes.c_int, _ip, _ip, _ip, _ip]),
    "a2d_rans_ilu_info": (ctypes.c_int, [_i64, _ip]),
    "a2d_rans_ilu_factor": (ctypes.c_int, [_i64, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _ip]),
    "a2d_rans_ilu_apply": (ctypes.c_int, [_i64, _dp, _dp]),
    "a2d_rans_newton_step": (ctypes.c_int, [_i64, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, _dp]),
    "a2d_rans_newton_dq": (ctypes.c_int, [_i64, _dp]),
    "a2d_rans_set_cases": (ctypes.c_int, [_i64, ctypes.c_int]),
    "a2d_rans_select": (ctypes.c_int, [_i64, ctypes.c_int]),
    "a2d_rans_newton_steps": (ctypes.c_int, [_i64, ctypes.c_int, _ip, _ip, _dp, _dp, _dp, _dp, _dp, ctypes.c_int,
                                             _ip, ctypes.c_int, _ip, _dp]),
    "a2d_rans_time_kernel": (ctypes.c_int, [_i64, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_int,
                                            _dp]),
    "a2d_rans_keep": (ctypes.c_int, [_i64, ctypes.c_int]),
}

_lib = None
_init_lock = threading.Lock()
_device = None


class HipError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, name, code, msg):
        super().__init__(f"{name} failed ({code}): {msg}")
        self.code = code


class BudgetExceeded(HipError):
    """a2d_case_add refused a case: the pool's HBM budget would be exceeded."""


def load():
    """Load the shared library (no device work). Raises when it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def lib(device: int | None = None):
    """The library, initialised on a HIP device (default: the current one)."""
    global _device
    L = load()
    with _init_lock:
        if _device is None:
            dev = 0 if device is None else device
            call("a2d_init", dev)
            _device = dev
    return L


def call(name, *args):
    L = load()
    rc = getattr(L, name)(*args)
    if rc != 0:
        msg = L.a2d_last_error().decode(errors="replace")
        cls = BudgetExceeded if (name == "a2d_case_add" and rc == -2) else HipError
        raise cls(name, rc, msg)
    return rc


def dptr(a: np.ndarray):
    """ctypes double* of a C-contiguous float64 array."""
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


def default_knobs() -> Knobs:
    """The reference's solver defaults. AERO2D_USE_GRAPH=0 in the environment
    switches HIP graph replay off (every Newton step then launches its kernels
    directly; the kernels are the same)."""
    k = Knobs()
    load().a2d_default_knobs(ctypes.byref(k))
    if os.environ.get("AERO2D_USE_GRAPH", "1") == "0":
        k.use_graph = 0
    return k


def case_bytes(ni: int, nc: int, precision: int, restart: int = 50) -> int:
    """Device footprint of one case (a2d_case_bytes). Raises HipError (-6)
    for a size the library refuses."""
    b = int(load().a2d_case_bytes(ni, nc, precision, restart))
    if b < 0:
        raise HipError("a2d_case_bytes", b, f"a {ni}x{nc} case is outside the library's limits")
    return b


def stencil_shapes():
    ti = (ctypes.c_int * 16)()
    tj = (ctypes.c_int * 16)()
    n = load().a2d_stencil_shapes(ti, tj, 16)
    return [(ti[k], tj[k]) for k in range(n)]


class Pool:
    """Owner of a HIP stream and device-resident cases (a2d_pool_*)."""

    def __init__(self, hbm_budget_bytes: int = 0, knobs: Knobs | None = None):
        lib()
        h = _i64(0)
        call("a2d_pool_create", int(hbm_budget_bytes), ctypes.byref(h))
        self.handle = h.value
        if knobs is not None:
            self.set_knobs(knobs)

    def set_knobs(self, k: Knobs):
        call("a2d_pool_set_knobs", self.handle, ctypes.byref(k))

    def knobs(self) -> Knobs:
        k = Knobs()
        call("a2d_pool_get_knobs", self.handle, ctypes.byref(k))
        return k

    def used_bytes(self) -> int:
        return int(load().a2d_pool_used_bytes(self.handle))

    STAT_NAMES = ("steps", "graph_replays", "graph_captures", "plans", "plan_tables_allocated",
                  "slabs_allocated", "slabs_reused", "bytes_held")

    def stats(self) -> dict:
        """a2d_pool_stats as a dict (graph_replay_fraction added)."""
        out = (_i64 * len(self.STAT_NAMES))()
        call("a2d_pool_stats", self.handle, out, len(self.STAT_NAMES))
        d = {k: int(v) for k, v in zip(self.STAT_NAMES, out)}
        d["graph_replay_fraction"] = d["graph_replays"] / d["steps"] if d["steps"] else 0.0
        return d

    def park(self, cid: int, parked: bool = True):
        call("a2d_case_park", self.handle, int(cid), int(bool(parked)))

    def add_case(self, x: np.ndarray, y: np.ndarray, mach: float, alpha: float, precision: int) -> int:
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        ni, nj = x.shape
        cid = ctypes.c_int(-1)
        call("a2d_case_add", self.handle, ni, nj - 1, dptr(x), dptr(y), float(mach), float(alpha),
             int(precision), ctypes.byref(cid))
        return cid.value

    def remove_case(self, cid: int):
        call("a2d_case_remove", self.handle, cid)

    def link_coarse(self, cid: int, coarse: int):
        """a2d_case_link_coarse: `coarse` becomes the next coarser grid level of `cid`."""
        call("a2d_case_link_coarse", self.handle, int(cid), int(coarse))

    def explicit_cycles(self, ids, ncycles: int, fresh: bool) -> np.ndarray:
        """a2d_explicit_cycles: (n, 5) of res, cl, cd, cm, Gamma per case."""
        ids = np.ascontiguousarray(ids, np.int32)
        out = np.empty((len(ids), 5))
        call("a2d_explicit_cycles", self.handle, ids.ctypes.data_as(_ip), len(ids), int(ncycles), int(bool(fresh)),
             dptr(out))
        return out

    def close(self):
        if getattr(self, "handle", None) is not None:
            call("a2d_pool_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
