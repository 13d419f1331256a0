"""RANS (Spalart-Allmaras, SA-neg) Newton-Krylov solver on an unstructured mesh.

Host side of the HIP library's RANS core (the a2d_rans_* entry points of
include/aero2d_hip.h). The class keeps the reference's API and defaults for
the fully turbulent SA model (aero2d/rans.py: RansSolver, RansResult,
Convergence, RansControl): geometry set-up on the host (NumPy, as
RansSolver._setup), the block-ILU structure on the host (reverse Cuthill-McKee
order and fill pattern, as BlockILU0.__init__), and every field operation on
the device: residual, exact Jacobian-vector product, first-order Jacobian,
ILU factorisation and solves, GMRES, step limit and update, forces. The
pseudo-transient continuation loop (CFL growth, first-order start, rollback,
step rejection, convergence rule, limiter freezing, turbulence step boost)
follows RansSolver.solve (aero2d/rans.py:1298-1900).

Out of this build: k-omega SST, the AFT transition model, SA-RC, laminar
regions, the unsteady solve, the exact LU preconditioner and the gmsh mesh
generator (aero2d_amd.rans_mesh builds hybrid meshes from O-grids instead).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from .rans_mesh import Mesh

__all__ = ["RansSolver", "RansResult", "Convergence", "RansControl", "ilu_structure"]

GAM = 1.4


@dataclass
class RansResult:
    """aero2d/rans.py:48-62."""
    alpha: float
    cl: float
    cd: float
    cdp: float
    cdf: float
    cm: float
    converged: bool
    iterations: int
    residual: list = field(default_factory=list)
    elements: list = field(default_factory=list)
    state: object = None
    status: str = ""
    history: dict = field(default_factory=dict)


@dataclass
class Convergence:
    """Stop rule between angles (aero2d/rans.py:65-80): the last `hold`
    second-order iterations at most 10^-orders of the freestream residual with
    Cl and Cd steady; or a deep drop with forces steady for 3 iterations; or,
    with stall_ok, a stalled residual with steady forces."""
    orders: float = 4.5
    hold: int = 20
    cl_tol: float = 2e-4
    cd_tol: float = 2e-5
    stall_ok: bool = True


class RansControl:
    """Live control of a running solve (aero2d/rans.py:83-99): monitor(info)
    per iteration; skip / stop end the solve; cfl_factor scales the CFL once;
    rollback returns to the best state; cfl_set sets the CFL (held with
    cfl_hold)."""

    def __init__(self, monitor=None):
        self.monitor = monitor
        self.skip = False
        self.stop = False
        self.cfl_factor = 1.0
        self.rollback = False
        self.pause = False
        self.cfl_set = None
        self.cfl_hold = True


_CFL_FUNCS = {"min": min, "max": max, "exp": math.exp, "log": math.log, "log10": math.log10,
              "sqrt": math.sqrt, "abs": abs, "pow": pow, "tanh": math.tanh, "floor": math.floor,
              "ceil": math.ceil, "where": lambda c, a, b: a if c else b, "pi": math.pi, "e": math.e}
_CFL_FUNCS["\u0433\u0434\u0435"] = _CFL_FUNCS["where"]     # the reference's name of where(c, a, b)
_CFL_VARS = ("n", "orders", "res", "dres", "cfl")


def _cfl_ast(expr):
    import ast
    text = str(expr).replace("^", "**").strip()
    try:
        tree = ast.parse(text, mode="eval")
    except SyntaxError as err:
        raise ValueError(f"CFL formula: syntax error ({err.msg})") from None
    allowed = (ast.Expression, ast.BinOp, ast.UnaryOp, ast.BoolOp, ast.Compare, ast.IfExp, ast.Call,
               ast.Name, ast.Load, ast.Constant, ast.operator, ast.unaryop, ast.boolop, ast.cmpop)
    for node in ast.walk(tree):
        if not isinstance(node, allowed):
            raise ValueError(f"CFL formula: {type(node).__name__} is not allowed")
        if isinstance(node, ast.Name) and node.id not in _CFL_FUNCS and node.id not in _CFL_VARS:
            raise ValueError(f"CFL formula: unknown name '{node.id}' (allowed: {', '.join(_CFL_VARS)} and "
                             "min, max, exp, log, log10, sqrt, abs, pow, tanh, floor, ceil, where, pi, e)")
        if isinstance(node, ast.Call) and not (isinstance(node.func, ast.Name) and node.func.id in _CFL_FUNCS
                                               and not node.keywords):
            raise ValueError("CFL formula: only the listed functions may be called")
        if isinstance(node, ast.Constant) and not isinstance(node.value, (int, float)):
            raise ValueError("CFL formula: only numbers are allowed as constants")
    return tree


def cfl_formula(expr, **v) -> float:
    """The CFL number given by a user expression (aero2d/rans.py:102-123): the
    variables n (iteration at this angle), orders (decades the residual has
    fallen since the first iteration), res (residual / first residual), dres
    (residual / previous residual) and cfl (the current CFL); the functions
    min, max, exp, log, log10, sqrt, abs, pow, tanh, floor, ceil and
    where(c, a, b) (also under the reference's name), the constants pi and e;
    '^' is a power. The expression is parsed and checked node by node (numbers,
    arithmetic, comparisons, the listed names and calls only). Raises
    ValueError on a bad expression or a value that is not a finite positive
    number."""
    tree = _cfl_ast(expr)
    names = dict(_CFL_FUNCS)
    names.update({k: v[k] for k in _CFL_VARS if k in v})
    try:
        val = float(eval(compile(tree, "<CFL>", "eval"), {"__builtins__": {}}, names))
    except (ArithmeticError, ValueError, TypeError, NameError) as err:
        raise ValueError(f"CFL formula: cannot evaluate ({err})") from None
    if not math.isfinite(val) or val <= 0:
        raise ValueError(f"CFL formula gave {val:g}: a positive number is needed")
    return val


def cfl_formula_recursive(expr) -> bool:
    """Whether the formula reads the current CFL or the residual ratio of
    neighbouring iterations (aero2d/rans.py:126-132)."""
    import ast
    try:
        tree = _cfl_ast(expr)
    except ValueError:
        return False
    return any(isinstance(nd, ast.Name) and nd.id in ("cfl", "dres") for nd in ast.walk(tree))


def ilu_structure(n, rows, cols, fill=0, perm=None):
    """Block ILU(k) structure as BlockILU0.__init__ forms it (aero2d/bilu.py:14-59):
    perm (reverse Cuthill-McKee of the symmetrised pattern unless given,
    perm[new] = old), the pattern in the new numbering grown `fill` times by
    P <- P + tril(P, -1) triu(P, 1), sorted by row then column, and the
    position of every given block in it. Returns (perm, r, c, src_pos)."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    rows, cols = np.asarray(rows), np.asarray(cols)
    if perm is None:
        G = sp.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(n, n))
        perm = reverse_cuthill_mckee((G + G.T).tocsr(), symmetric_mode=True)
    perm = np.asarray(perm)
    inv = np.empty(n, int)
    inv[perm] = np.arange(n)
    r0, c0 = inv[rows], inv[cols]
    P = sp.csr_matrix((np.ones(len(r0)), (r0, c0)), shape=(n, n))
    for _ in range(int(fill)):
        P = (P + sp.tril(P, -1).tocsr() @ sp.triu(P, 1).tocsr()).tocsr()
    P = P.tocoo()
    r, c = P.row.astype(np.int64), P.col.astype(np.int64)
    o = np.lexsort((c, r))
    r, c = r[o], c[o]
    key = r * n + c
    src_pos = np.searchsorted(key, r0.astype(np.int64) * n + c0)
    return perm, r, c, src_pos


def transfer_state(mesh_from: Mesh, Q, mesh_to: Mesh) -> np.ndarray:
    """A state (nc, 5) of mesh_from carried to mesh_to (aero2d/rans.py:2438-2454):
    each new cell takes the values of the nearest old cell centroid; where the
    new cell is closer to a wall than that old cell, its velocity and rho chi
    are scaled by the ratio of the wall distances (capped at 1) and the energy
    loses the removed kinetic energy, so density and pressure carry over."""
    from scipy.spatial import cKDTree
    Q = np.asarray(Q, np.float64)
    _, near = cKDTree(mesh_from.centroid).query(mesh_to.centroid)
    out = Q[near].copy()
    scale = np.clip(mesh_to.wall_dist / np.maximum(mesh_from.wall_dist[near], 1e-30), 0.0, 1.0)
    kinetic = 0.5 * (out[:, 1] ** 2 + out[:, 2] ** 2) / out[:, 0]
    out[:, 1] *= scale
    out[:, 2] *= scale
    out[:, 3] -= kinetic * (1.0 - scale ** 2)
    out[:, 4] *= scale
    return out


class RansSolver:
    """SA RANS on a Mesh (aero2d/rans.py:187-326 for the constructor and the
    defaults kept here)."""

    def __init__(self, mesh: Mesh, mach: float, re: float, chord: float = 1.0, xref=(0.25, 0.0),
                 chi_inf: float = 3.0, order: int = 2, vortex_bc: bool = True, turbulent: bool = True):
        self.m = mesh
        self.mach, self.re, self.chord = float(mach), float(re), float(chord)
        self.mu_inf = self.mach * chord / self.re
        self.nu_inf = self.mu_inf
        self.xref = tuple(xref)
        self.chi_inf = chi_inf if turbulent else 0.0
        self.order = order
        self.vortex_bc = vortex_bc
        self.turbulent = turbulent
        self.limiter_k = 5.0
        self.cfl_reset_on_order = True
        self.limiter_zone = "te"
        self.limiter_te_radius = 0.05
        self.cfl_mode = "grow"
        self.cfl_growth = 1.25
        self.cfl_formula = "min(5*1.2**n, 1000)"
        self.cfl_turb = None
        self.turb_boost = 300.0
        self.turb_boost_cfl = 10.0
        self.step_limit_cells = 100
        self.dt_visc_c = 1.0
        self.conv = Convergence()
        self.cfl_start = 5.0
        self.cfl_limit = 1e6
        self.first_order = 30
        self.rollback_factor = 100.0
        self.rollback_auto = True
        self.cfl_lin_back = 0.7
        self.rollback_first_order = 15
        self.step_reject = 3.0
        self.max_rejects = 6
        self.gmres_tol = 0.05
        self.max_rollbacks = 4
        self.freeze_limiter = True
        self.prec_fill = 3
        self.prec_lag = 1
        self.sa_raw_grad = False
        self.visc_raw_grad = False
        self.right_prec = False
        self.zmin = 0.0
        self.gamma_vortex = 0.0
        self.nv = 5
        self._h = None
        self._ilu_fill = None
        self._slot = 0
        self._setup()

    # ------------------------------------------------------------------ set-up
    def _setup(self):
        """Face sets, least-squares matrices and the limiter zone, as
        RansSolver._setup (aero2d/rans.py:328-425)."""
        m = self.m
        nc = m.n_cells
        self.nc = nc
        fi = np.where(m.face_tag == 0)[0]
        fw = np.where(m.face_tag == 1)[0]
        ff = np.where(m.face_tag == 2)[0]
        L, R = m.face_cells[fi, 0], m.face_cells[fi, 1]
        xc = m.centroid
        self.iL, self.iR = L, R
        self.n_i, self.A_i = m.face_normal[fi], m.face_area[fi]
        self.rL = m.face_center[fi] - xc[L]
        self.rR = m.face_center[fi] - xc[R]
        d = xc[R] - xc[L]
        self.d_i = d
        self.dist_i = np.hypot(d[:, 0], d[:, 1])
        cw = m.face_cells[fw, 0]
        self.cw = cw
        self.n_w, self.A_w = m.face_normal[fw], m.face_area[fw]
        self.r_w = m.face_center[fw] - xc[cw]
        self.dn_w = np.maximum(np.einsum("ij,ij->i", self.r_w, self.n_w), 1e-14)
        self.body_w = m.face_body[fw]
        self.fc_w = m.face_center[fw]
        cf = m.face_cells[ff, 0]
        self.cf = cf
        self.n_f, self.A_f = m.face_normal[ff], m.face_area[ff]
        self.r_f = m.face_center[ff] - xc[cf]
        self.fc_f = m.face_center[ff]
        w_i = 1.0 / self.dist_i**2
        Mxx, Mxy, Myy = np.zeros(nc), np.zeros(nc), np.zeros(nc)
        for idx in (L, R):
            np.add.at(Mxx, idx, w_i * d[:, 0] ** 2)
            np.add.at(Mxy, idx, w_i * d[:, 0] * d[:, 1])
            np.add.at(Myy, idx, w_i * d[:, 1] ** 2)
        dw = self.r_w
        ww = 1.0 / np.maximum(np.sum(dw**2, 1), 1e-30)
        np.add.at(Mxx, cw, ww * dw[:, 0] ** 2)
        np.add.at(Mxy, cw, ww * dw[:, 0] * dw[:, 1])
        np.add.at(Myy, cw, ww * dw[:, 1] ** 2)
        det = Mxx * Myy - Mxy**2
        det = np.where(np.abs(det) < 1e-300, 1e-300, det)
        self.Minv = np.stack([np.stack([Myy / det, -Mxy / det], -1),
                              np.stack([-Mxy / det, Mxx / det], -1)], -2)
        self.w_i, self.w_w = w_i, ww
        self.vol = m.volume
        self.wall_d = np.maximum(m.wall_dist, 1e-12)
        near = np.zeros(nc, bool)
        for b in np.unique(self.body_w):
            pts = self.fc_w[self.body_w == b]
            te = pts[np.argmax(pts[:, 0])]
            near |= np.hypot(xc[:, 0] - te[0], xc[:, 1] - te[1]) < self.limiter_te_radius * self.chord
        self.lim_te = near

    def _device(self):
        """The library's solver for this mesh and these constants (created once)."""
        if self._h is not None:
            return self._h
        _abi.lib()
        nc = self.nc
        fgeo = np.column_stack([self.n_i, self.A_i, self.rL, self.rR, self.d_i, self.dist_i, self.w_i])
        wgeo = np.column_stack([self.n_w, self.A_w, self.r_w, self.dn_w, self.w_w, self.fc_w])
        ffgeo = np.column_stack([self.n_f, self.A_f, self.r_f, self.fc_f])
        cgeo = np.column_stack([self.vol, self.wall_d, self.Minv[:, 0, 0], self.Minv[:, 0, 1],
                                self.Minv[:, 1, 1], self.lim_te.astype(float)])
        f64 = lambda a: np.ascontiguousarray(a, np.float64)   # noqa: E731
        i32 = lambda a: np.ascontiguousarray(a, np.int32)     # noqa: E731
        zone = {"off": 0, "te": 1, "all": 2}[self.limiter_zone]
        prm = _abi.RansParams(self.mach, self.re, self.chord, self.xref[0], self.xref[1], self.chi_inf,
                              self.limiter_k, self.zmin, int(self.order), int(bool(self.turbulent)),
                              int(bool(self.vortex_bc)), zone, int(bool(self.sa_raw_grad)),
                              int(bool(self.visc_raw_grad)))
        arrs = [i32(self.iL), i32(self.iR), f64(fgeo), i32(self.cw), f64(wgeo), i32(self.cf), f64(ffgeo),
                f64(cgeo)]
        self._keep = arrs
        h = _abi._i64(0)
        _abi.call("a2d_rans_create", nc, len(self.iL), len(self.cw), len(self.cf), _abi.iptr(arrs[0]),
                  _abi.iptr(arrs[1]), _abi.dptr(arrs[2]), _abi.iptr(arrs[3]), _abi.dptr(arrs[4]),
                  _abi.iptr(arrs[5]), _abi.dptr(arrs[6]), _abi.dptr(arrs[7]), ctypes.byref(prm), ctypes.byref(h))
        self._h = h.value
        return self._h

    def close(self):
        if self._h is not None:
            try:
                _abi.call("a2d_rans_destroy", self._h)
            finally:
                self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pattern(self):
        """Block rows/cols of the first-order Jacobian: the diagonal, then
        (iL, iR), then (iR, iL) (aero2d/rans.py:1128-1144)."""
        nc = self.nc
        return (np.concatenate([np.arange(nc), self.iL, self.iR]),
                np.concatenate([np.arange(nc), self.iR, self.iL]))

    def _ensure_ilu(self):
        if self._ilu_fill == int(self.prec_fill):
            return
        if self._ilu_fill is not None:
            raise RuntimeError("the ILU fill level of a solver is fixed once set")
        h = self._device()
        rows, cols = self.pattern()
        perm, r, c, pos = ilu_structure(self.nc, rows, cols, fill=int(self.prec_fill))
        i32 = lambda a: np.ascontiguousarray(a, np.int32)     # noqa: E731
        a = [i32(perm), i32(r), i32(c), i32(pos)]
        _abi.call("a2d_rans_ilu_setup", h, len(r), *[_abi.iptr(x) for x in a])
        self._ilu_fill = int(self.prec_fill)
        self.ilu_perm = perm

    def ilu_info(self) -> dict:
        self._ensure_ilu()
        out = np.zeros(4, np.int32)
        _abi.call("a2d_rans_ilu_info", self._device(), _abi.iptr(out))
        return dict(nnz=int(out[0]), levels=int(out[1]), levels_back=int(out[2]), factor_waves=int(out[3]))

    # ---------------------------------------------------------------- physics
    def prim(self, Q):
        """(rho, rho u, rho v, rho E, rho chi) -> (rho, u, v, p, chi) (aero2d/rans.py:455-461)."""
        Q = np.asarray(Q, float)
        r = Q[:, 0]
        u, v = Q[:, 1] / r, Q[:, 2] / r
        p = (GAM - 1) * (Q[:, 3] - 0.5 * r * (u**2 + v**2))
        return np.concatenate([np.stack([r, u, v, p], 1), Q[:, 4:] / r[:, None]], 1)

    def cons(self, w):
        """Inverse of prim (aero2d/rans.py:463-467)."""
        w = np.asarray(w, float)
        r, u, v, p = (w[:, k] for k in range(4))
        return np.concatenate([np.stack([r, r * u, r * v, p / (GAM - 1) + 0.5 * r * (u**2 + v**2)], 1),
                               w[:, 4:] * r[:, None]], 1)

    def freestream(self, alpha):
        """(1, M cos a, M sin a, 1/gamma, chi_inf) (aero2d/rans.py:469-475)."""
        a = np.radians(alpha)
        return np.array([1.0, self.mach * np.cos(a), self.mach * np.sin(a), 1.0 / GAM, self.chi_inf])

    # ------------------------------------------------------------- device calls
    def _set(self, Q):
        Q = np.ascontiguousarray(Q, np.float64).reshape(self.nc, 5)
        _abi.call("a2d_rans_set_state", self._device(), _abi.dptr(Q))

    def _get(self):
        Q = np.empty((self.nc, 5))
        _abi.call("a2d_rans_get_state", self._device(), _abi.dptr(Q))
        return Q

    def _freeze(self, phi):
        h = self._device()
        if phi is None:
            _abi.call("a2d_rans_freeze_limiter", h, None)
        else:
            p = np.ascontiguousarray(phi, np.float64).reshape(self.nc, 5)
            _abi.call("a2d_rans_freeze_limiter", h, _abi.dptr(p))

    def residual(self, Q, winf, gam=0.0, order=None, phi_frozen=None):
        """R (nc, 5) (aero2d/rans.py:860-964)."""
        order = self.order if order is None else int(order)
        self._set(Q)
        self._freeze(phi_frozen)
        R = np.empty((self.nc, 5))
        w = np.ascontiguousarray(winf, np.float64)
        _abi.call("a2d_rans_residual", self._device(), order, _abi.dptr(w), float(gam), _abi.dptr(R), None)
        return R

    def jvp(self, Q, v, winf, gam=0.0, order=None, phi_frozen=None):
        """jax.jvp of the residual along v (the jvp entries of _fns, aero2d/rans.py:1186-1191)."""
        order = self.order if order is None else int(order)
        self._set(Q)
        self._freeze(phi_frozen)
        out = np.empty((self.nc, 5))
        v = np.ascontiguousarray(v, np.float64).reshape(self.nc, 5)
        w = np.ascontiguousarray(winf, np.float64)
        _abi.call("a2d_rans_jvp", self._device(), order, _abi.dptr(w), float(gam), _abi.dptr(v), _abi.dptr(out))
        return out

    def phi(self, Q):
        """The Venkatakrishnan limiter of a state in every cell (aero2d/rans.py:1192-1194)."""
        self._set(Q)
        out = np.empty((self.nc, 5))
        _abi.call("a2d_rans_limiter", self._device(), _abi.dptr(out))
        return out

    def local_dt(self, Q, cfl):
        """Local pseudo-time step (aero2d/rans.py:1206-1249) for the flow
        equations; the turbulence equation's is this times the step factor."""
        self._set(Q)
        dt = np.empty(self.nc)
        _abi.call("a2d_rans_local_dt", self._device(), float(cfl), _abi.dptr(dt))
        return dt

    def _forces_dev(self, alpha):
        out = np.zeros(5)
        nw = len(self.cw)
        cp, cf = np.empty(nw), np.empty(nw)
        _abi.call("a2d_rans_forces", self._device(), float(alpha), _abi.dptr(out), _abi.dptr(cp), _abi.dptr(cf))
        return out, cp, cf

    def forces(self, Q, alpha):
        """Pressure and friction forces (aero2d/rans.py:1251-1296)."""
        self._set(Q)
        return self._forces_dict(alpha)

    def _forces_dict(self, alpha):
        out, cp, cf = self._forces_dev(alpha)
        q = 0.5 * self.mach**2
        per = []
        a = np.radians(alpha)
        for b in range(int(self.body_w.max()) + 1):
            k = self.body_w == b
            # per-body forces from the per-face coefficients
            n, A = self.n_w[k], self.A_w[k]
            fxp = cp[k] * q * n[:, 0] * A
            fyp = cp[k] * q * n[:, 1] * A
            fxv = cf[k] * q * (-n[:, 1]) * A
            fyv = cf[k] * q * n[:, 0] * A
            per.append(dict(x=self.fc_w[k, 0], y=self.fc_w[k, 1], cp=cp[k], cf=cf[k], gam=None,
                            cl=float(((-(fxp + fxv) * np.sin(a) + (fyp + fyv) * np.cos(a))).sum() / (q * self.chord)),
                            cd=float((((fxp + fxv) * np.cos(a) + (fyp + fyv) * np.sin(a))).sum() / (q * self.chord))))
        return dict(cl=float(out[0]), cd=float(out[1]), cdp=float(out[2]), cdf=float(out[3]), cm=float(out[4]),
                    cp=cp, cf=cf, elements=per)

    def jac1_blocks(self, Q, winf, gam):
        """Diagonal blocks (nc,5,5) without V/dt and the off-diagonal blocks of
        the pairs (iL,iR) then (iR,iL) (aero2d/rans.py:1107-1126)."""
        self._set(Q)
        D = np.empty((self.nc, 5, 5))
        off = np.empty((2 * len(self.iL), 5, 5))
        w = np.ascontiguousarray(winf, np.float64)
        _abi.call("a2d_rans_jac1", self._device(), _abi.dptr(w), float(gam), _abi.dptr(D), _abi.dptr(off))
        return D, off

    def ilu_factor(self, Q, winf, gam, cfl, turb_factor=1.0):
        """Factor vol/dt + J1 at the state (BlockILU0.factor); returns the count
        of non-finite inverse-diagonal entries."""
        self._ensure_ilu()
        self._set(Q)
        bad = ctypes.c_int(0)
        w = np.ascontiguousarray(winf, np.float64)
        _abi.call("a2d_rans_ilu_factor", self._device(), _abi.dptr(w), float(gam), float(cfl), float(turb_factor),
                  ctypes.byref(bad))
        return bad.value

    def ilu_apply(self, b):
        """BlockILU0.solve of b (nc, 5) with the last factorisation."""
        b = np.ascontiguousarray(b, np.float64).reshape(self.nc, 5)
        x = np.empty((self.nc, 5))
        _abi.call("a2d_rans_ilu_apply", self._device(), _abi.dptr(b), _abi.dptr(x))
        return x

    STEP_FIELDS = ("lin", "gmres_iterations", "info", "lam", "rel", "ilu_nonfinite", "applied", "nbig",
                   "per_cell", "factored")
    _INT_FIELDS = ("gmres_iterations", "info", "ilu_nonfinite", "applied", "nbig", "per_cell", "factored")

    def newton_step(self, order, winf, gam, cfl, gmres_tol, gmres_iter=60, turb_factor=1.0, right=None,
                    Q=None, reuse_ilu=False):
        """One Newton step at the device state (or at Q): the linear solve and
        update of RansSolver.solve (aero2d/rans.py:1738-1858). reuse_ilu keeps
        the slot's last factorisation as the preconditioner (prec_lag,
        aero2d/rans.py:1736-1752) when it is a Newton step's finite one.
        Returns a dict of the step's figures ("factored": whether it factored)."""
        self._ensure_ilu()
        if Q is not None:
            self._set(Q)
        out = np.zeros(12)
        w = np.ascontiguousarray(winf, np.float64)
        r = self.right_prec if right is None else right
        _abi.call("a2d_rans_newton_step", self._device(), int(order), _abi.dptr(w), float(gam), float(cfl),
                  float(turb_factor), float(gmres_tol), int(gmres_iter), int(bool(r)), int(self.step_limit_cells),
                  int(bool(reuse_ilu)), _abi.dptr(out))
        d = {k: float(v) for k, v in zip(self.STEP_FIELDS, out)}
        for k in self._INT_FIELDS:
            d[k] = int(d[k])
        return d

    KERNELS = ("ilu_solve", "ilu_factor", "matvec", "residual", "jacobian")

    def time_kernel(self, which, alpha=0.0, order=None, gam=0.0, reps=20):
        """Mean device time in ms of one of KERNELS at the device state
        (a2d_rans_time_kernel)."""
        self._ensure_ilu()
        k = self.KERNELS.index(which)
        ms = ctypes.c_double(0.0)
        w = np.ascontiguousarray(self.freestream(alpha), np.float64)
        _abi.call("a2d_rans_time_kernel", self._device(), k, int(self.order if order is None else order),
                  _abi.dptr(w), float(gam), int(reps), ctypes.byref(ms))
        return ms.value

    def newton_dq(self):
        dq = np.empty((self.nc, 5))
        _abi.call("a2d_rans_newton_dq", self._device(), _abi.dptr(dq))
        return dq

    def _rn(self, order, winf, gam):
        rn5 = np.zeros(5)
        w = np.ascontiguousarray(winf, np.float64)
        _abi.call("a2d_rans_residual", self._device(), int(order), _abi.dptr(w), float(gam), None, _abi.dptr(rn5))
        return rn5

    def _keep_state(self, what):
        _abi.call("a2d_rans_keep", self._device(), int(what))

    # ------------------------------------------------------------------ solve
    def initial_state(self, winf):
        """Freestream with the velocity faded to zero at the wall over 0.02 chord
        (aero2d/rans.py:1380-1385)."""
        f = np.clip(self.wall_d / 0.02, 0, 1)[:, None]
        w0 = np.tile(winf, (self.nc, 1))
        w0[:, 1:3] *= f
        return self.cons(w0)

    def solve(self, alpha, Q0=None, max_iter=300, tol=1e-7, cfl0=None, cfl_max=None, first_order_iters=None,
              verbose=False, callback=None, gmres_tol=None, gmres_iter=60, conv=None, control=None,
              alpha0=None) -> RansResult:
        """Pseudo-transient Newton-Krylov solve to the Convergence rule or
        max_iter iterations (aero2d/rans.py:1298-1900, the SA path)."""
        run = _Run(self, 0, alpha, Q0, max_iter, tol, cfl0, cfl_max, first_order_iters, verbose, callback,
                   gmres_tol, gmres_iter, conv, control, alpha0)
        return self._drive([run])[0]

    _SETTINGS = ("limiter_k", "cfl_reset_on_order", "limiter_zone", "limiter_te_radius", "cfl_mode",
                 "cfl_growth", "cfl_formula", "cfl_turb", "turb_boost", "turb_boost_cfl", "step_limit_cells",
                 "dt_visc_c", "conv", "cfl_start", "cfl_limit", "first_order", "rollback_factor",
                 "rollback_auto", "cfl_lin_back", "rollback_first_order", "step_reject", "max_rejects",
                 "gmres_tol", "max_rollbacks", "freeze_limiter", "prec_fill", "prec_lag", "right_prec", "zmin",
                 "sa_raw_grad", "visc_raw_grad")

    def clone(self) -> "RansSolver":
        """A new solver on the same mesh with the same constants and settings
        (its own device memory and stream)."""
        c = RansSolver(self.m, self.mach, self.re, self.chord, self.xref, self.chi_inf, self.order,
                       self.vortex_bc, self.turbulent)
        for k in self._SETTINGS:
            setattr(c, k, getattr(self, k))
        return c

    def solve_polar(self, alphas, Q0=None, max_iter=300, tol=1e-7, cfl0=None, cfl_max=None,
                    first_order_iters=None, verbose=False, gmres_tol=None, gmres_iter=60, conv=None, alpha0=None,
                    workers=1):
        """Every angle of a polar solved at once, one case slot per angle, each
        with the loop of `solve` from its own start; their Newton steps run as
        one batched device call per iteration (the factorisation and
        triangular sweeps of all angles in single launches). The reference runs
        a polar's angles in worker processes (rans_polar, aero2d/rans.py:2619);
        each result here equals that angle's `solve` alone. Q0: None or one
        state per angle; alpha0: the angle each Q0 was solved at (a number or
        one per angle), so its far field is turned to the new angle as
        `solve(..., alpha0=)` does. workers > 1: the angles are dealt round
        robin to that many solvers (clones of this one, each with its own
        stream), driven from as many host threads, so their launches run side
        by side on the device and each lockstep waits only for its own angles.
        Returns the RansResults in the order of alphas."""
        alphas = [float(a) for a in alphas]
        workers = max(1, min(int(workers), len(alphas)))
        if workers > 1:
            from concurrent.futures import ThreadPoolExecutor
            a0s = [alpha0] * len(alphas) if alpha0 is None or np.isscalar(alpha0) else list(alpha0)
            parts = [list(range(w, len(alphas), workers)) for w in range(workers)]
            solvers = [self] + [self.clone() for _ in range(workers - 1)]

            def one(w):
                idx = parts[w]
                return solvers[w].solve_polar([alphas[k] for k in idx], None if Q0 is None else [Q0[k] for k in idx],
                                              max_iter, tol, cfl0, cfl_max, first_order_iters, verbose, gmres_tol,
                                              gmres_iter, conv, [a0s[k] for k in idx])
            try:
                with ThreadPoolExecutor(max_workers=workers) as ex:
                    outs = list(ex.map(one, range(workers)))
            finally:
                for sv in solvers[1:]:
                    sv.close()
            res = [None] * len(alphas)
            for w, idx in enumerate(parts):
                for k, r in zip(idx, outs[w]):
                    res[k] = r
            return res
        self._ensure_ilu()
        _abi.call("a2d_rans_set_cases", self._device(), len(alphas))
        a0 = [alpha0] * len(alphas) if alpha0 is None or np.isscalar(alpha0) else list(alpha0)
        runs = [_Run(self, k, a, None if Q0 is None else Q0[k], max_iter, tol, cfl0, cfl_max, first_order_iters,
                     verbose, None, gmres_tol, gmres_iter, conv, None, a0[k]) for k, a in enumerate(alphas)]
        return self._drive(runs)

    def _select(self, slot):
        if self._slot != slot:
            _abi.call("a2d_rans_select", self._device(), int(slot))
            self._slot = slot

    def _drive(self, runs):
        """Advance the runs in lockstep: each run's iteration up to its Newton
        step on the host, then one batched Newton step for all that asked."""
        self._ensure_ilu()
        for r in runs:
            self._select(r.slot)
            r.start()
        while True:
            reqs = []
            for r in runs:
                if r.done:
                    continue
                self._select(r.slot)
                q = r.plan()
                if q is not None:
                    reqs.append((r, q))
            if not reqs:
                if all(r.done for r in runs):
                    break
                continue
            outs = self._newton_steps([q for _, q in reqs])
            for (r, _), st in zip(reqs, outs):
                self._select(r.slot)
                r.after(st)
        res = []
        for r in runs:
            self._select(r.slot)
            res.append(r.finish())
        self._select(0)
        return res

    def _newton_steps(self, qs):
        """a2d_rans_newton_steps over the requests (one per slot)."""
        n = len(qs)
        i32 = lambda a: np.ascontiguousarray(a, np.int32)     # noqa: E731
        f64 = lambda a: np.ascontiguousarray(a, np.float64)   # noqa: E731
        ids = i32([q["slot"] for q in qs])
        order = i32([q["order"] for q in qs])
        winf = f64(np.stack([q["winf"] for q in qs]))
        gam = f64([q["gam"] for q in qs])
        cfl = f64([q["cfl"] for q in qs])
        ct = f64([q["ct"] for q in qs])
        gtol = f64([q["gtol"] for q in qs])
        right = i32([int(bool(q["right"])) for q in qs])
        reuse = i32([int(bool(q.get("reuse", False))) for q in qs])
        out = np.zeros((n, 12))
        gi = int(qs[0]["gmres_iter"])
        if any(int(q["gmres_iter"]) != gi for q in qs):
            raise ValueError("a batch of Newton steps shares one GMRES size")
        _abi.call("a2d_rans_newton_steps", self._device(), n, _abi.iptr(ids), _abi.iptr(order), _abi.dptr(winf),
                  _abi.dptr(gam), _abi.dptr(cfl), _abi.dptr(ct), _abi.dptr(gtol), gi, _abi.iptr(right),
                  int(self.step_limit_cells), _abi.iptr(reuse), _abi.dptr(out))
        res = []
        for row in out:
            d = {k: float(v) for k, v in zip(self.STEP_FIELDS, row)}
            for k in self._INT_FIELDS:
                d[k] = int(d[k])
            res.append(d)
        return res

    def _rotate_state(self, Q, dalpha, inner=0.05, outer=0.5):
        """Velocity turned by dalpha degrees, fully beyond `outer` chords from the
        wall, fading to none at `inner` (aero2d/rans.py:1146-1158)."""
        f = np.clip((self.wall_d / self.chord - inner) / max(outer - inner, 1e-9), 0.0, 1.0)
        th = np.radians(dalpha) * f
        c, s_ = np.cos(th), np.sin(th)
        Q = Q.copy()
        mu, mv = Q[:, 1].copy(), Q[:, 2].copy()
        Q[:, 1] = c * mu - s_ * mv
        Q[:, 2] = s_ * mu + c * mv
        return Q


class _Run:
    """One case's pseudo-transient continuation loop (aero2d/rans.py:1298-1900,
    the SA path), split at its Newton step so that several cases can share one
    batched device step: plan() runs an iteration up to the step and returns the
    step's parameters (None when the iteration ends before a step: a rollback,
    an undone step, or the end of the loop), after() takes the step's figures."""

    def __init__(self, S, slot, alpha, Q0, max_iter, tol, cfl0, cfl_max, first_order_iters, verbose, callback,
                 gmres_tol, gmres_iter, conv, control, alpha0):
        self.S, self.slot, self.alpha = S, slot, float(alpha)
        self.Q0, self.max_iter, self.tol = Q0, max_iter, tol
        self.crit = conv or S.conv
        self.gmres_tol = S.gmres_tol if gmres_tol is None else gmres_tol
        self.gmres_iter = gmres_iter
        self.cfl0 = S.cfl_start if cfl0 is None else cfl0
        self.cfl_max = S.cfl_limit if cfl_max is None else cfl_max
        self.first_order_iters = S.first_order if first_order_iters is None else first_order_iters
        self.verbose, self.callback, self.control, self.alpha0 = verbose, callback, control, alpha0
        self.done = False

    def start(self):
        S = self.S
        self.winf = S.freestream(self.alpha)
        if self.Q0 is None:
            Q = S.initial_state(self.winf)
        else:
            Q = np.array(self.Q0, float)
            if self.alpha0 is not None and abs(self.alpha - float(self.alpha0)) > 1e-12:
                Q = S._rotate_state(Q, self.alpha - float(self.alpha0))
        # reference residual of the rule: that of the faded freestream (aero2d/rans.py:1408-1414)
        R_fs = S.residual(S.initial_state(self.winf), self.winf, 0.0, order=2)
        self.r_fs = float(np.sqrt(np.mean((R_fs / S.vol[:, None]) ** 2, 0)).sum())
        S._set(Q)
        S._freeze(None)
        self.warm = self.Q0 is not None
        self.turning = self.warm and self.alpha0 is not None and abs(self.alpha - float(self.alpha0)) > 1e-12
        self.peak = [0.0, 0, None]
        self.gamma = 0.0
        if self.warm and S.vortex_bc:
            self.gamma = 0.5 * S._forces_dict(self.alpha)["cl"] * S.mach * S.chord
        self.have_prev, self.prev_rn, self.prev_order = False, None, None
        self.n_rej = self.ok_streak = self.n_tiny = 0
        self.hist, self.fh = [], []
        self.r0 = None
        self.cfl = self.cfl0
        self.cfl_target, self.cfl_pen = None, 1.0
        if S.cfl_mode == "formula":      # checks the formula before the first step (aero2d/rans.py:1430-1431)
            self.cfl = min(self.cfl_max, cfl_formula(S.cfl_formula, n=0, orders=0.0, res=1.0, dres=1.0,
                                                     cfl=self.cfl0))
        self.last_ok, self.last_lin = True, 0.0
        self.H = dict(it=[], res=[], rn=[], cl=[], cd=[], cfl=[], gmres=[], lin=[], factored=[])
        self.events = []
        self.best = dict(rn=np.inf, has=self.turning, it=0, cfl=self.cfl0, order=0)
        if self.turning:
            S._keep_state(0)
        self.tb = 1.0
        self.first_until = 0
        self.n_back = 0
        self.frozen = False
        self.prev_order_used = None
        self.conv_ok = False
        self.status = ""
        self.it = 0
        self.cfl_hold = None
        self.ilu_state = None            # (iteration, CFL, order) of the last factorisation

    def _say(self, text):
        if self.verbose:
            print(f"  it {self.it:4d}: {text}", flush=True)

    def plan(self):
        S, c = self.S, self.control
        if self.it >= self.max_iter:
            self.done = True
            return None
        self.it += 1
        it = self.it
        if c is not None:
            v_ = c.cfl_set
            if v_ is not None and float(v_) > 0:
                c.cfl_set = None
                self.cfl = float(v_)
                self.cfl_hold = self.cfl if c.cfl_hold else None
            sk, stp = bool(c.skip), bool(c.stop)
            if sk or stp:
                self.status = "accepted by hand" if sk else "stopped"
                c.skip = False
                self.done = True
                return None
            f_ = float(c.cfl_factor)
            if f_ != 1.0:
                self.cfl = max(1.0, self.cfl * f_)
                c.cfl_factor = 1.0
        order = 1 if (it <= self.first_order_iters and self.Q0 is None) or it <= self.first_until else S.order
        if (order == 2 and self.prev_order_used == 1 and S.cfl_reset_on_order and self.cfl_hold is None):
            self.cfl = self.cfl0
            self.ilu_state = None
        self.prev_order_used = order
        self.order = order
        gm = self.gamma if S.vortex_bc else 0.0
        rn5 = S._rn(order, self.winf, gm)
        rn = float(rn5.sum())
        self.rn = rn
        manual = c is not None and c.rollback
        if manual:
            c.rollback = False
        best = self.best
        if best["order"] != order:
            best.update(rn=np.inf, order=order)
        blowup = (not np.isfinite(rn)) or (S.rollback_auto and rn > S.rollback_factor * best["rn"]
                                           and it - best["it"] >= 3 and not (self.warm and it <= 8))
        if (blowup or manual) and best["has"]:
            self.n_back += int(not manual)
            self.tb = 1.0
            why = "by hand" if manual else ("NaN" if not np.isfinite(rn) else
                                             f"residual x{rn / best['rn']:.3g} of the lowest")
            self.events.append((it, f"back to iteration {best['it']} ({why})", rn, []))
            self._say(f"back to iteration {best['it']} ({why})")
            S._keep_state(1)
            if self.n_back > S.max_rollbacks:
                self.status = f"unstable: {self.n_back - 1} rollbacks did not help (best state kept)"
                self.done = True
                return None
            self.cfl = max(self.cfl0, min(self.cfl, best["cfl"]) * 0.1)
            self.ilu_state = None
            self.first_until = it + (S.rollback_first_order if not np.isfinite(rn) else 0)
            if self.frozen:
                S._freeze(None)
                self.frozen = False
            best.update(rn=np.inf, it=it)
            return None
        if (S.rollback_auto and self.have_prev and self.prev_order == order and self.n_rej < S.max_rejects
                and (not np.isfinite(rn) or rn > S.step_reject * self.prev_rn)
                and not (self.warm and it <= 8) and not manual):
            self.n_rej += 1
            self.ilu_state = None
            S._keep_state(2)
            self.cfl = max(self.cfl0, self.cfl * 0.3)
            self.events.append((it, f"step undone (residual x{rn / self.prev_rn:.3g}), CFL -> {self.cfl:.3g}",
                                rn, []))
            self._say(f"step undone (residual x{rn / self.prev_rn:.3g}), CFL -> {self.cfl:.3g}")
            return None
        self.n_rej = 0
        self.ok_streak += 1
        if not np.isfinite(rn):
            self.status = "diverged"
            self.done = True
            return None
        peak = self.peak
        if peak[2] is None:
            peak[2] = rn
        if rn > peak[0]:
            peak[:2] = [rn, it]
        past_peak = (peak[0] >= 5.0 * peak[2] and it - peak[1] >= 5) or it > 40
        if rn < 0.95 * best["rn"] and (not self.turning or past_peak):
            best.update(rn=rn, has=True, it=it, cfl=self.cfl, order=order)
            S._keep_state(0)
        hist = self.hist
        if self.r0 is None or (order == 2 and len(hist) and hist[-1][1] == 1):
            self.r0 = rn
        hist.append((rn, order))
        fr_ = S._forces_dict(self.alpha)
        for k_, v_ in (("it", it), ("res", rn5), ("rn", rn), ("cl", fr_["cl"]), ("cd", fr_["cd"]),
                       ("cfl", self.cfl)):
            self.H[k_].append(v_)
        if order == 2:
            S._rref = max(getattr(S, "_rref", 0.0), rn)
        if c is not None and c.monitor is not None:
            c.monitor(dict(alpha=self.alpha, it=it, res=rn5, rn=rn, cl=fr_["cl"], cd=fr_["cd"], cfl=self.cfl,
                           rref=getattr(S, "_rref", rn), order=order, r_fs=self.r_fs, lin=self.last_lin))
        if order == 2 and rn < self.tol * self.r0:
            self.conv_ok = True
            self.done = True
            return None
        crit = self.crit
        if order == 2:
            fh = self.fh
            fh.append((fr_["cl"], fr_["cd"]))
            o2 = [h[0] for h in hist if h[1] == 2]
            rref = self.r_fs
            stalled = len(o2) > 60 and min(o2[-30:]) > 0.5 * min(o2[:-30])

            def steady(win, tol_cl, tol_cd):
                if len(fh) <= win:
                    return False
                cls = [f[0] for f in fh[-win - 1:]]
                cds = [f[1] for f in fh[-win - 1:]]
                return (max(cls) - min(cls) < tol_cl * max(0.2, abs(fh[-1][0]))
                        and max(cds) - min(cds) < tol_cd)
            hold = max(1, int(crit.hold))
            lvl = 10.0 ** (-crit.orders) * rref
            own = max(o2)
            dropped = (not self.warm) or rn <= 1e-2 * own
            end = None
            if (len(o2) >= hold and max(o2[-hold:]) <= lvl and dropped
                    and steady(hold, crit.cl_tol, crit.cd_tol)):
                end = "converged"
            elif (len(o2) >= 8 and max(o2[-3:]) <= 10.0 ** (-1.5) * lvl and rn <= 1e-3 * own
                    and steady(3, crit.cl_tol, crit.cd_tol)):
                end = "converged"
            elif (crit.stall_ok and len(o2) > 40 and stalled and rn < 10.0 ** (1.0 - crit.orders) * rref
                    and steady(max(hold, 30), crit.cl_tol, crit.cd_tol)):
                end = "converged (residual stalled, forces steady)"
            elif (crit.stall_ok and len(o2) > 60 and stalled and rn < 10.0 ** -2.5 * own
                    and steady(max(hold, 40), crit.cl_tol, crit.cd_tol)):
                end = "converged (forces steady)"
            if end is not None:
                self.conv_ok, self.status, self.done = True, end, True
                return None
            if (S.freeze_limiter and not self.frozen and S.limiter_k > 0 and S.limiter_zone != "off"
                    and stalled and steady(30, 0.005, 1e-3)):
                Qc = S._get()
                S._freeze(S.phi(Qc))
                S._set(Qc)
                self.frozen = True
                self._say(f"residual stalled at {rn:.2e}: limiter frozen")
        # CFL (aero2d/rans.py:1617-1680)
        if S.cfl_mode == "formula":
            # the user's expression; after a protective cut (rejected step,
            # rollback) the CFL returns to it x1.25 per good iteration unless the
            # expression itself reads cfl or dres (aero2d/rans.py:1643-1665)
            dres = rn / max(hist[-2][0], 1e-300) if len(hist) >= 2 else 1.0
            target = cfl_formula(S.cfl_formula, n=it, orders=float(np.log10(max(hist[0][0], 1e-300)
                                                                            / max(rn, 1e-300))),
                                 res=rn / max(hist[0][0], 1e-300), dres=dres, cfl=self.cfl)
            target = float(min(self.cfl_max, max(target, 1e-3)))
            if cfl_formula_recursive(S.cfl_formula):
                self.cfl_target, self.cfl_pen, self.cfl = target, 1.0, target
            else:
                last, pen = self.cfl_target, self.cfl_pen
                if last is not None and self.cfl < 0.99 * last * pen:
                    pen = self.cfl / last
                pen = min(1.0, pen * 1.25) if self.last_ok else pen
                self.cfl_target, self.cfl_pen = target, pen
                self.cfl = target * pen
        elif len(hist) >= 2 and hist[-2][1] == order:
            ok_, lin_ = self.last_ok, self.last_lin
            if S.cfl_mode == "fixed":
                self.cfl = min(self.cfl_max, self.cfl * S.cfl_growth) if ok_ and rn < 3.0 * hist[-2][0] \
                    else max(self.cfl0, 0.5 * self.cfl)
            elif S.cfl_mode == "ser":
                self.cfl = min(self.cfl_max, max(self.cfl0, self.cfl * min(2.0, hist[-2][0] / rn)))
            else:
                grow = (S.cfl_growth if rn < 1.2 * hist[-2][0] and ok_ and lin_ < 0.25 else
                        S.cfl_lin_back if lin_ >= 0.25 and ok_ and rn < 3.0 * hist[-2][0] else
                        1.0 if rn < 3.0 * hist[-2][0] and ok_ else 0.5)
                self.cfl = min(self.cfl_max, max(self.cfl0, self.cfl * grow))
        if self.cfl_hold is not None:
            self.cfl = float(self.cfl_hold)
        if S.turb_boost > 1.0:
            ok_ = self.last_ok
            fw_ = self.fh[-11:] if order >= 2 else []
            calm = (len(fw_) == 11 and max(f[0] for f in fw_) - min(f[0] for f in fw_)
                    < 2e-3 * max(0.2, abs(fw_[-1][0]))
                    and max(f[1] for f in fw_) - min(f[1] for f in fw_) < 5e-4)
            if calm and ok_ and self.cfl >= S.turb_boost_cfl and len(hist) >= 2 and rn < 1.2 * hist[-2][0]:
                self.tb = min(S.turb_boost, 1.5 * self.tb)
            elif not ok_ or (len(hist) >= 2 and rn >= 1.2 * hist[-2][0]):
                self.tb = max(1.0, self.tb / 4.0)
        ct = (S.cfl_turb if S.cfl_turb is not None else 1.0) * self.tb
        gtol = self.gmres_tol if (self.warm and self.n_back == 0) else min(self.gmres_tol, 0.01)
        # laminar solves take right preconditioning: left preconditioning
        # "converges" on |M^-1 r| while the true residual stays near 0.2
        # (aero2d/rans.py:1781-1790)
        # the last factorisation stays the preconditioner for up to prec_lag
        # iterations while the CFL is within a factor 2 of its own and the
        # order is the same (aero2d/rans.py:1736-1752); prec_lag 1 factors
        # every iteration
        st_ = self.ilu_state
        fresh = (st_ is None or it - st_[0] >= max(1, int(S.prec_lag)) or not (0.5 < self.cfl / st_[1] < 2.0)
                 or st_[2] != order)
        self.step_cfl = self.cfl
        return dict(slot=self.slot, order=order, winf=self.winf, gam=gm, cfl=self.cfl, ct=ct, gtol=gtol,
                    right=S.right_prec or not S.turbulent, gmres_iter=self.gmres_iter, reuse=not fresh)

    def after(self, st):
        S, it = self.S, self.it
        if st["ilu_nonfinite"]:
            self._say(f"ILU not finite, CFL {self.cfl:.3g} -> {max(self.cfl0, self.cfl / 10):.3g}")
            self.cfl = max(self.cfl0, self.cfl / 10)
            self.ilu_state = None
            return
        if st["factored"]:
            self.ilu_state = (it, self.step_cfl, self.order)
        lin = st["lin"]
        self.H["gmres"].append(st["gmres_iterations"])
        self.H["lin"].append(lin)
        self.H["factored"].append(st["factored"])
        self.last_ok = bool(np.isfinite(lin) and lin < 0.5)
        self.last_lin = lin
        if not st["applied"]:
            self._say(f"GMRES did not converge ({lin:.2g}), residual {self.rn:.3e}, CFL {self.cfl:.3g} -> "
                      f"{max(self.cfl0, self.cfl / 4):.3g}")
            self.cfl = max(self.cfl0, self.cfl / 4)
            return
        lam = st["lam"]
        if lam < 1.0 and not st["per_cell"]:
            self.cfl = max(self.cfl0 if self.n_tiny < 10 else 0.02 * self.cfl0, self.cfl * 0.5)
        self.n_tiny = self.n_tiny + 1 if lam < 0.05 and not st["per_cell"] else 0
        if self.n_tiny >= 60:
            self.status = "stuck: the Newton step was cut to almost nothing for 60 iterations"
            self.done = True
            return
        self.have_prev, self.prev_rn, self.prev_order = True, self.rn, self.order
        if S.vortex_bc and it % 5 == 0:
            self.gamma = 0.5 * S._forces_dict(self.alpha)["cl"] * S.mach * S.chord
        if self.verbose and (it % 10 == 0 or it < 5):
            fr = S._forces_dict(self.alpha)
            print(f"  it {it:4d} ord {self.order} res {self.rn:.3e} cfl {self.cfl:.2e} lam {lam:.2f} "
                  f"gmres {st['gmres_iterations']}:{lin:.1e} Cl {fr['cl']:.4f} Cd {fr['cd']:.5f}", flush=True)
        if self.callback and self.callback(it, self.rn):
            self.done = True

    def finish(self):
        S = self.S
        if self.status == "diverged" and self.best["has"]:
            S._keep_state(1)
            self.status = "diverged (best state kept)"
        fr = S._forces_dict(self.alpha)
        Hn = {k_: np.array(v_) for k_, v_ in self.H.items()}
        if not self.status:
            self.status = "converged" if self.conv_ok else "not converged within the iterations allowed"
        Hn["rref"] = float(getattr(S, "_rref", 0.0) or (max(Hn["rn"]) if len(Hn["rn"]) else 1.0))
        Hn["r_fs"] = self.r_fs
        Hn["events"] = self.events
        if self.n_back:
            self.status += f"; rollbacks: {self.n_back}"
        S.gamma_vortex = self.gamma
        return RansResult(self.alpha, fr["cl"], fr["cd"], fr["cdp"], fr["cdf"], fr["cm"], self.conv_ok, self.it,
                          [h[0] for h in self.hist], fr["elements"], S._get(), self.status, Hn)
