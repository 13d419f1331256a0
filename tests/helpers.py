"""Shared test helpers: small grids, perturbed states, oracle views."""
import numpy as np
import torch

from aero2d_amd import naca4, naca4_params, ogrid
from oracle import flow

DT = torch.float64


def small_grid(n_side=17, nj=9, ds0=2e-2, radius=5.0, code="0012"):
    """NACA O-grid with (2 n_side - 2) x (nj - 1) cells."""
    return ogrid(naca4(code, n_side), nj=nj, ds0=ds0, radius=radius)


def perturbed_state(ni, nc, mach, alpha, seed, amp=0.05):
    """Freestream plus a smooth random perturbation, (4, ni, nc) conserved."""
    rng = np.random.default_rng(seed)
    a = np.deg2rad(alpha)
    W = np.empty((4, ni, nc))
    W[0], W[1], W[2], W[3] = 1.0, mach * np.cos(a), mach * np.sin(a), 1 / 1.4
    i = np.arange(ni)[:, None] / ni
    j = np.arange(nc)[None, :] / nc
    for k in range(4):
        c = rng.normal(size=4)
        field = (c[0] * np.sin(2 * np.pi * (i + c[1])) * np.cos(np.pi * j * (1 + abs(c[2])))
                 + 0.3 * rng.normal(size=(ni, nc)))
        scale = amp * (1.0 if k in (0, 3) else mach)
        W[k] = W[k] + scale * field * (W[k] if k in (0, 3) else 1.0)
    rho, u, v, p = W
    return np.stack([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)])


def oracle_geo(g):
    return flow.case_geometry(g.x, g.y)


def winf_t(mach, alpha):
    return flow.freestream(torch.tensor(mach, dtype=DT), torch.tensor(alpha, dtype=DT))


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
