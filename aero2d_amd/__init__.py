"""aero2d_amd — MI355X-native implicit Euler path of aero2d.

Drop-in for aero2d's batched implicit Euler solve: the pseudo-transient
Newton-Krylov loop (aero2d/implicit.py) over the finite-volume residual
(aero2d/euler.py) on marching O-grids (aero2d/grid.py), with the numerics in
hand-written gfx950 HIP kernels behind a C ABI (include/aero2d_hip.h); also
the explicit scheme, the run_euler application entry, the airfoil readers,
and the RANS (Spalart-Allmaras) Newton-Krylov core on unstructured meshes
(aero2d/rans.py, bilu.py, krylov.py) with its C-grid meshes.

    from aero2d_amd import naca4, ogrid, EulerSolver, ImplicitEuler
    g = ogrid(naca4("0012", 65), nj=49, ds0=2e-3, radius=40.0)
    r = ImplicitEuler(EulerSolver(g, dtype=np.float64)).solve([0.5], [2.0], tol=1e-10)
"""
from .geometry import (Airfoil, naca4, naca4_params, cosine_spacing, normalize, repanel, smooth_airfoil,
                       load_airfoil)
from .grid import OGrid, ogrid, growth_ratio, self_similar_spacing
from .euler import (GAMMA, EulerResult, EulerSolver, compute_metrics, cons_to_prim, hllc_flux,
                    prim_to_cons)
from .implicit import ImplicitEuler
from .rans_mesh import CGrid, Mesh, cgrid, cgrid_mesh, connectivity, ogrid_mesh, refine_cwall, refine_wall
from .rans import Convergence, RansControl, RansResult, RansSolver, cfl_formula, transfer_state

__version__ = "0.1.0"

__all__ = ["Airfoil", "naca4", "naca4_params", "cosine_spacing", "normalize", "repanel", "smooth_airfoil",
           "load_airfoil", "OGrid", "ogrid", "growth_ratio",
           "self_similar_spacing", "GAMMA", "EulerResult", "EulerSolver", "compute_metrics",
           "cons_to_prim", "hllc_flux", "prim_to_cons", "ImplicitEuler", "Mesh", "connectivity", "ogrid_mesh",
           "refine_wall", "CGrid", "cgrid", "cgrid_mesh", "refine_cwall", "Convergence", "RansControl",
           "RansResult", "RansSolver", "cfl_formula", "transfer_state"]
