"""CPU oracle for the implicit Euler path — TEST INFRASTRUCTURE ONLY.

A PyTorch-CPU FP64 restatement of aero2d's euler.py residual and implicit.py
Newton-GMRES solve, written to check the HIP library and to time a CPU
baseline. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
may import it, and only as the checker: the product path (aero2d_amd) never
imports or calls anything here and fails loudly when its HIP library is missing.

Modules:
  flow    metrics, thermodynamics, HLLC, limiter, ghosts, residual, local dt,
          forces (aero2d/euler.py)
  newton  face Jacobians, assembly, block line solve, line-Jacobi
          preconditioner, JAX-"batched" GMRES, Newton step and host loop
          (aero2d/implicit.py, jax.scipy.sparse.linalg.gmres)
  explicit  four-stage RK with local time steps, implicit residual smoothing,
          FAS V-cycle and the chunked host loop (aero2d/euler.py:219-435)
  rans    SA-neg RANS on an unstructured mesh: LM-Roe, viscous fluxes, least-
          squares gradients, Venkatakrishnan limiter, SA source, far field with
          the point vortex, first-order Jacobian blocks, local step, forces,
          step limit (aero2d/rans.py); block ILU(k) (aero2d/bilu.py); GMRES
          with CGS2 (aero2d/krylov.py)

Pinning (what checks each function, all in tests/test_oracle_*.py):
  cons_to_prim/prim_to_cons  round trip (tests/test_euler.py:27-29 of the reference)
  hllc_flux                  consistency F(W,W) = physical flux, antisymmetry,
                             supersonic upwinding, stationary contact
  van_albada                 closed form, symmetry, phi(a,a) = a
  wall_ghost                 normal velocity reversed, tangential kept
  farfield_ghost             freestream in -> freestream out; closed-form vortex
  compute_metrics            closed cell contours, total area = annulus area
  residual                   freestream preservation, NACA 0012 alpha=0 mirror
                             symmetry, telescoping conservation
  forces                     uniform pressure -> zero force, rotation identities
  jvp of the residual        central finite differences, linearity, the circulation
                             coupling (wall perturbation reaches far-field rows)
  face Jacobians             finite differences
  line solve / precond       dense numpy.linalg.solve of the assembled system
  precond_rb                 dense block Gauss-Seidel in red-black line order;
                             many sweeps converge to the dense full solve
  gmres_batched              dense least squares on small systems
  solve                      reference goldens (benchmarks/results.json:83-104):
                             converged FP64 Cl/Cd at M 0.5 a 2 and M 0.8 a 1.25
                             on naca4("0012",65), ogrid(nj=49, ds0=2e-3, radius=40)
  explicit.smooth            the smoothing operators applied to its output
                             return the input (tests/test_oracle_explicit.py)
  explicit.sum4/prolong      volume-weighted mean of four cells; constant and
                             linear corrections reproduced
  explicit.iterate/cycle/    the solve drives R(U) to zero; that state is a
  solve                      fixed point of both; Newton steps of `newton`
                             from it leave it (same zero of the residual)
  rans.roe                   consistency F(w,w) = physical flux, antisymmetry,
                             supersonic upwinding (tests/test_oracle_rans.py)
  rans.visc_flux             closed-form shear layer stress and conduction
  rans.sa_terms/sa_source    fv1(chi = cv1) = 1/2, fw(r = 1) = 1, source at r = 1
  rans.gradients             exact for linear fields away from walls
  rans.residual              freestream preservation, mirror symmetry at alpha 0,
                             telescoping to boundary fluxes
  rans.jvp / jac1_parts      central differences, linearity
  rans.BlockILU              exact for block-tridiagonal matrices (the
                             reference's tests/test_rans.py:13-31 pattern) and
                             with complete fill; fill grows the pattern
  rans.gmres                 left and right preconditioning converge on a dense
                             system (tests/test_rans.py:201-209 pattern)
  rans.local_dt / forces     linear in CFL; uniform pressure gives no force
Parity unpinned: the per-iteration trajectory of JAX's GMRES (its internal
re-orthogonalisation test and normal-equation least squares are restated from
JAX's published source, not run); only converged answers are pinned.
"""
