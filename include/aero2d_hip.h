/* aero2d_hip.h — C ABI of the MI355X implicit Euler library (libaero2d_hip.so).
 *
 * Drop-in numerics for aero2d's batched implicit Euler path. The reference has
 * no FFI: its boundary is a Python class API (aero2d/euler.py EulerSolver,
 * aero2d/implicit.py ImplicitEuler). Each entry point below replaces the
 * device-side work behind one of those methods; the package aero2d_amd keeps
 * the Python classes and binds these symbols with ctypes.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative code on failure;
 *    a2d_last_error() then describes the failure (thread-local string).
 *  - Host arrays are caller-owned and copied. Field arrays use the
 *    reference's layout: (4, ni, nc) C order, j (wall-normal) fastest; node
 *    arrays (ni, nc + 1). Device memory is owned by the library.
 *  - A pool owns one HIP stream and the device memory of its cases. A pool is
 *    not thread-safe: drive it from one host thread. Pools are independent.
 *  - precision: 0 = FP64 throughout; 1 = mixed (FP32 Jacobian-vector product,
 *    line preconditioner and Krylov basis; FP64 residual, Newton update and
 *    force integration).
 */
#ifndef AERO2D_HIP_H
#define AERO2D_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define A2D_PREC_FP64 0
#define A2D_PREC_MIXED 1

/* Solver knobs. Defaults (a2d_default_knobs) are the reference's:
 * EulerSolver(cfl=2.5, order=2, vortex_correction=True, lim_k=0.02)
 * (aero2d/euler.py:219-221) and ImplicitEuler(gmres_restart=50,
 * gmres_tol=0.01, jacobi_sweeps=4, max_update=0.2) (aero2d/implicit.py:68-70).
 * early_exit = 0 keeps JAX's "batched" GMRES semantics (a fixed `restart`
 * Arnoldi steps, aero2d/implicit.py:184-185); 1 stops a case's Arnoldi loop once
 * its preconditioned residual estimate is below gmres_tol * |M b|. */
typedef struct a2d_knobs {
  double cfl_explicit;
  double lim_k;
  double max_update;
  double gmres_tol;
  int vortex;
  int order;
  int gmres_restart;
  int jacobi_sweeps;
  int early_exit;
  int use_graph;      /* replay captured HIP graphs for repeated batches */
  int precond_order;  /* 0: line Jacobi over i (reference, aero2d/implicit.py:150-157);
                         1: red-black line ordering (odd lines, then even lines
                         with the updated odd neighbours), same cost per sweep */
  int positivity;     /* 0: accept every finite step (reference); 1: also reject a
                         step that would leave a cell with rho <= 0 or p <= 0 */
  int precond_half;   /* mixed form only: 1 = the line sweeps read the factored
                         couplings (computed in FP32) from an FP16 copy, half the
                         bytes of the preconditioner's dominant traffic; 0 = FP32.
                         The factorisation writes both copies, so the knob may
                         change between factorisation and application. */
  int basis_half;     /* mixed form only: 1 = the Krylov basis vectors are stored
                         in FP16 (each scaled by a power of two to entries of
                         order one; Gram-Schmidt arithmetic stays FP32/FP64),
                         half the bytes of the Gram-Schmidt passes; 0 = FP32 */
  int stencil_march;  /* residual / Jacobian-vector product kernel: 0 = LDS tiles
                         (the case's tile shape), 2 = column march (a wavefront
                         walks a strip of columns up the rows), 1 = the march
                         for batches large enough to fill the device (default);
                         same results to round-off either way */
  int sweep_lds;      /* where a line sweep keeps the intermediate values of its
                         elimination between its two passes: 0 = HBM; 2 = LDS
                         when every block of the launch can still be resident;
                         1 = that only for batches that fill the device (32-line
                         blocks; default). Same results bit for bit */
  double irs_eps;     /* explicit scheme: residual smoothing coefficient
                         (EulerSolver.irs_eps, aero2d/euler.py:219-236; 0.6);
                         0 switches smoothing off */
} a2d_knobs;

/* Library set-up on one HIP device; idempotent. */
int a2d_init(int device);
const char* a2d_last_error(void);
int a2d_version(void);
void a2d_default_knobs(a2d_knobs* k);

/* Pools. hbm_budget_bytes bounds the sum of the pool's case footprints
 * (admission by HBM footprint for the serving loop). */
int a2d_pool_create(int64_t hbm_budget_bytes, int64_t* pool_out);
int a2d_pool_destroy(int64_t pool);
int a2d_pool_set_knobs(int64_t pool, const a2d_knobs* k);
int a2d_pool_get_knobs(int64_t pool, a2d_knobs* k);
int64_t a2d_pool_used_bytes(int64_t pool);
/* Counters of a pool's work since its creation (no counterpart in the
 * reference), up to n of: Newton steps, steps replayed from a captured HIP
 * graph, graph captures, batch plans built, plan tables allocated (the rest
 * reused released ones), case slabs allocated, case slabs reused from removed
 * cases, bytes of removed cases' slabs held for reuse. A removed case's slab
 * is kept for the next case of about its size; with a budget, held slabs are
 * freed as needed to keep live plus held bytes within it. */
int a2d_pool_stats(int64_t pool, int64_t* out, int n);
/* Device bytes one case of this size occupies (the admission footprint), or
 * -6 for a size a2d_case_add refuses (see there). */
int64_t a2d_case_bytes(int ni, int nc, int precision, int restart);

/* Cases. Replaces EulerSolver(grid, dtype) (aero2d/euler.py:219-250) for one
 * (grid, M, alpha): metrics are computed from the nodes (FP64,
 * aero2d/euler.py:44-60), freestream winf = (1, M cos a, M sin a, 1/gamma)
 * (aero2d/euler.py:253-256) and U = freestream (aero2d/euler.py:383-384).
 * x, y: node coordinates (ni, nc + 1). Fails with -2 when the pool's HBM
 * budget would be exceeded, and with -6 (before any device work) for a grid
 * whose preconditioner block storage, 80 * ni * nc values of the inner type,
 * would reach 4 GiB: kernels address inside a case with 32-bit offsets. That
 * caps a case at about 6.7 M cells in FP64 and 13.4 M cells in the mixed form. */
int a2d_case_add(int64_t pool, int ni, int nc, const double* x, const double* y, double mach,
                 double alpha_deg, int precision, int* case_id);
int a2d_case_remove(int64_t pool, int case_id);
/* Replace the case's freestream (M, alpha); the state is left as it is. */
int a2d_case_set_freestream(int64_t pool, int case_id, double mach, double alpha_deg);
int a2d_case_set_state(int64_t pool, int case_id, const double* U);  /* (4, ni, nc) */
int a2d_case_get_state(int64_t pool, int case_id, double* U);
/* This case's bound on the relative change of density and pressure per
 * Newton step (the step limiter omega = min(1, max_update / change));
 * 0 restores the pool's knob. No counterpart in the reference, whose bound is
 * one constant (ImplicitEuler.max_update, aero2d/implicit.py:68-78, 187-190):
 * the mixed form starts cases at a larger bound and drops a case to the
 * reference's after its first rejected step. Takes effect from the next
 * Newton step, also for replayed graphs. */
int a2d_case_set_step_limit(int64_t pool, int case_id, double max_update);
/* Park (1) or unpark (0) a case: a parked case may stay in the batches it is
 * named in, takes no Newton step there (no kernel does work for it, its state
 * is kept, its rows of a2d_newton_step's output are to be ignored). No
 * counterpart in the reference: it lets a serving loop keep a batch's
 * composition, and so its captured HIP graph, when a case finishes while no
 * request waits to take its place. */
int a2d_case_park(int64_t pool, int case_id, int parked);
/* Stencil tile shape used by this case's residual/JVP (index into the shape
 * table of a2d_stencil_shapes). */
int a2d_case_set_shape(int64_t pool, int case_id, int shape);
int a2d_stencil_shapes(int* ti, int* tj, int max_shapes);

/* Single-case operators.
 * a2d_residual: EulerSolver.residual(U, winf, circ) (aero2d/euler.py:259-287);
 *   circ NaN means Gamma = Cl |V_inf|/2 of U (ImplicitEuler._res,
 *   aero2d/implicit.py:166-167). rn = sqrt(mean((R_rho/vol)^2)).
 * a2d_forces: EulerSolver.forces (aero2d/euler.py:318-331) -> cl, cd, cm, cp(ni).
 * a2d_local_dt: EulerSolver.local_dt(W(U)) (aero2d/euler.py:289-296), (ni, nc).
 * a2d_jvp: d/de residual_with_circ(U + e v) (jax.jvp at aero2d/implicit.py:180),
 *   plus diag*v when with_diag (diag = vol/dtau at CFL `cfl`).
 * a2d_assemble: first-order blocks D, Bim, Bip, Bjm, Bjp at CFL `cfl`
 *   (ImplicitEuler._assemble, aero2d/implicit.py:81-110), each (ni, nc, 4, 4).
 * a2d_precond_setup / a2d_precond_apply: factor the lines at CFL `cfl`
 *   (aero2d/implicit.py:112-128), then x = M r (aero2d/implicit.py:150-157). */
int a2d_residual(int64_t pool, int case_id, double circ, double* R, double* rn);
int a2d_forces(int64_t pool, int case_id, double* clcdcm, double* cp, double* circ);
int a2d_local_dt(int64_t pool, int case_id, double* dt);
int a2d_jvp(int64_t pool, int case_id, const double* v, double cfl, int with_diag, double* out);
int a2d_assemble(int64_t pool, int case_id, double cfl, double* D, double* Bim, double* Bip,
                 double* Bjm, double* Bjp);
int a2d_precond_setup(int64_t pool, int case_id, double cfl);
int a2d_precond_apply(int64_t pool, int case_id, const double* r, double* x);

/* One pseudo-transient Newton step (ImplicitEuler.step, aero2d/implicit.py:169-192)
 * for n cases of one precision, each at its own CFL. The ids must be distinct
 * (-6 otherwise: a case's in-place updates would run twice). The update is applied
 * on the device only when rn, the step and its limiter are finite
 * (aero2d/implicit.py:218-225). out: 10 doubles per case:
 *   rn (before the update), omega, accepted (0/1), cl, cd, cm (after),
 *   Arnoldi steps taken, max relative change, beta, GMRES residual estimate. */
int a2d_newton_step(int64_t pool, const int* case_ids, int n, const double* cfl, double* out);

/* The explicit scheme (EulerSolver.solve's cycle, aero2d/euler.py:334-405).
 * a2d_case_link_coarse: make coarse_id the next coarser grid level of case_id
 * (the reference's EulerSolver.coarse, built from grid.x[::2, ::2],
 * aero2d/euler.py:238-250); both FP64 cases, the coarse one with half the cells
 * of the fine one along i and j (even counts, -6 otherwise); coarse_id -1
 * unlinks. The coarse case takes its freestream and circulation from the
 * finest case of the batch at every cycle; its node coordinates are the
 * caller's (every other node of the fine grid for the reference's scheme).
 * a2d_explicit_cycles: ncycles cycles on n distinct FP64 cases, each with the
 * same number of linked levels: a four-stage Runge-Kutta iteration with local
 * time steps at CFL cfl_explicit and residual smoothing (knob irs_eps), a FAS
 * V-cycle through the coarse levels (first order there) with a second
 * iteration after the correction, then Gamma = Cl |V_inf| / 2 of the new state
 * for the next cycle (aero2d/euler.py:396-405). fresh: Gamma = 0 before the
 * first cycle (the start of EulerSolver.solve). out: 5 doubles per case, res of
 * the last cycle (sqrt(mean(((U - U0)_rho / dt)^2)) of its first iteration),
 * cl, cd, cm, Gamma. -5 for a case of the mixed form, -6 for a parked case or
 * rows of more than 2048 cells (the smoothing kernel keeps a row in LDS). */
int a2d_case_link_coarse(int64_t pool, int case_id, int coarse_id);
int a2d_explicit_cycles(int64_t pool, const int* case_ids, int n, int ncycles, int fresh, double* out);

/* Kernel timing (no counterpart in the reference; for bench.py). Launches
 * kernel family `which` over the batch `reps` times back to back on the pool's
 * stream between two HIP events and returns the mean duration in ms:
 *   0 residual (FP64)   1 Jacobian-vector product   2 line sweep with i-coupling
 *   3 Gram-Schmidt dots (k = 25)   4 block assembly   5 line factorisation
 *   6 Gram-Schmidt update (k = 25)   7 plain vector copy (4 words/cell each way)
 *   8 the per-cell pass of a coupled line sweep   9 the line recurrences alone
 *   10 one whole preconditioner application (all sweeps, in the pool's order)
 *   11 explicit scheme: residual smoothing along j (FP64 cases)
 *   12 explicit scheme: a later Runge-Kutta stage, smoothing along i and the
 *      update (FP64 cases; the state is restored afterwards)
 * The cases must have taken a Newton step (their blocks and Krylov vectors
 * hold data); the state is not changed. The ids must be distinct. */
int a2d_time_kernel(int64_t pool, const int* case_ids, int n, int which, int reps, double* ms);

/* Test hook: the device HLLC flux on n faces. WL, WR (4, n), S (2, n), F (4, n). */
int a2d_hllc_flux(int n, const double* WL, const double* WR, const double* S, double* F);
/* The same without any device allocation: for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). No counterpart in the reference. Stages through a fixed device
 * array of the library (7.3 MB per device, held until the process ends), so
 * n is at most 65536 (-6 beyond that). Calls that overlap on one device race
 * on that array: do not call it from two threads at once. */
int a2d_hllc_flux_noalloc(int n, const double* WL, const double* WR, const double* S, double* F);

/* ---- RANS (Spalart-Allmaras, SA-neg) Newton-Krylov core on an unstructured mesh ----
 * Replaces the jitted functions of RansSolver._fns (aero2d/rans.py:1169-1204),
 * local_dt (1206-1249), forces (1251-1296), the linear solve and update of
 * solve (1738-1858), BlockILU0 (aero2d/bilu.py:13-166) and gmres
 * (aero2d/krylov.py:14-87). All arrays are FP64; per-cell arrays are (nc, 5)
 * cell-major as the reference's Q, and are copied. A solver is one device
 * context (mesh, state, Jacobian, ILU, Krylov basis) owned by the library. */
typedef struct a2d_rans_params {
  double mach, re, chord, xref_x, xref_y, chi_inf, limiter_k, zmin;
  int order;          /* 1 or 2 (RansSolver.order) */
  int turbulent;      /* 0: laminar, the rho chi equation is rho chi = 0 */
  int vortex_bc;      /* informational: the host passes gam = 0 without it */
  int limiter_zone;   /* 0 "off", 1 "te", 2 "all" */
  int sa_raw_grad;    /* 1: the SA source from the unlimited gradients (aero2d/rans.py:309,961) */
  int visc_raw_grad;  /* 1: the viscous fluxes from the unlimited gradients (aero2d/rans.py:310,897) */
} a2d_rans_params;

/* Mesh arrays as RansSolver._setup forms them (aero2d/rans.py:328-425):
 * interior faces: face_l, face_r (nfi), face_geom (nfi, 11) =
 *   nx ny A rLx rLy rRx rRy dx dy dist w_i   (d = xc[R] - xc[L], w_i = 1/dist^2);
 * wall faces: wall_cell (nfw), wall_geom (nfw, 9) = nx ny A rx ry dn w_w fcx fcy;
 * far faces: far_cell (nff), far_geom (nff, 7) = nx ny A rx ry fcx fcy;
 * cells: cell_geom (nc, 6) = vol wall_d Minv00 Minv01 Minv11 in_te_zone.
 * -6 for a bad index or size. */
int a2d_rans_create(int nc, int nfi, int nfw, int nff, const int* face_l, const int* face_r,
                    const double* face_geom, const int* wall_cell, const double* wall_geom, const int* far_cell,
                    const double* far_geom, const double* cell_geom, const a2d_rans_params* prm,
                    int64_t* handle_out);
int a2d_rans_destroy(int64_t h);
int a2d_rans_set_state(int64_t h, const double* Q);   /* (nc, 5) */
int a2d_rans_get_state(int64_t h, double* Q);
/* Frozen limiter for the second-order residual and JVP (res2/jvp2 of _fns);
 * NULL returns to the limiter of the state (res2f/jvp2f). */
int a2d_rans_freeze_limiter(int64_t h, const double* phi);
/* The Venkatakrishnan limiter of the state in every cell (the "phi" entry of _fns). */
int a2d_rans_limiter(int64_t h, double* phi);
/* residual(Q, winf, gam, order) (aero2d/rans.py:860-964): R (nc, 5) or NULL;
 * rn5 (5) = sqrt(mean((R_k / vol)^2)) per equation, or NULL. winf (5). */
int a2d_rans_residual(int64_t h, int order, const double* winf, double gam, double* R, double* rn5);
/* jax.jvp of the residual along v (nc, 5) (jvp1/jvp2/jvp2f of _fns). */
int a2d_rans_jvp(int64_t h, int order, const double* winf, double gam, const double* v, double* out);
/* local_dt(Q, cfl) (aero2d/rans.py:1206-1249), dt (nc). */
int a2d_rans_local_dt(int64_t h, double cfl, double* dt);
/* forces(Q, alpha) (aero2d/rans.py:1251-1296): out5 = cl cd cdp cdf cm; cp, cf
 * (nfw, per wall face) or NULL. */
int a2d_rans_forces(int64_t h, double alpha_deg, double* out5, double* cp, double* cf);
/* _jac1_blocks(Q, winf, gam) (aero2d/rans.py:1107-1126): D (nc, 5, 5) without
 * vol/dt and the off-diagonal blocks (2 nfi, 5, 5) for (iL,iR) then (iR,iL). */
int a2d_rans_jac1(int64_t h, const double* winf, double gam, double* D, double* off);
/* BlockILU0 structure (aero2d/bilu.py:14-130): perm (nc, perm[new] = old),
 * the pattern in the new numbering sorted by row then column (rows, cols:
 * nnz), and pos_src (nc + 2 nfi): the pattern entry of each Jacobian block in
 * the order of a2d_rans_jac1 (diagonal, (iL,iR), (iR,iL)). Once per solver. */
int a2d_rans_ilu_setup(int64_t h, int nnz, const int* perm, const int* rows, const int* cols, const int* pos_src);
/* nnz, forward levels, backward levels, wavefronts of the LDS-staged
 * factorisation (0: rows too long for LDS, the global kernel runs). */
int a2d_rans_ilu_info(int64_t h, int* out4);
/* Assemble vol/dt + J1 at the state and CFL (turbulence equation's step
 * times turb_factor) and factor it (BlockILU0.factor); nonfinite = count of
 * non-finite inverse-diagonal entries, or NULL. */
int a2d_rans_ilu_factor(int64_t h, const double* winf, double gam, double cfl, double turb_factor, int* nonfinite);
/* BlockILU0.solve of b (nc, 5) into x (nc, 5). */
int a2d_rans_ilu_apply(int64_t h, const double* b, double* x);
/* One Newton step of RansSolver.solve (aero2d/rans.py:1738-1858) at the
 * state: R, dt, ILU of vol/dt + J1, GMRES on (vol/dt + dR/dQ) dq = -R (left
 * preconditioning, retried with right preconditioning when the linear
 * residual is not below 0.5, or right only with right = 1), then, when the
 * linear residual is finite and <= 0.9, the step limit (rho, p fall by at
 * most 20 %; only the cells that ask for more when there are at most
 * step_limit_cells of them) and the update; the previous state is kept.
 * reuse_ilu = 1 keeps the slot's last factorisation as the preconditioner
 * (no Jacobian assembly, no factorisation: the reference's prec_lag,
 * aero2d/rans.py:1736-1752); it is honoured only when that factorisation was
 * made by a Newton step of this slot, was finite, and no other entry point
 * has factored since; otherwise the step factors.
 * out (12): lin, GMRES iterations, info, lam, rel, ilu_nonfinite, applied,
 * cells beyond 0.2, per-cell limit used, factored (1: this step factored,
 * 0: it reused the last factors), 0, 0. */
int a2d_rans_newton_step(int64_t h, int order, const double* winf, double gam, double cfl, double turb_factor,
                         double gmres_tol, int gmres_iter, int right, int step_limit_cells, int reuse_ilu,
                         double* out);
/* Bench hook: mean device time (ms, HIP events on the solver's stream) of one
 * of: 0 the ILU solve (BlockILU0.solve), 1 the ILU factorisation, 2 the
 * Newton matrix-vector product (Jacobian-vector product plus vol/dt), 3 the
 * residual, 4 the first-order Jacobian blocks; reps repetitions at the state.
 * 0 and 1 factor the Jacobian at CFL 10 first. No counterpart in the reference. */
int a2d_rans_time_kernel(int64_t h, int which, int order, const double* winf, double gam, int reps, double* ms);
/* Case slots (no counterpart in the reference, whose rans_polar runs angles
 * in separate processes): a solver holds ncases slots (grown, never shrunk)
 * that share its mesh, ILU structure and stream, each with its own state,
 * Jacobian, factors and Krylov basis. a2d_rans_select makes slot c the one
 * the single-case entry points above act on (slot 0 at creation). */
int a2d_rans_set_cases(int64_t h, int ncases);
int a2d_rans_select(int64_t h, int c);
/* a2d_rans_newton_step for n distinct slots at once: per slot its order,
 * winf (5), gam, CFL, turbulence step factor, GMRES tolerance, side and
 * reuse_ilu flag (reuse_ilu may be NULL: every slot factors); out (n, 12) as
 * a2d_rans_newton_step's. The factorisation of the slots that factor and
 * every triangular sweep of all slots run as one launch (a workgroup per
 * slot) and the Arnoldi steps go in lockstep; each slot's result equals its
 * step alone. */
int a2d_rans_newton_steps(int64_t h, int n, const int* case_ids, const int* order, const double* winf,
                          const double* gam, const double* cfl, const double* turb_factor, const double* gmres_tol,
                          int gmres_iter, const int* right, int step_limit_cells, const int* reuse_ilu,
                          double* out);
/* The last Newton step's dq before the step limit (nc, 5). */
int a2d_rans_newton_dq(int64_t h, double* dq);
/* 0: keep the state as the best; 1: return to the best; 2: undo the last step. */
int a2d_rans_keep(int64_t h, int what);

#ifdef __cplusplus
}
#endif

#endif /* AERO2D_HIP_H */
