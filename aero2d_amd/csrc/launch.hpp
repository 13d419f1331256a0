// Host-side launch wrappers for the device kernels (one translation unit per
// kernel family; the engine only sees these declarations).
//
// Plans: every wrapper takes a Plan whose blk_prefix lists, per batch entry,
// the blocks of that kernel family: per case (1), stencil tiles (per shape),
// cells (256 per block), vector chunks (VEC_CHUNK cells per block) or lines
// (LINES_PER_BLOCK per block).
#pragma once
#include <hip/hip_runtime.h>
#include "common.hpp"

namespace a2d {

#ifndef A2D_VEC_CHUNK
#define A2D_VEC_CHUNK 1024
#endif
constexpr int VEC_CHUNK = A2D_VEC_CHUNK;  // cells per block of the vector kernels
constexpr int CELL_BLOCK = 256;      // cells per block of the assembly kernel
constexpr int LINES_PER_BLOCK = 32;  // i-lines per block of the line factorisation
// Line sweeps run 8 lanes per line, so 8 lines per wavefront, with blocks of
// 32 lines (4 wavefronts) when the batch nearly fills the device (60 % of
// the CUs) and of 8 lines otherwise (a single 512-line case then spreads
// over 32 CUs instead of 8).
constexpr int SWEEP_LINES_BIG = 32, SWEEP_LINES_SMALL = 8;

struct KnobsDev {        // solver knobs the kernels read
  double cfl_explicit;   // EulerSolver.cfl (2.5)
  double lim_k;          // limiter scale factor (0.02)
  double max_update;     // step limiter (0.2)
  double gtol;           // GMRES tolerance (0.01)
  int vortex;            // far-field vortex correction on/off
  int order;             // 1 or 2
  int early_exit;        // stop Arnoldi once the estimate meets gtol
  int restart;           // Arnoldi steps per Newton step
  int precond_rb;        // 1: red-black line ordering (else line Jacobi)
  int positivity;        // 1: reject a step that leaves rho <= 0 or p <= 0 somewhere
  int precond_half;      // 1: mixed-form sweeps read the FP16 copy of the couplings
  int basis_half;        // 1: mixed-form Krylov basis stored in FP16
};

// ---- stencil.hip ------------------------------------------------------------
struct StencilShape { int ti, tj; };
int stencil_num_shapes();
StencilShape stencil_shape(int idx);
// Workgroups of one case's stencil launch (= its residual-norm partials).
int stencil_blocks(int shape, int ni, int nc);

// ---- march.hip ----------------------------------------------------------------
// The column-march stencil: a unit is one wavefront walking a strip of 60
// columns up a chunk of rows; four units per workgroup. Chunk heights: the
// FP64 residual and tangent 16 rows, the FP32 tangent 8 (twice the
// wavefronts, for latency hiding at its higher occupancy). The residual writes
// one norm partial per unit.
constexpr int MARCH_OWN = 60, MARCH_WAVES = 4;
constexpr int MARCH_ROWS_RES = 16, MARCH_ROWS_JVP64 = 16, MARCH_ROWS_JVP32 = 8;
__host__ __device__ inline int march_units(int ni, int nc, int rows) {
  return ((ni + MARCH_OWN - 1) / MARCH_OWN) * ((nc + rows - 1) / rows);
}
__host__ __device__ inline int march_blocks(int ni, int nc, int rows) {
  return (march_units(ni, nc, rows) + MARCH_WAVES - 1) / MARCH_WAVES;
}
void launch_march_residual(const Plan& p, int nblocks, const KnobsDev& k, int want_diag,
                           double circ_override, int refresh_ut, hipStream_t s);
void launch_march_jvp(const Plan& p, int nblocks, const KnobsDev& k, int vin, int vout, int mode,
                      int kidx, int use_diag, int gate, int prec, hipStream_t s);

// FP64 residual R(U); want_diag: also diag = vol/dtau at the case CFL;
// refresh_ut: also the FP32 copy of U for mixed-precision cases; rn partials
// always. circ_override: Gamma to use (NaN = the case's stored Gamma).
void launch_residual(const Plan& ps, int shape, int nblocks, const KnobsDev& k, int want_diag,
                     double circ_override, int refresh_ut, hipStream_t s);
// Tangent-linear in precision prec (0 fp64, 1 fp32), J including dGamma/dU:
// mode 0: out = scale * (diag v + J v), scale = s[kidx] (kidx >= 0) or 1;
// mode 1: out = -R - (diag v + J v). use_diag = 0 drops diag v.
void launch_jvp(const Plan& ps, int shape, int nblocks, const KnobsDev& k, int vin, int vout,
                int mode, int kidx, int use_diag, int gate, int prec, hipStream_t s);
void launch_forces(const Plan& pc, double* cp_out, hipStream_t s);
// kprev >= 0: finish Arnoldi step kprev first (as launch_normalize would).
void launch_dcirc(const Plan& pc, int vin, int gate, int kprev, const KnobsDev& kn, int prec,
                  hipStream_t s);
// march: the residual ran as the column march (its partials are per unit).
void launch_rn_finalize(const Plan& pc, int march, hipStream_t s);
void launch_hllc_test(int n, const double* WL, const double* WR, const double* S, double* F,
                      hipStream_t s);

// ---- precond.hip --------------------------------------------------------------
void launch_assemble(const Plan& pcell, int nblocks, const KnobsDev& k, int prec, hipStream_t s);
void launch_factor(const Plan& pl, int nblocks, int prec, hipStream_t s);
// One line-Jacobi sweep, one kernel: x_out = L^-1 (src - Bim x(i-1) - Bip x(i+1))
// with x = xprev, the neighbour terms only when coupled. src = VEC_NEG_R reads
// -R. The first sweep of an application keeps S0 src in the case's g0 buffer;
// reuse_g0 = 1 (a later sweep, same src) reads it instead of src and S0.
// want_norm: per-block partials of |xout|^2. hb (prec 1 only): couplings from
// the FP16 copy. Plan pl: pl.lines lines per block. Norm partials: one per
// 8-line group (wavefront), part[group].
void launch_line_sweep(const Plan& pl, int nblocks, int src, int xprev, int xout, int coupled,
                       int reuse_g0, int want_norm, int gate, int prec, int hb, hipStream_t s);
// One half sweep of the red-black ordering: lines i = color (mod 2) only,
// in place in xout (neighbours of the other colour are read from xout);
// coupled = 0 drops the i-coupling (the very first half sweep); reuse_g0 as
// above (the sweeps after the first). Norm partials go to
// part[(color * nlinerw + group)]. Plan pl: the colour's lines only.
void launch_line_sweep_rb(const Plan& pl, int nblocks, int src, int xout, int color, int coupled,
                          int reuse_g0, int want_norm, int gate, int prec, int hb, hipStream_t s);

// ---- krylov.hip -----------------------------------------------------------------
// bh (prec 1 only): the basis is stored in FP16 (wvec is then an FP32 work
// vector that the update reads; it writes V_{k+1} and vc).
void launch_dots(const Plan& pv, int nblocks, int k, int wvec, int gate, int prec, int bh, hipStream_t s);
// H[:, k] from the dot products' partials into the case scalars.
void launch_reduce_dots(const Plan& pc, int k, int bh, hipStream_t s);
// red: the update forms H[:, k] itself (no launch_reduce_dots before it).
void launch_update(const Plan& pv, int nblocks, int k, int wvec, int gate, int red, int prec, int bh,
                   hipStream_t s);
void launch_normalize(const Plan& pc, int k, const KnobsDev& kn, int prec, hipStream_t s);
void launch_krylov_begin(const Plan& pc, int stage, const KnobsDev& kn, int prec, hipStream_t s);
void launch_lstsq(const Plan& pc, hipStream_t s);
// split: x0 and the basis are colour-split (red-black line order).
void launch_dupdate(const Plan& pv, int nblocks, int prec, int bh, int split, hipStream_t s);
// FP16 basis: V_raw_0 = vsig * P (and vc) after the V_0 sweep.
void launch_store_v0(const Plan& pv, int nblocks, hipStream_t s);
void launch_limiter(const Plan& pc, const KnobsDev& kn, hipStream_t s);
// Positivity check of U + omega dU (knob `positivity`): per-block counts of
// cells with rho <= 0 or p <= 0, then a per-case pass that rejects the step
// when any cell would be non-positive.
void launch_positivity(const Plan& pv, int nblocks, const Plan& pc, hipStream_t s);
void launch_apply(const Plan& pv, int nblocks, hipStream_t s);
void launch_set_cfl(const Plan& pc, const double* cfl, hipStream_t s);
void launch_gather(const Plan& pc, double* out, hipStream_t s);
void launch_copy_vec(const Plan& pv, int nblocks, int src, int dst, int prec, hipStream_t s);

// ---- explicit.hip ---------------------------------------------------------------
// The explicit scheme's kernels (FP64 cases). Plans: prow, one block per four
// rows (nc blocks per case); pcol, NT-thread blocks over the 4 ni columns;
// cell plans CELL_BLOCK cells per block.
// Elements per lane of the row kernel for rows of up to ni_max cells, and the
// LDS bytes its workgroup then takes.
constexpr int EXPLICIT_BLOCK = 256;   // threads per workgroup of the explicit kernels
// 0 when rows are longer than 64 RK_ROWS_CM_MAX cells (not supported).
constexpr int RK_ROWS_CM_MAX = 32;
int rk_rows_cm(int ni_max);
size_t rk_rows_lds(int cm);
// Smoothing along j of R (+ the forcing in xa) into w. The segmented kernel
// (columns of up to 8 * IRS_J_SEG_MAX cells; plan pseg, 32 columns per block)
// keeps each column in registers; seg_len = irs_j_seg_len(largest nc), 0 when
// the columns are too long, then launch_irs_j (plan pcol) instead, which keeps
// the first pass's values in w.
constexpr int IRS_J_SEGMENTS = 8, IRS_J_COLS_PER_BLOCK = 32, IRS_J_SEG_MAX = 64;
int irs_j_seg_len(int nc_max);
void launch_irs_j(const Plan& pcol, int nblocks, double eps, int forcing, hipStream_t s);
void launch_irs_j_seg(const Plan& pseg, int nblocks, int seg_len, double eps, int forcing, hipStream_t s);
// One Runge-Kutta stage, U = U0 - a x / diag: x the smoothed residual (from w,
// smoothing along i; eps > 0) or R (+ P when forcing; eps = 0). first: U0 =
// U, kept in dU; want_res: the residual partials of the stage.
void launch_rk_rows(const Plan& prow, int nblocks, int cm, double eps, double a, int first, int want_res,
                    int forcing, hipStream_t s);
void launch_xres(const Plan& pc, hipStream_t s);
void launch_level_begin(const Plan& pc, const int* top_slots, double cfl, int zero_circ, hipStream_t s);
void launch_restrict(const Plan& pcc, int nblocks, const int* fine_slots, int fine_forcing, hipStream_t s);
void launch_forcing(const Plan& pcc, int nblocks, hipStream_t s);
void launch_prolong(const Plan& pfc, int nblocks, const int* coarse_slots, hipStream_t s);
void launch_xgather(const Plan& pc, double* out, hipStream_t s);

}  // namespace a2d
