// Kernels of the explicit scheme (aero2d/euler.py:298-381): the four-stage
// Runge-Kutta update with local time steps, implicit residual smoothing, and
// the FAS multigrid transfers. FP64 throughout (cases of the FP64 form).
//
// Residual smoothing solves (1 + 2 eps) x_i - eps (x_{i-1} + x_{i+1}) = r_i,
// periodic along i, and the same along j with 1 + eps on both end rows. The
// reference solves the i system by FFT; here the constant periodic operator is
// factored instead,
//   (1 + 2 eps) - eps (S + S^-1) = (eps / lam) (1 - lam S) (1 - lam S^-1),
//   lam = 2 eps / ((1 + 2 eps) + sqrt(1 + 4 eps))  (the root below one),
// so x = (lam / eps) (1 - lam S^-1)^-1 (1 - lam S)^-1 r: two periodic
// first-order recursions, y_i = r_i + lam y_{i-1} and z_i = y_i + lam z_{i+1}.
// One wavefront solves one row: each lane runs its chunk of the row with a zero
// carry, which gives the affine map carry-in -> carry-out of the chunk; a
// circular scan of those maps over the 64 lanes gives, for every lane, the
// map once around the whole row, whose fixed point is the exact periodic
// carry (division by 1 - lam^ni); the lanes then rerun their chunks with it.
// The result is the periodic solution to round-off, with no truncation.
// The j system is a plain tridiagonal elimination whose coefficients depend on
// j alone; they are formed once on the host and read from the case's slab.
#include "common.hpp"
#include "launch.hpp"

#include <cmath>

namespace a2d {

namespace {

constexpr int NT = EXPLICIT_BLOCK;
constexpr int ROWS = NT / WAVE;   // rows (one wavefront each) per workgroup
static_assert(ROWS == 4, "the row plan has one block per four rows");
static_assert(CELL_BLOCK == NT, "cell plans have one thread per cell");
constexpr int JSEG = IRS_J_SEGMENTS;        // row segments per column (segmented j smoothing)
constexpr int JCOLS = WAVE / JSEG;          // columns per wavefront
static_assert(ROWS * JCOLS == IRS_J_COLS_PER_BLOCK, "the segmented column plan has 32 columns per block");

// Smoothing along j, for columns of up to 8 L cells: each column is split into
// 8 segments of L rows held by 8 lanes (lane = segment * 8 + column, so that
// the 8 lanes of one segment read 8 neighbouring columns). Both passes of the
// elimination, y_j = m_j (r_j + eps y_{j-1}) and x_j = y_j - c_j x_{j+1}, are
// first-order linear recursions: each lane runs its segment with a zero carry
// (the affine map carry-in -> carry-out of the segment), a scan of those maps
// over the 8 segments gives every segment its carry, and the lanes rerun their
// segments with it. The column stays in registers between the passes: R (+ P)
// is read once and the smoothed residual written once. r = R (+ P: the FAS
// forcing held in xa, coarse levels).
template <int L>
__global__ void __launch_bounds__(NT) k_irs_j_seg(Plan plan, double eps, int forcing) {
  const int e = plan_entry(plan, blockIdx.x);
  const CaseDev& cd = plan.cases[plan.slots[e]];
  const int ni = cd.ni, nc = cd.nc, n = cd.n;
  const int wave = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
  const int seg = lane / JCOLS, col = lane - seg * JCOLS;
  const int t = (blockIdx.x - plan.blk_prefix[e]) * (ROWS * JCOLS) + wave * JCOLS + col;
  const bool live = t < 4 * ni;
  const int k = live ? t / ni : 0, i = live ? t - k * ni : 0;
  const size_t base = size_t(k) * n + i;
  const int j0 = seg * L;
  const int len = live ? max(0, min(L, nc - j0)) : 0;
  const gptr<const double> R = gp(cd.R), P = gp(ctp<double>(cd.xa));
  const gptr<double> w = gp(tp<double>(cd.w));
  const gptr<const double> m = gp(cd.irs), c = gp(cd.irs + nc);
  double r[L];
#pragma unroll
  for (int q = 0; q < L; ++q) {
    r[q] = 0.0;
    if (q < len) {
      r[q] = R[base + size_t(j0 + q) * ni];
      if (forcing) r[q] += P[base + size_t(j0 + q) * ni];
    }
  }
  // forward: y_j = m_j r_j + (eps m_j) y_{j-1}
  double A = 1.0, B = 0.0;
#pragma unroll
  for (int q = 0; q < L; ++q)
    if (q < len) {
      const double mj = m[j0 + q];
      B = mj * (r[q] + eps * B);
      A *= eps * mj;
    }
#pragma unroll
  for (int d = 1; d < JSEG; d <<= 1) {
    const int src = (lane - d * JCOLS) & (WAVE - 1);
    const double Ap = __shfl(A, src, WAVE), Bp = __shfl(B, src, WAVE);
    if (seg >= d) {
      B = A * Bp + B;
      A = A * Ap;
    }
  }
  double y = __shfl(B, (lane - JCOLS) & (WAVE - 1), WAVE);
  if (seg == 0) y = 0.0;
#pragma unroll
  for (int q = 0; q < L; ++q)
    if (q < len) {
      y = m[j0 + q] * (r[q] + eps * y);
      r[q] = y;
    }
  // backward: x_j = y_j - c_j x_{j+1}
  A = 1.0;
  B = 0.0;
#pragma unroll
  for (int q = L - 1; q >= 0; --q)
    if (q < len) {
      const double cj = c[j0 + q];
      B = r[q] - cj * B;
      A *= -cj;
    }
#pragma unroll
  for (int d = 1; d < JSEG; d <<= 1) {
    const int src = (lane + d * JCOLS) & (WAVE - 1);
    const double Ap = __shfl(A, src, WAVE), Bp = __shfl(B, src, WAVE);
    if (seg + d < JSEG) {
      B = A * Bp + B;
      A = A * Ap;
    }
  }
  double x = __shfl(B, (lane + JCOLS) & (WAVE - 1), WAVE);
  if (seg == JSEG - 1) x = 0.0;
#pragma unroll
  for (int q = L - 1; q >= 0; --q)
    if (q < len) {
      x = r[q] - c[j0 + q] * x;
      r[q] = x;
    }
#pragma unroll
  for (int q = 0; q < L; ++q)
    if (q < len) w[base + size_t(j0 + q) * ni] = r[q];
}

// Smoothing along j for longer columns: one thread per (component, column i),
// forward elimination into w, back substitution in place.
__global__ void __launch_bounds__(NT) k_irs_j(Plan plan, double eps, int forcing) {
  const int e = plan_entry(plan, blockIdx.x);
  const CaseDev& cd = plan.cases[plan.slots[e]];
  const int ni = cd.ni, nc = cd.nc, n = cd.n;
  const int t = (blockIdx.x - plan.blk_prefix[e]) * NT + threadIdx.x;
  if (t >= 4 * ni) return;
  const int k = t / ni, i = t - k * ni;
  const size_t base = size_t(k) * n + i;
  const gptr<const double> R = gp(cd.R), P = gp(ctp<double>(cd.xa));
  const gptr<double> w = gp(tp<double>(cd.w));
  const gptr<const double> m = gp(cd.irs), c = gp(cd.irs + nc);
  constexpr int B = 8;              // rows loaded together ahead of the recursion
  double y = 0.0;
  for (int j0 = 0; j0 < nc; j0 += B) {
    double r[B];
#pragma unroll
    for (int q = 0; q < B; ++q) {
      const int j = j0 + q;
      r[q] = 0.0;
      if (j < nc) {
        r[q] = R[base + size_t(j) * ni];
        if (forcing) r[q] += P[base + size_t(j) * ni];
      }
    }
#pragma unroll
    for (int q = 0; q < B; ++q) {
      const int j = j0 + q;
      if (j < nc) {
        y = (r[q] + eps * y) * m[j];
        w[base + size_t(j) * ni] = y;
      }
    }
  }
  double x = y;                     // row nc - 1
  for (int j1 = nc - 2; j1 >= 0; j1 -= B) {
    double d[B];
#pragma unroll
    for (int q = 0; q < B; ++q) {
      const int j = j1 - q;
      d[q] = j >= 0 ? w[base + size_t(j) * ni] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < B; ++q) {
      const int j = j1 - q;
      if (j >= 0) {
        x = d[q] - c[j] * x;
        w[base + size_t(j) * ni] = x;
      }
    }
  }
}

// Circular inclusive scan of affine maps v -> A v + B over the wavefront.
// DIR = +1: lane l composes the maps of lanes l, l-1, ..., l-63 (each applied
// after the one before it); DIR = -1: lanes l, l+1, ..., l+63.
template <int DIR>
__device__ __forceinline__ void circ_scan(double& A, double& B, int lane) {
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    const int src = (lane - DIR * d) & (WAVE - 1);
    const double Ap = __shfl(A, src, WAVE), Bp = __shfl(B, src, WAVE);
    B = A * Bp + B;
    A = A * Ap;
  }
}

// One Runge-Kutta stage over rows (component k, row j), one wavefront each:
// x = the smoothed residual (SMOOTH: the i solve of the j-smoothed w; else
// R + P), then U = U0 - a x / diag with diag = vol / dt, U0 read from U and
// kept in dU at the first stage. want_res (last stage): per-row partials of
// ((U - U0)_rho / dt)^2 in partial column 1. CM: elements per lane (a power of
// two with 64 CM >= ni); the row passes through LDS, padded one element per
// chunk, so that loads and stores stay coalesced.
template <int CM, bool SMOOTH>
__global__ void __launch_bounds__(NT) k_rk_rows(Plan plan, double lam, double scale, double a, int first,
                                               int want_res, int forcing) {
  extern __shared__ double rows_lds[];
  const int e = plan_entry(plan, blockIdx.x);
  const CaseDev& cd = plan.cases[plan.slots[e]];
  const int ni = cd.ni, nc = cd.nc, n = cd.n;
  const int wave = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
  const int row = (blockIdx.x - plan.blk_prefix[e]) * ROWS + wave;   // < 4 nc: 4 nc rows, nc blocks
  const int k = row / nc, j = row - k * nc;
  const size_t base = size_t(k) * n + size_t(j) * ni;
  double* s = rows_lds + wave * (CM + 1) * WAVE;

  if constexpr (SMOOTH) {
    const gptr<const double> src = gp(ctp<double>(cd.w));
    for (int i = lane; i < ni; i += WAVE) s[i + i / CM] = src[base + i];
    __syncthreads();
    const int beg = lane * CM;
    const int len = max(0, min(CM, ni - beg));
    double r[CM];
#pragma unroll
    for (int t = 0; t < CM; ++t) r[t] = t < len ? s[lane * (CM + 1) + t] : 0.0;
    // y_i = r_i + lam y_{i-1}, periodic
    double A = 1.0, B = 0.0;
#pragma unroll
    for (int t = 0; t < CM; ++t)
      if (t < len) {
        B = lam * B + r[t];
        A *= lam;
      }
    circ_scan<1>(A, B, lane);
    double y = __shfl(B / (1.0 - A), (lane - 1) & (WAVE - 1), WAVE);
#pragma unroll
    for (int t = 0; t < CM; ++t)
      if (t < len) {
        y = lam * y + r[t];
        r[t] = y;
      }
    // z_i = y_i + lam z_{i+1}, periodic
    A = 1.0;
    B = 0.0;
#pragma unroll
    for (int t = CM - 1; t >= 0; --t)
      if (t < len) {
        B = lam * B + r[t];
        A *= lam;
      }
    circ_scan<-1>(A, B, lane);
    double z = __shfl(B / (1.0 - A), (lane + 1) & (WAVE - 1), WAVE);
#pragma unroll
    for (int t = CM - 1; t >= 0; --t)
      if (t < len) {
        z = lam * z + r[t];
        r[t] = z;
      }
#pragma unroll
    for (int t = 0; t < CM; ++t)
      if (t < len) s[lane * (CM + 1) + t] = scale * r[t];
    __syncthreads();
  }

  const gptr<double> U = gp(cd.U), U0 = gp(cd.dU);
  const gptr<const double> R = gp(cd.R), P = gp(ctp<double>(cd.xa));
  const gptr<const double> dg = gp(ctp<double>(cd.diag)), vol = gp(cd.vol);
  const bool res = want_res && k == 0;
  double part = 0.0;
  for (int i = lane; i < ni; i += WAVE) {
    const size_t idx = base + i;
    const int cell = j * ni + i;
    double x;
    if constexpr (SMOOTH) x = s[i + i / CM];
    else x = forcing ? R[idx] + P[idx] : R[idx];
    const double u0 = first ? U[idx] : U0[idx];
    if (first) U0[idx] = u0;
    const double d = dg[cell];
    const double un = u0 - a * (x / d);
    U[idx] = un;
    if (res) {
      const double q = (un - u0) * d / vol[cell];
      part += q * q;
    }
  }
  if (res) {
    part = wave_sum(part);
    if (lane == 0) part_at(cd, j, 1) = part;
  }
}

// res = sqrt(mean(((U - U0)_rho / dt)^2)) of the last stage, per case.
__global__ void __launch_bounds__(WAVE) k_xres(Plan pc) {
  const CaseDev& cd = pc.cases[pc.slots[blockIdx.x]];
  const double s = case_sum(cd, cd.nc, 1);
  if (threadIdx.x == 0) cd.sc->xres = sqrt(s / cd.n);
}

// Start of a cycle on one grid level: the explicit CFL, and the freestream and
// circulation of the finest level's case (zero_circ: Gamma = 0, the start of a
// solve).
__global__ void k_level_begin(Plan pc, const int* top_slots, double cfl, int zero_circ) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= pc.nb) return;
  CaseScal* sc = pc.cases[pc.slots[e]].sc;
  const CaseScal* top = pc.cases[top_slots[e]].sc;
  const double circ = zero_circ ? 0.0 : top->circ;
  double w[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) w[k] = top->winf[k];
#pragma unroll
  for (int k = 0; k < 4; ++k) sc->winf[k] = w[k];
  sc->circ = circ;
  sc->cfl = cfl;
}

// Restriction to the next coarser level (aero2d/euler.py:372-374): the
// volume-weighted mean of the four fine cells as the coarse state (into U and,
// kept for the correction, xb) and the sum of the fine R + P (into xa; the
// coarse residual of that state is subtracted afterwards).
__global__ void __launch_bounds__(NT) k_restrict(Plan pcc, const int* fine_slots, int fine_forcing) {
  const int e = plan_entry(pcc, blockIdx.x);
  const CaseDev& cc = pcc.cases[pcc.slots[e]];
  const CaseDev& cf = pcc.cases[fine_slots[e]];
  const int c = (blockIdx.x - pcc.blk_prefix[e]) * CELL_BLOCK + threadIdx.x;
  if (c >= cc.n) return;
  const int J = c / cc.ni, I = c - J * cc.ni;
  const int nif = cf.ni, nf = cf.n, nn = cc.n;
  const int f00 = 2 * J * nif + 2 * I, f10 = f00 + 1, f01 = f00 + nif, f11 = f01 + 1;
  const double v00 = cf.vol[f00], v10 = cf.vol[f10], v01 = cf.vol[f01], v11 = cf.vol[f11];
  const double vs = v00 + v10 + v01 + v11;
  const double* Uf = cf.U;
  const double* Rf = cf.R;
  const double* Pf = tp<double>(cf.xa);
  double* xb = tp<double>(cc.xb);
  double* xa = tp<double>(cc.xa);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const size_t o = size_t(k) * nf;
    const double u = (Uf[o + f00] * v00 + Uf[o + f10] * v10 + Uf[o + f01] * v01 + Uf[o + f11] * v11) / vs;
    cc.U[size_t(k) * nn + c] = u;
    xb[size_t(k) * nn + c] = u;
    double r00 = Rf[o + f00], r10 = Rf[o + f10], r01 = Rf[o + f01], r11 = Rf[o + f11];
    if (fine_forcing) {
      r00 += Pf[o + f00];
      r10 += Pf[o + f10];
      r01 += Pf[o + f01];
      r11 += Pf[o + f11];
    }
    xa[size_t(k) * nn + c] = r00 + r10 + r01 + r11;
  }
}

// The coarse forcing P_c = sum4(R + P) - R_c(U_c0) (aero2d/euler.py:375-376).
__global__ void __launch_bounds__(NT) k_forcing(Plan pcc) {
  const int e = plan_entry(pcc, blockIdx.x);
  const CaseDev& cc = pcc.cases[pcc.slots[e]];
  const int c = (blockIdx.x - pcc.blk_prefix[e]) * CELL_BLOCK + threadIdx.x;
  if (c >= cc.n) return;
  double* xa = tp<double>(cc.xa);
#pragma unroll
  for (int k = 0; k < 4; ++k) xa[size_t(k) * cc.n + c] -= cc.R[size_t(k) * cc.n + c];
}

// Coarse correction (aero2d/euler.py:352-364, 378): U_f += bilinear
// interpolation (weights 3/4, 1/4; periodic along i, end rows held along j) of
// U_c - U_c0, interpolated along i first, then along j.
__global__ void __launch_bounds__(NT) k_prolong(Plan pfc, const int* coarse_slots) {
  const int e = plan_entry(pfc, blockIdx.x);
  const CaseDev& cf = pfc.cases[pfc.slots[e]];
  const CaseDev& cc = pfc.cases[coarse_slots[e]];
  const int c = (blockIdx.x - pfc.blk_prefix[e]) * CELL_BLOCK + threadIdx.x;
  if (c >= cf.n) return;
  const int nif = cf.ni, nic = cc.ni, ncc = cc.nc, nn = cc.n;
  const int j = c / nif, i = c - j * nif;
  const int I = i >> 1, J = j >> 1;
  const int In = (i & 1) ? (I + 1 == nic ? 0 : I + 1) : (I == 0 ? nic - 1 : I - 1);
  const int Jn = (j & 1) ? min(J + 1, ncc - 1) : max(J - 1, 0);
  const double* Uc = cc.U;
  const double* Uc0 = tp<double>(cc.xb);
  const int a0 = J * nic + I, a1 = J * nic + In, b0 = Jn * nic + I, b1 = Jn * nic + In;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const size_t o = size_t(k) * nn;
    const double dJ = 0.75 * (Uc[o + a0] - Uc0[o + a0]) + 0.25 * (Uc[o + a1] - Uc0[o + a1]);
    const double dN = 0.75 * (Uc[o + b0] - Uc0[o + b0]) + 0.25 * (Uc[o + b1] - Uc0[o + b1]);
    cf.U[size_t(k) * cf.n + c] += 0.75 * dJ + 0.25 * dN;
  }
}

// Per-case results of a run of cycles, 5 doubles per batch entry: res of the
// last cycle, Cl, Cd, Cm, Gamma.
__global__ void k_xgather(Plan pc, double* out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= pc.nb) return;
  const CaseScal* sc = pc.cases[pc.slots[e]].sc;
  double* o = out + size_t(e) * 5;
  o[0] = sc->xres;
  o[1] = sc->cl;
  o[2] = sc->cd;
  o[3] = sc->cm;
  o[4] = sc->circ;
}

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

template <int CM>
void run_rows(const Plan& p, int nblocks, double lam, double scale, double a, int first, int want_res,
              hipStream_t s) {
  const size_t lds = size_t(ROWS) * (CM + 1) * WAVE * sizeof(double);
  hipLaunchKernelGGL((k_rk_rows<CM, true>), dim3(nblocks), dim3(NT), lds, s, p, lam, scale, a, first, want_res,
                     0);
}

}  // namespace

int rk_rows_cm(int ni_max) {
  int cm = 1;
  while (cm * WAVE < ni_max) cm <<= 1;
  return cm <= RK_ROWS_CM_MAX ? cm : 0;
}
size_t rk_rows_lds(int cm) { return size_t(ROWS) * (cm + 1) * WAVE * sizeof(double); }

int irs_j_seg_len(int nc_max) {
  int l = 1;
  while (l * JSEG < nc_max) l <<= 1;
  return l <= IRS_J_SEG_MAX ? l : 0;
}

void launch_irs_j(const Plan& pcol, int nblocks, double eps, int forcing, hipStream_t s) {
  hipLaunchKernelGGL(k_irs_j, dim3(nblocks), dim3(NT), 0, s, pcol, eps, forcing);
}

void launch_irs_j_seg(const Plan& pseg, int nblocks, int seg_len, double eps, int forcing, hipStream_t s) {
  switch (seg_len) {
#define A2D_SEG(LL)                                                                                   \
  case LL: hipLaunchKernelGGL(k_irs_j_seg<LL>, dim3(nblocks), dim3(NT), 0, s, pseg, eps, forcing); break;
    A2D_SEG(1) A2D_SEG(2) A2D_SEG(4) A2D_SEG(8) A2D_SEG(16) A2D_SEG(32) A2D_SEG(64)
#undef A2D_SEG
    default: break;
  }
}

void launch_rk_rows(const Plan& prow, int nblocks, int cm, double eps, double a, int first, int want_res,
                    int forcing, hipStream_t s) {
  if (eps > 0.0) {
    const double lam = 2.0 * eps / ((1.0 + 2.0 * eps) + std::sqrt(1.0 + 4.0 * eps));
    const double scale = lam / eps;
    switch (cm) {
      case 1: run_rows<1>(prow, nblocks, lam, scale, a, first, want_res, s); break;
      case 2: run_rows<2>(prow, nblocks, lam, scale, a, first, want_res, s); break;
      case 4: run_rows<4>(prow, nblocks, lam, scale, a, first, want_res, s); break;
      case 8: run_rows<8>(prow, nblocks, lam, scale, a, first, want_res, s); break;
      case 16: run_rows<16>(prow, nblocks, lam, scale, a, first, want_res, s); break;
      case 32: run_rows<32>(prow, nblocks, lam, scale, a, first, want_res, s); break;
      default: break;   // the engine refuses longer rows before launching
    }
  } else {
    hipLaunchKernelGGL((k_rk_rows<1, false>), dim3(nblocks), dim3(NT), 0, s, prow, 0.0, 0.0, a, first, want_res,
                       forcing);
  }
}

void launch_xres(const Plan& pc, hipStream_t s) {
  hipLaunchKernelGGL(k_xres, dim3(pc.nb), dim3(WAVE), 0, s, pc);
}
void launch_level_begin(const Plan& pc, const int* top_slots, double cfl, int zero_circ, hipStream_t s) {
  hipLaunchKernelGGL(k_level_begin, dim3(cdiv(pc.nb, 64)), dim3(64), 0, s, pc, top_slots, cfl, zero_circ);
}
void launch_restrict(const Plan& pcc, int nblocks, const int* fine_slots, int fine_forcing, hipStream_t s) {
  hipLaunchKernelGGL(k_restrict, dim3(nblocks), dim3(NT), 0, s, pcc, fine_slots, fine_forcing);
}
void launch_forcing(const Plan& pcc, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_forcing, dim3(nblocks), dim3(NT), 0, s, pcc);
}
void launch_prolong(const Plan& pfc, int nblocks, const int* coarse_slots, hipStream_t s) {
  hipLaunchKernelGGL(k_prolong, dim3(nblocks), dim3(NT), 0, s, pfc, coarse_slots);
}
void launch_xgather(const Plan& pc, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_xgather, dim3(cdiv(pc.nb, 64)), dim3(64), 0, s, pc, out);
}

}  // namespace a2d
