// Cell-face stencil kernels of the implicit Euler path:
//   * the second-order finite-volume residual R(U) (aero2d/euler.py:259-287),
//     FP64, fused with the local pseudo-time step (aero2d/euler.py:289-296,
//     aero2d/implicit.py:173-177) and the density-residual norm partials;
//   * its exact tangent-linear J v (replaces jax.jvp at aero2d/implicit.py:180)
//     from the same code instantiated over Dual<T>;
//   * per-case wall reductions: forces / circulation (aero2d/euler.py:318-331,
//     aero2d/implicit.py:160-164) and the circulation tangent.
//
// Tiling: a workgroup of NTS threads owns TI x TJ cells. It stages the
// primitive state of the (TI+4) x (TJ+4) halo in LDS (MUSCL needs two cells
// each side; wall mirror ghosts and far-field ghosts are produced while
// loading), evaluates every i-face and j-face flux of the tile exactly once
// into LDS, then differences them. HBM traffic is one read of U (and v),
// the face metrics and one write of the result; neighbouring tiles re-read
// their 2-cell halos from L2.
#include "arnoldi.hpp"
#include "common.hpp"
#include "launch.hpp"
#include "physics.hpp"

#include <cmath>
#include <type_traits>

namespace a2d {

namespace {

constexpr int NT = 256;             // threads of the per-case wall reductions
#ifndef A2D_STENCIL_THREADS
#define A2D_STENCIL_THREADS 512
#endif
// Threads of a stencil tile's workgroup (a multiple of 256, so that its waves
// spread evenly over the CU's four SIMDs). The tile's passes (halo, slopes,
// faces, cells) are strided loops over this count. The LDS tile limits the
// workgroups per CU (2 for the FP64 JVP, 3 for the residual, 4 for the FP32
// JVP), so 512 threads doubles the resident waves of 256: measured on the
// default batch, FP64 JVP 286 -> 242 us, FP32 JVP 129 -> 112 us, residual
// unchanged (179 -> 184 us); 320, 384, 768 and 1024 threads were slower.
constexpr int NTS = A2D_STENCIL_THREADS;

template <class T>
struct CaseView {
  gptr<const T> U;     // base state (conserved)
  gptr<const T> vol;
  gptr<const T> Si;
  gptr<const T> Sj;
  gptr<const T> xf;
  gptr<const T> yf;
};

template <class T>
__device__ __forceinline__ CaseView<T> view(const CaseDev& cd) {
  if constexpr (std::is_same<T, double>::value) {
    return {gp(cd.U), gp(cd.vol), gp(cd.Si), gp(cd.Sj), gp(cd.xf), gp(cd.yf)};
  } else {
    return {gp(ctp<T>(cd.Ut)), gp(ctp<T>(cd.volt)), gp(ctp<T>(cd.Sit)), gp(ctp<T>(cd.Sjt)),
            gp(ctp<T>(cd.xft)), gp(ctp<T>(cd.yft))};
  }
}

// Primitive state of cell (i, j) as scalar type S (tangent from v when dual).
// idx: the cell in the state's natural order; vidx: the same cell in the
// tangent vector's order (colour-split under the red-black line order).
template <class S, class T>
__device__ __forceinline__ W4<S> load_prim(gptr<const T> U, gptr<const T> v, int n, int idx, int vidx) {
  if constexpr (std::is_same<S, T>::value) {
    return cons_to_prim<S>(U[idx], U[n + idx], U[2 * n + idx], U[3 * n + idx]);
  } else {
    return cons_to_prim<S>(S(U[idx], v[vidx]), S(U[n + idx], v[n + vidx]),
                           S(U[2 * n + idx], v[2 * n + vidx]), S(U[3 * n + idx], v[3 * n + vidx]));
  }
}

// The stencil kernel. TAN = false: FP64 residual. TAN = true: tangent-linear
// in precision T.
template <class T, bool TAN, int TI, int TJ>
__global__ void __launch_bounds__(NTS) k_stencil(Plan plan, KnobsDev kn, int vin, int vout,
                                                int mode, int kidx, int use_diag, int want_diag,
                                                double circ_override, int refresh_ut, int gate) {
  using S = typename std::conditional<TAN, Dual<T>, T>::type;
  constexpr int HI = TI + 4, HJ = TJ + 4, NH = HI * HJ;
  constexpr int NFI = (TI + 1) * TJ, NFJ = TI * (TJ + 1);
  __shared__ T sW[4][HJ][HI];
  __shared__ T sD[TAN ? 4 : 1][TAN ? HJ : 1][TAN ? HI : 1];
  __shared__ T sFi[4][TJ][TI + 1];
  __shared__ T sFj[4][TJ + 1][TI];
  // limited slopes per cell: i-direction for cells i0-1..i0+TI of the tile's
  // rows, j-direction for rows j0-1..j0+TJ of its columns (value, tangent)
  __shared__ T sSi[TAN ? 2 : 1][4][TJ][TI + 2];
  __shared__ T sSj[TAN ? 2 : 1][4][TJ + 2][TI];
  __shared__ double sred[NTS / WAVE];

  const int e = plan_entry(plan, blockIdx.x);
  const CaseDev& cd = plan.cases[plan.slots[e]];
  const CaseScal* sc = cd.sc;
  if (sc->parked || (gate && !sc->active)) return;
  const int ni = cd.ni, nc = cd.nc, n = cd.n;
  const int tiles_i = (ni + TI - 1) / TI;
  const int lt = blockIdx.x - plan.blk_prefix[e];
  const int i0 = (lt % tiles_i) * TI, j0 = (lt / tiles_i) * TJ;
  const int tid = threadIdx.x;

  const CaseView<T> cv = view<T>(cd);
  const gptr<const T> vt = TAN ? vec_gptr<T>(cd, vin) : gptr<const T>(nullptr);
  const Inf<T> inf = {T(sc->winf[0]), T(sc->winf[1]), T(sc->winf[2]), T(sc->winf[3])};
  const double circ_d = (circ_override == circ_override) ? circ_override : sc->circ;
  const T vort = kn.vortex ? T(1) : T(0);
  S circ;
  if constexpr (TAN) circ = S(T(circ_d) * vort, T(sc->dcirc) * vort);
  else circ = S(T(circ_d) * vort);

  // ---- stage primitives of the halo in LDS --------------------------------
  for (int c = tid; c < NH; c += NTS) {
    const int hi = c % HI, hj = c / HI;
    const int i = wrap(i0 - 2 + hi, ni);
    const int j = j0 - 2 + hj;
    const int vi = bcol(i, ni, kn.precond_rb);     // the column in the Krylov vectors
    W4<S> w;
    if (j >= 0 && j < nc) {
      w = load_prim<S, T>(cv.U, vt, n, j * ni + i, j * ni + vi);
    } else if (j < 0) {
      // mirror ghosts: j = -1 mirrors cell 0, j = -2 mirrors cell 1
      int jm = (j == -1) ? 0 : 1;
      if (jm >= nc) jm = nc - 1;
      const T sx = cv.Sj[i], sy = cv.Sj[(nc + 1) * ni + i];
      const T area = sqrt(sx * sx + sy * sy);
      w = mirror(load_prim<S, T>(cv.U, vt, n, jm * ni + i, jm * ni + vi), sx / area, sy / area);
    } else if (j <= nc + 1) {
      // both far-field ghost layers equal the characteristic boundary state
      const T sx = cv.Sj[nc * ni + i], sy = cv.Sj[(nc + 1) * ni + nc * ni + i];
      w = farfield_ghost<S, T>(load_prim<S, T>(cv.U, vt, n, (nc - 1) * ni + i, (nc - 1) * ni + vi), sx, sy, inf,
                               cv.xf[i], cv.yf[i], circ);
    } else {
      w = {K<S>(1.0), K<S>(0.0), K<S>(0.0), K<S>(1.0)};  // outside the grid, never read
    }
    sW[0][hj][hi] = val(w.r);
    sW[1][hj][hi] = val(w.u);
    sW[2][hj][hi] = val(w.v);
    sW[3][hj][hi] = val(w.p);
    if constexpr (TAN) {
      sD[0][hj][hi] = tan_of(w.r);
      sD[1][hj][hi] = tan_of(w.u);
      sD[2][hj][hi] = tan_of(w.v);
      sD[3][hj][hi] = tan_of(w.p);
    }
  }
  __syncthreads();

  // limiter thresholds eps_k = (lim_k * scale_k)^2 + 1e-20, scale = (rho, M, M, p)_inf
  const T mach = sqrt(inf.u * inf.u + inf.v * inf.v);
  T eps[4];
  {
    const T scl[4] = {inf.r, mach, mach, inf.p};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const T a = T(kn.lim_k) * scl[k];
      eps[k] = a * a + T(1e-20);
    }
  }

  // ---- limited slopes, once per cell and direction (MUSCL, van Albada) -------
  // The left state of a face is W0 + slope(W0)/2 and its right state
  // Wp - slope(Wp)/2 (aero2d/euler.py:211-215); each cell's slope serves the
  // faces on both of its sides.
  if (kn.order == 2) {
    constexpr int NSI = (TI + 2) * TJ, NSJ = TI * (TJ + 2);
    for (int c = tid; c < NSI + NSJ; c += NTS) {
      const bool isI = c < NSI;
      int hi, hj, a0, a1;
      if (isI) {
        a0 = c % (TI + 2);       // slope index, cell hi = a0 + 1
        a1 = c / (TI + 2);       // tile row, hj = a1 + 2
        hi = a0 + 1;
        hj = a1 + 2;
      } else {
        const int g = c - NSI;
        a0 = g % TI;             // tile column, hi = a0 + 2
        a1 = g / TI;             // slope index, cell hj = a1 + 1
        hi = a0 + 2;
        hj = a1 + 1;
      }
      const int di = isI ? 1 : 0, dj = isI ? 0 : 1;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        S wm, w0, wp;
        if constexpr (TAN) {
          wm = S(sW[k][hj - dj][hi - di], sD[k][hj - dj][hi - di]);
          w0 = S(sW[k][hj][hi], sD[k][hj][hi]);
          wp = S(sW[k][hj + dj][hi + di], sD[k][hj + dj][hi + di]);
        } else {
          wm = sW[k][hj - dj][hi - di];
          w0 = sW[k][hj][hi];
          wp = sW[k][hj + dj][hi + di];
        }
        const S sl = van_albada(w0 - wm, wp - w0, eps[k]);
        if (isI) {
          sSi[0][k][a1][a0] = val(sl);
          if constexpr (TAN) sSi[TAN ? 1 : 0][k][a1][a0] = tan_of(sl);
        } else {
          sSj[0][k][a1][a0] = val(sl);
          if constexpr (TAN) sSj[TAN ? 1 : 0][k][a1][a0] = tan_of(sl);
        }
      }
    }
    __syncthreads();
  }

  // ---- face fluxes ----------------------------------------------------------
  for (int f = tid; f < NFI + NFJ; f += NTS) {
    const bool isI = f < NFI;
    int li, lj, h0i, h0j, di, dj;
    T sx, sy;
    bool valid;
    if (isI) {
      li = f % (TI + 1);
      lj = f / (TI + 1);
      const int j = j0 + lj, i = i0 + li;
      valid = (j < nc) && (i <= ni);
      const int iw = (i >= ni) ? i - ni : i;
      h0i = li; h0j = lj + 2; di = 1; dj = 0;
      sx = valid ? cv.Si[j * ni + iw] : T(1);
      sy = valid ? cv.Si[n + j * ni + iw] : T(0);
    } else {
      const int g = f - NFI;
      li = g % TI;
      lj = g / TI;
      const int j = j0 + lj, i = i0 + li;
      valid = (i < ni) && (j <= nc);
      h0i = li + 2; h0j = lj; di = 0; dj = 1;
      sx = valid ? cv.Sj[j * ni + i] : T(1);
      sy = valid ? cv.Sj[(nc + 1) * ni + j * ni + i] : T(0);
    }
    if (!valid) continue;
    S wl[4], wr[4];
    const int hLi = h0i + di, hLj = h0j + dj;          // cell left of the face
    const int hRi = h0i + 2 * di, hRj = h0j + 2 * dj;  // cell right of the face
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      S q1, q2;
      if constexpr (TAN) {
        q1 = S(sW[k][hLj][hLi], sD[k][hLj][hLi]);
        q2 = S(sW[k][hRj][hRi], sD[k][hRj][hRi]);
      } else {
        q1 = sW[k][hLj][hLi];
        q2 = sW[k][hRj][hRi];
      }
      if (kn.order == 2) {
        // slope indices: i-face li uses i-slopes li (left) and li+1 (right) of
        // row lj; j-face lj uses j-slopes lj and lj+1 of column li
        S s1, s2;
        if (isI) {
          if constexpr (TAN) {
            s1 = S(sSi[0][k][lj][li], sSi[TAN ? 1 : 0][k][lj][li]);
            s2 = S(sSi[0][k][lj][li + 1], sSi[TAN ? 1 : 0][k][lj][li + 1]);
          } else {
            s1 = sSi[0][k][lj][li];
            s2 = sSi[0][k][lj][li + 1];
          }
        } else {
          if constexpr (TAN) {
            s1 = S(sSj[0][k][lj][li], sSj[TAN ? 1 : 0][k][lj][li]);
            s2 = S(sSj[0][k][lj + 1][li], sSj[TAN ? 1 : 0][k][lj + 1][li]);
          } else {
            s1 = sSj[0][k][lj][li];
            s2 = sSj[0][k][lj + 1][li];
          }
        }
        wl[k] = q1 + K<S>(0.5) * s1;
        wr[k] = q2 - K<S>(0.5) * s2;
      } else {
        wl[k] = q1;
        wr[k] = q2;
      }
    }
    const W4<S> L = {dmaxc(wl[0], T(1e-6)), wl[1], wl[2], dmaxc(wl[3], T(1e-6))};
    const W4<S> R = {dmaxc(wr[0], T(1e-6)), wr[1], wr[2], dmaxc(wr[3], T(1e-6))};
    S F[4];
    hllc_flux<S, T>(L, R, sx, sy, F);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const T fv = TAN ? tan_of(F[k]) : val(F[k]);
      if (isI) sFi[k][lj][li] = fv;
      else sFj[k][lj][li] = fv;
    }
  }
  __syncthreads();

  // ---- cell differences and outputs ----------------------------------------
  double rn_part = 0.0;
  for (int c = tid; c < TI * TJ; c += NTS) {
    const int li = c % TI, lj = c / TI;
    const int i = i0 + li, j = j0 + lj;
    if (i >= ni || j >= nc) continue;
    const int idx = j * ni + i;
    T r[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      r[k] = (sFi[k][lj][li + 1] - sFi[k][lj][li]) + (sFj[k][lj + 1][li] - sFj[k][lj][li]);
    if constexpr (!TAN) {
      const gptr<double> R = gp(cd.R);
#pragma unroll
      for (int k = 0; k < 4; ++k) R[k * n + idx] = r[k];
      const double vol = cd.vol[idx];
      const double q = r[0] / vol;
      rn_part += q * q;
      if (want_diag) {
        // local pseudo-time step from this cell's primitives (aero2d/euler.py:289-296)
        const double rho = sW[0][lj + 2][li + 2], u = sW[1][lj + 2][li + 2];
        const double v = sW[2][lj + 2][li + 2], p = sW[3][lj + 2][li + 2];
        const double cs = sqrt(GAMMA * p / rho);
        const int ip = (i + 1 == ni) ? 0 : i + 1;
        const double six = 0.5 * (cd.Si[idx] + cd.Si[j * ni + ip]);
        const double siy = 0.5 * (cd.Si[n + idx] + cd.Si[n + j * ni + ip]);
        const double sjx = 0.5 * (cd.Sj[(j + 1) * ni + i] + cd.Sj[idx]);
        const double sjy = 0.5 * (cd.Sj[(nc + 1) * ni + (j + 1) * ni + i] + cd.Sj[(nc + 1) * ni + idx]);
        const double li_ = fabs(u * six + v * siy) + cs * sqrt(six * six + siy * siy);
        const double lj_ = fabs(u * sjx + v * sjy) + cs * sqrt(sjx * sjx + sjy * sjy);
        const double dt = kn.cfl_explicit * vol / (li_ + lj_);
        const double dtl = dt / kn.cfl_explicit * sc->cfl;
        const double dti = 1.0 / dtl;
        if (cd.prec == 0) tp<double>(cd.diag)[idx] = vol * dti;
        else tp<float>(cd.diag)[idx] = float(vol * dti);
      }
      if (refresh_ut && cd.prec == 1) {
        float* Ut = tp<float>(cd.Ut);
        const double* U = cd.U;
#pragma unroll
        for (int k = 0; k < 4; ++k) Ut[k * n + idx] = float(U[k * n + idx]);
      }
    } else {
      const gptr<T> out = vec_gptr<T>(cd, vout);
      const int vidx = j * ni + bcol(i, ni, kn.precond_rb);
      const T dg = use_diag ? gp(ctp<T>(cd.diag))[idx] : T(0);
      if (mode == 0) {
        const T s = (kidx >= 0) ? T(sc->s[kidx]) : T(1);
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k * n + vidx] = s * (dg * vt[k * n + vidx] + r[k]);
      } else {
        const gptr<const double> R = gp(cd.R);
#pragma unroll
        for (int k = 0; k < 4; ++k)
          out[k * n + vidx] = T(-R[k * n + idx]) - (dg * vt[k * n + vidx] + r[k]);
      }
    }
  }
  if constexpr (!TAN) {
    rn_part = wave_sum(rn_part);
    if ((tid & (WAVE - 1)) == 0) sred[tid / WAVE] = rn_part;
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < NTS / WAVE; ++w) s += sred[w];
      part_at(cd, lt, 0) = s;
    }
  }
}

// ---- per-case wall reductions (one workgroup per case) ----------------------

// Cl, Cd, Cm and wall Cp of the FP64 state; Gamma = Cl |V_inf| / 2.
__global__ void __launch_bounds__(NT) k_forces(Plan plan, double* cp_out) {
  __shared__ double red[3][NT / WAVE];
  const int e = blockIdx.x;
  const CaseDev& cd = plan.cases[plan.slots[e]];
  CaseScal* sc = cd.sc;
  const int ni = cd.ni, n = cd.n, nc = cd.nc;
  const double* U = cd.U;
  const double q = 0.5 * (sc->winf[1] * sc->winf[1] + sc->winf[2] * sc->winf[2]);
  double fx = 0.0, fy = 0.0, cm = 0.0;
  for (int i = threadIdx.x; i < ni; i += NT) {
    const W4<double> w0 = cons_to_prim<double>(U[i], U[n + i], U[2 * n + i], U[3 * n + i]);
    const int i1 = ni + i;
    const W4<double> w1 = cons_to_prim<double>(U[i1], U[n + i1], U[2 * n + i1], U[3 * n + i1]);
    const double pw = 1.5 * w0.p - 0.5 * w1.p;
    const double cp = (pw - sc->winf[3]) / q;
    const double sx = cd.Sj[i], sy = cd.Sj[(nc + 1) * ni + i];
    fx += cp * sx;
    fy += cp * sy;
    cm += cp * ((cd.xw[i] - 0.25) * sy - cd.yw[i] * sx);
    if (cp_out) cp_out[(size_t)e * ni + i] = cp;  // caller sizes cp_out for the widest case
  }
  fx = wave_sum(fx);
  fy = wave_sum(fy);
  cm = wave_sum(cm);
  const int w = threadIdx.x / WAVE;
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    red[0][w] = fx;
    red[1][w] = fy;
    red[2][w] = cm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sx = 0, sy = 0, sm = 0;
    for (int k = 0; k < NT / WAVE; ++k) {
      sx += red[0][k];
      sy += red[1][k];
      sm += red[2][k];
    }
    fx = -sx;
    fy = -sy;
    const double mach = sqrt(2.0 * q);
    const double ca = sc->winf[1] / mach, sa = sc->winf[2] / mach;
    const double cl = fy * ca - fx * sa;
    sc->cl = cl;
    sc->cd = fx * ca + fy * sa;
    sc->cm = sm;
    sc->circ = 0.5 * cl * sqrt(sc->winf[1] * sc->winf[1] + sc->winf[2] * sc->winf[2]);
  }
}

// Tangent of Gamma along vector v: dGamma = |V_inf|/2 * dCl, dCl from the wall
// pressure extrapolation 1.5 p_0 - 0.5 p_1 (the circulation coupling of the
// implicit residual, aero2d/implicit.py:160-167).
// kprev >= 0: the block first finishes Arnoldi step kprev (whose vector v is),
// and goes on only if the case proceeds to the next step.
template <class T>
__global__ void __launch_bounds__(NT) k_dcirc(Plan plan, int vin, int gate, int kprev, KnobsDev kn) {
  __shared__ double red[2][NT / WAVE];
  __shared__ ArnoldiLds lds;
  const int e = blockIdx.x;
  const CaseDev& cd = plan.cases[plan.slots[e]];
  CaseScal* sc = cd.sc;
  // a step is finished only for a case still in its Arnoldi loop
  if (sc->parked || ((gate || kprev >= 0) && !sc->active)) return;
  if (kprev >= 0) {
    arnoldi_finish<T>(cd, kprev, kn, lds);
    if (!lds.live) return;
  }
  const int ni = cd.ni, n = cd.n, nc = cd.nc;
  const CaseView<T> cv = view<T>(cd);
  const gptr<const T> v = vec_gptr<T>(cd, vin);
  const double q = 0.5 * (sc->winf[1] * sc->winf[1] + sc->winf[2] * sc->winf[2]);
  double dfx = 0.0, dfy = 0.0;
  for (int i = threadIdx.x; i < ni; i += NT) {
    using D = Dual<T>;
    const int vi = bcol(i, ni, kn.precond_rb);
    const W4<D> w0 = load_prim<D, T>(cv.U, v, n, i, vi);
    const W4<D> w1 = load_prim<D, T>(cv.U, v, n, ni + i, ni + vi);
    const double dpw = 1.5 * double(w0.p.d) - 0.5 * double(w1.p.d);
    const double dcp = dpw / q;
    dfx += dcp * cd.Sj[i];
    dfy += dcp * cd.Sj[(nc + 1) * ni + i];
  }
  dfx = wave_sum(dfx);
  dfy = wave_sum(dfy);
  const int w = threadIdx.x / WAVE;
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    red[0][w] = dfx;
    red[1][w] = dfy;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sx = 0, sy = 0;
    for (int k = 0; k < NT / WAVE; ++k) {
      sx += red[0][k];
      sy += red[1][k];
    }
    const double mach = sqrt(2.0 * q);
    const double ca = sc->winf[1] / mach, sa = sc->winf[2] / mach;
    const double dcl = (-sy) * ca - (-sx) * sa;
    sc->dcirc = 0.5 * dcl * mach;
  }
}

// rn = sqrt(sum of the residual kernel's partials / n), one wavefront per case.
__global__ void __launch_bounds__(WAVE) k_rn_finalize(Plan pc, int march) {
  const CaseDev& cd = pc.cases[pc.slots[blockIdx.x]];
  const double s = case_sum(cd, march ? march_units(cd.ni, cd.nc, MARCH_ROWS_RES) : cd.ntiles, 0);
  if (threadIdx.x == 0) cd.sc->rn = sqrt(s / cd.n);
}

// ---- shape table --------------------------------------------------------------
#define A2D_SHAPES(X) X(16, 16) X(32, 8) X(64, 4) X(8, 32)

constexpr StencilShape kShapes[] = {
#define X(a, b) {a, b},
    A2D_SHAPES(X)
#undef X
};

template <int TI, int TJ>
void run_residual(const Plan& p, int nblocks, const KnobsDev& k, int want_diag, double circ_override,
                  int refresh_ut, hipStream_t s) {
  hipLaunchKernelGGL((k_stencil<double, false, TI, TJ>), dim3(nblocks), dim3(NTS), 0, s, p, k, 0, 0,
                     0, -1, 0, want_diag, circ_override, refresh_ut, 0);
}

template <class T, int TI, int TJ>
void run_jvp(const Plan& p, int nblocks, const KnobsDev& k, int vin, int vout, int mode, int kidx,
             int use_diag, int gate, hipStream_t s) {
  hipLaunchKernelGGL((k_stencil<T, true, TI, TJ>), dim3(nblocks), dim3(NTS), 0, s, p, k, vin, vout,
                     mode, kidx, use_diag, 0, double(NAN), 0, gate);
}

}  // namespace

int stencil_num_shapes() { return int(sizeof(kShapes) / sizeof(kShapes[0])); }
StencilShape stencil_shape(int idx) { return kShapes[idx]; }
int stencil_blocks(int shape, int ni, int nc) {
  return ((ni + kShapes[shape].ti - 1) / kShapes[shape].ti) * ((nc + kShapes[shape].tj - 1) / kShapes[shape].tj);
}

void launch_residual(const Plan& p, int shape, int nblocks, const KnobsDev& k, int want_diag,
                     double circ_override, int refresh_ut, hipStream_t s) {
  int idx = 0;
#define X(a, b)                                                                    \
  if (shape == idx) {                                                              \
    run_residual<a, b>(p, nblocks, k, want_diag, circ_override, refresh_ut, s);    \
    return;                                                                        \
  }                                                                                \
  ++idx;
  A2D_SHAPES(X)
#undef X
}

void launch_jvp(const Plan& p, int shape, int nblocks, const KnobsDev& k, int vin, int vout,
                int mode, int kidx, int use_diag, int gate, int prec, hipStream_t s) {
  int idx = 0;
#define X(a, b)                                                                      \
  if (shape == idx) {                                                                \
    if (prec == 0) run_jvp<double, a, b>(p, nblocks, k, vin, vout, mode, kidx, use_diag, gate, s); \
    else run_jvp<float, a, b>(p, nblocks, k, vin, vout, mode, kidx, use_diag, gate, s);    \
    return;                                                                          \
  }                                                                                  \
  ++idx;
  A2D_SHAPES(X)
#undef X
}

void launch_forces(const Plan& p, double* cp_out, hipStream_t s) {
  hipLaunchKernelGGL(k_forces, dim3(p.nb), dim3(NT), 0, s, p, cp_out);
}

void launch_dcirc(const Plan& p, int vin, int gate, int kprev, const KnobsDev& kn, int prec,
                  hipStream_t s) {
  if (prec == 0) hipLaunchKernelGGL(k_dcirc<double>, dim3(p.nb), dim3(NT), 0, s, p, vin, gate, kprev, kn);
  else hipLaunchKernelGGL(k_dcirc<float>, dim3(p.nb), dim3(NT), 0, s, p, vin, gate, kprev, kn);
}

void launch_rn_finalize(const Plan& pc, int march, hipStream_t s) {
  hipLaunchKernelGGL(k_rn_finalize, dim3(pc.nb), dim3(WAVE), 0, s, pc, march);
}

// Test hook: HLLC flux on n independent faces, arrays (4, n) / (2, n).
__global__ void k_hllc_test(int n, const double* WL, const double* WR, const double* S, double* F) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const W4<double> L = {WL[f], WL[n + f], WL[2 * n + f], WL[3 * n + f]};
  const W4<double> R = {WR[f], WR[n + f], WR[2 * n + f], WR[3 * n + f]};
  double out[4];
  hllc_flux<double, double>(L, R, S[f], S[n + f], out);
  for (int k = 0; k < 4; ++k) F[k * n + f] = out[k];
}

void launch_hllc_test(int n, const double* WL, const double* WR, const double* S, double* F,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_hllc_test, dim3((n + 255) / 256), dim3(256), 0, s, n, WL, WR, S, F);
}

}  // namespace a2d
