// Block-tridiagonal wall-normal line preconditioner (aero2d/implicit.py:81-157).
//
//   k_assemble  first-order HLLC face Jacobians by forward-mode duals (one unit
//               tangent per conserved variable, what jax.jacfwd computes at
//               aero2d/implicit.py:36-59), assembled into the blocks D, Bim,
//               Bip, Bjm, Bjp of each cell plus vol/dtau on the diagonal
//               (aero2d/implicit.py:81-110). Thread per cell.
//   k_factor    block Thomas elimination along j for every i-line
//               (aero2d/implicit.py:112-128): den = D - A c_prev, inv = den^-1,
//               c = inv C. Thread per line, 4x4 blocks in registers.
//   one line-Jacobi sweep (aero2d/implicit.py:130-157) is two kernels. The
//   reference's forward step d_j = inv_j (rhs_j - A_j d_{j-1}) is split as
//   d_j = g_j - P_j d_{j-1} with g_j = inv_j rhs_j and P_j = inv_j A_j:
//   k_sweep_rhs  fully parallel, thread per cell: rhs = r - Bim x(i-1) - Bip x(i+1)
//                (r alone in the first sweep), g = inv rhs.
//   k_sweep_seq  the j recurrences only: forward d_j = g_j - P_j d_{j-1}, backward
//                x_j = d_j - c_j x_{j+1}. Four lanes (a quad) per line, lane r
//                owns row r of each 4x4 product, rows exchanged in the quad;
//                loads for the next PF steps are in flight while the current PF
//                steps compute (double-buffered registers).
//   The same arithmetic as the reference in another association order.
//
// Block storage: T blk[5][16][nc][ni], slot 0 = D then inv, 1 = Bjp then c,
// 2 = A = Bjm then P = inv A, 3 = Bim, 4 = Bip; entry m = row * 4 + col.
#include "common.hpp"
#include "launch.hpp"
#include "physics.hpp"

#include <type_traits>

namespace a2d {

namespace {

template <class T>
__device__ __forceinline__ const T* base_U(const CaseDev& cd) {
  if constexpr (std::is_same<T, double>::value) return cd.U;
  else return ctp<T>(cd.Ut);
}
template <class T>
__device__ __forceinline__ const T* met_Si(const CaseDev& cd) {
  if constexpr (std::is_same<T, double>::value) return cd.Si;
  else return ctp<T>(cd.Sit);
}
template <class T>
__device__ __forceinline__ const T* met_Sj(const CaseDev& cd) {
  if constexpr (std::is_same<T, double>::value) return cd.Sj;
  else return ctp<T>(cd.Sjt);
}
template <class T>
__device__ __forceinline__ const T* met_xf(const CaseDev& cd) {
  if constexpr (std::is_same<T, double>::value) return cd.xf;
  else return ctp<T>(cd.xft);
}
template <class T>
__device__ __forceinline__ const T* met_yf(const CaseDev& cd) {
  if constexpr (std::is_same<T, double>::value) return cd.yf;
  else return ctp<T>(cd.yft);
}

template <class T>
__device__ __forceinline__ void load_cons(gptr<const T> U, int n, int idx, T q[4]) {
  q[0] = U[idx];
  q[1] = U[n + idx];
  q[2] = U[2 * n + idx];
  q[3] = U[3 * n + idx];
}

// Interior first-order face: flux of (qL, qR); tangent e_b on side `right`.
template <class T>
__device__ __forceinline__ void face_col(const T qL[4], const T qR[4], T sx, T sy, int b,
                                         bool right, T col[4]) {
  using D = Dual<T>;
  D l[4], r[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    l[m] = D(qL[m], (!right && m == b) ? T(1) : T(0));
    r[m] = D(qR[m], (right && m == b) ? T(1) : T(0));
  }
  const W4<D> WL = cons_to_prim<D>(l[0], l[1], l[2], l[3]);
  const W4<D> WR = cons_to_prim<D>(r[0], r[1], r[2], r[3]);
  D F[4];
  hllc_flux<D, T>(WL, WR, sx, sy, F);
#pragma unroll
  for (int m = 0; m < 4; ++m) col[m] = F[m].d;
}

// Wall face: flux(mirror(w), w) with w = prim(q) (aero2d/implicit.py:44-51).
template <class T>
__device__ __forceinline__ void wall_col(const T q[4], T sx, T sy, int b, T col[4]) {
  using D = Dual<T>;
  D u[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) u[m] = D(q[m], m == b ? T(1) : T(0));
  const W4<D> w = cons_to_prim<D>(u[0], u[1], u[2], u[3]);
  const T area = sqrt(sx * sx + sy * sy);
  const W4<D> g = mirror(w, sx / area, sy / area);
  D F[4];
  hllc_flux<D, T>(g, w, sx, sy, F);
#pragma unroll
  for (int m = 0; m < 4; ++m) col[m] = F[m].d;
}

// Far-field face: flux(w, ghost(w)) with the circulation frozen
// (aero2d/implicit.py:54-59).
template <class T>
__device__ __forceinline__ void far_col(const T q[4], T sx, T sy, const Inf<T>& inf, T xf, T yf,
                                        T circ, int b, T col[4]) {
  using D = Dual<T>;
  D u[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) u[m] = D(q[m], m == b ? T(1) : T(0));
  const W4<D> w = cons_to_prim<D>(u[0], u[1], u[2], u[3]);
  const W4<D> g = farfield_ghost<D, T>(w, sx, sy, inf, xf, yf, D(circ));
  D F[4];
  hllc_flux<D, T>(w, g, sx, sy, F);
#pragma unroll
  for (int m = 0; m < 4; ++m) col[m] = F[m].d;
}

constexpr int NT = 256;

template <class T>
__global__ void __launch_bounds__(NT) k_assemble(Plan plan, KnobsDev kn) {
  const int e = plan_entry(plan, blockIdx.x);
  const CaseDev& cd = plan.cases[plan.slots[e]];
  const int ni = cd.ni, nc = cd.nc, n = cd.n;
  const int idx = (blockIdx.x - plan.blk_prefix[e]) * NT + threadIdx.x;
  if (idx >= n) return;
  const int i = idx % ni, j = idx / ni;
  const int im = (i == 0) ? ni - 1 : i - 1, ip = (i + 1 == ni) ? 0 : i + 1;
  const gptr<const T> U = gp(base_U<T>(cd));
  const gptr<const T> Si = gp(met_Si<T>(cd));
  const gptr<const T> Sj = gp(met_Sj<T>(cd));
  const size_t sj1 = (size_t)(nc + 1) * ni;
  const gptr<T> blk = gp(tp<T>(cd.blk));
  const size_t plane = (size_t)cd.pstride;
  auto at = [&](int slot, int m) -> T& { return blk[((size_t)slot * 16 + m) * plane + idx]; };

  T qc[4], qim[4], qip[4];
  load_cons(U, n, idx, qc);
  load_cons(U, n, j * ni + im, qim);
  load_cons(U, n, j * ni + ip, qip);
  T Dm[16];
  const T dg = gp(ctp<T>(cd.diag))[idx];
#pragma unroll
  for (int m = 0; m < 16; ++m) Dm[m] = (m % 5 == 0) ? dg : T(0);

  // i-face i (cells im | i) and i-face i+1 (cells i | ip)
  const T six0 = Si[j * ni + i], siy0 = Si[n + j * ni + i];
  const T six1 = Si[j * ni + ip], siy1 = Si[n + j * ni + ip];
  for (int b = 0; b < 4; ++b) {
    T col[4];
    face_col<T>(qim, qc, six0, siy0, b, false, col);   // AL_i -> Bim = -AL_i
#pragma unroll
    for (int r = 0; r < 4; ++r) at(3, r * 4 + b) = -col[r];
    face_col<T>(qim, qc, six0, siy0, b, true, col);    // AR_i -> D -= AR_i
#pragma unroll
    for (int r = 0; r < 4; ++r) Dm[r * 4 + b] -= col[r];
    face_col<T>(qc, qip, six1, siy1, b, false, col);   // AL_{i+1} -> D += AL_{i+1}
#pragma unroll
    for (int r = 0; r < 4; ++r) Dm[r * 4 + b] += col[r];
    face_col<T>(qc, qip, six1, siy1, b, true, col);    // AR_{i+1} -> Bip
#pragma unroll
    for (int r = 0; r < 4; ++r) at(4, r * 4 + b) = col[r];
  }
  // j-face j (cells j-1 | j) or the wall
  const T sjx0 = Sj[j * ni + i], sjy0 = Sj[sj1 + j * ni + i];
  if (j >= 1) {
    T qjm[4];
    load_cons(U, n, idx - ni, qjm);
    for (int b = 0; b < 4; ++b) {
      T col[4];
      face_col<T>(qjm, qc, sjx0, sjy0, b, false, col);  // AL_j -> Bjm = -AL_j
#pragma unroll
      for (int r = 0; r < 4; ++r) at(2, r * 4 + b) = -col[r];
      face_col<T>(qjm, qc, sjx0, sjy0, b, true, col);   // AR_j -> D -= AR_j
#pragma unroll
      for (int r = 0; r < 4; ++r) Dm[r * 4 + b] -= col[r];
    }
  } else {
    for (int b = 0; b < 4; ++b) {
      T col[4];
      wall_col<T>(qc, sjx0, sjy0, b, col);              // D -= Aw
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        Dm[r * 4 + b] -= col[r];
        at(2, r * 4 + b) = T(0);
      }
    }
  }
  // j-face j+1 (cells j | j+1) or the far field
  const T sjx1 = Sj[(j + 1) * ni + i], sjy1 = Sj[sj1 + (j + 1) * ni + i];
  if (j + 1 < nc) {
    T qjp[4];
    load_cons(U, n, idx + ni, qjp);
    for (int b = 0; b < 4; ++b) {
      T col[4];
      face_col<T>(qc, qjp, sjx1, sjy1, b, false, col);  // AL_{j+1} -> D += AL_{j+1}
#pragma unroll
      for (int r = 0; r < 4; ++r) Dm[r * 4 + b] += col[r];
      face_col<T>(qc, qjp, sjx1, sjy1, b, true, col);   // AR_{j+1} -> Bjp
#pragma unroll
      for (int r = 0; r < 4; ++r) at(1, r * 4 + b) = col[r];
    }
  } else {
    const CaseScal* sc = cd.sc;
    const Inf<T> inf = {T(sc->winf[0]), T(sc->winf[1]), T(sc->winf[2]), T(sc->winf[3])};
    const T circ = kn.vortex ? T(sc->circ) : T(0);
    const T xf = met_xf<T>(cd)[i], yf = met_yf<T>(cd)[i];
    for (int b = 0; b < 4; ++b) {
      T col[4];
      far_col<T>(qc, sjx1, sjy1, inf, xf, yf, circ, b, col);  // D += Af
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        Dm[r * 4 + b] += col[r];
        at(1, r * 4 + b) = T(0);
      }
    }
  }
#pragma unroll
  for (int m = 0; m < 16; ++m) at(0, m) = Dm[m];
}

// 4x4 inverse by Gauss-Jordan elimination with partial pivoting; all indices
// compile-time so the blocks stay in registers.
template <class T>
__device__ __forceinline__ void inv4(const T a_in[16], T out[16]) {
  T a[16], b[16];
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    a[m] = a_in[m];
    b[m] = (m % 5 == 0) ? T(1) : T(0);
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    // pivot search in column c among rows c..3
    int p = c;
    T best = fabs(a[c * 4 + c]);
#pragma unroll
    for (int r = c + 1; r < 4; ++r) {
      const T v = fabs(a[r * 4 + c]);
      if (v > best) { best = v; p = r; }
    }
#pragma unroll
    for (int r = c + 1; r < 4; ++r) {
      const bool sw = (p == r);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const T ta = a[c * 4 + k], tb = b[c * 4 + k];
        a[c * 4 + k] = sw ? a[r * 4 + k] : ta;
        a[r * 4 + k] = sw ? ta : a[r * 4 + k];
        b[c * 4 + k] = sw ? b[r * 4 + k] : tb;
        b[r * 4 + k] = sw ? tb : b[r * 4 + k];
      }
    }
    const T piv = T(1) / a[c * 4 + c];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      a[c * 4 + k] *= piv;
      b[c * 4 + k] *= piv;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (r == c) continue;
      const T f = a[r * 4 + c];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        a[r * 4 + k] -= f * a[c * 4 + k];
        b[r * 4 + k] -= f * b[c * 4 + k];
      }
    }
  }
#pragma unroll
  for (int m = 0; m < 16; ++m) out[m] = b[m];
}

template <class T>
__global__ void __launch_bounds__(64) k_factor(Plan plan) {
  const int e = plan_entry(plan, blockIdx.x);
  const CaseDev& cd = plan.cases[plan.slots[e]];
  const int ni = cd.ni, nc = cd.nc, n = cd.n;
  const int i = (blockIdx.x - plan.blk_prefix[e]) * 64 + threadIdx.x;
  if (i >= ni) return;
  const gptr<T> blk = gp(tp<T>(cd.blk));
  const size_t plane = (size_t)cd.pstride;
  T cprev[16];
#pragma unroll
  for (int m = 0; m < 16; ++m) cprev[m] = T(0);
  for (int j = 0; j < nc; ++j) {
    const size_t idx = (size_t)j * ni + i;
    T den[16], A[16], C[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      den[m] = blk[(0 * 16 + m) * plane + idx];
      A[m] = blk[(2 * 16 + m) * plane + idx];
      C[m] = blk[(1 * 16 + m) * plane + idx];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        T s = T(0);
#pragma unroll
        for (int k = 0; k < 4; ++k) s += A[r * 4 + k] * cprev[k * 4 + c];
        den[r * 4 + c] -= s;
      }
    T inv[16];
    inv4<T>(den, inv);
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        T s = T(0);
#pragma unroll
        for (int k = 0; k < 4; ++k) s += inv[r * 4 + k] * C[k * 4 + c];
        cprev[r * 4 + c] = s;
      }
    T P[16];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        T s = T(0);
#pragma unroll
        for (int k = 0; k < 4; ++k) s += inv[r * 4 + k] * A[k * 4 + c];
        P[r * 4 + c] = s;
      }
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      blk[(0 * 16 + m) * plane + idx] = inv[m];
      blk[(1 * 16 + m) * plane + idx] = cprev[m];
      blk[(2 * 16 + m) * plane + idx] = P[m];
    }
  }
}

// Sweep part 1, thread per cell: g = inv (src - Bim x(i-1) - Bip x(i+1)) into d.
template <class T>
__global__ void __launch_bounds__(256) k_sweep_rhs(Plan plan, int src, int xprev, int sweep,
                                                   int gate) {
  const int e = plan_entry(plan, blockIdx.x);
  const CaseDev& cd = plan.cases[plan.slots[e]];
  if (gate && !cd.sc->active) return;
  const int ni = cd.ni, n = cd.n;
  const int idx = (blockIdx.x - plan.blk_prefix[e]) * 256 + threadIdx.x;
  if (idx >= n) return;
  const int i = idx % ni, j = idx / ni;
  const gptr<const T> blk = gp(ctp<T>(cd.blk));
  const size_t plane = (size_t)cd.pstride;
  T rhs[4];
  if (src == VEC_NEG_R) {
    const gptr<const double> Rg = gp(cd.R);
#pragma unroll
    for (int r = 0; r < 4; ++r) rhs[r] = T(-Rg[(size_t)r * n + idx]);
  } else {
    const gptr<const T> b = vec_gptr<T>(cd, src);
#pragma unroll
    for (int r = 0; r < 4; ++r) rhs[r] = b[(size_t)r * n + idx];
  }
  if (sweep > 0) {
    const gptr<const T> xp = vec_gptr<T>(cd, xprev);
    const int im = (i == 0) ? ni - 1 : i - 1, ip = (i + 1 == ni) ? 0 : i + 1;
    T xm[4], xq[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      xm[c] = xp[(size_t)c * n + j * ni + im];
      xq[c] = xp[(size_t)c * n + j * ni + ip];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 4; ++c)
        rhs[r] -= blk[(size_t)(3 * 16 + r * 4 + c) * plane + idx] * xm[c] +
                  blk[(size_t)(4 * 16 + r * 4 + c) * plane + idx] * xq[c];
  }
  const gptr<T> g = gp(tp<T>(cd.d));
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    T s = T(0);
#pragma unroll
    for (int c = 0; c < 4; ++c) s += blk[(size_t)(r * 4 + c) * plane + idx] * rhs[c];
    g[(size_t)r * n + idx] = s;
  }
}

// Sweep part 2: the two j recurrences of every line. Block = 64 lines x 4 rows.
template <class T, int PF>
__global__ void __launch_bounds__(256) k_sweep_seq(Plan plan, int xout, int want_norm, int gate) {
  __shared__ double red[4];
  const int e = plan_entry(plan, blockIdx.x);
  const CaseDev& cd = plan.cases[plan.slots[e]];
  if (gate && !cd.sc->active) return;
  const int ni = cd.ni, nc = cd.nc, n = cd.n;
  const int lb = blockIdx.x - plan.blk_prefix[e];
  const int r = threadIdx.x & 3;
  const int i = lb * 64 + (threadIdx.x >> 2);
  const bool live = i < ni;
  const int ic = live ? i : ni - 1;     // dead lanes shadow a real line and never store
  const size_t plane = (size_t)cd.pstride;
  const gptr<const T> P = gp(ctp<T>(cd.blk) + (size_t)(2 * 16 + r * 4) * plane + ic);
  const gptr<const T> C = gp(ctp<T>(cd.blk) + (size_t)(1 * 16 + r * 4) * plane + ic);
  const gptr<T> d = gp(tp<T>(cd.d) + (size_t)r * n + ic);
  const gptr<T> x = gp(vec_ptr<T>(cd, xout) + (size_t)r * n + ic);

  // forward: d_j = g_j - P_j d_{j-1}, in place over g
  T bm[2][PF][4], bv[2][PF];
  auto fload = [&](int b, int j0) {
#pragma unroll
    for (int q = 0; q < PF; ++q) {
      const int j = min(j0 + q, nc - 1);
      const size_t o = (size_t)j * ni;
      bv[b][q] = d[o];
#pragma unroll
      for (int c = 0; c < 4; ++c) bm[b][q][c] = P[(size_t)c * plane + o];
    }
  };
  T dr = T(0);
  auto fcomp = [&](int b, int j0) {
#pragma unroll
    for (int q = 0; q < PF; ++q) {
      const int j = j0 + q;
      if (j < nc) {
        const T t = bv[b][q] - bm[b][q][0] * quad_bcast<0>(dr) - bm[b][q][1] * quad_bcast<1>(dr) -
                    bm[b][q][2] * quad_bcast<2>(dr) - bm[b][q][3] * quad_bcast<3>(dr);
        dr = t;
        if (live) d[(size_t)j * ni] = dr;
      }
    }
  };
  fload(0, 0);
  for (int j0 = 0; j0 < nc; j0 += 2 * PF) {
    fload(1, j0 + PF);
    fcomp(0, j0);
    fload(0, j0 + 2 * PF);
    fcomp(1, j0 + PF);
  }
  // backward: x_j = d_j - c_j x_{j+1}
  auto bload = [&](int b, int j0) {
#pragma unroll
    for (int q = 0; q < PF; ++q) {
      const int j = max(j0 - q, 0);
      const size_t o = (size_t)j * ni;
      bv[b][q] = d[o];
#pragma unroll
      for (int c = 0; c < 4; ++c) bm[b][q][c] = C[(size_t)c * plane + o];
    }
  };
  T xr = T(0);
  double nrm = 0.0;
  auto bcomp = [&](int b, int j0) {
#pragma unroll
    for (int q = 0; q < PF; ++q) {
      const int j = j0 - q;
      if (j >= 0) {
        const T t = bv[b][q] - bm[b][q][0] * quad_bcast<0>(xr) - bm[b][q][1] * quad_bcast<1>(xr) -
                    bm[b][q][2] * quad_bcast<2>(xr) - bm[b][q][3] * quad_bcast<3>(xr);
        xr = t;
        if (live) {
          x[(size_t)j * ni] = xr;
          nrm += double(xr) * double(xr);
        }
      }
    }
  };
  bload(0, nc - 1);
  for (int j0 = nc - 1; j0 >= 0; j0 -= 2 * PF) {
    bload(1, j0 - PF);
    bcomp(0, j0);
    bload(0, j0 - 2 * PF);
    bcomp(1, j0 - PF);
  }
  if (want_norm) {
    nrm = wave_sum(nrm);
    if ((threadIdx.x & (WAVE - 1)) == 0) red[threadIdx.x / WAVE] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) cd.part[(size_t)lb * (cd.kmax + 2)] = red[0] + red[1] + red[2] + red[3];
  }
}

#ifndef A2D_SWEEP_PF
#define A2D_SWEEP_PF 8
#endif
constexpr int SWEEP_PF = A2D_SWEEP_PF;   // j steps loaded ahead by the line recurrences

}  // namespace

void launch_assemble(const Plan& pv, int nblocks, const KnobsDev& k, int prec, hipStream_t s) {
  if (prec == 0) hipLaunchKernelGGL(k_assemble<double>, dim3(nblocks), dim3(NT), 0, s, pv, k);
  else hipLaunchKernelGGL(k_assemble<float>, dim3(nblocks), dim3(NT), 0, s, pv, k);
}

void launch_factor(const Plan& pl, int nblocks, int prec, hipStream_t s) {
  if (prec == 0) hipLaunchKernelGGL(k_factor<double>, dim3(nblocks), dim3(64), 0, s, pl);
  else hipLaunchKernelGGL(k_factor<float>, dim3(nblocks), dim3(64), 0, s, pl);
}

void launch_line_sweep(const Plan& pcell, int ncell, const Plan& pl, int nblocks, int src, int xprev,
                       int xout, int sweep, int want_norm, int gate, int prec, hipStream_t s) {
  if (prec == 0) {
    hipLaunchKernelGGL(k_sweep_rhs<double>, dim3(ncell), dim3(256), 0, s, pcell, src, xprev, sweep, gate);
    hipLaunchKernelGGL((k_sweep_seq<double, SWEEP_PF>), dim3(nblocks), dim3(256), 0, s, pl, xout,
                       want_norm, gate);
  } else {
    hipLaunchKernelGGL(k_sweep_rhs<float>, dim3(ncell), dim3(256), 0, s, pcell, src, xprev, sweep, gate);
    hipLaunchKernelGGL((k_sweep_seq<float, SWEEP_PF>), dim3(nblocks), dim3(256), 0, s, pl, xout,
                       want_norm, gate);
  }
}

void launch_sweep_part(const Plan& pcell, int ncell, const Plan& pl, int nblocks, int part, int prec,
                       hipStream_t s) {
  if (part == 0) {
    if (prec == 0) hipLaunchKernelGGL(k_sweep_rhs<double>, dim3(ncell), dim3(256), 0, s, pcell, VEC_W, VEC_XA, 1, 0);
    else hipLaunchKernelGGL(k_sweep_rhs<float>, dim3(ncell), dim3(256), 0, s, pcell, VEC_W, VEC_XA, 1, 0);
  } else {
    if (prec == 0)
      hipLaunchKernelGGL((k_sweep_seq<double, SWEEP_PF>), dim3(nblocks), dim3(256), 0, s, pl, VEC_XB, 0, 0);
    else
      hipLaunchKernelGGL((k_sweep_seq<float, SWEEP_PF>), dim3(nblocks), dim3(256), 0, s, pl, VEC_XB, 0, 0);
  }
}

}  // namespace a2d
