// RANS (Spalart-Allmaras, SA-neg) Newton-Krylov core on an unstructured mesh:
// residual, exact Jacobian-vector product, first-order Jacobian blocks, block
// ILU(k) factorisation and triangular solves, GMRES, local pseudo-time step,
// step limit and forces; the C ABI declared in include/aero2d_hip.h
// (a2d_rans_*). Reference: aero2d/rans.py (RansSolver, SA path),
// aero2d/bilu.py (BlockILU0), aero2d/krylov.py (gmres).
//
// Device layout of one solver (all FP64):
//   Q, R, dq, x, ... [5][nc]        state and vectors, structure of arrays
//   g, gt            [10][nc]       least-squares gradients (k, d) -> 2k+d, and tangents
//   phi, phit        [5][nc]        limiter and its tangent
//   F, Ft            [5][nfi]       interior face fluxes times face length
//   B                [nnz][5][5]    ILU blocks in the permuted order, row-major blocks
//   y                [n][5]         ILU work vector in the permuted order
// Mesh arrays are structure-of-arrays per face set (interior, wall, far);
// each cell lists its faces (CSR), so every gather is per cell in a fixed
// order and results do not depend on scheduling (no atomics).
#include "../../include/aero2d_hip.h"
#include "api.hpp"
#include "rans_physics.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace a2d {
namespace rans {
namespace {

struct Fail {
  int code;
};
[[noreturn]] void fail(int code, const std::string& msg) {
  set_error_text(msg);
  throw Fail{code};
}
void hc(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(-10, std::string(what) + ": " + hipGetErrorString(e));
}
#define HC(x) hc((x), #x)

// Creation: no handle yet and no library-wide lock (the map insert locks the map).
template <class F>
int guarded_new(F&& f) {
  try {
    f();
    return 0;
  } catch (const Fail& e) {
    return e.code;
  } catch (const std::exception& e) {
    set_error_text(e.what());
    return -99;
  }
}

template <class F>
int guarded(F&& f) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  try {
    f();
    return 0;
  } catch (const Fail& e) {
    return e.code;
  } catch (const std::exception& e) {
    set_error_text(e.what());
    return -99;
  }
}

constexpr int NFG = 11;   // interior face: nx ny A rLx rLy rRx rRy dx dy dist w
constexpr int NWG = 9;    // wall face: nx ny A rx ry dn w fcx fcy
constexpr int NFF = 7;    // far face: nx ny A rx ry fcx fcy
constexpr int NCG = 6;    // cell: vol wd M00 M01 M11 in_limiter_zone
constexpr int TPB = 256;
constexpr int GMRES_MAX = 120;   // largest Krylov basis of one linear solve

struct MeshDev {
  int nc, nfi, nfw, nff;
  const int *iL, *iR;
  const double* fg;    // [NFG][nfi]
  const int* cw;
  const double* wg;    // [NWG][nfw]
  const int* cf;
  const double* ffg;   // [NFF][nff]
  const double* cg;    // [NCG][nc]
  const unsigned char* far_cell;
  const int *cptr, *cfac;   // interior faces of a cell: f (cell is left) or ~f (right)
  const int *wptr, *wfac;
  const int *fptr, *ffac;
};

// ---- loads and stores of plain or dual values ---------------------------------
template <class S> struct is_dual { static constexpr bool v = false; };
template <class T> struct is_dual<Dual<T>> { static constexpr bool v = true; };

template <class S>
__device__ __forceinline__ S ld(const double* a, const double* at, int i) {
  if constexpr (is_dual<S>::v) return S(a[i], at ? at[i] : 0.0);
  else return a[i];
}
template <class S>
__device__ __forceinline__ void st(double* a, double* at, int i, S x) {
  a[i] = val(x);
  if constexpr (is_dual<S>::v) { if (at) at[i] = x.d; }
}

template <class S>
__device__ __forceinline__ void load_w(const double* Q, const double* V, int nc, int c, S w[5]) {
  S q[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) q[k] = ld<S>(Q, V, k * nc + c);
  prim(q, w);
}

// Gradients of cell c as used by the fluxes: at second order multiplied by the
// limiter (when one is applied) and zero in cells on the far-field boundary
// (aero2d/rans.py:870-877).
template <class S>
__device__ __forceinline__ void load_g(const MeshDev& m, const double* g, const double* gt, const double* phi,
                                       const double* phit, int order, int c, S G[5][2]) {
#pragma unroll
  for (int k = 0; k < 5; ++k)
#pragma unroll
    for (int d = 0; d < 2; ++d) G[k][d] = ld<S>(g, gt, (2 * k + d) * m.nc + c);
  if (order >= 2) {
    if (phi) {
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        S p = ld<S>(phi, phit, k * m.nc + c);
        G[k][0] = G[k][0] * p;
        G[k][1] = G[k][1] * p;
      }
    }
    if (m.far_cell[c]) {
#pragma unroll
      for (int k = 0; k < 5; ++k) G[k][0] = G[k][1] = K<S>(0.0);
    }
  }
}

// ---- residual pieces ----------------------------------------------------------------
// Least-squares gradients of the primitives (aero2d/rans.py:809-819).
template <class S>
__global__ void k_grad(MeshDev m, const double* Q, const double* V, double* g, double* gt) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m.nc) return;
  S w[5];
  load_w<S>(Q, V, m.nc, c, w);
  S rhs[5][2];
#pragma unroll
  for (int k = 0; k < 5; ++k) rhs[k][0] = rhs[k][1] = K<S>(0.0);
  for (int e = m.cptr[c]; e < m.cptr[c + 1]; ++e) {
    const int code = m.cfac[e];
    const bool left = code >= 0;
    const int f = left ? code : ~code;
    const int nb = left ? m.iR[f] : m.iL[f];
    S wn[5];
    load_w<S>(Q, V, m.nc, nb, wn);
    const double wf = m.fg[10 * m.nfi + f], dx = m.fg[7 * m.nfi + f], dy = m.fg[8 * m.nfi + f];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      S dphi = left ? wn[k] - w[k] : w[k] - wn[k];       // w[iR] - w[iL]
      S a = K<S>(wf) * dphi;
      rhs[k][0] = rhs[k][0] + a * dx;
      rhs[k][1] = rhs[k][1] + a * dy;
    }
  }
  for (int e = m.wptr[c]; e < m.wptr[c + 1]; ++e) {
    const int f = m.wfac[e];
    const double ww = m.wg[6 * m.nfw + f], rx = m.wg[3 * m.nfw + f], ry = m.wg[4 * m.nfw + f];
    S wv[5] = {w[0], K<S>(0.0), K<S>(0.0), w[3], K<S>(0.0)};
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      S a = K<S>(ww) * (wv[k] - w[k]);
      rhs[k][0] = rhs[k][0] + a * rx;
      rhs[k][1] = rhs[k][1] + a * ry;
    }
  }
  const double M00 = m.cg[2 * m.nc + c], M01 = m.cg[3 * m.nc + c], M11 = m.cg[4 * m.nc + c];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    st<S>(g, gt, (2 * k) * m.nc + c, K<S>(M00) * rhs[k][0] + K<S>(M01) * rhs[k][1]);
    st<S>(g, gt, (2 * k + 1) * m.nc + c, K<S>(M01) * rhs[k][0] + K<S>(M11) * rhs[k][1]);
  }
}

// Venkatakrishnan limiter (aero2d/rans.py:983-1022) in the limiter zone, 1
// elsewhere (aero2d/rans.py:869-871).
template <class S>
__global__ void k_limit(MeshDev m, Params P, const double* Q, const double* V, const double* g,
                        const double* gt, double* phi, double* phit) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m.nc) return;
  const bool zone = P.limiter_zone == 2 || (P.limiter_zone == 1 && m.cg[5 * m.nc + c] != 0.0);
  if (!zone) {
#pragma unroll
    for (int k = 0; k < 5; ++k) st<S>(phi, phit, k * m.nc + c, K<S>(1.0));
    return;
  }
  S w[5];
  load_w<S>(Q, V, m.nc, c, w);
  Ext<S> mx[5], mn[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) mx[k] = mn[k] = Ext<S>{w[k], 1};
  for (int e = m.cptr[c]; e < m.cptr[c + 1]; ++e) {
    const int code = m.cfac[e];
    const int f = code >= 0 ? code : ~code;
    const int nb = code >= 0 ? m.iR[f] : m.iL[f];
    S wn[5];
    load_w<S>(Q, V, m.nc, nb, wn);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      ext_push_max(mx[k], wn[k]);
      ext_push_min(mn[k], wn[k]);
    }
  }
  for (int e = m.wptr[c]; e < m.wptr[c + 1]; ++e) {
    S wv[5] = {w[0], K<S>(0.0), K<S>(0.0), w[3], K<S>(0.0)};
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      ext_push_max(mx[k], wv[k]);
      ext_push_min(mn[k], wv[k]);
    }
  }
  S G[5][2];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    G[k][0] = ld<S>(g, gt, (2 * k) * m.nc + c);
    G[k][1] = ld<S>(g, gt, (2 * k + 1) * m.nc + c);
  }
  const double h = sqrt(m.cg[c]);
  const double kh = P.limiter_k * h;
  const double kh3 = kh * kh * kh;
  const double scale[5] = {1.0, P.mach, P.mach, 1.0 / GAM, P.chi_inf > 1.0 ? P.chi_inf : 1.0};
  S dmx[5], dmn[5];
  Ext<S> lim[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    dmx[k] = ext_value(mx[k]) - w[k];
    dmn[k] = ext_value(mn[k]) - w[k];
    lim[k] = Ext<S>{K<S>(1.0), 1};
  }
  for (int e = m.cptr[c]; e < m.cptr[c + 1]; ++e) {
    const int code = m.cfac[e];
    const bool left = code >= 0;
    const int f = left ? code : ~code;
    const double rx = m.fg[(left ? 3 : 5) * m.nfi + f], ry = m.fg[(left ? 4 : 6) * m.nfi + f];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      S d2 = G[k][0] * rx + G[k][1] * ry;
      ext_push_min(lim[k], venkat_psi(d2, dmx[k], dmn[k], kh3 * scale[k] * scale[k]));
    }
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) st<S>(phi, phit, k * m.nc + c, ext_value(lim[k]));
}

// Far-field state: freestream plus the Prandtl-Glauert point vortex of
// circulation gam at xref, isentropic rho and p (aero2d/rans.py:830-858).
__global__ void k_far_state(MeshDev m, Params P, double w0, double w1, double w2, double w3, double w4,
                            double gam, double* wgo) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.nff) return;
  const double x = m.ffg[5 * m.nff + f] - P.xref_x, y = m.ffg[6 * m.nff + f] - P.xref_y;
  const double M = P.mach;
  const double b = sqrt(1.0 - M * M);
  const double vm = sqrt(w1 * w1 + w2 * w2) + 1e-30;
  const double ca = w1 / vm, sa = w2 / vm;
  const double xr = x * ca + y * sa, yr = -x * sa + y * ca;
  const double th = atan2(yr, xr);
  const double r = sqrt(xr * xr + yr * yr);
  const double st_ = sin(th), ct_ = cos(th);
  const double fac = gam * b / (2.0 * M_PI * r * (1.0 - M * M * st_ * st_));
  const double dur = fac * st_, dvr = -fac * ct_;
  const double U = w1 + (dur * ca - dvr * sa), Vv = w2 + (dur * sa + dvr * ca);
  const double q2 = U * U + Vv * Vv;
  const double T = 1.0 + 0.5 * (GAM - 1.0) * (M * M - q2);
  const double rho = pow(T, 1.0 / (GAM - 1.0));
  wgo[f] = rho;
  wgo[m.nff + f] = U;
  wgo[2 * m.nff + f] = Vv;
  wgo[3 * m.nff + f] = rho * T / GAM;
  wgo[4 * m.nff + f] = w4;
  (void)w0;
  (void)w3;
}

// Interior face flux: reconstruction, LM-Roe, viscous flux with the face
// gradient (cell mean plus the correction along the line of centres), times
// the face length (aero2d/rans.py:878-919).
// VR (visc_raw_grad): the viscous part from the unlimited gradients.
template <class S, bool VR>
__global__ void k_face(MeshDev m, Params P, int order, const double* Q, const double* V, const double* g,
                       const double* gt, const double* phi, const double* phit, double* F, double* Ft) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.nfi) return;
  const int L = m.iL[f], R = m.iR[f];
  const int nf = m.nfi;
  const double nx = m.fg[f], ny = m.fg[nf + f], A = m.fg[2 * nf + f];
  S wL0[5], wR0[5], GL[5][2], GR[5][2];
  load_w<S>(Q, V, m.nc, L, wL0);
  load_w<S>(Q, V, m.nc, R, wR0);
  load_g<S>(m, g, gt, phi, phit, order, L, GL);
  load_g<S>(m, g, gt, phi, phit, order, R, GR);
  S wL[5], wR[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    wL[k] = wL0[k];
    wR[k] = wR0[k];
  }
  if (order >= 2) {
    const double rLx = m.fg[3 * nf + f], rLy = m.fg[4 * nf + f];
    const double rRx = m.fg[5 * nf + f], rRy = m.fg[6 * nf + f];
    S a[4], b[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      a[k] = wL0[k] + (GL[k][0] * rLx + GL[k][1] * rLy);
      b[k] = wR0[k] + (GR[k][0] * rRx + GR[k][1] * rRy);
    }
    if (val(a[0]) > 0.2 * val(wL0[0]) && val(a[3]) > 0.2 * val(wL0[3]))
#pragma unroll
      for (int k = 0; k < 4; ++k) wL[k] = a[k];
    if (val(b[0]) > 0.2 * val(wR0[0]) && val(b[3]) > 0.2 * val(wR0[3]))
#pragma unroll
      for (int k = 0; k < 4; ++k) wR[k] = b[k];
  }
  S Fc[5];
  roe(P, wL, wR, nx, ny, Fc);
  if constexpr (VR) {
    if (order >= 2) {               // the gradients before limiting and far-field zeroing
      load_g<S>(m, g, gt, nullptr, nullptr, 1, L, GL);
      load_g<S>(m, g, gt, nullptr, nullptr, 1, R, GR);
    }
  }
  S wf[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) wf[k] = K<S>(0.5) * (wL0[k] + wR0[k]);
  const S TL = K<S>(GAM) * wL0[3] / wL0[0], TR = K<S>(GAM) * wR0[3] / wR0[0];
  S gTL[2], gTR[2];
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    gTL[d] = K<S>(GAM) * (GL[3][d] / wL0[0] - (wL0[3] / (wL0[0] * wL0[0])) * GL[0][d]);
    gTR[d] = K<S>(GAM) * (GR[3][d] / wR0[0] - (wR0[3] / (wR0[0] * wR0[0])) * GR[0][d]);
  }
  const double dist = m.fg[9 * nf + f];
  const double ex = m.fg[7 * nf + f] / dist, ey = m.fg[8 * nf + f] / dist;
  auto fgrad = [&](S pl, S pr, const S gl[2], const S gr[2], S out[2]) {
    S ga0 = K<S>(0.5) * (gl[0] + gr[0]), ga1 = K<S>(0.5) * (gl[1] + gr[1]);
    S corr = (pr - pl) / K<S>(dist) - (ga0 * ex + ga1 * ey);
    out[0] = ga0 + corr * ex;
    out[1] = ga1 + corr * ey;
  };
  S gu[2], gv[2], gTf[2], gc[2];
  fgrad(wL0[1], wR0[1], GL[1], GR[1], gu);
  fgrad(wL0[2], wR0[2], GL[2], GR[2], gv);
  fgrad(TL, TR, gTL, gTR, gTf);
  fgrad(wL0[4], wR0[4], GL[4], GR[4], gc);
  S Fv[5];
  visc_flux(P, wf, gu, gv, gTf, gc, nx, ny, Fv);
#pragma unroll
  for (int k = 0; k < 5; ++k) st<S>(F, Ft, k * nf + f, (Fc[k] - Fv[k]) * K<S>(A));
}

// Cell residual: signed sum of interior face fluxes, wall and far-field faces,
// minus the SA source times the volume (aero2d/rans.py:920-964). With vdt
// non-null the output is the Newton matrix times V: the tangent plus
// vol/dt V (vol/(dt ct) on the turbulence equation). VR (visc_raw_grad): the
// wall's viscous flux from the unlimited gradients; SR (sa_raw_grad): the SA
// source from them (the wall pressure and the far-field state always take
// the limited ones).
template <class S, bool VR, bool SR>
__global__ void k_cell(MeshDev m, Params P, int order, const double* Q, const double* V, const double* g,
                       const double* gt, const double* phi, const double* phit, const double* F,
                       const double* Ft, const double* wgf, double* Rv, double* Rt, const double* vdt,
                       double ct_inv) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m.nc) return;
  S w[5], G[5][2];
  load_w<S>(Q, V, m.nc, c, w);
  load_g<S>(m, g, gt, phi, phit, order, c, G);
  S R[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) R[k] = K<S>(0.0);
  for (int e = m.cptr[c]; e < m.cptr[c + 1]; ++e) {
    const int code = m.cfac[e];
    if (code >= 0) {
#pragma unroll
      for (int k = 0; k < 5; ++k) R[k] = R[k] + ld<S>(F, Ft, k * m.nfi + code);
    } else {
#pragma unroll
      for (int k = 0; k < 5; ++k) R[k] = R[k] - ld<S>(F, Ft, k * m.nfi + ~code);
    }
  }
  const int nw = m.nfw;
  if (m.wptr[c] < m.wptr[c + 1]) {
    // VR: the limited G[3] kept for the wall pressure's reconstruction, G
    // reloaded unlimited for the viscous flux
    S g3[2] = {G[3][0], G[3][1]};
    if constexpr (VR) {
      if (order >= 2) load_g<S>(m, g, gt, nullptr, nullptr, 1, c, G);
    }
    S gT[2];
#pragma unroll
    for (int d = 0; d < 2; ++d) gT[d] = K<S>(GAM) * (G[3][d] / w[0] - (w[3] / (w[0] * w[0])) * G[0][d]);
    for (int e = m.wptr[c]; e < m.wptr[c + 1]; ++e) {
      const int f = m.wfac[e];
      const double nx = m.wg[f], ny = m.wg[nw + f], A = m.wg[2 * nw + f];
      const double rx = m.wg[3 * nw + f], ry = m.wg[4 * nw + f], dn = m.wg[5 * nw + f];
      S pw = w[3];
      if constexpr (VR) {
        if (order >= 2) pw = w[3] + (g3[0] * rx + g3[1] * ry);
      } else {
        if (order >= 2) pw = w[3] + (G[3][0] * rx + G[3][1] * ry);
      }
      if (!(val(pw) > 0.2 * val(w[3]))) pw = w[3];
      auto wgrad = [&](S phii, const S gi[2], S out[2]) {
        S corr = (K<S>(0.0) - phii) / K<S>(dn) - (gi[0] * nx + gi[1] * ny);
        out[0] = gi[0] + corr * nx;
        out[1] = gi[1] + corr * ny;
      };
      S gu[2], gv[2], gc[2], gTw[2];
      wgrad(w[1], G[1], gu);
      wgrad(w[2], G[2], gv);
      wgrad(w[4], G[4], gc);
      S tn = gT[0] * nx + gT[1] * ny;
      gTw[0] = gT[0] - tn * nx;
      gTw[1] = gT[1] - tn * ny;
      S wv[5] = {w[0], K<S>(0.0), K<S>(0.0), w[3], K<S>(0.0)};
      S Fv[5];
      visc_flux(P, wv, gu, gv, gTw, gc, nx, ny, Fv);
      R[0] = R[0] + (K<S>(0.0) - Fv[0]) * K<S>(A);
      R[1] = R[1] + (pw * nx - Fv[1]) * K<S>(A);
      R[2] = R[2] + (pw * ny - Fv[2]) * K<S>(A);
      R[3] = R[3] + (K<S>(0.0) - Fv[3]) * K<S>(A);
      R[4] = R[4] + (K<S>(0.0) - Fv[4]) * K<S>(A);
    }
    if constexpr (VR) {
      if (order >= 2) load_g<S>(m, g, gt, phi, phit, order, c, G);
    }
  }
  const int nff = m.nff;
  for (int e = m.fptr[c]; e < m.fptr[c + 1]; ++e) {
    const int f = m.ffac[e];
    const double nx = m.ffg[f], ny = m.ffg[nff + f], A = m.ffg[2 * nff + f];
    S wi[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) wi[k] = w[k];
    if (order >= 2) {
      const double rx = m.ffg[3 * nff + f], ry = m.ffg[4 * nff + f];
      S a[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) a[k] = w[k] + (G[k][0] * rx + G[k][1] * ry);
      if (val(a[0]) > 0.2 * val(w[0]) && val(a[3]) > 0.2 * val(w[3]))
#pragma unroll
        for (int k = 0; k < 5; ++k) wi[k] = a[k];
    }
    S wb[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) wb[k] = K<S>(wgf[k * nff + f]);
    S Ff[5];
    roe(P, wi, wb, nx, ny, Ff);
#pragma unroll
    for (int k = 0; k < 5; ++k) R[k] = R[k] + Ff[k] * K<S>(A);
  }
  const double vol = m.cg[c];
  if (P.turbulent) {
    if constexpr (SR) {
      if (order >= 2) load_g<S>(m, g, gt, nullptr, nullptr, 1, c, G);
    }
    R[4] = R[4] - sa_source(P, w, G, m.cg[m.nc + c]) * K<S>(vol);
  } else {
    R[4] = ld<S>(Q, V, 4 * m.nc + c) * K<S>(vol);
  }
  if constexpr (is_dual<S>::v) {
    if (vdt) {
      const double d = vdt[c];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const double dk = k == 4 ? d * ct_inv : d;
        Rt[k * m.nc + c] = R[k].d + dk * V[k * m.nc + c];
      }
      return;
    }
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) st<S>(Rv, Rt, k * m.nc + c, R[k]);
}

// Per-equation sums of (R_k / vol)^2 over blocks of cells, then in order.
__global__ void k_rnorm_part(int nc, const double* R, const double* vol, double* part) {
  __shared__ double sh[5][TPB];
  const int c = blockIdx.x * TPB + threadIdx.x;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    double v = 0.0;
    if (c < nc) {
      const double x = R[k * nc + c] / vol[c];
      v = x * x;
    }
    sh[k][threadIdx.x] = v;
  }
  __syncthreads();
  for (int s = TPB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
#pragma unroll
      for (int k = 0; k < 5; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 5) part[threadIdx.x * gridDim.x + blockIdx.x] = sh[threadIdx.x][0];
}

// Sums nparts partials of each of nsum rows in order (one lane per row).
__global__ void k_sum_rows(int nsum, int nparts, const double* part, double* out) {
  const int r = threadIdx.x;
  if (r >= nsum) return;
  double s = 0.0;
  for (int p = 0; p < nparts; ++p) s += part[r * nparts + p];
  out[r] = s;
}

// Local pseudo-time step (aero2d/rans.py:1206-1249, SA): vdt = vol / dt.
__global__ void k_local_dt(MeshDev m, Params P, const double* Q, double cfl, double* dt, double* vdt) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m.nc) return;
  double w[5];
  load_w<double>(Q, nullptr, m.nc, c, w);
  const double a = sqrt(GAM * w[3] / w[0]);
  double lam = 0.0, lv = 0.0;
  for (int e = m.cptr[c]; e < m.cptr[c + 1]; ++e) {
    const int code = m.cfac[e];
    const bool left = code >= 0;
    const int f = left ? code : ~code;
    double wn[5];
    load_w<double>(Q, nullptr, m.nc, left ? m.iR[f] : m.iL[f], wn);
    const double an = sqrt(GAM * wn[3] / wn[0]);
    const double uL = left ? w[1] : wn[1], uR = left ? wn[1] : w[1];
    const double vL = left ? w[2] : wn[2], vR = left ? wn[2] : w[2];
    const double aL = left ? a : an, aR = left ? an : a;
    const double nx = m.fg[f], ny = m.fg[m.nfi + f], A = m.fg[2 * m.nfi + f];
    const double un = fabs(0.5 * ((uL + uR) * nx + (vL + vR) * ny));
    lam += (un + 0.5 * (aL + aR)) * A;
    lv += A * A;
  }
  for (int e = m.fptr[c]; e < m.fptr[c + 1]; ++e) {
    const int f = m.ffac[e];
    const double nx = m.ffg[f], ny = m.ffg[m.nff + f], A = m.ffg[2 * m.nff + f];
    lam += (fabs(w[1] * nx + w[2] * ny) + a) * A;
  }
  for (int e = m.wptr[c]; e < m.wptr[c + 1]; ++e) {
    const double A = m.wg[2 * m.nfw + m.wfac[e]];
    lam += a * A;
    lv += A * A;
  }
  const double T = GAM * w[3] / w[0];
  const double mu = P.mu_inf * T * sqrt(T) * (1.0 + SUTH) / (T + SUTH);
  double nut = 0.0;
  if (P.turbulent) {
    const double nt = (w[4] > 0.0 ? w[4] : 0.0) * P.nu_inf;
    const double x = nt * w[0] / mu;
    const double X3 = x * x * x;
    nut = nt * X3 / (X3 + CV1 * CV1 * CV1);
  }
  const double vol = m.cg[c];
  lam += 1.4 * (mu / w[0] / PR + nut / PRT) * lv / vol;
  const double d = cfl * vol / lam;
  dt[c] = d;
  vdt[c] = vol / d;
}

// ---- first-order Jacobian blocks (aero2d/rans.py:1032-1126) --------------------------
// Interior face: F(qL, qR) of the first-order face flux with the two-point
// face gradient; columns by forward duals. Writes the off-diagonal blocks
// (iL,iR) = dF/dqR A and (iR,iL) = -dF/dqL A into the ILU storage and keeps
// both for the diagonal gather.
__global__ void k_jac_int(MeshDev m, Params P, const double* Q, double* J, double* B, const int* posLR,
                          const int* posRL) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.nfi) return;
  using D = Dual<double>;
  const int L = m.iL[f], R = m.iR[f];
  const int nf = m.nfi;
  const double nx = m.fg[f], ny = m.fg[nf + f], A = m.fg[2 * nf + f];
  const double dist = m.fg[9 * nf + f];
  const double ex = m.fg[7 * nf + f] / dist, ey = m.fg[8 * nf + f] / dist;
  double qL[5], qR[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    qL[k] = Q[k * m.nc + L];
    qR[k] = Q[k * m.nc + R];
  }
  double* jl = J + size_t(f) * 50;
  double* jr = jl + 25;
  for (int s = 0; s < 10; ++s) {
    D a[5], b[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      a[k] = D(qL[k], s == k ? 1.0 : 0.0);
      b[k] = D(qR[k], s == k + 5 ? 1.0 : 0.0);
    }
    D wL[5], wR[5];
    prim(a, wL);
    prim(b, wR);
    D Fc[5];
    roe(P, wL, wR, nx, ny, Fc);
    const D TL = K<D>(GAM) * wL[3] / wL[0], TR = K<D>(GAM) * wR[3] / wR[0];
    auto g2 = [&](D x, D y, D out[2]) {
      D t = (y - x) / K<D>(dist);
      out[0] = t * ex;
      out[1] = t * ey;
    };
    D gu[2], gv[2], gT[2], gc[2], wf[5];
    g2(wL[1], wR[1], gu);
    g2(wL[2], wR[2], gv);
    g2(TL, TR, gT);
    g2(wL[4], wR[4], gc);
#pragma unroll
    for (int k = 0; k < 5; ++k) wf[k] = K<D>(0.5) * (wL[k] + wR[k]);
    D Fv[5];
    visc_flux(P, wf, gu, gv, gT, gc, nx, ny, Fv);
    double* col = s < 5 ? jl : jr;
    const int j = s < 5 ? s : s - 5;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      double v = (Fc[i].d - Fv[i].d) * A;
      if (!P.turbulent && i == 4) v = 0.0;
      col[i * 5 + j] = v;
    }
  }
  if (B)
    for (int t = 0; t < 25; ++t) {
      B[size_t(posLR[f]) * 25 + t] = jr[t];
      B[size_t(posRL[f]) * 25 + t] = -jl[t];
    }
}

// Wall face flux Jacobian (no slip, adiabatic, one-sided gradients).
__global__ void k_jac_wall(MeshDev m, Params P, const double* Q, double* JW) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.nfw) return;
  using D = Dual<double>;
  const int c = m.cw[f];
  const int nw = m.nfw;
  const double nx = m.wg[f], ny = m.wg[nw + f], A = m.wg[2 * nw + f], dn = m.wg[5 * nw + f];
  for (int s = 0; s < 5; ++s) {
    D q[5], wi[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) q[k] = D(Q[k * m.nc + c], s == k ? 1.0 : 0.0);
    prim(q, wi);
    D p = wi[3];
    auto gw = [&](D x, D out[2]) {
      D t = (K<D>(0.0) - x) / K<D>(dn);
      out[0] = t * nx;
      out[1] = t * ny;
    };
    D gu[2], gv[2], gc[2];
    D gT[2] = {K<D>(0.0), K<D>(0.0)};
    gw(wi[1], gu);
    gw(wi[2], gv);
    gw(wi[4], gc);
    D wf[5] = {wi[0], K<D>(0.0), K<D>(0.0), p, K<D>(0.0)};
    D Fv[5];
    visc_flux(P, wf, gu, gv, gT, gc, nx, ny, Fv);
    D Fc[5] = {K<D>(0.0), p * nx, p * ny, K<D>(0.0), K<D>(0.0)};
#pragma unroll
    for (int i = 0; i < 5; ++i) JW[size_t(f) * 25 + i * 5 + s] = (Fc[i].d - Fv[i].d) * A;
  }
}

// Far-field face flux Jacobian (Roe flux against the fixed boundary state).
__global__ void k_jac_far(MeshDev m, Params P, const double* Q, const double* wgf, double* JF) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.nff) return;
  using D = Dual<double>;
  const int c = m.cf[f];
  const int n = m.nff;
  const double nx = m.ffg[f], ny = m.ffg[n + f], A = m.ffg[2 * n + f];
  D wb[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) wb[k] = K<D>(wgf[k * n + f]);
  for (int s = 0; s < 5; ++s) {
    D q[5], wi[5], F[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) q[k] = D(Q[k * m.nc + c], s == k ? 1.0 : 0.0);
    prim(q, wi);
    roe(P, wi, wb, nx, ny, F);
#pragma unroll
    for (int i = 0; i < 5; ++i) JF[size_t(f) * 25 + i * 5 + s] = F[i].d * A;
  }
}

// Diagonal blocks: sum of the cell's face blocks, the decreasing part of the
// SA source on (4,4) at frozen gradients, vol/dt on the diagonal (vol/(dt ct)
// on the turbulence equation) (aero2d/rans.py:1107-1144).
__global__ void k_jac_diag(MeshDev m, Params P, const double* Q, const double* g, const double* J,
                           const double* JW, const double* JF, const double* vdt, double ct_inv, double* B,
                           const int* posD, double* Dout) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m.nc) return;
  double Dm[25];
#pragma unroll
  for (int t = 0; t < 25; ++t) Dm[t] = 0.0;
  for (int e = m.cptr[c]; e < m.cptr[c + 1]; ++e) {
    const int code = m.cfac[e];
    if (code >= 0) {
      const double* jl = J + size_t(code) * 50;
#pragma unroll
      for (int t = 0; t < 25; ++t) Dm[t] += jl[t];
    } else {
      const double* jr = J + size_t(~code) * 50 + 25;
#pragma unroll
      for (int t = 0; t < 25; ++t) Dm[t] -= jr[t];
    }
  }
  for (int e = m.wptr[c]; e < m.wptr[c + 1]; ++e) {
    const double* jw = JW + size_t(m.wfac[e]) * 25;
#pragma unroll
    for (int t = 0; t < 25; ++t) Dm[t] += jw[t];
  }
  for (int e = m.fptr[c]; e < m.fptr[c + 1]; ++e) {
    const double* jf = JF + size_t(m.ffac[e]) * 25;
#pragma unroll
    for (int t = 0; t < 25; ++t) Dm[t] += jf[t];
  }
  const double vol = m.cg[c];
  if (P.turbulent) {
    using D = Dual<double>;
    D q[5], w[5], G[5][2];
#pragma unroll
    for (int k = 0; k < 5; ++k) q[k] = D(Q[k * m.nc + c], k == 4 ? 1.0 : 0.0);
    prim(q, w);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      G[k][0] = K<D>(g[(2 * k) * m.nc + c]);
      G[k][1] = K<D>(g[(2 * k + 1) * m.nc + c]);
    }
    const double ds = sa_source(P, w, G, m.cg[m.nc + c]).d;
    Dm[24] += vol * (-ds > 0.0 ? -ds : 0.0);
  } else {
#pragma unroll
    for (int j = 0; j < 5; ++j) Dm[20 + j] = 0.0;
    Dm[24] = vol;
  }
  if (Dout) {
#pragma unroll
    for (int t = 0; t < 25; ++t) Dout[size_t(c) * 25 + t] = Dm[t];
  }
  if (vdt) {
    const double d = vdt[c];
#pragma unroll
    for (int k = 0; k < 5; ++k) Dm[k * 6] += k == 4 ? d * ct_inv : d;
  }
  if (B) {
#pragma unroll
    for (int t = 0; t < 25; ++t) B[size_t(posD[c]) * 25 + t] = Dm[t];
  }
}

// ---- block ILU in level order ------------------------------------------------------
// One workgroup walks the levels; rows of a level are independent and the
// workgroup barrier separates levels. Level bounds are staged in LDS and each
// row's work is described by one 16-byte record (row, first entry, end entry),
// so a level costs a record load, the gathers and the barrier.
//
// Factorisation: one wavefront per row. Each half-wavefront holds the current
// lower block L_ik = B_ik invU_k, lane t < 25 owning element (t/5, t%5) (row
// elements exchanged by lane shuffles), and the two halves take alternate
// update pairs (i,j) -= L_ik (k,j). A fence orders the lanes' block updates
// before the next lower block and the diagonal are read; lane 0 inverts the
// diagonal block (Gauss-Jordan, partial pivoting).
struct IluDev {
  int n, nnz;
  const int *rdiag, *col;
  const int *pptr, *pt, *pe2;        // per lower entry: its update pairs (target, row-k entry)
  int nlev, nlevu;
  const int *lptr, *uptr;            // level bounds into frec / brec
  const int4 *frec, *brec;           // frec (row, row start, diagonal, row end) by forward
                                     // level; brec (row, first upper, row end, 0) by backward level
  const int* perm;                   // perm[new] = old
  int lds_levels;                    // level bounds fit in LDS
  int rowcap, paircap, lowcap;       // longest row (blocks), most update pairs and most lower
                                     // blocks of one row
  int lds_waves;                     // wavefronts of the LDS factorisation (0: use the global one)
  // Staged copy S of the factors for the LDS triangular sweeps: per level, the
  // blocks its rows read, contiguous (forward: each row's lower blocks;
  // backward: each row's upper blocks, then its inverse diagonal block).
  const int4 *fsrec, *bsrec;         // (row, first entry, end entry, the row's first block
                                     // in its level's run) by forward / backward level
  const int *flev, *blev;            // first staged block of each level (blev after flev's)
  const int* ssrc;                   // source of staged block b: entry e < nnz of B, or
                                     // nnz + i for invU of row i
  const int* scol;                   // column of staged block b (0 for an inverse)
  int nst;                           // staged blocks (nnz: every block once, inverses included)
  int scap;                          // longest level run, in doubles
  int ccap, rcap;                    // LDS slots per level for its columns (scap / 25, a
                                     // multiple of 4) and its row records
  int lds_solve;                     // 1: the sweeps stage levels in LDS (k_ilu_solve_lds)
};

constexpr int ILU_TPB = 1024;
constexpr int PAIR_SLOTS = 12;          // update pairs in flight per wavefront (5 lanes each)
constexpr int STAGE_K = 12;             // loads per lane in one batch of the row staging
constexpr int ILU_LDS_LEVELS = 12288;   // level bounds staged in LDS up to this many

__device__ __forceinline__ const int* stage_bounds(const int* g, int n, int* sh, bool use) {
  if (!use) return g;
  for (int t = threadIdx.x; t < n; t += blockDim.x) sh[t] = g[t];
  __syncthreads();
  return sh;
}

// Batched form: workgroup b works on case cases[b] (the cases share the
// structure; their blocks, inverses and vectors are sB, sI, sY apart), or on
// the arrays as given when cases is null.
__global__ __launch_bounds__(ILU_TPB) void k_ilu_factor(IluDev L, double* B, double* invU, const int* cases,
                                                        size_t sB, size_t sI) {
  if (cases) {
    B += size_t(cases[blockIdx.x]) * sB;
    invU += size_t(cases[blockIdx.x]) * sI;
  }
  extern __shared__ int sh_lev[];
  const int* lp = stage_bounds(L.lptr, L.nlev + 1, sh_lev, L.lds_levels);
  const int wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const int half = (threadIdx.x >> 5) & 1, hl = threadIdx.x & 31;
  const bool act = hl < 25;
  const int r = act ? hl / 5 : 4, cc = act ? hl % 5 : 4;
  for (int l = 0; l < L.nlev; ++l) {
    const int p1 = lp[l + 1];
    for (int p = lp[l] + wave; p < p1; p += nwave) {
      const int4 rec = L.frec[p];
      const int i = rec.x, d = L.rdiag[i];
      for (int e = rec.y; e < rec.z; ++e) {
        const int k = L.col[e];
        const double be = B[size_t(e) * 25 + r * 5 + cc];
        double s = 0.0;
#pragma unroll
        for (int mm = 0; mm < 5; ++mm) s += __shfl(be, r * 5 + mm, 32) * invU[size_t(k) * 25 + mm * 5 + cc];
        if (act && half == 0) B[size_t(e) * 25 + hl] = s;
        const int q1 = L.pptr[e + 1];
        for (int q = L.pptr[e] + half; q < q1; q += 2) {
          const size_t t = size_t(L.pt[q]) * 25, e2 = size_t(L.pe2[q]) * 25;
          double acc = 0.0;
#pragma unroll
          for (int mm = 0; mm < 5; ++mm) acc += __shfl(s, r * 5 + mm, 32) * B[e2 + mm * 5 + cc];
          if (act) B[t + hl] -= acc;
        }
        __threadfence_block();   // this row's block updates, by either half, before the next reads
      }
      // invU_i = B_ii^-1: Gauss-Jordan with partial pivoting, lane (r, cc)
      // holding element (r, cc) of the block and of the inverse being built
      double av = B[size_t(d) * 25 + r * 5 + cc];
      double iv = r == cc ? 1.0 : 0.0;
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        int piv = c;
        double best = fabs(__shfl(av, c * 5 + c, 32));
#pragma unroll
        for (int x = c + 1; x < 5; ++x) {
          const double v = fabs(__shfl(av, x * 5 + c, 32));
          if (v > best) {
            best = v;
            piv = x;
          }
        }
        if (piv != c) {
          const int src = r == c ? piv : (r == piv ? c : r);
          av = __shfl(av, src * 5 + cc, 32);
          iv = __shfl(iv, src * 5 + cc, 32);
        }
        const double inv = 1.0 / __shfl(av, c * 5 + c, 32);
        if (r == c) {
          av *= inv;
          iv *= inv;
        }
        const double f = __shfl(av, r * 5 + c, 32);
        const double ac = __shfl(av, c * 5 + cc, 32), ic = __shfl(iv, c * 5 + cc, 32);
        if (r != c) {
          av -= f * ac;
          iv -= f * ic;
        }
      }
      if (act && half == 0) invU[size_t(i) * 25 + hl] = iv;
    }
    __syncthreads();
  }
}

// Ordering of this wavefront's LDS accesses between lanes (LDS operations of
// one wavefront complete in order; the fences keep the compiler from moving
// accesses across).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The factorisation with each row staged in LDS: a wavefront copies its row's
// blocks and the row's update-pair list into LDS (coalesced), runs the
// elimination there (the same products in the same order as k_ilu_factor,
// with twelve update pairs in flight per wavefront: five lanes per pair, one
// column each), and writes the row back. The only global loads inside the elimination are
// the finished blocks and inverses of earlier rows, which no store of the
// row can alias, so they are issued early instead of waiting on the row's
// own updates.
__global__ __launch_bounds__(ILU_TPB) void k_ilu_factor_lds(IluDev L, double* B, double* invU, const int* cases,
                                                            size_t sB, size_t sI) {
  if (cases) {
    B += size_t(cases[blockIdx.x]) * sB;
    invU += size_t(cases[blockIdx.x]) * sI;
  }
  extern __shared__ double smem[];
  const int wave = threadIdx.x >> 6, nwave = blockDim.x >> 6, lane = threadIdx.x & 63;
  double* rowb = smem + size_t(wave) * L.rowcap * 25;
  double* rowi = smem + size_t(nwave) * L.rowcap * 25 + size_t(wave) * L.lowcap * 25;
  int* pairs_all = reinterpret_cast<int*>(smem + size_t(nwave) * (L.rowcap + L.lowcap) * 25);
  int* rowp = pairs_all + size_t(wave) * L.paircap * 2;
  int* rowq = pairs_all + size_t(nwave) * L.paircap * 2 + size_t(wave) * (L.lowcap + 1);
  int* lev_sh = pairs_all + size_t(nwave) * (L.paircap * 2 + L.lowcap + 1);
  const int* lp = stage_bounds(L.lptr, L.nlev + 1, lev_sh, L.lds_levels);
  const int half = lane >> 5, hl = lane & 31;
  const bool act = hl < 25;
  const int r = act ? hl / 5 : 4, cc = act ? hl % 5 : 4;
  const int pslot = lane / 5, pc = lane - 5 * pslot;
  for (int l = 0; l < L.nlev; ++l) {
    const int p1 = lp[l + 1];
    for (int p = lp[l] + wave; p < p1; p += nwave) {
      const int4 rec = L.frec[p];
      const int i = rec.x, rs = rec.y, d = rec.z, re = rec.w;
      const int nb25 = (re - rs) * 25, nl25 = (d - rs) * 25;
      const int q0 = L.pptr[rs], q1 = L.pptr[d];
      // stage the row: its blocks, the inverses of its lower entries' rows
      // (finished at earlier levels), its update pairs and their bounds. Each
      // batch issues all of a lane's loads (clamped indices, unconditional)
      // before its LDS stores, so a batch costs one memory latency.
      for (int t0 = 0; t0 < nb25; t0 += 64 * STAGE_K) {
        double v[STAGE_K];
#pragma unroll
        for (int k = 0; k < STAGE_K; ++k) v[k] = B[size_t(rs) * 25 + min(t0 + lane + 64 * k, nb25 - 1)];
#pragma unroll
        for (int k = 0; k < STAGE_K; ++k)
          if (t0 + lane + 64 * k < nb25) rowb[t0 + lane + 64 * k] = v[k];
      }
      for (int t0 = 0; t0 < nl25; t0 += 64 * STAGE_K) {
        int src[STAGE_K];
#pragma unroll
        for (int k = 0; k < STAGE_K; ++k) {
          const int t = min(t0 + lane + 64 * k, nl25 - 1);
          src[k] = L.col[rs + t / 25] * 25 + t % 25;
        }
        double v[STAGE_K];
#pragma unroll
        for (int k = 0; k < STAGE_K; ++k) v[k] = invU[size_t(src[k])];
#pragma unroll
        for (int k = 0; k < STAGE_K; ++k)
          if (t0 + lane + 64 * k < nl25) rowi[t0 + lane + 64 * k] = v[k];
      }
      for (int t0 = 0; t0 < q1 - q0; t0 += 64 * STAGE_K) {
        int a[STAGE_K], b[STAGE_K];
#pragma unroll
        for (int k = 0; k < STAGE_K; ++k) {
          const int t = min(t0 + lane + 64 * k, q1 - q0 - 1);
          a[k] = L.pt[q0 + t];
          b[k] = L.pe2[q0 + t];
        }
#pragma unroll
        for (int k = 0; k < STAGE_K; ++k)
          if (t0 + lane + 64 * k < q1 - q0) {
            rowp[2 * (t0 + lane + 64 * k)] = a[k] - rs;
            rowp[2 * (t0 + lane + 64 * k) + 1] = b[k];
          }
      }
      for (int t = lane; t <= d - rs; t += 64) rowq[t] = L.pptr[rs + t] - q0;
      wave_sync();
      for (int e = rs; e < d; ++e) {
        // row r of B_ik and column cc of invU_k straight from LDS (the values
        // the plain kernel moves by shuffles; all ten reads in flight at once)
        const double* bik = rowb + (e - rs) * 25 + r * 5;
        const double* ik = rowi + (e - rs) * 25 + cc;
        double sv = 0.0;
#pragma unroll
        for (int mm = 0; mm < 5; ++mm) sv += bik[mm] * ik[mm * 5];
        wave_sync();
        if (act && half == 0) rowb[(e - rs) * 25 + hl] = sv;
        wave_sync();
        // update pairs twelve at a time: lane (slot, column pc) forms column pc of
        // L_ik U_kj for its pair (L_ik read back from the row in LDS, the same
        // address for every lane), each element in the order of the plain kernel
        const double* lk = rowb + (e - rs) * 25;
        const int qb = rowq[e - rs + 1];
        if (lane < 5 * PAIR_SLOTS) {
          for (int q = rowq[e - rs] + pslot; q < qb; q += PAIR_SLOTS) {
            const int tl = rowp[2 * q], e2 = rowp[2 * q + 1];
            double u[5];
#pragma unroll
            for (int mm = 0; mm < 5; ++mm) u[mm] = B[size_t(e2) * 25 + mm * 5 + pc];
#pragma unroll
            for (int rr = 0; rr < 5; ++rr) {
              double acc = 0.0;
#pragma unroll
              for (int mm = 0; mm < 5; ++mm) acc += lk[rr * 5 + mm] * u[mm];
              rowb[tl * 25 + rr * 5 + pc] -= acc;
            }
          }
        }
        wave_sync();
      }
      double av = rowb[(d - rs) * 25 + r * 5 + cc];
      double iv = r == cc ? 1.0 : 0.0;
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        int piv = c;
        double best = fabs(__shfl(av, c * 5 + c, 32));
#pragma unroll
        for (int x = c + 1; x < 5; ++x) {
          const double v = fabs(__shfl(av, x * 5 + c, 32));
          if (v > best) {
            best = v;
            piv = x;
          }
        }
        if (piv != c) {
          const int src = r == c ? piv : (r == piv ? c : r);
          av = __shfl(av, src * 5 + cc, 32);
          iv = __shfl(iv, src * 5 + cc, 32);
        }
        const double inv = 1.0 / __shfl(av, c * 5 + c, 32);
        if (r == c) {
          av *= inv;
          iv *= inv;
        }
        const double f = __shfl(av, r * 5 + c, 32);
        const double ac = __shfl(av, c * 5 + cc, 32), ic = __shfl(iv, c * 5 + cc, 32);
        if (r != c) {
          av -= f * ac;
          iv -= f * ic;
        }
      }
      if (act && half == 0) invU[size_t(i) * 25 + hl] = iv;
      for (int t = lane; t < nb25; t += 64) B[size_t(rs) * 25 + t] = rowb[t];
      wave_sync();
    }
    __syncthreads();
  }
}

// x (SoA [5][n], original order) -> y ([n][5], permuted order).
__global__ void k_perm_in(int n, const int* perm, const double* b, double* y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = perm[i];
#pragma unroll
  for (int k = 0; k < 5; ++k) y[size_t(i) * 5 + k] = b[size_t(k) * n + o];
}
__global__ void k_perm_out(int n, const int* perm, const double* y, double* x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = perm[i];
#pragma unroll
  for (int k = 0; k < 5; ++k) x[size_t(k) * n + o] = y[size_t(i) * 5 + k];
}

// Forward (unit lower) and backward (upper, then invU) sweeps over the levels,
// in place on y (aero2d/bilu.py:148-166). Eight lanes share a row: each takes
// every eighth of the row's blocks and forms its 5x5 block-vector products,
// the eight partial sums are combined by a butterfly of lane shuffles, and
// lane k < 5 writes component k.
// Bb, colb: the block and the column of entry e0 (the row's follow them).
__device__ __forceinline__ void row_products(const double* Bb, const int* colb, const double* y, int e0, int e1,
                                             int lane, double a[5]) {
#pragma unroll
  for (int k = 0; k < 5; ++k) a[k] = 0.0;
  for (int e = e0 + lane; e < e1; e += 8) {
    const double* yc = y + size_t(colb[e - e0]) * 5;
    const double* b = Bb + size_t(e - e0) * 25;
    const double y0 = yc[0], y1 = yc[1], y2 = yc[2], y3 = yc[3], y4 = yc[4];
#pragma unroll
    for (int k = 0; k < 5; ++k)
      a[k] += b[k * 5] * y0 + b[k * 5 + 1] * y1 + b[k * 5 + 2] * y2 + b[k * 5 + 3] * y3 + b[k * 5 + 4] * y4;
  }
#pragma unroll
  for (int off = 4; off > 0; off >>= 1)
#pragma unroll
    for (int k = 0; k < 5; ++k) a[k] += __shfl_xor(a[k], off, 8);
}

__global__ __launch_bounds__(ILU_TPB) void k_ilu_solve(IluDev L, const double* B, const double* invU,
                                                       double* y, const int* cases, size_t sB, size_t sI,
                                                       size_t sY) {
  if (cases) {
    B += size_t(cases[blockIdx.x]) * sB;
    invU += size_t(cases[blockIdx.x]) * sI;
    y += size_t(cases[blockIdx.x]) * sY;
  }
  extern __shared__ int sh_lev[];
  const bool use = L.lds_levels;
  const int* lp = stage_bounds(L.lptr, L.nlev + 1, sh_lev, use);
  const int* up = stage_bounds(L.uptr, L.nlevu + 1, sh_lev + (use ? L.nlev + 1 : 0), use);
  const int grp = threadIdx.x >> 3, lane = threadIdx.x & 7, ngrp = blockDim.x >> 3;
  double a[5];
  for (int l = 0; l < L.nlev; ++l) {
    const int p1 = lp[l + 1];
    for (int p = lp[l] + grp; p < p1; p += ngrp) {
      const int4 rec = L.frec[p];
      row_products(B + size_t(rec.y) * 25, L.col + rec.y, y, rec.y, rec.z, lane, a);
      double mine = a[0];
#pragma unroll
      for (int k = 1; k < 5; ++k) mine = lane == k ? a[k] : mine;
      if (lane < 5) y[size_t(rec.x) * 5 + lane] -= mine;
    }
    __syncthreads();
  }
  for (int l = 0; l < L.nlevu; ++l) {
    const int p1 = up[l + 1];
    for (int p = up[l] + grp; p < p1; p += ngrp) {
      const int4 rec = L.brec[p];
      const int i = rec.x;
      row_products(B + size_t(rec.y) * 25, L.col + rec.y, y, rec.y, rec.z, lane, a);
      double t[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) t[k] = y[size_t(i) * 5 + k] - a[k];
      const int kk = lane < 5 ? lane : 4;
      const double* iu = invU + size_t(i) * 25 + kk * 5;
      const double xi = iu[0] * t[0] + iu[1] * t[1] + iu[2] * t[2] + iu[3] * t[3] + iu[4] * t[4];
      if (lane < 5) y[size_t(i) * 5 + lane] = xi;
    }
    __syncthreads();
  }
}

// S <- the factors' blocks in level order (a gather; grid.y over the cases).
__global__ void k_ilu_stage(IluDev L, const double* B, const double* invU, double* S, const int* cases,
                            size_t sB, size_t sI, size_t sS) {
  if (cases) {
    const size_t c = size_t(cases[blockIdx.y]);
    B += c * sB;
    invU += c * sI;
    S += c * sS;
  }
  const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= size_t(L.nst) * 25) return;
  const int b = int(t / 25), k = int(t - size_t(b) * 25);
  const int id = L.ssrc[b];
  S[t] = id < L.nnz ? B[size_t(id) * 25 + k] : invU[size_t(id - L.nnz) * 25 + k];
}

constexpr int SOLVE_K = 16;   // doubles a staging thread carries: a level run is at most
                              // SOLVE_K * ILU_TPB / 2 doubles

// Copy n doubles src -> dst by nt threads (t this thread's index): all loads
// first, then the LDS stores, so a thread has its loads in flight together.
__device__ __forceinline__ void stage_run(const double* __restrict__ src, int n, double* dst, int t, int nt) {
  double v[SOLVE_K];
#pragma unroll
  for (int k = 0; k < SOLVE_K; ++k) {
    const int idx = t + k * nt;
    v[k] = idx < n ? src[idx] : 0.0;
  }
#pragma unroll
  for (int k = 0; k < SOLVE_K; ++k) {
    const int idx = t + k * nt;
    if (idx < n) dst[idx] = v[k];
  }
}

// The sweeps of k_ilu_solve with the factors read from LDS: the upper half of
// the workgroup copies the next level's run of S, with its row records and
// block columns, into one set of LDS buffers while the lower half works on
// the current level from the other (eight lanes per row, the same products in
// the same order as k_ilu_solve), so a level costs the loads of y and the
// barrier; the loads of factors, records and columns (the record and column
// loads would otherwise be two dependent global round trips ahead of the y
// loads) overlap the previous level. Separate wavefronts keep the copy's
// loads out of the workers' wait counts.
__global__ __launch_bounds__(ILU_TPB) void k_ilu_solve_lds(IluDev L, const double* S, double* y, const int* cases,
                                                           size_t sS, size_t sY) {
  if (cases) {
    S += size_t(cases[blockIdx.x]) * sS;
    y += size_t(cases[blockIdx.x]) * sY;
  }
  extern __shared__ double smem[];
  double* buf0 = smem;
  double* buf1 = smem + L.scap;
  int4* rec0 = reinterpret_cast<int4*>(smem + 2 * size_t(L.scap));
  int4* rec1 = rec0 + L.rcap;
  int* col0 = reinterpret_cast<int*>(rec1 + L.rcap);
  int* col1 = col0 + L.ccap;
  int* lev_sh = col1 + L.ccap;
  const bool use = L.lds_levels;
  const int nb2 = use ? L.nlev + L.nlevu + 2 : 0;
  const int* lp = stage_bounds(L.lptr, L.nlev + 1, lev_sh, use);
  const int* up = stage_bounds(L.uptr, L.nlevu + 1, lev_sh + (use ? L.nlev + 1 : 0), use);
  // the staged runs' bounds too: a copier's first load of a level is then
  // from LDS, not a dependent global load ahead of its block loads
  const int* fl = stage_bounds(L.flev, L.nlev + 1, lev_sh + nb2, use);
  const int* bl = stage_bounds(L.blev, L.nlevu + 1, lev_sh + nb2 + (use ? L.nlev + 1 : 0), use);
  const int nw = blockDim.x >> 1;                 // workers: threads [0, nw); copiers: [nw, 2 nw)
  const bool copier = threadIdx.x >= nw;
  const int ct = threadIdx.x - nw;
  const int grp = threadIdx.x >> 3, lane = threadIdx.x & 7, ngrp = nw >> 3;
  const int nl = L.nlev + L.nlevu;
  // run of level q of the whole sequence (forward levels, then backward
  // ones): staged blocks [b0, b1), row records [p0, p1) of recs
  auto run = [&](int q, int& b0, int& b1, int& p0, int& p1, const int4*& recs) {
    if (q < L.nlev) {
      b0 = fl[q];
      b1 = fl[q + 1];
      p0 = lp[q];
      p1 = lp[q + 1];
      recs = L.fsrec;
    } else {
      const int l = q - L.nlev;
      b0 = bl[l];
      b1 = bl[l + 1];
      p0 = up[l];
      p1 = up[l + 1];
      recs = L.bsrec;
    }
  };
  // copiers: level q's blocks, columns and records into buffer set q & 1
  auto stage = [&](int q) {
    int b0, b1, p0, p1;
    const int4* recs;
    run(q, b0, b1, p0, p1, recs);
    int4* rd = (q & 1) ? rec1 : rec0;
    int* cd = (q & 1) ? col1 : col0;
    const int nb = b1 - b0, np = p1 - p0;
    // the small copies' loads first, so they are in flight with the blocks'
    int cv = 0;
    int4 rv = make_int4(0, 0, 0, 0);
    if (ct < nb) cv = L.scol[b0 + ct];
    if (ct < np) rv = recs[p0 + ct];
    stage_run(S + size_t(b0) * 25, nb * 25, (q & 1) ? buf1 : buf0, ct, nw);
    if (ct < nb) cd[ct] = cv;
    if (ct < np) rd[ct] = rv;
    for (int t = ct + nw; t < nb; t += nw) cd[t] = L.scol[b0 + t];
    for (int t = ct + nw; t < np; t += nw) rd[t] = recs[p0 + t];
  };
  if (copier) stage(0);
  __syncthreads();
  double a[5];
  for (int q = 0; q < nl; ++q) {
    const double* cur = (q & 1) ? buf1 : buf0;
    const int4* rcur = (q & 1) ? rec1 : rec0;
    const int* ccur = (q & 1) ? col1 : col0;
    if (copier) {
      if (q + 1 < nl) stage(q + 1);
    } else if (q < L.nlev) {
      const int p0 = lp[q], p1 = lp[q + 1];
      for (int p = p0 + grp; p < p1; p += ngrp) {
        const int4 rec = rcur[p - p0];
        row_products(cur + size_t(rec.w) * 25, ccur + rec.w, y, rec.y, rec.z, lane, a);
        double mine = a[0];
#pragma unroll
        for (int k = 1; k < 5; ++k) mine = lane == k ? a[k] : mine;
        if (lane < 5) y[size_t(rec.x) * 5 + lane] -= mine;
      }
    } else {
      const int l = q - L.nlev, p0 = up[l], p1 = up[l + 1];
      for (int p = p0 + grp; p < p1; p += ngrp) {
        const int4 rec = rcur[p - p0];
        const int i = rec.x;
        row_products(cur + size_t(rec.w) * 25, ccur + rec.w, y, rec.y, rec.z, lane, a);
        double t[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) t[k] = y[size_t(i) * 5 + k] - a[k];
        const int kk = lane < 5 ? lane : 4;
        const double* iu = cur + (size_t(rec.w) + size_t(rec.z - rec.y)) * 25 + kk * 5;
        const double xi = iu[0] * t[0] + iu[1] * t[1] + iu[2] * t[2] + iu[3] * t[3] + iu[4] * t[4];
        if (lane < 5) y[size_t(i) * 5 + lane] = xi;
      }
    }
    __syncthreads();
  }
}

size_t ilu_solve_lds_bytes(const IluDev& L) {
  return 2 * size_t(L.scap) * sizeof(double) + 2 * size_t(L.rcap) * sizeof(int4) +
         2 * size_t(L.ccap) * sizeof(int) + (L.lds_levels ? 2 * size_t(L.nlev + L.nlevu + 2) * sizeof(int) : 0);
}

// One factorisation launch over ncase workgroups (ncase = 1 with cases null):
// the LDS-staged kernel when the rows fit, else the global one; then, for the
// LDS sweeps, the staged copy S of the factors.
void launch_ilu_factor(const IluDev& L, double* B, double* invU, double* S, const int* cases, int ncase,
                       size_t sB, size_t sI, size_t sS, hipStream_t st) {
  const size_t lev = L.lds_levels ? size_t(L.nlev + 1) * sizeof(int) : 0;
  if (L.lds_waves > 0) {
    const size_t lds = size_t(L.lds_waves) * (size_t(L.rowcap + L.lowcap) * 25 * sizeof(double) +
                                              size_t(L.paircap) * 8 + size_t(L.lowcap + 1) * 4) + lev;
    hipLaunchKernelGGL(k_ilu_factor_lds, ncase, L.lds_waves * 64, lds, st, L, B, invU, cases, sB, sI);
  } else {
    hipLaunchKernelGGL(k_ilu_factor, ncase, ILU_TPB, lev, st, L, B, invU, cases, sB, sI);
  }
  if (L.lds_solve) {
    const size_t nt = size_t(L.nst) * 25;
    hipLaunchKernelGGL(k_ilu_stage, dim3(unsigned((nt + 255) / 256), unsigned(ncase)), 256, 0, st, L,
                       (const double*)B, (const double*)invU, S, cases, sB, sI, sS);
  }
}

// ---- vector kernels -----------------------------------------------------------------
constexpr int VCHUNK = 2048;

// part[j][b] = sum over block b's chunk of V_j . w, for j < nvec.
__global__ void k_mdot(int N, const double* Vb, size_t ldv, int nvec, const double* w, double* part) {
  __shared__ double sh[TPB];
  const int j = blockIdx.y;
  const double* v = Vb + size_t(j) * ldv;
  const int b0 = blockIdx.x * VCHUNK;
  const int b1 = min(N, b0 + VCHUNK);
  double s = 0.0;
  for (int i = b0 + threadIdx.x; i < b1; i += TPB) s += v[i] * w[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int t = TPB / 2; t > 0; t >>= 1) {
    if (threadIdx.x < t) sh[threadIdx.x] += sh[threadIdx.x + t];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[size_t(j) * gridDim.x + blockIdx.x] = sh[0];
  (void)nvec;
}

// w -= sum_j h_j V_j (the sum first, in order of j).
__global__ void k_gs_update(int N, const double* Vb, size_t ldv, int nvec, const double* h, double* w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double s = 0.0;
  for (int j = 0; j < nvec; ++j) s += h[j] * Vb[size_t(j) * ldv + i];
  w[i] -= s;
}

// out = a * x (scale) or out = x + a*y style helpers.
__global__ void k_scale(int N, double a, const double* x, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) out[i] = a * x[i];
}
__global__ void k_negate(int N, const double* x, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) out[i] = -x[i];
}
__global__ void k_add(int N, const double* x, const double* y, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) out[i] = x[i] + y[i];
}

// Relative decrease of rho and p under the full update, per cell
// (aero2d/rans.py:1822-1826); block partial max and count of cells >= 0.2.
__global__ void k_step_rel(int nc, const double* Q, const double* dq, double* reli, double* pmax,
                           double* pcnt) {
  __shared__ double smax[TPB], scnt[TPB];
  const int c = blockIdx.x * TPB + threadIdx.x;
  double rl = -INFINITY, cnt = 0.0;
  if (c < nc) {
    double q[5], qn[5], w[5], wn[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      q[k] = Q[k * nc + c];
      qn[k] = q[k] + dq[k * nc + c];
    }
    prim(q, w);
    prim(qn, wn);
    const double a = -(wn[0] - w[0]) / w[0], b = -(wn[3] - w[3]) / w[3];
    rl = a > b ? a : b;
    if (a != a || b != b) rl = NAN;
    reli[c] = rl;
    cnt = rl >= 0.2 ? 1.0 : 0.0;
  }
  smax[threadIdx.x] = rl;
  scnt[threadIdx.x] = cnt;
  __syncthreads();
  for (int t = TPB / 2; t > 0; t >>= 1) {
    if (threadIdx.x < t) {
      const double o = smax[threadIdx.x + t], s = smax[threadIdx.x];
      smax[threadIdx.x] = (o != o || s != s) ? NAN : (o > s ? o : s);
      scnt[threadIdx.x] += scnt[threadIdx.x + t];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pmax[blockIdx.x] = smax[0];
    pcnt[blockIdx.x] = scnt[0];
  }
}

__global__ void k_max_rows(int nparts, const double* pmax, const double* pcnt, double* out) {
  if (threadIdx.x != 0) return;
  double m = -INFINITY, c = 0.0;
  for (int p = 0; p < nparts; ++p) {
    const double x = pmax[p];
    m = (x != x || m != m) ? NAN : (x > m ? x : m);
    c += pcnt[p];
  }
  out[0] = m;
  out[1] = c;
}

// Q += lam dq (whole step) or Q += lam_c dq with lam_c = 1 below 0.2 and
// 0.2 / rel above (only the few cells that ask for more).
__global__ void k_apply(int nc, double* Q, const double* dq, const double* reli, double lam, int per_cell) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double s = lam;
  if (per_cell) {
    const double r = reli[c];
    s = r < 0.2 ? 1.0 : 0.2 / (r > 1e-300 ? r : 1e-300);
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) Q[k * nc + c] += s * dq[k * nc + c];
}

// Wall forces by faces (aero2d/rans.py:1251-1296): per face the pressure and
// friction forces and the moment arm, summed in order by one block.
__global__ void k_forces(MeshDev m, Params P, const double* Q, const double* g, double ca, double sa,
                         double* out, double* cp, double* cfo) {
  __shared__ double sh[6][TPB];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const double q = 0.5 * P.mach * P.mach;
  const int nw = m.nfw;
  for (int f = threadIdx.x; f < nw; f += TPB) {
    const int c = m.cw[f];
    double w[5];
    load_w<double>(Q, nullptr, m.nc, c, w);
    const double nx = m.wg[f], ny = m.wg[nw + f], A = m.wg[2 * nw + f];
    const double rx = m.wg[3 * nw + f], ry = m.wg[4 * nw + f], dn = m.wg[5 * nw + f];
    const double pw = w[3] + (g[6 * m.nc + c] * rx + g[7 * m.nc + c] * ry);
    const double T = GAM * w[3] / w[0];
    const double mu = P.mu_inf * T * sqrt(T) * (1.0 + SUTH) / (T + SUTH);
    const double ut = w[1] * (-ny) + w[2] * nx;
    const double tau = mu * ut / dn;
    const double fxp = (pw - 1.0 / GAM) * nx * A, fyp = (pw - 1.0 / GAM) * ny * A;
    const double fxv = tau * (-ny) * A, fyv = tau * nx * A;
    const double xm = m.wg[7 * nw + f] - P.xref_x, ym = m.wg[8 * nw + f] - P.xref_y;
    acc[0] += -fxp * sa + fyp * ca;
    acc[1] += fxp * ca + fyp * sa;
    acc[2] += -fxv * sa + fyv * ca;
    acc[3] += fxv * ca + fyv * sa;
    acc[4] += xm * (fyp + fyv) - ym * (fxp + fxv);
    if (cp) cp[f] = (pw - 1.0 / GAM) / q;
    if (cfo) cfo[f] = tau / q;
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) sh[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int t = TPB / 2; t > 0; t >>= 1) {
    if (threadIdx.x < t)
#pragma unroll
      for (int k = 0; k < 5; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + t];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double qc = q * P.chord;
    const double lp = sh[0][0] / qc, dp = sh[1][0] / qc, lv = sh[2][0] / qc, dv = sh[3][0] / qc;
    out[0] = lp + lv;
    out[1] = dp + dv;
    out[2] = dp;
    out[3] = dv;
    out[4] = -sh[4][0] / (qc * P.chord);
  }
}

__global__ void k_count_nonfinite(int n, const double* x, int* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !isfinite(x[i])) atomicAdd(out, 1);
}

// ---- host side ----------------------------------------------------------------------
template <class T>
T* dalloc(size_t n) {
  // A2D_RANS_POISON=1 fills every new array with all-ones bytes (NaN doubles,
  // -1 ints): a kernel that reads memory nothing wrote then shows NaN
  static const bool poison = [] {
    const char* e = std::getenv("A2D_RANS_POISON");
    return e && *e && *e != '0';
  }();
  T* p = nullptr;
  if (n == 0) n = 1;
  HC(hipMalloc(&p, n * sizeof(T)));
  if (poison) HC(hipMemset(p, 0xFF, n * sizeof(T)));
  return p;
}
template <class T>
T* dupload(const std::vector<T>& v) {
  T* p = dalloc<T>(v.size());
  if (!v.empty()) HC(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

struct Solver {
  MeshDev m;
  Params P;
  int order = 2;
  hipStream_t st = nullptr;
  std::vector<void*> owned;
  // state and work vectors [5][nc]
  double *Q, *Qprev, *Qbest, *R, *dq, *dq2, *tmp, *tmp2, *bvec;
  double *g, *gt, *phi, *phit, *phi_frozen, *zero5, *F, *Ft, *wgf;
  double *dt, *vdt, *reli, *part, *small, *cpbuf;
  int nparts_cell;
  bool frozen = false;
  std::vector<unsigned char> in_zone;   // host copy of the trailing-edge limiter zone
  // Jacobian and ILU
  double *J = nullptr, *JW = nullptr, *JF = nullptr;
  double *B = nullptr, *invU = nullptr, *y = nullptr, *S = nullptr;
  int *posD = nullptr, *posLR = nullptr, *posRL = nullptr;
  IluDev ilu{};
  bool has_ilu = false;
  // the factors in B/invU/S are a Newton step's finite factorisation of this
  // slot that no other entry point has overwritten since (what reuse_ilu may
  // keep)
  bool ilu_reusable = false;
  int* nonfinite = nullptr;
  // Krylov basis
  double* Vb = nullptr;
  int vcap = 0;
  double* hdev = nullptr;
  int N() const { return 5 * m.nc; }

  template <class T>
  T* own(T* p) {
    owned.push_back((void*)p);
    return p;
  }
  bool owns_stream = false;
  ~Solver() {
    if (Vb) (void)hipFree(Vb);
    for (void* p : owned) (void)hipFree(p);
    if (st && owns_stream) (void)hipStreamDestroy(st);
  }
};

// A solver handle: case slots sharing one mesh, its ILU structure and one
// stream. Slot 0 owns the mesh arrays; the ILU blocks, inverses and work
// vectors of all slots are contiguous, so the triangular kernels can take
// many cases in one launch (one workgroup each); the Arnoldi scalars of all
// slots are contiguous too (one copy to the host per Arnoldi step).
constexpr int HSTRIDE = 2 * (GMRES_MAX + 1) + 8;
struct Group {
  std::vector<std::unique_ptr<Solver>> cases;
  int cur = 0;
  std::recursive_mutex mu;   // one caller at a time per handle
  double *Ball = nullptr, *Iall = nullptr, *Yall = nullptr, *Sall = nullptr, *hall = nullptr, *smallall = nullptr;
  int *nfall = nullptr, *dcases = nullptr;
  // pinned host mirrors of the per-slot scalars and of the case list: copies
  // between them and the device are DMA on the group's stream (no staging
  // through the runtime's pageable-copy buffers, no null-stream copy)
  double *hall_h = nullptr, *small_h = nullptr;
  int *nf_h = nullptr, *cases_h = nullptr;
  size_t sB = 0, sI = 0, sY = 0, sS = 0;
  std::vector<int> dcases_host;
  ~Group() {
    while (cases.size() > 1) cases.pop_back();   // slot 0 (mesh, stream) last
    cases.clear();
    for (void* p : {(void*)Ball, (void*)Iall, (void*)Yall, (void*)Sall, (void*)hall, (void*)smallall,
                    (void*)nfall, (void*)dcases})
      if (p) (void)hipFree(p);
    for (void* p : {(void*)hall_h, (void*)small_h, (void*)nf_h, (void*)cases_h})
      if (p) (void)hipHostFree(p);
  }
};

// Handles. The map has its own lock; each call on a handle holds that
// handle's lock (guarded_h) and a reference to its group, so calls on
// different handles run concurrently from different host threads (each
// handle has its own stream), and a handle destroyed meanwhile lives until
// the calls on it end.
std::mutex g_map_mu;
std::map<int64_t, std::shared_ptr<Group>> g_solvers;
int64_t g_next = 1;

std::shared_ptr<Group> find_group(int64_t h) {
  std::lock_guard<std::mutex> lk(g_map_mu);
  auto it = g_solvers.find(h);
  if (it == g_solvers.end()) fail(-3, "unknown RANS solver handle");
  return it->second;
}

// The group of h, for a call inside guarded_h(h, ...) (which keeps it alive).
Group& get_group(int64_t h) {
  Group& g = *find_group(h);
  require_device();
  return g;
}

template <class F>
int guarded_h(int64_t h, F&& f) {
  try {
    std::shared_ptr<Group> keep = find_group(h);
    std::lock_guard<std::recursive_mutex> lk(keep->mu);
    f();
    return 0;
  } catch (const Fail& e) {
    return e.code;
  } catch (const std::exception& e) {
    set_error_text(e.what());
    return -99;
  }
}

Solver& get(int64_t h) {
  Group& g = get_group(h);
  return *g.cases[g.cur];
}

inline int nblk(int n, int t = TPB) { return (n + t - 1) / t; }

// Residual (or, with V, its tangent along V) of the state Q into Rv/Rt.
// Limiter: none at first order or with the zone off; frozen when set; else
// evaluated (and differentiated) from Q.
// k_face / k_cell in the instantiation the solver's raw-gradient options select.
template <class S, class... A>
void launch_face(Solver& s, int blocks, A... a) {
  if (s.P.visc_raw) hipLaunchKernelGGL((k_face<S, true>), blocks, TPB, 0, s.st, a...);
  else hipLaunchKernelGGL((k_face<S, false>), blocks, TPB, 0, s.st, a...);
}
template <class S, class... A>
void launch_cell(Solver& s, int blocks, A... a) {
  const int k = (s.P.visc_raw ? 1 : 0) + (s.P.sa_raw ? 2 : 0);
  switch (k) {
    case 0: hipLaunchKernelGGL((k_cell<S, false, false>), blocks, TPB, 0, s.st, a...); break;
    case 1: hipLaunchKernelGGL((k_cell<S, true, false>), blocks, TPB, 0, s.st, a...); break;
    case 2: hipLaunchKernelGGL((k_cell<S, false, true>), blocks, TPB, 0, s.st, a...); break;
    default: hipLaunchKernelGGL((k_cell<S, true, true>), blocks, TPB, 0, s.st, a...); break;
  }
}

void eval(Solver& s, int order, const double* V, double* Rv, double* Rt, const double* vdt, double ct_inv) {
  const MeshDev& m = s.m;
  const bool lim = order >= 2 && s.P.limiter_k > 0 && s.P.limiter_zone != 0;
  const double* phi = nullptr;
  const double* phit = nullptr;
  if (V) {
    hipLaunchKernelGGL(k_grad<Dual<double>>, nblk(m.nc), TPB, 0, s.st, m, s.Q, V, s.g, s.gt);
    if (lim && s.frozen) {
      phi = s.phi_frozen;
      phit = s.zero5;
    } else if (lim) {
      hipLaunchKernelGGL(k_limit<Dual<double>>, nblk(m.nc), TPB, 0, s.st, m, s.P, s.Q, V, s.g, s.gt, s.phi,
                         s.phit);
      phi = s.phi;
      phit = s.phit;
    }
    launch_face<Dual<double>>(s, nblk(m.nfi), m, s.P, order, s.Q, V, s.g, s.gt, phi, phit, s.F, s.Ft);
    launch_cell<Dual<double>>(s, nblk(m.nc), m, s.P, order, s.Q, V, s.g, s.gt, phi, phit, s.F, s.Ft, s.wgf, Rv,
                              Rt, vdt, ct_inv);
  } else {
    hipLaunchKernelGGL(k_grad<double>, nblk(m.nc), TPB, 0, s.st, m, s.Q, nullptr, s.g, nullptr);
    if (lim && s.frozen) {
      phi = s.phi_frozen;
    } else if (lim) {
      hipLaunchKernelGGL(k_limit<double>, nblk(m.nc), TPB, 0, s.st, m, s.P, s.Q, nullptr, s.g, nullptr, s.phi,
                         nullptr);
      phi = s.phi;
    }
    launch_face<double>(s, nblk(m.nfi), m, s.P, order, s.Q, (const double*)nullptr, s.g,
                        (const double*)nullptr, phi, (const double*)nullptr, s.F, (double*)nullptr);
    launch_cell<double>(s, nblk(m.nc), m, s.P, order, s.Q, (const double*)nullptr, s.g, (const double*)nullptr,
                        phi, (const double*)nullptr, s.F, (const double*)nullptr, s.wgf, Rv, (double*)nullptr,
                        (const double*)nullptr, 1.0);
  }
  HC(hipGetLastError());
}

void far_state(Solver& s, const double* winf, double gam) {
  hipLaunchKernelGGL(k_far_state, nblk(s.m.nff), TPB, 0, s.st, s.m, s.P, winf[0], winf[1], winf[2], winf[3],
                     winf[4], gam, s.wgf);
}

void rnorm(Solver& s, const double* R, double* out5) {
  hipLaunchKernelGGL(k_rnorm_part, s.nparts_cell, TPB, 0, s.st, s.m.nc, R, s.m.cg, s.part);
  hipLaunchKernelGGL(k_sum_rows, 1, 64, 0, s.st, 5, s.nparts_cell, s.part, s.small);
  double h[5];
  HC(hipMemcpyAsync(h, s.small, 5 * sizeof(double), hipMemcpyDeviceToHost, s.st));
  HC(hipStreamSynchronize(s.st));
  for (int k = 0; k < 5; ++k) out5[k] = std::sqrt(h[k] / s.m.nc);
}

// dot products V_j . w for j < nvec into s.hdev[off .. off+nvec)
void mdot(Solver& s, const double* Vb, size_t ldv, int nvec, const double* w, int off) {
  const int N = s.N();
  const int nb = (N + VCHUNK - 1) / VCHUNK;
  hipLaunchKernelGGL(k_mdot, dim3(nb, nvec), TPB, 0, s.st, N, Vb, ldv, nvec, w, s.part);
  hipLaunchKernelGGL(k_sum_rows, 1, 64 * ((nvec + 63) / 64), 0, s.st, nvec, nb, s.part, s.hdev + off);
}

void ilu_solve(Solver& s, const double* b, double* x) {
  const int n = s.m.nc;
  hipLaunchKernelGGL(k_perm_in, nblk(n), TPB, 0, s.st, n, s.ilu.perm, b, s.y);
  if (s.ilu.lds_solve) {
    hipLaunchKernelGGL(k_ilu_solve_lds, 1, ILU_TPB, ilu_solve_lds_bytes(s.ilu), s.st, s.ilu, (const double*)s.S,
                       s.y, nullptr, size_t(0), size_t(0));
  } else {
    const size_t lds = s.ilu.lds_levels ? size_t(s.ilu.nlev + s.ilu.nlevu + 2) * sizeof(int) : 0;
    hipLaunchKernelGGL(k_ilu_solve, 1, ILU_TPB, lds, s.st, s.ilu, s.B, s.invU, s.y, nullptr, size_t(0),
                       size_t(0), size_t(0));
  }
  hipLaunchKernelGGL(k_perm_out, nblk(n), TPB, 0, s.st, n, s.ilu.perm, s.y, x);
}

void matvec(Solver& s, int order, const double* v, double* out, double ct_inv) {
  eval(s, order, v, nullptr, out, s.vdt, ct_inv);
}

// Preconditioned GMRES without restarts (aero2d/krylov.py:14-87): left
// preconditioning minimises |M^-1 (b - A x)|, right solves A M^-1 u = b and
// returns x = M^-1 u. Arnoldi with CGS2; Givens rotations and the triangular
// solve on the host (one small copy per step). Returns the iterations done;
// info = 0 when the relative estimate reached rtol.
int gmres(Solver& s, int order, double ct_inv, const double* b, double* x, double rtol, int maxiter, bool right,
          int* info_out) {
  const int N = s.N();
  const int m = maxiter;
  if (m + 1 > s.vcap) {
    if (s.Vb) (void)hipFree(s.Vb);
    s.Vb = dalloc<double>(size_t(m + 1) * N);
    s.vcap = m + 1;
  }
  const size_t ld = size_t(N);
  auto V = [&](int j) { return s.Vb + size_t(j) * ld; };
  double* w = s.tmp;
  // r = P b (left) or b (right); x0 = 0
  if (right) HC(hipMemcpyAsync(V(0), b, N * sizeof(double), hipMemcpyDeviceToDevice, s.st));
  else ilu_solve(s, b, V(0));
  mdot(s, V(0), ld, 1, V(0), 0);
  double bb;
  HC(hipMemcpyAsync(&bb, s.hdev, sizeof(double), hipMemcpyDeviceToHost, s.st));
  HC(hipStreamSynchronize(s.st));
  const double beta = std::sqrt(bb);
  HC(hipMemsetAsync(x, 0, N * sizeof(double), s.st));
  if (beta == 0.0 || !std::isfinite(beta)) {
    *info_out = 0;
    return 0;
  }
  hipLaunchKernelGGL(k_scale, nblk(N), TPB, 0, s.st, N, 1.0 / beta, V(0), V(0));
  std::vector<double> H(size_t(m + 1) * m, 0.0), cs(m), sn(m), gv(m + 1, 0.0), hbuf(2 * (m + 1) + 1);
  auto Hij = [&](int i, int j) -> double& { return H[size_t(i) * m + j]; };
  gv[0] = beta;
  int k_done = 0, info = m;
  for (int j = 0; j < m; ++j) {
    if (right) {
      ilu_solve(s, V(j), s.tmp2);
      matvec(s, order, s.tmp2, w, ct_inv);
    } else {
      matvec(s, order, V(j), s.tmp2, ct_inv);
      ilu_solve(s, s.tmp2, w);
    }
    mdot(s, s.Vb, ld, j + 1, w, 0);
    hipLaunchKernelGGL(k_gs_update, nblk(N), TPB, 0, s.st, N, s.Vb, ld, j + 1, s.hdev, w);
    mdot(s, s.Vb, ld, j + 1, w, j + 1);
    hipLaunchKernelGGL(k_gs_update, nblk(N), TPB, 0, s.st, N, s.Vb, ld, j + 1, s.hdev + j + 1, w);
    mdot(s, w, ld, 1, w, 2 * (j + 1));
    HC(hipMemcpyAsync(hbuf.data(), s.hdev, (2 * (j + 1) + 1) * sizeof(double), hipMemcpyDeviceToHost, s.st));
    HC(hipStreamSynchronize(s.st));
    const double hn = std::sqrt(hbuf[2 * (j + 1)]);
    for (int i = 0; i <= j; ++i) Hij(i, j) = hbuf[i] + hbuf[j + 1 + i];
    Hij(j + 1, j) = hn;
    for (int i = 0; i < j; ++i) {
      const double t = cs[i] * Hij(i, j) + sn[i] * Hij(i + 1, j);
      Hij(i + 1, j) = -sn[i] * Hij(i, j) + cs[i] * Hij(i + 1, j);
      Hij(i, j) = t;
    }
    const double d = std::hypot(Hij(j, j), Hij(j + 1, j));
    if (d == 0.0) {
      k_done = j;
      break;
    }
    cs[j] = Hij(j, j) / d;
    sn[j] = Hij(j + 1, j) / d;
    Hij(j, j) = d;
    Hij(j + 1, j) = 0.0;
    gv[j + 1] = -sn[j] * gv[j];
    gv[j] = cs[j] * gv[j];
    k_done = j + 1;
    const double rel = std::fabs(gv[j + 1]) / beta;
    if (!std::isfinite(rel)) break;
    if (rel <= rtol || hn == 0.0) {
      info = 0;
      break;
    }
    hipLaunchKernelGGL(k_scale, nblk(N), TPB, 0, s.st, N, 1.0 / hn, w, V(j + 1));
  }
  *info_out = info;
  const int k = k_done;
  if (k == 0) return 0;
  std::vector<double> yv(k);
  for (int i = k - 1; i >= 0; --i) {
    double t = gv[i];
    for (int c = i + 1; c < k; ++c) t -= Hij(i, c) * yv[c];
    yv[i] = t / Hij(i, i);
  }
  // x = sum_j y_j V_j  (through k_gs_update with -y: x = 0 - sum(-y_j V_j))
  for (auto& v : yv) v = -v;
  HC(hipMemcpyAsync(s.hdev, yv.data(), k * sizeof(double), hipMemcpyHostToDevice, s.st));
  if (right) {
    HC(hipMemsetAsync(s.tmp2, 0, N * sizeof(double), s.st));
    hipLaunchKernelGGL(k_gs_update, nblk(N), TPB, 0, s.st, N, s.Vb, ld, k, s.hdev, s.tmp2);
    ilu_solve(s, s.tmp2, x);
  } else {
    hipLaunchKernelGGL(k_gs_update, nblk(N), TPB, 0, s.st, N, s.Vb, ld, k, s.hdev, x);
  }
  HC(hipStreamSynchronize(s.st));
  return k;
}

double norm2(Solver& s, const double* v) {
  mdot(s, v, size_t(s.N()), 1, v, 0);
  double h;
  HC(hipMemcpyAsync(&h, s.hdev, sizeof(double), hipMemcpyDeviceToHost, s.st));
  HC(hipStreamSynchronize(s.st));
  return std::sqrt(h);
}

// |A dq + R| / |R| with A the Newton matrix at the current step.
double lin_residual(Solver& s, int order, double ct_inv, const double* dq, double rnorm) {
  matvec(s, order, dq, s.tmp2, ct_inv);
  hipLaunchKernelGGL(k_add, nblk(s.N()), TPB, 0, s.st, s.N(), s.tmp2, s.R, s.tmp2);
  return norm2(s, s.tmp2) / std::max(rnorm, 1e-300);
}

void factor(Solver& s, const double* winf, double gam, double ct_inv, double* Dout, double* OffOut,
            bool launch_factor = true) {
  if (!Dout && !s.has_ilu) fail(-6, "a2d_rans_ilu_setup has not been called");
  const MeshDev& m = s.m;
  double* B = Dout ? nullptr : s.B;
  far_state(s, winf, gam);
  hipLaunchKernelGGL(k_grad<double>, nblk(m.nc), TPB, 0, s.st, m, s.Q, nullptr, s.g, nullptr);
  if (B) HC(hipMemsetAsync(B, 0, size_t(s.ilu.nnz) * 25 * sizeof(double), s.st));
  hipLaunchKernelGGL(k_jac_int, nblk(m.nfi), TPB, 0, s.st, m, s.P, s.Q, s.J, B, s.posLR, s.posRL);
  hipLaunchKernelGGL(k_jac_wall, nblk(m.nfw), TPB, 0, s.st, m, s.P, s.Q, s.JW);
  hipLaunchKernelGGL(k_jac_far, nblk(m.nff), TPB, 0, s.st, m, s.P, s.Q, s.wgf, s.JF);
  hipLaunchKernelGGL(k_jac_diag, nblk(m.nc), TPB, 0, s.st, m, s.P, s.Q, s.g, s.J, s.JW, s.JF,
                     Dout ? nullptr : s.vdt, ct_inv, B, s.posD, s.tmp2);
  HC(hipGetLastError());
  if (Dout) {
    std::vector<double> Jh(size_t(m.nfi) * 50);
    HC(hipMemcpyAsync(Dout, s.tmp2, size_t(m.nc) * 25 * sizeof(double), hipMemcpyDeviceToHost, s.st));
    HC(hipMemcpyAsync(Jh.data(), s.J, Jh.size() * sizeof(double), hipMemcpyDeviceToHost, s.st));
    HC(hipStreamSynchronize(s.st));
    for (int f = 0; f < m.nfi; ++f)
      for (int t = 0; t < 25; ++t) {
        OffOut[size_t(f) * 25 + t] = Jh[size_t(f) * 50 + 25 + t];
        OffOut[size_t(m.nfi + f) * 25 + t] = -Jh[size_t(f) * 50 + t];
      }
    return;
  }
  if (!launch_factor) return;
  launch_ilu_factor(s.ilu, s.B, s.invU, s.S, nullptr, 1, 0, 0, 0, s.st);
  HC(hipGetLastError());
}

int count_nonfinite(Solver& s, const double* x, int n) {
  HC(hipMemsetAsync(s.nonfinite, 0, sizeof(int), s.st));
  hipLaunchKernelGGL(k_count_nonfinite, nblk(n), TPB, 0, s.st, n, x, s.nonfinite);
  int h = 0;
  HC(hipMemcpyAsync(&h, s.nonfinite, sizeof(int), hipMemcpyDeviceToHost, s.st));
  HC(hipStreamSynchronize(s.st));
  return h;
}

void to_soa(int nc, const double* aos, std::vector<double>& soa) {
  soa.resize(size_t(nc) * 5);
  for (int c = 0; c < nc; ++c)
    for (int k = 0; k < 5; ++k) soa[size_t(k) * nc + c] = aos[size_t(c) * 5 + k];
}
void to_aos(int nc, const std::vector<double>& soa, double* aos) {
  for (int c = 0; c < nc; ++c)
    for (int k = 0; k < 5; ++k) aos[size_t(c) * 5 + k] = soa[size_t(k) * nc + c];
}

void upload_cells(Solver& s, const double* aos, double* dst) {
  std::vector<double> soa;
  to_soa(s.m.nc, aos, soa);
  HC(hipMemcpyAsync(dst, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice, s.st));
  HC(hipStreamSynchronize(s.st));
}
void download_cells(Solver& s, const double* src, double* aos) {
  std::vector<double> soa(size_t(s.m.nc) * 5);
  HC(hipMemcpyAsync(soa.data(), src, soa.size() * sizeof(double), hipMemcpyDeviceToHost, s.st));
  HC(hipStreamSynchronize(s.st));
  to_aos(s.m.nc, soa, aos);
}

// Device arrays of one case slot (everything but the group-level ones).
void alloc_case(Solver& S) {
  const int nc = S.m.nc, nfi = S.m.nfi, nfw = S.m.nfw, nff = S.m.nff;
  const size_t n5 = size_t(nc) * 5;
  S.Q = S.own(dalloc<double>(n5));
  S.Qprev = S.own(dalloc<double>(n5));
  S.Qbest = S.own(dalloc<double>(n5));
  S.R = S.own(dalloc<double>(n5));
  S.dq = S.own(dalloc<double>(n5));
  S.dq2 = S.own(dalloc<double>(n5));
  S.tmp = S.own(dalloc<double>(n5));
  S.tmp2 = S.own(dalloc<double>(std::max(n5, size_t(nc) * 25)));
  S.bvec = S.own(dalloc<double>(n5));
  S.g = S.own(dalloc<double>(2 * n5));
  S.gt = S.own(dalloc<double>(2 * n5));
  S.phi = S.own(dalloc<double>(n5));
  S.phit = S.own(dalloc<double>(n5));
  S.phi_frozen = S.own(dalloc<double>(n5));
  S.zero5 = S.own(dalloc<double>(n5));
  HC(hipMemset(S.zero5, 0, n5 * sizeof(double)));
  S.F = S.own(dalloc<double>(size_t(nfi) * 5));
  S.Ft = S.own(dalloc<double>(size_t(nfi) * 5));
  S.wgf = S.own(dalloc<double>(size_t(nff) * 5));
  S.dt = S.own(dalloc<double>(nc));
  S.vdt = S.own(dalloc<double>(nc));
  S.reli = S.own(dalloc<double>(nc));
  S.cpbuf = S.own(dalloc<double>(2 * size_t(nfw)));
  S.nparts_cell = nblk(nc);
  const int nbv = (5 * nc + VCHUNK - 1) / VCHUNK;
  S.part = S.own(dalloc<double>(std::max(size_t(5) * S.nparts_cell, size_t(GMRES_MAX + 1) * nbv) +
                                2 * S.nparts_cell));
  S.J = S.own(dalloc<double>(size_t(nfi) * 50));
  S.JW = S.own(dalloc<double>(size_t(nfw) * 25));
  S.JF = S.own(dalloc<double>(size_t(nff) * 25));
}

// (Re)allocates the group-level arrays for the current slots and points every
// slot at its part: Arnoldi scalars, small results, non-finite counters and,
// once the ILU structure is set, the blocks, inverses and work vectors.
void relink(Group& g) {
  const int nc = int(g.cases.size());
  Solver& s0 = *g.cases[0];
  for (void** p : {(void**)&g.hall, (void**)&g.smallall, (void**)&g.nfall, (void**)&g.dcases})
    if (*p) {
      HC(hipFree(*p));
      *p = nullptr;
    }
  for (void** p : {(void**)&g.hall_h, (void**)&g.small_h, (void**)&g.nf_h, (void**)&g.cases_h})
    if (*p) {
      HC(hipHostFree(*p));
      *p = nullptr;
    }
  g.hall = dalloc<double>(size_t(nc) * HSTRIDE);
  g.smallall = dalloc<double>(size_t(nc) * 64);
  g.nfall = dalloc<int>(nc);
  g.dcases = dalloc<int>(nc);
  HC(hipHostMalloc((void**)&g.hall_h, size_t(nc) * HSTRIDE * sizeof(double), hipHostMallocDefault));
  HC(hipHostMalloc((void**)&g.small_h, size_t(nc) * 64 * sizeof(double), hipHostMallocDefault));
  HC(hipHostMalloc((void**)&g.nf_h, size_t(nc) * sizeof(int), hipHostMallocDefault));
  HC(hipHostMalloc((void**)&g.cases_h, size_t(nc) * sizeof(int), hipHostMallocDefault));
  g.dcases_host.clear();
  for (int c = 0; c < nc; ++c) {
    Solver& s = *g.cases[c];
    s.hdev = g.hall + size_t(c) * HSTRIDE;
    s.small = g.smallall + size_t(c) * 64;
    s.nonfinite = g.nfall + c;
  }
  if (!s0.has_ilu) return;
  for (void** p : {(void**)&g.Ball, (void**)&g.Iall, (void**)&g.Yall, (void**)&g.Sall})
    if (*p) {
      HC(hipFree(*p));
      *p = nullptr;
    }
  g.sB = size_t(s0.ilu.nnz) * 25;
  g.sI = size_t(s0.ilu.n) * 25;
  g.sY = size_t(s0.ilu.n) * 5;
  g.sS = s0.ilu.lds_solve ? size_t(s0.ilu.nst) * 25 : 0;
  g.Ball = dalloc<double>(g.sB * nc);
  g.Iall = dalloc<double>(g.sI * nc);
  g.Yall = dalloc<double>(g.sY * nc);
  if (g.sS) g.Sall = dalloc<double>(g.sS * nc);
  for (int c = 0; c < nc; ++c) {
    Solver& s = *g.cases[c];
    s.ilu = s0.ilu;
    s.posD = s0.posD;
    s.posLR = s0.posLR;
    s.posRL = s0.posRL;
    s.has_ilu = true;
    s.ilu_reusable = false;      // fresh arrays hold no factors
    s.B = g.Ball + size_t(c) * g.sB;
    s.invU = g.Iall + size_t(c) * g.sI;
    s.y = g.Yall + size_t(c) * g.sY;
    s.S = g.Sall ? g.Sall + size_t(c) * g.sS : nullptr;
  }
}

// The device list of case slots for a batched launch.
const int* case_list(Group& g, const std::vector<int>& ids) {
  if (ids != g.dcases_host) {
    // the stream's earlier copy out of cases_h has run once the stream is idle
    HC(hipStreamSynchronize(g.cases[0]->st));
    std::copy(ids.begin(), ids.end(), g.cases_h);
    HC(hipMemcpyAsync(g.dcases, g.cases_h, ids.size() * sizeof(int), hipMemcpyHostToDevice, g.cases[0]->st));
    g.dcases_host = ids;
  }
  return g.dcases;
}

// ---- batched linear algebra over case slots -----------------------------------------
// The same operations as ilu_solve / gmres / the Newton step for several
// slots at once: per-slot kernels are launched in turn on the group's stream,
// the triangular sweeps and the factorisation of all slots in one launch (a
// workgroup per slot), and the Arnoldi scalars of all slots come back in one
// copy per step. Each slot's arithmetic is exactly that of a solve alone.

void ilu_solve_batch(Group& g, const std::vector<Solver*>& cs, const std::vector<int>& ids,
                     const std::vector<const double*>& b, const std::vector<double*>& x) {
  Solver& s0 = *g.cases[0];
  const int n = s0.m.nc;
  for (size_t i = 0; i < cs.size(); ++i)
    hipLaunchKernelGGL(k_perm_in, nblk(n), TPB, 0, s0.st, n, s0.ilu.perm, b[i], cs[i]->y);
  if (s0.ilu.lds_solve) {
    hipLaunchKernelGGL(k_ilu_solve_lds, int(ids.size()), ILU_TPB, ilu_solve_lds_bytes(s0.ilu), s0.st, s0.ilu,
                       (const double*)g.Sall, g.Yall, case_list(g, ids), g.sS, g.sY);
  } else {
    const size_t lds = s0.ilu.lds_levels ? size_t(s0.ilu.nlev + s0.ilu.nlevu + 2) * sizeof(int) : 0;
    hipLaunchKernelGGL(k_ilu_solve, int(ids.size()), ILU_TPB, lds, s0.st, s0.ilu, (const double*)g.Ball,
                       (const double*)g.Iall, g.Yall, case_list(g, ids), g.sB, g.sI, g.sY);
  }
  for (size_t i = 0; i < cs.size(); ++i)
    hipLaunchKernelGGL(k_perm_out, nblk(n), TPB, 0, s0.st, n, s0.ilu.perm, cs[i]->y, x[i]);
}

struct GmresState {
  std::vector<double> H, cs, sn, gv;
  double beta = 0.0;
  int k_done = 0, info = 0;
  bool active = false;
};

// gmres() for the slots in ids, all with the same preconditioning side; b[i],
// x[i] per slot; returns the iterations done per slot in its[i], info in info[i].
void gmres_batch(Group& g, const std::vector<int>& ids, const std::vector<int>& order,
                 const std::vector<double>& ct_inv, const std::vector<const double*>& b,
                 const std::vector<double*>& x, const std::vector<double>& rtol, int m, bool right,
                 std::vector<int>& its, std::vector<int>& info) {
  const size_t nb = ids.size();
  Solver& s0 = *g.cases[0];
  const int N = s0.N();
  const size_t ld = size_t(N);
  std::vector<Solver*> cs(nb);
  for (size_t i = 0; i < nb; ++i) {
    cs[i] = g.cases[ids[i]].get();
    Solver& s = *cs[i];
    if (m + 1 > s.vcap) {
      if (s.Vb) (void)hipFree(s.Vb);
      s.Vb = nullptr;
      s.Vb = dalloc<double>(size_t(m + 1) * N);
      s.vcap = m + 1;
    }
  }
  auto V = [&](size_t i, int j) { return cs[i]->Vb + size_t(j) * ld; };
  std::vector<const double*> pb(nb);
  std::vector<double*> px(nb);
  for (size_t i = 0; i < nb; ++i) px[i] = V(i, 0);
  if (right) {
    for (size_t i = 0; i < nb; ++i)
      HC(hipMemcpyAsync(V(i, 0), b[i], N * sizeof(double), hipMemcpyDeviceToDevice, s0.st));
  } else {
    ilu_solve_batch(g, cs, ids, b, px);
  }
  for (size_t i = 0; i < nb; ++i) mdot(*cs[i], V(i, 0), ld, 1, V(i, 0), 0);
  double* const hall = g.hall_h;
  auto pull = [&] {
    HC(hipMemcpyAsync(hall, g.hall, g.cases.size() * HSTRIDE * sizeof(double), hipMemcpyDeviceToHost, s0.st));
    HC(hipStreamSynchronize(s0.st));
  };
  pull();
  std::vector<GmresState> G(nb);
  its.assign(nb, 0);
  info.assign(nb, 0);
  for (size_t i = 0; i < nb; ++i) {
    const double* h = hall + size_t(ids[i]) * HSTRIDE;
    GmresState& q = G[i];
    q.beta = std::sqrt(h[0]);
    HC(hipMemsetAsync(x[i], 0, N * sizeof(double), s0.st));
    if (q.beta == 0.0 || !std::isfinite(q.beta)) continue;
    q.active = true;
    q.info = m;
    q.H.assign(size_t(m + 1) * m, 0.0);
    q.cs.assign(m, 0.0);
    q.sn.assign(m, 0.0);
    q.gv.assign(m + 1, 0.0);
    q.gv[0] = q.beta;
    hipLaunchKernelGGL(k_scale, nblk(N), TPB, 0, s0.st, N, 1.0 / q.beta, V(i, 0), V(i, 0));
  }
  for (int j = 0; j < m; ++j) {
    std::vector<size_t> act;
    for (size_t i = 0; i < nb; ++i)
      if (G[i].active) act.push_back(i);
    if (act.empty()) break;
    std::vector<Solver*> ac;
    std::vector<int> aid;
    std::vector<const double*> in;
    std::vector<double*> out;
    for (size_t i : act) {
      ac.push_back(cs[i]);
      aid.push_back(ids[i]);
    }
    if (right) {
      for (size_t i : act) {
        in.push_back(V(i, j));
        out.push_back(cs[i]->tmp2);
      }
      ilu_solve_batch(g, ac, aid, in, out);
      for (size_t i : act) matvec(*cs[i], order[i], cs[i]->tmp2, cs[i]->tmp, ct_inv[i]);
    } else {
      for (size_t i : act) {
        matvec(*cs[i], order[i], V(i, j), cs[i]->tmp2, ct_inv[i]);
        in.push_back(cs[i]->tmp2);
        out.push_back(cs[i]->tmp);
      }
      ilu_solve_batch(g, ac, aid, in, out);
    }
    for (size_t i : act) {
      Solver& s = *cs[i];
      double* w = s.tmp;
      mdot(s, s.Vb, ld, j + 1, w, 0);
      hipLaunchKernelGGL(k_gs_update, nblk(N), TPB, 0, s0.st, N, s.Vb, ld, j + 1, s.hdev, w);
      mdot(s, s.Vb, ld, j + 1, w, j + 1);
      hipLaunchKernelGGL(k_gs_update, nblk(N), TPB, 0, s0.st, N, s.Vb, ld, j + 1, s.hdev + j + 1, w);
      mdot(s, w, ld, 1, w, 2 * (j + 1));
    }
    pull();
    for (size_t i : act) {
      GmresState& q = G[i];
      const double* hb = hall + size_t(ids[i]) * HSTRIDE;
      auto Hij = [&](int a, int c) -> double& { return q.H[size_t(a) * m + c]; };
      const double hn = std::sqrt(hb[2 * (j + 1)]);
      for (int a = 0; a <= j; ++a) Hij(a, j) = hb[a] + hb[j + 1 + a];
      Hij(j + 1, j) = hn;
      for (int a = 0; a < j; ++a) {
        const double t = q.cs[a] * Hij(a, j) + q.sn[a] * Hij(a + 1, j);
        Hij(a + 1, j) = -q.sn[a] * Hij(a, j) + q.cs[a] * Hij(a + 1, j);
        Hij(a, j) = t;
      }
      const double d = std::hypot(Hij(j, j), Hij(j + 1, j));
      if (d == 0.0) {
        q.k_done = j;
        q.active = false;
        continue;
      }
      q.cs[j] = Hij(j, j) / d;
      q.sn[j] = Hij(j + 1, j) / d;
      Hij(j, j) = d;
      Hij(j + 1, j) = 0.0;
      q.gv[j + 1] = -q.sn[j] * q.gv[j];
      q.gv[j] = q.cs[j] * q.gv[j];
      q.k_done = j + 1;
      const double rel = std::fabs(q.gv[j + 1]) / q.beta;
      if (!std::isfinite(rel)) {
        q.active = false;
        continue;
      }
      if (rel <= rtol[i] || hn == 0.0) {
        q.info = 0;
        q.active = false;
        continue;
      }
      hipLaunchKernelGGL(k_scale, nblk(N), TPB, 0, s0.st, N, 1.0 / hn, cs[i]->tmp, V(i, j + 1));
    }
  }
  // x = V y (right: x = M^-1 V y)
  std::vector<Solver*> rc;
  std::vector<int> rid;
  std::vector<const double*> rin;
  std::vector<double*> rout;
  std::vector<std::vector<double>> ys(nb);
  for (size_t i = 0; i < nb; ++i) {
    GmresState& q = G[i];
    info[i] = q.beta == 0.0 || !std::isfinite(q.beta) ? 0 : q.info;
    its[i] = q.k_done;
    const int k = q.k_done;
    if (k == 0) continue;
    auto Hij = [&](int a, int c) { return q.H[size_t(a) * m + c]; };
    std::vector<double>& yv = ys[i];
    yv.assign(k, 0.0);
    for (int a = k - 1; a >= 0; --a) {
      double t = q.gv[a];
      for (int c = a + 1; c < k; ++c) t -= Hij(a, c) * yv[c];
      yv[a] = t / Hij(a, a);
    }
    for (auto& v : yv) v = -v;
    Solver& s = *cs[i];
    HC(hipMemcpyAsync(s.hdev, yv.data(), k * sizeof(double), hipMemcpyHostToDevice, s0.st));
    if (right) {
      HC(hipMemsetAsync(s.tmp2, 0, N * sizeof(double), s0.st));
      hipLaunchKernelGGL(k_gs_update, nblk(N), TPB, 0, s0.st, N, s.Vb, ld, k, s.hdev, s.tmp2);
      rc.push_back(cs[i]);
      rid.push_back(ids[i]);
      rin.push_back(s.tmp2);
      rout.push_back(x[i]);
    } else {
      hipLaunchKernelGGL(k_gs_update, nblk(N), TPB, 0, s0.st, N, s.Vb, ld, k, s.hdev, x[i]);
    }
    // the host vector must outlive the asynchronous copy
    HC(hipStreamSynchronize(s0.st));
  }
  if (!rc.empty()) ilu_solve_batch(g, rc, rid, rin, rout);
  HC(hipStreamSynchronize(s0.st));
}

}  // namespace
}  // namespace rans
}  // namespace a2d

using namespace a2d::rans;

extern "C" {

int a2d_rans_create(int nc, int nfi, int nfw, int nff, const int* face_l, const int* face_r,
                    const double* face_geom, const int* wall_cell, const double* wall_geom, const int* far_cell,
                    const double* far_geom, const double* cell_geom, const a2d_rans_params* prm,
                    int64_t* handle_out) {
  return guarded_new([&] {
    a2d::require_device();
    if (nc < 1 || nfi < 1 || nfw < 1 || nff < 1) fail(-6, "a RANS mesh needs cells, interior, wall and far faces");
    if (!prm || !handle_out) fail(-6, "null argument");
    if (prm->order != 1 && prm->order != 2) fail(-6, "order must be 1 or 2");
    if (prm->limiter_zone < 0 || prm->limiter_zone > 2) fail(-6, "limiter_zone must be 0, 1 or 2");
    if (!(prm->mach > 0.0 && prm->mach < 1.0)) fail(-6, "mach must be in (0, 1)");
    if (int64_t(nc) * 5 * 61 > (int64_t(1) << 31)) fail(-6, "mesh too large for 32-bit vector indexing");
    for (int f = 0; f < nfi; ++f)
      if (face_l[f] < 0 || face_l[f] >= nc || face_r[f] < 0 || face_r[f] >= nc || face_l[f] == face_r[f])
        fail(-6, "interior face with a bad cell index");
    for (int f = 0; f < nfw; ++f)
      if (wall_cell[f] < 0 || wall_cell[f] >= nc) fail(-6, "wall face with a bad cell index");
    for (int f = 0; f < nff; ++f)
      if (far_cell[f] < 0 || far_cell[f] >= nc) fail(-6, "far-field face with a bad cell index");
    auto s = std::make_unique<Solver>();
    HC(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
    MeshDev& m = s->m;
    m.nc = nc;
    m.nfi = nfi;
    m.nfw = nfw;
    m.nff = nff;
    auto soa = [](const double* aos, int n, int w) {
      std::vector<double> v(size_t(n) * w);
      for (int i = 0; i < n; ++i)
        for (int k = 0; k < w; ++k) v[size_t(k) * n + i] = aos[size_t(i) * w + k];
      return v;
    };
    m.iL = s->own(dupload(std::vector<int>(face_l, face_l + nfi)));
    m.iR = s->own(dupload(std::vector<int>(face_r, face_r + nfi)));
    m.fg = s->own(dupload(soa(face_geom, nfi, NFG)));
    m.cw = s->own(dupload(std::vector<int>(wall_cell, wall_cell + nfw)));
    m.wg = s->own(dupload(soa(wall_geom, nfw, NWG)));
    m.cf = s->own(dupload(std::vector<int>(far_cell, far_cell + nff)));
    m.ffg = s->own(dupload(soa(far_geom, nff, NFF)));
    m.cg = s->own(dupload(soa(cell_geom, nc, NCG)));
    // cell -> faces lists, in face order (interior, wall, far)
    std::vector<int> cnt(nc + 1, 0), wc(nc + 1, 0), fc(nc + 1, 0);
    for (int f = 0; f < nfi; ++f) {
      cnt[face_l[f] + 1]++;
      cnt[face_r[f] + 1]++;
    }
    for (int f = 0; f < nfw; ++f) wc[wall_cell[f] + 1]++;
    for (int f = 0; f < nff; ++f) fc[far_cell[f] + 1]++;
    for (int c = 0; c < nc; ++c) {
      cnt[c + 1] += cnt[c];
      wc[c + 1] += wc[c];
      fc[c + 1] += fc[c];
    }
    std::vector<int> cf(cnt[nc]), wf(wc[nc]), ff(fc[nc]);
    {
      std::vector<int> pos(cnt.begin(), cnt.end() - 1);
      for (int f = 0; f < nfi; ++f) {
        cf[pos[face_l[f]]++] = f;
        cf[pos[face_r[f]]++] = ~f;
      }
      std::vector<int> pw(wc.begin(), wc.end() - 1), pf(fc.begin(), fc.end() - 1);
      for (int f = 0; f < nfw; ++f) wf[pw[wall_cell[f]]++] = f;
      for (int f = 0; f < nff; ++f) ff[pf[far_cell[f]]++] = f;
    }
    s->in_zone.assign(nc, 0);
    for (int c = 0; c < nc; ++c) s->in_zone[c] = cell_geom[size_t(c) * NCG + 5] != 0.0;
    std::vector<unsigned char> farc(nc, 0);
    for (int f = 0; f < nff; ++f) farc[far_cell[f]] = 1;
    m.cptr = s->own(dupload(cnt));
    m.cfac = s->own(dupload(cf));
    m.wptr = s->own(dupload(wc));
    m.wfac = s->own(dupload(wf));
    m.fptr = s->own(dupload(fc));
    m.ffac = s->own(dupload(ff));
    m.far_cell = s->own(dupload(farc));
    Params& P = s->P;
    P.mach = prm->mach;
    P.mu_inf = prm->mach * prm->chord / prm->re;
    P.nu_inf = P.mu_inf;
    P.chi_inf = prm->turbulent ? prm->chi_inf : 0.0;
    P.limiter_k = prm->limiter_k;
    P.xref_x = prm->xref_x;
    P.xref_y = prm->xref_y;
    P.zmin = prm->zmin;
    P.chord = prm->chord;
    P.turbulent = prm->turbulent;
    P.limiter_zone = prm->limiter_zone;
    P.sa_raw = prm->sa_raw_grad != 0;
    P.visc_raw = prm->visc_raw_grad != 0;
    s->order = prm->order;
    s->owns_stream = true;
    alloc_case(*s);
    auto grp = std::make_shared<Group>();
    grp->cases.push_back(std::move(s));
    relink(*grp);
    std::lock_guard<std::mutex> lk(g_map_mu);
    const int64_t h = g_next++;
    g_solvers[h] = std::move(grp);
    *handle_out = h;
  });
}

int a2d_rans_destroy(int64_t h) {
  try {
    std::shared_ptr<Group> keep = find_group(h);
    std::lock_guard<std::recursive_mutex> lk(keep->mu);   // waits for calls in flight on h
    std::lock_guard<std::mutex> lm(g_map_mu);
    g_solvers.erase(h);
    return 0;   // the group is freed with the last reference, after the locks
  } catch (const Fail& e) {
    return e.code;
  }
}

int a2d_rans_set_cases(int64_t h, int ncases) {
  return guarded_h(h, [&] {
    Group& g = get_group(h);
    if (ncases < 1 || ncases > 1024) fail(-6, "ncases must be in [1, 1024]");
    if (ncases <= int(g.cases.size())) return;
    Solver& s0 = *g.cases[0];
    while (int(g.cases.size()) < ncases) {
      auto s = std::make_unique<Solver>();
      s->m = s0.m;
      s->P = s0.P;
      s->order = s0.order;
      s->st = s0.st;
      s->in_zone = s0.in_zone;
      alloc_case(*s);
      g.cases.push_back(std::move(s));
    }
    relink(g);
  });
}

int a2d_rans_select(int64_t h, int c) {
  return guarded_h(h, [&] {
    Group& g = get_group(h);
    if (c < 0 || c >= int(g.cases.size())) fail(-4, "no such case slot");
    g.cur = c;
  });
}

int a2d_rans_set_state(int64_t h, const double* Q) {
  return guarded_h(h, [&] {
    Solver& s = get(h);
    upload_cells(s, Q, s.Q);
  });
}

int a2d_rans_get_state(int64_t h, double* Q) {
  return guarded_h(h, [&] {
    Solver& s = get(h);
    download_cells(s, s.Q, Q);
  });
}

int a2d_rans_freeze_limiter(int64_t h, const double* phi) {
  return guarded_h(h, [&] {
    Solver& s = get(h);
    if (!phi) {
      s.frozen = false;
      return;
    }
    // outside the limiter zone the frozen limiter is 1, as for a live one
    // (aero2d/rans.py:868-871)
    std::vector<double> p(phi, phi + size_t(s.m.nc) * 5);
    if (s.P.limiter_zone == 1)
      for (int c = 0; c < s.m.nc; ++c)
        if (!s.in_zone[c])
          for (int k = 0; k < 5; ++k) p[size_t(c) * 5 + k] = 1.0;
    upload_cells(s, p.data(), s.phi_frozen);
    s.frozen = true;
  });
}

int a2d_rans_limiter(int64_t h, double* phi) {
  return guarded_h(h, [&] {
    Solver& s = get(h);
    hipLaunchKernelGGL(k_grad<double>, nblk(s.m.nc), TPB, 0, s.st, s.m, s.Q, nullptr, s.g, nullptr);
    Params P = s.P;
    P.limiter_zone = 2;
    hipLaunchKernelGGL(k_limit<double>, nblk(s.m.nc), TPB, 0, s.st, s.m, P, s.Q, nullptr, s.g, nullptr, s.phi,
                       nullptr);
    HC(hipGetLastError());
    download_cells(s, s.phi, phi);
  });
}

int a2d_rans_residual(int64_t h, int order, const double* winf, double gam, double* R, double* rn5) {
  return guarded_h(h, [&] {
    Solver& s = get(h);
    if (order != 1 && order != 2) fail(-6, "order must be 1 or 2");
    far_state(s, winf, gam);
    eval(s, order, nullptr, s.R, nullptr, nullptr, 1.0);
    if (rn5) rnorm(s, s.R, rn5);
    if (R) download_cells(s, s.R, R);
    HC(hipStreamSynchronize(s.st));
  });
}

int a2d_rans_jvp(int64_t h, int order, const double* winf, double gam, const double* v, double* out) {
  return guarded_h(h, [&] {
    Solver& s = get(h);
    if (order != 1 && order != 2) fail(-6, "order must be 1 or 2");
    upload_cells(s, v, s.bvec);
    far_state(s, winf, gam);
    eval(s, order, s.bvec, s.tmp, s.tmp, nullptr, 1.0);
    download_cells(s, s.tmp, out);
  });
}

int a2d_rans_local_dt(int64_t h, double cfl, double* dt) {
  return guarded_h(h, [&] {
    Solver& s = get(h);
    hipLaunchKernelGGL(k_local_dt, nblk(s.m.nc), TPB, 0, s.st, s.m, s.P, s.Q, cfl, s.dt, s.vdt);
    HC(hipGetLastError());
    HC(hipMemcpyAsync(dt, s.dt, s.m.nc * sizeof(double), hipMemcpyDeviceToHost, s.st));
    HC(hipStreamSynchronize(s.st));
  });
}

int a2d_rans_forces(int64_t h, double alpha_deg, double* out5, double* cp, double* cf) {
  return guarded_h(h, [&] {
    Solver& s = get(h);
    hipLaunchKernelGGL(k_grad<double>, nblk(s.m.nc), TPB, 0, s.st, s.m, s.Q, nullptr, s.g, nullptr);
    const double a = alpha_deg * M_PI / 180.0;
    double* dcp = cp ? s.cpbuf : nullptr;
    double* dcf = cf ? s.cpbuf + s.m.nfw : nullptr;
    hipLaunchKernelGGL(k_forces, 1, TPB, 0, s.st, s.m, s.P, s.Q, s.g, std::cos(a), std::sin(a), s.small, dcp,
                       dcf);
    HC(hipGetLastError());
    HC(hipMemcpyAsync(out5, s.small, 5 * sizeof(double), hipMemcpyDeviceToHost, s.st));
    if (cp) HC(hipMemcpyAsync(cp, dcp, s.m.nfw * sizeof(double), hipMemcpyDeviceToHost, s.st));
    if (cf) HC(hipMemcpyAsync(cf, dcf, s.m.nfw * sizeof(double), hipMemcpyDeviceToHost, s.st));
    HC(hipStreamSynchronize(s.st));
  });
}

int a2d_rans_jac1(int64_t h, const double* winf, double gam, double* D, double* off) {
  return guarded_h(h, [&] {
    Solver& s = get(h);
    factor(s, winf, gam, 1.0, D, off);
  });
}

int a2d_rans_ilu_setup(int64_t h, int nnz, const int* perm, const int* rows, const int* cols, const int* pos_src) {
  return guarded_h(h, [&] {
    Solver& s = *get_group(h).cases[0];
    const int n = s.m.nc;
    const int nfi = s.m.nfi;
    if (nnz < n) fail(-6, "pattern smaller than its diagonal");
    std::vector<int> rstart(n, 0), rend(n, 0), rdiag(n, -1);
    for (int e = 0; e < nnz; ++e) {
      if (rows[e] < 0 || rows[e] >= n || cols[e] < 0 || cols[e] >= n) fail(-6, "pattern index out of range");
      if (e > 0 && (rows[e] < rows[e - 1] || (rows[e] == rows[e - 1] && cols[e] <= cols[e - 1])))
        fail(-6, "pattern must be sorted by row, then column, without repeats");
    }
    {
      int e = 0;
      for (int i = 0; i < n; ++i) {
        rstart[i] = e;
        while (e < nnz && rows[e] == i) {
          if (cols[e] == i) rdiag[i] = e;
          ++e;
        }
        rend[i] = e;
        if (rdiag[i] < 0) fail(-6, "pattern misses a diagonal block");
      }
    }
    std::vector<int> permv(perm, perm + n), seen(n, 0);
    for (int i = 0; i < n; ++i) {
      if (permv[i] < 0 || permv[i] >= n || seen[permv[i]]) fail(-6, "perm is not a permutation");
      seen[permv[i]] = 1;
    }
    for (int q = 0; q < n + 2 * nfi; ++q)
      if (pos_src[q] < 0 || pos_src[q] >= nnz) fail(-6, "block position out of range");
    // update pairs of every lower entry e = (i,k): (i,j) -= (i,k)(k,j) for the
    // upper entries (k,j) whose (i,j) is in the pattern
    std::vector<int> pptr(nnz + 1, 0), pt, pe2;
    for (int i = 0; i < n; ++i) {
      for (int e = rstart[i]; e < rdiag[i]; ++e) {
        const int k = cols[e];
        for (int e2 = rdiag[k] + 1; e2 < rend[k]; ++e2) {
          const int j = cols[e2];
          const int* b = cols + rstart[i];
          const int* en = cols + rend[i];
          const int* it = std::lower_bound(b, en, j);
          if (it != en && *it == j) {
            pt.push_back(int(it - cols));
            pe2.push_back(e2);
          }
        }
        pptr[e + 1] = int(pt.size());
      }
      for (int e = rdiag[i]; e < rend[i]; ++e) pptr[e + 1] = int(pt.size());
    }
    // levels: forward by the lower pattern, backward by the upper pattern
    std::vector<int> lev(n, 0), levu(n, 0);
    int nlev = 0, nlevu = 0;
    for (int i = 0; i < n; ++i) {
      int l = 0;
      for (int e = rstart[i]; e < rdiag[i]; ++e) l = std::max(l, lev[cols[e]] + 1);
      lev[i] = l;
      nlev = std::max(nlev, l + 1);
    }
    for (int i = n - 1; i >= 0; --i) {
      int l = 0;
      for (int e = rdiag[i] + 1; e < rend[i]; ++e) l = std::max(l, levu[cols[e]] + 1);
      levu[i] = l;
      nlevu = std::max(nlevu, l + 1);
    }
    auto group = [&](const std::vector<int>& lv, int nl, std::vector<int>& ptr, std::vector<int>& rws) {
      ptr.assign(nl + 1, 0);
      for (int i = 0; i < n; ++i) ptr[lv[i] + 1]++;
      for (int l = 0; l < nl; ++l) ptr[l + 1] += ptr[l];
      rws.assign(n, 0);
      std::vector<int> pos(ptr.begin(), ptr.end() - 1);
      for (int i = 0; i < n; ++i) rws[pos[lv[i]]++] = i;
    };
    std::vector<int> lptr, lrows, uptr, urows;
    group(lev, nlev, lptr, lrows);
    group(levu, nlevu, uptr, urows);
    // within a level, the rows with the most elimination work first: the
    // factorisation deals rows to wavefronts in this order, so the longest
    // rows start in the first round (rows of a level are independent: the
    // results do not depend on the order)
    for (int l = 0; l < nlev; ++l)
      std::stable_sort(lrows.begin() + lptr[l], lrows.begin() + lptr[l + 1], [&](int a, int b) {
        const int wa = pptr[rdiag[a]] - pptr[rstart[a]] + (rdiag[a] - rstart[a]);
        const int wb = pptr[rdiag[b]] - pptr[rstart[b]] + (rdiag[b] - rstart[b]);
        return wa > wb;
      });
    std::vector<int4> frec(n), brec(n);
    for (int p = 0; p < n; ++p) {
      const int i = lrows[p], j = urows[p];
      frec[p] = make_int4(i, rstart[i], rdiag[i], rend[i]);
      brec[p] = make_int4(j, rdiag[j] + 1, rend[j], 0);
    }
    // release an earlier setup
    if (s.has_ilu) fail(-6, "the ILU pattern is already set");
    IluDev& L = s.ilu;
    L.n = n;
    L.nnz = nnz;
    L.rdiag = s.own(dupload(rdiag));
    L.col = s.own(dupload(std::vector<int>(cols, cols + nnz)));
    L.pptr = s.own(dupload(pptr));
    L.pt = s.own(dupload(pt));
    L.pe2 = s.own(dupload(pe2));
    L.nlev = nlev;
    L.nlevu = nlevu;
    L.lptr = s.own(dupload(lptr));
    L.uptr = s.own(dupload(uptr));
    L.frec = s.own(dupload(frec));
    L.brec = s.own(dupload(brec));
    L.perm = s.own(dupload(permv));
    L.lds_levels = nlev + nlevu + 2 <= ILU_LDS_LEVELS;
    int rowcap = 1, paircap = 1, lowcap = 1;
    for (int i = 0; i < n; ++i) {
      rowcap = std::max(rowcap, rend[i] - rstart[i]);
      paircap = std::max(paircap, pptr[rdiag[i]] - pptr[rstart[i]]);
      lowcap = std::max(lowcap, rdiag[i] - rstart[i]);
    }
    L.rowcap = rowcap;
    L.paircap = paircap;
    L.lowcap = lowcap;
    {
      int lds_max = 65536;
      HC(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, a2d::require_device()));
      const size_t lev = L.lds_levels ? size_t(nlev + 1) * sizeof(int) : 0;
      const size_t per = size_t(rowcap + lowcap) * 25 * sizeof(double) + size_t(paircap) * 8 +
                         size_t(lowcap + 1) * 4;
      const long fit = (long(lds_max) - long(lev)) / long(per);
      L.lds_waves = fit >= 4 ? int(std::min<long>(fit, ILU_TPB / 64)) : 0;
      if (getenv("A2D_RANS_FACTOR_GLOBAL")) L.lds_waves = 0;
    }
    {
      // staged copy for the LDS sweeps: each level's blocks contiguous
      std::vector<int> ssrc, flev(nlev + 1), blev(nlevu + 1);
      std::vector<int4> fsrec(n), bsrec(n);
      ssrc.reserve(size_t(nnz));
      int scap = 0;
      for (int l = 0; l < nlev; ++l) {
        flev[l] = int(ssrc.size());
        for (int p = lptr[l]; p < lptr[l + 1]; ++p) {
          const int i = lrows[p];
          fsrec[p] = make_int4(i, rstart[i], rdiag[i], int(ssrc.size()) - flev[l]);
          for (int e = rstart[i]; e < rdiag[i]; ++e) ssrc.push_back(e);
        }
        scap = std::max(scap, (int(ssrc.size()) - flev[l]) * 25);
      }
      flev[nlev] = int(ssrc.size());
      for (int l = 0; l < nlevu; ++l) {
        blev[l] = int(ssrc.size());
        for (int p = uptr[l]; p < uptr[l + 1]; ++p) {
          const int j = urows[p];
          bsrec[p] = make_int4(j, rdiag[j] + 1, rend[j], int(ssrc.size()) - blev[l]);
          for (int e = rdiag[j] + 1; e < rend[j]; ++e) ssrc.push_back(e);
          ssrc.push_back(nnz + j);
        }
        scap = std::max(scap, (int(ssrc.size()) - blev[l]) * 25);
      }
      blev[nlevu] = int(ssrc.size());
      L.nst = int(ssrc.size());
      L.scap = scap;
      L.ccap = (scap / 25 + 3) & ~3;
      int rcap = 0;
      for (int l = 0; l < nlev; ++l) rcap = std::max(rcap, lptr[l + 1] - lptr[l]);
      for (int l = 0; l < nlevu; ++l) rcap = std::max(rcap, uptr[l + 1] - uptr[l]);
      L.rcap = rcap;
      std::vector<int> scol(ssrc.size());
      for (size_t b = 0; b < ssrc.size(); ++b) scol[b] = ssrc[b] < nnz ? cols[ssrc[b]] : 0;
      int lds_max = 65536;
      HC(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, a2d::require_device()));
      L.lds_solve = scap <= SOLVE_K * (ILU_TPB / 2) && ilu_solve_lds_bytes(L) <= size_t(lds_max) &&
                    !getenv("A2D_RANS_SOLVE_GLOBAL");
      if (L.lds_solve) {
        L.fsrec = s.own(dupload(fsrec));
        L.bsrec = s.own(dupload(bsrec));
        L.flev = s.own(dupload(flev));
        L.blev = s.own(dupload(blev));
        L.ssrc = s.own(dupload(ssrc));
        L.scol = s.own(dupload(scol));
      }
    }
    s.posD = s.own(dupload(std::vector<int>(pos_src, pos_src + n)));
    s.posLR = s.own(dupload(std::vector<int>(pos_src + n, pos_src + n + nfi)));
    s.posRL = s.own(dupload(std::vector<int>(pos_src + n + nfi, pos_src + n + 2 * nfi)));
    s.has_ilu = true;
    relink(get_group(h));
  });
}

int a2d_rans_ilu_info(int64_t h, int* out4) {
  return guarded_h(h, [&] {
    Solver& s = get(h);
    if (!s.has_ilu) fail(-6, "a2d_rans_ilu_setup has not been called");
    out4[0] = s.ilu.nnz;
    out4[1] = s.ilu.nlev;
    out4[2] = s.ilu.nlevu;
    out4[3] = s.ilu.lds_waves;
  });
}

int a2d_rans_ilu_factor(int64_t h, const double* winf, double gam, double cfl, double turb_factor, int* nonfinite) {
  return guarded_h(h, [&] {
    Solver& s = get(h);
    s.ilu_reusable = false;
    hipLaunchKernelGGL(k_local_dt, nblk(s.m.nc), TPB, 0, s.st, s.m, s.P, s.Q, cfl, s.dt, s.vdt);
    factor(s, winf, gam, 1.0 / turb_factor, nullptr, nullptr);
    if (nonfinite) *nonfinite = count_nonfinite(s, s.invU, s.m.nc * 25);
    HC(hipStreamSynchronize(s.st));
  });
}

int a2d_rans_ilu_apply(int64_t h, const double* b, double* x) {
  return guarded_h(h, [&] {
    Solver& s = get(h);
    if (!s.has_ilu) fail(-6, "a2d_rans_ilu_setup has not been called");
    upload_cells(s, b, s.bvec);
    ilu_solve(s, s.bvec, s.tmp);
    download_cells(s, s.tmp, x);
  });
}

int a2d_rans_newton_step(int64_t h, int order, const double* winf, double gam, double cfl, double turb_factor,
                         double gmres_tol, int gmres_iter, int right, int step_limit_cells, int reuse_ilu,
                         double* out) {
  return guarded_h(h, [&] {
    Solver& s = get(h);
    if (order != 1 && order != 2) fail(-6, "order must be 1 or 2");
    if (gmres_iter < 1 || gmres_iter > GMRES_MAX) fail(-6, "gmres_iter must be in [1, 120]");
    if (!(turb_factor > 0.0)) fail(-6, "turb_factor must be positive");
    const int N = s.N();
    const double ct_inv = 1.0 / turb_factor;
    for (int k = 0; k < 12; ++k) out[k] = 0.0;
    // residual of the state (the Newton right-hand side)
    far_state(s, winf, gam);
    eval(s, order, nullptr, s.R, nullptr, nullptr, 1.0);
    // pseudo-time step, first-order Jacobian + vol/dt, ILU (or the last
    // factors kept: only the preconditioner lags, the Newton matrix-vector
    // product uses this step's vol/dt)
    hipLaunchKernelGGL(k_local_dt, nblk(s.m.nc), TPB, 0, s.st, s.m, s.P, s.Q, cfl, s.dt, s.vdt);
    if (!(reuse_ilu && s.ilu_reusable)) {
      s.ilu_reusable = false;
      factor(s, winf, gam, ct_inv, nullptr, nullptr);
      out[9] = 1.0;
      if (count_nonfinite(s, s.invU, s.m.nc * 25) > 0) {
        out[5] = 1.0;                     // ILU not finite: no step
        return;
      }
      s.ilu_reusable = true;
    }
    const double rnorm = norm2(s, s.R);
    hipLaunchKernelGGL(k_negate, nblk(N), TPB, 0, s.st, N, s.R, s.bvec);
    far_state(s, winf, gam);
    int info = 0;
    int its = gmres(s, order, ct_inv, s.bvec, s.dq, gmres_tol, gmres_iter, right != 0, &info);
    double lin = lin_residual(s, order, ct_inv, s.dq, rnorm);
    int total = its;
    if (!right && !(std::isfinite(lin) && lin < 0.5)) {
      // left preconditioning can stop on |M^-1 r| while the true residual is large:
      // retry with right preconditioning and keep the better (aero2d/rans.py:1791-1801)
      int info2 = 0;
      int its2 = gmres(s, order, ct_inv, s.bvec, s.dq2, gmres_tol, gmres_iter, true, &info2);
      const double lin2 = lin_residual(s, order, ct_inv, s.dq2, rnorm);
      total += its2;
      if (std::isfinite(lin2) && (!std::isfinite(lin) || lin2 < lin)) {
        HC(hipMemcpyAsync(s.dq, s.dq2, N * sizeof(double), hipMemcpyDeviceToDevice, s.st));
        lin = lin2;
        info = info2;
      }
    }
    out[0] = lin;
    out[1] = double(total);
    out[2] = double(info);
    if (!std::isfinite(lin) || lin > 0.9) return;   // linear solve failed: no step
    // step limit (aero2d/rans.py:1822-1847)
    const int nb = s.nparts_cell;
    double* pmax = s.part;
    double* pcnt = s.part + nb;
    hipLaunchKernelGGL(k_step_rel, nb, TPB, 0, s.st, s.m.nc, s.Q, s.dq, s.reli, pmax, pcnt);
    hipLaunchKernelGGL(k_max_rows, 1, 64, 0, s.st, nb, pmax, pcnt, s.small);
    double mc[2];
    HC(hipMemcpyAsync(mc, s.small, 2 * sizeof(double), hipMemcpyDeviceToHost, s.st));
    HC(hipStreamSynchronize(s.st));
    const double rel = mc[0];
    const double lam = rel < 0.2 ? 1.0 : 0.2 / rel;
    const int nbig = int(mc[1]);
    const bool per_cell = lam < 1.0 && step_limit_cells > 0 && nbig <= step_limit_cells;
    HC(hipMemcpyAsync(s.Qprev, s.Q, N * sizeof(double), hipMemcpyDeviceToDevice, s.st));
    hipLaunchKernelGGL(k_apply, nblk(s.m.nc), TPB, 0, s.st, s.m.nc, s.Q, s.dq, s.reli, lam, per_cell ? 1 : 0);
    HC(hipGetLastError());
    HC(hipStreamSynchronize(s.st));
    out[3] = lam;
    out[4] = rel;
    out[6] = 1.0;                        // update applied
    out[7] = double(nbig);
    out[8] = per_cell ? 1.0 : 0.0;
  });
}

int a2d_rans_time_kernel(int64_t h, int which, int order, const double* winf, double gam, int reps, double* ms) {
  return guarded_h(h, [&] {
    Solver& s = get(h);
    if (reps < 1) fail(-6, "reps must be >= 1");
    if (which < 0 || which > 4) fail(-6, "which must be 0..4");
    if ((which == 0 || which == 1) && !s.has_ilu) fail(-6, "a2d_rans_ilu_setup has not been called");
    far_state(s, winf, gam);
    // inputs: the state's residual as the vector, a factorisation in place
    eval(s, order, nullptr, s.R, nullptr, nullptr, 1.0);
    if (which == 0 || which == 1) {
      s.ilu_reusable = false;
      hipLaunchKernelGGL(k_local_dt, nblk(s.m.nc), TPB, 0, s.st, s.m, s.P, s.Q, 10.0, s.dt, s.vdt);
      factor(s, winf, gam, 1.0, nullptr, nullptr);
    }
    auto once = [&] {
      switch (which) {
        case 0: ilu_solve(s, s.R, s.tmp); break;
        case 1: launch_ilu_factor(s.ilu, s.B, s.invU, s.S, nullptr, 1, 0, 0, 0, s.st); break;
        case 2: eval(s, order, s.R, nullptr, s.tmp, s.vdt, 1.0); break;
        case 3: eval(s, order, nullptr, s.tmp, nullptr, nullptr, 1.0); break;
        case 4: {
          hipLaunchKernelGGL(k_jac_int, nblk(s.m.nfi), TPB, 0, s.st, s.m, s.P, s.Q, s.J, nullptr, s.posLR, s.posRL);
          hipLaunchKernelGGL(k_jac_wall, nblk(s.m.nfw), TPB, 0, s.st, s.m, s.P, s.Q, s.JW);
          hipLaunchKernelGGL(k_jac_far, nblk(s.m.nff), TPB, 0, s.st, s.m, s.P, s.Q, s.wgf, s.JF);
          hipLaunchKernelGGL(k_jac_diag, nblk(s.m.nc), TPB, 0, s.st, s.m, s.P, s.Q, s.g, s.J, s.JW, s.JF, s.vdt,
                             1.0, nullptr, s.posD, s.tmp2);
          break;
        }
      }
    };
    once();
    hipEvent_t e0, e1;
    HC(hipEventCreate(&e0));
    HC(hipEventCreate(&e1));
    HC(hipEventRecord(e0, s.st));
    for (int r = 0; r < reps; ++r) once();
    HC(hipEventRecord(e1, s.st));
    HC(hipEventSynchronize(e1));
    float t = 0.f;
    HC(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HC(hipGetLastError());
    *ms = double(t) / reps;
  });
}

int a2d_rans_newton_steps(int64_t h, int n, const int* case_ids, const int* order, const double* winf,
                          const double* gam, const double* cfl, const double* turb_factor, const double* gmres_tol,
                          int gmres_iter, const int* right, int step_limit_cells, const int* reuse_ilu,
                          double* out) {
  return guarded_h(h, [&] {
    Group& g = get_group(h);
    Solver& s0 = *g.cases[0];
    if (!s0.has_ilu) fail(-6, "a2d_rans_ilu_setup has not been called");
    if (n < 1) fail(-6, "n must be >= 1");
    if (gmres_iter < 1 || gmres_iter > GMRES_MAX) fail(-6, "gmres_iter must be in [1, 120]");
    std::vector<int> ids(case_ids, case_ids + n), sorted(ids);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end()) fail(-6, "a batch must not name a case twice");
    for (int i = 0; i < n; ++i) {
      if (ids[i] < 0 || ids[i] >= int(g.cases.size())) fail(-4, "no such case slot");
      if (order[i] != 1 && order[i] != 2) fail(-6, "order must be 1 or 2");
      if (!(turb_factor[i] > 0.0)) fail(-6, "turb_factor must be positive");
    }
    const int N = s0.N();
    std::vector<Solver*> cs(n);
    std::vector<double> ct(n);
    for (int i = 0; i < n; ++i) {
      cs[i] = g.cases[ids[i]].get();
      ct[i] = 1.0 / turb_factor[i];
      for (int k = 0; k < 12; ++k) out[size_t(i) * 12 + k] = 0.0;
    }
    // residual, pseudo-time step and first-order Jacobian of every case, then
    // one factorisation launch for all that factor (the others keep their
    // last factors)
    std::vector<int> fids;
    std::vector<char> fresh(n);
    for (int i = 0; i < n; ++i) {
      Solver& s = *cs[i];
      far_state(s, winf + 5 * i, gam[i]);
      eval(s, order[i], nullptr, s.R, nullptr, nullptr, 1.0);
      hipLaunchKernelGGL(k_local_dt, nblk(s.m.nc), TPB, 0, s.st, s.m, s.P, s.Q, cfl[i], s.dt, s.vdt);
      fresh[i] = !(reuse_ilu && reuse_ilu[i] && s.ilu_reusable);
      if (!fresh[i]) continue;
      s.ilu_reusable = false;
      factor(s, winf + 5 * i, gam[i], ct[i], nullptr, nullptr, false);
      fids.push_back(ids[i]);
      out[size_t(i) * 12 + 9] = 1.0;
    }
    if (!fids.empty()) {
      launch_ilu_factor(s0.ilu, g.Ball, g.Iall, g.Sall, case_list(g, fids), int(fids.size()), g.sB, g.sI, g.sS,
                        s0.st);
      HC(hipGetLastError());
    }
    HC(hipMemsetAsync(g.nfall, 0, g.cases.size() * sizeof(int), s0.st));
    for (int i = 0; i < n; ++i)
      if (fresh[i])
        hipLaunchKernelGGL(k_count_nonfinite, nblk(s0.m.nc * 25), TPB, 0, s0.st, s0.m.nc * 25, cs[i]->invU,
                           cs[i]->nonfinite);
    for (int i = 0; i < n; ++i) mdot(*cs[i], cs[i]->R, size_t(N), 1, cs[i]->R, 0);
    int* const nf = g.nf_h;
    double* const hall = g.hall_h;
    const size_t hall_bytes = g.cases.size() * HSTRIDE * sizeof(double);
    HC(hipMemcpyAsync(nf, g.nfall, g.cases.size() * sizeof(int), hipMemcpyDeviceToHost, s0.st));
    HC(hipMemcpyAsync(hall, g.hall, hall_bytes, hipMemcpyDeviceToHost, s0.st));
    HC(hipStreamSynchronize(s0.st));
    std::vector<double> rnorm(n);
    std::vector<int> live;
    for (int i = 0; i < n; ++i) {
      rnorm[i] = std::sqrt(hall[size_t(ids[i]) * HSTRIDE]);
      if (nf[ids[i]] > 0) {
        out[size_t(i) * 12 + 5] = 1.0;   // ILU not finite: no step
        continue;
      }
      if (fresh[i]) cs[i]->ilu_reusable = true;
      live.push_back(i);
      hipLaunchKernelGGL(k_negate, nblk(N), TPB, 0, s0.st, N, cs[i]->R, cs[i]->bvec);
    }
    // GMRES, grouped by preconditioning side
    std::vector<double> lin(n, NAN);
    std::vector<int> total(n, 0), infos(n, 0);
    auto run = [&](const std::vector<int>& sel, bool rt, bool into_dq2, std::vector<int>& itsv,
                   std::vector<int>& infv) {
      std::vector<int> sid, so;
      std::vector<double> sct, stol;
      std::vector<const double*> sb;
      std::vector<double*> sx;
      for (int i : sel) {
        sid.push_back(ids[i]);
        so.push_back(order[i]);
        sct.push_back(ct[i]);
        stol.push_back(gmres_tol[i]);
        sb.push_back(cs[i]->bvec);
        sx.push_back(into_dq2 ? cs[i]->dq2 : cs[i]->dq);
      }
      for (int i : sel) far_state(*cs[i], winf + 5 * i, gam[i]);
      gmres_batch(g, sid, so, sct, sb, sx, stol, gmres_iter, rt, itsv, infv);
    };
    auto lins = [&](const std::vector<int>& sel, bool use_dq2) {
      for (int i : sel) {
        Solver& s = *cs[i];
        matvec(s, order[i], use_dq2 ? s.dq2 : s.dq, s.tmp2, ct[i]);
        hipLaunchKernelGGL(k_add, nblk(N), TPB, 0, s0.st, N, s.tmp2, s.R, s.tmp2);
        mdot(s, s.tmp2, size_t(N), 1, s.tmp2, 0);
      }
      HC(hipMemcpyAsync(hall, g.hall, hall_bytes, hipMemcpyDeviceToHost, s0.st));
      HC(hipStreamSynchronize(s0.st));
      std::vector<double> r(sel.size());
      for (size_t q = 0; q < sel.size(); ++q)
        r[q] = std::sqrt(hall[size_t(ids[sel[q]]) * HSTRIDE]) / std::max(rnorm[sel[q]], 1e-300);
      return r;
    };
    for (int side = 0; side < 2; ++side) {
      std::vector<int> sel;
      for (int i : live)
        if ((right[i] != 0) == (side == 1)) sel.push_back(i);
      if (sel.empty()) continue;
      std::vector<int> itsv, infv;
      run(sel, side == 1, false, itsv, infv);
      const std::vector<double> l = lins(sel, false);
      for (size_t q = 0; q < sel.size(); ++q) {
        lin[sel[q]] = l[q];
        total[sel[q]] = itsv[q];
        infos[sel[q]] = infv[q];
      }
    }
    // left preconditioning that did not reach a true residual below 0.5:
    // retry with right preconditioning, keep the better (aero2d/rans.py:1791-1801)
    std::vector<int> retry;
    for (int i : live)
      if (!right[i] && !(std::isfinite(lin[i]) && lin[i] < 0.5)) retry.push_back(i);
    if (!retry.empty()) {
      std::vector<int> itsv, infv;
      run(retry, true, true, itsv, infv);
      const std::vector<double> l2 = lins(retry, true);
      for (size_t q = 0; q < retry.size(); ++q) {
        const int i = retry[q];
        total[i] += itsv[q];
        if (std::isfinite(l2[q]) && (!std::isfinite(lin[i]) || l2[q] < lin[i])) {
          HC(hipMemcpyAsync(cs[i]->dq, cs[i]->dq2, N * sizeof(double), hipMemcpyDeviceToDevice, s0.st));
          lin[i] = l2[q];
          infos[i] = infv[q];
        }
      }
    }
    // step limit and update of the cases whose linear solve succeeded
    std::vector<int> upd;
    for (int i : live) {
      double* o = out + size_t(i) * 12;
      o[0] = lin[i];
      o[1] = double(total[i]);
      o[2] = double(infos[i]);
      if (!std::isfinite(lin[i]) || lin[i] > 0.9) continue;
      upd.push_back(i);
      Solver& s = *cs[i];
      const int nbp = s.nparts_cell;
      hipLaunchKernelGGL(k_step_rel, nbp, TPB, 0, s0.st, s.m.nc, s.Q, s.dq, s.reli, s.part, s.part + nbp);
      hipLaunchKernelGGL(k_max_rows, 1, 64, 0, s0.st, nbp, s.part, s.part + nbp, s.small);
    }
    if (upd.empty()) return;
    double* const sm = g.small_h;
    HC(hipMemcpyAsync(sm, g.smallall, g.cases.size() * 64 * sizeof(double), hipMemcpyDeviceToHost, s0.st));
    HC(hipStreamSynchronize(s0.st));
    for (int i : upd) {
      Solver& s = *cs[i];
      const double rel = sm[size_t(ids[i]) * 64];
      const double lam = rel < 0.2 ? 1.0 : 0.2 / rel;
      const int nbig = int(sm[size_t(ids[i]) * 64 + 1]);
      const bool per_cell = lam < 1.0 && step_limit_cells > 0 && nbig <= step_limit_cells;
      HC(hipMemcpyAsync(s.Qprev, s.Q, N * sizeof(double), hipMemcpyDeviceToDevice, s0.st));
      hipLaunchKernelGGL(k_apply, nblk(s.m.nc), TPB, 0, s0.st, s.m.nc, s.Q, s.dq, s.reli, lam, per_cell ? 1 : 0);
      double* o = out + size_t(i) * 12;
      o[3] = lam;
      o[4] = rel;
      o[6] = 1.0;
      o[7] = double(nbig);
      o[8] = per_cell ? 1.0 : 0.0;
    }
    HC(hipGetLastError());
    HC(hipStreamSynchronize(s0.st));
  });
}

int a2d_rans_newton_dq(int64_t h, double* dq) {
  return guarded_h(h, [&] {
    Solver& s = get(h);
    download_cells(s, s.dq, dq);
  });
}

int a2d_rans_keep(int64_t h, int what) {
  return guarded_h(h, [&] {
    Solver& s = get(h);
    const size_t bytes = size_t(s.N()) * sizeof(double);
    switch (what) {
      case 0: HC(hipMemcpyAsync(s.Qbest, s.Q, bytes, hipMemcpyDeviceToDevice, s.st)); break;   // save best
      case 1: HC(hipMemcpyAsync(s.Q, s.Qbest, bytes, hipMemcpyDeviceToDevice, s.st)); break;   // restore best
      case 2: HC(hipMemcpyAsync(s.Q, s.Qprev, bytes, hipMemcpyDeviceToDevice, s.st)); break;   // undo the step
      default: fail(-6, "what must be 0 (save best), 1 (restore best) or 2 (undo the last step)");
    }
    HC(hipStreamSynchronize(s.st));
  });
}

}  // extern "C"
