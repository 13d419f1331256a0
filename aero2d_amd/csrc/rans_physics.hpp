// Pointwise physics of the RANS (Spalart-Allmaras, SA-neg) residual, written
// once over a scalar type S (double, or Dual<double> for the exact Jacobian-
// vector product and the first-order Jacobian columns). Data-dependent selects
// compare values only, so the dual instantiation follows the primal's branches
// as jax.jvp / jax.jacfwd do for the reference (aero2d/rans.py).
//
// Unknowns per cell: Q = (rho, rho u, rho v, rho E, rho chi), chi = nu~/nu_inf;
// primitives w = (rho, u, v, p, chi). Units: rho_inf = 1, a_inf = 1,
// p_inf = 1/gamma, T_inf = 1, chord 1, mu_inf = M/Re (aero2d/rans.py:1-24).
#pragma once
#include "dual.hpp"

namespace a2d {
namespace rans {

constexpr double GAM = 1.4;
constexpr double PR = 0.72, PRT = 0.9;
constexpr double SUTH = 110.4 / 288.15;
constexpr double CB1 = 0.1355, SIG = 2.0 / 3.0, CB2 = 0.622, KAP = 0.41;
constexpr double CW1 = CB1 / (KAP * KAP) + (1.0 + CB2) / SIG;
constexpr double CW2 = 0.3, CW3 = 2.0, CV1 = 7.1, CV2 = 0.7, CV3 = 0.9, CN1 = 16.0;

// |x| with derivative sign(x) (0 at 0), as jnp.abs.
__host__ __device__ __forceinline__ double dabs(double x) { return fabs(x); }
template <class T> __host__ __device__ __forceinline__ Dual<T> dabs(Dual<T> x) {
  return {x.v < T(0) ? -x.v : x.v, x.v > T(0) ? x.d : (x.v < T(0) ? -x.d : T(0))};
}
// min(x, c) for a constant c, balanced at a tie.
template <class T> __host__ __device__ __forceinline__ T dminc(T a, T c) { return a < c ? a : c; }
template <class T> __host__ __device__ __forceinline__ Dual<T> dminc(Dual<T> a, T c) {
  T dd = a.v < c ? a.d : (a.v > c ? T(0) : T(0.5) * a.d);
  return {a.v < c ? a.v : c, dd};
}

// Constants of one solver (the RansSolver fields the residual reads).
struct Params {
  double mach, mu_inf, nu_inf, chi_inf, limiter_k, xref_x, xref_y, zmin, chord;
  int turbulent;      // 0: laminar (rho chi == 0 equation)
  int limiter_zone;   // 0 off, 1 trailing-edge zone, 2 everywhere
  int sa_raw;         // 1: SA source from the unlimited gradients
  int visc_raw;       // 1: viscous fluxes from the unlimited gradients
};

// prim (aero2d/rans.py:455-461)
template <class S>
__host__ __device__ __forceinline__ void prim(const S q[5], S w[5]) {
  S r = q[0];
  S ri = drcp(r);
  S u = q[1] * ri, v = q[2] * ri;
  w[0] = r;
  w[1] = u;
  w[2] = v;
  w[3] = K<S>(GAM - 1.0) * (q[3] - K<S>(0.5) * r * (u * u + v * v));
  w[4] = q[4] * ri;
}

// Sutherland's law (aero2d/rans.py:477-478); T^1.5 as T sqrt(T).
template <class S>
__host__ __device__ __forceinline__ S mu_lam(const Params& P, S T) {
  return K<S>(P.mu_inf * (1.0 + SUTH)) * T * dsqrt(T) / (T + K<S>(SUTH));
}

// sa_terms (aero2d/rans.py:526-535): mu, mu_t, X, fv1, fn.
template <class S>
__host__ __device__ __forceinline__ void sa_terms(const Params& P, S rho, S chi, S T, S& mu, S& mut, S& X,
                                                  S& fv1, S& fn) {
  mu = mu_lam(P, T);
  S nut = chi * K<S>(P.nu_inf);
  X = rho * nut / mu;
  S X3 = X * X * X;
  fv1 = X3 / (X3 + K<S>(CV1 * CV1 * CV1));
  mut = sel(val(chi) > 0.0, rho * nut * fv1, K<S>(0.0));
  fn = sel(val(chi) >= 0.0, K<S>(1.0), (K<S>(CN1) + X3) / (K<S>(CN1) - X3));
}

// LM-Roe flux with entropy fix and mass-flux upwinding of chi (aero2d/rans.py:480-524).
template <class S>
__host__ __device__ __forceinline__ void roe(const Params& P, const S wL[5], const S wR[5], double nx,
                                             double ny, S F[5]) {
  const S rL = wL[0], uL = wL[1], vL = wL[2], pL = wL[3];
  const S rR = wR[0], uR = wR[1], vR = wR[2], pR = wR[3];
  const double g1 = GAM / (GAM - 1.0);
  S unL = uL * nx + vL * ny, unR = uR * nx + vR * ny;
  S qL2 = uL * uL + vL * vL, qR2 = uR * uR + vR * vR;
  S HL = K<S>(g1) * pL / rL + K<S>(0.5) * qL2;
  S HR = K<S>(g1) * pR / rR + K<S>(0.5) * qR2;
  S sL = dsqrt(rL), sR = dsqrt(rR);
  S rh = sL * sR;
  S den = sL + sR;
  S uh = (sL * uL + sR * uR) / den;
  S vh = (sL * vL + sR * vR) / den;
  S Hh = (sL * HL + sR * HR) / den;
  S q2 = uh * uh + vh * vh;
  S ah = dsqrt(dmaxc(K<S>(GAM - 1.0) * (Hh - K<S>(0.5) * q2), 1e-12));
  S unh = uh * nx + vh * ny;
  const double tx = -ny, ty = nx;
  S uth = uh * tx + vh * ty;
  S dr = rR - rL, dp = pR - pL;
  S dun = unR - unL;
  S dut = (uR - uL) * tx + (vR - vL) * ty;
  S aL = dsqrt(K<S>(GAM) * pL / rL), aR = dsqrt(K<S>(GAM) * pR / rR);
  S z = dmin(K<S>(1.0), dmax(dsqrt(qL2) / aL, dsqrt(qR2) / aR));
  z = dmaxc(z, P.zmin);
  S ah2 = ah * ah;
  S a1 = (dp - rh * ah * z * dun) / (K<S>(2.0) * ah2);
  S a2 = dr - dp / ah2;
  S a3 = (dp + rh * ah * z * dun) / (K<S>(2.0) * ah2);
  S at = rh * dut;
  S dl = K<S>(0.05) * ah;
  auto fix = [&](S l) {
    S al = dabs(l);
    return sel(val(al) < val(dl), (l * l + dl * dl) / (K<S>(2.0) * dl), al);
  };
  S l1 = fix(unh - ah), l2 = dabs(unh), l3 = fix(unh + ah);
  S d0 = l1 * a1 + l2 * a2 + l3 * a3;
  S d1 = l1 * a1 * (uh - ah * nx) + l2 * (a2 * uh + at * tx) + l3 * a3 * (uh + ah * nx);
  S d2 = l1 * a1 * (vh - ah * ny) + l2 * (a2 * vh + at * ty) + l3 * a3 * (vh + ah * ny);
  S d3 = l1 * a1 * (Hh - ah * unh) + l2 * (a2 * K<S>(0.5) * q2 + at * uth) + l3 * a3 * (Hh + ah * unh);
  S FL0 = rL * unL, FR0 = rR * unR;
  F[0] = K<S>(0.5) * (FL0 + FR0) - K<S>(0.5) * d0;
  F[1] = K<S>(0.5) * ((rL * uL * unL + pL * nx) + (rR * uR * unR + pR * nx)) - K<S>(0.5) * d1;
  F[2] = K<S>(0.5) * ((rL * vL * unL + pL * ny) + (rR * vR * unR + pR * ny)) - K<S>(0.5) * d2;
  F[3] = K<S>(0.5) * (rL * HL * unL + rR * HR * unR) - K<S>(0.5) * d3;
  S mf = F[0];
  F[4] = K<S>(0.5) * mf * (wL[4] + wR[4]) - K<S>(0.5) * dabs(mf) * (wR[4] - wL[4]);
}

// Viscous flux from face values and face gradients (aero2d/rans.py:537-571, SA).
template <class S>
__host__ __device__ __forceinline__ void visc_flux(const Params& P, const S wf[5], const S gu[2], const S gv[2],
                                                   const S gT[2], const S gc[2], double nx, double ny, S F[5]) {
  S rho = wf[0], u = wf[1], v = wf[2], p = wf[3], chi = wf[4];
  S T = K<S>(GAM) * p / rho;
  S mu, mut, X, fv1, fn;
  sa_terms(P, rho, chi, T, mu, mut, X, fv1, fn);
  S me = mu + mut;
  S div = gu[0] + gv[1];
  S txx = me * (K<S>(2.0) * gu[0] - K<S>(2.0 / 3.0) * div);
  S tyy = me * (K<S>(2.0) * gv[1] - K<S>(2.0 / 3.0) * div);
  S txy = me * (gu[1] + gv[0]);
  S fx = txx * nx + txy * ny;
  S fy = txy * nx + tyy * ny;
  S k = (mu / K<S>(PR) + mut / K<S>(PRT)) / K<S>(GAM - 1.0);
  F[0] = K<S>(0.0);
  F[1] = fx;
  F[2] = fy;
  F[3] = u * fx + v * fy + k * (gT[0] * nx + gT[1] * ny);
  F[4] = (mu + rho * chi * K<S>(P.nu_inf) * fn) / K<S>(SIG) * (gc[0] * nx + gc[1] * ny);
}

// SA-neg source per unit volume in rho chi (aero2d/rans.py:609-629); g[k][d]
// is d w_k / dx_d.
template <class S>
__host__ __device__ __forceinline__ S sa_source(const Params& P, const S w[5], const S g[5][2], double wd) {
  S rho = w[0], p = w[3], chi = w[4];
  S T = K<S>(GAM) * p / rho;
  S mu, mut, X, fv1, fn;
  sa_terms(P, rho, chi, T, mu, mut, X, fv1, fn);
  S Om = dabs(g[2][0] - g[1][1]);
  S nut = chi * K<S>(P.nu_inf);
  S fv2 = K<S>(1.0) - X / (K<S>(1.0) + X * fv1);
  const double kd2 = KAP * KAP * wd * wd;
  S Sbar = nut * fv2 / K<S>(kd2);
  S St = sel(val(Sbar) >= -CV2 * val(Om), Om + Sbar,
             Om + Om * (K<S>(CV2 * CV2) * Om + K<S>(CV3) * Sbar) /
                      (K<S>(CV3 - 2.0 * CV2) * Om - Sbar + K<S>(1e-30)));
  St = dmaxc(St, 1e-16);
  S r = dminc(nut / (St * K<S>(kd2)), 10.0);
  S r2 = r * r;
  S r6 = r2 * r2 * r2;
  S gg = r + K<S>(CW2) * (r6 - r);
  S g2 = gg * gg;
  S g6 = g2 * g2 * g2;
  const double cw36 = CW3 * CW3 * CW3 * CW3 * CW3 * CW3;
  S fw = gg * dpow(K<S>(1.0 + cw36) / (g6 + K<S>(cw36)), 1.0 / 6.0);
  const bool pos = val(chi) >= 0.0;
  S prod = pos ? K<S>(CB1) * St * chi : K<S>(CB1) * Om * chi;
  S dest = pos ? K<S>(CW1) * fw * nut * chi / K<S>(wd * wd) : K<S>(-CW1) * nut * chi / K<S>(wd * wd);
  S cross = K<S>(CB2 / SIG * P.nu_inf) * (g[4][0] * g[4][0] + g[4][1] * g[4][1]);
  return rho * (prod - dest + cross);
}

// Venkatakrishnan's psi of one face side (aero2d/rans.py:1012-1016).
template <class S>
__host__ __device__ __forceinline__ S venkat_psi(S d2, S dmx, S dmn, double eps2) {
  S d1 = sel(val(d2) > 0.0, dmx, dmn);
  S num = (d1 * d1 + K<S>(eps2)) * d2 + K<S>(2.0) * d2 * d2 * d1;
  S den = d2 * (d1 * d1 + K<S>(2.0) * d2 * d2 + d1 * d2 + K<S>(eps2)) + K<S>(1e-300);
  S c = dmin(dmax(num / den, K<S>(0.0)), K<S>(1.0));
  return sel(fabs(val(d2)) < 1e-14, K<S>(1.0), c);
}

// Running extremum with JAX's scatter rule: the derivative is the mean of the
// tangents of every candidate equal to the extremum.
template <class S> struct Ext {
  S v;
  int n;
};
template <class T>
__host__ __device__ __forceinline__ void ext_push_max(Ext<T>& e, T x) { if (x > e.v) e.v = x; }
template <class T>
__host__ __device__ __forceinline__ void ext_push_max(Ext<Dual<T>>& e, Dual<T> x) {
  if (x.v > e.v.v) { e.v = x; e.n = 1; }
  else if (x.v == e.v.v) { e.v.d += x.d; e.n += 1; }
}
template <class T>
__host__ __device__ __forceinline__ void ext_push_min(Ext<T>& e, T x) { if (x < e.v) e.v = x; }
template <class T>
__host__ __device__ __forceinline__ void ext_push_min(Ext<Dual<T>>& e, Dual<T> x) {
  if (x.v < e.v.v) { e.v = x; e.n = 1; }
  else if (x.v == e.v.v) { e.v.d += x.d; e.n += 1; }
}
template <class T> __host__ __device__ __forceinline__ T ext_value(const Ext<T>& e) { return e.v; }
template <class T> __host__ __device__ __forceinline__ Dual<T> ext_value(const Ext<Dual<T>>& e) {
  return {e.v.v, e.v.d / T(e.n)};
}

}  // namespace rans
}  // namespace a2d
