// Pointwise physics of the finite-volume Euler residual, written once over a
// scalar type S (T = float/double for the residual, Dual<T> for its exact
// tangent-linear). Geometry (face vectors, positions) is always plain T.
//
// Non-dimensionalisation: rho_inf = 1, a_inf = 1, p_inf = 1/gamma, |V_inf| = M.
#pragma once
#include "dual.hpp"

namespace a2d {

constexpr double GAMMA = 1.4;                 // aero2d/euler.py:27
constexpr double GM1 = GAMMA - 1.0;
constexpr double PI_D = 3.141592653589793;

template <class S>
struct W4 {  // primitive state (rho, u, v, p)
  S r, u, v, p;
};

// (rho, rho u, rho v, E) -> (rho, u, v, p)   (aero2d/euler.py:66-71)
template <class S>
__host__ __device__ __forceinline__ W4<S> cons_to_prim(S q0, S q1, S q2, S q3) {
  const S ir = drcp(q0);
  S u = q1 * ir;
  S v = q2 * ir;
  S p = K<S>(GM1) * (q3 - K<S>(0.5) * q0 * (u * u + v * v));
  return {q0, u, v, p};
}

// HLLC flux through a face with area vector (sx, sy) (aero2d/euler.py:79-121).
// Einfeldt bounds from Roe averages; the four-way select of the reference
//   SL >= 0 -> F_L, SM >= 0 -> F*_L, SR > 0 -> F*_R, else F_R
// is evaluated as: pick side L when SL >= 0 or SM >= 0 (else R), then add the
// star correction unless that side is supersonic. Only the chosen side's flux
// is formed, so every lane runs one path.
//
// Divisions are the expensive FP64 operations on this path, so the flux is
// arranged around six of them: 1/area, 1/rho_L, 1/rho_R, the Roe-average
// denominator, S_M, and 1/(S_k - S_M), which gives the whole star state: with
// d = 1/(S_k - S_M) and f = rho (S_k - q) d,
//   f E/rho = (S_k - q) E d,   f p/(rho (S_k - q)) = p d.
template <class S, class T>
__host__ __device__ __forceinline__ void hllc_flux(const W4<S>& L, const W4<S>& R, T sx, T sy, S F[4]) {
  const T area = fsqrt(sx * sx + sy * sy);
  const T ia = frcp(area);
  const T nx = sx * ia, ny = sy * ia;
  const S irL = drcp(L.r), irR = drcp(R.r);
  const S qL = L.u * nx + L.v * ny;
  const S qR = R.u * nx + R.v * ny;
  const S cL = dsqrt(K<S>(GAMMA) * L.p * irL);
  const S cR = dsqrt(K<S>(GAMMA) * R.p * irR);
  const S EL = L.p * K<S>(1.0 / GM1) + K<S>(0.5) * L.r * (L.u * L.u + L.v * L.v);
  const S ER = R.p * K<S>(1.0 / GM1) + K<S>(0.5) * R.r * (R.u * R.u + R.v * R.v);
  const S sl = dsqrt(L.r), sr = dsqrt(R.r);
  const S iden = drcp(sl + sr);
  const S qt = (sl * qL + sr * qR) * iden;
  const S HL = (EL + L.p) * irL;
  const S HR = (ER + R.p) * irR;
  const S Ht = (sl * HL + sr * HR) * iden;
  const S ut = (sl * L.u + sr * R.u) * iden;
  const S vt = (sl * L.v + sr * R.v) * iden;
  const S ct = dsqrt(dmaxc(K<S>(GM1) * (Ht - K<S>(0.5) * (ut * ut + vt * vt)), base_t<S>(1e-12)));
  const S SL = dmin(qL - cL, qt - ct);
  const S SR = dmax(qR + cR, qt + ct);
  const S SM = (R.p - L.p + L.r * qL * (SL - qL) - R.r * qR * (SR - qR)) /
               (L.r * (SL - qL) - R.r * (SR - qR));
  const bool left = (val(SL) >= 0) || (val(SM) >= 0);
  const bool star = left ? !(val(SL) >= 0) : (val(SR) > 0);
  const S r = sel(left, L.r, R.r), u = sel(left, L.u, R.u), v = sel(left, L.v, R.v);
  const S p = sel(left, L.p, R.p), E = sel(left, EL, ER), q = sel(left, qL, qR);
  const S Sk = sel(left, SL, SR);
  // physical flux of the chosen side
  S f0 = r * q;
  S f1 = r * u * q + p * nx;
  S f2 = r * v * q + p * ny;
  S f3 = (E + p) * q;
  if (star) {
    const S skq = Sk - q;
    const S d = drcp(Sk - SM);
    const S f = r * skq * d;
    const S smq = SM - q;
    f0 = f0 + Sk * (f - r);
    f1 = f1 + Sk * (f * (u + smq * nx) - r * u);
    f2 = f2 + Sk * (f * (v + smq * ny) - r * v);
    f3 = f3 + Sk * ((skq * E * d + f * smq * SM + smq * p * d) - E);
  }
  F[0] = f0 * area;
  F[1] = f1 * area;
  F[2] = f2 * area;
  F[3] = f3 * area;
}

// Smooth van Albada limiter, LIMITER="smooth" (aero2d/euler.py:124-137);
// eps already includes the +1e-20.
template <class S, class T>
__host__ __device__ __forceinline__ S van_albada(S a, S b, T eps) {
  return ((b * b + eps) * a + (a * a + eps) * b) / (a * a + b * b + T(2) * eps);
}

// Mirror ghost about the wall normal (aero2d/euler.py:143-153).
template <class S, class T>
__host__ __device__ __forceinline__ W4<S> mirror(const W4<S>& w, T nx, T ny) {
  S vn = w.u * nx + w.v * ny;
  return {w.r, w.u - K<S>(2.0) * vn * nx, w.v - K<S>(2.0) * vn * ny, w.p};
}

// Per-case freestream constants (winf = (1, M cos a, M sin a, 1/gamma)).
template <class T>
struct Inf {
  T r, u, v, p;
};

// Characteristic far-field ghost with the compressible point-vortex
// correction about (0.25, 0) (aero2d/euler.py:156-205). Geometry and the
// freestream are constants; the interior state and the circulation carry
// tangents in the dual instantiation.
template <class S, class T>
__host__ __device__ __forceinline__ W4<S> farfield_ghost(const W4<S>& wi, T sx, T sy, const Inf<T>& f,
                                                T xf, T yf, S circ) {
  const T area = sqrt(sx * sx + sy * sy);
  const T nx = sx / area, ny = sy / area;
  const T minf2 = (f.u * f.u + f.v * f.v) / (T(GAMMA) * f.p / f.r);
  const T dxv = xf - T(0.25), dyv = yf;
  const T alpha = atan2(f.v, f.u);
  const T phi = atan2(dyv, dxv);
  const T rr = sqrt(dxv * dxv + dyv * dyv);
  const T beta = sqrt(dmaxc(T(1) - minf2, T(1e-6)));
  const T sp = sin(phi - alpha);
  const T kfac = beta / (T(2.0 * PI_D) * rr * (T(1) - minf2 * sp * sp));
  const S fac = (minf2 < T(1)) ? circ * kfac : K<S>(0.0);
  const S ub = f.u + fac * T(sin(phi));
  const S vb = f.v - fac * T(cos(phi));
  const T q2inf = f.u * f.u + f.v * f.v;
  const S q2b = ub * ub + vb * vb;
  const T cinf2 = T(GAMMA) * f.p / f.r;
  const S cb2 = dmaxc(cinf2 + T(0.5 * GM1) * (q2inf - q2b), T(1e-6));
  const S pb = f.p * dpow(cb2 / cinf2, GAMMA / GM1);
  const S rb = f.r * dpow(cb2 / cinf2, 1.0 / GM1);
  const S ci = dsqrt(K<S>(GAMMA) * wi.p / wi.r);
  const S cb = dsqrt(K<S>(GAMMA) * pb / rb);
  const S qni = wi.u * nx + wi.v * ny;
  const S qnb = ub * nx + vb * ny;
  const S Rp = qni + K<S>(2.0) * ci / K<S>(GM1);
  const S Rm = qnb - K<S>(2.0) * cb / K<S>(GM1);
  const S qn = K<S>(0.5) * (Rp + Rm);
  const S c = K<S>(0.25 * GM1) * (Rp - Rm);
  const bool outflow = val(qn) > 0;
  const S ut = outflow ? wi.u - qni * nx : ub - qnb * nx;
  const S vt = outflow ? wi.v - qni * ny : vb - qnb * ny;
  const S s = outflow ? wi.p / dpow(wi.r, GAMMA) : pb / dpow(rb, GAMMA);
  const S rho = dpow(c * c / (K<S>(GAMMA) * s), 1.0 / GM1);
  const S pres = rho * c * c / K<S>(GAMMA);
  W4<S> wb = {rho, ut + qn * nx, vt + qn * ny, pres};
  const T qnorm_inf = f.u * nx + f.v * ny;
  const bool sup_in = (qnorm_inf < T(0)) && (fabs(qnorm_inf) > sqrt(cinf2));
  const bool sup_out = (val(qni) > 0) && (val(qni) > val(ci));
  if (sup_in) wb = {rb, ub, vb, pb};
  if (sup_out) wb = wi;
  return wb;
}

}  // namespace a2d
