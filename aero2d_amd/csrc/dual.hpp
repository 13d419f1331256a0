// Forward-mode dual numbers for the exact tangent-linear residual.
//
// Dual<T> carries a value and one directional derivative. The flux code in
// physics.hpp is written once over a scalar type S; instantiating it with
// S = Dual<T> yields the Jacobian-vector product with the same branch choices
// as the primal (every data-dependent select compares values only), which is
// what jax.jvp gives the reference (aero2d/implicit.py:180).
//
// Ties in min/max split the derivative evenly between both arguments, the rule
// JAX uses for lax.min/lax.max.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

namespace a2d {

// Reciprocal and square root of the working precision. FP64 keeps IEEE
// division and square root (the outer residual, update and forces, and the
// whole FP64 form). FP32 is only the inner precision of the mixed form (the
// Jacobian-vector product and preconditioner blocks that feed the Krylov
// solve), where the hardware's single-instruction v_rcp_f32 / v_sqrt_f32
// (1 ulp) replace the ten-instruction correctly rounded sequences.
__host__ __device__ __forceinline__ double frcp(double x) { return 1.0 / x; }
__host__ __device__ __forceinline__ double fsqrt(double x) { return sqrt(x); }
__host__ __device__ __forceinline__ float frcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}
__host__ __device__ __forceinline__ float fsqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sqrtf(x);
#else
  return sqrtf(x);
#endif
}

template <class T>
struct Dual {
  T v, d;
  __host__ __device__ Dual() {}
  __host__ __device__ Dual(T x) : v(x), d(T(0)) {}
  __host__ __device__ Dual(T x, T dx) : v(x), d(dx) {}
};

template <class S> struct base_of { using type = S; };
template <class T> struct base_of<Dual<T>> { using type = T; };
template <class S> using base_t = typename base_of<S>::type;

// A constant of scalar type S (plain or dual) from a double literal.
template <class S> __host__ __device__ __forceinline__ S K(double x) { return S(base_t<S>(x)); }

template <class T> __host__ __device__ __forceinline__ T val(T x) { return x; }
template <class T> __host__ __device__ __forceinline__ T val(Dual<T> x) { return x.v; }
template <class T> __host__ __device__ __forceinline__ T tan_of(T) { return T(0); }
template <class T> __host__ __device__ __forceinline__ T tan_of(Dual<T> x) { return x.d; }

// ---- arithmetic -------------------------------------------------------------
template <class T> __host__ __device__ __forceinline__ Dual<T> operator+(Dual<T> a, Dual<T> b) { return {a.v + b.v, a.d + b.d}; }
template <class T> __host__ __device__ __forceinline__ Dual<T> operator-(Dual<T> a, Dual<T> b) { return {a.v - b.v, a.d - b.d}; }
template <class T> __host__ __device__ __forceinline__ Dual<T> operator*(Dual<T> a, Dual<T> b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
template <class T> __host__ __device__ __forceinline__ Dual<T> operator/(Dual<T> a, Dual<T> b) {
  const T r = frcp(b.v);
  const T q = a.v * r;
  return {q, (a.d - q * b.d) * r};
}
template <class T> __host__ __device__ __forceinline__ Dual<T> operator-(Dual<T> a) { return {-a.v, -a.d}; }

template <class T> __host__ __device__ __forceinline__ Dual<T> operator+(Dual<T> a, T b) { return {a.v + b, a.d}; }
template <class T> __host__ __device__ __forceinline__ Dual<T> operator+(T a, Dual<T> b) { return {a + b.v, b.d}; }
template <class T> __host__ __device__ __forceinline__ Dual<T> operator-(Dual<T> a, T b) { return {a.v - b, a.d}; }
template <class T> __host__ __device__ __forceinline__ Dual<T> operator-(T a, Dual<T> b) { return {a - b.v, -b.d}; }
template <class T> __host__ __device__ __forceinline__ Dual<T> operator*(Dual<T> a, T b) { return {a.v * b, a.d * b}; }
template <class T> __host__ __device__ __forceinline__ Dual<T> operator*(T a, Dual<T> b) { return {a * b.v, a * b.d}; }
template <class T> __host__ __device__ __forceinline__ Dual<T> operator/(Dual<T> a, T b) {
  const T r = frcp(b);
  return {a.v * r, a.d * r};
}
template <class T> __host__ __device__ __forceinline__ Dual<T> operator/(T a, Dual<T> b) {
  const T r = frcp(b.v);
  const T q = a * r;
  return {q, -q * b.d * r};
}

// Reciprocal 1/x (one division for value and tangent).
__host__ __device__ __forceinline__ double drcp(double x) { return 1.0 / x; }
__host__ __device__ __forceinline__ float drcp(float x) { return frcp(x); }
template <class T> __host__ __device__ __forceinline__ Dual<T> drcp(Dual<T> x) {
  const T r = frcp(x.v);
  return {r, -r * r * x.d};
}

// ---- elementary functions ---------------------------------------------------
__host__ __device__ __forceinline__ double dsqrt(double x) { return sqrt(x); }
__host__ __device__ __forceinline__ float dsqrt(float x) { return fsqrt(x); }
template <class T> __host__ __device__ __forceinline__ Dual<T> dsqrt(Dual<T> x) {
  T s = dsqrt(x.v);
  return {s, T(0.5) * x.d * frcp(s)};
}

__host__ __device__ __forceinline__ double dpow(double x, double a) { return pow(x, a); }
__host__ __device__ __forceinline__ float dpow(float x, double a) { return powf(x, float(a)); }
// d(x^a) = a x^(a-1) dx, written as a (x^a / x) dx.
template <class T> __host__ __device__ __forceinline__ Dual<T> dpow(Dual<T> x, double a) {
  T p = dpow(x.v, a);
  return {p, T(a) * (p / x.v) * x.d};
}

// min / max with JAX's balanced tie rule.
template <class T> __host__ __device__ __forceinline__ T dmin(T a, T b) { return a < b ? a : b; }
template <class T> __host__ __device__ __forceinline__ T dmax(T a, T b) { return a > b ? a : b; }
template <class T> __host__ __device__ __forceinline__ Dual<T> dmin(Dual<T> a, Dual<T> b) {
  T dd = a.v < b.v ? a.d : (b.v < a.v ? b.d : T(0.5) * (a.d + b.d));
  return {a.v < b.v ? a.v : b.v, dd};
}
template <class T> __host__ __device__ __forceinline__ Dual<T> dmax(Dual<T> a, Dual<T> b) {
  T dd = a.v > b.v ? a.d : (b.v > a.v ? b.d : T(0.5) * (a.d + b.d));
  return {a.v > b.v ? a.v : b.v, dd};
}
// max(x, c) for a constant c.
template <class T> __host__ __device__ __forceinline__ T dmaxc(T a, T c) { return a > c ? a : c; }
template <class T> __host__ __device__ __forceinline__ Dual<T> dmaxc(Dual<T> a, T c) {
  T dd = a.v > c ? a.d : (a.v < c ? T(0) : T(0.5) * a.d);
  return {a.v > c ? a.v : c, dd};
}

// Branch select on a value-only predicate.
template <class S> __host__ __device__ __forceinline__ S sel(bool c, S a, S b) { return c ? a : b; }

}  // namespace a2d
