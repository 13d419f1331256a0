// Block-tridiagonal wall-normal line preconditioner (aero2d/implicit.py:81-157).
//
//   k_assemble  first-order HLLC face Jacobians by forward-mode duals (one unit
//               tangent per conserved variable, what jax.jacfwd computes at
//               aero2d/implicit.py:36-59), assembled into the blocks D, Bim,
//               Bip, Bjm, Bjp of each cell plus vol/dtau on the diagonal
//               (aero2d/implicit.py:81-110). Thread per cell.
//   k_factor    block elimination along j for every i-line, two-sided: the
//               reference's top-down Thomas recurrence (aero2d/implicit.py:112-128)
//               on the wall half, the mirror recurrence on the far-field half,
//               joined at the middle row. Eight lanes per line (a quad per
//               half, lane r = row r of every 4x4 block): products by DPP
//               broadcasts in the quad, the inverse by Gauss-Jordan with
//               partial pivoting across the quad.
//               The same pass also stores S0 Bim and S0 Bip (S0 = inverse of
//               the eliminated diagonal) over Bim and Bip, so that a sweep
//               forms g = S0 rhs without a second dependent product.
//   k_sweep     one line-Jacobi sweep (aero2d/implicit.py:130-157) or one
//               red-black half sweep in a single pass: the right-hand side
//               with the neighbour-line coupling, and the two-sided block
//               recurrences of every line (eight lanes per line, lane r owns
//               row r of each 4x4 product, components exchanged in the quad).
//   The reference's forward step d_j = inv_j (rhs_j - A_j d_{j-1}) is
//   regrouped as d_j = g_j - P_j d_{j-1} with g_j = inv_j rhs_j and
//   P_j = inv_j A_j: the same direct solve as the reference in another
//   elimination order, so the same answer up to rounding.
//
// Block storage: T blk[5][pstride][4][4] = slot, cell, block row, column
// (cell-major): a block row is one 16-byte load in FP32 (two in FP64) and the
// four rows a quad of lanes loads for one cell are one contiguous 64-byte
// run. Slots after assembly: 0 = D, 1 = Bjp, 2 = Bjm, 3 = Bim, 4 = Bip; after
// factorisation 0 = S0, 1 = backward coupling (c or e), 2 = forward coupling
// (P or Q), 3 = S0 Bim, 4 = S0 Bip. With the red-black line order the cell
// index is colour-split (column of line i = (i % 2) ni/2 + i/2).
// Mixed form: the factorisation (in FP32) also writes slots 1-4 to an FP16
// copy, _Float16 hb[4][pstride][4][4] (slot - 1, cell, row, column); with the
// knob precond_half the sweeps read their couplings from it (an 8-byte load
// per block row, half the bytes of the sweeps' dominant traffic). The
// couplings are products with the inverse of the eliminated diagonal, entries
// of order one; conversion saturates at +-65504 and keeps NaN, so a factor
// that failed in FP32 still fails. Slot 0 (S0, unbounded) stays FP32.
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

// Row-wise access to the block storage (layout in the file header).
template <class T>
struct Blk {
  gptr<T> p;
  size_t ps;
  // 32-bit element index from the (wave-uniform) base: the compiler can then
  // address with a scalar base and one 32-bit vector offset
  __device__ __forceinline__ gptr<T> row(int slot, int r, unsigned cell) const {
    return at(p, ((unsigned)slot * (unsigned)ps + cell) * 16u + (unsigned)r * 4u);
  }
};

typedef float a2d_f4 __attribute__((ext_vector_type(4)));
typedef double a2d_d2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void load_row(gptr<const float> p, float v[4]) {
  const a2d_f4 t = *(gptr<const a2d_f4>)p;
  v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
}
__device__ __forceinline__ void load_row(gptr<const double> p, double v[4]) {
  const a2d_d2 a = *(gptr<const a2d_d2>)p;
  const a2d_d2 b = *(gptr<const a2d_d2>)(p + 2);
  v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
__device__ __forceinline__ void store_row(gptr<float> p, const float v[4]) {
  a2d_f4 t;
  t.x = v[0]; t.y = v[1]; t.z = v[2]; t.w = v[3];
  *(gptr<a2d_f4>)p = t;
}
__device__ __forceinline__ void store_row(gptr<double> p, const double v[4]) {
  a2d_d2 a, b;
  a.x = v[0]; a.y = v[1]; b.x = v[2]; b.y = v[3];
  *(gptr<a2d_d2>)p = a;
  *(gptr<a2d_d2>)(p + 2) = b;
}

typedef _Float16 a2d_h4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ _Float16 to_half_sat(float v) {
  const float c = fminf(fmaxf(v, -65504.f), 65504.f);
  return (_Float16)(v == v ? c : v);
}
// Row r of slot `slot` (1-4) of the FP16 coupling copy.
__device__ __forceinline__ gptr<const _Float16> hrow(gptr<const _Float16> H, size_t ps, int slot,
                                                     int r, unsigned cell) {
  return at(H, ((unsigned)(slot - 1) * (unsigned)ps + cell) * 16u + (unsigned)r * 4u);
}

template <class T>
__device__ __forceinline__ void store_blk(const Blk<T>& B, int slot, size_t cell, const T a[16]) {
#pragma unroll
  for (int r = 0; r < 4; ++r) store_row(B.row(slot, r, cell), a + 4 * r);
}

template <class T>
__global__ void __launch_bounds__(NT) k_assemble(Plan plan, KnobsDev kn) {
  const int e = plan_entry(plan, blockIdx.x);
  const CaseDev& cd = plan.cases[plan.slots[e]];
  if (cd.sc->parked) return;
  const int ni = cd.ni, nc = cd.nc, n = cd.n;
  const int idx = (blockIdx.x - plan.blk_prefix[e]) * NT + threadIdx.x;
  if (idx >= n) return;
  const int i = idx % ni, j = idx / ni;
  const int im = (i == 0) ? ni - 1 : i - 1, ip = (i + 1 == ni) ? 0 : i + 1;
  const gptr<const T> U = gp(base_U<T>(cd));
  const gptr<const T> Si = gp(met_Si<T>(cd));
  const gptr<const T> Sj = gp(met_Sj<T>(cd));
  const size_t sj1 = (size_t)(nc + 1) * ni;
  const size_t bidx = (size_t)j * ni + bcol(i, ni, kn.precond_rb);
  const Blk<T> B{gp(tp<T>(cd.blk)), (size_t)cd.pstride};

  T qc[4], qim[4], qip[4];
  load_cons(U, n, idx, qc);
  load_cons(U, n, j * ni + im, qim);
  load_cons(U, n, j * ni + ip, qip);
  T Dm[16], Ba[16], Bb[16];
  const T dg = gp(ctp<T>(cd.diag))[idx];
#pragma unroll
  for (int m = 0; m < 16; ++m) Dm[m] = (m % 5 == 0) ? dg : T(0);

  // i-face i (cells im | i) and i-face i+1 (cells i | ip)
  const T six0 = Si[j * ni + i], siy0 = Si[n + j * ni + i];
  const T six1 = Si[j * ni + ip], siy1 = Si[n + j * ni + ip];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    T col[4];
    face_col<T>(qim, qc, six0, siy0, b, false, col);   // AL_i -> Bim = -AL_i
#pragma unroll
    for (int r = 0; r < 4; ++r) Ba[r * 4 + b] = -col[r];
    face_col<T>(qim, qc, six0, siy0, b, true, col);    // AR_i -> D -= AR_i
#pragma unroll
    for (int r = 0; r < 4; ++r) Dm[r * 4 + b] -= col[r];
    face_col<T>(qc, qip, six1, siy1, b, false, col);   // AL_{i+1} -> D += AL_{i+1}
#pragma unroll
    for (int r = 0; r < 4; ++r) Dm[r * 4 + b] += col[r];
    face_col<T>(qc, qip, six1, siy1, b, true, col);    // AR_{i+1} -> Bip
#pragma unroll
    for (int r = 0; r < 4; ++r) Bb[r * 4 + b] = col[r];
  }
  store_blk(B, 3, bidx, Ba);
  store_blk(B, 4, bidx, Bb);
  // j-face j (cells j-1 | j) or the wall
  const T sjx0 = Sj[j * ni + i], sjy0 = Sj[sj1 + j * ni + i];
  if (j >= 1) {
    T qjm[4];
    load_cons(U, n, idx - ni, qjm);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      T col[4];
      face_col<T>(qjm, qc, sjx0, sjy0, b, false, col);  // AL_j -> Bjm = -AL_j
#pragma unroll
      for (int r = 0; r < 4; ++r) Ba[r * 4 + b] = -col[r];
      face_col<T>(qjm, qc, sjx0, sjy0, b, true, col);   // AR_j -> D -= AR_j
#pragma unroll
      for (int r = 0; r < 4; ++r) Dm[r * 4 + b] -= col[r];
    }
  } else {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      T col[4];
      wall_col<T>(qc, sjx0, sjy0, b, col);              // D -= Aw
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        Dm[r * 4 + b] -= col[r];
        Ba[r * 4 + b] = T(0);
      }
    }
  }
  store_blk(B, 2, bidx, Ba);
  // j-face j+1 (cells j | j+1) or the far field
  const T sjx1 = Sj[(j + 1) * ni + i], sjy1 = Sj[sj1 + (j + 1) * ni + i];
  if (j + 1 < nc) {
    T qjp[4];
    load_cons(U, n, idx + ni, qjp);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      T col[4];
      face_col<T>(qc, qjp, sjx1, sjy1, b, false, col);  // AL_{j+1} -> D += AL_{j+1}
#pragma unroll
      for (int r = 0; r < 4; ++r) Dm[r * 4 + b] += col[r];
      face_col<T>(qc, qjp, sjx1, sjy1, b, true, col);   // AR_{j+1} -> Bjp
#pragma unroll
      for (int r = 0; r < 4; ++r) Bb[r * 4 + b] = col[r];
    }
  } else {
    const CaseScal* sc = cd.sc;
    const Inf<T> inf = {T(sc->winf[0]), T(sc->winf[1]), T(sc->winf[2]), T(sc->winf[3])};
    const T circ = kn.vortex ? T(sc->circ) : T(0);
    const T xf = met_xf<T>(cd)[i], yf = met_yf<T>(cd)[i];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      T col[4];
      far_col<T>(qc, sjx1, sjy1, inf, xf, yf, circ, b, col);  // D += Af
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        Dm[r * 4 + b] += col[r];
        Bb[r * 4 + b] = T(0);
      }
    }
  }
  store_blk(B, 1, bidx, Bb);
  store_blk(B, 0, bidx, Dm);
}

// Two-sided (twisted) block elimination of every j-line. With m = nc / 2, the
// top half eliminates j = 0 .. m-1 downward from the wall exactly as
// aero2d/implicit.py:112-128 does (den = D - A c_prev, inv = den^-1,
// c = inv C, P = inv A); the bottom half eliminates j = nc-1 .. m+1 upward
// from the far field (den = D - C e_prev, inv = den^-1, e = inv A, Q = inv C);
// row m closes both: M = D - A c_{m-1} - C e_{m+1}, slot 0 = M^-1,
// slot 2 = M^-1 A, slot 1 = M^-1 C. In every row slot 0 maps the right-hand
// side to g, slot 2 multiplies the recurrence value arriving from the line end
// and slot 1 the one arriving from the middle on the way back. The same direct
// solve as the reference's with half the chain length.
//
// Eight lanes per line, as in the sweeps: lanes 0-3 the top half, 4-7 the
// bottom, lane r owning row r of every 4x4 block. A product's row r takes the
// other rows by DPP broadcast inside the quad, and the inverse is Gauss-Jordan
// with partial pivoting across the quad (the pivot row is fetched from its lane
// by a lane shuffle), the same operations in the same order as a one-thread
// inversion. The rows of the next steps' five blocks load while a step computes.

// quad_bcast with the source lane as a run-time value that unrolling makes a
// constant.
template <class T>
__device__ __forceinline__ T qbc(int c, T x) {
  switch (c) {
    case 0: return quad_bcast<0>(x);
    case 1: return quad_bcast<1>(x);
    case 2: return quad_bcast<2>(x);
    default: return quad_bcast<3>(x);
  }
}
// Row r (this lane's) of X Y, Y's rows held by the quad's lanes.
template <class T>
__device__ __forceinline__ void qmul(const T x[4], const T y[4], T out[4]) {
#pragma unroll
  for (int c = 0; c < 4; ++c)
    out[c] = x[0] * quad_bcast<0>(y[c]) + x[1] * quad_bcast<1>(y[c]) + x[2] * quad_bcast<2>(y[c]) +
             x[3] * quad_bcast<3>(y[c]);
}
// Row r of A^-1 (A's rows held by the quad's lanes): Gauss-Jordan, pivot =
// first row of largest magnitude in the column among the rows not yet used.
template <class T>
__device__ __forceinline__ void qinv(const T a_in[4], int r, T out[4]) {
  T a[4], b[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    a[k] = a_in[k];
    b[k] = (k == r) ? T(1) : T(0);
  }
  const int base = (int)(threadIdx.x & ~3u);
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const T v0 = quad_bcast<0>(fabs(a[c])), v1 = quad_bcast<1>(fabs(a[c])),
            v2 = quad_bcast<2>(fabs(a[c])), v3 = quad_bcast<3>(fabs(a[c]));
    const T vv[4] = {v0, v1, v2, v3};
    int p = c;
    T best = vv[c];
#pragma unroll
    for (int q = c + 1; q < 4; ++q)
      if (vv[q] > best) { best = vv[q]; p = q; }
    // swap rows c and p: lane c takes row p, lane p takes row c
    T ra[4], rb[4], ca[4], cb[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      ra[k] = __shfl(a[k], base + p, WAVE);
      rb[k] = __shfl(b[k], base + p, WAVE);
      ca[k] = qbc<T>(c, a[k]);
      cb[k] = qbc<T>(c, b[k]);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      a[k] = (r == c) ? ra[k] : ((r == p) ? ca[k] : a[k]);
      b[k] = (r == c) ? rb[k] : ((r == p) ? cb[k] : b[k]);
    }
    const T piv = T(1) / qbc<T>(c, a[c]);
    T pa[4], pb[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      pa[k] = qbc<T>(c, a[k]) * piv;
      pb[k] = qbc<T>(c, b[k]) * piv;
    }
    const T f = a[c];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      a[k] = (r == c) ? pa[k] : a[k] - f * pa[k];
      b[k] = (r == c) ? pb[k] : b[k] - f * pb[k];
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) out[k] = b[k];
}

template <class T>
__global__ void __launch_bounds__(8 * LINES_PER_BLOCK) k_factor(Plan plan) {
  const int e = plan_entry(plan, blockIdx.x);
  const CaseDev& cd = plan.cases[plan.slots[e]];
  if (cd.sc->parked) return;
  const int ni = cd.ni, nc = cd.nc;
  const int r = threadIdx.x & 3;
  const int half = (threadIdx.x >> 2) & 1;
  const int q = (blockIdx.x - plan.blk_prefix[e]) * LINES_PER_BLOCK + (threadIdx.x >> 3);
  const bool live = q < ni;
  const int col = live ? q : ni - 1;   // dead lanes shadow a real line and never store
  const int m = nc >> 1;
  const int steps = half ? nc - 1 - m : m;
  const int j0 = half ? nc - 1 : 0, dj = half ? -1 : 1;
  const Blk<T> B{gp(uniform_ptr(tp<T>(cd.blk))), (size_t)cd.pstride};
  const Blk<const T> Bc{gp(uniform_ptr(ctp<T>(cd.blk))), (size_t)cd.pstride};
  const gptr<_Float16> hbw = gp(uniform_ptr((_Float16*)cd.hb));
  // row r of a block: T storage, and for slots 1-4 the FP16 copy (mixed form)
  auto keep = [&](int slot, size_t cell, const T v[4]) {
    store_row(B.row(slot, r, cell), v);
    if constexpr (std::is_same<T, float>::value) {
      if (slot > 0) {
        a2d_h4 t;
        t.x = to_half_sat(v[0]); t.y = to_half_sat(v[1]);
        t.z = to_half_sat(v[2]); t.w = to_half_sat(v[3]);
        *(gptr<a2d_h4>)at(hbw, ((unsigned)(slot - 1) * (unsigned)B.ps + (unsigned)cell) * 16u + (unsigned)r * 4u) = t;
      }
    }
  };
  T cprev[4] = {T(0), T(0), T(0), T(0)};
  constexpr int FPF = 2;
  T rb[FPF][5][4];
  auto fload = [&](int k, int s) {
    const size_t cell = (size_t)(j0 + dj * max(0, min(s, steps - 1))) * ni + col;
#pragma unroll
    for (int b = 0; b < 5; ++b) load_row(Bc.row(b, r, cell), rb[k][b]);
  };
#pragma unroll
  for (int k = 0; k < FPF; ++k) fload(k, k);
  for (int s0 = 0; s0 < m; s0 += FPF) {
#pragma unroll
    for (int k = 0; k < FPF; ++k) {
      const int s = s0 + k;
      const bool valid = s < steps;
      const size_t cell = (size_t)(j0 + dj * max(0, min(s, steps - 1))) * ni + col;
      T D[4], C[4], A[4], Bm[4], Bp[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        D[i] = rb[k][0][i];
        C[i] = rb[k][1][i];
        A[i] = rb[k][2][i];
        Bm[i] = rb[k][3][i];
        Bp[i] = rb[k][4][i];
      }
      fload(k, s + FPF);
      T L[4], F[4], t[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        L[i] = half ? C[i] : A[i];    // couples to the row already eliminated
        F[i] = half ? A[i] : C[i];    // couples to the row eliminated next
      }
      qmul<T>(L, cprev, t);
#pragma unroll
      for (int i = 0; i < 4; ++i) D[i] -= t[i];
      T inv[4], P[4], cn[4];
      qinv<T>(D, r, inv);
      qmul<T>(inv, F, cn);
      qmul<T>(inv, L, P);
      T s3[4], s4[4];
      qmul<T>(inv, Bm, s3);
      qmul<T>(inv, Bp, s4);
      if (live && valid) {
        keep(0, cell, inv);
        keep(1, cell, cn);
        keep(2, cell, P);
        keep(3, cell, s3);
        keep(4, cell, s4);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) cprev[i] = valid ? cn[i] : cprev[i];
    }
  }
  // ---- row m: the top's c_{m-1} and the bottom's e_{m+1} --------------------
  T other[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) other[i] = __shfl_xor(cprev[i], 4, WAVE);
  {
    const size_t cell = (size_t)m * ni + col;
    T D[4], A[4], C[4], Bm[4], Bp[4], t[4];
    load_row(Bc.row(0, r, cell), D);
    load_row(Bc.row(2, r, cell), A);
    load_row(Bc.row(1, r, cell), C);
    load_row(Bc.row(3, r, cell), Bm);
    load_row(Bc.row(4, r, cell), Bp);
    const T* ctop = half ? other : cprev;   // c_{m-1}
    const T* ebot = half ? cprev : other;   // e_{m+1} (zero when the bottom part is empty)
    qmul<T>(A, ctop, t);
#pragma unroll
    for (int i = 0; i < 4; ++i) D[i] -= t[i];
    qmul<T>(C, ebot, t);
#pragma unroll
    for (int i = 0; i < 4; ++i) D[i] -= t[i];
    T inv[4], P[4], Q[4], s3[4], s4[4];
    qinv<T>(D, r, inv);
    qmul<T>(inv, A, P);
    qmul<T>(inv, C, Q);
    qmul<T>(inv, Bm, s3);
    qmul<T>(inv, Bp, s4);
    if (half == 0 && live) {
      keep(0, cell, inv);
      keep(1, cell, Q);
      keep(2, cell, P);
      keep(3, cell, s3);
      keep(4, cell, s4);
    }
  }
}

// A block row as loaded: four T, or four FP16 kept packed until the step that
// consumes them (a conversion placed next to its load would wait for the load).
template <class T> struct RowT { T v[4]; };
template <class T, bool HB> struct CplRow { using type = RowT<T>; };
template <> struct CplRow<float, true> { using type = a2d_h4; };
template <class T>
__device__ __forceinline__ void unpack(const RowT<T>& a, T v[4]) {
  v[0] = a.v[0]; v[1] = a.v[1]; v[2] = a.v[2]; v[3] = a.v[3];
}
__device__ __forceinline__ void unpack(const a2d_h4& a, float v[4]) {
  v[0] = float(a.x); v[1] = float(a.y); v[2] = float(a.z); v[3] = float(a.w);
}

// One line sweep in a single pass. For every line of the sweep (every line
// when color < 0, line Jacobi; lines i = color mod 2 of a red-black half
// sweep otherwise):
//   x = L^-1 (src - Bim x(i-1) - Bip x(i+1))      (neighbour terms when COUP)
// with the factored blocks (slot 0 = S0, the inverse of the eliminated
// diagonal; slots 3, 4 = S0 Bim, S0 Bip; slot 2 = the forward coupling; slot 1
// the backward one):
//   g_j = S0_j src_j - (S0 Bim)_j xm_j - (S0 Bip)_j xp_j
//   d_j = g_j - S2_j d_prev                      toward the middle row m
//   x_m = g_m - S2_m d_top - S1_m f_bot
//   x_j = d_j - S1_j x_prev                      back out from the middle.
// Eight lanes per line: lanes 0-3 (one per block row) own the wall half,
// lanes 4-7 the far-field half. Each lane loads row r of each block (one
// 16-byte load in FP32; a quad's four rows are one contiguous 64-byte run in
// the cell-major block storage) and component r of each vector, and the quad
// exchanges components by DPP. g is formed from loads already in flight, off
// the recurrence chain. The loads of PF steps (PFB on the way back) are in
// flight at any time: a register ring whose slot is refilled with the step
// PF further on as soon as it is consumed. Loads are unconditional (rows past
// the line's end are clamped to its last row, an L2 hit): a load under a
// branch makes the compiler wait for every load in flight where the branch
// joins. The d values of the way in are kept for the way back: in LDS when
// the plan grants it (DL: dsteps slots per lane, [step][thread], each lane
// reading back only what it wrote), else in the case's d buffer ([cell][4]),
// an HBM round trip of 8 words per cell.
// Source of g: GS_VEC / GS_NEGR form S0 src from a vector / from -R (read in
// FP64) and keep it in the case's g0 buffer; GS_G0 reads it back. Every sweep
// of one preconditioner application has the same source, so only the first
// sweep over a line reads src and S0.
// HB (FP32 only): the couplings (slots 1-4) come from the FP16 copy.
enum GSrc : int { GS_VEC = 0, GS_NEGR = 1, GS_G0 = 2 };

template <class T, bool COUP, int GS, int PF, int PFB, bool HB, int SL, bool DL>
__global__ void __launch_bounds__(8 * SL) k_sweep(Plan plan, int src, int xnb, int xout, int want_norm,
                                               int gate, int color) {
  extern __shared__ unsigned char sweep_lds[];
  T* const sd = reinterpret_cast<T*>(sweep_lds);
  const int e = plan_entry(plan, blockIdx.x);
  const CaseDev& cd = plan.cases[plan.slots[e]];
  if (cd.sc->parked || (gate && !cd.sc->active)) return;
  const int ni = cd.ni, nc = cd.nc, n = cd.n;
  const int lb = blockIdx.x - plan.blk_prefix[e];
  // color < 0: every line, natural block layout; else line q of colour
  // `color` (i = 2q + color), whose blocks sit in column color * ni/2 + q
  // (the plan then holds nlinerb blocks per case: the colour's lines only)
  const int nq = color < 0 ? ni : (ni >> 1);
  if (lb * SL >= nq) return;   // guard: plans hold no such block
  const int r = threadIdx.x & 3;
  const int half = (threadIdx.x >> 2) & 1;
  const int q = lb * SL + (threadIdx.x >> 3);
  const bool live = q < nq;
  const int qc = live ? q : nq - 1;     // dead lanes shadow a real line and never store
  const int ic = color < 0 ? qc : 2 * qc + color;
  const int col = color < 0 ? qc : color * (ni >> 1) + qc;
  const int im = (ic == 0) ? ni - 1 : ic - 1, ip = (ic + 1 == ni) ? 0 : ic + 1;
  const int m = nc >> 1;
  const int steps = half ? nc - 1 - m : m;
  const int j0 = half ? nc - 1 : 0, dj = half ? -1 : 1;
  const Blk<const T> B{gp(uniform_ptr(ctp<T>(cd.blk))), (size_t)cd.pstride};
  // Every array is addressed as its case's (wave-uniform) base plus a 32-bit
  // element index, so that an access costs one 32-bit offset rather than
  // 64-bit address arithmetic. Component r of a vector at row j and line i:
  // r n + j ni + i; g0 and d are [cell][4].
  const unsigned rn = (unsigned)r * (unsigned)n, niu = (unsigned)ni, colu = (unsigned)col;
  const gptr<T> gd = gp(uniform_ptr(tp<T>(cd.d)));
  const gptr<T> xv = gp(uniform_ptr(vec_ptr<T>(cd, xout)));
  const gptr<const double> sR = gp(uniform_ptr((const double*)cd.R));
  const gptr<const T> sV = gp(uniform_ptr((const T*)vec_ptr<T>(cd, GS == GS_VEC ? src : VEC_W)));
  const gptr<const T> xnbp = gp(uniform_ptr((const T*)vec_ptr<T>(cd, COUP ? xnb : VEC_W)));
  const gptr<T> g0p = gp(uniform_ptr(tp<T>(cd.g0)));
  const gptr<const _Float16> hbp = gp(uniform_ptr((const _Float16*)cd.hb));
  // Per-lane constant parts of the element indices; a step adds its row's
  // first cell (rowc = j ni) to each: g0 and d rowc 4 + cgd, neighbours
  // rowc + cxm / cxp, the source rowc + csrc, block row (slot, r) of the
  // step's cell rowc 16 + cblk[slot] (T storage) or + chb[slot] (FP16 copy).
  const unsigned psu = (unsigned)B.ps, ru = (unsigned)r;
  const unsigned cgd = colu * 4u + ru;
  // vectors are colour-split in the red-black order (column bcol), so a half
  // sweep's own lines sit at its block column and the neighbours in the other
  // colour's contiguous run; -R keeps the natural order
  const int split = color >= 0;
  const unsigned cxm = rn + (unsigned)bcol(im, ni, split), cxp = rn + (unsigned)bcol(ip, ni, split);
  const unsigned csrc = rn + (unsigned)(split ? col : ic), csrcR = rn + (unsigned)ic;
  auto cblk = [&](int slot) -> unsigned { return ((unsigned)slot * psu + colu) * 16u + ru * 4u; };
  auto chb = [&](int slot) -> unsigned { return ((unsigned)(slot - 1) * psu + colu) * 16u + ru * 4u; };
  // row r of block `slot` of the cell in row rowc, from the T storage or
  // (slots 1-4, HB) the FP16 copy
  auto brow = [&](int slot, unsigned rowc) { return at(B.p, rowc * 16u + cblk(slot)); };
  // row r of coupling slot `slot` (1-4) of a cell, as loaded (CR) or as T
  using CR = typename CplRow<T, HB>::type;
  auto crow = [&](int slot, unsigned cell, CR& a) {
    if constexpr (HB) a = *(gptr<const a2d_h4>)hrow(hbp, B.ps, slot, r, cell);
    else load_row(B.row(slot, r, cell), a.v);
  };
  auto crow_t = [&](int slot, unsigned cell, T v[4]) {
    CR a;
    crow(slot, cell, a);
    unpack(a, v);
  };
  // the source as loaded (-R stays FP64 until consumed) and as T
  using SR = typename std::conditional<GS == GS_NEGR, double, T>::type;
  auto load_src = [&](int jj) -> SR {
    if constexpr (GS == GS_NEGR) return *at(sR, (unsigned)jj * niu + csrcR);
    else return *at(sV, (unsigned)jj * niu + csrc);
  };
  auto src_val = [&](SR v) -> T {
    if constexpr (GS == GS_NEGR) return T(-v);
    else return v;
  };
  // S0 src (row r) from component r of src and row r of S0
  auto s0src = [&](T sv, const T* s0) -> T {
    return s0[0] * quad_bcast<0>(sv) + s0[1] * quad_bcast<1>(sv) + s0[2] * quad_bcast<2>(sv) +
           s0[3] * quad_bcast<3>(sv);
  };
  // g = S0 src - (S0 Bim) xm - (S0 Bip) xp, with gs = S0 src
  auto gform = [&](T gs, T xmv, T xpv, const T* s3, const T* s4) -> T {
    T g = gs;
    if constexpr (COUP) {
      g -= s3[0] * quad_bcast<0>(xmv) + s3[1] * quad_bcast<1>(xmv) + s3[2] * quad_bcast<2>(xmv) +
           s3[3] * quad_bcast<3>(xmv);
      g -= s4[0] * quad_bcast<0>(xpv) + s4[1] * quad_bcast<1>(xpv) + s4[2] * quad_bcast<2>(xpv) +
           s4[3] * quad_bcast<3>(xpv);
    }
    return g;
  };

  // ---- toward the middle -------------------------------------------------------
  // ring of PF steps in flight: step s lives in slot s % PF; consuming a slot
  // issues the loads of the step PF further on into it
  SR bs[PF];
  T bxm[PF], bxp[PF], b0[PF][4];
  CR b3[PF], b4[PF], b2[PF];
  auto frow = [&](int s) -> int { return j0 + dj * max(0, min(s, steps - 1)); };
  unsigned rowr[PF];   // each slot's row offset, for the stores of its step
  auto cload = [&](int slot, unsigned rowc, CR& a) {
    if constexpr (HB) a = *(gptr<const a2d_h4>)at(hbp, rowc * 16u + chb(slot));
    else load_row(brow(slot, rowc), a.v);
  };
  auto fload = [&](int k, int s) {
    const unsigned rowc = (unsigned)frow(s) * niu;
    rowr[k] = rowc;
    if constexpr (GS == GS_G0) {
      bs[k] = *at(g0p, rowc * 4u + cgd);
    } else {
      if constexpr (GS == GS_NEGR) bs[k] = *at(sR, rowc + csrcR);
      else bs[k] = *at(sV, rowc + csrc);
      load_row(brow(0, rowc), b0[k]);
    }
    if constexpr (COUP) {
      bxm[k] = *at(xnbp, rowc + cxm);
      bxp[k] = *at(xnbp, rowc + cxp);
      cload(3, rowc, b3[k]);
      cload(4, rowc, b4[k]);
    }
    cload(2, rowc, b2[k]);
  };
  T dr = T(0);
  // Fill the ring in slot order (the scheduling barriers keep slot 0's loads
  // oldest, so that the first wait in the loop can leave the later slots in
  // flight). Every pass then consumes every slot, and the loop has no per-lane
  // branch: a step past the line's half computes a value that the select
  // drops and stores, to the half's last row (where the clamped loads read),
  // the value that row already holds; dead lanes store what the line they
  // shadow stores. (Around a per-lane branch the compiler reuses consumed
  // slots as temporaries and then waits for nearly every load in flight.)
#pragma unroll
  for (int k = 0; k < PF; ++k) {
    fload(k, k);
    __builtin_amdgcn_sched_barrier(0);
  }
  for (int s0 = 0; s0 < m; s0 += PF) {
#pragma unroll
    for (int k = 0; k < PF; ++k) {
      const int st = s0 + k;
      const bool valid = st < steps;
      const unsigned rowc = rowr[k];
      T gs;
      if constexpr (GS == GS_G0) {
        gs = bs[k];
      } else {
        gs = s0src(src_val(bs[k]), b0[k]);
        *at(g0p, rowc * 4u + cgd) = gs;
      }
      T p2[4], p3[4], p4[4];
      if constexpr (COUP) {
        unpack(b3[k], p3);
        unpack(b4[k], p4);
      }
      unpack(b2[k], p2);
      const T g = gform(gs, bxm[k], bxp[k], p3, p4);
      const T dn = g - p2[0] * quad_bcast<0>(dr) - p2[1] * quad_bcast<1>(dr) -
                   p2[2] * quad_bcast<2>(dr) - p2[3] * quad_bcast<3>(dr);
      dr = valid ? dn : dr;
      if constexpr (DL) sd[(unsigned)max(0, min(st, steps - 1)) * (8u * SL) + threadIdx.x] = dr;
      else *at(gd, rowc * 4u + cgd) = dr;
      fload(k, st + PF);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  // ---- the middle row: the top's d_{m-1} and the bottom's f_{m+1} meet ---------
  const T other = __shfl_xor(dr, 4, WAVE);
  const T dtop = half ? other : dr, fbot = half ? dr : other;
  T xr;
  {
    const unsigned cell = (unsigned)m * niu + colu;
    T s0r[4], s3r[4], s4r[4], pm[4], qm[4];
    T xmv = T(0), xpv = T(0), gs;
    if constexpr (GS == GS_G0) {
      gs = *at(g0p, cell * 4u + (unsigned)r);
    } else {
      const T sv = src_val(load_src(m));
      load_row(B.row(0, r, cell), s0r);
      gs = s0src(sv, s0r);
      if (live && half == 0) *at(g0p, cell * 4u + (unsigned)r) = gs;
    }
    if constexpr (COUP) {
      xmv = *at(xnbp, (unsigned)m * niu + cxm);
      xpv = *at(xnbp, (unsigned)m * niu + cxp);
      crow_t(3, cell, s3r);
      crow_t(4, cell, s4r);
    }
    crow_t(2, cell, pm);
    crow_t(1, cell, qm);
    const T g = gform(gs, xmv, xpv, s3r, s4r);
    xr = g - pm[0] * quad_bcast<0>(dtop) - pm[1] * quad_bcast<1>(dtop) - pm[2] * quad_bcast<2>(dtop) -
         pm[3] * quad_bcast<3>(dtop) - qm[0] * quad_bcast<0>(fbot) - qm[1] * quad_bcast<1>(fbot) -
         qm[2] * quad_bcast<2>(fbot) - qm[3] * quad_bcast<3>(fbot);
  }
  double nrm = 0.0;
  if (live && half == 0) {
    *at(xv, (unsigned)m * niu + csrc) = xr;
    nrm += double(xr) * double(xr);
  }
  // ---- away from the middle: x_j = d_j - S1_j x_prev (ring of PFB steps) -------
  T cv[PFB];
  CR cm[PFB];
  auto bj = [&](int s) -> int {
    const int sc = max(0, min(s, steps - 1));
    return steps > 0 ? m - dj * (1 + sc) : m;
  };
  unsigned rowb[PFB];   // each slot's row offset, for the store of its step
  auto bload = [&](int k, int s) {
    const unsigned rowc = (unsigned)bj(s) * niu;
    rowb[k] = rowc;
    if constexpr (DL) {
      // the way in stored step t of this half in slot t; row bj(s) is step steps - 1 - s
      const int sc = max(0, min(s, steps - 1));
      cv[k] = sd[(unsigned)max(0, steps - 1 - sc) * (8u * SL) + threadIdx.x];
    } else {
      cv[k] = *at(gd, rowc * 4u + cgd);
    }
    cload(1, rowc, cm[k]);
  };
  // no per-lane branch, as on the way in: a step past the half's end keeps xr
  // and stores to the half's last row the value it already holds; dead lanes
  // store what the line they shadow stores
#pragma unroll
  for (int k = 0; k < PFB; ++k) bload(k, k);
  for (int s0 = 0; s0 < m; s0 += PFB) {
#pragma unroll
    for (int k = 0; k < PFB; ++k) {
      const bool valid = s0 + k < steps;
      T c1[4];
      unpack(cm[k], c1);
      const T xn = cv[k] - c1[0] * quad_bcast<0>(xr) - c1[1] * quad_bcast<1>(xr) -
                   c1[2] * quad_bcast<2>(xr) - c1[3] * quad_bcast<3>(xr);
      xr = valid ? xn : xr;
      *at(xv, rowb[k] + csrc) = xr;
      nrm += (live && valid) ? double(xr) * double(xr) : 0.0;
      bload(k, s0 + k + PFB);
    }
  }
  if (want_norm) {
    // one partial per wavefront (8 lines), whatever the block size
    nrm = wave_sum(nrm);
    const int grp = lb * (SL / 8) + (int)(threadIdx.x / WAVE);
    if ((threadIdx.x & (WAVE - 1)) == 0 && grp * 8 < nq) {
      const int slot = color < 0 ? grp : color * cd.nlinerw + grp;
      part_at(cd, slot, 0) = nrm;
    }
  }
}

// Steps in flight on the way in (PF) and back out (PFB); the way in loads up
// to seven rows and components per step, the way back two.
template <class T> struct SweepDepth;
template <> struct SweepDepth<float> { static constexpr int in = 8, out = 16; };
template <> struct SweepDepth<double> { static constexpr int in = 4, out = 8; };

// Kernel variant for a sweep: coupled or not; g from src (a vector or -R,
// kept in g0) or from g0 (reuse_g0: a later sweep of the same application);
// couplings from the FP16 copy (hb).
template <class T, bool HB, int SL, bool DL>
void run_sweep_d(const Plan& pl, int nblocks, int src, int xnb, int xout, int coupled, int reuse_g0,
                 int want_norm, int gate, int color, hipStream_t s) {
  constexpr int PF = SweepDepth<T>::in, PFB = SweepDepth<T>::out;
  const dim3 grid(nblocks), block(8 * SL);
  const size_t lds = DL ? size_t(8 * SL) * size_t(pl.dsteps) * sizeof(T) : 0;
  if (coupled && reuse_g0)
    hipLaunchKernelGGL((k_sweep<T, true, GS_G0, PF, PFB, HB, SL, DL>), grid, block, lds, s, pl, src, xnb,
                       xout, want_norm, gate, color);
  else if (coupled && src == VEC_NEG_R)
    hipLaunchKernelGGL((k_sweep<T, true, GS_NEGR, PF, PFB, HB, SL, DL>), grid, block, lds, s, pl, src, xnb,
                       xout, want_norm, gate, color);
  else if (coupled)
    hipLaunchKernelGGL((k_sweep<T, true, GS_VEC, PF, PFB, HB, SL, DL>), grid, block, lds, s, pl, src, xnb,
                       xout, want_norm, gate, color);
  else if (src == VEC_NEG_R)
    hipLaunchKernelGGL((k_sweep<T, false, GS_NEGR, PF, PFB, HB, SL, DL>), grid, block, lds, s, pl, src, xnb,
                       xout, want_norm, gate, color);
  else
    hipLaunchKernelGGL((k_sweep<T, false, GS_VEC, PF, PFB, HB, SL, DL>), grid, block, lds, s, pl, src, xnb,
                       xout, want_norm, gate, color);
}

template <class T, bool HB, int SL>
void run_sweep_v(const Plan& pl, int nblocks, int src, int xnb, int xout, int coupled, int reuse_g0,
                 int want_norm, int gate, int color, hipStream_t s) {
  if (pl.dsteps > 0)
    run_sweep_d<T, HB, SL, true>(pl, nblocks, src, xnb, xout, coupled, reuse_g0, want_norm, gate, color, s);
  else
    run_sweep_d<T, HB, SL, false>(pl, nblocks, src, xnb, xout, coupled, reuse_g0, want_norm, gate, color, s);
}

template <int SL>
void run_sweep_sl(const Plan& pl, int nblocks, int src, int xnb, int xout, int coupled, int reuse_g0,
                  int want_norm, int gate, int color, int prec, int hb, hipStream_t s) {
  if (prec == 0)
    run_sweep_v<double, false, SL>(pl, nblocks, src, xnb, xout, coupled, reuse_g0, want_norm, gate, color, s);
  else if (hb)
    run_sweep_v<float, true, SL>(pl, nblocks, src, xnb, xout, coupled, reuse_g0, want_norm, gate, color, s);
  else
    run_sweep_v<float, false, SL>(pl, nblocks, src, xnb, xout, coupled, reuse_g0, want_norm, gate, color, s);
}

void run_sweep(const Plan& pl, int nblocks, int src, int xnb, int xout, int coupled, int reuse_g0,
               int want_norm, int gate, int color, int prec, int hb, hipStream_t s) {
  if (pl.lines == SWEEP_LINES_SMALL)
    run_sweep_sl<SWEEP_LINES_SMALL>(pl, nblocks, src, xnb, xout, coupled, reuse_g0, want_norm, gate, color,
                                    prec, hb, s);
  else
    run_sweep_sl<SWEEP_LINES_BIG>(pl, nblocks, src, xnb, xout, coupled, reuse_g0, want_norm, gate, color,
                                  prec, hb, s);
}

}  // namespace

void launch_assemble(const Plan& pv, int nblocks, const KnobsDev& k, int prec, hipStream_t s) {
  if (prec == 0) hipLaunchKernelGGL(k_assemble<double>, dim3(nblocks), dim3(NT), 0, s, pv, k);
  else hipLaunchKernelGGL(k_assemble<float>, dim3(nblocks), dim3(NT), 0, s, pv, k);
}

void launch_factor(const Plan& pl, int nblocks, int prec, hipStream_t s) {
  if (prec == 0) hipLaunchKernelGGL(k_factor<double>, dim3(nblocks), dim3(8 * LINES_PER_BLOCK), 0, s, pl);
  else hipLaunchKernelGGL(k_factor<float>, dim3(nblocks), dim3(8 * LINES_PER_BLOCK), 0, s, pl);
}

void launch_line_sweep(const Plan& pl, int nblocks, int src, int xprev, int xout, int coupled,
                       int reuse_g0, int want_norm, int gate, int prec, int hb, hipStream_t s) {
  run_sweep(pl, nblocks, src, xprev, xout, coupled, reuse_g0, want_norm, gate, -1, prec, hb, s);
}

void launch_line_sweep_rb(const Plan& pl, int nblocks, int src, int xout, int color, int coupled,
                          int reuse_g0, int want_norm, int gate, int prec, int hb, hipStream_t s) {
  // neighbours of the other colour are read from xout itself
  run_sweep(pl, nblocks, src, xout, xout, coupled, reuse_g0, want_norm, gate, color, prec, hb, s);
}

}  // namespace a2d
