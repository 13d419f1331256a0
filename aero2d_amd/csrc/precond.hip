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
//               joined at the middle row. Two threads per line, 4x4 blocks in
//               registers.
//   one line-Jacobi sweep (aero2d/implicit.py:130-157) is two kernels. The
//   reference's forward step d_j = inv_j (rhs_j - A_j d_{j-1}) is split as
//   d_j = g_j - P_j d_{j-1} with g_j = inv_j rhs_j and P_j = inv_j A_j:
//   k_sweep_rhs  fully parallel, thread per cell: rhs = r - Bim x(i-1) - Bip x(i+1)
//                (r alone in the first sweep), g = inv rhs.
//   k_sweep_seq  the j recurrences only, both halves of each line at once
//                (eight lanes per line, lane r owns row r of each 4x4 product,
//                rows exchanged in the quad; loads for the next PF steps are in
//                flight while the current PF steps compute).
//   The same direct solve as the reference in another elimination order, so
//   the same answer up to rounding.
//
// Block storage: T blk[5][4][pstride][4] = slot, block row, cell, column, so a
// block row is one 16-byte load in FP32 (two in FP64). Slots after assembly:
// 0 = D, 1 = Bjp, 2 = Bjm, 3 = Bim, 4 = Bip; after factorisation 0 = inverse
// of the eliminated diagonal, 1 = backward coupling (c or e), 2 = forward
// coupling (P or Q). With the red-black line order the cell index is
// colour-split (column of line i = (i % 2) ni/2 + i/2).
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
  const size_t bidx = (size_t)j * ni + bcol(i, ni, kn.precond_rb);
  auto at = [&](int slot, int m) -> T& { return blk[blk_off(slot, m, bidx, plane)]; };

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

// Row-wise access to the block storage (layout in the file header).
template <class T>
struct Blk {
  gptr<T> p;
  size_t ps;
  __device__ __forceinline__ gptr<T> row(int slot, int r, size_t cell) const {
    return p + (((size_t)slot * 4 + r) * ps + cell) * 4;
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

template <class T>
__device__ __forceinline__ void load_blk(const Blk<const T>& B, int slot, size_t cell, T a[16]) {
#pragma unroll
  for (int r = 0; r < 4; ++r) load_row(B.row(slot, r, cell), a + 4 * r);
}
template <class T>
__device__ __forceinline__ void store_blk(const Blk<T>& B, int slot, size_t cell, const T a[16]) {
#pragma unroll
  for (int r = 0; r < 4; ++r) store_row(B.row(slot, r, cell), a + 4 * r);
}
template <class T>
__device__ __forceinline__ void mul4(const T a[16], const T b[16], T out[16]) {
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      T s = T(0);
#pragma unroll
      for (int k = 0; k < 4; ++k) s += a[r * 4 + k] * b[k * 4 + c];
      out[r * 4 + c] = s;
    }
}

// Two-sided (twisted) block elimination of every j-line, two threads per line.
// With m = nc / 2, the top thread eliminates j = 0 .. m-1 downward from the
// wall exactly as aero2d/implicit.py:112-128 does (den = D - A c_prev,
// inv = den^-1, c = inv C, P = inv A); the bottom thread eliminates
// j = nc-1 .. m+1 upward from the far field (den = D - C e_prev, inv = den^-1,
// e = inv A, Q = inv C); row m closes both: M = D - A c_{m-1} - C e_{m+1},
// slot 0 = M^-1, slot 2 = M^-1 A, slot 1 = M^-1 C. In every row slot 0 maps
// the right-hand side to g, slot 2 multiplies the recurrence value arriving
// from the line end and slot 1 the one arriving from the middle on the way
// back. The same direct solve as the reference's with half the chain length.
template <class T>
__global__ void __launch_bounds__(64) k_factor(Plan plan) {
  const int e = plan_entry(plan, blockIdx.x);
  const CaseDev& cd = plan.cases[plan.slots[e]];
  const int ni = cd.ni, nc = cd.nc;
  const int half = threadIdx.x & 1;
  const int q = (blockIdx.x - plan.blk_prefix[e]) * LINES_PER_BLOCK + (threadIdx.x >> 1);
  const bool live = q < ni;
  const int col = live ? q : ni - 1;   // dead lanes shadow a real line and never store
  const int m = nc >> 1;
  const int steps = half ? nc - 1 - m : m;
  const int j0 = half ? nc - 1 : 0, dj = half ? -1 : 1;
  const Blk<T> B{gp(tp<T>(cd.blk)), (size_t)cd.pstride};
  const Blk<const T> Bc{gp(ctp<T>(cd.blk)), (size_t)cd.pstride};
  T cprev[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) cprev[k] = T(0);
  for (int s = 0; s < m; ++s) {
    if (s < steps) {
      const size_t cell = (size_t)(j0 + dj * s) * ni + col;
      T D[16], A[16], C[16];
      load_blk(Bc, 0, cell, D);
      load_blk(Bc, 2, cell, A);
      load_blk(Bc, 1, cell, C);
      T L[16], F[16], t[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        L[k] = half ? C[k] : A[k];    // couples to the row already eliminated
        F[k] = half ? A[k] : C[k];    // couples to the row eliminated next
      }
      mul4<T>(L, cprev, t);
#pragma unroll
      for (int k = 0; k < 16; ++k) D[k] -= t[k];
      T inv[16], P[16];
      inv4<T>(D, inv);
      mul4<T>(inv, F, cprev);
      mul4<T>(inv, L, P);
      if (live) {
        store_blk(B, 0, cell, inv);
        store_blk(B, 1, cell, cprev);
        store_blk(B, 2, cell, P);
      }
    }
  }
  T other[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) other[k] = __shfl_xor(cprev[k], 1, WAVE);
  if (half == 0 && live) {
    const size_t cell = (size_t)m * ni + col;
    T D[16], A[16], C[16], t[16];
    load_blk(Bc, 0, cell, D);
    load_blk(Bc, 2, cell, A);
    load_blk(Bc, 1, cell, C);
    mul4<T>(A, cprev, t);           // A c_{m-1}
#pragma unroll
    for (int k = 0; k < 16; ++k) D[k] -= t[k];
    mul4<T>(C, other, t);           // C e_{m+1} (zero when the bottom part is empty)
#pragma unroll
    for (int k = 0; k < 16; ++k) D[k] -= t[k];
    T inv[16], P[16], Q[16];
    inv4<T>(D, inv);
    mul4<T>(inv, A, P);
    mul4<T>(inv, C, Q);
    store_blk(B, 0, cell, inv);
    store_blk(B, 1, cell, Q);
    store_blk(B, 2, cell, P);
  }
}

// Sweep part 1, thread per cell: g = slot0 (src - Bim x(i-1) - Bip x(i+1))
// into the work buffer d (four components per cell, contiguous).
template <class T>
__global__ void __launch_bounds__(256) k_sweep_rhs(Plan plan, int src, int xprev, int sweep,
                                                   int gate, int color) {
  const int e = plan_entry(plan, blockIdx.x);
  const CaseDev& cd = plan.cases[plan.slots[e]];
  if (gate && !cd.sc->active) return;
  const int ni = cd.ni, n = cd.n;
  const int t = (blockIdx.x - plan.blk_prefix[e]) * 256 + threadIdx.x;
  int i, j, col;
  if (color < 0) {            // every cell, natural block layout
    if (t >= n) return;
    i = t % ni;
    j = t / ni;
    col = i;
  } else {                    // cells of lines i = color (mod 2), colour-split blocks
    const int nh = ni >> 1;
    if (t >= nh * cd.nc) return;
    const int q = t % nh;
    j = t / nh;
    i = 2 * q + color;
    col = color * nh + q;
  }
  const int idx = j * ni + i;                  // natural position (vectors)
  const size_t bidx = (size_t)j * ni + col;    // block storage and g
  const Blk<const T> B{gp(ctp<T>(cd.blk)), (size_t)cd.pstride};
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
    for (int r = 0; r < 4; ++r) {
      T bm[4], bp[4];
      load_row(B.row(3, r, bidx), bm);
      load_row(B.row(4, r, bidx), bp);
#pragma unroll
      for (int c = 0; c < 4; ++c) rhs[r] -= bm[c] * xm[c] + bp[c] * xq[c];
    }
  }
  T g[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    T iv[4];
    load_row(B.row(0, r, bidx), iv);
    T s = T(0);
#pragma unroll
    for (int c = 0; c < 4; ++c) s += iv[c] * rhs[c];
    g[r] = s;
  }
  store_row(gp(tp<T>(cd.d)) + bidx * 4, g);
}

// Sweep part 2: the j recurrences of the two-sided solve. Eight lanes per
// line: lanes 0-3 (one per block row) run j = 0 .. m-1, lanes 4-7 run
// j = nc-1 .. m+1, each d_j = g_j - slot2_j d_prev in place over g; the two
// halves meet at row m (x_m = g_m - slot2_m d_{m-1} - slot1_m f_{m+1}), then
// walk back out, x_j = d_j - slot1_j x_prev. Rows are exchanged inside each
// quad of lanes; loads for the next PF steps are in flight while the current
// PF steps compute (double-buffered registers).
template <class T, int PF>
__global__ void __launch_bounds__(256) k_sweep_seq(Plan plan, int xout, int want_norm, int gate,
                                                   int color) {
  __shared__ double red[4];
  const int e = plan_entry(plan, blockIdx.x);
  const CaseDev& cd = plan.cases[plan.slots[e]];
  if (gate && !cd.sc->active) return;
  const int ni = cd.ni, nc = cd.nc, n = cd.n;
  const int lb = blockIdx.x - plan.blk_prefix[e];
  const int r = threadIdx.x & 3;
  const int half = (threadIdx.x >> 2) & 1;
  const int q = lb * LINES_PER_BLOCK + (threadIdx.x >> 3);
  // colour < 0: every line, natural block layout; else line q of colour
  // `color` (i = 2q + color), whose blocks sit in column color * ni/2 + q
  const int nq = color < 0 ? ni : (ni >> 1);
  if (lb * LINES_PER_BLOCK >= nq) {
    // a whole block past the colour's lines: nothing to solve (the plan is
    // sized for every line); its norm partial reads zero
    if (want_norm && threadIdx.x == 0) {
      const int slot = color < 0 ? lb : color * cd.nlineb + lb;
      cd.part[(size_t)slot * (cd.kmax + 2)] = 0.0;
    }
    return;
  }
  const bool live = q < nq;
  const int qc = live ? q : nq - 1;     // dead lanes shadow a real line and never store
  const int ic = color < 0 ? qc : 2 * qc + color;
  const int col = color < 0 ? qc : color * (ni >> 1) + qc;
  const int m = nc >> 1;
  const int steps = half ? nc - 1 - m : m;
  const int j0 = half ? nc - 1 : 0, dj = half ? -1 : 1;
  const Blk<const T> B{gp(ctp<T>(cd.blk)), (size_t)cd.pstride};
  const gptr<T> gd = gp(tp<T>(cd.d));
  const gptr<T> x = gp(vec_ptr<T>(cd, xout) + (size_t)r * n + ic);

  T bm[2][PF][4], bv[2][PF];
  // toward the middle: d_j = g_j - slot2_j d_prev
  auto fcell = [&](int s) -> size_t {
    const int sc = max(0, min(s, steps - 1));
    return (size_t)(j0 + dj * sc) * ni + col;
  };
  auto fload = [&](int b, int s0) {
#pragma unroll
    for (int k = 0; k < PF; ++k) {
      const size_t cell = fcell(s0 + k);
      bv[b][k] = gd[cell * 4 + r];
      load_row(B.row(2, r, cell), bm[b][k]);
    }
  };
  T dr = T(0);
  auto fcomp = [&](int b, int s0) {
#pragma unroll
    for (int k = 0; k < PF; ++k) {
      if (s0 + k < steps) {
        const T t = bv[b][k] - bm[b][k][0] * quad_bcast<0>(dr) - bm[b][k][1] * quad_bcast<1>(dr) -
                    bm[b][k][2] * quad_bcast<2>(dr) - bm[b][k][3] * quad_bcast<3>(dr);
        dr = t;
        if (live) gd[fcell(s0 + k) * 4 + r] = dr;
      }
    }
  };
  fload(0, 0);
  for (int s0 = 0; s0 < m; s0 += 2 * PF) {
    fload(1, s0 + PF);
    fcomp(0, s0);
    fload(0, s0 + 2 * PF);
    fcomp(1, s0 + PF);
  }
  // row m: both quads combine the top's d_{m-1} with the bottom's f_{m+1}
  const T other = __shfl_xor(dr, 4, WAVE);
  const T dtop = half ? other : dr, fbot = half ? dr : other;
  T xr;
  {
    const size_t cell = (size_t)m * ni + col;
    T pm[4], qm[4];
    load_row(B.row(2, r, cell), pm);
    load_row(B.row(1, r, cell), qm);
    xr = gd[cell * 4 + r] - pm[0] * quad_bcast<0>(dtop) - pm[1] * quad_bcast<1>(dtop) -
         pm[2] * quad_bcast<2>(dtop) - pm[3] * quad_bcast<3>(dtop) - qm[0] * quad_bcast<0>(fbot) -
         qm[1] * quad_bcast<1>(fbot) - qm[2] * quad_bcast<2>(fbot) - qm[3] * quad_bcast<3>(fbot);
  }
  double nrm = 0.0;
  if (live && half == 0) {
    x[(size_t)m * ni] = xr;
    nrm += double(xr) * double(xr);
  }
  // away from the middle: x_j = d_j - slot1_j x_prev
  auto bj = [&](int s) -> int {
    const int sc = max(0, min(s, steps - 1));
    return steps > 0 ? m - dj * (1 + sc) : m;
  };
  auto bload = [&](int b, int s0) {
#pragma unroll
    for (int k = 0; k < PF; ++k) {
      const size_t cell = (size_t)bj(s0 + k) * ni + col;
      bv[b][k] = gd[cell * 4 + r];
      load_row(B.row(1, r, cell), bm[b][k]);
    }
  };
  auto bcomp = [&](int b, int s0) {
#pragma unroll
    for (int k = 0; k < PF; ++k) {
      if (s0 + k < steps) {
        const T t = bv[b][k] - bm[b][k][0] * quad_bcast<0>(xr) - bm[b][k][1] * quad_bcast<1>(xr) -
                    bm[b][k][2] * quad_bcast<2>(xr) - bm[b][k][3] * quad_bcast<3>(xr);
        xr = t;
        if (live) {
          x[(size_t)bj(s0 + k) * ni] = xr;
          nrm += double(xr) * double(xr);
        }
      }
    }
  };
  bload(0, 0);
  for (int s0 = 0; s0 < m; s0 += 2 * PF) {
    bload(1, s0 + PF);
    bcomp(0, s0);
    bload(0, s0 + 2 * PF);
    bcomp(1, s0 + PF);
  }
  if (want_norm) {
    nrm = wave_sum(nrm);
    if ((threadIdx.x & (WAVE - 1)) == 0) red[threadIdx.x / WAVE] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) {
      const int slot = color < 0 ? lb : color * cd.nlineb + lb;
      cd.part[(size_t)slot * (cd.kmax + 2)] = red[0] + red[1] + red[2] + red[3];
    }
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
    hipLaunchKernelGGL(k_sweep_rhs<double>, dim3(ncell), dim3(256), 0, s, pcell, src, xprev, sweep, gate,
                       -1);
    hipLaunchKernelGGL((k_sweep_seq<double, SWEEP_PF>), dim3(nblocks), dim3(256), 0, s, pl, xout,
                       want_norm, gate, -1);
  } else {
    hipLaunchKernelGGL(k_sweep_rhs<float>, dim3(ncell), dim3(256), 0, s, pcell, src, xprev, sweep, gate,
                       -1);
    hipLaunchKernelGGL((k_sweep_seq<float, SWEEP_PF>), dim3(nblocks), dim3(256), 0, s, pl, xout,
                       want_norm, gate, -1);
  }
}

void launch_line_sweep_rb(const Plan& pcell, int ncell, const Plan& pl, int nblocks, int src, int xout,
                          int color, int coupled, int want_norm, int gate, int prec, hipStream_t s) {
  // the plans are sized for every cell / line; blocks past the colour's half return at once
  const int sweep = coupled ? 1 : 0;
  if (prec == 0) {
    hipLaunchKernelGGL(k_sweep_rhs<double>, dim3(ncell), dim3(256), 0, s, pcell, src, xout, sweep, gate,
                       color);
    hipLaunchKernelGGL((k_sweep_seq<double, SWEEP_PF>), dim3(nblocks), dim3(256), 0, s, pl, xout,
                       want_norm, gate, color);
  } else {
    hipLaunchKernelGGL(k_sweep_rhs<float>, dim3(ncell), dim3(256), 0, s, pcell, src, xout, sweep, gate,
                       color);
    hipLaunchKernelGGL((k_sweep_seq<float, SWEEP_PF>), dim3(nblocks), dim3(256), 0, s, pl, xout,
                       want_norm, gate, color);
  }
}

void launch_sweep_part(const Plan& pcell, int ncell, const Plan& pl, int nblocks, int part, int prec,
                       hipStream_t s) {
  if (part == 0) {
    if (prec == 0) hipLaunchKernelGGL(k_sweep_rhs<double>, dim3(ncell), dim3(256), 0, s, pcell, VEC_W, VEC_XA, 1, 0, -1);
    else hipLaunchKernelGGL(k_sweep_rhs<float>, dim3(ncell), dim3(256), 0, s, pcell, VEC_W, VEC_XA, 1, 0, -1);
  } else {
    if (prec == 0)
      hipLaunchKernelGGL((k_sweep_seq<double, SWEEP_PF>), dim3(nblocks), dim3(256), 0, s, pl, VEC_XB, 0, 0, -1);
    else
      hipLaunchKernelGGL((k_sweep_seq<float, SWEEP_PF>), dim3(nblocks), dim3(256), 0, s, pl, VEC_XB, 0, 0, -1);
  }
}

}  // namespace a2d
