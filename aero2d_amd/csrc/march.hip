// Column-march form of the cell-face stencil (residual and its exact
// tangent-linear, aero2d/euler.py:259-287 and the jax.jvp at
// aero2d/implicit.py:180), an alternative to the LDS tiles of stencil.hip.
//
// A wavefront owns a strip of columns (i, around the contour) and walks up a
// chunk of rows (j, wall-normal) one row per step. Lane l holds column
// i0 - 2 + l; lanes 2..61 own their column, lanes 0, 1, 62, 63 only supply
// the two-cell i-stencil of the owned ones. Per step the wavefront loads one
// row (64 consecutive columns: one coalesced access per array), converts it
// to primitives, forms the j-slope of the row below and the j-face flux
// between the two rows below; the row two below is then complete in j and
// gets its i-slopes, i-face fluxes and its output. The j-stencil lives in
// registers (three rows of primitives, two rows of slopes, the previous
// j-flux); i-neighbours come from the adjacent lanes by DPP whole-wave
// shifts. No LDS, no barriers, each face flux evaluated once per strip.
//
// Workgroups of four wavefronts take four consecutive (strip, chunk) units of
// one case; workgroup indices are remapped so that consecutive units run on
// the same XCD (the dispatcher deals workgroups to the XCDs in turn), which
// keeps the shared halo columns and rows in that XCD's L2.
#include "arnoldi.hpp"
#include "common.hpp"
#include "launch.hpp"
#include "physics.hpp"

#include <type_traits>

namespace a2d {

namespace {

constexpr int MW = WAVE;          // lanes of a strip
constexpr int MOWN = MARCH_OWN;   // columns a strip owns
constexpr int MWAVES = MARCH_WAVES;  // wavefronts (units) per workgroup
constexpr int MNT = MW * MWAVES;
constexpr int XCDS = 8;
static_assert(MOWN == MW - 4, "a strip owns all but two lanes on each side");

// lane l <- lane l - 1 / lane l + 1 (DPP wave_shr:1 / wave_shl:1). The edge
// lanes receive an unspecified value; they are never owned.
__device__ __forceinline__ float lane_left(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float lane_right(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x130, 0xF, 0xF, false));
}
__device__ __forceinline__ double lane_left(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), 0x138, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), 0x138, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_right(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), 0x130, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), 0x130, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <class T>
__device__ __forceinline__ Dual<T> lane_left(Dual<T> x) { return {lane_left(x.v), lane_left(x.d)}; }
template <class T>
__device__ __forceinline__ Dual<T> lane_right(Dual<T> x) { return {lane_right(x.v), lane_right(x.d)}; }
template <class S>
__device__ __forceinline__ W4<S> lane_left(const W4<S>& w) {
  return {lane_left(w.r), lane_left(w.u), lane_left(w.v), lane_left(w.p)};
}
template <class S>
__device__ __forceinline__ W4<S> lane_right(const W4<S>& w) {
  return {lane_right(w.r), lane_right(w.u), lane_right(w.v), lane_right(w.p)};
}

// Workgroup index -> logical index such that the workgroups an XCD receives
// (b, b + 8, b + 16, ...) form one contiguous logical range. A bijection on
// [0, n).
__device__ __forceinline__ int xcd_remap(int b, int n) {
  const int q = n / XCDS, r = n % XCDS, x = b % XCDS;
  return x * q + (x < r ? x : r) + b / XCDS;
}

// Raw conserved values of one cell (and the tangent direction when dual).
template <class T, bool TAN>
struct Raw {
  T q[4];
  T d[TAN ? 4 : 1];
};

// idx: the cell in the state's natural order; vidx: the same cell in the
// tangent vector's order (colour-split under the red-black line order).
template <class T, bool TAN>
__device__ __forceinline__ void fetch(Raw<T, TAN>& a, gptr<const T> U, gptr<const T> v, int n, int idx, int vidx) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    a.q[k] = U[k * n + idx];
    if constexpr (TAN) a.d[k] = v[k * n + vidx];
  }
}

template <class S, class T, bool TAN>
__device__ __forceinline__ W4<S> prim_of(const Raw<T, TAN>& a) {
  if constexpr (TAN) return cons_to_prim<S>(S(a.q[0], a.d[0]), S(a.q[1], a.d[1]), S(a.q[2], a.d[2]), S(a.q[3], a.d[3]));
  else return cons_to_prim<S>(a.q[0], a.q[1], a.q[2], a.q[3]);
}

template <class S, class T>
__device__ __forceinline__ W4<S> slope(const W4<S>& m, const W4<S>& c, const W4<S>& p, const T eps[4]) {
  return {van_albada(c.r - m.r, p.r - c.r, eps[0]), van_albada(c.u - m.u, p.u - c.u, eps[1]),
          van_albada(c.v - m.v, p.v - c.v, eps[2]), van_albada(c.p - m.p, p.p - c.p, eps[3])};
}

// Face states from the two cells' primitives and limited slopes, clamped
// (aero2d/euler.py:211-215, 274-275).
template <class S, class T>
__device__ __forceinline__ void face(const W4<S>& a, const W4<S>& sa, const W4<S>& b, const W4<S>& sb,
                                     bool second, T sx, T sy, S F[4]) {
  W4<S> L = a, R = b;
  if (second) {
    const S h = K<S>(0.5);
    L = {a.r + h * sa.r, a.u + h * sa.u, a.v + h * sa.v, a.p + h * sa.p};
    R = {b.r - h * sb.r, b.u - h * sb.u, b.v - h * sb.v, b.p - h * sb.p};
  }
  L.r = dmaxc(L.r, T(1e-6));
  L.p = dmaxc(L.p, T(1e-6));
  R.r = dmaxc(R.r, T(1e-6));
  R.p = dmaxc(R.p, T(1e-6));
  hllc_flux<S, T>(L, R, sx, sy, F);
}

// The march kernel. TAN = false: FP64 residual (T = double) with the fused
// local time step, FP32 state refresh and rn partials; TAN = true: the
// tangent-linear in precision T with the same outputs as k_stencil.
// MH rows per chunk; MINW: wavefronts per SIMD the register allocation must
// allow (the FP32 tangent runs at four: 128 VGPRs, measured faster than the
// unconstrained 148 at three despite a few spilled dwords).
template <class T, bool TAN, int MH, int MINW>
__global__ void __launch_bounds__(MNT, MINW) k_march(Plan plan, KnobsDev kn, int vin, int vout, int mode, int kidx,
                                               int use_diag, int want_diag, double circ_override,
                                               int refresh_ut, int gate) {
  using S = typename std::conditional<TAN, Dual<T>, T>::type;
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int e = plan_entry(plan, blk);
  const CaseDev& cd = plan.cases[plan.slots[e]];
  const CaseScal* sc = cd.sc;
  if (sc->parked || (gate && !sc->active)) return;
  const int ni = cd.ni, nc = cd.nc, n = cd.n;
  const int nstrip = (ni + MOWN - 1) / MOWN;
  const int nunit = nstrip * ((nc + MH - 1) / MH);
  const int wave = threadIdx.x / MW, lane = threadIdx.x % MW;
  const int unit = (blk - plan.blk_prefix[e]) * MWAVES + wave;
  if (unit >= nunit) return;                       // a whole wavefront
  const int strip = unit % nstrip, chunk = unit / nstrip;
  const int j0 = chunk * MH;
  const int jend = min(j0 + MH, nc);               // rows j0 .. jend-1 are output
  int col = (strip * MOWN - 2 + lane) % ni;
  if (col < 0) col += ni;
  const int vcol = bcol(col, ni, kn.precond_rb);   // the column in the Krylov vectors
  const bool own = lane >= 2 && lane < MW - 2 && strip * MOWN + lane - 2 < ni;

  gptr<const T> U, vol, Si, Sj, xf, yf;
  if constexpr (std::is_same<T, double>::value) {
    U = gp(cd.U); vol = gp(cd.vol); Si = gp(cd.Si); Sj = gp(cd.Sj); xf = gp(cd.xf); yf = gp(cd.yf);
  } else {
    U = gp(ctp<T>(cd.Ut)); vol = gp(ctp<T>(cd.volt)); Si = gp(ctp<T>(cd.Sit)); Sj = gp(ctp<T>(cd.Sjt));
    xf = gp(ctp<T>(cd.xft)); yf = gp(ctp<T>(cd.yft));
  }
  const gptr<const T> vt = TAN ? vec_gptr<T>(cd, vin) : gptr<const T>(nullptr);
  const Inf<T> inf = {T(sc->winf[0]), T(sc->winf[1]), T(sc->winf[2]), T(sc->winf[3])};
  const double circ_d = (circ_override == circ_override) ? circ_override : sc->circ;
  const T vort = kn.vortex ? T(1) : T(0);
  S circ;
  if constexpr (TAN) circ = S(T(circ_d) * vort, T(sc->dcirc) * vort);
  else circ = S(T(circ_d) * vort);
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
  const bool second = kn.order == 2;
  const int sjy = (nc + 1) * ni;                   // offset of the y components of Sj

  // wall normal for the mirror ghosts (rows -1, -2)
  T wnx = T(0), wny = T(0);
  if (j0 == 0) {
    const T sx = Sj[col], sy = Sj[sjy + col];
    const T a = sqrt(sx * sx + sy * sy);
    wnx = sx / a;
    wny = sy / a;
  }
  // the row state of row r (wave-uniform branch on the row kind); prev = row r-1
  auto src_row = [&](int r) { return r < 0 ? min(-1 - r, nc - 1) : min(r, nc - 1); };
  auto make = [&](int r, const Raw<T, TAN>& a, const W4<S>& prev) -> W4<S> {
    if (r < 0) return mirror(prim_of<S, T, TAN>(a), wnx, wny);
    if (r < nc) return prim_of<S, T, TAN>(a);
    if (r == nc) {
      const T sx = Sj[nc * ni + col], sy = Sj[sjy + nc * ni + col];
      return farfield_ghost<S, T>(prev, sx, sy, inf, xf[col], yf[col], circ);
    }
    return prev;                                   // both far-field ghost layers are alike
  };

  W4<S> A, B, C;                                   // rows r-2, r-1, r
  W4<S> sA, sB;                                    // j-slopes of rows r-2, r-1
  S Fprev[4];                                      // j-flux through the face below row r-2 + 1
  T sjx_prev = T(0), sjy_prev = T(0);              // that face's metric (local time step)
  double rn_part = 0.0;
  const W4<S> zero = {K<S>(0.0), K<S>(0.0), K<S>(0.0), K<S>(0.0)};
  A = B = C = sA = sB = zero;
#pragma unroll
  for (int k = 0; k < 4; ++k) Fprev[k] = K<S>(0.0);

  const int rfirst = j0 - 2, rlast = jend + 1;
  Raw<T, TAN> nxt;
  if (rfirst < nc) fetch<T, TAN>(nxt, U, vt, n, src_row(rfirst) * ni + col, src_row(rfirst) * ni + vcol);
  for (int r = rfirst; r <= rlast; ++r) {
    const Raw<T, TAN> cur = nxt;
    // prefetch the next row's cells (a row below the far field reuses registers)
    if (r + 1 <= rlast && r + 1 < nc)
      fetch<T, TAN>(nxt, U, vt, n, src_row(r + 1) * ni + col, src_row(r + 1) * ni + vcol);
    A = B;
    B = C;
    C = make(r, cur, B);
    sA = sB;
    if (r - 1 >= rfirst + 1 && second) sB = slope(A, B, C, eps);
    if (r - 1 < rfirst + 2) continue;              // no complete j-face yet
    // j-face f = r - 1 between rows r - 2 (A) and r - 1 (B)
    const int f = r - 1;
    const T fsx = Sj[f * ni + col], fsy = Sj[sjy + f * ni + col];
    S Fj[4];
    face<S, T>(A, sA, B, sB, second, fsx, fsy, Fj);
    const int c = r - 2;                           // row A is complete in j
    if (c >= j0) {
      // i-direction of row c: slopes, the face between col-1 and col, the next face by a lane shift
      const W4<S> Al = lane_left(A), Ar = lane_right(A);
      W4<S> si = zero;
      if (second) si = slope(Al, A, Ar, eps);
      const W4<S> sil = lane_left(si);
      const int idx = c * ni + col;
      const T isx = Si[idx], isy = Si[n + idx];
      S Fi[4];
      face<S, T>(Al, sil, A, si, second, isx, isy, Fi);
      T res[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const T fi = TAN ? tan_of(Fi[k]) : val(Fi[k]);
        const T fj = TAN ? tan_of(Fj[k]) : val(Fj[k]);
        const T fp = TAN ? tan_of(Fprev[k]) : val(Fprev[k]);
        res[k] = (lane_right(fi) - fi) + (fj - fp);
      }
      if constexpr (!TAN) {
        const T isxr = lane_right(isx), isyr = lane_right(isy);
        if (own) {
          const gptr<double> R = gp(cd.R);
#pragma unroll
          for (int k = 0; k < 4; ++k) R[k * n + idx] = res[k];
          const double cv = vol[idx];
          const double q = res[0] / cv;
          rn_part += q * q;
          if (want_diag) {
            // local pseudo-time step of the cell (aero2d/euler.py:289-296)
            const double rho = val(A.r), u = val(A.u), v = val(A.v), p = val(A.p);
            const double cs = sqrt(GAMMA * p / rho);
            const double six = 0.5 * (isx + isxr), siy = 0.5 * (isy + isyr);
            const double sjx = 0.5 * (fsx + sjx_prev), sjy2 = 0.5 * (fsy + sjy_prev);
            const double li_ = fabs(u * six + v * siy) + cs * sqrt(six * six + siy * siy);
            const double lj_ = fabs(u * sjx + v * sjy2) + cs * sqrt(sjx * sjx + sjy2 * sjy2);
            const double dt = kn.cfl_explicit * cv / (li_ + lj_);
            const double dtl = dt / kn.cfl_explicit * sc->cfl;
            const double dti = 1.0 / dtl;
            if (cd.prec == 0) tp<double>(cd.diag)[idx] = cv * dti;
            else tp<float>(cd.diag)[idx] = float(cv * dti);
          }
          if (refresh_ut && cd.prec == 1) {
            float* Ut = tp<float>(cd.Ut);
#pragma unroll
            for (int k = 0; k < 4; ++k) Ut[k * n + idx] = float(U[k * n + idx]);
          }
        }
      } else {
        if (own) {
          const gptr<T> out = vec_gptr<T>(cd, vout);
          const int vidx = c * ni + vcol;
          const T dg = use_diag ? gp(ctp<T>(cd.diag))[idx] : T(0);
          if (mode == 0) {
            const T s = (kidx >= 0) ? T(sc->s[kidx]) : T(1);
#pragma unroll
            for (int k = 0; k < 4; ++k) out[k * n + vidx] = s * (dg * vt[k * n + vidx] + res[k]);
          } else {
            const gptr<const double> R = gp(cd.R);
#pragma unroll
            for (int k = 0; k < 4; ++k)
              out[k * n + vidx] = T(-R[k * n + idx]) - (dg * vt[k * n + vidx] + res[k]);
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) Fprev[k] = Fj[k];
    sjx_prev = fsx;
    sjy_prev = fsy;
  }
  if constexpr (!TAN) {
    rn_part = wave_sum(rn_part);
    if (lane == 0) part_at(cd, unit, 0) = rn_part;
  }
}

}  // namespace

void launch_march_residual(const Plan& p, int nblocks, const KnobsDev& k, int want_diag, double circ_override,
                           int refresh_ut, hipStream_t s) {
  hipLaunchKernelGGL((k_march<double, false, MARCH_ROWS_RES, 1>), dim3(nblocks), dim3(MNT), 0, s, p, k, 0, 0, 0,
                     -1, 0, want_diag, circ_override, refresh_ut, 0);
}

void launch_march_jvp(const Plan& p, int nblocks, const KnobsDev& k, int vin, int vout, int mode, int kidx,
                      int use_diag, int gate, int prec, hipStream_t s) {
  if (prec == 0)
    hipLaunchKernelGGL((k_march<double, true, MARCH_ROWS_JVP64, 1>), dim3(nblocks), dim3(MNT), 0, s, p, k, vin,
                       vout, mode, kidx, use_diag, 0, double(NAN), 0, gate);
  else
    hipLaunchKernelGGL((k_march<float, true, MARCH_ROWS_JVP32, 4>), dim3(nblocks), dim3(MNT), 0, s, p, k, vin,
                       vout, mode, kidx, use_diag, 0, double(NAN), 0, gate);
}

}  // namespace a2d
