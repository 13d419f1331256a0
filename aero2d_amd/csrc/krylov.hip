// GMRES and Newton-update kernels (aero2d/implicit.py:184-191 and the
// "batched" GMRES of jax.scipy.sparse.linalg that it calls).
//
// Basis vectors are stored raw: V_true_j = s_j * V_raw_j with the scale s_j
// kept in the case scalars. Normalising is then a scalar update, the JVP
// applies s_j to its output, and no separate scaling pass touches HBM.
//
// FP16 basis (mixed form, knob basis_half; template flag BH): the
// preconditioner writes its output to an FP32 work vector P; V_raw_j is stored
// in FP16 as sigma_j times the orthogonalised vector, sigma_j a power of two
// that brings the entries to order one (norm sqrt(4n)), and s_j = 1/|V_raw_j|
// is taken from the rounded values themselves. The update also writes the
// rounded vector in FP32 (vc), which the next Jacobian-vector product reads,
// so the Arnoldi relation holds for the stored basis. Dot products and the
// update accumulate in FP64 as before.
//
// Reductions are two-level and case-local: every block writes its partial
// sums to the case's partial array, then the case's sums are formed from them
// in a fixed order (by a small kernel per case, or by every block of the next
// kernel of the chain that needs them). Results never depend on which other
// cases share a launch.
#include "arnoldi.hpp"
#include "common.hpp"
#include "launch.hpp"
#include "physics.hpp"

#include <cfloat>
#include <type_traits>

namespace a2d {

namespace {

constexpr int NT = 256;
constexpr int PER = VEC_CHUNK / NT;   // cells per thread in the vector kernels
#ifndef A2D_UPDATE_JB64
#define A2D_UPDATE_JB64 2             // basis vectors per pass of the CGS update, FP64
#endif
#ifndef A2D_UPDATE_JB32
#define A2D_UPDATE_JB32 4             // and FP32
#endif

// Power-of-two storage scale for an FP16 basis vector of (estimated) norm nrm
// over 4n entries: the scaled norm is ~sqrt(4n) / 64, so typical entries are
// ~1/64 (FP16 normals reach down to 6.1e-5) and an entry overflows only if the
// true norm exceeds the estimate ~5000-fold for n = 2^17. Conversion also
// saturates at +-65504 (NaN kept).
__device__ __forceinline__ double basis_sig(double nrm, int n) {
  if (!(nrm > 0.0) || !isfinite(nrm)) return 1.0;
  return exp2(rint(log2(sqrt(4.0 * n) / (64.0 * nrm))));
}
__device__ __forceinline__ _Float16 half_sat(double v) {
  const double c = fmin(fmax(v, -65504.0), 65504.0);
  return (_Float16)float(v == v ? c : v);
}

// Element type of the stored basis.
template <class T, bool BH> using BasisT = typename std::conditional<BH, _Float16, T>::type;

// h_j = <V_j, w> partials for j = 0..k (times s_j at reduction), and <w, w>.
template <class T, bool BH>
__global__ void __launch_bounds__(NT) k_dots(Plan plan, int k, int wvec, int gate) {
  using VT = BasisT<T, BH>;
  __shared__ double sred[KMAX + 2][NT / WAVE];
  const int e = plan_entry(plan, blockIdx.x);
  const CaseDev& cd = plan.cases[plan.slots[e]];
  if (gate && !cd.sc->active) return;
  const int n = cd.n;
  const int lb = blockIdx.x - plan.blk_prefix[e];
  const int c0 = lb * VEC_CHUNK + threadIdx.x;
  const gptr<const T> w = vec_gptr<T>(cd, wvec);
  const gptr<const VT> Vb = gp((const VT*)cd.V);
  const size_t vs = (size_t)cd.vstride;
  // indices are clamped into the case and out-of-range cells contribute zero
  // through selects, so every load is unconditional (no per-element branches)
  int cc[PER];
  bool in[PER];
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    const int c = c0 + m * NT;
    in[m] = c < n;
    cc[m] = in[m] ? c : n - 1;
  }
  T wv[4][PER];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const T x = w[(size_t)v * n + cc[m]];
      wv[v][m] = in[m] ? x : T(0);
    }
  const int wave = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
  // four basis vectors per pass so that their loads are in flight together
  constexpr int JB = 4;
  for (int j0 = 0; j0 <= k; j0 += JB) {
    double acc[JB];
#pragma unroll
    for (int q = 0; q < JB; ++q) {
      acc[q] = 0.0;
      const int j = min(j0 + q, k);
      const gptr<const VT> V = Vb + (size_t)j * vs;
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int m = 0; m < PER; ++m) acc[q] += double(T(V[(size_t)v * n + cc[m]])) * double(wv[v][m]);
    }
#pragma unroll
    for (int q = 0; q < JB; ++q) {
      const double a = wave_sum(acc[q]);
      if (lane == 0 && j0 + q <= k) sred[j0 + q][wave] = a;
    }
  }
  {
    double acc = 0.0;
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int m = 0; m < PER; ++m) acc += double(wv[v][m]) * double(wv[v][m]);
    acc = wave_sum(acc);
    if (lane == 0) sred[k + 1][wave] = acc;
  }
  __syncthreads();
  for (int j = threadIdx.x; j <= k + 1; j += NT) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < NT / WAVE; ++q) s += sred[j][q];
    part_at(cd, lb, j) = s;
  }
}

// The case's sums of the dot products' partials, a wavefront per column in
// block order: hcol[j] = H[j][k] = s_j * sum_b part[b][j] for j <= k, and
// hcol[k + 1] = |w|. Every block that forms them gets the same values.
__device__ __forceinline__ void dots_sums(const CaseDev& cd, int k, double* hcol) {
  for (int j = threadIdx.x / WAVE; j <= k + 1; j += NT / WAVE) {
    const double s = case_sum(cd, cd.nvecb, j);
    if ((threadIdx.x & (WAVE - 1)) == 0) hcol[j] = j <= k ? cd.sc->s[j] * s : sqrt(s);
  }
}

// FP16 basis: the storage scale of V_{k+1} from the norm the orthogonalised
// vector will have, |w|^2 - sum_j H[j][k]^2 (clamped below at (1e-3 |w|)^2:
// the estimate's error from the basis's departure from orthonormality is far
// smaller than that, so the true norm is at most about twice the estimate and
// no stored entry can overflow).
__device__ __forceinline__ double next_sig(const double* hcol, int k, int n) {
  const double vn0 = hcol[k + 1], w2 = vn0 * vn0;
  double h2 = 0.0;
  for (int j = 0; j <= k; ++j) h2 += hcol[j] * hcol[j];
  return basis_sig(sqrt(fmax(w2 - h2, 1e-6 * w2)), n);
}

// One block per case: H[:, k], |w| and (bh) the next storage scale into the
// case scalars, for the update to read.
__global__ void __launch_bounds__(NT) k_reduce_dots(Plan pc, int k, int bh) {
  __shared__ double hcol[KMAX + 2];
  const CaseDev& cd = pc.cases[pc.slots[blockIdx.x]];
  CaseScal* sc = cd.sc;
  if (!sc->active) return;
  dots_sums(cd, k, hcol);
  __syncthreads();
  for (int j = threadIdx.x; j <= k; j += NT) sc->H[j][k] = hcol[j];
  if (threadIdx.x == 0) {
    sc->vn0 = hcol[k + 1];
    if (bh) sc->vsig = next_sig(hcol, k, cd.n);
  }
}

// w <- w - sum_j H[j][k] s_j V_j (classical Gram-Schmidt), in place in V_{k+1};
// partial |w|^2 per block. BH: w is read from wvec, V_raw_{k+1} = sigma w is
// stored in FP16 and in FP32 (vc), and the partials are of the stored values.
//
// red: every block first forms the case's sums of the dot products' partials
// itself (the case's first block stores them), in place of k_reduce_dots: one
// launch fewer, for batches whose blocks are all resident at once; a larger
// launch would pay the serial reduction once per round of blocks.
template <class T, bool BH>
__global__ void __launch_bounds__(NT) k_update(Plan plan, int k, int wvec, int gate, int red) {
  using VT = BasisT<T, BH>;
  __shared__ double sred[NT / WAVE];
  __shared__ double hcol[KMAX + 2];
  __shared__ double coef[KMAX + 1];
  __shared__ double ssig;
  const int e = plan_entry(plan, blockIdx.x);
  const CaseDev& cd = plan.cases[plan.slots[e]];
  CaseScal* sc = cd.sc;
  if (gate && !sc->active) return;
  const int n = cd.n;
  const int lb = blockIdx.x - plan.blk_prefix[e];
  const int c0 = lb * VEC_CHUNK + threadIdx.x;
  if (red) {
    dots_sums(cd, k, hcol);
    __syncthreads();
    if (threadIdx.x == 0) {
      double sig = 1.0;
      if constexpr (BH) sig = next_sig(hcol, k, n);
      ssig = sig;
      if (lb == 0) {
        sc->vn0 = hcol[k + 1];
        if constexpr (BH) sc->vsig = sig;
      }
    }
    for (int j = threadIdx.x; j <= k; j += NT) {
      coef[j] = hcol[j] * sc->s[j];
      if (lb == 0) sc->H[j][k] = hcol[j];
    }
  } else {
    for (int j = threadIdx.x; j <= k; j += NT) coef[j] = sc->H[j][k] * sc->s[j];
    if (threadIdx.x == 0) ssig = BH ? sc->vsig : 1.0;
  }
  __syncthreads();
  const gptr<T> w = vec_gptr<T>(cd, wvec);
  const gptr<const VT> Vb = gp((const VT*)cd.V);
  const size_t vs = (size_t)cd.vstride;
  // all PER cells of a thread at once, JB basis vectors per pass, so that
  // JB x 4 x PER loads are in flight; indices are clamped into the case and
  // only in-range cells are stored
  int cc[PER];
  bool in[PER];
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    const int c = c0 + m * NT;
    in[m] = c < n;
    cc[m] = in[m] ? c : n - 1;
  }
  double acc[4][PER];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int m = 0; m < PER; ++m) acc[v][m] = double(w[(size_t)v * n + cc[m]]);
  constexpr int JB = std::is_same<T, double>::value ? A2D_UPDATE_JB64 : A2D_UPDATE_JB32;
  int j = 0;
  for (; j + JB - 1 <= k; j += JB) {
    VT vv[JB][4][PER];
#pragma unroll
    for (int q = 0; q < JB; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int m = 0; m < PER; ++m) vv[q][v][m] = Vb[(size_t)(j + q) * vs + (size_t)v * n + cc[m]];
#pragma unroll
    for (int q = 0; q < JB; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int m = 0; m < PER; ++m) acc[v][m] -= coef[j + q] * double(T(vv[q][v][m]));
  }
  for (; j <= k; ++j) {
    const double cj = coef[j];
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int m = 0; m < PER; ++m) acc[v][m] -= cj * double(T(Vb[(size_t)j * vs + (size_t)v * n + cc[m]]));
  }
  double nrm = 0.0;
  if constexpr (BH) {
    const double sig = ssig;
    const gptr<_Float16> Vo = gp((_Float16*)cd.V + (size_t)(k + 1) * vs);
    const gptr<T> vc = vec_gptr<T>(cd, VEC_VC);
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      if (!in[m]) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const _Float16 h = half_sat(sig * acc[v][m]);
        const T r = T(h);
        Vo[(size_t)v * n + cc[m]] = h;
        vc[(size_t)v * n + cc[m]] = r;
        nrm += double(r) * double(r);
      }
    }
  } else {
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      if (!in[m]) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const T r = T(acc[v][m]);
        w[(size_t)v * n + cc[m]] = r;
        nrm += double(r) * double(r);
      }
    }
  }
  nrm = wave_sum(nrm);
  if ((threadIdx.x & (WAVE - 1)) == 0) sred[threadIdx.x / WAVE] = nrm;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < NT / WAVE; ++q) s += sred[q];
    part_at(cd, lb, update_norm_col(cd)) = s;
  }
}

// FP16 basis: V_raw_0 = vsig * P in FP16, and its rounded values in FP32 (vc).
__global__ void __launch_bounds__(NT) k_store_v0(Plan plan) {
  const int e = plan_entry(plan, blockIdx.x);
  const CaseDev& cd = plan.cases[plan.slots[e]];
  const CaseScal* sc = cd.sc;
  if (!sc->run) return;
  const int n = cd.n;
  const int lb = blockIdx.x - plan.blk_prefix[e];
  const double sig = sc->vsig;
  const gptr<const float> pv = vec_gptr<float>(cd, VEC_P);
  const gptr<float> vc = vec_gptr<float>(cd, VEC_VC);
  const gptr<_Float16> V0 = gp((_Float16*)cd.V);
  for (int m = 0; m < PER; ++m) {
    const int c = lb * VEC_CHUNK + threadIdx.x + m * NT;
    if (c >= n) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const size_t o = (size_t)v * n + c;
      const _Float16 h = half_sat(sig * double(pv[o]));
      V0[o] = h;
      vc[o] = float(h);
    }
  }
}

// One wavefront per case: finish Arnoldi step k (the last one of a cycle; the
// others are finished by the kernel that opens the next step).
template <class T>
__global__ void __launch_bounds__(WAVE) k_normalize(Plan pc, int k, KnobsDev kn) {
  __shared__ ArnoldiLds lds;
  const CaseDev& cd = pc.cases[pc.slots[blockIdx.x]];
  if (!cd.sc->active) return;
  arnoldi_finish<T>(cd, k, kn, lds);
}

// stage 0: |M b| from the x0 sweep's partials.
// stage 1: beta = |M(b - A x0)| from the V_0 sweep's partials; run flag.
template <class T>
__global__ void __launch_bounds__(WAVE) k_krylov_begin(Plan pc, int stage, KnobsDev kn) {
  const CaseDev& cd = pc.cases[pc.slots[blockIdx.x]];
  CaseScal* sc = cd.sc;
  const int nblk = kn.precond_rb ? 2 * cd.nlinerw : cd.nlinew;
  const double nrm = sqrt(case_sum(cd, nblk, 0));
  if (threadIdx.x != 0) return;
  if (stage == 0) {
    sc->mbnorm = nrm;
    return;
  }
  // FP64 keeps JAX's absolute threshold (machine eps on the norm); the FP32
  // Krylov of the mixed form scales it by |M b| so that deep convergence of the
  // FP64 outer residual is not cut off by the size of its FP32 corrections.
  const double thresh = std::is_same<T, double>::value ? eps_of<T>() : eps_of<T>() * sc->mbnorm;
  const double beta = (nrm > thresh) ? nrm : 0.0;
  const double ptol = sc->mbnorm * fmin(1.0, kn.gtol);
  sc->beta = beta;
  // FP16 basis: V_raw_0 = vsig * P is stored by k_store_v0
  const double vsig = (std::is_same<T, float>::value && kn.basis_half) ? basis_sig(beta, cd.n) : 1.0;
  sc->vsig = vsig;
  sc->s[0] = beta > 0 ? 1.0 / (vsig * beta) : 0.0;
  sc->g[0] = beta;
  sc->gres = beta;
  sc->run = beta > ptol && !sc->parked;
  sc->active = sc->run && kn.restart > 0;
  sc->kdone = 0;
}

// y = R^-1 g by back substitution over the completed Arnoldi steps, one
// wavefront per case: the triangle and g are staged in LDS by all lanes, then
// the lanes form each row's sum together (fixed-order wave reduction).
__global__ void __launch_bounds__(WAVE) k_lstsq(Plan pc) {
  __shared__ double sH[KMAX][KMAX + 1];
  __shared__ double sy[KMAX];
  CaseScal* sc = pc.cases[pc.slots[blockIdx.x]].sc;
  const int lane = threadIdx.x;
  const int kd = sc->run ? sc->kdone : 0;
  for (int i = 0; i < kd; ++i)
    for (int l = i + lane; l < kd; l += WAVE) sH[i][l] = sc->H[i][l];
  __syncthreads();
  for (int i = kd - 1; i >= 0; --i) {
    double part = 0.0;
    for (int l = i + 1 + lane; l < kd; l += WAVE) part += sH[i][l] * sy[l];
    const double acc = sc->g[i] - wave_sum(part);
    const double yi = acc / sH[i][i];
    if (lane == 0) sy[i] = yi;
    __syncthreads();
  }
  for (int i = lane; i < kd; i += WAVE) sc->y[i] = sy[i];
  if (lane == 0) sc->kdone = kd;
}

// dU = x0 + sum_j y_j s_j V_j (stored FP64), with the step-limiter partials:
// max over cells of |rho_new - rho|/rho and |p_new - p|/p for U + dU, and the
// count of non-finite entries of dU (aero2d/implicit.py:186-189). Cells run
// in the natural order of U and dU; split: x0 and the basis are colour-split
// (red-black line order), read at column bcol (two contiguous runs per
// wavefront).
template <class T, bool BH>
__global__ void __launch_bounds__(NT) k_dupdate(Plan plan, int split) {
  using VT = BasisT<T, BH>;
  __shared__ double coef[KMAX];
  __shared__ double smax[NT / WAVE], scnt[NT / WAVE];
  const int e = plan_entry(plan, blockIdx.x);
  const CaseDev& cd = plan.cases[plan.slots[e]];
  const CaseScal* sc = cd.sc;
  if (sc->parked) return;
  const int n = cd.n, kd = sc->kdone, ni = cd.ni;
  const int lb = blockIdx.x - plan.blk_prefix[e];
  for (int j = threadIdx.x; j < kd; j += NT) coef[j] = sc->y[j] * sc->s[j];
  __syncthreads();
  const gptr<const T> x0 = gp(ctp<T>(cd.x0));
  const gptr<const VT> Vb = gp((const VT*)cd.V);
  const size_t vs = (size_t)cd.vstride;
  const gptr<const double> U = gp(cd.U);
  const gptr<double> dU = gp(cd.dU);
  double rel = 0.0, bad = 0.0;
  for (int m = 0; m < PER; ++m) {
    const int c = lb * VEC_CHUNK + threadIdx.x + m * NT;
    if (c >= n) continue;
    int cv = c;
    if (split) {
      const int row = c / ni, i = c - row * ni;
      cv = row * ni + bcol(i, ni, 1);
    }
    double du[4], u[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) du[v] = double(x0[(size_t)v * n + cv]);
    int j = 0;
    for (; j + 4 <= kd; j += 4) {
      VT vv[4][4];
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int v = 0; v < 4; ++v) vv[q][v] = Vb[(size_t)(j + q) * vs + (size_t)v * n + cv];
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int v = 0; v < 4; ++v) du[v] += coef[j + q] * double(T(vv[q][v]));
    }
    for (; j < kd; ++j) {
      const double cj = coef[j];
#pragma unroll
      for (int v = 0; v < 4; ++v) du[v] += cj * double(T(Vb[(size_t)j * vs + (size_t)v * n + cv]));
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const size_t o = (size_t)v * n + c;
      dU[o] = du[v];
      u[v] = U[o];
      if (!isfinite(du[v])) bad += 1.0;
    }
    const W4<double> w0 = cons_to_prim<double>(u[0], u[1], u[2], u[3]);
    const W4<double> w1 =
        cons_to_prim<double>(u[0] + du[0], u[1] + du[1], u[2] + du[2], u[3] + du[3]);
    const double q = fmax(fabs(w1.r - w0.r) / w0.r, fabs(w1.p - w0.p) / w0.p);
    if (q == q) rel = fmax(rel, q); else bad += 1.0;
  }
  rel = wave_max(rel);
  bad = wave_sum(bad);
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    smax[threadIdx.x / WAVE] = rel;
    scnt[threadIdx.x / WAVE] = bad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double mx = 0.0, ct = 0.0;
#pragma unroll
    for (int q = 0; q < NT / WAVE; ++q) {
      mx = fmax(mx, smax[q]);
      ct += scnt[q];
    }
    part_at(cd, lb, 0) = mx;
    part_at(cd, lb, 1) = ct;
  }
}

// One wavefront per case: omega = min(1, max_update / (rel + 1e-30)); accept
// the step when rn, rel and dU are finite (aero2d/implicit.py:190, 218-219).
// max_update is the case's own bound when set (sc->max_update > 0), else the
// pool's.
__global__ void __launch_bounds__(WAVE) k_limiter(Plan pc, KnobsDev kn) {
  const CaseDev& cd = pc.cases[pc.slots[blockIdx.x]];
  CaseScal* sc = cd.sc;
  const double mx = case_max(cd, cd.nvecb, 0);
  const double ct = case_sum(cd, cd.nvecb, 1);
  if (threadIdx.x != 0) return;
  sc->rel = mx;
  sc->nonfinite = int(ct);
  const double mu = sc->max_update > 0.0 ? sc->max_update : kn.max_update;
  const double om = fmin(1.0, mu / (mx + 1e-30));
  sc->omega = om;
  sc->accept = (ct == 0.0) && isfinite(mx) && isfinite(sc->rn) && !sc->frozen && !sc->parked;
}

// Cells of U + omega dU with rho <= 0 or p <= 0 (or not finite), per block,
// for accepted cases.
__global__ void __launch_bounds__(NT) k_poscount(Plan plan) {
  __shared__ double scnt[NT / WAVE];
  const int e = plan_entry(plan, blockIdx.x);
  const CaseDev& cd = plan.cases[plan.slots[e]];
  const CaseScal* sc = cd.sc;
  const int lb = blockIdx.x - plan.blk_prefix[e];
  double bad = 0.0;
  if (sc->accept) {
    const double om = sc->omega;
    const int n = cd.n;
    const gptr<const double> U = gp(cd.U), dU = gp(cd.dU);
    for (int m = 0; m < PER; ++m) {
      const int c = lb * VEC_CHUNK + threadIdx.x + m * NT;
      if (c >= n) continue;
      double u[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) u[v] = U[(size_t)v * n + c] + om * dU[(size_t)v * n + c];
      const W4<double> w = cons_to_prim<double>(u[0], u[1], u[2], u[3]);
      if (!(w.r > 0.0) || !(w.p > 0.0)) bad += 1.0;
    }
  }
  bad = wave_sum(bad);
  if ((threadIdx.x & (WAVE - 1)) == 0) scnt[threadIdx.x / WAVE] = bad;
  __syncthreads();
  if (threadIdx.x == 0) {
    double ct = 0.0;
#pragma unroll
    for (int q = 0; q < NT / WAVE; ++q) ct += scnt[q];
    part_at(cd, lb, 0) = ct;
  }
}

// One wavefront per case: reject the step when any cell would be non-positive.
__global__ void __launch_bounds__(WAVE) k_posreject(Plan pc) {
  const CaseDev& cd = pc.cases[pc.slots[blockIdx.x]];
  CaseScal* sc = cd.sc;
  const double ct = case_sum(cd, cd.nvecb, 0);
  if (threadIdx.x == 0 && ct > 0.0) {
    sc->accept = 0;
    sc->nonfinite += int(ct);
  }
}

// U += omega dU for accepted cases.
__global__ void __launch_bounds__(NT) k_apply(Plan plan) {
  const int e = plan_entry(plan, blockIdx.x);
  const CaseDev& cd = plan.cases[plan.slots[e]];
  const CaseScal* sc = cd.sc;
  if (!sc->accept) return;
  const double om = sc->omega;
  const int n = cd.n;
  const int lb = blockIdx.x - plan.blk_prefix[e];
  for (int m = 0; m < PER; ++m) {
    const int c = lb * VEC_CHUNK + threadIdx.x + m * NT;
    if (c >= n) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) cd.U[(size_t)v * n + c] += om * cd.dU[(size_t)v * n + c];
  }
}

// Start of a Newton step: per-case CFL from the host, flags reset.
__global__ void k_set_cfl(Plan pc, const double* cfl) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= pc.nb) return;
  CaseScal* sc = pc.cases[pc.slots[e]].sc;
  sc->cfl = cfl[e];
  sc->active = !sc->parked;
  sc->run = 0;
  sc->kdone = 0;
  sc->accept = 0;
  sc->nonfinite = 0;
  sc->dcirc = 0.0;
}

// Per-case results of a step, 10 doubles per batch entry.
__global__ void k_gather(Plan pc, double* out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= pc.nb) return;
  const CaseScal* sc = pc.cases[pc.slots[e]].sc;
  double* o = out + (size_t)e * 10;
  o[0] = sc->rn;
  o[1] = sc->omega;
  o[2] = sc->accept;
  o[3] = sc->cl;
  o[4] = sc->cd;
  o[5] = sc->cm;
  o[6] = sc->kdone;
  o[7] = sc->rel;
  o[8] = sc->beta;
  o[9] = sc->gres;
}

template <class T>
__global__ void __launch_bounds__(NT) k_copy(Plan plan, int src, int dst) {
  const int e = plan_entry(plan, blockIdx.x);
  const CaseDev& cd = plan.cases[plan.slots[e]];
  const int n = cd.n;
  const int lb = blockIdx.x - plan.blk_prefix[e];
  const T* a = vec_ptr<T>(cd, src);
  T* b = vec_ptr<T>(cd, dst);
  for (int m = 0; m < PER; ++m) {
    const int c = lb * VEC_CHUNK + threadIdx.x + m * NT;
    if (c >= n) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) b[(size_t)v * n + c] = a[(size_t)v * n + c];
  }
}

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

}  // namespace

void launch_dots(const Plan& pv, int nblocks, int k, int wvec, int gate, int prec, int bh, hipStream_t s) {
  if (prec == 0) hipLaunchKernelGGL((k_dots<double, false>), dim3(nblocks), dim3(NT), 0, s, pv, k, wvec, gate);
  else if (bh) hipLaunchKernelGGL((k_dots<float, true>), dim3(nblocks), dim3(NT), 0, s, pv, k, wvec, gate);
  else hipLaunchKernelGGL((k_dots<float, false>), dim3(nblocks), dim3(NT), 0, s, pv, k, wvec, gate);
}
void launch_reduce_dots(const Plan& pc, int k, int bh, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_dots, dim3(pc.nb), dim3(NT), 0, s, pc, k, bh);
}
void launch_update(const Plan& pv, int nblocks, int k, int wvec, int gate, int red, int prec, int bh,
                   hipStream_t s) {
  if (prec == 0)
    hipLaunchKernelGGL((k_update<double, false>), dim3(nblocks), dim3(NT), 0, s, pv, k, wvec, gate, red);
  else if (bh)
    hipLaunchKernelGGL((k_update<float, true>), dim3(nblocks), dim3(NT), 0, s, pv, k, wvec, gate, red);
  else
    hipLaunchKernelGGL((k_update<float, false>), dim3(nblocks), dim3(NT), 0, s, pv, k, wvec, gate, red);
}
void launch_normalize(const Plan& pc, int k, const KnobsDev& kn, int prec, hipStream_t s) {
  if (prec == 0) hipLaunchKernelGGL(k_normalize<double>, dim3(pc.nb), dim3(WAVE), 0, s, pc, k, kn);
  else hipLaunchKernelGGL(k_normalize<float>, dim3(pc.nb), dim3(WAVE), 0, s, pc, k, kn);
}
void launch_krylov_begin(const Plan& pc, int stage, const KnobsDev& kn, int prec, hipStream_t s) {
  if (prec == 0)
    hipLaunchKernelGGL(k_krylov_begin<double>, dim3(pc.nb), dim3(WAVE), 0, s, pc, stage, kn);
  else
    hipLaunchKernelGGL(k_krylov_begin<float>, dim3(pc.nb), dim3(WAVE), 0, s, pc, stage, kn);
}
void launch_lstsq(const Plan& pc, hipStream_t s) {
  hipLaunchKernelGGL(k_lstsq, dim3(pc.nb), dim3(WAVE), 0, s, pc);
}
void launch_dupdate(const Plan& pv, int nblocks, int prec, int bh, int split, hipStream_t s) {
  if (prec == 0) hipLaunchKernelGGL((k_dupdate<double, false>), dim3(nblocks), dim3(NT), 0, s, pv, split);
  else if (bh) hipLaunchKernelGGL((k_dupdate<float, true>), dim3(nblocks), dim3(NT), 0, s, pv, split);
  else hipLaunchKernelGGL((k_dupdate<float, false>), dim3(nblocks), dim3(NT), 0, s, pv, split);
}
void launch_store_v0(const Plan& pv, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_store_v0, dim3(nblocks), dim3(NT), 0, s, pv);
}
void launch_limiter(const Plan& pc, const KnobsDev& kn, hipStream_t s) {
  hipLaunchKernelGGL(k_limiter, dim3(pc.nb), dim3(WAVE), 0, s, pc, kn);
}
void launch_positivity(const Plan& pv, int nblocks, const Plan& pc, hipStream_t s) {
  hipLaunchKernelGGL(k_poscount, dim3(nblocks), dim3(NT), 0, s, pv);
  hipLaunchKernelGGL(k_posreject, dim3(pc.nb), dim3(WAVE), 0, s, pc);
}
void launch_apply(const Plan& pv, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_apply, dim3(nblocks), dim3(NT), 0, s, pv);
}
void launch_set_cfl(const Plan& pc, const double* cfl, hipStream_t s) {
  hipLaunchKernelGGL(k_set_cfl, dim3(cdiv(pc.nb, 64)), dim3(64), 0, s, pc, cfl);
}
void launch_gather(const Plan& pc, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_gather, dim3(cdiv(pc.nb, 64)), dim3(64), 0, s, pc, out);
}
void launch_copy_vec(const Plan& pv, int nblocks, int src, int dst, int prec, hipStream_t s) {
  if (prec == 0) hipLaunchKernelGGL(k_copy<double>, dim3(nblocks), dim3(NT), 0, s, pv, src, dst);
  else hipLaunchKernelGGL(k_copy<float>, dim3(nblocks), dim3(NT), 0, s, pv, src, dst);
}

}  // namespace a2d
