// End of one Arnoldi step (aero2d/implicit.py:184-191 through the batched
// GMRES of jax.scipy.sparse.linalg): norm of the orthogonalised vector,
// breakdown test with threshold eps * |w| as in JAX, Givens update of the
// Hessenberg column and of the residual estimate, optional early exit.
//
// Shared by the stand-alone finishing kernel (last step of a restart cycle)
// and by the far-field circulation kernel that opens the next step, which
// finishes the previous step first so that no launch of its own is spent on
// it. Called by every thread of a block of any size; every thread must reach
// it (it synchronises the block).
#pragma once
#include "common.hpp"
#include "launch.hpp"

#include <cfloat>
#include <type_traits>

namespace a2d {

template <class T> __device__ __forceinline__ double eps_of() {
  return std::is_same<T, double>::value ? DBL_EPSILON : double(FLT_EPSILON);
}

// Partial column of the orthogonalised vector's |w|^2 per vector block: the
// last one, which the dot products (columns 0..k+1 <= kmax) never use.
__device__ __forceinline__ int update_norm_col(const CaseDev& cd) { return cd.kmax + 1; }

// LDS staging for arnoldi_finish: the column and the rotations, so that the
// sequential rotation chain never waits on memory; live = the case goes on to
// step k + 1 (valid for every thread on return).
struct ArnoldiLds {
  double h[KMAX + 1], gcs[KMAX], gsn[KMAX];
  int live;
};

// Finish step k of the case whose scalars the caller found active.
template <class T>
__device__ void arnoldi_finish(const CaseDev& cd, int k, const KnobsDev& kn, ArnoldiLds& L) {
  CaseScal* sc = cd.sc;
  const int t = threadIdx.x, nt = blockDim.x;
  // every wave forms the same sum; thread 0 uses wave 0's
  const double nrm = sqrt(case_sum(cd, cd.nvecb, update_norm_col(cd)));
  for (int i = t; i <= k; i += nt) L.h[i] = sc->H[i][k];
  for (int i = t; i < k; i += nt) {
    L.gcs[i] = sc->gc[i];
    L.gsn[i] = sc->gs[i];
  }
  __syncthreads();
  if (t == 0) {
    // FP16 basis: nrm is the norm of the stored sigma-scaled vector
    const double sig = (std::is_same<T, float>::value && kn.basis_half) ? sc->vsig : 1.0;
    const double hn = nrm / sig;
    const bool brk = !(hn > eps_of<T>() * sc->vn0);
    const double hk1 = brk ? 0.0 : hn;
    sc->s[k + 1] = brk ? 0.0 : 1.0 / nrm;
    // apply the previous rotations to column k; `a` carries H[i][k]
    double a = L.h[0];
    for (int i = 0; i < k; ++i) {
      const double b = L.h[i + 1];
      L.h[i] = L.gcs[i] * a + L.gsn[i] * b;
      a = -L.gsn[i] * a + L.gcs[i] * b;
    }
    const double r = hypot(a, hk1);
    const double c = (r > 0) ? a / r : 1.0, sn = (r > 0) ? hk1 / r : 0.0;
    sc->gc[k] = c;
    sc->gs[k] = sn;
    L.h[k] = r;
    sc->H[k + 1][k] = 0.0;
    sc->g[k + 1] = -sn * sc->g[k];
    sc->g[k] = c * sc->g[k];
    sc->gres = fabs(sc->g[k + 1]);
    sc->kdone = k + 1;
    bool stop = brk || (k + 1 >= kn.restart);
    if (kn.early_exit && sc->gres <= fmin(1.0, kn.gtol) * sc->mbnorm) stop = true;
    if (stop) sc->active = 0;
    L.live = !stop;
  }
  __syncthreads();
  for (int i = t; i <= k; i += nt) sc->H[i][k] = L.h[i];
}

}  // namespace a2d
