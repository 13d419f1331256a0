// Device-side case descriptors, batch plans and small helpers shared by every
// kernel of the implicit Euler path.
//
// HBM layout of one case (ni cells around, nc cells wall-normal, n = ni*nc):
// every field is structure-of-arrays with i (around the contour) fastest,
// [var][j][i]. A wavefront then walks 64 neighbouring i-lines in lock step,
// which is what the j-marching line solves need for coalesced loads.
//   U, R, dU              double [4][nc][ni]
//   vol                   double [nc][ni]
//   Si                    double [2][nc][ni]     face between cells i-1 and i
//   Sj                    double [2][nc+1][ni]   face j; j = 0 wall, j = nc far
//   xf, yf, xw, yw        double [ni]            outer / wall face midpoints
//   Ut, volt, Sit, Sjt... T      copies in the inner precision (alias the
//                                double arrays when T = double)
//   blocks                T [5][nc][ni][16]       cell-major 4x4 blocks, five slots
//   V                     T [K+1][4][nc][ni]      raw Krylov basis
// The explicit scheme (FP64 cases only) reuses the Newton buffers: dU holds
// the Runge-Kutta stage's starting state, w the j-smoothed residual, xa the
// FAS forcing and xb the restricted state of a coarse level, diag vol / dt.
// Consecutive block slots sit 16 pstride elements apart and basis vectors
// vstride elements apart: the sizes plus a pad, so that the many arrays a
// kernel walks in step do not all start at the same offset modulo a large
// power of two.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace a2d {

constexpr int KMAX = 64;          // largest Krylov dimension a case may use
constexpr int WAVE = 64;

// Per-case scalars kept in HBM next to the case's arrays. Everything the
// Newton step decides per case lives here, so a whole step can be launched
// (or replayed from a graph) without host round trips.
struct CaseScal {
  double winf[4];      // rho, u, v, p of the freestream
  double cfl;          // pseudo-time CFL for this Newton step
  double circ;         // Gamma = Cl |V_inf| / 2 at the step's base state
  double cl, cd, cm;
  double dcirc;        // tangent of Gamma for the vector in flight
  double rn;           // sqrt(mean((R_rho / vol)^2))
  double rel;          // max relative change of rho, p under the full step
  double omega;        // step limiter
  double max_update;   // this case's bound on the relative change of rho, p (0: the pool's knob)
  double beta;         // |M(b - A x0)|
  double mbnorm;       // |M b|
  double bnorm;        // |b|
  double gres;         // GMRES residual estimate of the current iterate
  double vn0;          // |w| before orthogonalisation
  double vsig;         // FP16 basis: power-of-two storage scale of V_0
  double xres;         // explicit scheme: sqrt(mean(((U - U0)_rho / dt)^2)) of the last cycle
  int run;             // the Krylov cycle runs (|M(b - A x0)| > ptol)
  int active;          // still taking Arnoldi steps
  int kdone;           // Arnoldi steps completed
  int accept;          // the Newton update was applied
  int nonfinite;       // count of non-finite entries in the update
  int frozen;          // host says: do not update this case
  int parked;          // host says: this case stays in its batch but takes no step
  int pad;
  double s[KMAX + 1];          // scale of raw basis vector j (V_true = s_j V_raw)
  double H[KMAX + 1][KMAX];    // Hessenberg, H[row][col]
  double gc[KMAX], gs[KMAX];   // Givens rotations
  double g[KMAX + 1];          // rotated right-hand side
  double y[KMAX];              // least-squares solution
};

struct CaseDev {
  int ni, nc, n;        // n = ni * nc
  int prec;             // 0 = fp64 inner, 1 = fp32 inner
  int kmax;             // Krylov dimension allocated
  int npart;            // partial-sum slots available per reduction
  int ntiles;           // stencil workgroups (= residual-norm partials) of this case's tile shape
  int nvecb;            // vector-kernel blocks (VEC_CHUNK cells each)
  int nlineb;           // line-factorisation blocks (LINES_PER_BLOCK lines each)
  int nlinew;           // 8-line groups (one wavefront of a sweep each): norm partials
  int nlinerw;          // 8-line groups of one colour (red-black half sweeps)
  int shape;            // stencil tile shape index
  int vstride;          // elements between consecutive Krylov basis vectors (>= 4n)
  int pstride;          // cells per block slot (>= n); a slot is 16 pstride elements
  double* U;
  double* R;
  double* dU;
  const double* vol;
  const double* Si;
  const double* Sj;
  const double* xf;
  const double* yf;
  const double* xw;
  const double* yw;
  void* Ut;             // inner-precision copies (T*)
  void* volt;
  void* Sit;
  void* Sjt;
  void* xft;
  void* yft;
  void* diag;           // T [nc][ni]: vol / dtau
  void* blk;            // T [5][pstride][16] (cell-major blocks)
  void* V;              // T [kmax+1][4][nc][ni]
  void* x0;             // T [4][nc][ni]
  void* w;              // T [4][nc][ni]
  void* d;              // T [cell][4] line-solve values of the way in
  void* g0;             // T [cell][4] S0 src of the current preconditioner application
  void* hb;             // mixed form: _Float16 [4][pstride][16], slots 1-4 of blk in FP16
  void* xa;             // T [4][nc][ni] sweep ping-pong
  void* xb;
  void* p;              // mixed form, FP16 basis: FP32 preconditioner output [4][nc][ni]
  void* vc;             // and the FP32 copy of the newest (rounded) basis vector
  double* part;         // [kmax + 2][npart] reduction partials (column-major)
  double* irs;          // explicit scheme: [2][nc] elimination coefficients of the j smoothing
  CaseScal* sc;
};

// A batch of cases and the block ranges of one kernel family over it.
// blk_prefix[c] is the first block of batch entry c; blk_prefix[nb] the total.
struct Plan {
  const CaseDev* cases;   // device array indexed by slot
  const int* slots;       // nb slot indices
  const int* blk_prefix;  // nb + 1
  int nb;
  int lines = 0;          // line sweeps: lines per block (8 or 32)
  int dsteps = 0;         // line sweeps: d values per lane kept in LDS (0: in HBM)
};

// Map a flat block index to (batch entry, local block) by binary search over
// the prefix array; every lane of the block computes the same answer.
__device__ __forceinline__ int plan_entry(const Plan& p, int b) {
  int lo = 0, hi = p.nb;  // invariant: prefix[lo] <= b < prefix[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (p.blk_prefix[mid] <= b) lo = mid; else hi = mid;
  }
  return lo;
}

template <class T> __device__ __forceinline__ T* tp(void* p) { return reinterpret_cast<T*>(p); }
template <class T> __device__ __forceinline__ const T* ctp(const void* p) { return reinterpret_cast<const T*>(p); }

// Global-address-space views of device pointers. Pointers read out of the
// case descriptors are generic to the compiler, which then emits flat memory
// instructions; those share the wait counter of LDS and cross-lane traffic,
// so a shuffle would also wait for every outstanding prefetch. Casting to the
// global address space gives global_load/global_store instead.
#if defined(__HIP_DEVICE_COMPILE__)
#define A2D_GLOBAL __attribute__((address_space(1)))
#else
#define A2D_GLOBAL
#endif
template <class T> using gptr = A2D_GLOBAL T*;
// A pointer every lane of the wavefront holds, moved to scalar registers, so
// that loads can take it as their scalar base with a 32-bit vector offset.
template <class T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
  const unsigned long long v = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (T*)(((unsigned long long)hi << 32) | lo);
}
template <class T> __device__ __forceinline__ gptr<T> gp(T* p) { return (gptr<T>)p; }
// Element idx of base with the byte offset formed in 32 bits (a case's arrays
// are far below 4 GB): with a wave-uniform base the load then takes the base
// from scalar registers and a single 32-bit vector offset.
template <class T>
__device__ __forceinline__ gptr<T> at(gptr<T> base, unsigned idx) {
  return (gptr<T>)((A2D_GLOBAL char*)base + idx * (unsigned)sizeof(T));
}
template <class T> __device__ __forceinline__ gptr<const T> gp(const T* p) { return (gptr<const T>)p; }

// Value of x held by lane C of this lane's quad (DPP quad permutation: a VALU
// operation, no LDS traffic). A quad permutation always reads a lane of the
// same quad, so no "old" value is needed (mov_dpp, bound_ctrl set): this
// saves the zeroing move that update_dpp(0, ...) puts before every broadcast.
template <int C>
__device__ __forceinline__ double quad_bcast(double x) {
  constexpr int ctrl = C | (C << 2) | (C << 4) | (C << 6);
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), ctrl, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), ctrl, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}
template <int C>
__device__ __forceinline__ float quad_bcast(float x) {
  constexpr int ctrl = C | (C << 2) | (C << 4) | (C << 6);
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), ctrl, 0xF, 0xF, true));
}

// Wavefront sum with cross-lane exchanges (fixed order, deterministic).
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, WAVE);
  return x;
}
__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, WAVE));
  return x;
}

// Reductions over the first nblk partial rows of a case's partial array
// (column col), by one wavefront: lane l takes rows l, l + 64, ... in order,
// then a fixed butterfly. Every lane gets the result; it depends on the case
// alone, never on the batch it runs in.
// Partial col of block blk. Columns are contiguous, so the lanes of a case sum
// read consecutive blocks' partials of one column (one or two cache lines per
// load instead of one line per lane).
__device__ __forceinline__ double& part_at(const CaseDev& cd, int blk, int col) {
  return cd.part[(size_t)col * cd.npart + blk];
}
__device__ __forceinline__ double case_sum(const CaseDev& cd, int nblk, int col) {
  const int lane = threadIdx.x & (WAVE - 1);
  double s = 0.0;
  for (int b = lane; b < nblk; b += WAVE) s += part_at(cd, b, col);
  return wave_sum(s);
}
__device__ __forceinline__ double case_max(const CaseDev& cd, int nblk, int col) {
  const int lane = threadIdx.x & (WAVE - 1);
  double s = 0.0;
  for (int b = lane; b < nblk; b += WAVE) s = fmax(s, part_at(cd, b, col));
  return wave_max(s);
}

// Named per-case vectors (ids >= 0 select raw Krylov basis vector V_id).
enum VecId : int {
  VEC_X0 = -1, VEC_W = -2, VEC_D = -3, VEC_XA = -4, VEC_XB = -5, VEC_NEG_R = -6, VEC_P = -7, VEC_VC = -8
};

template <class T>
__device__ __forceinline__ T* vec_ptr(const CaseDev& cd, int id);
template <class T>
__device__ __forceinline__ gptr<T> vec_gptr(const CaseDev& cd, int id) { return gp(vec_ptr<T>(cd, id)); }

template <class T>
__device__ __forceinline__ T* vec_ptr(const CaseDev& cd, int id) {
  switch (id) {
    case VEC_X0: return tp<T>(cd.x0);
    case VEC_W: return tp<T>(cd.w);
    case VEC_D: return tp<T>(cd.d);
    case VEC_XA: return tp<T>(cd.xa);
    case VEC_XB: return tp<T>(cd.xb);
    case VEC_P: return tp<T>(cd.p);
    case VEC_VC: return tp<T>(cd.vc);
    default: return tp<T>(cd.V) + (size_t)id * cd.vstride;
  }
}

// Offset of entry m (= row * 4 + column) of block `slot` of cell `cell` in the
// preconditioner's cell-major block storage T blk[5][pstride][16].
__host__ __device__ __forceinline__ size_t blk_off(int slot, int m, size_t cell, size_t pstride) {
  return (size_t(slot) * pstride + cell) * 16 + m;
}

// Column of line i in the preconditioner's block storage and in the Krylov
// vectors (x0, w, basis, P, vc, xa, xb). With the red-black line order both
// are kept colour-split (even lines, then odd lines, in every row) so that a
// half sweep reads its blocks, its own lines' vector values and the other
// colour's neighbours as contiguous runs and writes whole cache lines;
// otherwise column = i. The state, residual, update and vol/dtau keep the
// natural order.
__host__ __device__ __forceinline__ int bcol(int i, int ni, int split) {
  return split ? (i & 1) * (ni >> 1) + (i >> 1) : i;
}

// Positive modulo for the periodic i direction.
__device__ __forceinline__ int wrap(int i, int ni) {
  i %= ni;
  return i < 0 ? i + ni : i;
}

}  // namespace a2d
