// Host engine of the implicit Euler library: pools of device-resident cases,
// batch plans, the launch sequence of one Newton step (optionally captured
// into a HIP graph and replayed), and the C ABI declared in
// include/aero2d_hip.h.
#include "../../include/aero2d_hip.h"
#include "common.hpp"
#include "launch.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace a2d {
namespace {

thread_local std::string g_err;
int g_device = -1;
int g_num_cu = 256;   // compute units of the device (sweep block sizing)
int g_lds_block = 65536, g_lds_cu = 65536;   // LDS bytes per workgroup / per CU
// 32-line sweep blocks from this many per 100 CUs, else 8-line blocks.
// Measured crossover (preconditioner application, 512x256 mixed / 256x128
// FP64): 8-line blocks win at 128 32-line blocks (16 cases, -15..-22 %),
// 32-line blocks from 192 on (24 cases and more, -3..-12 %).
#ifndef A2D_SWEEP_BIG_PCT
#define A2D_SWEEP_BIG_PCT 60
#endif
// The CGS update forms the dot products' sums itself (no separate reduction
// launch) up to this many of its blocks per CU: 2 = all resident at once at
// its register use. 0 keeps the separate launch always.
#ifndef A2D_UPDATE_FUSE_PER_CU
#define A2D_UPDATE_FUSE_PER_CU 2
#endif
std::recursive_mutex g_mu;

struct Fail {
  int code;
};

void fail(int code, const std::string& msg) {
  g_err = msg;
  throw Fail{code};
}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(-10, std::string(what) + ": " + hipGetErrorString(e));
}
#define HC(x) hip_check((x), #x)

constexpr int MAX_SLOTS = 4096;
constexpr size_t ALIGN = 256;

size_t align_up(size_t x) { return (x + ALIGN - 1) / ALIGN * ALIGN; }
int cdiv(int a, int b) { return (a + b - 1) / b; }

KnobsDev to_dev(const a2d_knobs& k) {
  KnobsDev d;
  d.cfl_explicit = k.cfl_explicit;
  d.lim_k = k.lim_k;
  d.max_update = k.max_update;
  d.gtol = k.gmres_tol;
  d.vortex = k.vortex;
  d.order = k.order;
  d.early_exit = k.early_exit;
  d.restart = k.gmres_restart;
  d.precond_rb = k.precond_order == 1;
  d.positivity = k.positivity;
  d.precond_half = k.precond_half;
  d.basis_half = k.basis_half;
  return d;
}

// Byte layout of one case's slab; offsets of every sub-buffer.
struct Layout {
  size_t U, R, dU, vol, Si, Sj, xf, yf, xw, yw;
  size_t Ut, volt, Sit, Sjt, xft, yft;
  size_t diag, blk, hb, V, x0, w, d, g0, xa, xb, pv, vc, part, cp, irs, sc;
  size_t total;
  int npart;
};

// Residual-norm partials of the largest stencil launch (any tile shape, or
// the column march).
int max_tiles(int ni, int nc) {
  int m = march_units(ni, nc, MARCH_ROWS_RES);
  for (int s = 0; s < stencil_num_shapes(); ++s) m = std::max(m, stencil_blocks(s, ni, nc));
  return m;
}

#ifndef A2D_PAD_BYTES
#define A2D_PAD_BYTES 0
#endif
// Padded strides of the block planes and basis vectors (elements).
size_t plane_stride(size_t n, size_t t) { return n + A2D_PAD_BYTES / t; }
size_t vec_stride(size_t n, size_t t) { return 4 * n + A2D_PAD_BYTES / t; }

// Kernels form element offsets inside a case's arrays in 32-bit unsigned
// arithmetic (byte offsets from a wave-uniform base). The largest such array is
// the preconditioner's block storage, 80 * pstride elements of the inner type:
// a case whose block storage reaches 4 GiB is refused (FP64: about 6.7 M
// cells, mixed: about 13.4 M cells).
constexpr uint64_t ADDR32_LIMIT = uint64_t(1) << 32;
uint64_t addr32_bytes(int ni, int nc, int prec) {
  const size_t t = prec == 0 ? 8 : 4;
  return uint64_t(80) * plane_stride(size_t(ni) * size_t(nc), t) * t;
}
bool addr32_ok(int ni, int nc, int prec) { return addr32_bytes(ni, nc, prec) < ADDR32_LIMIT; }

Layout layout(int ni, int nc, int prec, int kmax) {
  Layout L{};
  const size_t n = size_t(ni) * nc;
  const size_t t = prec == 0 ? 8 : 4;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    size_t at = o;
    o = align_up(o + bytes);
    return at;
  };
  L.U = take(4 * n * 8);
  L.R = take(4 * n * 8);
  L.dU = take(4 * n * 8);
  L.vol = take(n * 8);
  L.Si = take(2 * n * 8);
  L.Sj = take(2 * (n + ni) * 8);
  L.xf = take(ni * 8);
  L.yf = take(ni * 8);
  L.xw = take(ni * 8);
  L.yw = take(ni * 8);
  if (prec == 1) {
    L.Ut = take(4 * n * 4);
    L.volt = take(n * 4);
    L.Sit = take(2 * n * 4);
    L.Sjt = take(2 * (n + ni) * 4);
    L.xft = take(ni * 4);
    L.yft = take(ni * 4);
  }
  L.diag = take(n * t);
  L.blk = take(80 * plane_stride(n, t) * t);
  if (prec == 1) L.hb = take(64 * plane_stride(n, t) * 2);
  L.V = take(size_t(kmax + 1) * vec_stride(n, t) * t);
  L.x0 = take(4 * n * t);
  L.w = take(4 * n * t);
  L.d = take(4 * n * t);
  L.g0 = take(4 * n * t);
  L.xa = take(4 * n * t);
  L.xb = take(4 * n * t);
  if (prec == 1) {
    L.pv = take(4 * n * 4);
    L.vc = take(4 * n * 4);
  }
  // partial rows: stencil/march units, vector blocks, sweep groups, and the
  // explicit scheme's per-row residual partials (nc)
  L.npart = std::max({max_tiles(ni, nc), cdiv(int(n), VEC_CHUNK), 2 * cdiv(ni, 8) + 2, nc, 1});
  L.part = take(size_t(L.npart) * (kmax + 2) * 8);
  L.cp = take(size_t(ni) * 8);
  L.irs = take(size_t(2) * nc * 8);
  L.sc = take(sizeof(CaseScal));
  L.total = o;
  return L;
}

struct CaseHost {
  bool live = false;
  int ni = 0, nc = 0, prec = 0, kmax = 0;
  char* slab = nullptr;
  size_t bytes = 0;     // the slab's capacity (>= the layout's total)
  double* cp = nullptr;  // wall Cp buffer (ni) for a2d_forces
  int coarse = -1;       // explicit scheme: the case of the next coarser grid level
  double irs_eps = -1;   // smoothing coefficient the case's irs table was formed for
  bool parked = false;
  CaseDev dev{};
  double winf[4];
};

// Device buffers of released plan tables, by capacity in ints, for reuse.
using TableCache = std::multimap<size_t, int*>;

struct BatchPlan {
  std::vector<int> slots;
  int prec = 0, nb = 0;
  int* dbuf = nullptr;     // prefix tables (from / back to the pool's table cache)
  size_t dbuf_cap = 0;
  TableCache* tables = nullptr;
  double* dcfl = nullptr;  // the pool's step buffers (shared by its plans)
  double* dout = nullptr;
  double* hcfl = nullptr;
  double* hout = nullptr;
  Plan pc{}, pcell{}, pvec{}, pline{}, psw{}, prb{};
  int ncell = 0, nvec = 0, nline = 0, nsw = 0, nrb = 0;
  // column-march stencil: residual (pmr) and tangent (pmj) block ranges, and
  // the batch's unit counts that decide whether it is used
  Plan pmr{}, pmj{};
  int nmr = 0, nmj = 0;
  // explicit scheme: rows (four per block) and columns (256 per block, or 32
  // per block for the segmented j smoothing)
  Plan prow{}, pcol{}, pseg{};
  int nrow = 0, ncol = 0, nseg = 0, ni_max = 0, nc_max = 0;
  struct Group {
    Plan p;
    int shape, nblocks;
  };
  std::vector<Group> sten;
  hipGraphExec_t exec = nullptr;
  int uses = 0;
  uint64_t knob_version = 0;
  ~BatchPlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (dbuf) tables->emplace(dbuf_cap, dbuf);
  }
};

// Counters of a pool's work (a2d_pool_stats).
struct PoolStats {
  int64_t steps = 0, replays = 0, captures = 0, plans = 0, tables_new = 0;
  int64_t slabs_new = 0, slabs_reused = 0;
};

struct Pool {
  int64_t id = 0;
  hipStream_t stream = nullptr;
  int64_t budget = 0, used = 0;
  std::vector<CaseHost> cases;
  CaseDev* dcases = nullptr;
  a2d_knobs knobs{};
  uint64_t knob_version = 1;
  // released case slabs kept for reuse, by capacity (bytes); with a budget,
  // used + cached stays within it
  std::multimap<size_t, char*> free_slabs;
  int64_t cached = 0;
  TableCache tables;
  // step buffers shared by every plan of the pool (one step runs at a time):
  // per-case CFL in, 10 results per case out; pinned host staging
  double* dcfl = nullptr;
  double* dout = nullptr;
  double* hcfl = nullptr;
  double* hout = nullptr;
  PoolStats stats;
  // declared last: its plans return their tables to `tables` when destroyed
  std::map<std::vector<int>, std::unique_ptr<BatchPlan>> plans;
  std::vector<double> stage_d;
  std::vector<float> stage_f;
};

std::map<int64_t, std::unique_ptr<Pool>> g_pools;
int64_t g_next_pool = 1;

// Staging memory of the allocation-free HLLC entry (a2d_hllc_flux_noalloc):
// WL, WR (4 n each), S (2 n), F (4 n) for at most NOALLOC_MAX_FACES faces,
// 7.3 MB held on every device the library runs on until the process ends.
// Host code reaches it only through the per-device address from
// hipGetSymbolAddress (the array's name is not a device address on the host).
constexpr int NOALLOC_MAX_FACES = 1 << 16;
static __device__ __align__(256) double scratch[14 * NOALLOC_MAX_FACES];
std::map<int, double*> g_scratch;   // device -> address of its copy of `scratch`

void default_knobs(a2d_knobs* k) {
  k->cfl_explicit = 2.5;
  k->lim_k = 0.02;
  k->max_update = 0.2;
  k->gmres_tol = 0.01;
  k->vortex = 1;
  k->order = 2;
  k->gmres_restart = 50;
  k->jacobi_sweeps = 4;
  k->early_exit = 0;
  k->use_graph = 1;
  k->precond_order = 0;
  k->positivity = 0;
  k->precond_half = 0;
  k->basis_half = 0;
  k->stencil_march = 1;
  k->sweep_lds = 1;
  k->irs_eps = 0.6;
}

Pool& get_pool(int64_t id) {
  auto it = g_pools.find(id);
  if (it == g_pools.end()) fail(-3, "unknown pool handle");
  return *it->second;
}

CaseHost& get_case(Pool& p, int id) {
  if (id < 0 || id >= int(p.cases.size()) || !p.cases[id].live) fail(-4, "unknown case id");
  return p.cases[id];
}

void ensure_init() {
  if (g_device < 0) fail(-1, "a2d_init has not been called");
  HC(hipSetDevice(g_device));
}

// A removed case's slab goes to the pool's cache for the next admission; with
// a budget, cached slabs are freed (largest first) so used + cached fits it.
void release_slab(Pool& p, CaseHost& c) {
  p.used -= int64_t(c.bytes);
  p.free_slabs.emplace(c.bytes, c.slab);
  p.cached += int64_t(c.bytes);
  while (p.budget > 0 && p.used + p.cached > p.budget && !p.free_slabs.empty()) {
    auto big = std::prev(p.free_slabs.end());
    (void)hipFree(big->second);
    p.cached -= int64_t(big->first);
    p.free_slabs.erase(big);
  }
  c.slab = nullptr;
}

// Every id names a live case, and no id appears twice: kernels that update a
// case in place (state update, Krylov vectors) would otherwise run twice over
// the same arrays within one launch.
void check_batch(Pool& p, const std::vector<int>& slots) {
  for (int s : slots) get_case(p, s);
  std::vector<int> sorted(slots);
  std::sort(sorted.begin(), sorted.end());
  if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
    fail(-6, "a batch must not name a case twice");
}

// ---- plans --------------------------------------------------------------------
void drop_plans_with(Pool& p, int slot) {
  for (auto it = p.plans.begin(); it != p.plans.end();) {
    if (std::find(it->first.begin(), it->first.end(), slot) != it->first.end()) it = p.plans.erase(it);
    else ++it;
  }
}

BatchPlan& get_plan(Pool& p, const std::vector<int>& slots) {
  auto it = p.plans.find(slots);
  if (it != p.plans.end()) return *it->second;
  if (p.plans.size() >= 64) p.plans.clear();
  auto bp = std::make_unique<BatchPlan>();
  bp->slots = slots;
  bp->nb = int(slots.size());
  bp->prec = p.cases[slots[0]].prec;
  for (int s : slots)
    if (p.cases[s].prec != bp->prec) fail(-5, "a batch must hold cases of one precision");
  const int nb = bp->nb;
  // host image: slots | pc | pcell | pvec | pline | per group: slots_g | prefix_g
  std::vector<int> h;
  h.insert(h.end(), slots.begin(), slots.end());
  auto add_prefix = [&](auto count) {
    size_t at = h.size();
    int acc = 0;
    for (int e = 0; e < nb; ++e) {
      h.push_back(acc);
      acc += count(p.cases[slots[e]]);
    }
    h.push_back(acc);
    return std::make_pair(at, acc);
  };
  auto pcase = add_prefix([](const CaseHost&) { return 1; });
  auto pcell = add_prefix([](const CaseHost& c) { return cdiv(c.dev.n, CELL_BLOCK); });
  auto pvec = add_prefix([](const CaseHost& c) { return c.dev.nvecb; });
  auto pline = add_prefix([](const CaseHost& c) { return c.dev.nlineb; });
  // sweep block size: 32 lines when such blocks fill the device, else 8
  int big_sw = 0, big_rb = 0;
  for (int s : slots) {
    big_sw += cdiv(p.cases[s].dev.ni, SWEEP_LINES_BIG);
    big_rb += cdiv(p.cases[s].dev.ni / 2, SWEEP_LINES_BIG);
  }
  const int big_min = g_num_cu * A2D_SWEEP_BIG_PCT / 100;
  const int lsw = big_sw >= big_min ? SWEEP_LINES_BIG : SWEEP_LINES_SMALL;
  const int lrb = big_rb >= big_min ? SWEEP_LINES_BIG : SWEEP_LINES_SMALL;
  auto psw = add_prefix([&](const CaseHost& c) { return cdiv(c.dev.ni, lsw); });
  auto prb = add_prefix([&](const CaseHost& c) { return cdiv(c.dev.ni / 2, lrb); });
  const int rows_jvp = bp->prec == 0 ? MARCH_ROWS_JVP64 : MARCH_ROWS_JVP32;
  auto prow = add_prefix([](const CaseHost& c) { return c.nc; });
  auto pcol = add_prefix([](const CaseHost& c) { return cdiv(4 * c.ni, EXPLICIT_BLOCK); });
  auto pseg = add_prefix([](const CaseHost& c) { return cdiv(4 * c.ni, IRS_J_COLS_PER_BLOCK); });
  auto pmr = add_prefix([](const CaseHost& c) { return march_blocks(c.ni, c.nc, MARCH_ROWS_RES); });
  auto pmj = add_prefix([&](const CaseHost& c) { return march_blocks(c.ni, c.nc, rows_jvp); });
  std::map<int, std::vector<int>> groups;
  for (int s : slots) groups[p.cases[s].dev.shape].push_back(s);
  struct G {
    size_t slots_at, prefix_at;
    int nb, shape, total;
  };
  std::vector<G> gs;
  for (auto& kv : groups) {
    G g;
    g.shape = kv.first;
    g.nb = int(kv.second.size());
    g.slots_at = h.size();
    h.insert(h.end(), kv.second.begin(), kv.second.end());
    g.prefix_at = h.size();
    int acc = 0;
    for (int s : kv.second) {
      h.push_back(acc);
      acc += p.cases[s].dev.ntiles;
    }
    h.push_back(acc);
    g.total = acc;
    gs.push_back(g);
  }
  bp->tables = &p.tables;
  {
    // a released table of at least this size (within 2x), else a new one
    auto it = p.tables.lower_bound(h.size());
    if (it != p.tables.end() && it->first <= 2 * h.size()) {
      bp->dbuf = it->second;
      bp->dbuf_cap = it->first;
      p.tables.erase(it);
    } else {
      HC(hipMalloc(&bp->dbuf, h.size() * sizeof(int)));
      bp->dbuf_cap = h.size();
      p.stats.tables_new++;
    }
  }
  p.stats.plans++;
  HC(hipMemcpy(bp->dbuf, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice));
  bp->dcfl = p.dcfl;
  bp->dout = p.dout;
  bp->hcfl = p.hcfl;
  bp->hout = p.hout;
  auto mk = [&](size_t prefix_at) {
    Plan q;
    q.cases = p.dcases;
    q.slots = bp->dbuf;
    q.blk_prefix = bp->dbuf + prefix_at;
    q.nb = nb;
    return q;
  };
  bp->pc = mk(pcase.first);
  bp->pcell = mk(pcell.first);
  bp->ncell = pcell.second;
  bp->pvec = mk(pvec.first);
  bp->nvec = pvec.second;
  bp->pline = mk(pline.first);
  bp->nline = pline.second;
  bp->psw = mk(psw.first);
  bp->psw.lines = lsw;
  bp->nsw = psw.second;
  bp->prb = mk(prb.first);
  bp->prb.lines = lrb;
  bp->nrb = prb.second;
  // Sweeps of 32-line blocks (batches that fill the device) keep their d
  // values in LDS when every block of the launch can then still be resident at
  // once (the LDS takes no residency from the launch): the red-black half
  // sweep of 32 transonic 512 x 256 cases 97 -> 90 us. Small batches' 8-line
  // blocks are latency bound, and measured slower with LDS (one 512 x 256
  // case: an application 111 -> 119 us), so they keep the HBM round trip.
  int mmax = 0;
  for (int s : slots) mmax = std::max(mmax, p.cases[s].nc >> 1);
  const size_t tsz = bp->prec == 0 ? 8 : 4;
  auto lds_steps = [&](int lines, int nblocks) {
    const size_t bytes = size_t(8 * lines) * size_t(mmax) * tsz;
    if (bytes == 0 || bytes > size_t(g_lds_block)) return 0;
    const size_t per_cu = size_t(g_lds_cu) / bytes;
    return size_t(nblocks) <= per_cu * size_t(g_num_cu) ? mmax : 0;
  };
  bp->psw.dsteps = lds_steps(lsw, bp->nsw);
  bp->prb.dsteps = lds_steps(lrb, bp->nrb);
  bp->prow = mk(prow.first);
  bp->nrow = prow.second;
  bp->pcol = mk(pcol.first);
  bp->ncol = pcol.second;
  bp->pseg = mk(pseg.first);
  bp->nseg = pseg.second;
  for (int s : slots) {
    bp->ni_max = std::max(bp->ni_max, p.cases[s].ni);
    bp->nc_max = std::max(bp->nc_max, p.cases[s].nc);
  }
  bp->pmr = mk(pmr.first);
  bp->nmr = pmr.second;
  bp->pmj = mk(pmj.first);
  bp->nmj = pmj.second;
  for (auto& g : gs) {
    BatchPlan::Group gr;
    gr.p.cases = p.dcases;
    gr.p.slots = bp->dbuf + g.slots_at;
    gr.p.blk_prefix = bp->dbuf + g.prefix_at;
    gr.p.nb = g.nb;
    gr.shape = g.shape;
    gr.nblocks = g.total;
    bp->sten.push_back(gr);
  }
  BatchPlan& ref = *bp;
  p.plans[slots] = std::move(bp);
  return ref;
}

// A sweep plan as the knob sweep_lds has it: 0 the d values through HBM, 2 in
// LDS wherever the plan found room, 1 (default) that only for 32-line blocks.
Plan sweep_plan(const Pool& p, Plan q) {
  const int k = p.knobs.sweep_lds;
  if (k == 0 || (k == 1 && q.lines != SWEEP_LINES_BIG)) q.dsteps = 0;
  return q;
}

// ---- launch sequences ---------------------------------------------------------
void enq_precond(Pool& p, BatchPlan& b, int src, int dest, int want_norm, int gate) {
  const int S = std::max(1, p.knobs.jacobi_sweeps);
  const int hb = b.prec == 1 && p.knobs.precond_half;
  const Plan psw = sweep_plan(p, b.psw), prb = sweep_plan(p, b.prb);
  if (p.knobs.precond_order == 1) {
    // red-black: odd lines then even lines, S times, in place in dest
    for (int s = 0; s < S; ++s)
      for (int color = 1; color >= 0; --color) {
        const int coupled = !(s == 0 && color == 1);
        launch_line_sweep_rb(prb, b.nrb, src, dest, color, coupled, s > 0,
                             (s == S - 1) ? want_norm : 0, gate, b.prec, hb, p.stream);
      }
    return;
  }
  for (int s = 0; s < S; ++s) {
    const int out = (s == S - 1) ? dest : (s % 2 == 0 ? VEC_XA : VEC_XB);
    const int prev = (s == 0) ? VEC_XB : ((s - 1) % 2 == 0 ? VEC_XA : VEC_XB);
    launch_line_sweep(psw, b.nsw, src, prev, out, s > 0, s > 0, (s == S - 1) ? want_norm : 0,
                      gate, b.prec, hb, p.stream);
  }
}

// Stencil kernel choice for a batch (knob stencil_march: 0 the LDS tiles, 2
// the column march, 1 the march when the batch has enough of its units to
// fill the device: it trades the tiles' halo work for a serial walk up each
// chunk, which wins only with several wavefronts per SIMD. Measured, 512 x 256
// cases: FP64 residual 359 -> 298 us at 32 cases (march units: 4.5 per
// SIMD), 165 -> 142 us at 16, but 51 -> 70 us at 4; the FP32 tangent (8-row
// chunks) 205 -> 169 us at 32 cases, slower at 4; the FP64 tangent is slower
// as a march at every size (its registers allow one wavefront per SIMD).
bool march_res(const Pool& p, const BatchPlan& b) {
  const int k = p.knobs.stencil_march;
  return k == 2 || (k == 1 && b.nmr * MARCH_WAVES >= 2 * 4 * g_num_cu);
}
bool march_jvp(const Pool& p, const BatchPlan& b) {
  const int k = p.knobs.stencil_march;
  return k == 2 || (k == 1 && b.prec == 1 && b.nmj * MARCH_WAVES >= 4 * 4 * g_num_cu);
}

void enq_residual(Pool& p, BatchPlan& b, const KnobsDev& kn, int want_diag, double circ, int refresh_ut,
                  bool finalize = true) {
  const bool m = march_res(p, b);
  if (m) launch_march_residual(b.pmr, b.nmr, kn, want_diag, circ, refresh_ut, p.stream);
  else
    for (auto& g : b.sten) launch_residual(g.p, g.shape, g.nblocks, kn, want_diag, circ, refresh_ut, p.stream);
  if (finalize) launch_rn_finalize(b.pc, m, p.stream);
}

void enq_jvp(Pool& p, BatchPlan& b, const KnobsDev& kn, int vin, int vout, int mode, int kidx, int use_diag,
             int gate) {
  if (march_jvp(p, b)) launch_march_jvp(b.pmj, b.nmj, kn, vin, vout, mode, kidx, use_diag, gate, b.prec, p.stream);
  else
    for (auto& g : b.sten)
      launch_jvp(g.p, g.shape, g.nblocks, kn, vin, vout, mode, kidx, use_diag, gate, b.prec, p.stream);
}

// Residual with the local time step at each case's CFL (sc->cfl), rn.
void enq_prologue(Pool& p, BatchPlan& b, const KnobsDev& kn) {
  launch_forces(b.pc, nullptr, p.stream);
  enq_residual(p, b, kn, 1, NAN, 1);
}

void enq_newton(Pool& p, BatchPlan& b) {
  const KnobsDev kn = to_dev(p.knobs);
  const int prec = b.prec;
  hipStream_t s = p.stream;
  launch_set_cfl(b.pc, b.dcfl, s);
  enq_prologue(p, b, kn);
  launch_assemble(b.pcell, b.ncell, kn, prec, s);
  launch_factor(b.pline, b.nline, prec, s);
  // x0 = M b with b = -R
  enq_precond(p, b, VEC_NEG_R, VEC_X0, 1, 0);
  launch_krylov_begin(b.pc, 0, kn, prec, s);
  // V_0 = M(b - A x0)
  launch_dcirc(b.pc, VEC_X0, 0, -1, kn, prec, s);
  enq_jvp(p, b, kn, VEC_X0, VEC_W, 1, -1, 1, 0);
  // FP16 basis (mixed form): the preconditioner writes the FP32 vector P, the
  // basis vectors are stored from it in FP16, and the operator reads their
  // rounded values in FP32 from vc
  const int bh = prec == 1 && p.knobs.basis_half;
  enq_precond(p, b, VEC_W, bh ? VEC_P : 0, 1, 0);
  launch_krylov_begin(b.pc, 1, kn, prec, s);
  if (bh) launch_store_v0(b.pvec, b.nvec, s);
  const int K = std::min(p.knobs.gmres_restart, KMAX);
  // the update forms the dot products' sums itself when all its blocks are
  // resident at once (two per compute unit at its register use)
  const int red = b.nvec <= A2D_UPDATE_FUSE_PER_CU * g_num_cu;
  for (int k = 0; k < K; ++k) {
    const int vin = bh ? VEC_VC : k, wout = bh ? VEC_P : k + 1;
    // the circulation kernel first finishes step k - 1 (a launch fewer per
    // step: a small batch's step is latency bound)
    launch_dcirc(b.pc, vin, 1, k - 1, kn, prec, s);
    enq_jvp(p, b, kn, vin, VEC_W, 0, k, 1, 1);
    enq_precond(p, b, VEC_W, wout, 0, 1);
    launch_dots(b.pvec, b.nvec, k, wout, 1, prec, bh, s);
    if (!red) launch_reduce_dots(b.pc, k, bh, s);
    launch_update(b.pvec, b.nvec, k, wout, 1, red, prec, bh, s);
  }
  if (K > 0) launch_normalize(b.pc, K - 1, kn, prec, s);
  launch_lstsq(b.pc, s);
  launch_dupdate(b.pvec, b.nvec, prec, bh, kn.precond_rb, s);
  launch_limiter(b.pc, kn, s);
  if (kn.positivity) launch_positivity(b.pvec, b.nvec, b.pc, s);
  launch_apply(b.pvec, b.nvec, s);
  launch_forces(b.pc, nullptr, s);
  launch_gather(b.pc, b.dout, s);
}

// ---- explicit scheme ------------------------------------------------------------
constexpr double RK_STAGES[4] = {0.25, 1.0 / 3.0, 0.5, 1.0};   // aero2d/euler.py:340

// Elimination coefficients of the j smoothing, (1 + 2 eps) on the diagonal,
// 1 + eps on both end rows, -eps off it (aero2d/euler.py:305-311): m_j =
// 1 / pivot_j and c_j = -eps m_j, formed in FP64 on the host.
void irs_table(Pool& p, CaseHost& c, double eps) {
  if (c.irs_eps == eps) return;
  const int nc = c.nc;
  std::vector<double> t(2 * size_t(nc));
  double* m = t.data();
  double* cc = t.data() + nc;
  for (int j = 0; j < nc; ++j) {
    const double b = (j == 0 || j == nc - 1) ? 1.0 + eps : 1.0 + 2.0 * eps;
    m[j] = 1.0 / (j == 0 ? b : b + eps * cc[j - 1]);
    cc[j] = -eps * m[j];
  }
  HC(hipMemcpyAsync(c.dev.irs, t.data(), t.size() * 8, hipMemcpyHostToDevice, p.stream));
  HC(hipStreamSynchronize(p.stream));
  c.irs_eps = eps;
}

void enq_irs_j(Pool& p, BatchPlan& b, double eps, int forcing) {
  const int sl = irs_j_seg_len(b.nc_max);
  if (sl > 0) launch_irs_j_seg(b.pseg, b.nseg, sl, eps, forcing, p.stream);
  else launch_irs_j(b.pcol, b.ncol, eps, forcing, p.stream);
}

// One Runge-Kutta iteration on level l (aero2d/euler.py:334-345): the local
// time step from the starting state (the first stage's residual kernel), four
// stages of residual, smoothing and update; coarse levels add their forcing
// and run first order. want_res: the res of the iteration per case.
void enq_iterate(Pool& p, std::vector<BatchPlan*>& L, int l, const KnobsDev& kn, bool want_res) {
  BatchPlan& b = *L[l];
  KnobsDev k = kn;
  if (l > 0) k.order = 1;
  const double eps = p.knobs.irs_eps;
  const int forcing = l > 0;
  const int cm = rk_rows_cm(b.ni_max);
  for (int st = 0; st < 4; ++st) {
    enq_residual(p, b, k, st == 0, NAN, 0, false);
    if (eps > 0) enq_irs_j(p, b, eps, forcing);
    launch_rk_rows(b.prow, b.nrow, cm, eps, RK_STAGES[st], st == 0, want_res && st == 3, forcing, p.stream);
  }
  if (want_res) launch_xres(b.pc, p.stream);
}

// FAS V-cycle from level l down (aero2d/euler.py:367-381).
void enq_cycle(Pool& p, std::vector<BatchPlan*>& L, int l, const KnobsDev& kn) {
  enq_iterate(p, L, l, kn, l == 0);
  if (l + 1 >= int(L.size())) return;
  BatchPlan& f = *L[l];
  BatchPlan& c = *L[l + 1];
  KnobsDev kf = kn, kc = kn;
  if (l > 0) kf.order = 1;
  kc.order = 1;
  enq_residual(p, f, kf, 0, NAN, 0, false);
  launch_restrict(c.pcell, c.ncell, f.pc.slots, l > 0, p.stream);
  enq_residual(p, c, kc, 0, NAN, 0, false);
  launch_forcing(c.pcell, c.ncell, p.stream);
  enq_cycle(p, L, l + 1, kn);
  launch_prolong(f.pcell, f.ncell, c.pc.slots, p.stream);
  enq_iterate(p, L, l, kn, false);
}

// ---- host <-> device field transfer -------------------------------------------
// Reference layout (4, ni, nc) j-fastest <-> device [4][nc][ni] i-fastest;
// split: a Krylov-space vector under the red-black line order (column bcol).
void upload_field(Pool& p, const CaseHost& c, const double* src, void* dst, int prec, int nv, int split = 0) {
  const int ni = c.ni, nc = c.nc;
  const size_t n = size_t(ni) * nc;
  if (prec == 0) {
    p.stage_d.resize(nv * n);
    for (int v = 0; v < nv; ++v)
      for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nc; ++j)
          p.stage_d[v * n + size_t(j) * ni + bcol(i, ni, split)] = src[v * n + size_t(i) * nc + j];
    HC(hipMemcpyAsync(dst, p.stage_d.data(), nv * n * 8, hipMemcpyHostToDevice, p.stream));
  } else {
    p.stage_f.resize(nv * n);
    for (int v = 0; v < nv; ++v)
      for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nc; ++j)
          p.stage_f[v * n + size_t(j) * ni + bcol(i, ni, split)] = float(src[v * n + size_t(i) * nc + j]);
    HC(hipMemcpyAsync(dst, p.stage_f.data(), nv * n * 4, hipMemcpyHostToDevice, p.stream));
  }
  HC(hipStreamSynchronize(p.stream));
}

void download_field(Pool& p, const CaseHost& c, const void* src, double* dst, int prec, int nv, int split = 0) {
  const int ni = c.ni, nc = c.nc;
  const size_t n = size_t(ni) * nc;
  if (prec == 0) {
    p.stage_d.resize(nv * n);
    HC(hipMemcpyAsync(p.stage_d.data(), src, nv * n * 8, hipMemcpyDeviceToHost, p.stream));
    HC(hipStreamSynchronize(p.stream));
    for (int v = 0; v < nv; ++v)
      for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nc; ++j)
          dst[v * n + size_t(i) * nc + j] = p.stage_d[v * n + size_t(j) * ni + bcol(i, ni, split)];
  } else {
    p.stage_f.resize(nv * n);
    HC(hipMemcpyAsync(p.stage_f.data(), src, nv * n * 4, hipMemcpyDeviceToHost, p.stream));
    HC(hipStreamSynchronize(p.stream));
    for (int v = 0; v < nv; ++v)
      for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nc; ++j)
          dst[v * n + size_t(i) * nc + j] = p.stage_f[v * n + size_t(j) * ni + bcol(i, ni, split)];
  }
}

void set_cfl_one(Pool& p, BatchPlan& b, double cfl) {
  b.hcfl[0] = cfl;
  HC(hipMemcpyAsync(b.dcfl, b.hcfl, sizeof(double), hipMemcpyHostToDevice, p.stream));
  launch_set_cfl(b.pc, b.dcfl, p.stream);
}

template <class F>
int guarded(F&& f) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  try {
    f();
    return 0;
  } catch (const Fail& e) {
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -99;
  }
}

}  // namespace

// Shared with the other translation units of the library (api.hpp): the
// thread-local error text behind a2d_last_error, the library lock and the
// device chosen by a2d_init.
void set_error_text(const std::string& s) { g_err = s; }
std::recursive_mutex& api_mutex() { return g_mu; }
int require_device() {
  ensure_init();
  return g_device;
}
}  // namespace a2d

using namespace a2d;

extern "C" {

int a2d_version(void) { return 1; }

const char* a2d_last_error(void) { return g_err.c_str(); }

void a2d_default_knobs(a2d_knobs* k) { default_knobs(k); }

int a2d_init(int device) {
  return guarded([&] {
    int count = 0;
    HC(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) fail(-1, "no such HIP device");
    HC(hipSetDevice(device));
    HC(hipDeviceGetAttribute(&g_num_cu, hipDeviceAttributeMultiprocessorCount, device));
    HC(hipDeviceGetAttribute(&g_lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    HC(hipDeviceGetAttribute(&g_lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device));
    if (!g_scratch.count(device)) {
      double* ptr = nullptr;
      HC(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
      g_scratch[device] = ptr;
    }
    g_device = device;
  });
}

int a2d_pool_create(int64_t hbm_budget_bytes, int64_t* pool_out) {
  return guarded([&] {
    ensure_init();
    auto p = std::make_unique<Pool>();
    p->id = g_next_pool++;
    p->budget = hbm_budget_bytes;
    default_knobs(&p->knobs);
    HC(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    HC(hipMalloc(&p->dcases, MAX_SLOTS * sizeof(CaseDev)));
    HC(hipMalloc(&p->dcfl, MAX_SLOTS * sizeof(double)));
    HC(hipMalloc(&p->dout, MAX_SLOTS * 10 * sizeof(double)));
    HC(hipHostMalloc(&p->hcfl, MAX_SLOTS * sizeof(double), hipHostMallocDefault));
    HC(hipHostMalloc(&p->hout, MAX_SLOTS * 10 * sizeof(double), hipHostMallocDefault));
    *pool_out = p->id;
    g_pools[p->id] = std::move(p);
  });
}

int a2d_pool_destroy(int64_t pool) {
  return guarded([&] {
    Pool& p = get_pool(pool);
    (void)hipStreamSynchronize(p.stream);
    p.plans.clear();
    for (auto& c : p.cases)
      if (c.live) (void)hipFree(c.slab);
    for (auto& kv : p.free_slabs) (void)hipFree(kv.second);
    for (auto& kv : p.tables) (void)hipFree(kv.second);
    (void)hipFree(p.dcases);
    (void)hipFree(p.dcfl);
    (void)hipFree(p.dout);
    (void)hipHostFree(p.hcfl);
    (void)hipHostFree(p.hout);
    (void)hipStreamDestroy(p.stream);
    g_pools.erase(pool);
  });
}

int a2d_pool_set_knobs(int64_t pool, const a2d_knobs* k) {
  return guarded([&] {
    Pool& p = get_pool(pool);
    if (k->gmres_restart < 0 || k->gmres_restart > KMAX)
      fail(-6, "gmres_restart must be in [0, " + std::to_string(KMAX) + "]");
    if (k->jacobi_sweeps < 1) fail(-6, "jacobi_sweeps must be >= 1");
    if (k->precond_order != 0 && k->precond_order != 1) fail(-6, "precond_order must be 0 or 1");
    if (k->positivity != 0 && k->positivity != 1) fail(-6, "positivity must be 0 or 1");
    if (k->precond_half != 0 && k->precond_half != 1) fail(-6, "precond_half must be 0 or 1");
    if (k->basis_half != 0 && k->basis_half != 1) fail(-6, "basis_half must be 0 or 1");
    if (k->stencil_march < 0 || k->stencil_march > 2) fail(-6, "stencil_march must be 0, 1 or 2");
    if (k->sweep_lds < 0 || k->sweep_lds > 2) fail(-6, "sweep_lds must be 0, 1 or 2");
    if (!(k->irs_eps >= 0.0) || !std::isfinite(k->irs_eps)) fail(-6, "irs_eps must be finite and >= 0");
    if (!(k->cfl_explicit > 0.0)) fail(-6, "cfl_explicit must be > 0");
    for (auto& c : p.cases)
      if (c.live && k->precond_order == 1 && (c.ni % 2) != 0)
        fail(-6, "red-black line ordering needs an even number of cells around");
    if (k->order != 1 && k->order != 2) fail(-6, "order must be 1 or 2");
    for (auto& c : p.cases)
      if (c.live && c.kmax < k->gmres_restart)
        fail(-6, "a live case was allocated for a smaller Krylov dimension");
    p.knobs = *k;
    p.knob_version++;
    for (auto& kv : p.plans)
      if (kv.second->exec) {
        (void)hipGraphExecDestroy(kv.second->exec);
        kv.second->exec = nullptr;
      }
  });
}

int a2d_pool_get_knobs(int64_t pool, a2d_knobs* k) {
  return guarded([&] { *k = get_pool(pool).knobs; });
}

int64_t a2d_pool_used_bytes(int64_t pool) {
  int64_t v = -1;
  guarded([&] { v = get_pool(pool).used; });
  return v;
}

int64_t a2d_case_bytes(int ni, int nc, int precision, int restart) {
  if (ni < 4 || nc < 2 || (precision != 0 && precision != 1)) return -6;
  if (!addr32_ok(ni, nc, precision)) return -6;
  return int64_t(layout(ni, nc, precision, std::max(restart, 1)).total);
}

int a2d_stencil_shapes(int* ti, int* tj, int max_shapes) {
  const int n = stencil_num_shapes();
  for (int s = 0; s < std::min(n, max_shapes); ++s) {
    ti[s] = stencil_shape(s).ti;
    tj[s] = stencil_shape(s).tj;
  }
  return n;
}

int a2d_case_add(int64_t pool, int ni, int nc, const double* x, const double* y, double mach,
                 double alpha_deg, int precision, int* case_id) {
  return guarded([&] {
    ensure_init();
    Pool& p = get_pool(pool);
    if (ni < 4 || nc < 2) fail(-6, "grid too small (need ni >= 4, nc >= 2)");
    if (precision != 0 && precision != 1) fail(-6, "precision must be 0 (fp64) or 1 (mixed)");
    if (!addr32_ok(ni, nc, precision))
      fail(-6, "grid too large: the case's block storage (" + std::to_string(addr32_bytes(ni, nc, precision)) +
                   " bytes) must stay below 4 GiB for 32-bit in-case addressing");
    if (p.knobs.precond_order == 1 && (ni % 2) != 0)
      fail(-6, "red-black line ordering needs an even number of cells around");
    const int kmax = std::max(1, std::min(p.knobs.gmres_restart, KMAX));
    const Layout L = layout(ni, nc, precision, kmax);
    if (p.budget > 0 && p.used + int64_t(L.total) > p.budget)
      fail(-2, "HBM budget exceeded: case needs " + std::to_string(L.total) + " bytes, " +
                   std::to_string(p.budget - p.used) + " free");
    // a released slab of this size (up to 25 % larger), else a new one; cached
    // slabs are freed first where a new one would pass the budget
    char* slab = nullptr;
    size_t cap = L.total;
    {
      auto it = p.free_slabs.lower_bound(L.total);
      if (it != p.free_slabs.end() && it->first <= L.total + L.total / 4 &&
          (p.budget <= 0 || p.used + int64_t(it->first) <= p.budget)) {
        slab = it->second;
        cap = it->first;
        p.cached -= int64_t(cap);
        p.free_slabs.erase(it);
        p.stats.slabs_reused++;
      } else {
        while (!p.free_slabs.empty() &&
               (p.budget > 0 && p.used + p.cached + int64_t(L.total) > p.budget)) {
          auto big = std::prev(p.free_slabs.end());
          (void)hipFree(big->second);
          p.cached -= int64_t(big->first);
          p.free_slabs.erase(big);
        }
        hipError_t e = hipMalloc(&slab, L.total);
        if (e != hipSuccess && !p.free_slabs.empty()) {
          // out of device memory: release every cached slab and try once more
          (void)hipGetLastError();
          for (auto& kv : p.free_slabs) (void)hipFree(kv.second);
          p.free_slabs.clear();
          p.cached = 0;
          e = hipMalloc(&slab, L.total);
        }
        hip_check(e, "hipMalloc(case slab)");
        p.stats.slabs_new++;
      }
    }
    int slot = -1;
    for (int s = 0; s < int(p.cases.size()); ++s)
      if (!p.cases[s].live) { slot = s; break; }
    if (slot < 0) {
      if (int(p.cases.size()) >= MAX_SLOTS) fail(-7, "pool is full");
      p.cases.emplace_back();
      slot = int(p.cases.size()) - 1;
    }
    CaseHost& c = p.cases[slot];
    c = CaseHost{};
    c.ni = ni;
    c.nc = nc;
    c.prec = precision;
    c.kmax = kmax;
    c.bytes = cap;
    c.slab = slab;
    c.live = true;
    p.used += int64_t(cap);
    const size_t n = size_t(ni) * nc;
    char* b = c.slab;
    CaseDev& d = c.dev;
    d.ni = ni;
    d.nc = nc;
    d.n = int(n);
    d.prec = precision;
    d.kmax = kmax;
    d.npart = L.npart;
    d.shape = 0;
    d.ntiles = stencil_blocks(0, ni, nc);
    d.nvecb = cdiv(int(n), VEC_CHUNK);
    d.nlineb = cdiv(ni, LINES_PER_BLOCK);
    d.nlinew = cdiv(ni, 8);
    d.nlinerw = cdiv(ni / 2, 8);
    d.vstride = int(vec_stride(n, precision == 0 ? 8 : 4));
    d.pstride = int(plane_stride(n, precision == 0 ? 8 : 4));
    d.U = reinterpret_cast<double*>(b + L.U);
    d.R = reinterpret_cast<double*>(b + L.R);
    d.dU = reinterpret_cast<double*>(b + L.dU);
    double* vol = reinterpret_cast<double*>(b + L.vol);
    double* Si = reinterpret_cast<double*>(b + L.Si);
    double* Sj = reinterpret_cast<double*>(b + L.Sj);
    double* xf = reinterpret_cast<double*>(b + L.xf);
    double* yf = reinterpret_cast<double*>(b + L.yf);
    double* xw = reinterpret_cast<double*>(b + L.xw);
    double* yw = reinterpret_cast<double*>(b + L.yw);
    d.vol = vol;
    d.Si = Si;
    d.Sj = Sj;
    d.xf = xf;
    d.yf = yf;
    d.xw = xw;
    d.yw = yw;
    if (precision == 1) {
      d.Ut = b + L.Ut;
      d.volt = b + L.volt;
      d.Sit = b + L.Sit;
      d.Sjt = b + L.Sjt;
      d.xft = b + L.xft;
      d.yft = b + L.yft;
    } else {
      d.Ut = d.U;
      d.volt = vol;
      d.Sit = Si;
      d.Sjt = Sj;
      d.xft = xf;
      d.yft = yf;
    }
    d.diag = b + L.diag;
    d.blk = b + L.blk;
    d.V = b + L.V;
    d.x0 = b + L.x0;
    d.w = b + L.w;
    d.d = b + L.d;
    d.g0 = b + L.g0;
    d.hb = precision == 1 ? b + L.hb : nullptr;
    d.xa = b + L.xa;
    d.xb = b + L.xb;
    d.p = precision == 1 ? b + L.pv : nullptr;
    d.vc = precision == 1 ? b + L.vc : nullptr;
    d.part = reinterpret_cast<double*>(b + L.part);
    d.sc = reinterpret_cast<CaseScal*>(b + L.sc);
    c.cp = reinterpret_cast<double*>(b + L.cp);
    d.irs = reinterpret_cast<double*>(b + L.irs);

    // metrics from the nodes, FP64 (aero2d/euler.py:44-60); node (i, j) at x[i*(nc+1)+j]
    const int nj = nc + 1;
    auto X = [&](int i, int j) { return x[size_t(i) * nj + j]; };
    auto Y = [&](int i, int j) { return y[size_t(i) * nj + j]; };
    std::vector<double> hvol(n), hSi(2 * n), hSj(2 * (n + ni)), hxf(ni), hyf(ni), hxw(ni), hyw(ni);
    for (int i = 0; i < ni; ++i) {
      const int i1 = (i + 1) % ni;
      for (int j = 0; j < nc; ++j) {
        const double d1x = X(i1, j + 1) - X(i, j), d1y = Y(i1, j + 1) - Y(i, j);
        const double d2x = X(i, j + 1) - X(i1, j), d2y = Y(i, j + 1) - Y(i1, j);
        hvol[size_t(j) * ni + i] = -0.5 * (d1x * d2y - d1y * d2x);
        hSi[size_t(j) * ni + i] = -(Y(i, j + 1) - Y(i, j));
        hSi[n + size_t(j) * ni + i] = X(i, j + 1) - X(i, j);
      }
      for (int j = 0; j < nj; ++j) {
        hSj[size_t(j) * ni + i] = Y(i1, j) - Y(i, j);
        hSj[n + ni + size_t(j) * ni + i] = -(X(i1, j) - X(i, j));
      }
      hxw[i] = 0.5 * (X(i, 0) + X(i1, 0));
      hyw[i] = 0.5 * (Y(i, 0) + Y(i1, 0));
      hxf[i] = 0.5 * (X(i, nc) + X(i1, nc));
      hyf[i] = 0.5 * (Y(i, nc) + Y(i1, nc));
    }
    for (double v : hvol)
      if (!(v > 0)) {
        release_slab(p, c);
        c = CaseHost{};
        fail(-8, "grid has a cell of non-positive area");
      }
    auto put = [&](void* dst, const std::vector<double>& src) {
      HC(hipMemcpy(dst, src.data(), src.size() * 8, hipMemcpyHostToDevice));
    };
    auto putf = [&](void* dst, const std::vector<double>& src) {
      std::vector<float> f(src.begin(), src.end());
      HC(hipMemcpy(dst, f.data(), f.size() * 4, hipMemcpyHostToDevice));
    };
    put(vol, hvol);
    put(Si, hSi);
    put(Sj, hSj);
    put(xf, hxf);
    put(yf, hyf);
    put(xw, hxw);
    put(yw, hyw);
    if (precision == 1) {
      putf(d.volt, hvol);
      putf(d.Sit, hSi);
      putf(d.Sjt, hSj);
      putf(d.xft, hxf);
      putf(d.yft, hyf);
    }
    // freestream and initial state (aero2d/euler.py:253-256, 383-384)
    const double a = alpha_deg * (3.141592653589793 / 180.0);
    c.winf[0] = 1.0;
    c.winf[1] = mach * std::cos(a);
    c.winf[2] = mach * std::sin(a);
    c.winf[3] = 1.0 / 1.4;
    CaseScal sc{};
    for (int k = 0; k < 4; ++k) sc.winf[k] = c.winf[k];
    sc.cfl = p.knobs.cfl_explicit;
    HC(hipMemcpy(d.sc, &sc, sizeof(CaseScal), hipMemcpyHostToDevice));
    const double r = c.winf[0], u = c.winf[1], v = c.winf[2], pr = c.winf[3];
    const double Uc[4] = {r, r * u, r * v, pr / (1.4 - 1.0) + 0.5 * r * (u * u + v * v)};
    std::vector<double> hU(4 * n);
    for (int k = 0; k < 4; ++k) std::fill(hU.begin() + k * n, hU.begin() + (k + 1) * n, Uc[k]);
    put(d.U, hU);
    if (precision == 1) {
      std::vector<float> f(hU.begin(), hU.end());
      HC(hipMemcpy(d.Ut, f.data(), f.size() * 4, hipMemcpyHostToDevice));
    }
    HC(hipMemcpy(p.dcases + slot, &d, sizeof(CaseDev), hipMemcpyHostToDevice));
    *case_id = slot;
  });
}

int a2d_case_remove(int64_t pool, int case_id) {
  return guarded([&] {
    Pool& p = get_pool(pool);
    CaseHost& c = get_case(p, case_id);
    HC(hipStreamSynchronize(p.stream));
    drop_plans_with(p, case_id);
    release_slab(p, c);
    c = CaseHost{};
    for (auto& o : p.cases)
      if (o.coarse == case_id) o.coarse = -1;
  });
}

int a2d_case_set_freestream(int64_t pool, int case_id, double mach, double alpha_deg) {
  return guarded([&] {
    Pool& p = get_pool(pool);
    CaseHost& c = get_case(p, case_id);
    const double a = alpha_deg * (3.141592653589793 / 180.0);
    c.winf[0] = 1.0;
    c.winf[1] = mach * std::cos(a);
    c.winf[2] = mach * std::sin(a);
    c.winf[3] = 1.0 / 1.4;
    HC(hipMemcpy(c.dev.sc->winf, c.winf, sizeof(c.winf), hipMemcpyHostToDevice));
  });
}

int a2d_case_set_step_limit(int64_t pool, int case_id, double max_update) {
  return guarded([&] {
    Pool& p = get_pool(pool);
    CaseHost& c = get_case(p, case_id);
    if (!(max_update >= 0.0)) fail(-6, "max_update must be >= 0 (0 = the pool's knob)");
    HC(hipMemcpyAsync(&c.dev.sc->max_update, &max_update, sizeof(double), hipMemcpyHostToDevice,
                      p.stream));
    HC(hipStreamSynchronize(p.stream));
  });
}

int a2d_case_park(int64_t pool, int case_id, int parked) {
  return guarded([&] {
    Pool& p = get_pool(pool);
    CaseHost& c = get_case(p, case_id);
    if (parked != 0 && parked != 1) fail(-6, "parked must be 0 or 1");
    c.parked = parked != 0;
    HC(hipMemcpyAsync(&c.dev.sc->parked, &parked, sizeof(int), hipMemcpyHostToDevice, p.stream));
    HC(hipStreamSynchronize(p.stream));
  });
}

int a2d_pool_stats(int64_t pool, int64_t* out, int n) {
  return guarded([&] {
    Pool& p = get_pool(pool);
    const int64_t v[] = {p.stats.steps,     p.stats.replays,     p.stats.captures,     p.stats.plans,
                         p.stats.tables_new, p.stats.slabs_new,  p.stats.slabs_reused, p.cached};
    for (int k = 0; k < n && k < int(sizeof(v) / sizeof(v[0])); ++k) out[k] = v[k];
  });
}

int a2d_case_set_shape(int64_t pool, int case_id, int shape) {
  return guarded([&] {
    Pool& p = get_pool(pool);
    CaseHost& c = get_case(p, case_id);
    if (shape < 0 || shape >= stencil_num_shapes()) fail(-6, "no such stencil shape");
    c.dev.shape = shape;
    c.dev.ntiles = stencil_blocks(shape, c.ni, c.nc);
    HC(hipMemcpy(p.dcases + case_id, &c.dev, sizeof(CaseDev), hipMemcpyHostToDevice));
    drop_plans_with(p, case_id);
  });
}

int a2d_case_set_state(int64_t pool, int case_id, const double* U) {
  return guarded([&] {
    Pool& p = get_pool(pool);
    CaseHost& c = get_case(p, case_id);
    upload_field(p, c, U, c.dev.U, 0, 4);
    if (c.prec == 1) upload_field(p, c, U, c.dev.Ut, 1, 4);
  });
}

int a2d_case_get_state(int64_t pool, int case_id, double* U) {
  return guarded([&] {
    Pool& p = get_pool(pool);
    CaseHost& c = get_case(p, case_id);
    download_field(p, c, c.dev.U, U, 0, 4);
  });
}

int a2d_residual(int64_t pool, int case_id, double circ, double* R, double* rn) {
  return guarded([&] {
    Pool& p = get_pool(pool);
    CaseHost& c = get_case(p, case_id);
    BatchPlan& b = get_plan(p, {case_id});
    const KnobsDev kn = to_dev(p.knobs);
    if (std::isnan(circ)) launch_forces(b.pc, nullptr, p.stream);
    enq_residual(p, b, kn, 0, circ, 0);
    if (R) download_field(p, c, c.dev.R, R, 0, 4);
    if (rn) {
      HC(hipMemcpyAsync(rn, &c.dev.sc->rn, sizeof(double), hipMemcpyDeviceToHost, p.stream));
      HC(hipStreamSynchronize(p.stream));
    }
  });
}

int a2d_forces(int64_t pool, int case_id, double* clcdcm, double* cp, double* circ) {
  return guarded([&] {
    Pool& p = get_pool(pool);
    CaseHost& c = get_case(p, case_id);
    BatchPlan& b = get_plan(p, {case_id});
    launch_forces(b.pc, c.cp, p.stream);
    CaseScal sc;
    HC(hipMemcpyAsync(&sc, c.dev.sc, offsetof(CaseScal, s), hipMemcpyDeviceToHost, p.stream));
    if (cp) HC(hipMemcpyAsync(cp, c.cp, c.ni * sizeof(double), hipMemcpyDeviceToHost, p.stream));
    HC(hipStreamSynchronize(p.stream));
    if (clcdcm) {
      clcdcm[0] = sc.cl;
      clcdcm[1] = sc.cd;
      clcdcm[2] = sc.cm;
    }
    if (circ) *circ = sc.circ;
  });
}

int a2d_local_dt(int64_t pool, int case_id, double* dt) {
  return guarded([&] {
    Pool& p = get_pool(pool);
    CaseHost& c = get_case(p, case_id);
    BatchPlan& b = get_plan(p, {case_id});
    const KnobsDev kn = to_dev(p.knobs);
    set_cfl_one(p, b, p.knobs.cfl_explicit);
    enq_prologue(p, b, kn);
    const size_t n = size_t(c.ni) * c.nc;
    std::vector<double> diag(n), vol(n);
    if (c.prec == 0) {
      HC(hipMemcpyAsync(diag.data(), c.dev.diag, n * 8, hipMemcpyDeviceToHost, p.stream));
    } else {
      std::vector<float> f(n);
      HC(hipMemcpyAsync(f.data(), c.dev.diag, n * 4, hipMemcpyDeviceToHost, p.stream));
      HC(hipStreamSynchronize(p.stream));
      std::copy(f.begin(), f.end(), diag.begin());
    }
    HC(hipMemcpyAsync(vol.data(), c.dev.vol, n * 8, hipMemcpyDeviceToHost, p.stream));
    HC(hipStreamSynchronize(p.stream));
    // diag = vol / dt at CFL = cfl_explicit
    for (int i = 0; i < c.ni; ++i)
      for (int j = 0; j < c.nc; ++j) dt[size_t(i) * c.nc + j] = vol[size_t(j) * c.ni + i] / diag[size_t(j) * c.ni + i];
  });
}

int a2d_jvp(int64_t pool, int case_id, const double* v, double cfl, int with_diag, double* out) {
  return guarded([&] {
    Pool& p = get_pool(pool);
    CaseHost& c = get_case(p, case_id);
    BatchPlan& b = get_plan(p, {case_id});
    const KnobsDev kn = to_dev(p.knobs);
    set_cfl_one(p, b, cfl);
    enq_prologue(p, b, kn);
    upload_field(p, c, v, c.dev.xa, c.prec, 4, kn.precond_rb);
    launch_dcirc(b.pc, VEC_XA, 0, -1, kn, c.prec, p.stream);
    enq_jvp(p, b, kn, VEC_XA, VEC_W, 0, -1, with_diag, 0);
    download_field(p, c, c.dev.w, out, c.prec, 4, kn.precond_rb);
  });
}

int a2d_assemble(int64_t pool, int case_id, double cfl, double* D, double* Bim, double* Bip,
                 double* Bjm, double* Bjp) {
  return guarded([&] {
    Pool& p = get_pool(pool);
    CaseHost& c = get_case(p, case_id);
    BatchPlan& b = get_plan(p, {case_id});
    const KnobsDev kn = to_dev(p.knobs);
    set_cfl_one(p, b, cfl);
    enq_prologue(p, b, kn);
    launch_assemble(b.pcell, b.ncell, kn, c.prec, p.stream);
    const size_t t = c.prec == 0 ? 8 : 4;
    const size_t ps = size_t(c.dev.pstride);
    std::vector<char> raw(80 * ps * t);
    HC(hipMemcpyAsync(raw.data(), c.dev.blk, raw.size(), hipMemcpyDeviceToHost, p.stream));
    HC(hipStreamSynchronize(p.stream));
    auto get = [&](int slot, int m, size_t idx) -> double {
      const size_t o = blk_off(slot, m, idx, ps);
      return c.prec == 0 ? reinterpret_cast<double*>(raw.data())[o] : reinterpret_cast<float*>(raw.data())[o];
    };
    double* outs[5] = {D, Bjp, Bjm, Bim, Bip};
    for (int slot = 0; slot < 5; ++slot) {
      if (!outs[slot]) continue;
      for (int i = 0; i < c.ni; ++i)
        for (int j = 0; j < c.nc; ++j)
          for (int m = 0; m < 16; ++m)
            outs[slot][(size_t(i) * c.nc + j) * 16 + m] =
                get(slot, m, size_t(j) * c.ni + bcol(i, c.ni, kn.precond_rb));
    }
  });
}

int a2d_precond_setup(int64_t pool, int case_id, double cfl) {
  return guarded([&] {
    Pool& p = get_pool(pool);
    CaseHost& c = get_case(p, case_id);
    BatchPlan& b = get_plan(p, {case_id});
    const KnobsDev kn = to_dev(p.knobs);
    set_cfl_one(p, b, cfl);
    enq_prologue(p, b, kn);
    launch_assemble(b.pcell, b.ncell, kn, c.prec, p.stream);
    launch_factor(b.pline, b.nline, c.prec, p.stream);
    HC(hipStreamSynchronize(p.stream));
  });
}

int a2d_precond_apply(int64_t pool, int case_id, const double* r, double* x) {
  return guarded([&] {
    Pool& p = get_pool(pool);
    CaseHost& c = get_case(p, case_id);
    BatchPlan& b = get_plan(p, {case_id});
    const int split = p.knobs.precond_order == 1;
    upload_field(p, c, r, c.dev.w, c.prec, 4, split);
    enq_precond(p, b, VEC_W, VEC_X0, 0, 0);
    download_field(p, c, c.dev.x0, x, c.prec, 4, split);
  });
}

int a2d_newton_step(int64_t pool, const int* case_ids, int n, const double* cfl, double* out) {
  return guarded([&] {
    Pool& p = get_pool(pool);
    if (n <= 0) fail(-6, "empty batch");
    std::vector<int> slots(case_ids, case_ids + n);
    check_batch(p, slots);
    BatchPlan& b = get_plan(p, slots);
    std::memcpy(b.hcfl, cfl, n * sizeof(double));
    HC(hipMemcpyAsync(b.dcfl, b.hcfl, n * sizeof(double), hipMemcpyHostToDevice, p.stream));
    b.uses++;
    p.stats.steps++;
    const bool graph = p.knobs.use_graph && b.uses >= 2;
    if (graph && (!b.exec || b.knob_version != p.knob_version)) {
      p.stats.captures++;
      if (b.exec) {
        (void)hipGraphExecDestroy(b.exec);
        b.exec = nullptr;
      }
      hipGraph_t gr;
      HC(hipStreamBeginCapture(p.stream, hipStreamCaptureModeThreadLocal));
      enq_newton(p, b);
      HC(hipStreamEndCapture(p.stream, &gr));
      HC(hipGraphInstantiate(&b.exec, gr, nullptr, nullptr, 0));
      (void)hipGraphDestroy(gr);
      b.knob_version = p.knob_version;
    }
    if (graph) {
      HC(hipGraphLaunch(b.exec, p.stream));
      p.stats.replays++;
    } else {
      enq_newton(p, b);
    }
    HC(hipMemcpyAsync(b.hout, b.dout, n * 10 * sizeof(double), hipMemcpyDeviceToHost, p.stream));
    HC(hipStreamSynchronize(p.stream));
    std::memcpy(out, b.hout, n * 10 * sizeof(double));
  });
}

int a2d_case_link_coarse(int64_t pool, int case_id, int coarse_id) {
  return guarded([&] {
    Pool& p = get_pool(pool);
    CaseHost& f = get_case(p, case_id);
    if (coarse_id < 0) {
      f.coarse = -1;
      return;
    }
    CaseHost& c = get_case(p, coarse_id);
    if (coarse_id == case_id) fail(-6, "a case cannot be its own coarse level");
    if (f.prec != 0 || c.prec != 0) fail(-6, "the explicit scheme's grid levels must be FP64 cases");
    if ((f.ni % 2) != 0 || (f.nc % 2) != 0 || 2 * c.ni != f.ni || 2 * c.nc != f.nc)
      fail(-6, "the coarse case must have half the cells of the fine case along i and j (even counts)");
    f.coarse = coarse_id;
  });
}

int a2d_explicit_cycles(int64_t pool, const int* case_ids, int n, int ncycles, int fresh, double* out) {
  return guarded([&] {
    Pool& p = get_pool(pool);
    if (n <= 0) fail(-6, "empty batch");
    if (ncycles < 0) fail(-6, "ncycles must be >= 0");
    std::vector<std::vector<int>> levels{std::vector<int>(case_ids, case_ids + n)};
    check_batch(p, levels[0]);
    for (;;) {
      const std::vector<int>& cur = levels.back();
      int linked = 0;
      for (int s : cur) linked += p.cases[s].coarse >= 0;
      if (linked == 0) break;
      if (linked != int(cur.size())) fail(-6, "every case of a batch needs the same number of grid levels");
      std::vector<int> nxt;
      for (int s : cur) {
        const CaseHost& f = p.cases[s];
        const CaseHost& c = get_case(p, f.coarse);
        if (2 * c.ni != f.ni || 2 * c.nc != f.nc || c.prec != 0)
          fail(-6, "a coarse level no longer matches its fine case");
        nxt.push_back(f.coarse);
      }
      if (levels.size() >= 16) fail(-6, "too many grid levels (do the coarse links form a cycle?)");
      levels.push_back(nxt);
    }
    std::vector<int> all;
    int ni_max = 0;
    for (auto& lv : levels)
      for (int s : lv) {
        const CaseHost& c = p.cases[s];
        if (c.prec != 0) fail(-5, "the explicit scheme runs FP64 cases only");
        if (c.parked) fail(-6, "a parked case takes no steps");
        ni_max = std::max(ni_max, c.ni);
        all.push_back(s);
      }
    check_batch(p, all);   // no case on two levels or twice on one
    const double eps = p.knobs.irs_eps;
    if (eps > 0 && (rk_rows_cm(ni_max) == 0 || rk_rows_lds(rk_rows_cm(ni_max)) > size_t(g_lds_block)))
      fail(-6, "rows too long for the smoothing kernel (" + std::to_string(ni_max) + " cells around)");
    if (eps > 0)
      for (int s : all) irs_table(p, p.cases[s], eps);
    // fetch every level's plan with no eviction in between
    if (p.plans.size() + levels.size() >= 64) p.plans.clear();
    std::vector<BatchPlan*> L;
    for (auto& lv : levels) L.push_back(&get_plan(p, lv));
    const KnobsDev kn = to_dev(p.knobs);
    const double cfl = p.knobs.cfl_explicit;
    BatchPlan& top = *L[0];
    for (int it = 0; it < ncycles; ++it) {
      for (size_t l = 0; l < L.size(); ++l)
        launch_level_begin(L[l]->pc, top.pc.slots, cfl, fresh && it == 0 && l == 0, p.stream);
      enq_cycle(p, L, 0, kn);
      // Gamma = Cl |V_inf| / 2 of the new state (aero2d/euler.py:400-405)
      launch_forces(top.pc, nullptr, p.stream);
    }
    if (ncycles == 0 && fresh) launch_level_begin(top.pc, top.pc.slots, cfl, 1, p.stream);
    launch_xgather(top.pc, top.dout, p.stream);
    HC(hipMemcpyAsync(top.hout, top.dout, size_t(n) * 5 * sizeof(double), hipMemcpyDeviceToHost, p.stream));
    HC(hipStreamSynchronize(p.stream));
    std::memcpy(out, top.hout, size_t(n) * 5 * sizeof(double));
  });
}

int a2d_time_kernel(int64_t pool, const int* case_ids, int n, int which, int reps, double* ms) {
  return guarded([&] {
    Pool& p = get_pool(pool);
    if (n <= 0 || reps <= 0) fail(-6, "empty batch or no repetitions");
    std::vector<int> slots(case_ids, case_ids + n);
    check_batch(p, slots);
    BatchPlan& b = get_plan(p, slots);
    const KnobsDev kn = to_dev(p.knobs);
    const int k = std::min(25, std::max(0, std::min(p.knobs.gmres_restart, KMAX) - 2));
    const int hb = b.prec == 1 && p.knobs.precond_half;
    const int bh = b.prec == 1 && p.knobs.basis_half;
    if ((which == 11 || which == 12) && b.prec != 0) fail(-5, "the explicit scheme's kernels run FP64 cases only");
    const double eps = p.knobs.irs_eps > 0 ? p.knobs.irs_eps : 0.6;
    if (which == 11 || which == 12)
      for (int s : slots) irs_table(p, p.cases[s], eps);
    if (which == 12 && (rk_rows_cm(b.ni_max) == 0 || rk_rows_lds(rk_rows_cm(b.ni_max)) > size_t(g_lds_block)))
      fail(-6, "rows too long for the smoothing kernel");
    // a Runge-Kutta stage overwrites U: keep it in xb (FP64 scratch) meanwhile
    if (which == 12)
      for (int s : slots) {
        const CaseHost& c = p.cases[s];
        HC(hipMemcpyAsync(c.dev.xb, c.dev.U, size_t(4) * c.ni * c.nc * 8, hipMemcpyDeviceToDevice, p.stream));
      }
    auto launch = [&] {
      switch (which) {
        case 0:  // as launched in a Newton step: with the local time step and U copy
          enq_residual(p, b, kn, 1, NAN, 1);
          break;
        case 1:
          enq_jvp(p, b, kn, VEC_X0, VEC_W, 0, -1, 1, 0);
          break;
        case 2:  // a coupled line-Jacobi sweep over every line (sweeps 2.. of an application)
          launch_line_sweep(sweep_plan(p, b.psw), b.nsw, VEC_W, VEC_XA, VEC_XB, 1, 1, 0, 0,
                            b.prec, hb, p.stream);
          break;
        case 3: launch_dots(b.pvec, b.nvec, k, bh ? VEC_P : k + 1, 0, b.prec, bh, p.stream); break;
        case 4: launch_assemble(b.pcell, b.ncell, kn, b.prec, p.stream); break;
        case 5: launch_factor(b.pline, b.nline, b.prec, p.stream); break;
        case 6: launch_update(b.pvec, b.nvec, k, VEC_XA, 0, 0, b.prec, bh, p.stream); break;
        case 7: launch_copy_vec(b.pvec, b.nvec, VEC_X0, VEC_XB, b.prec, p.stream); break;
        case 8:  // the first (uncoupled) sweep of an application
          launch_line_sweep(sweep_plan(p, b.psw), b.nsw, VEC_W, VEC_XA, VEC_XB, 0, 0, 0, 0,
                            b.prec, hb, p.stream);
          break;
        case 9:  // a coupled red-black half sweep (odd lines) of a later sweep
          launch_line_sweep_rb(sweep_plan(p, b.prb), b.nrb, VEC_W, VEC_XB, 1, 1, 1, 0, 0,
                               b.prec, hb, p.stream);
          break;
        case 10: enq_precond(p, b, VEC_W, VEC_XA, 0, 0); break;
        case 11: enq_irs_j(p, b, eps, 0); break;
        case 12:  // a later Runge-Kutta stage: smoothing along i and the update
          launch_rk_rows(b.prow, b.nrow, rk_rows_cm(b.ni_max), eps, 0.5, 0, 0, 0, p.stream);
          break;
        default: fail(-6, "unknown kernel id");
      }
    };
    launch();  // warm
    hipEvent_t e0, e1;
    HC(hipEventCreate(&e0));
    HC(hipEventCreate(&e1));
    HC(hipEventRecord(e0, p.stream));
    for (int r = 0; r < reps; ++r) launch();
    HC(hipEventRecord(e1, p.stream));
    HC(hipEventSynchronize(e1));
    float t = 0.f;
    HC(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = double(t) / reps;
    if (which == 12) {
      for (int s : slots) {
        const CaseHost& c = p.cases[s];
        HC(hipMemcpyAsync(c.dev.U, c.dev.xb, size_t(4) * c.ni * c.nc * 8, hipMemcpyDeviceToDevice, p.stream));
      }
      HC(hipStreamSynchronize(p.stream));
    }
    if (which == 4 || which == 5) {
      // assembly/factorisation overwrote the factored blocks: refactor so the
      // cases stay consistent for a following Newton step
      launch_assemble(b.pcell, b.ncell, kn, b.prec, p.stream);
      launch_factor(b.pline, b.nline, b.prec, p.stream);
      HC(hipStreamSynchronize(p.stream));
    }
  });
}

int a2d_hllc_flux_noalloc(int n, const double* WL, const double* WR, const double* S, double* F) {
  return guarded([&] {
    ensure_init();
    if (n < 0 || n > NOALLOC_MAX_FACES)
      fail(-6, "a2d_hllc_flux_noalloc takes at most " + std::to_string(NOALLOC_MAX_FACES) + " faces");
    if (n == 0) return;
    auto it = g_scratch.find(g_device);
    if (it == g_scratch.end()) fail(-1, "no scratch address for this device: call a2d_init");
    double* d = it->second;
    const size_t nn = size_t(n);
    HC(hipMemcpy(d, WL, nn * 4 * 8, hipMemcpyHostToDevice));
    HC(hipMemcpy(d + 4 * nn, WR, nn * 4 * 8, hipMemcpyHostToDevice));
    HC(hipMemcpy(d + 8 * nn, S, nn * 2 * 8, hipMemcpyHostToDevice));
    launch_hllc_test(n, d, d + 4 * nn, d + 8 * nn, d + 10 * nn, nullptr);
    HC(hipGetLastError());
    HC(hipMemcpy(F, d + 10 * nn, nn * 4 * 8, hipMemcpyDeviceToHost));
  });
}

int a2d_hllc_flux(int n, const double* WL, const double* WR, const double* S, double* F) {
  return guarded([&] {
    ensure_init();
    double* d = nullptr;
    const size_t bytes = size_t(n) * 14 * sizeof(double);
    HC(hipMalloc(&d, bytes));
    HC(hipMemcpy(d, WL, n * 4 * 8, hipMemcpyHostToDevice));
    HC(hipMemcpy(d + 4 * size_t(n), WR, n * 4 * 8, hipMemcpyHostToDevice));
    HC(hipMemcpy(d + 8 * size_t(n), S, n * 2 * 8, hipMemcpyHostToDevice));
    launch_hllc_test(n, d, d + 4 * size_t(n), d + 8 * size_t(n), d + 10 * size_t(n), nullptr);
    HC(hipGetLastError());
    HC(hipMemcpy(F, d + 10 * size_t(n), n * 4 * 8, hipMemcpyDeviceToHost));
    (void)hipFree(d);
  });
}

}  // extern "C"
