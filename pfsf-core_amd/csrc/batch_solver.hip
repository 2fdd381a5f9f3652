// Batched multi-island solve: many independent islands, one set of launches.
//
// The reference tick solves dirty islands one after another, each with its
// own fixed schedule of small dispatches (NP/src/pfsf_engine.cpp:424-600);
// small islands leave the GPU idle between launches.  Here every island keeps
// the single-system numerics of mg_solver.hip (FP32 MG-preconditioned PCG
// with an FP64 iterate, FP64 restarts, its own scalars and convergence test),
// but each stencil launch covers the tiles of all islands at once: a tile
// record names the island's level descriptor and the tile origin, and a
// converged island's work-groups return at once.  Islands have between 2 and
// 5 levels; level-l launches list the tiles of the islands deeper than l+1,
// and one launch solves every island's coarsest level (one work-group each).
// Two iterations plus a live-island count are captured in one HIP graph.
// Per-island set-up, the FP64 residual and the solution copies run one
// work-group per island.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "batch_solver.h"
#include "mg_kernels.h"
#include "mg_stream.h"

namespace pfsf {
namespace {

constexpr int kChunk = 1024;  // cells per work-group of the per-cell passes
// Coarsest-level Chebyshev steps in a batch: a batched sweep costs ~5 ms of
// fine-level work against ~50 us for the coarse solve, so the steps that
// cut the iteration count of the slowest island pay (12 steps: 35 iterations
// at most on the 4,096-island batch; 6 steps: 37).
constexpr int kBatchCoarseSweeps = 12;

template <typename KT>
__device__ __forceinline__ Lv<KT> lv_of(const BatchLevel& D) {
  Lv<KT> L;
  L.d = D.d;
  L.kap = (const KT*)D.kap;
  L.dg = D.dg;
  L.am = D.am;
  L.msk = D.msk;
  L.scale = D.scale;
  L.omega = D.omega;
  return L;
}

// ── batched stencil launches (one work-group per tile record) ──────────
// UPDATE (finest level, iteration c): r[c] -> r[1-c] with p[1-c], q.
// Otherwise the down-sweep of the level's own right-hand side.
template <int TZ, typename KT, bool UPDATE, bool MASK = false>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(PFSF_WPE)))
    kb_down(const BatchTile* __restrict__ tiles, const BatchLevel* __restrict__ lv,
            const BatchIsland* __restrict__ isl, PcgScalars* sc, const double* __restrict__ pq_part,
            double* rr_part, int c) {
  const BatchTile t = tiles[blockIdx.x];
  if (t.level < 0) return;
  const BatchLevel& D = lv[t.level];
  const BatchIsland& I = isl[D.island];
  const float* rin = UPDATE ? (c ? I.r1 : I.r0) : D.r;
  float* rout = UPDATE ? (c ? I.r0 : I.r1) : nullptr;
  down_body<TZ, KT, UPDATE, MASK>(lv_of<KT>(D), true, t.x0, t.y0, t.z0, D.c, rin, rout, I.q,
                               c ? I.p0 : I.p1, I.x, D.rc, sc + D.island, pq_part + I.tb,
                               (unsigned)I.tn, t.slot == I.tb, rr_part, (unsigned)t.slot);
}

// FINEST: z = M r[ridx] and the r.z partial.  Otherwise the up-sweep of a
// coarse level into its correction.
template <int TZ, typename KT, bool FINEST>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(PFSF_WPE)))
    kb_up(const BatchTile* __restrict__ tiles, const BatchLevel* __restrict__ lv,
          const BatchIsland* __restrict__ isl, const PcgScalars* sc, double* rz_part, int ridx) {
  const BatchTile t = tiles[blockIdx.x];
  if (t.level < 0) return;
  const BatchLevel& D = lv[t.level];
  const BatchIsland& I = isl[D.island];
  const float* r = FINEST ? (ridx ? I.r1 : I.r0) : D.r;
  float* out = FINEST ? I.z : D.e;
  // every level below the finest carries a coupling mask
  up_body<TZ, KT, FINEST, !FINEST>(lv_of<KT>(D), true, t.x0, t.y0, t.z0, D.c, r, D.ec, out,
                                      sc + D.island, rz_part, (unsigned)t.slot);
}

// Finest level: the plane-streaming bodies (one 32 x 8 x TZ column per
// tile record): the SpMV with the new direction, the update + down-sweep
// and the up-sweep with r.z.
template <int TZ>
__global__ void __launch_bounds__(BS)
    kbs_spmv(const BatchTile* __restrict__ tiles, const BatchLevel* __restrict__ lv,
             const BatchIsland* __restrict__ isl, PcgScalars* sc,
             const double* __restrict__ rz_part, double* pq_part, int c) {
  const BatchTile t = tiles[blockIdx.x];
  if (t.level < 0) return;
  const BatchLevel& D = lv[t.level];
  const BatchIsland& I = isl[D.island];
  spmv_stream_body<TZ>(lv_of<float>(D), OneColumn{t.x0, t.y0, t.z0}, I.z, c ? I.p1 : I.p0,
                       c ? I.p0 : I.p1, I.q, sc + D.island, rz_part + I.tb, (unsigned)I.tn, c,
                       t.slot == I.tb, pq_part, (unsigned)t.slot);
}

// Iterations of a pair (c = 0, 1; passes start at c = 0) update the FP64
// iterate as the single solver's do: the even one leaves alpha p pending,
// the odd one applies both directions (kXDefer / kXPair, see ks_down).
template <int TZ, bool UPDATE, int XM = kXNow>
__global__ void __launch_bounds__(BS)
    kbs_down(const BatchTile* __restrict__ tiles, const BatchLevel* __restrict__ lv,
             const BatchIsland* __restrict__ isl, PcgScalars* sc,
             const double* __restrict__ pq_part, double* rr_part, int c) {
  const BatchTile t = tiles[blockIdx.x];
  if (t.level < 0) return;
  const BatchLevel& D = lv[t.level];
  const BatchIsland& I = isl[D.island];
  const float* rin = UPDATE ? (c ? I.r1 : I.r0) : D.r;
  float* rout = UPDATE ? (c ? I.r0 : I.r1) : nullptr;
  down_stream_body<TZ, UPDATE, XM>(lv_of<float>(D), OneColumn{t.x0, t.y0, t.z0}, D.c, rin, rout,
                                   I.q, c ? I.p0 : I.p1, I.p1, I.x, D.rc, sc + D.island,
                                   pq_part + I.tb, (unsigned)I.tn, t.slot == I.tb, rr_part,
                                   (unsigned)t.slot);
}

template <int TZ>
__global__ void __launch_bounds__(BS)
    kbs_up(const BatchTile* __restrict__ tiles, const BatchLevel* __restrict__ lv,
           const BatchIsland* __restrict__ isl, const PcgScalars* sc, double* rz_part, int ridx) {
  const BatchTile t = tiles[blockIdx.x];
  if (t.level < 0) return;
  const BatchLevel& D = lv[t.level];
  const BatchIsland& I = isl[D.island];
  up_stream_body<TZ, true>(lv_of<float>(D), OneColumn{t.x0, t.y0, t.z0}, D.c, ridx ? I.r1 : I.r0, D.ec,
                           I.z, sc + D.island, rz_part, (unsigned)t.slot);
}

// Every island's coarsest level, one work-group each; with `check` the
// island's convergence test from its r.r partials runs first.
template <typename KT>
__global__ void __launch_bounds__(kMaxLdsCoarse)
    kb_cheb(const BatchLevel* __restrict__ lv, const BatchIsland* __restrict__ isl,
            const float* __restrict__ coef, int steps, PcgScalars* sc,
            const double* __restrict__ rr_part, int check) {
  const BatchIsland& I = isl[blockIdx.x];
  const BatchLevel& D = lv[I.coarsest];
  cheb_lds_body<KT>(lv_of<KT>(D), D.r, D.e, coef, steps, sc + blockIdx.x, rr_part + I.tb,
                       check ? (unsigned)I.tn : 0u);
}

// ── per-cell passes over chunks of every island ────────────────────────
// ||b||^2 partials over the free set and the FP64 iterate from the caller's
// guess (or 0) on the free set.
__global__ void __launch_bounds__(BS)
    kb_begin_part(const BatchChunk* __restrict__ ch, const BatchIsland* __restrict__ isl,
                  double* part) {
  __shared__ double sh[1][BS / 64];
  const BatchChunk c = ch[blockIdx.x];
  const BatchIsland& I = isl[c.island];
  double acc = 0.0;
  for (int i = c.begin + (int)threadIdx.x; i < c.end; i += BS) {
    const bool f = I.fr[i] != 0;
    I.x[i] = (f && I.xio) ? I.xio[i] : 0.0;
    const double bv = f ? (double)I.b[i] : 0.0;
    acc += bv * bv;
  }
  write_partial_at(acc, part, sh, blockIdx.x);
}

// Per island: ||b||^2 from its chunk partials; state reset.
__global__ void __launch_bounds__(BS)
    kb_begin_fin(const BatchIsland* __restrict__ isl, const double* __restrict__ part,
                 BatchState* st, PcgScalars* sc) {
  const BatchIsland& I = isl[blockIdx.x];
  const double bb = block_total(part + I.cb, (unsigned)I.cn);
  if (threadIdx.x == 0) {
    BatchState& S = st[blockIdx.x];
    S.bb = bb;
    S.rel = S.prev_rel = INFINITY;
    S.total = S.passes = S.finished = S.converged = 0;
    sc[blockIdx.x].iters = 0;
    sc[blockIdx.x].done = 3;
  }
}

// True FP64 residual of the unfinished islands into their r[0], one finest
// tile per work-group (every island's tiles, in XCD order), with the ||r||^2
// partial in the tile's slot.
template <int TZ>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(PFSF_WPE)))
    kb_residual_tiles(const BatchTile* __restrict__ tiles, const BatchLevel* __restrict__ lv,
                      const BatchIsland* __restrict__ isl, const BatchState* __restrict__ st,
                      double* part) {
  const BatchTile t = tiles[blockIdx.x];
  if (t.level < 0) return;
  const BatchLevel& D = lv[t.level];
  const BatchState& S = st[D.island];
  if (S.finished) return;
  const BatchIsland& I = isl[D.island];
  // before the first pass an island without an initial guess has x = 0
  const bool xzero = S.passes == 0 && !I.xio;
  residual_stream_body<TZ>(I.d, OneColumn{t.x0, t.y0, t.z0}, I.type, I.kappa, I.fr, I.b, I.x,
                           I.r0, part, (unsigned)t.slot, xzero);
}

// The iterate update an unfinished island's pass left pending (its last
// iteration, or the last one before a breakdown, was an even one: an odd
// number of convergence tests): x += xa p1 over the island's finest tiles,
// the FMA of a single update (p1 is zero off the free set).  Runs before
// each residual phase and once after the last pass; an island is finished
// only after a residual phase, so no update is applied twice.
template <int TZ>
__global__ void __launch_bounds__(BS)
    kb_x_pending(const BatchTile* __restrict__ tiles, const BatchLevel* __restrict__ lv,
                 const BatchIsland* __restrict__ isl, const BatchState* __restrict__ st,
                 const PcgScalars* __restrict__ sc) {
  const BatchTile t = tiles[blockIdx.x];
  if (t.level < 0) return;
  const int k = lv[t.level].island;
  if (st[k].finished || (sc[k].iters & 1) == 0) return;
  const BatchIsland& I = isl[k];
  const double a = sc[k].xa;
  const int x = t.x0 + (int)threadIdx.x % TX, y = t.y0 + (int)threadIdx.x / TX;
  if (x >= I.d.lx || y >= I.d.ly) return;
  for (int z = t.z0; z < t.z0 + TZ && z < I.d.lz; z++) {
    const int i = flat32(x, y, z, I.d);
    I.x[i] = fma(a, (double)I.p1[i], I.x[i]);
  }
}

// Pass boundary, per island: an island at the tolerance, out of iterations
// or stalled finishes; the others restart PCG from the new residual (the pass
// stops when the recursive residual has dropped by max(0.85 tol / rel, 2e-6)).
// *live counts the restarts.
__global__ void __launch_bounds__(BS)
    kb_residual_fin(const BatchIsland* __restrict__ isl, const double* __restrict__ part,
                    BatchState* st, PcgScalars* sc, double tol, int max_iters, int* live) {
  const int k = blockIdx.x;
  const BatchIsland& I = isl[k];
  BatchState& S = st[k];
  if (S.finished) return;
  const double rr = block_total(part + I.tb, (unsigned)I.tn);
  if (threadIdx.x != 0) return;
  PcgScalars& P = sc[k];
  const int pass_iters = P.iters;
  const int pass_done = P.done;
  const double rel = S.bb > 0.0 ? sqrt(rr / S.bb) : sqrt(rr);
  const bool first = S.passes == 0;
  S.total += pass_iters;
  const bool stalled = !first && (pass_iters == 0 || (pass_done == 2 && rel >= S.rel));
  // the first pass's drift from the true residual sets the next solve's
  // first-pass floor (see MgSolver::solve)
  if (S.passes == 1 && pass_done == 1 && pass_iters > 0 && P.rr > 0.0 && S.rr0 > 0.0 &&
      S.rel > 0.0) {
    const double rec = S.rel * sqrt(P.rr / S.rr0);
    const double drift = rel / rec;
    const double f0 = S.floor0 > 0.0 ? S.floor0 : PFSF_PASS_FLOOR;
    if (drift > 3.0)
      S.floor0 = fmin(fmax(0.5 * rel / S.rel, PFSF_PASS_FLOOR), 1e-2);
    else if (drift < 1.5)
      S.floor0 = fmax(0.5 * f0, PFSF_PASS_FLOOR);
  }
  S.prev_rel = S.rel;
  S.rel = rel;
  if (rel <= tol || stalled || S.total >= max_iters || S.passes >= 16) {
    S.finished = 1;
    S.converged = rel <= tol;
    P.done = 3;
    return;
  }
  const double floor = (first && S.floor0 > 0.0) ? S.floor0 : PFSF_PASS_FLOOR;
  const double stop_rel = fmax(0.85 * tol / rel, floor);
  S.rr0 = rr;
  P.done = 0;
  P.iters = 0;
  P.have_dir = 0;
  P.rz = P.beta = P.alpha = P.pq = P.xa = 0.0;
  P.rzs[0] = P.rzs[1] = 0.0;
  P.rr = rr;
  P.stop2 = rr * stop_rel * stop_rel;
  P.max_iters = max_iters - S.total;
  S.passes += 1;
  atomicAdd(live, 1);
}

// Islands whose PCG pass is still running: flags[k] and their count.
__global__ void __launch_bounds__(1024)
    kb_live(const PcgScalars* __restrict__ sc, int n, int* flags, int* live) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int c = 0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const int f = sc[k].done == 0;
    flags[k] = f;
    c += f;
  }
  atomicAdd(&cnt, c);
  __syncthreads();
  if (threadIdx.x == 0) *live = cnt;
}

// End of a solve: FP64 and FP32 solutions to the caller's arrays.
__global__ void __launch_bounds__(BS)
    kb_finish_part(const BatchChunk* __restrict__ ch, const BatchIsland* __restrict__ isl) {
  const BatchChunk c = ch[blockIdx.x];
  const BatchIsland& I = isl[c.island];
  for (int i = c.begin + (int)threadIdx.x; i < c.end; i += BS) {
    const double v = I.x[i];
    if (I.xio) I.xio[i] = v;
    if (I.phio) I.phio[i] = (float)v;
  }
}

}  // namespace

// ══════════════════════════════ host side ══════════════════════════════

static inline bool chk(hipError_t e, hipError_t& keep) {
  if (e == hipSuccess) return false;
  if (keep == hipSuccess) keep = e;
  return true;
}

BatchSolver::~BatchSolver() { release(); }

void BatchSolver::release() {
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  graph_exec_ = nullptr;
  graph_ = nullptr;
  if (stream_) (void)hipStreamSynchronize(stream_);
  sol_.clear();
  isl_.clear();
  for (void* p : owned_) (void)hipFree(p);
  owned_.clear();
  if (pool_mem_) (void)hipFree(pool_mem_);
  pool_mem_ = nullptr;
  if (live_host_) (void)hipHostFree(live_host_);
  live_host_ = nullptr;
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  ev0_ = ev1_ = nullptr;
  if (stream_) (void)hipStreamDestroy(stream_);
  stream_ = nullptr;
}

template <typename T>
static T* balloc(size_t n, std::vector<void*>& owned, hipError_t& err) {
  void* p = nullptr;
  if (chk(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)), err)) return nullptr;
  owned.push_back(p);
  return (T*)p;
}

// Launch order of a tile list: cut into 8 contiguous chunks, work-group g
// takes chunk g % 8 (work-groups are dealt to the 8 XCDs in turn), so each
// XCD walks a contiguous run of tiles; padding records have level -1.
static std::vector<BatchTile> xcd_order(const std::vector<BatchTile>& v) {
  const int T = (int)v.size();
  const int chunk = (T + 7) / 8;
  std::vector<BatchTile> o((size_t)8 * chunk, BatchTile{-1, 0, 0, 0, 0});
  for (int g = 0; g < 8 * chunk; g++) {
    const int src = (g % 8) * chunk + g / 8;
    if (src < T) o[g] = v[src];
  }
  return o;
}

// Island-size classes of the shape tuner (cells lx*ly*lz): [bound k, bound k+1).
constexpr int64_t kClassBound[PFSF_BATCH_CLASSES + 1] = {0, int64_t(1) << 12, int64_t(1) << 15,
                                                         int64_t(1) << 18, INT64_MAX};
constexpr int kTuneSample = 32;  // islands timed per class and candidate
constexpr int kTuneIters = 16;   // PCG iterations per island of a tuning solve

hipError_t BatchSolver::create(const pfsf_batch_desc& desc) {
  hipError_t err = hipSuccess;
  const int n = desc.count;
  std::vector<int> cls(n, 0);
  for (int k = 0; k < PFSF_BATCH_CLASSES; k++) {
    pfsf_batch_class& C = classes_[k];
    C = pfsf_batch_class{};
    C.min_cells = kClassBound[k];
    C.max_cells = kClassBound[k + 1];
    const int c0 = desc.shape >= 0 ? desc.shape : 0;
    C.tz_fine = kShapeFine[c0];
    C.tz_coarse = kShapeCoarse[c0];
  }
  for (int i = 0; i < n; i++) {
    const int64_t c = nvox(desc.islands[i].dims);
    int k = 0;
    while (k + 1 < PFSF_BATCH_CLASSES && c >= kClassBound[k + 1]) k++;
    cls[i] = k;
    classes_[k].islands++;
  }
  if (desc.shape == PFSF_BATCH_AUTOTUNE && chk(tune(desc, cls), err)) return err;
  std::vector<int> shape(n, 0);
  for (int i = 0; i < n; i++) {
    const pfsf_batch_class& C = classes_[cls[i]];
    for (int c = 0; c < PFSF_BATCH_SHAPES; c++)
      if (kShapeFine[c] == C.tz_fine && kShapeCoarse[c] == C.tz_coarse) shape[i] = c;
  }
  return build(desc, shape);
}

// Solve time (ms) of the islands `pick` of desc built with island pick[j]'s
// shape shape[j]: a right-hand side of ones, a tolerance no solve reaches and
// kTuneIters PCG iterations per island, timed after a warm-up solve (graph
// capture).
static hipError_t time_shapes(const pfsf_batch_desc& desc, const std::vector<int>& pick,
                              const std::vector<int>& shape, float* ms) {
  hipError_t err = hipSuccess;
  const int m = (int)pick.size();
  std::vector<pfsf_batch_island> sub(m);
  int64_t cells = 0;
  for (int j = 0; j < m; j++) {
    sub[j] = desc.islands[pick[j]];
    cells += nvox(sub[j].dims);
  }
  float* ones = nullptr;
  if (chk(hipMalloc((void**)&ones, sizeof(float) * (size_t)cells), err)) return err;
  std::vector<const float*> bp(m);
  int64_t off = 0;
  for (int j = 0; j < m; j++) {
    bp[j] = ones + off;
    off += nvox(sub[j].dims);
  }
  chk(hipMemsetD32((hipDeviceptr_t)ones, 0x3f800000, (size_t)cells), err);  // 1.0f
  pfsf_batch_desc td = desc;
  td.count = m;
  td.islands = sub.data();
  td.shape = 0;
  td.tol = 1e-300;
  td.max_iters = kTuneIters;
  {
    BatchSolver t;
    pfsf_batch_info bi{};
    if (err == hipSuccess) chk(t.build(td, shape), err);
    if (err == hipSuccess) chk(t.solve(bp.data(), nullptr, nullptr, &bi, nullptr), err);
    if (err == hipSuccess) chk(t.solve(bp.data(), nullptr, nullptr, &bi, nullptr), err);
    *ms = (float)bi.solve_ms;
  }
  (void)hipFree(ones);
  return err;
}

// Up to `count` of `order`'s islands, evenly spaced, the last included.
static std::vector<int> evenly(const std::vector<int>& order, int count) {
  const int m = std::min<int>((int)order.size(), count);
  std::vector<int> out(m);
  for (int j = 0; j < m; j++)
    out[j] = order[m == 1 ? (int)order.size() - 1
                          : (int)((int64_t)j * ((int64_t)order.size() - 1) / (m - 1))];
  return out;
}

// Per class present: every candidate shape on up to kTuneSample of its
// islands (evenly spaced in size order).  Then the whole-batch decision on a
// stratified sample of kTuneStrata islands: one shape for every island (the
// default and each class's fastest) and, when the classes disagree, the
// per-class mix.  An option replaces the default only when it is faster by
// kTuneMargin, more than the run-to-run spread of these timings (1-2 %).
// Measured on the 4,096-island batch: the small classes prefer 4-plane
// columns in isolation (7-13 % on their samples), but on the stratified
// sample the mix was 0.2 % and 4-plane columns everywhere 6.5 % slower than
// 8x2, and on the whole batch the mix took 144.2 ms against 142.6 ms.
constexpr int kTuneStrata = 256;
constexpr float kTuneMargin = 0.03f;

hipError_t BatchSolver::tune(const pfsf_batch_desc& desc, const std::vector<int>& cls) {
  hipError_t err = hipSuccess;
  const int n = desc.count;
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return nvox(desc.islands[a].dims) < nvox(desc.islands[b].dims);
  });
  int best[PFSF_BATCH_CLASSES] = {};
  for (int k = 0; k < PFSF_BATCH_CLASSES && err == hipSuccess; k++) {
    std::vector<int> mem;
    for (int i : order)
      if (cls[i] == k) mem.push_back(i);
    if (mem.empty()) continue;
    const std::vector<int> pick = evenly(mem, kTuneSample);
    for (int c = 0; c < PFSF_BATCH_SHAPES && err == hipSuccess; c++) {
      chk(time_shapes(desc, pick, std::vector<int>(pick.size(), c), &classes_[k].ms[c]), err);
      if (classes_[k].ms[c] < classes_[k].ms[best[k]]) best[k] = c;
    }
    classes_[k].sample = (int)pick.size();
  }
  if (err != hipSuccess) return err;
  const std::vector<int> strata = evenly(order, kTuneStrata);
  tuning_ = pfsf_batch_tuning{};
  tuning_.sample = (int)strata.size();
  bool agree = true;
  int first = -1;
  std::vector<bool> uniform(PFSF_BATCH_SHAPES, false);
  uniform[0] = true;
  for (int k = 0; k < PFSF_BATCH_CLASSES; k++)
    if (classes_[k].sample > 0) {
      uniform[best[k]] = true;
      if (first < 0) first = best[k];
      agree = agree && best[k] == first;
    }
  int choice = 0;
  float choice_ms = INFINITY;
  for (int c = 0; c < PFSF_BATCH_SHAPES && err == hipSuccess; c++) {
    if (!uniform[c]) continue;
    chk(time_shapes(desc, strata, std::vector<int>(strata.size(), c), &tuning_.uniform_ms[c]),
        err);
    // candidate 0 (the default) is timed first
    if (c == 0 || tuning_.uniform_ms[c] < (1.0f - kTuneMargin) * choice_ms) {
      choice_ms = tuning_.uniform_ms[c];
      choice = c;
    }
  }
  if (err == hipSuccess && !agree) {
    std::vector<int> sh(strata.size());
    for (size_t j = 0; j < strata.size(); j++) sh[j] = best[cls[strata[j]]];
    chk(time_shapes(desc, strata, sh, &tuning_.mixed_ms), err);
    tuning_.mixed = tuning_.mixed_ms < (1.0f - kTuneMargin) * choice_ms ? 1 : 0;
  }
  for (int k = 0; k < PFSF_BATCH_CLASSES; k++) {
    const int c = tuning_.mixed && classes_[k].sample > 0 ? best[k] : choice;
    classes_[k].tz_fine = kShapeFine[c];
    classes_[k].tz_coarse = kShapeCoarse[c];
  }
  return err;
}

hipError_t BatchSolver::build(const pfsf_batch_desc& desc, const std::vector<int>& shape) {
  hipError_t err = hipSuccess;
  desc_ = desc;
  coarse_sweeps_ = desc.coarse_sweeps > 0 ? std::min(desc.coarse_sweeps, 63) : kBatchCoarseSweeps;
  bf16_ = desc.precision == PFSF_PREC_BF16_COARSE;
  const int n = desc.count;
  if (chk(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), err)) return err;
  chk(hipEventCreate(&ev0_), err);
  chk(hipEventCreate(&ev1_), err);
  if (err != hipSuccess) return err;

  // shapes, then one measured pool for every island's hierarchy
  isl_.resize(n);
  nlev_.assign(n, 0);
  for (int i = 0; i < n; i++) {
    const pfsf_batch_island& bi = desc.islands[i];
    if (nvox(bi.dims) > kMaxSolverCells) return hipErrorInvalidValue;
    pfsf_solver_desc sd{};
    sd.dims = bi.dims;
    sd.type = bi.type;
    sd.kappa = bi.kappa;
    sd.precision = desc.precision;
    sd.coarse_sweeps = coarse_sweeps_;
    sd.tol = desc.tol;
    sd.max_iters = desc.max_iters;
    isl_[i].reset(new MgSolver());
    isl_[i]->configure_pooled(sd, kShapeFine[shape[i]], kShapeCoarse[shape[i]]);
    const MgLevel& C = isl_[i]->level(isl_[i]->levels() - 1);
    if (C.n > kMaxLdsCoarse || padded_cells(C.d) > kMaxLdsPadded) return hipErrorInvalidValue;
  }
  DevPool measure;
  for (auto& s : isl_) s->allocate_pooled(measure);
  if (chk(hipMalloc(&pool_mem_, std::max<size_t>(measure.used, 1)), err)) return err;
  DevPool pool;
  pool.base = (char*)pool_mem_;
  for (auto& s : isl_) s->allocate_pooled(pool);
  for (int i = 0; i < n && err == hipSuccess; i++) chk(isl_[i]->build_pooled(stream_), err);
  if (err != hipSuccess) return err;
  sol_.resize(n);
  for (int i = 0; i < n; i++) sol_[i] = isl_[i].get();
  return build_tables();
}

hipError_t BatchSolver::attach(const std::vector<MgSolver*>& solvers, double tol, int max_iters,
                               hipStream_t caller) {
  hipError_t err = hipSuccess;
  desc_ = pfsf_batch_desc{};
  desc_.count = (int32_t)solvers.size();
  desc_.tol = tol;
  desc_.max_iters = max_iters;
  desc_.stream = caller;
  if (solvers.empty()) return hipErrorInvalidValue;
  coarse_sweeps_ = std::min(solvers[0]->coarse_sweeps(), 63);
  if (solvers[0]->levels() < 2) return hipErrorInvalidValue;
  bf16_ = solvers[0]->level(1).bf16;
  desc_.precision = bf16_ ? PFSF_PREC_BF16_COARSE : PFSF_PREC_FP32;
  for (MgSolver* m : solvers) {
    const int tf = m->level(0).tiles.tz;
    if (m->levels() < 2 || m->coarse_sweeps() != solvers[0]->coarse_sweeps() ||
        (tf != 4 && tf != 8 && tf != 16))
      return hipErrorInvalidValue;
    for (int l = 1; l < m->levels(); l++) {
      const int tc = m->level(l).tiles.tz;
      if ((tc != 2 && tc != 4) || m->level(l).bf16 != bf16_) return hipErrorInvalidValue;
    }
    const MgLevel& C = m->level(m->levels() - 1);
    if (C.n > kMaxLdsCoarse || padded_cells(C.d) > kMaxLdsPadded) return hipErrorInvalidValue;
  }
  if (chk(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), err)) return err;
  chk(hipEventCreate(&ev0_), err);
  chk(hipEventCreate(&ev1_), err);
  if (err != hipSuccess) return err;
  sol_ = solvers;
  return build_tables();
}

// Descriptor tables, tile lists, chunks and scalars over sol_.
hipError_t BatchSolver::build_tables() {
  hipError_t err = hipSuccess;
  const int n = (int)sol_.size();
  nlev_.assign(n, 0);

  // descriptor tables and tile lists
  std::vector<BatchLevel> lv;
  host_isl_.assign(n, BatchIsland{});
  lmax_ = 0;
  for (int i = 0; i < n; i++) lmax_ = std::max(lmax_, sol_[i]->levels());
  all_tiles_.assign(lmax_, {});
  tile_begin_.assign(lmax_, std::vector<int>(n + 1, 0));
  std::vector<BatchChunk> chunks;
  int slot = 0;
  for (int i = 0; i < n; i++) {
    MgSolver& S = *sol_[i];
    const int L = S.levels();
    nlev_[i] = L;
    const int base = (int)lv.size();
    for (int l = 0; l < L; l++) {
      const MgLevel& M = S.level(l);
      BatchLevel D{};
      D.d = M.d;
      // the finest level is read by the plane-streaming bodies only: the
      // anchor-coded kappa, no anchor mass
      D.kap = l == 0 ? (const void*)M.ks : (const void*)M.kap;
      D.dg = M.dg;
      D.am = l == 0 ? nullptr : M.am;
      D.msk = M.msk;
      D.scale = M.scale;
      D.omega = M.omega;
      if (l + 1 < L) {
        const MgLevel& C = S.level(l + 1);
        D.c = C.d;
        D.rc = C.r;
        D.ec = C.e;
      } else {
        D.c = pfsf_dims{1, 1, 1};
      }
      D.r = l == 0 ? S.residual(0) : M.r;
      D.e = M.e;
      D.island = i;
      lv.push_back(D);
    }
    BatchIsland& I = host_isl_[i];
    I.d = S.dims();
    I.n = S.cells();
    I.type = S.type();
    I.kappa = S.kappa();
    I.fr = S.level(0).fr;
    I.r0 = S.residual(0);
    I.r1 = S.residual(1);
    I.p0 = S.direction(0);
    I.p1 = S.direction(1);
    I.q = S.product();
    I.z = S.preconditioned();
    I.x = S.iterate();
    I.coarsest = base + L - 1;
    I.tb = slot;
    // levels 0 .. L-2 take part in the down/up sweeps (L >= 2)
    for (int l = 0; l <= L - 2; l++) {
      const TileList& T = S.level(l).tiles;
      for (int t : T.host_active) {
        BatchTile r;
        r.level = base + l;
        r.x0 = (t % T.ntx) * kTX;
        r.y0 = ((t / T.ntx) % T.nty) * kTY;
        r.z0 = (t / (T.ntx * T.nty)) * T.tz;
        r.slot = l == 0 ? slot++ : 0;
        all_tiles_[l].push_back(r);
      }
    }
    for (int l = 0; l < lmax_; l++) tile_begin_[l][i + 1] = (int)all_tiles_[l].size();
    I.tn = slot - I.tb;
    I.cb = (int)chunks.size();
    for (int64_t c0 = 0; c0 < I.n; c0 += kChunk)
      chunks.push_back(BatchChunk{i, (int)c0, (int)std::min<int64_t>(I.n, c0 + kChunk)});
    I.cn = (int)chunks.size() - I.cb;
  }
  nslots_ = slot;
  nchunks_ = (int)chunks.size();
  levels_ = balloc<BatchLevel>(lv.size(), owned_, err);
  isl_d_ = balloc<BatchIsland>(n, owned_, err);
  tiles_.assign(lmax_, nullptr);
  ntiles_.assign(lmax_, 0);
  for (int l = 0; l < lmax_; l++)
    if (!all_tiles_[l].empty())
      tiles_[l] = balloc<BatchTile>(all_tiles_[l].size() + 8 * 4, owned_, err);  // group padding
  chunks_ = balloc<BatchChunk>(nchunks_, owned_, err);
  pcell_ = balloc<double>(nchunks_, owned_, err);
  pres_ = balloc<double>(nslots_, owned_, err);
  groups_.assign(lmax_, {});
  {
    std::vector<int> every(n, 1);
    const std::vector<BatchTile> o = grouped(0, every.data(), groups0_);
    nall0_ = (int)o.size();
    all0_ = balloc<BatchTile>(o.size(), owned_, err);
    if (err == hipSuccess)
      chk(hipMemcpy(all0_, o.data(), o.size() * sizeof(BatchTile), hipMemcpyHostToDevice), err);
  }
  flags_ = balloc<int>(n, owned_, err);
  sc_ = balloc<PcgScalars>(n, owned_, err);
  st_ = balloc<BatchState>(n, owned_, err);
  if (st_) chk(hipMemset(st_, 0, n * sizeof(BatchState)), err);  // floor0 = 0 (the default)
  ppq_ = balloc<double>(nslots_, owned_, err);
  prr_ = balloc<double>(nslots_, owned_, err);
  prz_ = balloc<double>(nslots_, owned_, err);
  coef_ = balloc<float>(2 * 64, owned_, err);
  live_ = balloc<int>(2, owned_, err);
  if (err != hipSuccess) return err;
  if (chk(hipHostMalloc((void**)&live_host_, 2 * sizeof(int)), err)) return err;
  flags_host_.assign(n, 1);
  hipStream_t s = stream_;
  chk(hipMemcpyAsync(levels_, lv.data(), lv.size() * sizeof(BatchLevel), hipMemcpyHostToDevice,
                     s),
      err);
  chk(hipMemcpyAsync(chunks_, chunks.data(), chunks.size() * sizeof(BatchChunk),
                     hipMemcpyHostToDevice, s),
      err);
  // the island records hold every pointer the V-cycle kernels follow (the
  // caller's b / x / phi come with each solve), so apply_preconditioner()
  // may run straight after attach
  chk(hipMemcpyAsync(isl_d_, host_isl_.data(), n * sizeof(BatchIsland), hipMemcpyHostToDevice,
                     s),
      err);
  // every island uses the same coarse-solve polynomial (same bound, steps)
  chk(hipMemcpyAsync(coef_, sol_[0]->chebyshev_coefficients(), 2 * 64 * sizeof(float),
                     hipMemcpyHostToDevice, s),
      err);
  chk(hipMemsetAsync(sc_, 0, n * sizeof(PcgScalars), s), err);
  chk(hipMemsetAsync(ppq_, 0, nslots_ * sizeof(double), s), err);
  chk(hipMemsetAsync(prr_, 0, nslots_ * sizeof(double), s), err);
  chk(hipMemsetAsync(prz_, 0, nslots_ * sizeof(double), s), err);
  chk(hipStreamSynchronize(s), err);
  if (err != hipSuccess) return err;
  return select_islands(flags_host_.data());
}

// Level l's tiles of the islands with live[i] != 0, one run per tile depth
// (ascending), each run in XCD order; `g` receives the runs.
std::vector<BatchTile> BatchSolver::grouped(int l, const int* live,
                                            std::vector<TileGroup>& g) const {
  std::vector<BatchTile> out;
  g.clear();
  const int n = (int)sol_.size();
  for (int tz : {2, 4, 8, 16}) {
    std::vector<BatchTile> v;
    for (int i = 0; i < n; i++)
      if (live[i] && l < sol_[i]->levels() && sol_[i]->level(l).tiles.tz == tz)
        v.insert(v.end(), all_tiles_[l].begin() + tile_begin_[l][i],
                 all_tiles_[l].begin() + tile_begin_[l][i + 1]);
    if (v.empty()) continue;
    const std::vector<BatchTile> o = xcd_order(v);
    g.push_back(TileGroup{tz, (int)out.size(), (int)o.size()});
    out.insert(out.end(), o.begin(), o.end());
  }
  return out;
}

// Launch lists holding only the tiles of islands with live[i] != 0, in
// XCD-grouped order; the graph is re-captured at its next use.
hipError_t BatchSolver::select_islands(const int* live) {
  hipError_t err = hipSuccess;
  const int n = (int)sol_.size();
  listed_.assign(live, live + n);
  listed_tiles_ = 0;
  std::vector<std::vector<BatchTile>> order(lmax_);  // alive until the copies are done
  for (int l = 0; l < lmax_; l++) {
    order[l] = grouped(l, live, groups_[l]);
    if (l == 0)
      for (int i = 0; i < n; i++)
        if (live[i]) listed_tiles_ += tile_begin_[0][i + 1] - tile_begin_[0][i];
    ntiles_[l] = (int)order[l].size();
    if (!order[l].empty())
      chk(hipMemcpyAsync(tiles_[l], order[l].data(), order[l].size() * sizeof(BatchTile),
                         hipMemcpyHostToDevice, stream_),
          err);
  }
  chk(hipStreamSynchronize(stream_), err);
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  graph_exec_ = nullptr;
  graph_ = nullptr;
  return err;
}

// Two iterations plus the live-island flags, captured for replay.
hipError_t BatchSolver::capture() {
  hipError_t err = hipSuccess;
  hipStream_t s = stream_;
  const int n = (int)sol_.size();
  if (chk(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed), err)) return err;
  (void)iteration(0);
  (void)iteration(1);
  kb_live<<<1, 1024, 0, s>>>(sc_, n, flags_, live_ + 1);
  hipGraph_t g = nullptr;
  chk(hipStreamEndCapture(s, &g), err);
  if (err == hipSuccess && !chk(hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0), err))
    graph_ = g;
  else if (g)
    (void)hipGraphDestroy(g);
  return err;
}

// Run f(integral_constant<int, TZ>) for a finest-level (4, 8, 16) or coarse
// (2, 4) tile depth, and g(KT) for the coarse kappa type.
template <typename F>
static void with_fine_tz(int tz, F&& f) {
  if (tz == 4)
    f(std::integral_constant<int, 4>{});
  else if (tz == 16)
    f(std::integral_constant<int, 16>{});
  else
    f(std::integral_constant<int, 8>{});
}
template <typename F>
static void with_coarse_tz(int tz, F&& f) {
  if (tz == 4)
    f(std::integral_constant<int, 4>{});
  else
    f(std::integral_constant<int, 2>{});
}
template <typename F>
static void with_kt(bool bf16, F&& f) {
  if (bf16)
    f(uint16_t{});
  else
    f(float{});
}

// Coarse part of every island's V-cycle below its finest level.
hipError_t BatchSolver::vcycle_coarse(bool check) {
  hipStream_t s = stream_;
  const int n = (int)sol_.size();
  with_kt(bf16_, [&](auto kt) {
    using KT = decltype(kt);
    for (int l = 1; l + 1 < lmax_; l++)
      for (const TileGroup& g : groups_[l])
        with_coarse_tz(g.tz, [&](auto tz) {
          kb_down<decltype(tz)::value, KT, false, true><<<g.count, BS, 0, s>>>(
              tiles_[l] + g.off, levels_, isl_d_, sc_, nullptr, nullptr, 0);
        });
    kb_cheb<KT><<<n, kMaxLdsCoarse, 0, s>>>(levels_, isl_d_, coef_, coarse_sweeps_, sc_, prr_,
                                            check ? 1 : 0);
    for (int l = lmax_ - 2; l >= 1; l--)
      for (const TileGroup& g : groups_[l])
        with_coarse_tz(g.tz, [&](auto tz) {
          kb_up<decltype(tz)::value, KT, false><<<g.count, BS, 0, s>>>(
              tiles_[l] + g.off, levels_, isl_d_, sc_, nullptr, 0);
        });
  });
  return hipGetLastError();
}

// z = M r[0] and the r.z partials for the islands starting a pass.
hipError_t BatchSolver::precondition() {
  hipStream_t s = stream_;
  if (ntiles_[0] == 0) return hipSuccess;
  for (const TileGroup& g : groups_[0])
    with_fine_tz(g.tz, [&](auto tz) {
      kbs_down<decltype(tz)::value, false><<<g.count, BS, 0, s>>>(tiles_[0] + g.off, levels_,
                                                                   isl_d_, sc_, nullptr, nullptr, 0);
    });
  (void)vcycle_coarse(false);
  for (const TileGroup& g : groups_[0])
    with_fine_tz(g.tz, [&](auto tz) {
      kbs_up<decltype(tz)::value><<<g.count, BS, 0, s>>>(tiles_[0] + g.off, levels_, isl_d_, sc_,
                                                         prz_, 0);
    });
  return hipGetLastError();
}

// One PCG iteration of every live island, reading r[c], p[c].
hipError_t BatchSolver::apply_pending() {
  hipStream_t s = stream_;
  for (const TileGroup& g : groups0_)
    with_fine_tz(g.tz, [&](auto tz) {
      kb_x_pending<decltype(tz)::value><<<g.count, BS, 0, s>>>(all0_ + g.off, levels_, isl_d_,
                                                               st_, sc_);
    });
  return hipGetLastError();
}

hipError_t BatchSolver::iteration(int c) {
  hipStream_t s = stream_;
  if (ntiles_[0] == 0) return hipSuccess;
  for (const TileGroup& g : groups_[0])
    with_fine_tz(g.tz, [&](auto tz) {
      kbs_spmv<decltype(tz)::value><<<g.count, BS, 0, s>>>(tiles_[0] + g.off, levels_, isl_d_,
                                                           sc_, prz_, ppq_, c);
    });
  for (const TileGroup& g : groups_[0])
    with_fine_tz(g.tz, [&](auto tz) {
      if (c)
        kbs_down<decltype(tz)::value, true, kXPair><<<g.count, BS, 0, s>>>(
            tiles_[0] + g.off, levels_, isl_d_, sc_, ppq_, prr_, c);
      else
        kbs_down<decltype(tz)::value, true, kXDefer><<<g.count, BS, 0, s>>>(
            tiles_[0] + g.off, levels_, isl_d_, sc_, ppq_, prr_, c);
    });
  (void)vcycle_coarse(true);
  for (const TileGroup& g : groups_[0])
    with_fine_tz(g.tz, [&](auto tz) {
      kbs_up<decltype(tz)::value><<<g.count, BS, 0, s>>>(tiles_[0] + g.off, levels_, isl_d_, sc_,
                                                         prz_, 1 - c);
    });
  return hipGetLastError();
}

hipError_t BatchSolver::solve(const float* const* b, double* const* x64, float* const* phi32,
                              pfsf_batch_info* info, pfsf_solve_info* per_island) {
  hipError_t err = hipSuccess;
  hipStream_t s = stream_;
  hipStream_t caller = (hipStream_t)desc_.stream;
  const int n = (int)sol_.size();
  const double tol = desc_.tol > 0 ? desc_.tol : 1e-6;
  const int max_it = desc_.max_iters > 0 ? desc_.max_iters : 1000;
  for (int i = 0; i < n; i++) {
    host_isl_[i].b = b[i];
    host_isl_[i].xio = x64 ? x64[i] : nullptr;
    host_isl_[i].phio = phi32 ? phi32[i] : nullptr;
  }
  chk(hipEventRecord(ev0_, caller), err);  // order after the caller's stream
  chk(hipStreamWaitEvent(s, ev0_, 0), err);
  chk(hipEventRecord(ev0_, s), err);
  chk(hipMemcpyAsync(isl_d_, host_isl_.data(), n * sizeof(BatchIsland), hipMemcpyHostToDevice, s),
      err);
  kb_begin_part<<<nchunks_, BS, 0, s>>>(chunks_, isl_d_, pcell_);
  kb_begin_fin<<<n, BS, 0, s>>>(isl_d_, pcell_, st_, sc_);
  int sweeps = 0;
  const int pairs_per_check = 4;
  for (int pass = 0; pass < 18 && err == hipSuccess; pass++) {
    chk(hipMemsetAsync(live_, 0, sizeof(int), s), err);
    chk(apply_pending(), err);
    for (const TileGroup& g : groups0_)
      with_fine_tz(g.tz, [&](auto tz) {
        kb_residual_tiles<decltype(tz)::value><<<g.count, BS, 0, s>>>(all0_ + g.off, levels_,
                                                                      isl_d_, st_, pres_);
      });
    kb_residual_fin<<<n, BS, 0, s>>>(isl_d_, pres_, st_, sc_, tol, max_it, live_);
    kb_live<<<1, 1024, 0, s>>>(sc_, n, flags_, live_ + 1);
    chk(hipMemcpyAsync(flags_host_.data(), flags_, n * sizeof(int), hipMemcpyDeviceToHost, s),
        err);
    chk(hipMemcpyAsync(live_host_, live_, 2 * sizeof(int), hipMemcpyDeviceToHost, s), err);
    chk(hipStreamSynchronize(s), err);
    if (err != hipSuccess || live_host_[0] == 0) break;
    if (flags_host_ != listed_) chk(select_islands(flags_host_.data()), err);
    chk(precondition(), err);
    // every island's pass ends within max_it iterations (device cap)
    for (int it = 0; it < max_it + 2 * pairs_per_check && err == hipSuccess;
         it += 2 * pairs_per_check) {
      if (!graph_exec_) chk(capture(), err);
      for (int k = 0; k < pairs_per_check && err == hipSuccess; k++)
        chk(hipGraphLaunch(graph_exec_, s), err);
      sweeps += 2 * pairs_per_check;
      chk(hipMemcpyAsync(flags_host_.data(), flags_, n * sizeof(int), hipMemcpyDeviceToHost, s),
          err);
      chk(hipMemcpyAsync(live_host_ + 1, live_ + 1, sizeof(int), hipMemcpyDeviceToHost, s), err);
      chk(hipStreamSynchronize(s), err);
      if (live_host_[1] == 0) break;
      // drop finished islands from the launches once they hold a quarter of the tiles
      int live_tiles = 0;
      for (int i = 0; i < n; i++)
        if (flags_host_[i]) live_tiles += host_isl_[i].tn;
      if (4 * live_tiles < 3 * listed_tiles_) chk(select_islands(flags_host_.data()), err);
    }
  }
  chk(apply_pending(), err);  // a pass cut off by the pass limit
  kb_finish_part<<<nchunks_, BS, 0, s>>>(chunks_, isl_d_);
  chk(hipGetLastError(), err);
  chk(hipEventRecord(ev1_, s), err);
  chk(hipEventSynchronize(ev1_), err);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, ev0_, ev1_);
  chk(hipStreamWaitEvent(caller, ev1_, 0), err);
  std::vector<BatchState> st(n);
  chk(hipMemcpy(st.data(), st_, n * sizeof(BatchState), hipMemcpyDeviceToHost), err);
  if (err != hipSuccess) return err;
  pfsf_batch_info bi{};
  bi.islands = n;
  bi.sweeps = sweeps;
  bi.solve_ms = ms;
  for (int i = 0; i < n; i++) {
    const BatchState& S = st[i];
    bi.converged += S.converged;
    bi.max_iterations = std::max(bi.max_iterations, S.total);
    bi.total_iterations += S.total;
    bi.max_rel_residual = std::max(bi.max_rel_residual, S.rel);
    if (per_island) {
      pfsf_solve_info& o = per_island[i];
      o.iterations = S.total;
      o.refinements = S.passes;
      o.converged = S.converged;
      o.levels = nlev_[i];
      o.rel_residual = S.rel;
      o.solve_ms = ms;
    }
  }
  if (info) *info = bi;
  return err;
}

}  // namespace pfsf
