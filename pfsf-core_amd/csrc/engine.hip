// libpfsf C ABI on HIP: the reference engine surface (pfsf.h) and the additive
// device-pointer surface (pfsf_hip.h).
//
// Engine behaviour follows NP/src/pfsf_engine.cpp (argument validation
// :156-285, dirty-island tick loop :424-600, failure tuple wire format
// NP/src/core/island_buffer.cpp:930-962) with the solve replaced: each dirty
// island is solved to a relative-residual tolerance by MG-preconditioned PCG
// warm-started from its previous phi, then the history field and the failure
// scan run on the device and failures are compacted on the device.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <system_error>
#include <thread>
#include <vector>

#include "batch_solver.h"
#include "mg_solver.h"
#include "pfsf/pfsf.h"
#include "pfsf_internal.h"
#include "soa6.h"

using namespace pfsf;

namespace {

constexpr int kMaxSparse = 512;     // MAX_SPARSE_UPDATES_PER_TICK
constexpr int kSparseRecord = 48;   // bytes per record
constexpr float kPhiOrphan = 1e6f;  // PFSFConstants.PHI_ORPHAN_THRESHOLD
constexpr int kRbgsSweeps = 16;     // STEPS_MAJOR (NP/src/core/constants.h:37)
// phase-field parameters of the reference engine (NP/src/solver/dispatcher.cpp:
// 240-248, NP/src/core/constants.h:81-86): l0 = max(PHASE_FIELD_L0, 2),
// G_C_CONCRETE, PHASE_FIELD_RELAX, spectral split on

struct HostBufs {
  bool registered = false;
  void* phi = nullptr;
  void* source = nullptr;
  void* cond = nullptr;
  void* type = nullptr;
  void* rcomp = nullptr;
  void* rtens = nullptr;
  void* max_phi = nullptr;
  void* kappa = nullptr;  // optional exact per-voxel kappa
  void* stress = nullptr;
  int64_t stress_bytes = 0;
  bool lookups = false;
  void* curing = nullptr;  // hydration per voxel (registered lookups), or none
};

struct Island {
  pfsf_island_desc desc{};
  pfsf_dims dims{0, 0, 0};
  int64_t n = 0;
  bool dirty = true;
  bool topo_dirty = true;  // type or kappa changed since the hierarchy was built
  bool have_phi = false;
  // operator mode: SoA-6 faces (no per-voxel kappa registered) or kappa
  bool soa6 = false;
  HostBufs host;
  // device state
  uint8_t* type = nullptr;
  float* kappa = nullptr;  // registered kappa, or the SoA-6 proxy (largest face)
  float* cond = nullptr;
  float* source = nullptr;
  float* rcomp = nullptr;
  float* rtens = nullptr;
  float* max_phi = nullptr;
  float* phi = nullptr;
  double* x64 = nullptr;
  float* H = nullptr;
  float* dmg[2] = {nullptr, nullptr};  // phase-field damage (current, next)
  float* hyd = nullptr;                // hydration (allocated with registered lookups)
  uint8_t* code = nullptr;
  int32_t* scratch = nullptr;
  int64_t scratch_elems = 0;
  uint32_t* sparse_dev = nullptr;
  void* sparse_host = nullptr;
  uint32_t* labels = nullptr;     // component labels (N), allocated on first use
  uint32_t* lab_rec = nullptr;    // kMaxLabelComponents records of 12 words
  uint32_t* lab_scratch = nullptr;
  double* cw = nullptr;           // SoA-6 coupling table (13 N), allocated on first use
  std::unique_ptr<MgSolver> solver;
  std::unique_ptr<Soa6Solver> s6;
  int64_t extra_bytes = 0;        // SoA-6 table and solver, counted in vram_used
  int32_t builds = 0;             // hierarchies built for this island
  pfsf_solve_info last{};         // its last solve
  // type and kappa the solver's hierarchy was built from: a re-upload of
  // identical arrays keeps the hierarchy
  uint8_t* built_type = nullptr;
  float* built_kappa = nullptr;
  bool built_valid = false;
  std::vector<void*> owned;
};

}  // namespace

struct pfsf_engine_t {
  pfsf_config cfg{};
  bool available = false;
  hipStream_t stream = nullptr;
  std::map<int32_t, std::unique_ptr<Island>> islands;
  bool pcg = true;
  double tol = 1e-6;
  int32_t max_iters = 500;
  pfsf_material_fn material_fn = nullptr;
  void* material_ud = nullptr;
  pfsf_anchor_fn anchor_fn = nullptr;
  void* anchor_ud = nullptr;
  pfsf_fill_ratio_fn fill_fn = nullptr;
  void* fill_ud = nullptr;
  pfsf_curing_fn curing_fn = nullptr;
  void* curing_ud = nullptr;
  pfsf_vec3 wind{0, 0, 0};
  mutable std::mutex stats_mtx;
  float last_tick_ms = 0.f;
  int64_t vram_used = 0;
  // per tick: topology flags of the re-uploaded islands and the failure
  // tuples of all reported islands, each read back in one copy
  int32_t* flags_dev = nullptr;
  int32_t* flags_host = nullptr;  // pinned
  size_t flags_cap = 0;
  int32_t* ev_dev = nullptr;
  int32_t* ev_host = nullptr;  // pinned
  int64_t ev_cap = 0;          // tuples
};

namespace {

pfsf_config default_config() {
  pfsf_config c{};
  c.max_island_size = 50000;
  c.tick_budget_ms = 8;
  c.vram_budget_bytes = 512LL * 1024 * 1024;
  c.enable_phase_field = true;
  c.enable_multigrid = true;
  return c;
}

void free_island(Island& is) {
  is.solver.reset();
  is.s6.reset();
  for (void* p : is.owned) (void)hipFree(p);
  is.owned.clear();
  if (is.sparse_host) (void)hipHostFree(is.sparse_host);
  is.sparse_host = nullptr;
}

void free_tick_buffers(pfsf_engine_t* E) {
  if (E->flags_dev) (void)hipFree(E->flags_dev);
  if (E->flags_host) (void)hipHostFree(E->flags_host);
  if (E->ev_dev) (void)hipFree(E->ev_dev);
  if (E->ev_host) (void)hipHostFree(E->ev_host);
  E->flags_dev = E->flags_host = E->ev_dev = E->ev_host = nullptr;
  E->flags_cap = 0;
  E->ev_cap = 0;
}

template <typename T>
bool dev_alloc(Island& is, T** out, int64_t n) {
  void* p = nullptr;
  if (hipMalloc(&p, (size_t)(n > 0 ? n : 1) * sizeof(T)) != hipSuccess) return false;
  is.owned.push_back(p);
  *out = (T*)p;
  return true;
}

int64_t island_bytes(int64_t n) {
  // type, kappa, cond(6), source, rcomp, rtens, max_phi, phi, x64, H, damage(2), code, the
  // type and kappa the hierarchy was built from + solver (~2.3x fine)
  return n * (1 + 4 + 24 + 4 + 4 + 4 + 4 + 4 + 8 + 4 + 8 + 1 + 5) + n * 60;
}

bool allocate(pfsf_engine_t* E, Island& is) {
  int64_t n = is.n;
  bool ok = dev_alloc(is, &is.type, n) && dev_alloc(is, &is.kappa, n) &&
            dev_alloc(is, &is.cond, 6 * n) && dev_alloc(is, &is.source, n) &&
            dev_alloc(is, &is.rcomp, n) && dev_alloc(is, &is.rtens, n) &&
            dev_alloc(is, &is.max_phi, n) && dev_alloc(is, &is.phi, n) &&
            dev_alloc(is, &is.x64, n) && dev_alloc(is, &is.H, n) && dev_alloc(is, &is.code, n) &&
            dev_alloc(is, &is.dmg[0], n) && dev_alloc(is, &is.dmg[1], n) &&
            dev_alloc(is, &is.built_type, n) && dev_alloc(is, &is.built_kappa, n);
  if (!ok) return false;
  is.scratch_elems = compact_scratch_elems(is.dims);
  if (!dev_alloc(is, &is.scratch, is.scratch_elems)) return false;
  hipStream_t s = E->stream;
  ok = hipMemsetAsync(is.type, 0, n, s) == hipSuccess &&
       hipMemsetAsync(is.kappa, 0, 4 * n, s) == hipSuccess &&
       hipMemsetAsync(is.cond, 0, 24 * n, s) == hipSuccess &&
       hipMemsetAsync(is.source, 0, 4 * n, s) == hipSuccess &&
       hipMemsetAsync(is.rcomp, 0, 4 * n, s) == hipSuccess &&
       hipMemsetAsync(is.rtens, 0, 4 * n, s) == hipSuccess &&
       hipMemsetAsync(is.max_phi, 0, 4 * n, s) == hipSuccess &&
       hipMemsetAsync(is.phi, 0, 4 * n, s) == hipSuccess &&
       hipMemsetAsync(is.x64, 0, 8 * n, s) == hipSuccess &&
       hipMemsetAsync(is.H, 0, 4 * n, s) == hipSuccess &&
       hipMemsetAsync(is.dmg[0], 0, 4 * n, s) == hipSuccess &&
       hipMemsetAsync(is.code, 0, n, s) == hipSuccess;
  return ok && hipStreamSynchronize(s) == hipSuccess;
}

Island* find(pfsf_engine_t* E, int32_t id) {
  auto it = E->islands.find(id);
  return it == E->islands.end() ? nullptr : it->second.get();
}

// Warm start: the registered phi widened to FP64 (a non-finite entry starts
// at 0).
__global__ void k_widen(int64_t n, const float* __restrict__ a, double* __restrict__ b) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    const float v = a[i];
    b[i] = isfinite(v) ? (double)v : 0.0;
  }
}

// A non-finite entry of an uploaded source or kappa counts as 0 (no load, no
// coupling), so one bad value cannot spread NaN through a solve.
__global__ void k_finite_or_zero(int64_t n, float* __restrict__ a) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n && !isfinite(a[i])) a[i] = 0.0f;
}

// The operator data of the uploaded arrays: the registered kappa (finite),
// or from the SoA-6 faces the coupling table and the isotropic proxy kappa
// of the preconditioner.
bool derive(pfsf_engine_t* E, Island& is) {
  hipStream_t s = E->stream;
  const int64_t n = is.n;
  const unsigned nb = (unsigned)((n + 255) / 256);
  k_finite_or_zero<<<nb, 256, 0, s>>>(n, is.source);
  if (is.host.kappa) {
    is.soa6 = false;
    k_finite_or_zero<<<nb, 256, 0, s>>>(n, is.kappa);
  } else {
    is.soa6 = true;
    if (!is.cw) {
      if (!dev_alloc(is, &is.cw, kSoa6Slots * n)) return false;
      is.extra_bytes += 8 * kSoa6Slots * n;
      E->vram_used += 8 * kSoa6Slots * n;
    }
    if (launch_kappa_from_soa6(is.cond, n, is.kappa, s) != hipSuccess ||
        launch_soa6_couplings(is.dims, is.type, is.cond, is.cw, s) != hipSuccess)
      return false;
  }
  return hipGetLastError() == hipSuccess;
}

// Host DBB -> device (NP/src/core/island_buffer.cpp:427-598).
bool upload(pfsf_engine_t* E, Island& is) {
  hipStream_t s = E->stream;
  int64_t n = is.n;
  HostBufs& h = is.host;
  bool ok = hipMemcpyAsync(is.type, h.type, n, hipMemcpyHostToDevice, s) == hipSuccess &&
            hipMemcpyAsync(is.source, h.source, 4 * n, hipMemcpyHostToDevice, s) == hipSuccess &&
            hipMemcpyAsync(is.cond, h.cond, 24 * n, hipMemcpyHostToDevice, s) == hipSuccess &&
            hipMemcpyAsync(is.rcomp, h.rcomp, 4 * n, hipMemcpyHostToDevice, s) == hipSuccess &&
            hipMemcpyAsync(is.rtens, h.rtens, 4 * n, hipMemcpyHostToDevice, s) == hipSuccess &&
            hipMemcpyAsync(is.max_phi, h.max_phi, 4 * n, hipMemcpyHostToDevice, s) == hipSuccess &&
            hipMemcpyAsync(is.phi, h.phi, 4 * n, hipMemcpyHostToDevice, s) == hipSuccess;
  if (!ok) return false;
  if (h.kappa && hipMemcpyAsync(is.kappa, h.kappa, 4 * n, hipMemcpyHostToDevice, s) != hipSuccess)
    return false;
  if (h.curing) {
    if (!is.hyd && !dev_alloc(is, &is.hyd, n)) return false;
    if (hipMemcpyAsync(is.hyd, h.curing, 4 * n, hipMemcpyHostToDevice, s) != hipSuccess)
      return false;
  }
  if (!derive(E, is)) return false;
  // the registered phi is the warm start of the solve
  k_widen<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, is.phi, is.x64);
  is.topo_dirty = true;
  return hipGetLastError() == hipSuccess;
}

// Sets *diff when type or kappa (compared as bits) differs from the copy the
// hierarchy was built from.
__global__ void k_topo_diff(int64_t n, const uint8_t* __restrict__ type,
                            const uint32_t* __restrict__ kappa,
                            const uint8_t* __restrict__ built_type,
                            const uint32_t* __restrict__ built_kappa, int32_t* diff) {
  bool d = false;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256)
    d |= (type[i] != built_type[i]) | (kappa[i] != built_kappa[i]);
  if (__syncthreads_or(d) && threadIdx.x == 0) *diff = 1;
}

bool has_solver(const Island& is) { return is.soa6 ? (bool)is.s6 : (bool)is.solver; }

// Clears topo_dirty of each island whose device type and kappa equal, bit for
// bit, the arrays its hierarchy was built from: one pass over 5 B per voxel
// per island, all flags in one device array, read back with one copy and one
// stream synchronisation (against a full hierarchy rebuild per island).
bool settle_topo(pfsf_engine_t* E, const std::vector<Island*>& todo) {
  hipStream_t s = E->stream;
  std::vector<Island*> chk;
  for (Island* is : todo)
    if (has_solver(*is) && is->topo_dirty && is->built_valid) chk.push_back(is);
  if (chk.empty()) return true;
  if (chk.size() > E->flags_cap) {
    if (E->flags_dev) (void)hipFree(E->flags_dev);
    if (E->flags_host) (void)hipHostFree(E->flags_host);
    E->flags_dev = E->flags_host = nullptr;
    E->flags_cap = 0;
    const size_t cap = std::max<size_t>(64, 2 * chk.size());
    if (hipMalloc((void**)&E->flags_dev, cap * sizeof(int32_t)) != hipSuccess) return false;
    if (hipHostMalloc((void**)&E->flags_host, cap * sizeof(int32_t)) != hipSuccess) return false;
    E->flags_cap = cap;
  }
  if (hipMemsetAsync(E->flags_dev, 0, chk.size() * sizeof(int32_t), s) != hipSuccess)
    return false;
  for (size_t k = 0; k < chk.size(); k++) {
    Island& is = *chk[k];
    const unsigned blocks = (unsigned)std::min<int64_t>((is.n + 255) / 256, 2048);
    k_topo_diff<<<blocks, 256, 0, s>>>(is.n, is.type, (const uint32_t*)is.kappa, is.built_type,
                                       (const uint32_t*)is.built_kappa, E->flags_dev + k);
  }
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(E->flags_host, E->flags_dev, chk.size() * sizeof(int32_t),
                     hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return false;
  for (size_t k = 0; k < chk.size(); k++)
    if (E->flags_host[k] == 0) chk[k]->topo_dirty = false;
  return true;
}

// Records the type and kappa a hierarchy was just built from.
bool keep_built_topo(pfsf_engine_t* E, Island& is) {
  hipStream_t s = E->stream;
  is.built_valid =
      hipMemcpyAsync(is.built_type, is.type, is.n, hipMemcpyDeviceToDevice, s) == hipSuccess &&
      hipMemcpyAsync(is.built_kappa, is.kappa, 4 * is.n, hipMemcpyDeviceToDevice, s) ==
          hipSuccess;
  return is.built_valid;
}

pfsf_result from_hip(hipError_t e) {
  if (e == hipSuccess) return PFSF_OK;
  (void)hipGetLastError();
  return e == hipErrorOutOfMemory ? PFSF_ERROR_OUT_OF_VRAM : PFSF_ERROR_VULKAN;
}

// The island's solver, created or rebuilt when its voxels changed.  With
// multigrid on, kappa-mode solvers are batchable so that several dirty
// islands can be solved in one batched solve.  A SoA-6 island's solver
// re-reads the coupling table on every solve (the faces may change while the
// type and the proxy kappa, which fix the hierarchy, do not).
pfsf_result ensure_solver(pfsf_engine_t* E, Island& is) {
  if (is.soa6) {
    is.solver.reset();
    if (!is.s6) {
      auto sv = std::make_unique<Soa6Solver>();
      hipError_t e = sv->create(is.dims, is.type, is.cond, is.cw, is.kappa,
                                E->cfg.enable_multigrid != 0, E->tol, E->max_iters, E->stream,
                                /*batchable=*/true);
      if (e != hipSuccess) return from_hip(e);
      const int64_t b = sv->device_bytes();
      is.extra_bytes += b;
      E->vram_used += b;
      is.s6 = std::move(sv);
      is.builds++;
    } else if (is.topo_dirty) {
      hipError_t e = is.s6->rebuild();
      if (e != hipSuccess) return from_hip(e);
      is.builds++;
    }
    if (is.topo_dirty && !keep_built_topo(E, is)) return PFSF_ERROR_VULKAN;
    is.topo_dirty = false;
    return from_hip(is.s6->refresh());
  }
  is.s6.reset();
  if (is.solver && !is.topo_dirty) return PFSF_OK;
  if (!is.solver) {
    pfsf_solver_desc d{};
    d.dims = is.dims;
    d.type = is.type;
    d.kappa = is.kappa;
    d.mg_levels = E->cfg.enable_multigrid ? 0 : 1;
    d.precision = PFSF_PREC_FP32;
    d.tol = E->tol;
    d.max_iters = E->max_iters;
    d.stream = E->stream;
    auto sv = std::make_unique<MgSolver>();
    hipError_t e = sv->create(d, E->cfg.enable_multigrid != 0);
    if (e != hipSuccess) return e == hipErrorOutOfMemory ? PFSF_ERROR_OUT_OF_VRAM : from_hip(e);
    is.solver = std::move(sv);
  } else if (is.solver->update(is.type, is.kappa) != hipSuccess) {
    return PFSF_ERROR_VULKAN;
  }
  is.builds++;
  if (!keep_built_topo(E, is)) return PFSF_ERROR_VULKAN;
  is.topo_dirty = false;
  return PFSF_OK;
}

// History, phase field and failure scan of a solved island, in the
// reference's order (NP/src/pfsf_engine.cpp:484-486): H = max(H, psi_e); the
// damage step (voxels with last tick's failure codes frozen); the scan; then
// CANTILEVER for voxels the scan leaves at 0 whose damage passed 0.95.
pfsf_result post_solve(pfsf_engine_t* E, Island& is) {
  hipStream_t s = E->stream;
  hipError_t e = hipSuccess;
  const bool pf = E->cfg.enable_phase_field;
  if (pf) {
    e = is.soa6 ? launch_soa6_history(is.dims, is.type, is.cond, is.phi, is.H, nullptr, s)
                : launch_history(is.dims, is.type, is.kappa, is.phi, is.H, nullptr, s);
    if (e == hipSuccess)
      e = launch_phase_field(is.dims, is.type, is.cond, is.phi, is.H, is.hyd, is.code, is.dmg[0],
                             is.dmg[1], kPfL0, kPfGc, kPfRelax, 1, s);
    if (e != hipSuccess) return PFSF_ERROR_VULKAN;
    std::swap(is.dmg[0], is.dmg[1]);
  }
  e = is.soa6 ? launch_soa6_failure_scan(is.dims, is.type, is.cw, is.phi, is.rcomp, is.rtens,
                                         kPhiOrphan, is.code, s)
              : launch_failure_scan(is.dims, is.type, is.kappa, is.phi, is.rcomp, is.rtens,
                                    kPhiOrphan, is.code, s);
  if (e == hipSuccess && pf) e = launch_mark_fractured(is.n, is.type, is.dmg[0], is.code, s);
  if (e != hipSuccess) return PFSF_ERROR_VULKAN;
  is.have_phi = true;
  return PFSF_OK;
}

pfsf_result solve_island(pfsf_engine_t* E, Island& is) {
  hipStream_t s = E->stream;
  pfsf_solve_info info{};
  if (E->pcg) {
    pfsf_result r = ensure_solver(E, is);
    if (r != PFSF_OK) return r;
    hipError_t e = is.soa6 ? is.s6->solve(is.source, is.x64, is.phi, &info)
                           : is.solver->solve(is.source, is.x64, is.phi, &info);
    if (e != hipSuccess) return PFSF_ERROR_VULKAN;
  } else {
    // the reference's fixed smoothing schedule: no tolerance applies, so the
    // island is done for this tick (reported as converged, 0 iterations)
    hipError_t e = is.soa6 ? launch_soa6_rbgs(is.dims, is.type, is.cw, is.source, is.phi,
                                              kRbgsSweeps, s)
                           : launch_rbgs_exact(is.dims, is.type, is.kappa, is.source, is.phi,
                                               kRbgsSweeps, s);
    if (e != hipSuccess) return PFSF_ERROR_VULKAN;
    k_widen<<<(unsigned)((is.n + 255) / 256), 256, 0, s>>>(is.n, is.phi, is.x64);
    info.converged = 1;
  }
  is.last = info;
  return post_solve(E, is);
}

// The SoA-6 islands of one tick without multigrid (with it, several are
// solved as one Soa6Batch).  Each has its own solver, vectors and stream and
// its solve touches nothing else, so up to kSoa6Workers host threads run the
// solves side by side and their small kernels overlap on the GPU's queues.
// Hierarchies are prepared before, and the post-solve steps (history,
// damage, scan) run after, on the calling thread in island order.
constexpr int kSoa6Workers = 8;

pfsf_result solve_soa6_concurrent(pfsf_engine_t* E, const std::vector<Island*>& list) {
  for (Island* is : list) {
    pfsf_result r = ensure_solver(E, *is);
    if (r != PFSF_OK) return r;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return PFSF_ERROR_VULKAN;
  std::atomic<int> next{0};
  std::vector<hipError_t> errs(list.size(), hipSuccess);
  auto work = [&](bool worker) {
    if (worker && hipSetDevice(dev) != hipSuccess) return;  // the others take its share
    for (;;) {
      const int k = next.fetch_add(1);
      if (k >= (int)list.size()) return;
      Island& is = *list[k];
      pfsf_solve_info info{};
      errs[k] = is.s6->solve(is.source, is.x64, is.phi, &info);
      is.last = info;
    }
  };
  std::vector<std::thread> th;
  const int nt = std::min<int>(kSoa6Workers, (int)list.size());
  try {
    for (int t = 1; t < nt; t++) th.emplace_back(work, true);
  } catch (const std::system_error&) {
    // fewer threads: the calling thread solves what is left
  }
  work(false);
  for (std::thread& t : th) t.join();
  for (hipError_t e : errs)
    if (e != hipSuccess) return PFSF_ERROR_VULKAN;
  for (Island* is : list) {
    pfsf_result r = post_solve(E, *is);
    if (r != PFSF_OK) return r;
  }
  return PFSF_OK;
}

// Several SoA-6 dirty islands in one batched solve (Soa6Batch: the FP64 PCG
// of each island, every vector pass and the proxies' V-cycles batched).
pfsf_result solve_soa6_batch(pfsf_engine_t* E, const std::vector<Island*>& list) {
  std::vector<Soa6Solver*> sv;
  std::vector<const float*> b;
  std::vector<double*> x;
  std::vector<float*> phi;
  for (Island* is : list) {
    pfsf_result r = ensure_solver(E, *is);
    if (r != PFSF_OK) return r;
    sv.push_back(is->s6.get());
    b.push_back(is->source);
    x.push_back(is->x64);
    phi.push_back(is->phi);
  }
  Soa6Batch batch;
  hipError_t e = batch.attach(sv, E->tol, E->max_iters, E->stream);
  if (e == hipErrorInvalidValue) return solve_soa6_concurrent(E, list);  // not batchable
  if (e != hipSuccess) return from_hip(e);
  std::vector<pfsf_solve_info> per(list.size());
  e = batch.solve(b.data(), x.data(), phi.data(), per.data());
  if (e != hipSuccess) return from_hip(e);
  for (size_t k = 0; k < list.size(); k++) {
    list[k]->last = per[k];
    pfsf_result r = post_solve(E, *list[k]);
    if (r != PFSF_OK) return r;
  }
  return PFSF_OK;
}

// Several kappa-mode dirty islands in one batched solve (same numerics per
// island).
pfsf_result solve_batch(pfsf_engine_t* E, const std::vector<Island*>& todo) {
  std::vector<MgSolver*> sv;
  std::vector<const float*> b;
  std::vector<double*> x;
  std::vector<float*> phi;
  for (Island* is : todo) {
    pfsf_result r = ensure_solver(E, *is);
    if (r != PFSF_OK) return r;
    sv.push_back(is->solver.get());
    b.push_back(is->source);
    x.push_back(is->x64);
    phi.push_back(is->phi);
  }
  BatchSolver batch;
  if (batch.attach(sv, E->tol, E->max_iters, E->stream) != hipSuccess) return PFSF_ERROR_VULKAN;
  pfsf_batch_info info{};
  std::vector<pfsf_solve_info> per(todo.size());
  if (batch.solve(b.data(), x.data(), phi.data(), &info, per.data()) != hipSuccess)
    return PFSF_ERROR_VULKAN;
  for (size_t k = 0; k < todo.size(); k++) {
    todo[k]->last = per[k];
    pfsf_result r = post_solve(E, *todo[k]);
    if (r != PFSF_OK) return r;
  }
  return PFSF_OK;
}

// Failure tuples of every solved island, appended on the device in tick
// order into one {count, tuples} buffer, and the stress readbacks: all queued,
// one copy of the tuples, one stream synchronisation.  out: host {count,
// tuples} (count at out[0]) with room for `cap` tuples, or NULL.
pfsf_result report(pfsf_engine_t* E, const std::vector<Island*>& todo, int32_t* out, int64_t cap) {
  hipStream_t s = E->stream;
  if (out && cap > 0) {
    int64_t vox = 0;
    for (Island* is : todo) vox += is->n;
    cap = std::min<int64_t>(cap, vox);  // never more failures than voxels
  }
  const bool events = out && cap > 0;
  if (events && cap > E->ev_cap) {
    if (E->ev_dev) (void)hipFree(E->ev_dev);
    if (E->ev_host) (void)hipHostFree(E->ev_host);
    E->ev_dev = E->ev_host = nullptr;
    E->ev_cap = 0;
    const int64_t c = std::max<int64_t>(cap, 1024);
    if (hipMalloc((void**)&E->ev_dev, sizeof(int32_t) * (1 + 4 * (size_t)c)) != hipSuccess ||
        hipHostMalloc((void**)&E->ev_host, sizeof(int32_t) * (1 + 4 * (size_t)c)) != hipSuccess)
      return PFSF_ERROR_OUT_OF_VRAM;
    E->ev_cap = c;
  }
  if (events) {
    if (hipMemsetAsync(E->ev_dev, 0, sizeof(int32_t), s) != hipSuccess) return PFSF_ERROR_VULKAN;
    for (Island* is : todo)
      if (launch_compact(is->dims, is->code, is->desc.origin, E->ev_dev, (int32_t)cap,
                         is->scratch, is->scratch_elems, s, true) != hipSuccess)
        return PFSF_ERROR_VULKAN;
  }
  for (Island* is : todo)
    if (is->host.stress && is->host.stress_bytes > 0) {
      const int64_t cnt = std::min<int64_t>(is->host.stress_bytes / 4, is->n);
      if (hipMemcpyAsync(is->host.stress, is->phi, 4 * cnt, hipMemcpyDeviceToHost, s) !=
          hipSuccess)
        return PFSF_ERROR_VULKAN;
    }
  if (events && hipMemcpyAsync(E->ev_host, E->ev_dev, sizeof(int32_t) * (1 + 4 * (size_t)cap),
                               hipMemcpyDeviceToHost, s) != hipSuccess)
    return PFSF_ERROR_VULKAN;
  if (hipStreamSynchronize(s) != hipSuccess) return PFSF_ERROR_VULKAN;
  if (events) {
    const int32_t got = std::min<int32_t>(E->ev_host[0], (int32_t)cap);
    out[0] = got;
    if (got > 0) memcpy(out + 1, E->ev_host + 1, 16 * (size_t)got);
  }
  // a solve that stopped short of the tolerance leaves the island dirty: the
  // next tick continues from its phi
  for (Island* is : todo) is->dirty = !is->last.converged;
  return PFSF_OK;
}

// The reference walks the dirty islands one at a time, uploading and solving
// each within the tick budget (NP/src/pfsf_engine.cpp:424-600).  Here the
// islands that fit the budget are uploaded first; with PCG and multigrid on,
// two or more kappa-mode islands are then solved in one batched solve, two or
// more SoA-6 islands in another, the rest one by one; each then gets its history update and failure scan, and
// the reports of all islands come back together.
pfsf_result tick_impl(pfsf_engine_t* E, const int32_t* ids, int32_t count, int32_t* out,
                      int64_t cap) {
  if (!E->available) return PFSF_ERROR_NOT_INIT;
  if (count > 0 && !ids) return PFSF_ERROR_INVALID_ARG;
  auto t0 = std::chrono::steady_clock::now();
  if (out) out[0] = 0;
  std::vector<Island*> todo;
  for (int32_t k = 0; k < count; k++) {
    Island* is = find(E, ids[k]);
    if (!is || !is->dirty) continue;
    if (std::find(todo.begin(), todo.end(), is) != todo.end()) continue;
    auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(
                  std::chrono::steady_clock::now() - t0)
                  .count();
    if (ms >= E->cfg.tick_budget_ms) break;  // the rest stay dirty
    if (is->host.registered && !upload(E, *is)) return PFSF_ERROR_VULKAN;
    todo.push_back(is);
  }
  // a re-upload that left type and kappa unchanged (the host-buffer path
  // uploads every array on every tick) keeps the island's hierarchy, a
  // function of those two arrays only
  if (E->pcg && !settle_topo(E, todo)) return PFSF_ERROR_VULKAN;
  std::vector<Island*> batch, single;
  for (Island* is : todo) (is->soa6 ? single : batch).push_back(is);
  if (E->pcg && E->cfg.enable_multigrid && batch.size() >= 2) {
    pfsf_result r = solve_batch(E, batch);
    if (r != PFSF_OK) return r;
  } else {
    single.insert(single.begin(), batch.begin(), batch.end());
  }
  if (E->pcg && single.size() >= 2) {
    std::vector<Island*> s6;
    for (Island* is : single)
      if (is->soa6) s6.push_back(is);
    single.erase(std::remove_if(single.begin(), single.end(), [](Island* is) { return is->soa6; }),
                 single.end());
    if (!s6.empty()) {
      pfsf_result r = E->cfg.enable_multigrid && s6.size() >= 2 ? solve_soa6_batch(E, s6)
                                                                : solve_soa6_concurrent(E, s6);
      if (r != PFSF_OK) return r;
    }
  }
  for (Island* is : single) {
    pfsf_result r = solve_island(E, *is);
    if (r != PFSF_OK) return r;
  }
  pfsf_result r = report(E, todo, out, cap);
  if (r != PFSF_OK) return r;
  std::lock_guard<std::mutex> lk(E->stats_mtx);
  E->last_tick_ms =
      std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return PFSF_OK;
}

}  // namespace

#define CHECK_ENGINE(h)                        \
  if (!(h)) return PFSF_ERROR_INVALID_ARG;     \
  if (!(h)->available) return PFSF_ERROR_NOT_INIT;

extern "C" {

// ═════════════════════════════ lifecycle ═════════════════════════════════

pfsf_engine pfsf_create(const pfsf_config* config) {
  pfsf_engine_t* E = new (std::nothrow) pfsf_engine_t();
  if (!E) return nullptr;
  E->cfg = config ? *config : default_config();
  return E;
}

pfsf_engine pfsf_create_with_vulkan(const pfsf_config*, const pfsf_vulkan_handles*) {
  return nullptr;
}

pfsf_result pfsf_init(pfsf_engine E) {
  if (!E) return PFSF_ERROR_INVALID_ARG;
  if (E->available) return PFSF_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PFSF_ERROR_NO_DEVICE;
  if (hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking) != hipSuccess)
    return PFSF_ERROR_VULKAN;
  E->available = true;
  return PFSF_OK;
}

void pfsf_shutdown(pfsf_engine E) {
  if (!E || !E->available) return;
  (void)hipStreamSynchronize(E->stream);
  for (auto& kv : E->islands) free_island(*kv.second);
  E->islands.clear();
  free_tick_buffers(E);
  (void)hipStreamDestroy(E->stream);
  E->stream = nullptr;
  E->available = false;
  E->vram_used = 0;
}

void pfsf_destroy(pfsf_engine E) {
  if (!E) return;
  pfsf_shutdown(E);
  delete E;
}

bool pfsf_is_available(pfsf_engine E) { return E && E->available; }

pfsf_result pfsf_get_stats(pfsf_engine E, pfsf_stats* out) {
  CHECK_ENGINE(E);
  if (!out) return PFSF_ERROR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(E->stats_mtx);
  out->island_count = (int32_t)E->islands.size();
  int64_t v = 0;
  for (auto& kv : E->islands) v += kv.second->n;
  out->total_voxels = v;
  out->vram_used_bytes = E->vram_used;
  out->vram_budget_bytes = E->cfg.vram_budget_bytes;
  out->last_tick_ms = E->last_tick_ms;
  return PFSF_OK;
}

// ═════════════════════════════ configuration ═════════════════════════════

void pfsf_set_material_lookup(pfsf_engine E, pfsf_material_fn fn, void* ud) {
  if (E) E->material_fn = fn, E->material_ud = ud;
}
void pfsf_set_anchor_lookup(pfsf_engine E, pfsf_anchor_fn fn, void* ud) {
  if (E) E->anchor_fn = fn, E->anchor_ud = ud;
}
void pfsf_set_fill_ratio_lookup(pfsf_engine E, pfsf_fill_ratio_fn fn, void* ud) {
  if (E) E->fill_fn = fn, E->fill_ud = ud;
}
void pfsf_set_curing_lookup(pfsf_engine E, pfsf_curing_fn fn, void* ud) {
  if (E) E->curing_fn = fn, E->curing_ud = ud;
}
void pfsf_set_wind(pfsf_engine E, const pfsf_vec3* wind) {
  if (E) E->wind = wind ? *wind : pfsf_vec3{0, 0, 0};
}
void pfsf_set_pcg_enabled(pfsf_engine E, int enabled) {
  if (E) E->pcg = enabled != 0;
}
void pfsf_set_solve_tolerance(pfsf_engine E, double tol, int32_t max_iters) {
  if (!E) return;
  if (tol > 0) E->tol = tol;
  if (max_iters > 0) E->max_iters = max_iters;
  for (auto& kv : E->islands) {  // rebuilt with the new settings
    Island& is = *kv.second;
    if (is.s6) {
      E->vram_used -= is.s6->device_bytes();
      is.extra_bytes -= is.s6->device_bytes();
    }
    is.solver.reset();
    is.s6.reset();
    is.topo_dirty = true;
  }
}

// ═════════════════════════════ islands ═══════════════════════════════════

pfsf_result pfsf_add_island(pfsf_engine E, const pfsf_island_desc* desc) {
  CHECK_ENGINE(E);
  if (!desc) return PFSF_ERROR_INVALID_ARG;
  if (desc->lx <= 0 || desc->ly <= 0 || desc->lz <= 0) return PFSF_ERROR_ISLAND_FULL;
  int64_t mx = E->cfg.max_island_size;
  if (desc->lx > mx || desc->ly > mx || desc->lz > mx) return PFSF_ERROR_ISLAND_FULL;
  int64_t n = (int64_t)desc->lx * desc->ly * desc->lz;
  if (n > mx) return PFSF_ERROR_ISLAND_FULL;
  if (Island* old = find(E, desc->island_id)) {
    // same id: keep the existing buffers when the grid is unchanged
    if (old->dims.lx == desc->lx && old->dims.ly == desc->ly && old->dims.lz == desc->lz) {
      old->desc = *desc;
      return PFSF_OK;
    }
    E->vram_used -= island_bytes(old->n) + old->extra_bytes;
    free_island(*old);
    E->islands.erase(desc->island_id);
  }
  if (E->vram_used + island_bytes(n) > E->cfg.vram_budget_bytes) return PFSF_ERROR_OUT_OF_VRAM;
  auto is = std::make_unique<Island>();
  is->desc = *desc;
  is->dims = {desc->lx, desc->ly, desc->lz};
  is->n = n;
  if (!allocate(E, *is)) {
    free_island(*is);
    return PFSF_ERROR_OUT_OF_VRAM;
  }
  E->vram_used += island_bytes(n);
  E->islands[desc->island_id] = std::move(is);
  return PFSF_OK;
}

void pfsf_remove_island(pfsf_engine E, int32_t id) {
  if (!E) return;
  Island* is = find(E, id);
  if (!is) return;
  (void)hipStreamSynchronize(E->stream);
  E->vram_used -= island_bytes(is->n) + is->extra_bytes;
  free_island(*is);
  E->islands.erase(id);
}

pfsf_result pfsf_notify_block_change(pfsf_engine E, int32_t id, const pfsf_voxel_update* u) {
  CHECK_ENGINE(E);
  if (!u) return PFSF_ERROR_INVALID_ARG;
  Island* is = find(E, id);
  if (!is) return PFSF_ERROR_INVALID_ARG;
  if (u->flat_index < 0 || (int64_t)u->flat_index >= is->n) return PFSF_ERROR_INVALID_ARG;
  is->dirty = true;
  return PFSF_OK;
}

void pfsf_mark_full_rebuild(pfsf_engine E, int32_t id) {
  if (!E) return;
  if (Island* is = find(E, id)) is->dirty = true;
}

// ═════════════════════════════ buffers ═══════════════════════════════════

pfsf_result pfsf_register_island_buffers(pfsf_engine E, int32_t id,
                                         const pfsf_island_buffers* b) {
  CHECK_ENGINE(E);
  if (!b) return PFSF_ERROR_INVALID_ARG;
  if (!b->phi_addr || !b->source_addr || !b->conductivity_addr || !b->voxel_type_addr ||
      !b->rcomp_addr || !b->rtens_addr || !b->max_phi_addr)
    return PFSF_ERROR_INVALID_ARG;
  Island* is = find(E, id);
  if (!is) return PFSF_ERROR_INVALID_ARG;
  int64_t n = is->n;
  if (b->phi_bytes < 4 * n || b->source_bytes < 4 * n || b->conductivity_bytes < 24 * n ||
      b->voxel_type_bytes < n || b->rcomp_bytes < 4 * n || b->rtens_bytes < 4 * n ||
      b->max_phi_bytes < 4 * n)
    return PFSF_ERROR_INVALID_ARG;
  HostBufs& h = is->host;
  h.phi = b->phi_addr;
  h.source = b->source_addr;
  h.cond = b->conductivity_addr;
  h.type = b->voxel_type_addr;
  h.rcomp = b->rcomp_addr;
  h.rtens = b->rtens_addr;
  h.max_phi = b->max_phi_addr;
  h.registered = true;
  if (!upload(E, *is) || hipStreamSynchronize(E->stream) != hipSuccess) return PFSF_ERROR_VULKAN;
  is->dirty = true;
  return PFSF_OK;
}

pfsf_result pfsf_register_island_kappa(pfsf_engine E, int32_t id, void* kappa, int64_t bytes) {
  CHECK_ENGINE(E);
  Island* is = find(E, id);
  if (!is || !kappa || bytes < 4 * is->n) return PFSF_ERROR_INVALID_ARG;
  is->host.kappa = kappa;
  is->dirty = true;
  is->topo_dirty = true;
  return PFSF_OK;
}

pfsf_result pfsf_register_island_lookups(pfsf_engine E, int32_t id,
                                         const pfsf_island_lookups* l) {
  CHECK_ENGINE(E);
  if (!l) return PFSF_ERROR_INVALID_ARG;
  if (!l->material_id_addr || !l->anchor_bitmap_addr || !l->fluid_pressure_addr ||
      !l->curing_addr)
    return PFSF_ERROR_INVALID_ARG;
  Island* is = find(E, id);
  if (!is) return PFSF_ERROR_INVALID_ARG;
  int64_t n = is->n;
  if (l->material_id_bytes < 4 * n || l->anchor_bitmap_bytes < 8 * n ||
      l->fluid_pressure_bytes < 4 * n || l->curing_bytes < 4 * n)
    return PFSF_ERROR_INVALID_ARG;
  is->host.lookups = true;
  is->host.curing = l->curing_addr;  // hydration of the phase field
  return PFSF_OK;
}

pfsf_result pfsf_register_stress_readback(pfsf_engine E, int32_t id, void* addr, int64_t bytes) {
  CHECK_ENGINE(E);
  if (!addr || bytes <= 0) return PFSF_ERROR_INVALID_ARG;
  Island* is = find(E, id);
  if (!is) return PFSF_ERROR_INVALID_ARG;
  is->host.stress = addr;
  is->host.stress_bytes = bytes;
  return PFSF_OK;
}

// ═════════════════════════════ tick ══════════════════════════════════════

pfsf_result pfsf_tick(pfsf_engine E, const int32_t* ids, int32_t count, int64_t,
                      pfsf_tick_result* result) {
  CHECK_ENGINE(E);
  if (result) result->count = 0;
  if (result && result->events && result->capacity > 0) {
    // pfsf_failure_event is {x, y, z, type}: the tuple layout after a count
    std::vector<int32_t> buf(1 + 4 * (size_t)result->capacity);
    pfsf_result r = tick_impl(E, ids, count, buf.data(), result->capacity);
    if (r == PFSF_OK) {
      result->count = buf[0];
      memcpy(result->events, buf.data() + 1, 16 * (size_t)buf[0]);
    }
    return r;
  }
  return tick_impl(E, ids, count, nullptr, 0);
}

pfsf_result pfsf_tick_dbb(pfsf_engine E, const int32_t* ids, int32_t count, int64_t,
                          void* fail_addr, int64_t fail_bytes) {
  CHECK_ENGINE(E);
  if (fail_addr && fail_bytes >= 4) *(int32_t*)fail_addr = 0;
  const int64_t cap = (fail_addr && fail_bytes >= 20) ? (fail_bytes - 4) / 16 : 0;
  return tick_impl(E, ids, count, cap > 0 ? (int32_t*)fail_addr : nullptr, cap);
}

pfsf_result pfsf_read_stress(pfsf_engine E, int32_t id, float* out, int32_t cap,
                             int32_t* out_count) {
  CHECK_ENGINE(E);
  if (!out || !out_count) return PFSF_ERROR_INVALID_ARG;
  Island* is = find(E, id);
  if (!is) return PFSF_ERROR_INVALID_ARG;
  int64_t cnt = cap < is->n ? cap : is->n;
  if (cnt < 0) cnt = 0;
  if (cnt > 0 && (hipMemcpyAsync(out, is->phi, 4 * cnt, hipMemcpyDeviceToHost, E->stream) !=
                      hipSuccess ||
                  hipStreamSynchronize(E->stream) != hipSuccess))
    return PFSF_ERROR_VULKAN;
  *out_count = (int32_t)cnt;
  return PFSF_OK;
}

pfsf_result pfsf_get_island_status(pfsf_engine E, int32_t id, pfsf_island_status* out) {
  CHECK_ENGINE(E);
  if (!out) return PFSF_ERROR_INVALID_ARG;
  Island* is = find(E, id);
  if (!is) return PFSF_ERROR_INVALID_ARG;
  *out = pfsf_island_status{};
  out->last_solve = is->last;
  out->solver_builds = is->builds;
  out->operator_mode = is->soa6 ? 1 : 0;
  out->dirty = is->dirty ? 1 : 0;
  return PFSF_OK;
}

pfsf_result pfsf_read_damage(pfsf_engine E, int32_t id, float* out, int32_t cap,
                             int32_t* out_count) {
  CHECK_ENGINE(E);
  if (!out || !out_count) return PFSF_ERROR_INVALID_ARG;
  Island* is = find(E, id);
  if (!is) return PFSF_ERROR_INVALID_ARG;
  int64_t cnt = cap < is->n ? cap : is->n;
  if (cnt < 0) cnt = 0;
  if (cnt > 0 && (hipMemcpyAsync(out, is->dmg[0], 4 * cnt, hipMemcpyDeviceToHost, E->stream) !=
                      hipSuccess ||
                  hipStreamSynchronize(E->stream) != hipSuccess))
    return PFSF_ERROR_VULKAN;
  *out_count = (int32_t)cnt;
  return PFSF_OK;
}

pfsf_result pfsf_read_history(pfsf_engine E, int32_t id, float* out, int32_t cap,
                              int32_t* out_count) {
  CHECK_ENGINE(E);
  if (!out || !out_count) return PFSF_ERROR_INVALID_ARG;
  Island* is = find(E, id);
  if (!is) return PFSF_ERROR_INVALID_ARG;
  int64_t cnt = cap < is->n ? cap : is->n;
  if (cnt < 0) cnt = 0;
  if (cnt > 0 && (hipMemcpyAsync(out, is->H, 4 * cnt, hipMemcpyDeviceToHost, E->stream) !=
                      hipSuccess ||
                  hipStreamSynchronize(E->stream) != hipSuccess))
    return PFSF_ERROR_VULKAN;
  *out_count = (int32_t)cnt;
  return PFSF_OK;
}

pfsf_result pfsf_island_components(pfsf_engine E, int32_t id, int32_t max_components,
                                   uint32_t* header, uint32_t* records, uint32_t* labels_out) {
  CHECK_ENGINE(E);
  if (!header || !records || max_components < 1 || max_components > kMaxLabelComponents)
    return PFSF_ERROR_INVALID_ARG;
  Island* is = find(E, id);
  if (!is || !is->host.registered) return PFSF_ERROR_INVALID_ARG;
  if (!is->labels) {
    if (!dev_alloc(*is, &is->labels, is->n) ||
        !dev_alloc(*is, &is->lab_rec, 12 * (int64_t)kMaxLabelComponents) ||
        !dev_alloc(*is, &is->lab_scratch, label_scratch_words(is->dims)))
      return PFSF_ERROR_OUT_OF_VRAM;
  }
  hipStream_t s = E->stream;
  // the registered host buffer is the source of truth, as at every tick
  if (hipMemcpyAsync(is->type, is->host.type, is->n, hipMemcpyHostToDevice, s) != hipSuccess)
    return PFSF_ERROR_VULKAN;
  uint32_t hdr[4] = {0, 0, 0, 0};
  if (label_components(is->dims, is->type, 26, is->labels, is->lab_rec, max_components,
                       is->lab_scratch, hdr, s) != hipSuccess)
    return PFSF_ERROR_VULKAN;
  if (hipMemcpyAsync(records, is->lab_rec, sizeof(uint32_t) * 12 * (size_t)max_components,
                     hipMemcpyDeviceToHost, s) != hipSuccess ||
      (labels_out && hipMemcpyAsync(labels_out, is->labels, sizeof(uint32_t) * (size_t)is->n,
                                    hipMemcpyDeviceToHost, s) != hipSuccess) ||
      hipStreamSynchronize(s) != hipSuccess)
    return PFSF_ERROR_VULKAN;
  header[0] = hdr[0];
  header[1] = hdr[1];
  return PFSF_OK;
}

int32_t pfsf_drain_callbacks(pfsf_engine E, int32_t* out, int32_t cap) {
  (void)E;
  (void)out;
  (void)cap;
  return 0;
}

pfsf_result pfsf_get_sparse_upload_buffer(pfsf_engine E, int32_t id, void** out_addr,
                                          int64_t* out_bytes) {
  CHECK_ENGINE(E);
  if (!out_addr || !out_bytes) return PFSF_ERROR_INVALID_ARG;
  Island* is = find(E, id);
  if (!is) return PFSF_ERROR_INVALID_ARG;
  if (!is->sparse_host) {
    if (hipHostMalloc(&is->sparse_host, kMaxSparse * kSparseRecord) != hipSuccess)
      return PFSF_ERROR_VULKAN;
    if (!dev_alloc(*is, &is->sparse_dev, kMaxSparse * 12)) return PFSF_ERROR_OUT_OF_VRAM;
  }
  *out_addr = is->sparse_host;
  *out_bytes = (int64_t)kMaxSparse * kSparseRecord;
  return PFSF_OK;
}

pfsf_result pfsf_notify_sparse_updates(pfsf_engine E, int32_t id, int32_t count) {
  CHECK_ENGINE(E);
  if (count <= 0) return PFSF_OK;
  Island* is = find(E, id);
  if (!is || !is->sparse_host) return PFSF_ERROR_INVALID_ARG;
  if (count > kMaxSparse) count = kMaxSparse;
  hipStream_t s = E->stream;
  if (hipMemcpyAsync(is->sparse_dev, is->sparse_host, (size_t)count * kSparseRecord,
                     hipMemcpyHostToDevice, s) != hipSuccess ||
      launch_sparse_scatter(is->sparse_dev, count, is->n, is->source, is->cond, is->type,
                            is->max_phi, is->rcomp, is->rtens, s) != hipSuccess)
    return PFSF_ERROR_VULKAN;
  if (!derive(E, *is)) return PFSF_ERROR_VULKAN;
  if (hipStreamSynchronize(s) != hipSuccess) return PFSF_ERROR_VULKAN;
  is->dirty = true;
  is->topo_dirty = true;
  return PFSF_OK;
}

// ═════════════════════════════ version ═══════════════════════════════════

const char* pfsf_version(void) { return "0.1.0-hip"; }
uint32_t pfsf_abi_version(void) { return (1u << 16) | (6u << 8) | 0u; }
bool pfsf_has_feature(const char* name) {
  if (!name) return false;
  static const char* kFeat[] = {"hip.gfx950",        "solve.mg_pcg",      "solve.batched",
                                "solve.mixed_bf16",  "solve.adjoint",     "scan.compact",
                                "labels.union_find", "operator.soa6",     "smooth.chebyshev",
                                "compute.normalize", "compute.downsample"};
  for (const char* f : kFeat)
    if (strcmp(f, name) == 0) return true;
  return false;
}
const char* pfsf_build_info(void) { return "hipcc -O3 --offload-arch=gfx950 (HIP " PFSF_HIP_VER ")"; }
const char* pfsf_abi_contract_version(void) { return "1.6.0"; }

// ═════════════════════════════ device-pointer surface ════════════════════

#define DIMS_OR_FAIL(d) \
  if (!dims_ok(d)) return PFSF_ERROR_INVALID_ARG;
#define RET(e) return (e) == hipSuccess ? PFSF_OK : PFSF_ERROR_VULKAN

pfsf_result pfsf_hip_jacobi_step(pfsf_dims d, const uint8_t* type, const float* kappa,
                                 const float* source, const float* prev, float* out, float omega,
                                 void* stream) {
  DIMS_OR_FAIL(d);
  if (!type || !kappa || !source || !prev || !out || prev == out) return PFSF_ERROR_INVALID_ARG;
  RET(launch_jacobi_exact(d, type, kappa, source, prev, out, omega, (hipStream_t)stream));
}

pfsf_result pfsf_hip_rbgs(pfsf_dims d, const uint8_t* type, const float* kappa,
                          const float* source, float* phi, int32_t sweeps, void* stream) {
  DIMS_OR_FAIL(d);
  if (!type || !kappa || !source || !phi || sweeps < 0) return PFSF_ERROR_INVALID_ARG;
  RET(launch_rbgs_exact(d, type, kappa, source, phi, sweeps, (hipStream_t)stream));
}

pfsf_result pfsf_hip_failure_scan(pfsf_dims d, const uint8_t* type, const float* kappa,
                                  const float* phi, const float* rcomp, const float* rtens,
                                  float phi_orphan, uint8_t* code, void* stream) {
  DIMS_OR_FAIL(d);
  if (!type || !kappa || !phi || !rcomp || !rtens || !code) return PFSF_ERROR_INVALID_ARG;
  RET(launch_failure_scan(d, type, kappa, phi, rcomp, rtens, phi_orphan, code,
                          (hipStream_t)stream));
}

pfsf_result pfsf_hip_history_update(pfsf_dims d, const uint8_t* type, const float* kappa,
                                    const float* phi, float* H, float* psi, void* stream) {
  DIMS_OR_FAIL(d);
  if (!type || !kappa || !phi || (!H && !psi)) return PFSF_ERROR_INVALID_ARG;
  RET(launch_history(d, type, kappa, phi, H, psi, (hipStream_t)stream));
}

pfsf_result pfsf_hip_phase_field_evolve(pfsf_dims d, const uint8_t* type, const float* cond,
                                        const float* phi, const float* H, const float* hydration,
                                        uint8_t* code, const float* damage_in, float* damage_out,
                                        float l0, float gc, float relax, int32_t spectral_split,
                                        void* stream) {
  DIMS_OR_FAIL(d);
  if (!type || !cond || !phi || !H || !code || !damage_in || !damage_out ||
      damage_in == damage_out)
    return PFSF_ERROR_INVALID_ARG;
  RET(launch_phase_field(d, type, cond, phi, H, hydration, code, damage_in, damage_out, l0, gc,
                         relax, spectral_split, (hipStream_t)stream));
}

pfsf_result pfsf_hip_macro_residual(pfsf_dims d, const uint8_t* type, const float* kappa,
                                    const float* phi, const float* b, float* block_max,
                                    void* stream) {
  DIMS_OR_FAIL(d);
  if (!type || !kappa || !phi || !b || !block_max) return PFSF_ERROR_INVALID_ARG;
  RET(launch_macro_residual(d, type, kappa, phi, b, block_max, (hipStream_t)stream));
}

pfsf_result pfsf_hip_apply_f64(pfsf_dims d, const uint8_t* type, const float* kappa,
                               const double* x, double* y, void* stream) {
  DIMS_OR_FAIL(d);
  if (!type || !kappa || !x || !y || (const void*)x == (void*)y) return PFSF_ERROR_INVALID_ARG;
  RET(launch_apply_f64(d, type, kappa, x, y, (hipStream_t)stream));
}

pfsf_result pfsf_hip_compact_failures(pfsf_dims d, const uint8_t* code, pfsf_pos origin,
                                      int32_t* out, int32_t capacity, void* stream) {
  DIMS_OR_FAIL(d);
  if (!code || !out || capacity < 0) return PFSF_ERROR_INVALID_ARG;
  hipStream_t s = (hipStream_t)stream;
  int64_t ne = compact_scratch_elems(d);
  int32_t* scratch = nullptr;
  if (hipMallocAsync((void**)&scratch, sizeof(int32_t) * ne, s) != hipSuccess)
    return PFSF_ERROR_OUT_OF_VRAM;
  hipError_t e = launch_compact(d, code, origin, out, capacity, scratch, ne, s);
  (void)hipFreeAsync(scratch, s);
  RET(e);
}

pfsf_result pfsf_hip_downsample_2to1(pfsf_dims f, const float* field, const uint8_t* type,
                                     float* coarse, uint8_t* ctype, void* stream) {
  DIMS_OR_FAIL(f);
  if (!field || !type || !coarse) return PFSF_ERROR_INVALID_ARG;
  RET(launch_downsample(f, field, type, coarse, ctype, (hipStream_t)stream));
}

pfsf_result pfsf_hip_normalize_soa6(float* source, float* rcomp, float* rtens, float* cond,
                                    int64_t n, float* sigma_max_host, void* stream) {
  if (!source || !rcomp || !rtens || !cond || n <= 0) {
    if (sigma_max_host) *sigma_max_host = 1.0f;
    return (n <= 0) ? PFSF_OK : PFSF_ERROR_INVALID_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  unsigned* mx = nullptr;
  if (hipMallocAsync((void**)&mx, sizeof(unsigned), s) != hipSuccess) return PFSF_ERROR_OUT_OF_VRAM;
  unsigned one = 0x3f800000u;  // 1.0f: the maximum starts at 1
  hipError_t e = hipMemcpyAsync(mx, &one, 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = launch_max_soa6(cond, 6 * n, mx, s);
  if (e == hipSuccess) e = launch_scale3(source, rcomp, rtens, cond, n, mx, s);
  unsigned hb = one;
  if (e == hipSuccess) e = hipMemcpyAsync(&hb, mx, 4, hipMemcpyDeviceToHost, s);
  (void)hipFreeAsync(mx, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (sigma_max_host) memcpy(sigma_max_host, &hb, 4);
  RET(e);
}

pfsf_result pfsf_hip_build_island_data(pfsf_dims d, const int32_t* material_index,
                                      const pfsf_material* materials, int32_t n_materials,
                                      const uint8_t* indestructible, const uint8_t* anchor,
                                      const float* curing, const float* fill, uint8_t* type,
                                      float* source, float* conductivity, float* rcomp,
                                      float* rtens, float* max_phi, void* stream) {
  DIMS_OR_FAIL(d);
  if (!material_index || !type || !source || !conductivity || !rcomp || !rtens || !max_phi ||
      n_materials < 0 || (n_materials > 0 && !materials))
    return PFSF_ERROR_INVALID_ARG;
  RET(launch_build_island_data(d, material_index, materials, n_materials, indestructible, anchor,
                               curing, fill, type, source, conductivity, rcomp, rtens, max_phi,
                               (hipStream_t)stream));
}

pfsf_result pfsf_hip_chebyshev_jacobi(pfsf_dims d, const uint8_t* type, const float* kappa,
                                      const float* source, float* phi, float* tmp, int32_t steps,
                                      float rho_spec, void* stream) {
  DIMS_OR_FAIL(d);
  if (!type || !kappa || !source || !phi || !tmp || phi == tmp || steps < 0)
    return PFSF_ERROR_INVALID_ARG;
  RET(launch_chebyshev_jacobi(d, type, kappa, source, phi, tmp, steps, rho_spec,
                              (hipStream_t)stream));
}

pfsf_result pfsf_hip_soa6_couplings(pfsf_dims d, const uint8_t* type, const float* cond,
                                    double* cw, void* stream) {
  DIMS_OR_FAIL(d);
  if (!type || !cond || !cw) return PFSF_ERROR_INVALID_ARG;
  RET(launch_soa6_couplings(d, type, cond, cw, (hipStream_t)stream));
}

pfsf_result pfsf_hip_soa6_apply_f64(pfsf_dims d, const uint8_t* type, const double* cw,
                                    const double* x, double* y, void* stream) {
  DIMS_OR_FAIL(d);
  if (!type || !cw || !x || !y || x == y) return PFSF_ERROR_INVALID_ARG;
  RET(launch_soa6_apply_f64(d, type, cw, x, y, (hipStream_t)stream));
}

pfsf_result pfsf_hip_soa6_failure_scan(pfsf_dims d, const uint8_t* type, const double* cw,
                                       const float* phi, const float* rcomp, const float* rtens,
                                       float phi_orphan, uint8_t* code, void* stream) {
  DIMS_OR_FAIL(d);
  if (!type || !cw || !phi || !rcomp || !rtens || !code) return PFSF_ERROR_INVALID_ARG;
  RET(launch_soa6_failure_scan(d, type, cw, phi, rcomp, rtens, phi_orphan, code,
                               (hipStream_t)stream));
}

pfsf_result pfsf_hip_soa6_history_update(pfsf_dims d, const uint8_t* type, const float* cond,
                                         const float* phi, float* H, float* psi, void* stream) {
  DIMS_OR_FAIL(d);
  if (!type || !cond || !phi || (!H && !psi)) return PFSF_ERROR_INVALID_ARG;
  RET(launch_soa6_history(d, type, cond, phi, H, psi, (hipStream_t)stream));
}

pfsf_result pfsf_hip_soa6_rbgs(pfsf_dims d, const uint8_t* type, const double* cw,
                               const float* source, float* phi, int32_t sweeps, void* stream) {
  DIMS_OR_FAIL(d);
  if (!type || !cw || !source || !phi || sweeps < 0) return PFSF_ERROR_INVALID_ARG;
  RET(launch_soa6_rbgs(d, type, cw, source, phi, sweeps, (hipStream_t)stream));
}

pfsf_result pfsf_hip_kappa_from_soa6(const float* cond, int64_t n, float* kappa, void* stream) {
  if (!cond || !kappa || n <= 0) return PFSF_ERROR_INVALID_ARG;
  RET(launch_kappa_from_soa6(cond, n, kappa, (hipStream_t)stream));
}

pfsf_result pfsf_hip_anchored_mask(pfsf_dims d, const uint8_t* type, uint8_t* reach,
                                   void* stream) {
  DIMS_OR_FAIL(d);
  if (!type || !reach) return PFSF_ERROR_INVALID_ARG;
  if (nvox(d) > kMaxLabelCells) return PFSF_ERROR_INVALID_ARG;
  hipStream_t s = (hipStream_t)stream;
  uint32_t* scratch = nullptr;
  if (hipMalloc((void**)&scratch, sizeof(uint32_t) * (size_t)(nvox(d) + 1)) != hipSuccess)
    return PFSF_ERROR_OUT_OF_VRAM;
  hipError_t e = anchored_mask(d, type, reach, scratch, (int*)(scratch + nvox(d)), s);
  (void)hipFree(scratch);
  RET(e);
}

pfsf_result pfsf_hip_label_components(pfsf_dims d, const uint8_t* type, int32_t connectivity,
                                      uint32_t* island_id, uint32_t* components,
                                      int32_t max_components, int32_t* num_components,
                                      int32_t* overflow, void* stream) {
  DIMS_OR_FAIL(d);
  if (!type || !island_id || !components || (connectivity != 26 && connectivity != 6) ||
      max_components < 1 || max_components > kMaxLabelComponents || nvox(d) > kMaxLabelCells)
    return PFSF_ERROR_INVALID_ARG;
  hipStream_t s = (hipStream_t)stream;
  uint32_t* scratch = nullptr;
  if (hipMalloc((void**)&scratch, sizeof(uint32_t) * (size_t)label_scratch_words(d)) !=
      hipSuccess)
    return PFSF_ERROR_OUT_OF_VRAM;
  uint32_t hdr[4] = {0, 0, 0, 0};
  hipError_t e = label_components(d, type, connectivity, island_id, components, max_components,
                                  scratch, hdr, s);
  (void)hipFree(scratch);
  if (e == hipSuccess) {
    if (num_components) *num_components = (int32_t)hdr[0];
    if (overflow) *overflow = (int32_t)hdr[1];
  }
  RET(e);
}

// ── solver object ──────────────────────────────────────────────────────

struct pfsf_solver_t {
  MgSolver mg;
};

pfsf_result pfsf_solver_create(const pfsf_solver_desc* desc, pfsf_solver* out) {
  if (!desc || !out || !desc->type || !desc->kappa || !dims_ok(desc->dims))
    return PFSF_ERROR_INVALID_ARG;
  if (desc->precision != PFSF_PREC_FP32 && desc->precision != PFSF_PREC_BF16_COARSE)
    return PFSF_ERROR_INVALID_ARG;
  *out = nullptr;
  pfsf_solver_t* s = new (std::nothrow) pfsf_solver_t();
  if (!s) return PFSF_ERROR_OUT_OF_VRAM;
  hipError_t e = s->mg.create(*desc);
  if (e != hipSuccess) {
    delete s;
    if (e == hipErrorInvalidValue) return PFSF_ERROR_INVALID_ARG;  // grid too large
    return e == hipErrorOutOfMemory ? PFSF_ERROR_OUT_OF_VRAM : PFSF_ERROR_VULKAN;
  }
  *out = s;
  return PFSF_OK;
}

pfsf_result pfsf_solver_solve(pfsf_solver s, const float* b, double* x64, float* phi32,
                              pfsf_solve_info* info) {
  if (!s || !b) return PFSF_ERROR_INVALID_ARG;
  RET(s->mg.solve(b, x64, phi32, info));
}

pfsf_result pfsf_solver_update(pfsf_solver s, const uint8_t* type, const float* kappa) {
  if (!s || !type || !kappa) return PFSF_ERROR_INVALID_ARG;
  RET(s->mg.update(type, kappa));
}

void pfsf_solver_destroy(pfsf_solver s) { delete s; }

pfsf_result pfsf_solver_profile(pfsf_solver s, int32_t which, int32_t reps, double* avg_ms) {
  if (!s || !avg_ms || reps <= 0 || which < 0 || which > 5) return PFSF_ERROR_INVALID_ARG;
  float ms = 0.f;
  hipError_t e = s->mg.profile(which, reps, &ms);
  if (e == hipErrorInvalidValue) return PFSF_ERROR_INVALID_ARG;
  *avg_ms = ms;
  RET(e);
}

// ═════════════════════════════ batched solve ═════════════════════════════

struct pfsf_batch_t {
  pfsf::BatchSolver b;
};

pfsf_result pfsf_batch_create(const pfsf_batch_desc* desc, pfsf_batch* out) {
  if (!desc || !out || desc->count <= 0 || !desc->islands) return PFSF_ERROR_INVALID_ARG;
  if (desc->precision != PFSF_PREC_FP32 && desc->precision != PFSF_PREC_BF16_COARSE)
    return PFSF_ERROR_INVALID_ARG;
  if (desc->shape != PFSF_BATCH_AUTOTUNE && (desc->shape < 0 || desc->shape >= PFSF_BATCH_SHAPES))
    return PFSF_ERROR_INVALID_ARG;
  for (int i = 0; i < desc->count; i++) {
    const pfsf_batch_island& bi = desc->islands[i];
    if (!bi.type || !bi.kappa || !dims_ok(bi.dims)) return PFSF_ERROR_INVALID_ARG;
  }
  *out = nullptr;
  pfsf_batch_t* s = new (std::nothrow) pfsf_batch_t();
  if (!s) return PFSF_ERROR_OUT_OF_VRAM;
  hipError_t e = s->b.create(*desc);
  if (e != hipSuccess) {
    delete s;
    if (e == hipErrorInvalidValue) return PFSF_ERROR_INVALID_ARG;
    return e == hipErrorOutOfMemory ? PFSF_ERROR_OUT_OF_VRAM : PFSF_ERROR_VULKAN;
  }
  *out = s;
  return PFSF_OK;
}

pfsf_result pfsf_batch_solve(pfsf_batch s, const float* const* b, double* const* x64,
                             float* const* phi32, pfsf_batch_info* info,
                             pfsf_solve_info* per_island) {
  if (!s || !b) return PFSF_ERROR_INVALID_ARG;
  for (int i = 0; i < s->b.islands(); i++)
    if (!b[i]) return PFSF_ERROR_INVALID_ARG;
  RET(s->b.solve(b, x64, phi32, info, per_island));
}

int32_t pfsf_batch_classes(pfsf_batch s, pfsf_batch_class* out, int32_t cap) {
  if (!s) return -1;
  for (int k = 0; k < PFSF_BATCH_CLASSES && k < cap && out; k++) out[k] = s->b.shape_class(k);
  return PFSF_BATCH_CLASSES;
}

pfsf_result pfsf_batch_tuning_info(pfsf_batch s, pfsf_batch_tuning* out) {
  if (!s || !out) return PFSF_ERROR_INVALID_ARG;
  *out = s->b.tuning();
  return PFSF_OK;
}

void pfsf_batch_destroy(pfsf_batch s) { delete s; }

// ═════════════════════════════ load path / adjoint ═══════════════════════

static bool load_path_ok(const pfsf_load_path_desc* d) {
  return d && d->type && d->kappa && d->source && d->rcomp && dims_ok(d->dims) &&
         nvox(d->dims) <= pfsf::kMaxSolverCells && d->steps >= 0 && d->checkpoint_every >= 0;
}

pfsf_result pfsf_load_path_loss(const pfsf_load_path_desc* desc, int32_t fp64,
                                pfsf_load_path_info* info) {
  if (!load_path_ok(desc) || !info) return PFSF_ERROR_INVALID_ARG;
  hipError_t e = pfsf::load_path_run(*desc, fp64 != 0, nullptr, info);
  return e == hipSuccess ? PFSF_OK
                         : (e == hipErrorOutOfMemory ? PFSF_ERROR_OUT_OF_VRAM : PFSF_ERROR_VULKAN);
}

pfsf_result pfsf_load_path_grad(const pfsf_load_path_desc* desc, double* grad,
                                pfsf_load_path_info* info) {
  if (!load_path_ok(desc) || !grad || !info) return PFSF_ERROR_INVALID_ARG;
  hipError_t e = pfsf::load_path_run(*desc, false, grad, info);
  return e == hipSuccess ? PFSF_OK
                         : (e == hipErrorOutOfMemory ? PFSF_ERROR_OUT_OF_VRAM : PFSF_ERROR_VULKAN);
}

}  // extern "C"
