// Multigrid-preconditioned PCG solver object (host side of mg_solver.hip).
#pragma once

#include <algorithm>
#include <vector>

#include "pfsf_internal.h"

namespace pfsf {

// Device-resident scalars of the inner PCG; written only by kernels.
struct PcgScalars {
  double rz, pq, alpha, beta, rr, stop2;
  double rzs[2];  // r.z of the last two iterations (slot = iteration parity)
  // alpha of the iterate update an even iteration of a pass leaves pending
  // (the odd iteration after it applies both; k_x_pending one left at a stop)
  double xa;
  int done;  // 1 converged / capped, 2 breakdown
  int iters;
  int max_iters;
  int have_dir;  // a search direction exists (set after the first SpMV)
};

struct IrScalars {
  double bb, rr, rel;
};

// Stencil tiles: 32 x 8 x tz voxels per 256-thread work-group.
constexpr int kTX = 32, kTY = 8;

// Chebyshev steps of the coarsest-level solve (interval [lmax / 100, lmax])
// when it runs inside one work-group's LDS (batched and batchable solvers):
// each step costs ~1 us per iteration there.  Measured on the 128^3 island:
// 6 steps need 40 PCG iterations, 8 need 41 and 12 need 39.
constexpr int kCoarseSweeps = 6;
// Degree of the coarsest-level polynomial when it is applied as a dense
// matrix (standalone solvers whose coarsest level fits one work-group): the
// degree costs nothing per iteration there, only at set-up.  Measured PCG
// iterations for degrees 6 / 8 / 12 / 16 / 32: 128^3 38 / 36 / 35 / 35 / 35,
// 256^3 80 / 78 / 71 / 76 / 76.
constexpr int kCoarseSweepsDense = 12;
// Largest level folded into the dense coarsest solve (its cells are the
// columns of the folded matrix; 16^3 = 4,096 with an 8^3 coarsest level:
// 512 x 4,096 FP32, 8 MB), and that matrix's row length: the next of 1,024,
// 2,048, 4,096 or 8,192 columns.
constexpr int64_t kMaxFoldCells = 8192;
inline int fold_ld(int64_t n) { return n <= 1024 ? 1024 : n <= 2048 ? 2048 : n <= 4096 ? 4096 : 8192; }

// Smallest relative drop of the recursive residual a solve pass aims for
// before the true FP64 residual is checked (a pass stops at
// max(0.85 tol / rel, this); a first pass at max(0.85 tol / rel, floor0_),
// floor0_ starting here and adapting to the FP32 recursion's drift).
#ifndef PFSF_PASS_FLOOR
#define PFSF_PASS_FLOOR 2e-6
#endif

// Launch shapes of the finest level's plane-streaming kernels in a single
// solve.  Each consumer of a partial sum reads one partial per work-group of
// its producer, so with one work-group per column the partials read per
// launch grow with the square of the column count (8,192 columns at 256^3:
// 0.5 GB of L2 reads per kernel).  Lists of up to kStreamOneShot columns
// (the 128^3 island: 952) take one column per work-group, as many as 256 CUs
// hold at four 256-thread work-groups each; longer lists are walked in
// strides by kStreamGroups work-groups (a multiple of 8: three per CU, what
// the column-loop kernels' register use allows resident at once).
#ifndef PFSF_STREAM_GROUPS
#define PFSF_STREAM_GROUPS 768
#endif
constexpr int kStreamOneShot = 1024;
constexpr int kStreamGroups = PFSF_STREAM_GROUPS;
static_assert(kStreamGroups % 8 == 0, "XCD chunks need a multiple of 8 work-groups");

// Tile depths of a batchable hierarchy (finest level: 8-plane streaming
// columns; coarse levels).
constexpr int kBatchTZFine = 8, kBatchTZCoarse = 2;

struct TileList {
  int tz = 8;                     // tile depth (8, 4 or 2) chosen per level
  int ntx = 0, nty = 0, ntz = 0;  // tiles per axis
  int nlaunch = 0;                // launch size (multiple of 8, padded with -1)
  int nactive = 0;                // tiles holding at least one free voxel
  // launches cover every tile of the level in XCD-chunked order, computed
  // from the work-group index (no list read); coarse levels of a standalone
  // solver whose tiles are nearly all active
  bool dense = false;
  int* ids = nullptr;             // device: tile id per work-group, XCD-grouped
  std::vector<int> host_active;   // host: active tile ids, ascending
  // finest level of a standalone solver (columns of at most 14 planes):
  // kCellWords words per list slot saying which cells of the column the
  // streaming kernels must read (see k_cell_bits)
  uint32_t* cells = nullptr;
};

// Grids of the solver stay below 2^29 cells, so every array (FP64 included)
// is under 4 GiB and 32-bit byte offsets address it.
constexpr int64_t kMaxSolverCells = int64_t(1) << 29;

// Bump allocator over one device allocation (the batch solver's pooled
// hierarchies).  With base == nullptr it only measures.
struct DevPool {
  char* base = nullptr;
  size_t used = 0;
  void* take(size_t bytes);
};

struct MgLevel {
  pfsf_dims d{0, 0, 0};
  int64_t n = 0;
  bool bf16 = false;
  void* kap = nullptr;    // float or bf16 kappa on the free set, -1 elsewhere
  float* dg = nullptr;    // diagonal (0 outside the free set)
  float* am = nullptr;    // anchor mass (couplings to Dirichlet voxels)
  // finest level of a hierarchy (>= 2 levels), read by the plane-streaming
  // kernels in place of kap and am: kappa on the free set, -kappa (at most
  // -FLT_MIN) on anchors, -inf elsewhere, so the stencil forms the couplings
  // to anchors itself
  float* ks = nullptr;
  uint8_t* fr = nullptr;  // free mask
  uint32_t* msk = nullptr;  // coupling mask (levels > 0; bit nbit(dx,dy,dz))
  float* r = nullptr;     // restricted residual (levels > 0)
  float* e = nullptr;     // correction returned to the finer level (levels > 0)
  float* xb = nullptr;    // Chebyshev ping-pong (coarsest)
  float* dv = nullptr;    // Chebyshev direction (coarsest)
  // coarsest level of a standalone solver: its Chebyshev polynomial as a
  // dense symmetric matrix, n rows of pm_ld floats (k_cheb_dense)
  float* pm = nullptr;
  int pm_ld = 0;
  // with the dense matrix, hierarchies of at least four levels fold the
  // down-sweep of the level above into it: fg = P R (n rows of fg_ld
  // floats), R = P^T (I - A S) of that level, built from the column table
  // fpar / fval (8 parent slots per cell of that level); k_fold_dense
  float* fg = nullptr;
  int fg_ld = 0;
  int* fpar = nullptr;
  float* fval = nullptr;
  float scale = 1.f;      // coupling scale 2^level
  float omega = 1.f;      // damped-Jacobi weight
  float lmax = 2.f;       // estimate of lambda_max(D^-1 A)
  TileList tiles;
};

class MgSolver {
 public:
  ~MgSolver();
  // batchable: at least two levels and the batch solver's tile depths (4 on
  // the finest level, 2 below), so the solver can join a BatchSolver::attach.
  hipError_t create(const pfsf_solver_desc& desc, bool batchable = false);
  // Pooled hierarchy for the batch solver: shapes first (host), then arrays
  // carved from `pool` (a measuring pass first), then the set-up kernels on
  // the batch's stream.  solve() and profile() are not available.
  void configure_pooled(const pfsf_solver_desc& desc, int tz_fine, int tz_coarse);
  void allocate_pooled(DevPool& pool);
  hipError_t build_pooled(hipStream_t s);
  hipError_t update(const uint8_t* type, const float* kappa);
  // New kappa, same type: every level's operator is rebuilt on the device
  // with the tile lists and captured iteration graphs kept (the free set
  // follows type, and the couplings' sign pattern with kappa's).  Where the
  // free set does move (a coupling that vanishes or appears), a full
  // update() follows.  Ordered after the caller's stream; standalone only.
  hipError_t refresh_kappa(const float* kappa);
  hipError_t solve(const float* b, double* x64, float* phi32, pfsf_solve_info* info);
  hipError_t profile(int which, int reps, float* avg_ms);
  int levels() const { return nlev_; }
  const MgLevel& level(int l) const { return levels_[l]; }
  pfsf_dims dims() const { return d0_; }
  int64_t cells() const { return n0_; }
  const float* chebyshev_coefficients() const { return coef_host_.data(); }
  int coarse_sweeps() const { return coarse_sweeps_; }
  double* iterate() const { return x_; }
  float* residual(int k) const { return r_[k]; }
  float* direction(int k) const { return p_[k]; }
  float* product() const { return q_; }
  float* preconditioned() const { return z_; }
  const uint8_t* type() const { return type_; }
  const float* kappa() const { return kappa_; }
  const uint8_t* free_mask() const { return levels_.empty() ? nullptr : levels_[0].fr; }
  // solid voxels 26-connected to an anchor (the orphan mask of the last
  // build, on stream())
  const uint8_t* reach() const { return reach_; }
  hipStream_t stream() const { return stream_; }
  // Preconditioner-only use by an outer Krylov solve with another operator:
  // precond_begin() clears the stop flag, then each precond_apply() writes
  // z = M r into precond_out() from the residual in precond_in() (FP32, zero
  // outside free_mask()), on stream().
  float* precond_in() const { return r_[0]; }
  const float* precond_out() const { return z_; }
  hipError_t precond_begin();
  hipError_t precond_apply() { return precondition(0); }

 private:
  void configure(const pfsf_solver_desc& desc, int min_levels, int tz_fine, int tz_coarse);
  hipError_t allocate(bool standalone);
  template <typename T>
  T* alloc(int64_t n, hipError_t& err);
  void zero_vectors();
  hipError_t build();
  void setup_operator(hipStream_t s);  // kappa-dependent part of build()
  hipError_t wait_caller();            // stream_ after the caller's stream
  hipError_t build_tiles(MgLevel& L);
  hipError_t precondition(int c);       // z = M r[c], r.z
  hipError_t true_residual(const float* b, double* rel, int c, int seq = 0);  // b - A x into r[c]
  // the finest streaming kernels walk their columns in strides
  bool fine_loop() const { return levels_[0].tiles.nlaunch > kStreamOneShot; }
  // their launch size (and partial-sum count)
  unsigned fine_groups() const {
    return fine_loop() ? (unsigned)kStreamGroups : (unsigned)levels_[0].tiles.nlaunch;
  }
  hipError_t iteration(int c);         // one PCG iteration reading r[c], p[c]
  hipError_t vcycle_below(int from, bool check);  // levels from..L-1 given level `from` residual
  // n = 1, 2 or 4 iterations starting at parity c (graph replay once captured)
  hipError_t launch_iters(int n, int c);
  // Spins until the mirrored progress shows `target` convergence tests or a
  // stop; a drained stream ends the wait with the device scalars.
  hipError_t wait_progress(int target, int* iters, int* done, double* rr);
  void drop_graphs();
  void release();

  pfsf_solver_desc desc_{};
  hipStream_t stream_ = nullptr;
  bool owns_stream_ = false;
  DevPool* pool_ = nullptr;
  pfsf_dims d0_{0, 0, 0};
  int64_t n0_ = 0;
  const uint8_t* type_ = nullptr;
  const float* kappa_ = nullptr;
  bool bf16_ = false;
  int nlev_ = 1;
  int coarse_sweeps_ = 12;
  // standalone, not batchable: coarse levels launch over all their tiles and
  // the coarsest solve is a dense matrix when it fits one work-group
  bool dense_ok_ = false;
  std::vector<MgLevel> levels_;
  std::vector<float> coef_host_;
  std::vector<void*> owned_;
  PcgScalars* sc_ = nullptr;
  IrScalars* ir_ = nullptr;
  PcgScalars* sc_host_ = nullptr;  // pinned: PcgScalars followed by IrScalars
  // pinned, coherent: {(iterations << 8) | done, r.r bits}, stored by the
  // convergence test of every iteration
  unsigned long long* mirror_ = nullptr;
  double* partials_ = nullptr;  // work-group partial sums; regions below
  int64_t npartials_ = 0;
  double* ppq_ = nullptr;  // p.q (SpMV)
  double* prr_ = nullptr;  // r.r (update) [+ r.z of the Jacobi update]
  double* prz_ = nullptr;  // r.z (preconditioner)
  uint8_t* reach_ = nullptr;
  int* flag_ = nullptr;
  uint8_t* fr_snap_ = nullptr;  // refresh_kappa: every level's free set before
  double* x_ = nullptr;  // FP64 iterate
  float* r_[2] = {nullptr, nullptr};
  float* p_[2] = {nullptr, nullptr};
  float *q_ = nullptr, *z_ = nullptr;
  float* coef_ = nullptr;
  int* nz_flag_ = nullptr;  // last solve (seq) whose initial guess held a nonzero
  int* guess_flags_ = nullptr;  // per block of the guess load: a nonzero seen
  int solve_seq_ = 0;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  // PCG iteration sequences: index 2 * log2(length) + starting parity
  hipGraph_t graph_[6] = {};
  hipGraphExec_t graph_exec_[6] = {};
  double rho_pass_[4] = {};  // per-iteration r.r contraction of each pass of the last solve
  // floor of the first pass (see solve()): PFSF_PASS_FLOOR, raised while the
  // first passes of earlier solves ended well above their recursive residual
  double floor0_ = PFSF_PASS_FLOOR;
  bool use_graph_ = true;
};

// Load-path loss and adjoint gradient (adjoint.hip).
hipError_t load_path_run(const pfsf_load_path_desc& desc, bool fp64, double* grad,
                         pfsf_load_path_info* info);

}  // namespace pfsf
