// Multi-island batch solver (host side of batch_solver.hip).
#pragma once

#include <memory>
#include <vector>

#include "mg_solver.h"

namespace pfsf {

// One island at one multigrid level, as the batched kernels see it.
struct BatchLevel {
  pfsf_dims d;        // this level
  const void* kap;    // kappa on the free set, -1 elsewhere: float, or bf16 on the coarse levels of a
                      // BF16-coarse batch (finest level: anchor-coded float, MgLevel::ks)
  const float* dg;    // diagonal
  const float* am;    // anchor mass (null on the finest level)
  const uint32_t* msk;  // coupling mask (levels >= 1), null on the finest level
  float scale, omega;
  pfsf_dims c;        // next coarser level
  float* r;           // this level's right-hand side (level 0: the island's r[0])
  float* e;           // this level's correction (levels >= 1)
  float* rc;          // next level's right-hand side
  const float* ec;    // next level's correction
  int island;
  int pad_;
};

// Per-island vectors and the caller's arrays for the current solve.
struct BatchIsland {
  pfsf_dims d;
  int64_t n;
  const uint8_t* type;
  const float* kappa;
  const uint8_t* fr;  // free mask of the finest level
  float *r0, *r1, *p0, *p1, *q, *z;
  double* x;          // FP64 iterate
  const float* b;     // caller: right-hand side
  double* xio;        // caller: initial guess in, solution out (may be null)
  float* phio;        // caller: FP32 solution out (may be null)
  int coarsest;       // index of the island's coarsest BatchLevel
  int tb, tn;         // the island's finest tiles: partial-sum slots [tb, tb + tn)
  int cb, cn;         // the island's cell chunks [cb, cb + cn)
};

// A run of cells of one island (the per-cell set-up and copy-out passes).
struct BatchChunk {
  int island;
  int begin, end;
};

// One work-group of a batched stencil launch.
struct BatchTile {
  int level;  // BatchLevel index, -1 for padding
  int x0, y0, z0;
  int slot;   // partial-sum slot (finest level)
};

// A run of one launch list whose tiles share one depth (islands of
// different size classes may have different tile depths): records
// [off, off + count) of the level's list.
struct TileGroup {
  int tz, off, count;
};

// Tile-shape candidates (finest column depth, coarse tile depth), in the
// order of pfsf_batch_class::ms; candidate 0 is the default.
constexpr int kShapeFine[PFSF_BATCH_SHAPES] = {8, 4, 16, 8, 4, 16};
constexpr int kShapeCoarse[PFSF_BATCH_SHAPES] = {2, 2, 2, 4, 4, 4};

// Per-island outcome of a batched solve.
struct BatchState {
  double bb;        // ||b||^2 over the free set
  double rel;       // last true relative residual
  double prev_rel;
  int total;        // PCG iterations over all passes
  int passes;
  int finished;     // 1: converged or given up
  int converged;
  double rr0;       // ||r||^2 of the true residual a pass started from
  // the first pass's floor (0: PFSF_PASS_FLOOR), kept across solves of the
  // batch: raised while first passes end well above their recursive
  // residual (as MgSolver::solve does)
  double floor0;
};

class BatchSolver {
 public:
  ~BatchSolver();
  hipError_t create(const pfsf_batch_desc& desc);
  // The batch with island i built in tile shape kShape*[shape[i]] (create()
  // picks the shapes from desc.shape, or by tuning).
  hipError_t build(const pfsf_batch_desc& desc, const std::vector<int>& shape);
  // A batch over solvers owned by the caller (each built with two or more
  // levels and the batch tile depths, see MgSolver::create(desc, true)); the
  // solvers must outlive the batch, and their vectors are the batch's.
  hipError_t attach(const std::vector<MgSolver*>& solvers, double tol, int max_iters,
                    hipStream_t caller);
  hipError_t solve(const float* const* b, double* const* x64, float* const* phi32,
                   pfsf_batch_info* info, pfsf_solve_info* per_island);
  int islands() const { return (int)sol_.size(); }
  // Preconditioner-only use by an outer Krylov solve over the attached
  // islands (another operator per island): z = M r into each solver's
  // precond_out() from its precond_in(), every island at once, on stream().
  // An island whose scalars (island_scalars()[i].done != 0) say so is
  // skipped; kernels only, no synchronisation, capturable.
  hipError_t apply_preconditioner() { return precondition(); }
  PcgScalars* island_scalars() const { return sc_; }
  hipStream_t stream() const { return stream_; }
  pfsf_batch_class shape_class(int k) const { return classes_[k]; }
  pfsf_batch_tuning tuning() const { return tuning_; }

 private:
  // times every candidate shape on a sample of each class (`cls`: class per island)
  hipError_t tune(const pfsf_batch_desc& desc, const std::vector<int>& cls);
  hipError_t build_tables();
  hipError_t iteration(int c);
  // the FP64 iterate updates that unfinished islands' passes left pending
  hipError_t apply_pending();
  hipError_t capture();
  hipError_t select_islands(const int* live);  // tile lists of the live islands only
  std::vector<BatchTile> grouped(int l, const int* live, std::vector<TileGroup>& g) const;
  hipError_t vcycle_coarse(bool check);
  hipError_t precondition();
  void release();

  pfsf_batch_desc desc_{};
  hipStream_t stream_ = nullptr;
  std::vector<std::unique_ptr<MgSolver>> isl_;  // owned hierarchies (create)
  std::vector<MgSolver*> sol_;                   // the islands, owned or attached
  std::vector<BatchIsland> host_isl_;
  std::vector<int> nlev_;
  int lmax_ = 0;
  void* pool_mem_ = nullptr;
  std::vector<void*> owned_;
  BatchLevel* levels_ = nullptr;
  BatchIsland* isl_d_ = nullptr;
  std::vector<BatchTile*> tiles_;  // per level: launch list (live islands)
  std::vector<int> ntiles_;        // per level: launch size
  std::vector<std::vector<TileGroup>> groups_;  // per level: the list's runs of one depth
  bool bf16_ = false;              // coarse kappa in BF16 (every island)
  pfsf_batch_class classes_[PFSF_BATCH_CLASSES] = {};
  pfsf_batch_tuning tuning_{};
  std::vector<std::vector<BatchTile>> all_tiles_;  // per level: every island's tiles, by island
  std::vector<std::vector<int>> tile_begin_;       // per level: island -> first index (n + 1)
  std::vector<int> listed_;        // islands in the current launch lists
  int listed_tiles_ = 0;           // their finest tiles
  BatchChunk* chunks_ = nullptr;
  int nchunks_ = 0;
  double* pcell_ = nullptr;        // per-chunk partial sums
  double* pres_ = nullptr;         // per-tile partial sums of the FP64 residual
  BatchTile* all0_ = nullptr;      // every island's finest tiles, XCD order per depth
  int nall0_ = 0;
  std::vector<TileGroup> groups0_;  // all0_'s runs of one depth
  int* flags_ = nullptr;           // device: per island 1 while its PCG pass runs
  std::vector<int> flags_host_;
  int nslots_ = 0;
  PcgScalars* sc_ = nullptr;
  BatchState* st_ = nullptr;
  double *ppq_ = nullptr, *prr_ = nullptr, *prz_ = nullptr;
  float* coef_ = nullptr;
  int coarse_sweeps_ = 12;
  int* live_ = nullptr;       // device: islands still iterating / restarted
  int* live_host_ = nullptr;  // pinned
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
};

}  // namespace pfsf
