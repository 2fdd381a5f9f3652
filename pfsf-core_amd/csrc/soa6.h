// SoA-6 operator: the symmetric 26-point operator built from six face
// conductivities per voxel (the arrays pfsf_register_island_buffers hands
// over), its bit-exact kernels, and a PCG solver for it.
#pragma once

#include <memory>

#include "mg_solver.h"
#include "pfsf_internal.h"

namespace pfsf {

// Coupling table: 13 FP64 couplings per voxel, cw[s * n + i] for the forward
// offsets k = 2 s of the stencil order.
constexpr int kSoa6Slots = 13;

hipError_t launch_soa6_couplings(pfsf_dims d, const uint8_t* type, const float* cond,
                                 double* cw, hipStream_t s);
hipError_t launch_soa6_apply_f64(pfsf_dims d, const uint8_t* type, const double* cw,
                                 const double* x, double* y, hipStream_t s);
hipError_t launch_soa6_failure_scan(pfsf_dims d, const uint8_t* type, const double* cw,
                                    const float* phi, const float* rcomp, const float* rtens,
                                    float phi_orphan, uint8_t* code, hipStream_t s);
hipError_t launch_soa6_history(pfsf_dims d, const uint8_t* type, const float* cond,
                               const float* phi, float* H, float* psi, hipStream_t s);
hipError_t launch_soa6_rbgs(pfsf_dims d, const uint8_t* type, const double* cw,
                            const float* src, float* phi, int sweeps, hipStream_t s);

// Device scalars of the SoA-6 PCG.
struct S6Scalars {
  double bb, rr, rz, pq, alpha, beta, stop2;
  int done;  // 1 converged / capped, 2 breakdown; batched solve: 3 finished
  int iters, max_iters, first;
  int pass_end;  // batched solve: `done` as the pass ended (before its true residual)
};

// PCG for the SoA-6 operator, in FP64, preconditioned by the multigrid
// V-cycle of an isotropic proxy (kappa = largest face conductivity of each
// voxel).  The proxy V-cycle is a fixed symmetric positive-definite operator,
// so PCG converges to the SoA-6 solution; the proxy only sets the rate.
// Unknowns: solid voxels 26-connected to an anchor whose SoA-6 diagonal is
// non-zero.  Voxels of that set outside the proxy's free set take a Jacobi
// step instead of the V-cycle, so the preconditioner stays definite there.
class Soa6Solver {
 public:
  ~Soa6Solver();
  // type, cond (6n), cw (13n, filled by launch_soa6_couplings) and the proxy
  // kappa are device arrays owned by the caller and read again by refresh()
  // and rebuild().
  // batchable (with multigrid): the proxy hierarchy is built with the batch
  // tile depths, so that several solvers can be solved together (Soa6Batch).
  hipError_t create(pfsf_dims d, const uint8_t* type, const float* cond, const double* cw,
                    const float* proxy_kappa, bool multigrid, double tol, int max_iters,
                    hipStream_t caller, bool batchable = false);
  // The hierarchy of the proxy and the orphan mask, after type or proxy
  // kappa changed.
  hipError_t rebuild();
  // Diagonal and free set from the current coupling table (every solve
  // after the couplings were rebuilt).
  hipError_t refresh();
  // x64: initial guess (may be NULL) and FP64 solution; phi32 may be NULL.
  // Synchronous.
  hipError_t solve(const float* b, double* x64, float* phi32, pfsf_solve_info* info);
  int64_t device_bytes() const;

 private:
  friend class Soa6Batch;
  hipError_t residual(const float* b, double* x, double* rel);  // r = b - A x, synchronises
  hipError_t queue_iteration(double* x);
  void release();

  pfsf_dims d_{0, 0, 0};
  int64_t n_ = 0;
  const uint8_t* type_ = nullptr;
  const float* cond_ = nullptr;
  const double* cw_ = nullptr;
  const float* kappa_ = nullptr;
  double tol_ = 1e-6;
  int max_iters_ = 500;
  hipStream_t caller_ = nullptr;
  std::unique_ptr<MgSolver> mg_;
  std::vector<void*> owned_;
  double *dg_ = nullptr, *r_ = nullptr, *p_ = nullptr, *q_ = nullptr, *z_ = nullptr;
  double* xo_ = nullptr;  // own iterate when the caller passes no x64
  uint8_t* free_ = nullptr;
  double* part_ = nullptr;
  unsigned nblk_ = 0;
  S6Scalars* sc_ = nullptr;
  S6Scalars* sc_host_ = nullptr;  // pinned
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
};

// One island of a batched SoA-6 solve as its kernels see it.
struct S6Isl {
  pfsf_dims d;
  int64_t n;
  const uint8_t* fr;    // SoA-6 unknowns
  const uint8_t* mgfr;  // the proxy V-cycle's free set
  const double* cw;     // coupling table
  const double* dg;     // SoA-6 diagonal
  double *x, *r, *p, *q, *z;
  const float* b;       // right-hand side
  const double* guess;  // initial guess (may be null)
  double* xio;          // solution out (FP64, may be null)
  float* phio;          // solution out (FP32, may be null)
  float* rmg;           // the proxy V-cycle's input
  const float* zmg;     // and output
  int cb, cn;           // the island's 256-cell chunks [cb, cb + cn)
};

struct S6Chunk {
  int island;
  int begin;  // first cell (256 cells from here, clipped at the island's end)
};

class BatchSolver;

// Several SoA-6 islands solved together: the same FP64 PCG per island as
// Soa6Solver (passes from the true residual, per-island scalars and stop),
// with every vector pass one launch over all islands' cells and the proxy
// V-cycles of all islands one batched preconditioner application
// (BatchSolver::apply_preconditioner), instead of ~16 launches per
// iteration per island.
class Soa6Batch {
 public:
  ~Soa6Batch();
  // solvers: created batchable with multigrid, outliving the batch; their
  // vectors are the batch's.
  hipError_t attach(const std::vector<Soa6Solver*>& solvers, double tol, int max_iters,
                    hipStream_t caller);
  // Synchronous, like Soa6Solver::solve for each island (x64[i] may be null).
  hipError_t solve(const float* const* b, double* const* x64, float* const* phi32,
                   pfsf_solve_info* per_island);

 private:
  hipError_t iteration();
  void release();

  std::vector<Soa6Solver*> sol_;
  std::unique_ptr<BatchSolver> vc_;
  hipStream_t caller_ = nullptr;
  double tol_ = 1e-6;
  int max_iters_ = 500;
  std::vector<S6Isl> host_;
  std::vector<void*> owned_;
  S6Isl* isl_ = nullptr;
  S6Chunk* ch_ = nullptr;
  int nch_ = 0;
  double* part_ = nullptr;
  S6Scalars* sc_ = nullptr;
  S6Scalars* sc_host_ = nullptr;  // pinned, one per island
  int* flags_ = nullptr;          // device: 1 = the island takes part in the pass
  int* flags_host_ = nullptr;     // pinned
  int* live_ = nullptr;           // device: islands still iterating
  int* live_host_ = nullptr;      // pinned
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
};

}  // namespace pfsf
