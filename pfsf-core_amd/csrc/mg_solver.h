// Multigrid-preconditioned PCG solver object (host side of mg_solver.hip).
#pragma once

#include <vector>

#include "pfsf_internal.h"

namespace pfsf {

// Device-resident scalars of the inner PCG; written only by kernels.
struct PcgScalars {
  double rz, pq, alpha, beta, rr, stop2;
  int done;  // 1 converged / capped, 2 breakdown
  int iters;
  int max_iters;
  int have_dir;
  unsigned counter[4];
};

struct IrScalars {
  double bb, rr, rel;
  unsigned counter;
  unsigned pad;
};

struct MgLevel {
  pfsf_dims d{0, 0, 0};
  int64_t n = 0;
  bool bf16 = false;
  void* kap = nullptr;  // float or bf16, 0 outside the free set
  void* sq = nullptr;   // sqrt(kappa), same storage type
  float* dg = nullptr;  // diagonal (0 outside the free set)
  float* dinv = nullptr;
  float* am = nullptr;   // anchor mass
  uint8_t* fr = nullptr; // free mask
  float* r = nullptr;    // restricted residual (levels > 0)
  float* e = nullptr;    // correction returned to the finer level (levels > 0)
  float* xs = nullptr;   // pre-smoothed iterate
  float* xb = nullptr;   // Chebyshev ping-pong (coarsest)
  float* dv = nullptr;   // Chebyshev direction (coarsest)
  float scale = 1.f;
  float omega = 1.f;
  float lmax = 2.f;
};

class MgSolver {
 public:
  ~MgSolver();
  hipError_t create(const pfsf_solver_desc& desc);
  hipError_t update(const uint8_t* type, const float* kappa);
  hipError_t solve(const float* b, double* x64, float* phi32, pfsf_solve_info* info);
  hipError_t profile(int which, int reps, float* avg_ms);
  int levels() const { return nlev_; }
  const uint8_t* free_mask() const { return levels_.empty() ? nullptr : levels_[0].fr; }
  hipStream_t stream() const { return stream_; }

 private:
  hipError_t build();
  hipError_t precondition();
  hipError_t iterate();
  void release();

  pfsf_solver_desc desc_{};
  hipStream_t stream_ = nullptr;
  pfsf_dims d0_{0, 0, 0};
  int64_t n0_ = 0;
  const uint8_t* type_ = nullptr;
  const float* kappa_ = nullptr;
  bool bf16_ = false;
  int nlev_ = 1;
  int coarse_sweeps_ = 12;
  std::vector<MgLevel> levels_;
  std::vector<float> coef_host_;
  std::vector<void*> owned_;
  PcgScalars* sc_ = nullptr;
  IrScalars* ir_ = nullptr;
  PcgScalars* sc_host_ = nullptr;  // pinned: PcgScalars followed by IrScalars
  double* partials_ = nullptr;
  uint8_t* reach_ = nullptr;
  int* flag_ = nullptr;
  float *x32_ = nullptr, *r32_ = nullptr, *p_ = nullptr, *q_ = nullptr, *z_ = nullptr;
  double *xa_ = nullptr, *xb_ = nullptr, *last_xc_ = nullptr;
  float* coef_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
};

}  // namespace pfsf
