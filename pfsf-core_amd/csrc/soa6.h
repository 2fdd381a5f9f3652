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
  int done;  // 1 converged / capped, 2 breakdown
  int iters, max_iters, first;
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
  hipError_t create(pfsf_dims d, const uint8_t* type, const float* cond, const double* cw,
                    const float* proxy_kappa, bool multigrid, double tol, int max_iters,
                    hipStream_t caller);
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
  uint8_t* reach_ = nullptr;
  double* part_ = nullptr;
  unsigned nblk_ = 0;
  S6Scalars* sc_ = nullptr;
  S6Scalars* sc_host_ = nullptr;  // pinned
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
};

}  // namespace pfsf
