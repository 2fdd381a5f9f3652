// Internal declarations shared by the HIP translation units of libpfsf.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pfsf/pfsf_hip.h"

namespace pfsf {

constexpr float kEdgeP = 0.5f;          // PFSFStencil.java:42
constexpr float kCornerP = 0.1666667f;  // PFSFStencil.java:43

// Phase-field damage step of a tick (the dispatcher's values,
// NP/src/solver/dispatcher.cpp:243-246): length scale l0, fracture energy Gc,
// relaxation.
constexpr float kPfL0 = 2.0f, kPfGc = 100.0f, kPfRelax = 0.3f;

// Stencil offsets in PFSFStencil.NEIGHBOR_OFFSETS order (PFSFStencil.java:70-84):
// 6 faces, 12 edges (XY, XZ, YZ), 8 corners.  Encoded as dx,dy,dz in {-1,0,1}.
#define PFSF_OFFSETS                                                              \
  {{1, 0, 0},  {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},   {0, 0, -1},   \
   {1, 1, 0},  {1, -1, 0},  {-1, 1, 0},  {-1, -1, 0}, {1, 0, 1},   {1, 0, -1},   \
   {-1, 0, 1}, {-1, 0, -1}, {0, 1, 1},   {0, 1, -1},  {0, -1, 1},  {0, -1, -1},  \
   {1, 1, 1},  {1, 1, -1},  {1, -1, 1},  {1, -1, -1}, {-1, 1, 1},  {-1, 1, -1},  \
   {-1, -1, 1}, {-1, -1, -1}}

__host__ __device__ inline int64_t nvox(pfsf_dims d) { return (int64_t)d.lx * d.ly * d.lz; }

inline bool dims_ok(pfsf_dims d) { return d.lx > 0 && d.ly > 0 && d.lz > 0; }

inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ── k_exact.hip: bit-exact kernels (FP64 accumulation, no contraction) ──
hipError_t launch_jacobi_exact(pfsf_dims d, const uint8_t* type, const float* kappa,
                               const float* src, const float* prev, float* out, float omega,
                               hipStream_t s);
hipError_t launch_chebyshev_jacobi(pfsf_dims d, const uint8_t* type, const float* kappa,
                                   const float* src, float* phi, float* tmp, int steps,
                                   float rho, hipStream_t s);
hipError_t launch_rbgs_exact(pfsf_dims d, const uint8_t* type, const float* kappa,
                             const float* src, float* phi, int sweeps, hipStream_t s);
hipError_t launch_failure_scan(pfsf_dims d, const uint8_t* type, const float* kappa,
                               const float* phi, const float* rcomp, const float* rtens,
                               float phi_orphan, uint8_t* code, hipStream_t s);
hipError_t launch_history(pfsf_dims d, const uint8_t* type, const float* kappa,
                          const float* phi, float* H, float* psi, hipStream_t s);
// Phase-field damage step (d_in -> d_out; code: previous codes in, 1 where
// the voxel fractures) and the CANTILEVER merge after a failure scan.
hipError_t launch_phase_field(pfsf_dims d, const uint8_t* type, const float* cond,
                              const float* phi, const float* H, const float* hydration,
                              uint8_t* code, const float* d_in, float* d_out, float l0, float gc,
                              float relax, int split, hipStream_t s);
hipError_t launch_mark_fractured(int64_t n, const uint8_t* type, const float* dmg, uint8_t* code,
                                 hipStream_t s);
// 8^3 block maxima of |b - A phi| (float per block, overwritten).
hipError_t launch_macro_residual(pfsf_dims d, const uint8_t* type, const float* kappa,
                                const float* phi, const float* b, float* blk, hipStream_t s);
hipError_t launch_apply_f64(pfsf_dims d, const uint8_t* type, const float* kappa,
                            const double* x, double* y, hipStream_t s);
// Failure tuples of `code` into out = {count, capacity x {x,y,z,code}} in
// ascending flat order.  append: the tuples follow the out[0] already there
// (several islands into one buffer, in launch order); out[0] becomes the
// clamped running count.
hipError_t launch_compact(pfsf_dims d, const uint8_t* code, pfsf_pos origin, int32_t* out,
                          int32_t capacity, int32_t* scratch, int64_t scratch_elems,
                          hipStream_t s, bool append = false);
int64_t compact_scratch_elems(pfsf_dims d);
hipError_t launch_downsample(pfsf_dims fine, const float* field, const uint8_t* type,
                             float* coarse, uint8_t* ctype, hipStream_t s);
hipError_t launch_kappa_from_soa6(const float* cond, int64_t n, float* kappa, hipStream_t s);
hipError_t launch_build_island_data(pfsf_dims d, const int32_t* mat_idx,
                                   const pfsf_material* mats, int32_t nmat,
                                   const uint8_t* indestructible, const uint8_t* anchor,
                                   const float* curing, const float* fill, uint8_t* type,
                                   float* source, float* cond, float* rcomp, float* rtens,
                                   float* max_phi, hipStream_t s);
hipError_t launch_max_soa6(const float* cond, int64_t n6, unsigned* max_bits, hipStream_t s);
hipError_t launch_scale3(float* a, float* b, float* c, float* cond, int64_t n,
                         const unsigned* max_bits, hipStream_t s);
hipError_t launch_sparse_scatter(const uint32_t* rec, int count, int64_t n, float* source,
                                 float* cond, uint8_t* type, float* max_phi, float* rcomp,
                                 float* rtens, hipStream_t s);
// ── labels.hip: connected components by union-find ──────────────────────
constexpr int64_t kMaxLabelCells = int64_t(1) << 30;
constexpr int kMaxLabelComponents = 1024;
// Words of scratch label_components needs (device, uint32).
int64_t label_scratch_words(pfsf_dims d);
// Labels (1 + smallest flat index of the component, air 0xFFFFFFFF) of the
// solid/anchor voxels, conn 26 or 6, and maxc 12-word component records;
// hdr_host[0..3] = {records, overflow, components, error}.  Synchronous.
hipError_t label_components(pfsf_dims d, const uint8_t* type, int conn, uint32_t* id,
                            uint32_t* rec, int maxc, uint32_t* scratch, uint32_t* hdr_host,
                            hipStream_t s);
// Solid voxels 26-connected through non-air voxels to an anchor: reach u8 N.
// scratch: N uint32 (overwritten); dflag: one int.  Synchronous.
hipError_t anchored_mask(pfsf_dims d, const uint8_t* type, uint8_t* reach, uint32_t* scratch,
                         int* dflag, hipStream_t s);

}  // namespace pfsf
