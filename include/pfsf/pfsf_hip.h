/*
 * pfsf_hip.h — additive device-pointer surface of libpfsf (ABI contract 1.6).
 *
 * The reference ABI (pfsf.h) hands host buffers to the engine each tick.
 * These entry points operate on device-resident arrays (hipMalloc'd, or owned
 * by any allocator) and a caller-chosen stream, so a caller that keeps its
 * islands in HBM pays no PCIe transfer.  Every array is on the dense grid
 * lx*ly*lz with flat index x + lx*(y + ly*z); voxel types are uint8
 * {0 air, 1 solid, 2 anchor}; `stream` is a hipStream_t (NULL = default
 * stream).  Calls are asynchronous on `stream` unless stated otherwise.
 *
 * The operator is the symmetric 26-point form of the Java CPU reference
 * (VoxelPhysicsCpuReference.java:66-118): for in-bounds non-air neighbours
 *   face a = (k_i + k_j)/2, edge a = sqrt(k_i k_j)*0.5f,
 *   corner a = sqrt(k_i k_j)*0.1666667f, k = max(0, kappa),
 * anchors are Dirichlet phi = 0 and air is skipped.
 */
#ifndef PFSF_HIP_H
#define PFSF_HIP_H

#include "pfsf_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t lx, ly, lz;
} pfsf_dims;

/* ── Kernel-level entry points ────────────────────────────────────────── */

/* One Jacobi step, bit-identical to VoxelPhysicsCpuReference.jacobiStep
 * (java:66-118): FP64 accumulation in PFSFStencil offset order, FP32 store,
 * phi_out = (float)(omega*(phiJ - prev) + prev).  Replaces the reference's
 * jacobi_smooth dispatch (SH/jacobi_smooth.comp.glsl:1-227). */
PFSF_API pfsf_result pfsf_hip_jacobi_step(pfsf_dims d, const uint8_t* type, const float* kappa,
                                          const float* source, const float* phi_prev,
                                          float* phi_out, float omega, void* stream);

/* `sweeps` in-place 8-colour Gauss-Seidel sweeps, colours 0..7 with colour
 * (x&1)|(y&1)<<1|(z&1)<<2 (SH/rbgs_smooth.comp.glsl:86-87,
 * NP/src/solver/jacobi_solver.cpp:163-175), each voxel updated with the
 * jacobiStep formula (omega 1).  Bit-identical to the CPU restatement.  On
 * a stream not under capture all sweeps run in one launch with the
 * couplings formed once, and the call synchronises the stream: planes of at
 * most 32 x 32 with no more planes than CUs stay in LDS (a temporary of 8 B
 * per voxel), other grids of up to 2^21 voxels read a coupling table (a
 * temporary of 224 B per voxel).  Larger grids, and streams under capture,
 * take eight colour launches per sweep, asynchronously. */
PFSF_API pfsf_result pfsf_hip_rbgs(pfsf_dims d, const uint8_t* type, const float* kappa,
                                   const float* source, float* phi, int32_t sweeps, void* stream);

/* Failure scan (SH/failure_scan.comp.glsl:36-176 precedence on the symmetric
 * operator): code 0 on non-solid; NO_SUPPORT(3) when the voxel has no
 * coupling or phi > phi_orphan; CRUSHING(2) when rcomp > 0 and
 * flux_in > rcomp; TENSION(4) when rtens > 0 and flux_out > rtens.  Fluxes
 * are FP64 sums of a_ij |phi_i - phi_j| split by sign, in stencil order. */
PFSF_API pfsf_result pfsf_hip_failure_scan(pfsf_dims d, const uint8_t* type, const float* kappa,
                                           const float* phi, const float* rcomp,
                                           const float* rtens, float phi_orphan, uint8_t* code,
                                           void* stream);

/* Elastic energy density psi_e = 0.5*kappa*|grad phi|^2 (per-axis central
 * difference over occupied face neighbours, one-sided at air/boundary, 0 if
 * neither) and the history update H = max(H, psi_e) on solid voxels.  Both
 * FP32 with a fixed expression order (bit-exact vs the CPU restatement).
 * psi may be NULL. */
PFSF_API pfsf_result pfsf_hip_history_update(pfsf_dims d, const uint8_t* type,
                                             const float* kappa, const float* phi, float* H,
                                             float* psi, void* stream);

/* Phase-field damage step (SH/phase_field_evolve.comp.glsl:50-168, the
 * hybrid AT2 form of Ambati et al. 2015).  For each solid voxel whose code is
 * 0 (a voxel with a failure code from the previous scan is frozen), over its
 * in-grid non-air face neighbours (order -x, +x, -y, +y, -z, +z):
 * lap_phi += phi_j - phi_i and sumSigma += cond[d * N + i] for finite phi_j,
 * lap_d += d_j - d_i for finite d_j; with H > 0,
 *   h = hydration clamped to [0.01, 1] (1 when hydration is NULL),
 *   Gc = gc * h * sqrt(h),
 *   H_eff = spectral_split ? max(H, max(lap_phi, 0) / max(sumSigma, 1e-12)) : H,
 *   d' = clamp((2 H_eff + Gc l0 lap_d) / max(2 H_eff + Gc / l0, 1e-8), d, 1),
 *   d' = max(clamp(d (1 - relax) + d' relax, 0, 1), d) (d if not finite: the
 *   mix can round below d by an ulp, and damage never decreases);
 * FP32 in this order.  Reads damage_in, writes damage_out (must differ: every
 * voxel reads its neighbours' old damage), and sets code = 1 (CANTILEVER)
 * where d' > 0.95.  The engine uses l0 = 2, gc = 100, relax = 0.3 and the
 * spectral split (NP/src/solver/dispatcher.cpp:240-248). */
PFSF_API pfsf_result pfsf_hip_phase_field_evolve(pfsf_dims d, const uint8_t* type,
                                                 const float* cond, const float* phi,
                                                 const float* H, const float* hydration,
                                                 uint8_t* code, const float* damage_in,
                                                 float* damage_out, float l0, float gc,
                                                 float relax, int32_t spectral_split,
                                                 void* stream);

/* Macro-block residual map (SH/failure_scan.comp.glsl:183-195, blocks of
 * 8^3 voxels): block_max[(x/8) + bx * ((y/8) + by * (z/8))], bx = ceil(lx/8),
 * by = ceil(ly/8), receives the largest |b_i - (A phi)_i| over the block's
 * solid voxels (the residual in FP64 with pfsf_hip_apply_f64's operator and
 * phi widened from FP32, rounded to FP32), 0 for blocks without one.  The
 * reference seeds its RBGS block skip from this map; a tolerance solve
 * cannot skip blocks (a local change moves phi in nearly every block), so
 * here it is a diagnostic of where a solution is still off. */
PFSF_API pfsf_result pfsf_hip_macro_residual(pfsf_dims d, const uint8_t* type,
                                             const float* kappa, const float* phi,
                                             const float* b, float* block_max, void* stream);

/* y = A x in FP64 (exact coefficients), 0 on non-solid rows. */
PFSF_API pfsf_result pfsf_hip_apply_f64(pfsf_dims d, const uint8_t* type, const float* kappa,
                                        const double* x, double* y, void* stream);

/* Compact non-zero failure codes into the DBB wire format in device memory:
 * out[0] = count (int32), then count x {x+ox, y+oy, z+oz, code} int32, in
 * ascending flat-index order, at most `capacity` tuples (the count is
 * clamped).  Replaces the host loop of NP/src/core/island_buffer.cpp:945-960
 * and SH/failure_compact.comp.glsl:27-45. */
PFSF_API pfsf_result pfsf_hip_compact_failures(pfsf_dims d, const uint8_t* code, pfsf_pos origin,
                                               int32_t* out, int32_t capacity, void* stream);

/* No counterpart in the reference: pfsf_hip_compact_failures for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same output, bit for
 * bit; it calls no allocator.  Its count/scan scratch is a library-owned
 * static device array of 1 MiB on each GPU (held until the process ends), so
 * grids of more than 268,433,408 voxels return PFSF_ERROR_INVALID_ARG.  Every
 * call on one GPU shares that array: calls that overlap on one GPU (on
 * different streams) race, and the caller must order them. */
PFSF_API pfsf_result pfsf_hip_compact_failures_noalloc(pfsf_dims d, const uint8_t* code,
                                                       pfsf_pos origin, int32_t* out,
                                                       int32_t capacity, void* stream);

/* 2:1 downsample of a per-voxel field and its types (NP/src/compute/
 * downsample.cpp:39-98): mean over in-bounds children; type anchor if
 * anchors >= solids and >= airs, else solid if solids >= airs, else air. */
PFSF_API pfsf_result pfsf_hip_downsample_2to1(pfsf_dims fine, const float* field,
                                              const uint8_t* type, float* coarse,
                                              uint8_t* coarse_type, void* stream);

/* sigma-max normalisation (PFSFDataBuilder.java:222-278,
 * NP/src/compute/normalize.cpp:29-63): s = max(1, max cond); when s > 1 the
 * source, rcomp, rtens (n each) and cond (6n) are multiplied by 1/s in place.
 * *sigma_max_host receives s (this call synchronises the stream). */
PFSF_API pfsf_result pfsf_hip_normalize_soa6(float* source, float* rcomp, float* rtens,
                                             float* cond, int64_t n, float* sigma_max_host,
                                             void* stream);

/* Island arrays built on the device from a per-voxel material index grid;
 * replaces the Java data prep PFSFDataBuilder.buildFromLookups
 * (PFSFDataBuilder.java:77-133) with the conductivity rule
 * PFSFConductivity.sigma (PFSFConductivity.java:36-97, force arm 0, no wind
 * bias).  material_index[i] in [0, n_materials) selects materials[] (device
 * array), anything else is air.  indestructible (n_materials flags), anchor
 * (per voxel), curing (per voxel hydration, clamped to [0, 1]) and fill (per
 * voxel fill ratio) may be NULL: none, none, 1 and 1.  Writes type (0 air,
 * 1 solid, 2 anchor), source = density * fill * 9.81 * 1 m^3 (FP64 rounded
 * to FP32), the SoA-6 conductivity (6n; dir * n + i; -x, +x, -y, +y, -z, +z;
 * y vertical) scaled by sqrt(curing), rcomp, rtens and max_phi (+inf); air
 * voxels get zeros.  Fluid-pressure source terms are not added. */
PFSF_API pfsf_result pfsf_hip_build_island_data(
    pfsf_dims d, const int32_t* material_index, const pfsf_material* materials,
    int32_t n_materials, const uint8_t* indestructible, const uint8_t* anchor,
    const float* curing, const float* fill, uint8_t* type, float* source, float* conductivity,
    float* rcomp, float* rtens, float* max_phi, void* stream);

/* Per-voxel kappa from SoA-6 face conductivities: kappa_i = max_d cond[d*n+i]
 * over the finite faces (0 when none is positive and finite). */
PFSF_API pfsf_result pfsf_hip_kappa_from_soa6(const float* cond, int64_t n, float* kappa,
                                              void* stream);

/* `steps` Chebyshev-accelerated Jacobi steps (SH/jacobi_smooth.comp.glsl:
 * 215-226, schedule NP/src/compute/chebyshev.cpp:30-57): step k is
 * pfsf_hip_jacobi_step with omega = pfsf_chebyshev_omega(k, rho_spec) from
 * the current phi; phi (in/out) and tmp (N floats, scratch) alternate and the
 * result ends in phi.  Bit-identical to the same steps of the CPU
 * restatement. */
PFSF_API pfsf_result pfsf_hip_chebyshev_jacobi(pfsf_dims d, const uint8_t* type,
                                               const float* kappa, const float* source,
                                               float* phi, float* tmp, int32_t steps,
                                               float rho_spec, void* stream);

/* ── SoA-6 operator ──────────────────────────────────────────────────────
 * The operator the drop-in tick uses when an island registers only its six
 * face conductivities (cond, 6N, SoA d * N + i, d = -x, +x, -y, +y, -z, +z;
 * PFSFDataBuilder.java:124-132).  The reference's shader form
 * (SH/pcg_matvec.comp.glsl:42-128) builds each weight from voxel i's faces
 * alone and is not symmetric; here a pair takes the mean of its endpoints'
 * half weights (symmetric; equal to the shader's weight when both agree):
 *   face   a = 0.5 (c_i[d] + c_j[d'])
 *   edge   a = 0.5 * 0.5 (sqrt(c_i[da] c_i[db]) + sqrt(c_j[da'] c_j[db']))
 *   corner a = 0.1666667 * 0.5 (cbrt(c_i[dx] c_i[dy] c_i[dz]) + cbrt(...))
 * with ' the opposite direction and c = cond if 0 < cond <= FLT_MAX, else 0;
 * FP64, products left to right, the cube root from correctly rounded basic
 * operations (bit-identical on any IEEE FP64 implementation).  Air and
 * out-of-grid neighbours have no coupling; anchors are Dirichlet 0.
 *
 * The coupling table cw (device, 13N doubles) holds the weights to the 13
 * forward neighbours (even stencil offsets k, slot k / 2). */
PFSF_API pfsf_result pfsf_hip_soa6_couplings(pfsf_dims d, const uint8_t* type, const float* cond,
                                             double* cw, void* stream);

/* y = A x in FP64 with the SoA-6 couplings, 0 on non-solid rows. */
PFSF_API pfsf_result pfsf_hip_soa6_apply_f64(pfsf_dims d, const uint8_t* type, const double* cw,
                                             const double* x, double* y, void* stream);

/* pfsf_hip_failure_scan on the SoA-6 couplings (same precedence, FP64 fluxes
 * in stencil order). */
PFSF_API pfsf_result pfsf_hip_soa6_failure_scan(pfsf_dims d, const uint8_t* type,
                                                const double* cw, const float* phi,
                                                const float* rcomp, const float* rtens,
                                                float phi_orphan, uint8_t* code, void* stream);

/* History update with the anisotropic energy density
 * psi = 0.5 (kx gx^2 + ky gy^2 + kz gz^2), k_a = 0.5 (c[-a] + c[+a]) (FP32,
 * this order; gradients as pfsf_hip_history_update). */
PFSF_API pfsf_result pfsf_hip_soa6_history_update(pfsf_dims d, const uint8_t* type,
                                                  const float* cond, const float* phi, float* H,
                                                  float* psi, void* stream);

/* `sweeps` 8-colour Gauss-Seidel sweeps (pfsf_hip_rbgs) with the SoA-6
 * couplings: phi_i = (b_i + sum a_ij phi_j) / sum a_ij in FP64, stencil
 * order, kept when sum a_ij <= 1e-15. */
PFSF_API pfsf_result pfsf_hip_soa6_rbgs(pfsf_dims d, const uint8_t* type, const double* cw,
                                        const float* source, float* phi, int32_t sweeps,
                                        void* stream);

/* ── Solver object: PCG to a relative-residual tolerance ─────────────── */

typedef struct pfsf_solver_t* pfsf_solver;

enum {
  PFSF_PREC_FP32 = 0,       /* FP32 smoothing/SpMV, FP32 kappa on every level */
  PFSF_PREC_BF16_COARSE = 1 /* as FP32, coarse-level kappa stored as BF16 */
};

typedef struct {
  pfsf_dims dims;
  const uint8_t* type;      /* device, N */
  const float* kappa;       /* device, N */
  int32_t mg_levels;        /* 1 = Jacobi-PCG; >= 2 geometric V-cycle levels; 0 = auto */
  int32_t precision;        /* PFSF_PREC_* */
  int32_t coarse_sweeps;    /* Chebyshev degree of the coarsest-level solve (0: 12 when that
                               level fits one work-group and is applied as a dense matrix,
                               else 6) */
  double tol;               /* ||b - A x||_2 / ||b||_2 target (FP64 residual) */
  int32_t max_iters;        /* cap on total PCG iterations */
  void* stream;
} pfsf_solver_desc;

typedef struct {
  int32_t iterations;   /* total inner PCG iterations */
  int32_t refinements;  /* FP64 iterative-refinement passes */
  int32_t converged;    /* 1 when rel_residual <= tol */
  int32_t levels;       /* multigrid levels actually built */
  double rel_residual;  /* final FP64 ||b - A x|| / ||b|| */
  double solve_ms;      /* device time of the solve (hipEvents) */
} pfsf_solve_info;

/* Builds the multigrid hierarchy and workspaces (allocates device memory).
 * Orphan components (solid voxels with no 26-connected path to an anchor) and
 * isolated voxels are excluded from the solve and keep phi = 0. */
PFSF_API pfsf_result pfsf_solver_create(const pfsf_solver_desc* desc, pfsf_solver* out);

/* Solves A x = b.  x64 (device, N doubles) is the initial guess on entry
 * (NULL = start from 0) and the FP64 solution on exit (may be NULL);
 * phi32 (device, N floats, may be NULL) receives the FP32-rounded solution.
 * Synchronous: returns after convergence or max_iters. */
PFSF_API pfsf_result pfsf_solver_solve(pfsf_solver s, const float* b, double* x64, float* phi32,
                                       pfsf_solve_info* info);

/* Re-reads kappa/type (same dims) and rebuilds the hierarchy in place. */
PFSF_API pfsf_result pfsf_solver_update(pfsf_solver s, const uint8_t* type, const float* kappa);

PFSF_API void pfsf_solver_destroy(pfsf_solver s);

/* Measurement hook: average device time (ms) of one launch of a solver
 * kernel over `reps` launches on the solver's stream with the convergence
 * stop disabled.  which: 0 finest-level direction update + SpMV + p.Ap,
 * 1 finest vector update + r.r + down-sweep (smooth, residual, restrict) of
 * an odd PCG iteration, which also moves the FP64 iterate by the last two
 * directions, 5 the same of an even iteration (the iterate update left
 * pending), 2 second-level down-sweep, 3 finest up-sweep (prolong, smooth,
 * r.z) -- each timed by the kernel's own start/stop event pair -- and 4 one
 * PCG iteration including the whole V-cycle (the average over back-to-back
 * iterations launched as the solve launches them, between events).  Use after
 * pfsf_solver_solve; it overwrites the solver's scratch vectors and its
 * internal FP64 iterate, not the caller's arrays. */
PFSF_API pfsf_result pfsf_solver_profile(pfsf_solver s, int32_t which, int32_t reps,
                                         double* avg_ms);

/* ── Batched multi-island solve ──────────────────────────────────────────
 * Replaces the reference tick's sequential per-island loop
 * (NP/src/pfsf_engine.cpp:424-600: one island's fixed 16-step schedule after
 * another) with one solve of many independent islands: every stencil launch
 * covers the tiles of all islands still iterating, each island keeps its own
 * PCG scalars, multigrid hierarchy and convergence test (the same numerics
 * as pfsf_solver_*), and the iteration launches are replayed from one HIP
 * graph.  Every island gets at least two multigrid levels. */
typedef struct {
  pfsf_dims dims;
  const uint8_t* type; /* device, N */
  const float* kappa;  /* device, N */
} pfsf_batch_island;

typedef struct {
  int32_t count;
  const pfsf_batch_island* islands; /* host array of `count` entries */
  int32_t coarse_sweeps;            /* Chebyshev sweeps on each coarsest level (0 = 12) */
  double tol;                       /* per-island ||b - A x|| / ||b|| target */
  int32_t max_iters;                /* cap on PCG iterations per island */
  void* stream;
  int32_t precision; /* PFSF_PREC_FP32 or PFSF_PREC_BF16_COARSE, for every island */
  int32_t shape;     /* tile shape candidate 0..PFSF_BATCH_SHAPES-1 for every island (0: the
                        default), or PFSF_BATCH_AUTOTUNE: timed per island-size class */
} pfsf_batch_desc;

/* Tile shapes of a batch, per island-size class (cells lx*ly*lz in
 * [min_cells, max_cells)).  Candidates 0..5, in `ms` order: (finest column
 * depth, coarse tile depth) = (8,2) (4,2) (16,2) (8,4) (4,4) (16,4); 0 is the
 * default.  With PFSF_BATCH_AUTOTUNE each class present in the batch times
 * every candidate on a sample of up to 32 of its islands (a sub-batch built
 * with that shape, 16 PCG iterations per island after a warm-up solve) and
 * keeps the fastest. */
#define PFSF_BATCH_SHAPES 6
#define PFSF_BATCH_CLASSES 4
#define PFSF_BATCH_AUTOTUNE (-1)
typedef struct {
  int64_t min_cells, max_cells;
  int32_t islands;    /* islands of the batch in this class */
  int32_t sample;     /* islands timed per candidate (0: not tuned) */
  int32_t tz_fine;    /* chosen finest-level column depth (planes) */
  int32_t tz_coarse;  /* chosen coarse-level tile depth */
  float ms[PFSF_BATCH_SHAPES]; /* tuning solve time per candidate (0: not timed) */
} pfsf_batch_class;

typedef struct {
  int32_t islands;
  int32_t converged;       /* islands that reached tol */
  int32_t max_iterations;  /* largest per-island PCG iteration count */
  int32_t sweeps;          /* batched iterations executed (graph steps) */
  int64_t total_iterations;
  double max_rel_residual;
  double solve_ms;
} pfsf_batch_info;

typedef struct pfsf_batch_t* pfsf_batch;

/* Builds every island's hierarchy in one pooled device allocation. */
PFSF_API pfsf_result pfsf_batch_create(const pfsf_batch_desc* desc, pfsf_batch* out);

/* b, x64, phi32: host arrays of `count` device pointers (N_i elements each).
 * x64[i] is island i's initial guess and FP64 solution (NULL entry or NULL
 * array: start from 0, no FP64 output); phi32[i] receives the FP32 solution
 * (may be NULL).  per_island (may be NULL) receives `count` records.
 * Synchronous. */
PFSF_API pfsf_result pfsf_batch_solve(pfsf_batch s, const float* const* b, double* const* x64,
                                      float* const* phi32, pfsf_batch_info* info,
                                      pfsf_solve_info* per_island);

/* The batch's size classes and their tile shapes (PFSF_BATCH_CLASSES
 * records; at most `cap` written).  Returns the number of classes, or -1 for
 * a NULL batch. */
PFSF_API int32_t pfsf_batch_classes(pfsf_batch s, pfsf_batch_class* out, int32_t cap);

/* The tuner's whole-batch decision.  Islands of different tile depths cannot
 * share a launch, so mixing shapes splits every stage of an iteration into
 * one launch per depth; per-class timings alone do not predict that.  The
 * tuner therefore times, on a stratified sample of the batch (islands evenly
 * spaced in size order, the largest included), one shape for every island
 * (the default and each class's fastest) and, when the classes disagree,
 * the per-class mix; an option replaces the default, and the mix the best
 * single shape, only when it is at least 3 % faster. */
typedef struct {
  int32_t sample;  /* islands of the stratified sample (0: not tuned) */
  int32_t mixed;   /* 1: every class keeps its own fastest shape */
  float mixed_ms;  /* sample time with the per-class shapes (0: not timed) */
  float uniform_ms[PFSF_BATCH_SHAPES]; /* sample time with one shape for all (0: not timed) */
} pfsf_batch_tuning;

PFSF_API pfsf_result pfsf_batch_tuning_info(pfsf_batch s, pfsf_batch_tuning* out);

PFSF_API void pfsf_batch_destroy(pfsf_batch s);

/* ── Load-path loss and its adjoint gradient dL/dkappa ───────────────────
 * No counterpart in the reference (it has no adjoint); the loss is built on
 * the failure scan's crushing flux (SH/failure_scan.comp.glsl:116-158).
 * Load steps b_t = (t/steps) * source, t = 1..steps, each solved with the
 * multigrid PCG solver to `tol`;
 *   L = sum_t sum_{free i} softplus(f_i(phi_t) - rcomp_i),
 *   f_i = sum_j a_ij (phi_j - phi_i)^+ over the non-air neighbours of i.
 * The gradient uses one adjoint solve per step with the same (symmetric)
 * operator; states are kept every `checkpoint_every` steps and each segment
 * is re-solved from its checkpoint during the backward pass. */
typedef struct {
  pfsf_dims dims;
  const uint8_t* type;  /* device, N */
  const float* kappa;   /* device, N */
  const float* source;  /* device, N: the full load b */
  const float* rcomp;   /* device, N */
  int32_t steps;            /* 0 = 50 */
  int32_t checkpoint_every; /* 0 = 10 */
  double tol;               /* per-solve relative residual (0 = 1e-6) */
  int32_t max_iters;        /* per solve (0 = 2000) */
  void* stream;
  /* 1: damage couples the steps.  After each state the phase-field history
   * H = max(H, psi_e) and one damage step (the engine's hybrid AT2 step:
   * l0 2, Gc 100, relaxation 0.3, spectral split, faces of conductivity
   * kappa_t) update the state, and the next step solves with
   * kappa_t = kappa * ((1 - 1e-3)(1 - d)^2 + 1e-3).  The gradient then runs
   * through the damage, history and psi of every step as well (FP64 state),
   * with (phi, d, H) checkpointed and each segment recomputed. */
  int32_t damage;
} pfsf_load_path_desc;

typedef struct {
  double loss;
  int32_t forward_solves, recompute_solves, adjoint_solves, checkpoints;
  int64_t pcg_iterations;
  double ms; /* device time of the whole evaluation */
} pfsf_load_path_info;

/* Loss only.  fp64 = 0: FP32 flux and softplus (the gradient's forward);
 * fp64 = 1: FP64 throughout (for finite-difference checks).  Synchronous. */
PFSF_API pfsf_result pfsf_load_path_loss(const pfsf_load_path_desc* desc, int32_t fp64,
                                         pfsf_load_path_info* info);

/* Loss (FP32 forward) and dL/dkappa into grad (device, N doubles; voxels
 * whose kappa does not enter the operator get 0).  Synchronous. */
PFSF_API pfsf_result pfsf_load_path_grad(const pfsf_load_path_desc* desc, double* grad,
                                         pfsf_load_path_info* info);

/* Orphan mask used by the solver: 1 for solid voxels 26-connected to an
 * anchor through non-air voxels, else 0 (device, N bytes).  Synchronous.
 * N <= 2^30.  Computed from the same union-find components as
 * pfsf_hip_label_components, over non-air voxels. */
PFSF_API pfsf_result pfsf_hip_anchored_mask(pfsf_dims d, const uint8_t* type, uint8_t* reach,
                                            void* stream);

/* Connected components of the live voxels (type 1 solid or 2 anchor), with
 * connectivity 26 (PFSFStencil offsets) or 6 (faces).  Replaces the
 * label-propagation pipeline label_prop_{init,iterate,summarise_alloc,
 * summarise_aggregate}.comp.glsl (PFSFLabelPropRecorder.java:75-190) and its
 * CPU oracle LabelPropagation.shiloachVishkin (LabelPropagation.java:199-270):
 *   island_id (device, N uint32): 1 + the smallest flat index of the voxel's
 *     component; 0xFFFFFFFF for every other type.  This is the fixed point of
 *     the reference's hook/jump iteration, reached here in one union-find pass
 *     whatever the component diameter (the reference stops after 12 rounds).
 *   components (device, max_components x 12 uint32): one record per root in
 *     ascending root order, {label, voxel count, anchored 0/1, 0, min x,
 *     min y, min z, 0, max x, max y, max z, 0}; slots past the last record
 *     are zero (PFSFLabelPropCpuSimulator.java:164-238 layout).
 *   *num_components = records written (<= max_components); *overflow = 1 when
 *     the grid has more components than max_components.
 * 1 <= max_components <= 1024 (the reference uses 64); N <= 2^30.
 * Synchronous (the counts are returned to host memory). */
PFSF_API pfsf_result pfsf_hip_label_components(pfsf_dims d, const uint8_t* type,
                                               int32_t connectivity, uint32_t* island_id,
                                               uint32_t* components, int32_t max_components,
                                               int32_t* num_components, int32_t* overflow,
                                               void* stream);

/* ── Engine extension: exact per-voxel kappa ────────────────────────────
 * By default a registered island is solved with the SoA-6 operator of its
 * face conductivities (above).  This host buffer (float32 x N, valid for the
 * island's lifetime) switches the island to the per-voxel kappa operator of
 * the Java CPU reference; the island's SoA-6 faces are then not read by the
 * solve, the history update or the failure scan.  A non-finite entry of
 * kappa or of the source counts as 0, as does a non-finite registered phi
 * (the warm start). */
PFSF_API pfsf_result pfsf_register_island_kappa(pfsf_engine engine, int32_t island_id,
                                                void* kappa_addr, int64_t bytes);

/* Engine solve controls: relative tolerance (default 1e-6) and the cap on
 * PCG iterations per island per tick (default 500). */
PFSF_API void pfsf_set_solve_tolerance(pfsf_engine engine, double tol, int32_t max_iters);

/* Connected components of a registered island: its registered voxel-type
 * buffer is uploaded, labelled as by pfsf_hip_label_components (26-connected)
 * and the summary returned to host memory -- the work of
 * PFSFLabelPropRecorder.recordLabelProp + readbackAfterFence
 * (PFSFLabelPropRecorder.java:75-266) in one call.  header[0] = records,
 * header[1] = overflow; records: max_components x 12 uint32 (host);
 * labels_out: N uint32 (host) or NULL.  1 <= max_components <= 1024. */
PFSF_API pfsf_result pfsf_island_components(pfsf_engine engine, int32_t island_id,
                                            int32_t max_components, uint32_t* header,
                                            uint32_t* records, uint32_t* labels_out);

/* Per-island state after a tick.  A solve that stops short of the tolerance
 * (iteration cap, breakdown) leaves the island dirty, so the next tick
 * continues from its phi; its failure records are still reported. */
typedef struct {
  pfsf_solve_info last_solve; /* the island's last solve (zeros before the first); the
                                 fixed RBGS schedule (PCG off) reports converged = 1 */
  int32_t solver_builds;      /* multigrid hierarchies built for the island so far */
  int32_t operator_mode;      /* 0: registered per-voxel kappa; 1: SoA-6 faces */
  int32_t dirty;              /* 1 while the island waits for a (further) solve */
  int32_t reserved;
} pfsf_island_status;

PFSF_API pfsf_result pfsf_get_island_status(pfsf_engine engine, int32_t island_id,
                                            pfsf_island_status* out);

/* Device damage field of an island (N floats, the phase-field d after the
 * last tick; 0 before any), for inspection. */
PFSF_API pfsf_result pfsf_read_damage(pfsf_engine engine, int32_t island_id, float* out,
                                      int32_t capacity, int32_t* out_count);

/* Device history field of an island (N floats), for inspection. */
PFSF_API pfsf_result pfsf_read_history(pfsf_engine engine, int32_t island_id, float* out,
                                       int32_t capacity, int32_t* out_count);

#ifdef __cplusplus
}
#endif

#endif /* PFSF_HIP_H */
