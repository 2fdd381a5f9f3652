/*
 * pfsf_diagnostics.h — the reference's Chebyshev scheduling primitives
 * (reference L1-native/libpfsf/include/pfsf/pfsf_diagnostics.h:24-62).
 * Host functions by the reference's contract; the device smoother that uses
 * the schedule is pfsf_hip_chebyshev_jacobi (pfsf_hip.h).
 */
#ifndef PFSF_DIAGNOSTICS_H
#define PFSF_DIAGNOSTICS_H

#include "pfsf_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Scheduler constants (pfsf_diagnostics.h:24-26, PFSFConstants.java). */
#define PFSF_OMEGA_TABLE_SIZE 64
#define PFSF_OMEGA_DENOM_EPSILON 0.01f
#define PFSF_MAX_OMEGA 1.98f

/* pfsf_diagnostics.h:43 (PFSFScheduler.computeOmega): the Chebyshev
 * semi-iteration weight of step `iter` for spectral radius rho (Wang 2015,
 * Eq. 12): 1 for iter <= 0 or rho^2 >= 1, 2 / (2 - rho^2) for iter 1, then
 * w_k = 4 / (4 - rho^2 w_{k-1}); the recurrence stops at the last good value
 * when the denominator falls below PFSF_OMEGA_DENOM_EPSILON or a step leaves
 * (0, 2]; results past iter 1 are capped at PFSF_MAX_OMEGA.  FP32
 * throughout. */
PFSF_API float pfsf_chebyshev_omega(int32_t iter, float rho_spec);

/* pfsf_diagnostics.h:51 (PFSFScheduler.precomputeOmegaTable): out[0] = 1,
 * out[1] = 2 / (2 - rho^2) (1 when rho^2 >= 1), then each entry from the
 * previous one: the previous value again when 4 - rho^2 w < epsilon, else
 * 4 / (4 - rho^2 w) capped at PFSF_MAX_OMEGA (1 if NaN).  Returns capacity,
 * or PFSF_ERROR_INVALID_ARG for a NULL table or capacity <= 0. */
PFSF_API int32_t pfsf_precompute_omega_table(float rho_spec, float* out, int32_t capacity);

/* pfsf_diagnostics.h:60 (PFSFScheduler.estimateSpectralRadius):
 * cos(pi / l_max) * safety_margin (FP64, rounded to FP32); 0.5 for
 * l_max <= 1. */
PFSF_API float pfsf_estimate_spectral_radius(int32_t l_max, float safety_margin);

#ifdef __cplusplus
}
#endif

#endif /* PFSF_DIAGNOSTICS_H */
