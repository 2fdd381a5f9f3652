/*
 * pfsf_compute.h — the reference's stateless host primitives that sit beside
 * the tick path (reference L1-native/libpfsf/include/pfsf/pfsf_compute.h).
 *
 * These two are CPU functions by the reference's contract (host buffers, no
 * device, callable before an engine exists: the Java data builder calls them
 * on its direct buffers).  Their device counterparts on HBM-resident arrays
 * are pfsf_hip_normalize_soa6 and pfsf_hip_downsample_2to1 (pfsf_hip.h).
 */
#ifndef PFSF_COMPUTE_H
#define PFSF_COMPUTE_H

#include "pfsf_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* pfsf_compute.h:37-43 (PFSFDataBuilder.java:222-278 normalizeSoA6):
 * s = max(1, largest of the 6n conductivities); when s > 1 the source,
 * rcomp, rtens (n each) and conductivity (6n) are scaled by 1/s in place
 * (FP32 reciprocal, then one FP32 product per element).  hydration is
 * accepted and not read.  *out_sigma_max (may be NULL) receives s; with a
 * NULL array or n <= 0 nothing is scaled and s = 1. */
PFSF_API void pfsf_normalize_soa6(float* source, float* rcomp, float* rtens, float* conductivity,
                                  const float* hydration, int32_t n, float* out_sigma_max);

/* pfsf_compute.h:129-133 (PFSFDataBuilder.java:430-471 downsample): 2:1
 * coarsening of a per-voxel field on the lxf x lyf x lzf grid to
 * ceil(l/2) per axis: the FP32 mean over the in-grid children (summed in
 * x-fastest child order); the coarse type is anchor when anchors are at least
 * as many as solids and as airs, else solid when solids are at least as many
 * as airs, else air (children counted from fine_type; with fine_type NULL
 * every count is 0).  coarse_type may be NULL. */
PFSF_API void pfsf_downsample_2to1(const float* fine, const uint8_t* fine_type, int32_t lxf,
                                   int32_t lyf, int32_t lzf, float* coarse,
                                   uint8_t* coarse_type);

#ifdef __cplusplus
}
#endif

#endif /* PFSF_COMPUTE_H */
