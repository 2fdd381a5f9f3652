/*
 * pfsf_version.h — ABI version and feature probes
 * (reference L1-native/libpfsf/include/pfsf/pfsf_version.h:24-58).
 */
#ifndef PFSF_VERSION_H
#define PFSF_VERSION_H

#include "pfsf_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* pfsf_version.h:25.  (MAJOR << 16) | (MINOR << 8) | PATCH. */
PFSF_API uint32_t pfsf_abi_version(void);

/* pfsf_version.h:34.  Known names here: "hip.gfx950", "solve.mg_pcg",
 * "solve.batched", "solve.mixed_bf16", "solve.adjoint", "scan.compact",
 * "labels.union_find", "operator.soa6", "smooth.chebyshev",
 * "compute.normalize", "compute.downsample". */
PFSF_API bool pfsf_has_feature(const char* name);

/* pfsf_version.h:40.  Static "<compiler> <flags> <arch>" string. */
PFSF_API const char* pfsf_build_info(void);

/* pfsf_version.h:57.  Contract semver of pfsf_v1.abi.json; this library
 * reports "1.6.0" (1.5.0 plus the additive pfsf_hip.h surface). */
PFSF_API const char* pfsf_abi_contract_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PFSF_VERSION_H */
