/*
 * pfsf.h — libpfsf public C ABI, HIP (gfx950) implementation.
 *
 * Drop-in for the reference surface L1-native/libpfsf/include/pfsf/pfsf.h:43-346
 * (the tick path the JNI bridge NP/src/jni/native_pfsf_bridge.cpp:84-640
 * binds).  Every declaration keeps the reference's name, argument meaning and
 * error convention; the line cited on each is the declaration it replaces.
 * The reference umbrella also pulls in pfsf_compute/diagnostics/extension/
 * plan/trace (CPU helper and diagnostics modules); those are not part of the
 * structural-solve path and are not provided here.
 *
 * Threading: all calls from one thread, except pfsf_get_stats (any thread).
 * Lifecycle: pfsf_create -> pfsf_init -> [pfsf_tick*] -> pfsf_shutdown ->
 * pfsf_destroy.
 */
#ifndef PFSF_H
#define PFSF_H

#include "pfsf_types.h"
#include "pfsf_version.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ── Lifecycle ─────────────────────────────────────────────────────────── */

/* pfsf.h:43.  NULL config = defaults {50000 voxels, 8 ms, 512 MiB, phase
 * field on, multigrid on} (NP/src/pfsf_api.cpp:23-31).  No device work. */
PFSF_API pfsf_engine pfsf_create(const pfsf_config* config);

/* pfsf.h:56-62 / :69.  Vulkan handle adoption has no HIP meaning: this always
 * returns NULL, which the Java runtime already treats as "fall back to
 * pfsf_create" (NativePFSFRuntime.java:157-170). */
typedef struct pfsf_vulkan_handles {
  uint64_t vk_instance;
  uint64_t vk_physical_device;
  uint64_t vk_device;
  uint32_t compute_queue_family;
  uint64_t vk_compute_queue;
} pfsf_vulkan_handles;
PFSF_API pfsf_engine pfsf_create_with_vulkan(const pfsf_config* config,
                                             const pfsf_vulkan_handles* handles);

/* pfsf.h:78.  Selects the current HIP device and creates the engine stream.
 * PFSF_ERROR_NO_DEVICE when no GPU is visible. */
PFSF_API pfsf_result pfsf_init(pfsf_engine engine);

/* pfsf.h:85 / :93 / :98 / :103 */
PFSF_API void pfsf_shutdown(pfsf_engine engine);
PFSF_API void pfsf_destroy(pfsf_engine engine);
PFSF_API bool pfsf_is_available(pfsf_engine engine);
PFSF_API pfsf_result pfsf_get_stats(pfsf_engine engine, pfsf_stats* out);

/* ── Configuration (pfsf.h:116-144) ────────────────────────────────────── */

PFSF_API void pfsf_set_material_lookup(pfsf_engine engine, pfsf_material_fn fn, void* user_data);
PFSF_API void pfsf_set_anchor_lookup(pfsf_engine engine, pfsf_anchor_fn fn, void* user_data);
PFSF_API void pfsf_set_fill_ratio_lookup(pfsf_engine engine, pfsf_fill_ratio_fn fn,
                                         void* user_data);
PFSF_API void pfsf_set_curing_lookup(pfsf_engine engine, pfsf_curing_fn fn, void* user_data);
PFSF_API void pfsf_set_wind(pfsf_engine engine, const pfsf_vec3* wind);
/* pfsf.h:144.  Enabled (default): MG-preconditioned PCG to tolerance.
 * Disabled: a fixed budget of 8-colour RBGS sweeps per tick. */
PFSF_API void pfsf_set_pcg_enabled(pfsf_engine engine, int enabled);

/* ── Islands (pfsf.h:156-182) ─────────────────────────────────────────── */

/* pfsf.h:156.  ISLAND_FULL when a dimension is <= 0 or lx*ly*lz exceeds
 * max_island_size (NP/src/pfsf_engine.cpp:156-173). */
PFSF_API pfsf_result pfsf_add_island(pfsf_engine engine, const pfsf_island_desc* desc);
PFSF_API void pfsf_remove_island(pfsf_engine engine, int32_t island_id);
PFSF_API pfsf_result pfsf_notify_block_change(pfsf_engine engine, int32_t island_id,
                                              const pfsf_voxel_update* update);
PFSF_API void pfsf_mark_full_rebuild(pfsf_engine engine, int32_t island_id);

/* ── Tick ─────────────────────────────────────────────────────────────── */

/* pfsf.h:198.  Solves every dirty island, updates the history field and runs
 * the failure scan; failures are appended to result->events (up to capacity). */
PFSF_API pfsf_result pfsf_tick(pfsf_engine engine, const int32_t* dirty_island_ids,
                               int32_t dirty_count, int64_t current_epoch,
                               pfsf_tick_result* result);

/* pfsf.h:217.  Raw phi of the island (not a ratio), N floats. */
PFSF_API pfsf_result pfsf_read_stress(pfsf_engine engine, int32_t island_id, float* out_stress,
                                      int32_t capacity, int32_t* out_count);

/* ── Zero-copy host buffer registration (pfsf.h:232-287) ─────────────── */

typedef struct pfsf_island_buffers {
  void* phi_addr; /* float32 x N */
  int64_t phi_bytes;
  void* source_addr; /* float32 x N */
  int64_t source_bytes;
  void* conductivity_addr; /* float32 x 6N, SoA: dir*N + i */
  int64_t conductivity_bytes;
  void* voxel_type_addr; /* uint8 x N */
  int64_t voxel_type_bytes;
  void* rcomp_addr; /* float32 x N */
  int64_t rcomp_bytes;
  void* rtens_addr; /* float32 x N */
  int64_t rtens_bytes;
  void* max_phi_addr; /* float32 x N */
  int64_t max_phi_bytes;
} pfsf_island_buffers;

typedef struct pfsf_island_lookups {
  void* material_id_addr; /* int32 x N */
  int64_t material_id_bytes;
  void* anchor_bitmap_addr; /* int64 x N */
  int64_t anchor_bitmap_bytes;
  void* fluid_pressure_addr; /* float32 x N */
  int64_t fluid_pressure_bytes;
  void* curing_addr; /* float32 x N */
  int64_t curing_bytes;
} pfsf_island_lookups;

/* pfsf.h:264.  INVALID_ARG on NULL pointers or any buffer smaller than its
 * N-scaled size (NP/src/pfsf_engine.cpp:205-259).  Uploads immediately and
 * marks the island dirty. */
PFSF_API pfsf_result pfsf_register_island_buffers(pfsf_engine engine, int32_t island_id,
                                                  const pfsf_island_buffers* bufs);
PFSF_API pfsf_result pfsf_register_island_lookups(pfsf_engine engine, int32_t island_id,
                                                  const pfsf_island_lookups* lookups);
/* pfsf.h:272.  phi is copied here after each solved tick (Java registers the
 * phi buffer itself, which makes it the warm start of the next tick). */
PFSF_API pfsf_result pfsf_register_stress_readback(pfsf_engine engine, int32_t island_id,
                                                   void* addr, int64_t bytes);

/* pfsf.h:282.  failure_addr wire format: int32 count, then count x
 * {int32 x, y, z, type} in world coordinates; the header is zeroed first. */
PFSF_API pfsf_result pfsf_tick_dbb(pfsf_engine engine, const int32_t* dirty_island_ids,
                                   int32_t dirty_count, int64_t current_epoch,
                                   void* failure_addr, int64_t failure_bytes);

/* pfsf.h:294.  No native->Java events are produced; returns 0. */
PFSF_API int32_t pfsf_drain_callbacks(pfsf_engine engine, int32_t* out_events, int32_t capacity);

/* ── Sparse re-upload (pfsf.h:319-339) ──────────────────────────────────
 * Records are 48 bytes: {int32 flat_index; float source; uint32 type;
 * float max_phi; float rcomp; float rtens; float cond[6]}, at most 512 per
 * call (SH/sparse_scatter.comp.glsl:11-19,48-67). */
PFSF_API pfsf_result pfsf_get_sparse_upload_buffer(pfsf_engine engine, int32_t island_id,
                                                   void** out_addr, int64_t* out_bytes);
PFSF_API pfsf_result pfsf_notify_sparse_updates(pfsf_engine engine, int32_t island_id,
                                                int32_t update_count);

/* pfsf.h:346 */
PFSF_API const char* pfsf_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PFSF_H */
