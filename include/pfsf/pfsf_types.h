/*
 * pfsf_types.h — POD types and enums of the libpfsf C ABI.
 *
 * Layout- and value-compatible with the reference header
 * L1-native/libpfsf/include/pfsf/pfsf_types.h:19-148, so a caller built
 * against the reference (the JNI shim, pfsf_cli) links against this library
 * unchanged.  Every enum value and struct field order below is part of that
 * frozen contract (L1-native/libpfsf/abi/pfsf_v1.abi.json).
 */
#ifndef PFSF_TYPES_H
#define PFSF_TYPES_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFSF_API __attribute__((visibility("default")))

/* Opaque engine handle (reference pfsf_types.h:30). */
typedef struct pfsf_engine_t* pfsf_engine;

/* Result codes (reference pfsf_types.h:33-41).  PFSF_ERROR_VULKAN keeps its
 * value and now reports any device (HIP runtime) failure. */
typedef enum {
  PFSF_OK = 0,
  PFSF_ERROR_VULKAN = -1,
  PFSF_ERROR_NO_DEVICE = -2,
  PFSF_ERROR_OUT_OF_VRAM = -3,
  PFSF_ERROR_INVALID_ARG = -4,
  PFSF_ERROR_NOT_INIT = -5,
  PFSF_ERROR_ISLAND_FULL = -6,
} pfsf_result;

/* Voxel type markers (reference pfsf_types.h:44-48). */
typedef enum {
  PFSF_VOXEL_AIR = 0,
  PFSF_VOXEL_SOLID = 1,
  PFSF_VOXEL_ANCHOR = 2,
} pfsf_voxel_type;

/* Failure codes (reference pfsf_types.h:51-57). */
typedef enum {
  PFSF_FAIL_OK = 0,
  PFSF_FAIL_CANTILEVER = 1,
  PFSF_FAIL_CRUSHING = 2,
  PFSF_FAIL_NO_SUPPORT = 3,
  PFSF_FAIL_TENSION = 4,
} pfsf_failure_type;

/* Face directions of the SoA-6 conductivity layout (reference :60-67). */
typedef enum {
  PFSF_DIR_NEG_X = 0,
  PFSF_DIR_POS_X = 1,
  PFSF_DIR_NEG_Y = 2,
  PFSF_DIR_POS_Y = 3,
  PFSF_DIR_NEG_Z = 4,
  PFSF_DIR_POS_Z = 5,
} pfsf_direction;

typedef struct {
  int32_t x, y, z;
} pfsf_pos;

typedef struct {
  float x, y, z;
} pfsf_vec3;

/* Island grid descriptor (reference :80-84). */
typedef struct {
  int32_t island_id;
  pfsf_pos origin; /* AABB minimum corner, world coordinates */
  int32_t lx, ly, lz;
} pfsf_island_desc;

/* Per-voxel material record (reference :87-95). */
typedef struct {
  float density;
  float rcomp;
  float rtens;
  float youngs_gpa;
  float poisson;
  float gc;
  bool is_anchor;
} pfsf_material;

/* Sparse voxel update (reference :98-105). */
typedef struct {
  int32_t flat_index;
  float source;
  pfsf_voxel_type type;
  float max_phi;
  float rcomp;
  float cond[6];
} pfsf_voxel_update;

/* Failure event, 16 bytes {x,y,z,type} (reference :108-111). */
typedef struct {
  pfsf_pos pos;
  pfsf_failure_type type;
} pfsf_failure_event;

/* Engine configuration (reference :114-120). */
typedef struct {
  int32_t max_island_size;
  int32_t tick_budget_ms;
  int64_t vram_budget_bytes;
  bool enable_phase_field;
  bool enable_multigrid;
} pfsf_config;

/* Engine statistics (reference :123-129). */
typedef struct {
  int32_t island_count;
  int64_t total_voxels;
  int64_t vram_used_bytes;
  int64_t vram_budget_bytes;
  float last_tick_ms;
} pfsf_stats;

/* Tick result (reference :132-136). */
typedef struct {
  pfsf_failure_event* events;
  int32_t capacity;
  int32_t count;
} pfsf_tick_result;

/* Host callbacks (reference :139-148). */
typedef pfsf_material (*pfsf_material_fn)(pfsf_pos pos, void* user_data);
typedef bool (*pfsf_anchor_fn)(pfsf_pos pos, void* user_data);
typedef float (*pfsf_fill_ratio_fn)(pfsf_pos pos, void* user_data);
typedef float (*pfsf_curing_fn)(pfsf_pos pos, void* user_data);

#ifdef __cplusplus
}
#endif

#endif /* PFSF_TYPES_H */
