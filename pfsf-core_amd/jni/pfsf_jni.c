/*
 * libblockreality_pfsf: the JNI shim the Java side loads
 * (NativePFSFBridge.java:186-600), forwarding each native of the tick path,
 * the version probes and the host primitives to this library's C ABI.  It
 * replaces the reference's NP/src/jni/native_pfsf_bridge.cpp:84-940 for those
 * natives and keeps its marshalling conventions:
 *   - handles are the pfsf_engine pointer as a jlong (0 = none);
 *   - nativeGetStats returns long[5] {islands, voxels, vram used, vram budget,
 *     last tick in microseconds};
 *   - nativeTick writes int[] {count, x, y, z, type, ...}; nativeTickDbb uses
 *     the direct failure buffer's wire format directly;
 *   - direct buffers are passed by address and capacity; a missing or
 *     non-direct buffer is PFSF_ERROR_INVALID_ARG;
 *   - nativeRegisterShader accepts SPIR-V and ignores it (no Vulkan here).
 * Built against a JDK's jni.h (make -C pfsf-core_amd/csrc jni).
 */
#include <jni.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pfsf/pfsf.h"
#include "pfsf/pfsf_compute.h"
#include "pfsf/pfsf_diagnostics.h"

#define BRIDGE(ret, name) \
  JNIEXPORT ret JNICALL Java_com_blockreality_api_physics_pfsf_NativePFSFBridge_##name

static pfsf_engine eng(jlong h) { return (pfsf_engine)(intptr_t)h; }
static jlong handle_of(pfsf_engine e) { return (jlong)(intptr_t)e; }

static pfsf_config make_config(jint max_island, jint budget_ms, jlong vram, jboolean pf,
                               jboolean mg) {
  pfsf_config c;
  memset(&c, 0, sizeof(c));
  c.max_island_size = max_island;
  c.tick_budget_ms = budget_ms;
  c.vram_budget_bytes = vram;
  c.enable_phase_field = pf == JNI_TRUE;
  c.enable_multigrid = mg == JNI_TRUE;
  return c;
}

/* A direct buffer's address and capacity; NULL / 0 when absent or not direct. */
static void* dbb(JNIEnv* env, jobject buf, int64_t* bytes) {
  *bytes = 0;
  if (!buf) return NULL;
  void* a = (*env)->GetDirectBufferAddress(env, buf);
  jlong n = (*env)->GetDirectBufferCapacity(env, buf);
  if (!a || n < 0) return NULL;
  *bytes = (int64_t)n;
  return a;
}

/* ── lifecycle ────────────────────────────────────────────────────────── */

BRIDGE(jlong, nativeCreate)(JNIEnv* env, jclass cls, jint max_island, jint budget_ms,
                            jlong vram, jboolean pf, jboolean mg) {
  (void)env;
  (void)cls;
  pfsf_config c = make_config(max_island, budget_ms, vram, pf, mg);
  return handle_of(pfsf_create(&c));
}

BRIDGE(jlong, nativeCreateWithVulkan)(JNIEnv* env, jclass cls, jint max_island, jint budget_ms,
                                      jlong vram, jboolean pf, jboolean mg, jlong instance,
                                      jlong physical, jlong device, jint family, jlong queue) {
  (void)env;
  (void)cls;
  pfsf_config c = make_config(max_island, budget_ms, vram, pf, mg);
  pfsf_vulkan_handles h;
  memset(&h, 0, sizeof(h));
  h.vk_instance = (uint64_t)instance;
  h.vk_physical_device = (uint64_t)physical;
  h.vk_device = (uint64_t)device;
  h.compute_queue_family = family >= 0 ? (uint32_t)family : 0u;
  h.vk_compute_queue = (uint64_t)queue;
  /* NULL (0): no Vulkan device context here; the Java side takes its
   * fallback and calls nativeCreate */
  return handle_of(pfsf_create_with_vulkan(&c, &h));
}

BRIDGE(jint, nativeInit)(JNIEnv* env, jclass cls, jlong h) {
  (void)env;
  (void)cls;
  return h ? (jint)pfsf_init(eng(h)) : PFSF_ERROR_INVALID_ARG;
}

BRIDGE(void, nativeShutdown)(JNIEnv* env, jclass cls, jlong h) {
  (void)env;
  (void)cls;
  if (h) pfsf_shutdown(eng(h));
}

BRIDGE(void, nativeDestroy)(JNIEnv* env, jclass cls, jlong h) {
  (void)env;
  (void)cls;
  if (h) pfsf_destroy(eng(h));
}

BRIDGE(jboolean, nativeIsAvailable)(JNIEnv* env, jclass cls, jlong h) {
  (void)env;
  (void)cls;
  return (h && pfsf_is_available(eng(h))) ? JNI_TRUE : JNI_FALSE;
}

BRIDGE(jlongArray, nativeGetStats)(JNIEnv* env, jclass cls, jlong h) {
  (void)cls;
  pfsf_stats s;
  if (!h || pfsf_get_stats(eng(h), &s) != PFSF_OK) return NULL;
  jlong v[5] = {(jlong)s.island_count, (jlong)s.total_voxels, (jlong)s.vram_used_bytes,
                (jlong)s.vram_budget_bytes, (jlong)(s.last_tick_ms * 1000.0f)};
  jlongArray out = (*env)->NewLongArray(env, 5);
  if (out) (*env)->SetLongArrayRegion(env, out, 0, 5, v);
  return out;
}

/* ── configuration and islands ────────────────────────────────────────── */

BRIDGE(void, nativeSetWind)(JNIEnv* env, jclass cls, jlong h, jfloat wx, jfloat wy, jfloat wz) {
  (void)env;
  (void)cls;
  pfsf_vec3 w = {wx, wy, wz};
  if (h) pfsf_set_wind(eng(h), &w);
}

BRIDGE(void, nativeSetPCGEnabled)(JNIEnv* env, jclass cls, jlong h, jboolean on) {
  (void)env;
  (void)cls;
  if (h) pfsf_set_pcg_enabled(eng(h), on == JNI_TRUE);
}

BRIDGE(jint, nativeAddIsland)(JNIEnv* env, jclass cls, jlong h, jint id, jint ox, jint oy,
                              jint oz, jint lx, jint ly, jint lz) {
  (void)env;
  (void)cls;
  if (!h) return PFSF_ERROR_INVALID_ARG;
  pfsf_island_desc d;
  memset(&d, 0, sizeof(d));
  d.island_id = id;
  d.origin.x = ox;
  d.origin.y = oy;
  d.origin.z = oz;
  d.lx = lx;
  d.ly = ly;
  d.lz = lz;
  return (jint)pfsf_add_island(eng(h), &d);
}

BRIDGE(void, nativeRemoveIsland)(JNIEnv* env, jclass cls, jlong h, jint id) {
  (void)env;
  (void)cls;
  if (h) pfsf_remove_island(eng(h), id);
}

BRIDGE(void, nativeMarkFullRebuild)(JNIEnv* env, jclass cls, jlong h, jint id) {
  (void)env;
  (void)cls;
  if (h) pfsf_mark_full_rebuild(eng(h), id);
}

BRIDGE(jint, nativeNotifyBlockChange)(JNIEnv* env, jclass cls, jlong h, jint id, jint flat,
                                      jfloat source, jint type, jfloat max_phi, jfloat rcomp,
                                      jfloatArray cond6) {
  (void)cls;
  if (!h) return PFSF_ERROR_INVALID_ARG;
  if (!cond6 || (*env)->GetArrayLength(env, cond6) < 6) return PFSF_ERROR_INVALID_ARG;
  pfsf_voxel_update u;
  memset(&u, 0, sizeof(u));
  u.flat_index = flat;
  u.source = source;
  u.type = (pfsf_voxel_type)type;
  u.max_phi = max_phi;
  u.rcomp = rcomp;
  jfloat* c = (*env)->GetFloatArrayElements(env, cond6, NULL);
  if (!c) return PFSF_ERROR_INVALID_ARG;
  for (int k = 0; k < 6; k++) u.cond[k] = c[k];
  (*env)->ReleaseFloatArrayElements(env, cond6, c, JNI_ABORT);
  return (jint)pfsf_notify_block_change(eng(h), id, &u);
}

/* ── buffers and ticks ────────────────────────────────────────────────── */

BRIDGE(jint, nativeRegisterIslandBuffers)(JNIEnv* env, jclass cls, jlong h, jint id, jobject phi,
                                          jobject source, jobject cond, jobject type,
                                          jobject rcomp, jobject rtens, jobject max_phi) {
  (void)cls;
  if (!h) return PFSF_ERROR_NOT_INIT;
  pfsf_island_buffers b;
  memset(&b, 0, sizeof(b));
  b.phi_addr = dbb(env, phi, &b.phi_bytes);
  b.source_addr = dbb(env, source, &b.source_bytes);
  b.conductivity_addr = dbb(env, cond, &b.conductivity_bytes);
  b.voxel_type_addr = dbb(env, type, &b.voxel_type_bytes);
  b.rcomp_addr = dbb(env, rcomp, &b.rcomp_bytes);
  b.rtens_addr = dbb(env, rtens, &b.rtens_bytes);
  b.max_phi_addr = dbb(env, max_phi, &b.max_phi_bytes);
  return (jint)pfsf_register_island_buffers(eng(h), id, &b);
}

BRIDGE(jint, nativeRegisterIslandLookups)(JNIEnv* env, jclass cls, jlong h, jint id,
                                          jobject material, jobject anchor, jobject fluid,
                                          jobject curing) {
  (void)cls;
  if (!h) return PFSF_ERROR_NOT_INIT;
  pfsf_island_lookups l;
  memset(&l, 0, sizeof(l));
  l.material_id_addr = dbb(env, material, &l.material_id_bytes);
  l.anchor_bitmap_addr = dbb(env, anchor, &l.anchor_bitmap_bytes);
  l.fluid_pressure_addr = dbb(env, fluid, &l.fluid_pressure_bytes);
  l.curing_addr = dbb(env, curing, &l.curing_bytes);
  return (jint)pfsf_register_island_lookups(eng(h), id, &l);
}

BRIDGE(jint, nativeRegisterStressReadback)(JNIEnv* env, jclass cls, jlong h, jint id,
                                           jobject stress) {
  (void)cls;
  if (!h) return PFSF_ERROR_NOT_INIT;
  int64_t bytes;
  void* a = dbb(env, stress, &bytes);
  if (!a) return PFSF_ERROR_INVALID_ARG;
  return (jint)pfsf_register_stress_readback(eng(h), id, a, bytes);
}

/* The dirty-id array pinned for the call (NULL / 0 ids for a null array). */
static jint* pin_ids(JNIEnv* env, jintArray ids, jsize* n) {
  *n = ids ? (*env)->GetArrayLength(env, ids) : 0;
  return *n > 0 ? (*env)->GetIntArrayElements(env, ids, NULL) : NULL;
}

BRIDGE(jint, nativeTickDbb)(JNIEnv* env, jclass cls, jlong h, jintArray ids, jlong epoch,
                            jobject failures) {
  (void)cls;
  if (!h) return PFSF_ERROR_NOT_INIT;
  jsize n;
  jint* p = pin_ids(env, ids, &n);
  if (n > 0 && !p) return PFSF_ERROR_INVALID_ARG;
  int64_t bytes;
  void* f = dbb(env, failures, &bytes);
  pfsf_result r = pfsf_tick_dbb(eng(h), (const int32_t*)p, (int32_t)n, (int64_t)epoch, f, bytes);
  if (p) (*env)->ReleaseIntArrayElements(env, ids, p, JNI_ABORT);
  return (jint)r;
}

BRIDGE(jint, nativeTick)(JNIEnv* env, jclass cls, jlong h, jintArray ids, jlong epoch,
                         jintArray out) {
  (void)cls;
  if (!h) return PFSF_ERROR_INVALID_ARG;
  jsize n;
  jint* p = pin_ids(env, ids, &n);
  if (n > 0 && !p) return PFSF_ERROR_INVALID_ARG;
  jsize cap = out ? (*env)->GetArrayLength(env, out) : 0;
  int32_t room = cap > 1 ? (int32_t)((cap - 1) / 4) : 0;
  pfsf_failure_event* ev = room > 0 ? calloc((size_t)room, sizeof(pfsf_failure_event)) : NULL;
  pfsf_tick_result res = {ev, ev ? room : 0, 0};
  pfsf_result r = pfsf_tick(eng(h), (const int32_t*)p, (int32_t)n, (int64_t)epoch,
                            ev ? &res : NULL);
  if (p) (*env)->ReleaseIntArrayElements(env, ids, p, JNI_ABORT);
  if (out && cap > 0) {
    /* the event record is {x, y, z, type}: the int[] tuples as they lie */
    jint count = res.count;
    (*env)->SetIntArrayRegion(env, out, 0, 1, &count);
    if (ev && count > 0) (*env)->SetIntArrayRegion(env, out, 1, 4 * count, (const jint*)ev);
  }
  free(ev);
  return (jint)r;
}

BRIDGE(jint, nativeReadStress)(JNIEnv* env, jclass cls, jlong h, jint id, jfloatArray out) {
  (void)cls;
  if (!h || !out) return PFSF_ERROR_INVALID_ARG;
  jsize cap = (*env)->GetArrayLength(env, out);
  if (cap <= 0) return PFSF_ERROR_INVALID_ARG;
  jfloat* buf = (*env)->GetFloatArrayElements(env, out, NULL);
  if (!buf) return PFSF_ERROR_INVALID_ARG;
  int32_t got = 0;
  pfsf_result r = pfsf_read_stress(eng(h), id, buf, (int32_t)cap, &got);
  (*env)->ReleaseFloatArrayElements(env, out, buf, 0);
  return r == PFSF_OK ? (jint)got : (jint)r;
}

BRIDGE(jint, nativeDrainCallbacks)(JNIEnv* env, jclass cls, jlong h, jintArray out) {
  (void)cls;
  if (!h || !out) return 0;
  jsize cap = (*env)->GetArrayLength(env, out);
  if (cap <= 0) return 0;
  jint* buf = (*env)->GetIntArrayElements(env, out, NULL);
  if (!buf) return 0;
  int32_t n = pfsf_drain_callbacks(eng(h), (int32_t*)buf, (int32_t)cap);
  (*env)->ReleaseIntArrayElements(env, out, buf, 0);
  return (jint)n;
}

BRIDGE(jobject, nativeGetSparseUploadBuffer)(JNIEnv* env, jclass cls, jlong h, jint id) {
  (void)cls;
  void* a = NULL;
  int64_t bytes = 0;
  if (!h || pfsf_get_sparse_upload_buffer(eng(h), id, &a, &bytes) != PFSF_OK || !a || bytes <= 0)
    return NULL;
  return (*env)->NewDirectByteBuffer(env, a, (jlong)bytes);
}

BRIDGE(jint, nativeNotifySparseUpdates)(JNIEnv* env, jclass cls, jlong h, jint id, jint count) {
  (void)env;
  (void)cls;
  return h ? (jint)pfsf_notify_sparse_updates(eng(h), id, count) : PFSF_ERROR_NOT_INIT;
}

BRIDGE(void, nativeRegisterShader)(JNIEnv* env, jclass cls, jstring name, jobject spirv) {
  (void)env;
  (void)cls;
  (void)name;
  (void)spirv;
}

/* ── version and feature probes ───────────────────────────────────────── */

static jstring str_or(JNIEnv* env, const char* s, const char* dflt) {
  return (*env)->NewStringUTF(env, s ? s : dflt);
}

BRIDGE(jstring, nativeVersion)(JNIEnv* env, jclass cls) {
  (void)cls;
  return str_or(env, pfsf_version(), "unknown");
}

BRIDGE(jint, nativeAbiVersion)(JNIEnv* env, jclass cls) {
  (void)env;
  (void)cls;
  return (jint)pfsf_abi_version();
}

BRIDGE(jboolean, nativeHasFeature)(JNIEnv* env, jclass cls, jstring name) {
  (void)cls;
  if (!name) return JNI_FALSE;
  const char* c = (*env)->GetStringUTFChars(env, name, NULL);
  if (!c) return JNI_FALSE;
  jboolean r = pfsf_has_feature(c) ? JNI_TRUE : JNI_FALSE;
  (*env)->ReleaseStringUTFChars(env, name, c);
  return r;
}

BRIDGE(jstring, nativeBuildInfo)(JNIEnv* env, jclass cls) {
  (void)cls;
  return str_or(env, pfsf_build_info(), "n/a");
}

BRIDGE(jstring, nativeAbiContractVersion)(JNIEnv* env, jclass cls) {
  (void)cls;
  return str_or(env, pfsf_abi_contract_version(), "0.0.0");
}

/* ── host primitives ──────────────────────────────────────────────────── */

BRIDGE(jfloat, nativeNormalizeSoA6)(JNIEnv* env, jclass cls, jfloatArray source,
                                    jfloatArray rcomp, jfloatArray rtens, jfloatArray cond,
                                    jfloatArray hydration, jint n) {
  (void)cls;
  if (!source || !rcomp || !rtens || !cond || n <= 0) return 1.0f;
  jfloatArray arr[5] = {source, rcomp, rtens, cond, hydration};
  float* p[5] = {NULL, NULL, NULL, NULL, NULL};
  for (int k = 0; k < 5; k++)
    if (arr[k]) p[k] = (float*)(*env)->GetPrimitiveArrayCritical(env, arr[k], NULL);
  float smax = 1.0f;
  if (p[0] && p[1] && p[2] && p[3]) pfsf_normalize_soa6(p[0], p[1], p[2], p[3], p[4], n, &smax);
  for (int k = 4; k >= 0; k--)
    if (p[k]) (*env)->ReleasePrimitiveArrayCritical(env, arr[k], p[k], 0);
  return smax;
}

BRIDGE(jint, nativeDownsample2to1)(JNIEnv* env, jclass cls, jfloatArray fine,
                                   jbyteArray fine_type, jint lxf, jint lyf, jint lzf,
                                   jfloatArray coarse, jbyteArray coarse_type) {
  (void)cls;
  if (!fine || !coarse || lxf <= 0 || lyf <= 0 || lzf <= 0) return PFSF_ERROR_INVALID_ARG;
  float* f = (float*)(*env)->GetPrimitiveArrayCritical(env, fine, NULL);
  uint8_t* ft = fine_type ? (uint8_t*)(*env)->GetPrimitiveArrayCritical(env, fine_type, NULL)
                          : NULL;
  float* c = (float*)(*env)->GetPrimitiveArrayCritical(env, coarse, NULL);
  uint8_t* ct = coarse_type
                    ? (uint8_t*)(*env)->GetPrimitiveArrayCritical(env, coarse_type, NULL)
                    : NULL;
  if (f && c) pfsf_downsample_2to1(f, ft, lxf, lyf, lzf, c, ct);
  if (ct) (*env)->ReleasePrimitiveArrayCritical(env, coarse_type, ct, 0);
  if (c) (*env)->ReleasePrimitiveArrayCritical(env, coarse, c, 0);
  if (ft) (*env)->ReleasePrimitiveArrayCritical(env, fine_type, ft, JNI_ABORT);
  if (f) (*env)->ReleasePrimitiveArrayCritical(env, fine, f, JNI_ABORT);
  return PFSF_OK;
}

BRIDGE(jfloat, nativeChebyshevOmega)(JNIEnv* env, jclass cls, jint iter, jfloat rho) {
  (void)env;
  (void)cls;
  return pfsf_chebyshev_omega(iter, rho);
}

BRIDGE(jint, nativePrecomputeOmegaTable)(JNIEnv* env, jclass cls, jfloat rho, jfloatArray out) {
  (void)cls;
  if (!out) return PFSF_ERROR_INVALID_ARG;
  jsize cap = (*env)->GetArrayLength(env, out);
  if (cap <= 0) return PFSF_ERROR_INVALID_ARG;
  float* p = (float*)(*env)->GetPrimitiveArrayCritical(env, out, NULL);
  if (!p) return 0;
  int32_t n = pfsf_precompute_omega_table(rho, p, (int32_t)cap);
  (*env)->ReleasePrimitiveArrayCritical(env, out, p, 0);
  return (jint)n;
}

BRIDGE(jfloat, nativeEstimateSpectralRadius)(JNIEnv* env, jclass cls, jint lmax,
                                             jfloat margin) {
  (void)env;
  (void)cls;
  return pfsf_estimate_spectral_radius(lmax, margin);
}
