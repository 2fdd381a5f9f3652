/*
 * TEST INFRASTRUCTURE ONLY: a JNIEnv stand-in for running the JNI shim
 * (pfsf-core_amd/jni/pfsf_jni.c) from Python tests.  Objects are tagged
 * structs: direct buffers (address, capacity), primitive arrays (data,
 * length; Get*Elements hands out the data itself, as a JVM that pins arrays
 * does) and strings.  It also counts element pins and releases, so a test
 * can check that the shim releases what it pins.
 */
#include <stdlib.h>
#include <string.h>

#include "jni.h"

enum { K_DIRECT = 1, K_INT, K_LONG, K_FLOAT, K_BYTE, K_STRING };

struct _jobject {
  int kind;
  void* data;
  int64_t len;  /* elements, or capacity in bytes for direct buffers */
  int owned;
};

static int g_pins = 0;

static jobject make(int kind, void* data, int64_t len, int owned) {
  jobject o = (jobject)calloc(1, sizeof(struct _jobject));
  o->kind = kind;
  o->data = data;
  o->len = len;
  o->owned = owned;
  return o;
}

static jsize arr_len(JNIEnv* e, jarray a) { (void)e; return a ? (jsize)a->len : 0; }
static jint* int_elems(JNIEnv* e, jintArray a, jboolean* c) {
  (void)e; if (c) *c = JNI_FALSE; g_pins++; return (jint*)a->data;
}
static void int_release(JNIEnv* e, jintArray a, jint* p, jint m) { (void)e; (void)a; (void)p; (void)m; g_pins--; }
static void int_set(JNIEnv* e, jintArray a, jsize s, jsize n, const jint* b) {
  (void)e; memcpy((jint*)a->data + s, b, sizeof(jint) * (size_t)n);
}
static jfloat* float_elems(JNIEnv* e, jfloatArray a, jboolean* c) {
  (void)e; if (c) *c = JNI_FALSE; g_pins++; return (jfloat*)a->data;
}
static void float_release(JNIEnv* e, jfloatArray a, jfloat* p, jint m) { (void)e; (void)a; (void)p; (void)m; g_pins--; }
static jlongArray new_long(JNIEnv* e, jsize n) {
  (void)e; return make(K_LONG, calloc((size_t)n + 1, sizeof(jlong)), n, 1);
}
static void long_set(JNIEnv* e, jlongArray a, jsize s, jsize n, const jlong* b) {
  (void)e; memcpy((jlong*)a->data + s, b, sizeof(jlong) * (size_t)n);
}
static jstring new_str(JNIEnv* e, const char* u) {
  (void)e; size_t n = strlen(u); char* d = (char*)malloc(n + 1); memcpy(d, u, n + 1);
  return make(K_STRING, d, (int64_t)n, 1);
}
static const char* str_chars(JNIEnv* e, jstring s, jboolean* c) {
  (void)e; if (c) *c = JNI_FALSE; g_pins++; return (const char*)s->data;
}
static void str_release(JNIEnv* e, jstring s, const char* p) { (void)e; (void)s; (void)p; g_pins--; }
static void* crit(JNIEnv* e, jarray a, jboolean* c) { (void)e; if (c) *c = JNI_FALSE; g_pins++; return a->data; }
static void crit_release(JNIEnv* e, jarray a, void* p, jint m) { (void)e; (void)a; (void)p; (void)m; g_pins--; }
static jobject new_direct(JNIEnv* e, void* a, jlong n) { (void)e; return make(K_DIRECT, a, n, 0); }
static void* direct_addr(JNIEnv* e, jobject b) { (void)e; return b && b->kind == K_DIRECT ? b->data : NULL; }
static jlong direct_cap(JNIEnv* e, jobject b) { (void)e; return b && b->kind == K_DIRECT ? b->len : -1; }

static const struct JNINativeInterface_ g_table = {
    arr_len, int_elems, int_release, int_set, float_elems, float_release, new_long, long_set,
    new_str, str_chars, str_release, crit, crit_release, new_direct, direct_addr, direct_cap};
static JNIEnv g_env = &g_table;

JNIEXPORT JNIEnv* mock_env(void) { return &g_env; }
JNIEXPORT int mock_pins(void) { return g_pins; }
JNIEXPORT jobject mock_direct(void* addr, int64_t bytes) { return make(K_DIRECT, addr, bytes, 0); }
/* kind: 'I' int32, 'J' int64, 'F' float, 'B' byte; the array aliases data */
JNIEXPORT jobject mock_array(int kind, void* data, int64_t len) {
  int k = kind == 'I' ? K_INT : kind == 'J' ? K_LONG : kind == 'F' ? K_FLOAT : K_BYTE;
  return make(k, data, len, 0);
}
JNIEXPORT jobject mock_string(const char* s) { return new_str(&g_env, s); }
JNIEXPORT void* mock_data(jobject o) { return o ? o->data : NULL; }
JNIEXPORT int64_t mock_len(jobject o) { return o ? o->len : -1; }
JNIEXPORT void mock_free(jobject o) {
  if (!o) return;
  if (o->owned) free(o->data);
  free(o);
}
