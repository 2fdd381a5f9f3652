/*
 * TEST INFRASTRUCTURE ONLY: the subset of the JNI C interface that
 * pfsf-core_amd/jni/pfsf_jni.c uses, so the shim compiles and runs in tests
 * against tests/jni/mock_env.c (this image has no JDK).  Names and
 * signatures follow the JNI specification; the function table holds only the
 * members the shim calls, so this header is not a jni.h for any JVM.  The
 * product build uses a JDK's own jni.h.
 */
#ifndef PFSF_TEST_JNI_H
#define PFSF_TEST_JNI_H

#include <stdint.h>

#define JNIEXPORT __attribute__((visibility("default")))
#define JNICALL

typedef int32_t jint;
typedef int64_t jlong;
typedef int8_t jbyte;
typedef uint8_t jboolean;
typedef float jfloat;
typedef jint jsize;

typedef struct _jobject* jobject;
typedef jobject jclass;
typedef jobject jstring;
typedef jobject jarray;
typedef jarray jintArray;
typedef jarray jlongArray;
typedef jarray jfloatArray;
typedef jarray jbyteArray;

#define JNI_FALSE 0
#define JNI_TRUE 1
#define JNI_COMMIT 1
#define JNI_ABORT 2

struct JNINativeInterface_;
typedef const struct JNINativeInterface_* JNIEnv;

struct JNINativeInterface_ {
  jsize (*GetArrayLength)(JNIEnv* env, jarray array);
  jint* (*GetIntArrayElements)(JNIEnv* env, jintArray array, jboolean* is_copy);
  void (*ReleaseIntArrayElements)(JNIEnv* env, jintArray array, jint* elems, jint mode);
  void (*SetIntArrayRegion)(JNIEnv* env, jintArray array, jsize start, jsize len,
                            const jint* buf);
  jfloat* (*GetFloatArrayElements)(JNIEnv* env, jfloatArray array, jboolean* is_copy);
  void (*ReleaseFloatArrayElements)(JNIEnv* env, jfloatArray array, jfloat* elems, jint mode);
  jlongArray (*NewLongArray)(JNIEnv* env, jsize len);
  void (*SetLongArrayRegion)(JNIEnv* env, jlongArray array, jsize start, jsize len,
                             const jlong* buf);
  jstring (*NewStringUTF)(JNIEnv* env, const char* utf);
  const char* (*GetStringUTFChars)(JNIEnv* env, jstring str, jboolean* is_copy);
  void (*ReleaseStringUTFChars)(JNIEnv* env, jstring str, const char* chars);
  void* (*GetPrimitiveArrayCritical)(JNIEnv* env, jarray array, jboolean* is_copy);
  void (*ReleasePrimitiveArrayCritical)(JNIEnv* env, jarray array, void* carray, jint mode);
  jobject (*NewDirectByteBuffer)(JNIEnv* env, void* address, jlong capacity);
  void* (*GetDirectBufferAddress)(JNIEnv* env, jobject buf);
  jlong (*GetDirectBufferCapacity)(JNIEnv* env, jobject buf);
};

#endif
