/* JNI shim between the Scala facade (Csmi.scala, `object Csmi`) and libcsmi.so.
 *
 * It replaces LWJGL's OpenCL bindings under `trait OpenCL` (OpenCL.scala:1139-1433): every
 * function forwards one `cs_*` entry point of include/csmi.h. Conventions:
 *   - native handles (cs_device*, cs_stream*, ...) and device pointers cross as jlong;
 *   - out-parameters are one-element jlongArray / jintArray / jfloatArray / jdoubleArray;
 *   - host memory crosses as direct java.nio.ByteBuffer (what LWJGL's MemoryStack and
 *     memAlloc give the reference), never as a copied Java array;
 *   - the return value is the cs_status, equal to OpenCL's status codes, so the Scala side
 *     keeps `OpenCL.checkErrorCode` (OpenCL.scala:143-312) to raise the same exceptions;
 *   - a serialised expression DAG is a direct ByteBuffer of cs_node records (632 bytes each,
 *     little endian, the layout of include/csmi.h) written by HipTensors.scala.
 *
 * Built against a JDK's include/ and include/linux/ directories and linked with -lcsmi:
 *   cc -shared -fPIC -I$JAVA_HOME/include -I$JAVA_HOME/include/linux -I../../include \
 *      csmi_jni.c -L../lib -lcsmi -Wl,-rpath,'$ORIGIN' -o libcsmi_jni.so
 * (there is no JDK in the build image, so this file is not compiled there). */
#include <jni.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "csmi.h"

#define H(type, v) ((type *)(intptr_t)(v))
#define P(v) ((void *)(intptr_t)(v))
#define J(name) JNICALL Java_com_thoughtworks_compute_Csmi_00024_##name

static void put_long(JNIEnv *env, jlongArray out, const void *handle) {
  jlong h = (jlong)(intptr_t)handle;
  (*env)->SetLongArrayRegion(env, out, 0, 1, &h);
}

static void *direct(JNIEnv *env, jobject buffer) {
  return buffer ? (*env)->GetDirectBufferAddress(env, buffer) : NULL;
}

/* ------------------------------------------------------------------ status and strings */
JNIEXPORT jstring J(statusString)(JNIEnv *env, jobject self, jint status) {
  (void)self;
  return (*env)->NewStringUTF(env, cs_status_string(status));
}

JNIEXPORT jstring J(lastError)(JNIEnv *env, jobject self) {
  (void)self;
  char buf[4096];
  cs_last_error(buf, sizeof buf);
  return (*env)->NewStringUTF(env, buf);
}

/* ------------------------------------------------------------------ devices and streams */
JNIEXPORT jint J(deviceCount)(JNIEnv *env, jobject self, jintArray out) {
  (void)self;
  int n = 0;
  cs_status s = cs_device_count(&n);
  jint v = n;
  (*env)->SetIntArrayRegion(env, out, 0, 1, &v);
  return s;
}

JNIEXPORT jint J(deviceOpen)(JNIEnv *env, jobject self, jint ordinal, jlongArray out) {
  (void)self;
  cs_device *d = NULL;
  cs_status s = cs_device_open(ordinal, &d);
  put_long(env, out, d);
  return s;
}

JNIEXPORT jint J(deviceClose)(JNIEnv *env, jobject self, jlong device) {
  (void)env, (void)self;
  return cs_device_close(H(cs_device, device));
}

JNIEXPORT jint J(deviceInfo)(JNIEnv *env, jobject self, jlong device, jint key, jlongArray out) {
  (void)self;
  int64_t v = 0;
  cs_status s = cs_device_info(H(cs_device, device), key, &v);
  jlong j = v;
  (*env)->SetLongArrayRegion(env, out, 0, 1, &j);
  return s;
}

JNIEXPORT jint J(streamCreate)(JNIEnv *env, jobject self, jlong device, jlongArray out) {
  (void)self;
  cs_stream *st = NULL;
  cs_status s = cs_stream_create(H(cs_device, device), &st);
  put_long(env, out, st);
  return s;
}

JNIEXPORT jint J(streamDestroy)(JNIEnv *env, jobject self, jlong stream) {
  (void)env, (void)self;
  return cs_stream_destroy(H(cs_stream, stream));
}

JNIEXPORT jint J(streamSynchronize)(JNIEnv *env, jobject self, jlong stream) {
  (void)env, (void)self;
  return cs_stream_synchronize(H(cs_stream, stream));
}

/* ------------------------------------------------------------------ events */
JNIEXPORT jint J(eventRecord)(JNIEnv *env, jobject self, jlong stream, jlongArray out) {
  (void)self;
  cs_event *e = NULL;
  cs_status s = cs_event_record(H(cs_stream, stream), &e);
  put_long(env, out, e);
  return s;
}

JNIEXPORT jint J(eventRelease)(JNIEnv *env, jobject self, jlong event) {
  (void)env, (void)self;
  return cs_event_release(H(cs_event, event));
}

JNIEXPORT jint J(eventWait)(JNIEnv *env, jobject self, jlong event) {
  (void)env, (void)self;
  return cs_event_wait(H(cs_event, event));
}

JNIEXPORT jint J(eventElapsedMs)(JNIEnv *env, jobject self, jlong start, jlong end, jfloatArray out) {
  (void)self;
  float ms = 0.0f;
  cs_status s = cs_event_elapsed_ms(H(cs_event, start), H(cs_event, end), &ms);
  jfloat v = ms;
  (*env)->SetFloatArrayRegion(env, out, 0, 1, &v);
  return s;
}

/* Completion callbacks: cs_event_on_complete runs `complete` on the library's completion
 * thread, which attaches itself to the JVM (once; it is a daemon thread of the library) and
 * runs the continuation -- the role of Event.eventCallback and clSetEventCallback
 * (OpenCL.scala:1246-1263). The Runnable must not call back into HIP on that thread: it
 * re-posts its work to an executor, as OpenCL.scala:393-412 does. */
static JavaVM *g_vm;

struct continuation {
  jobject runnable; /* global reference */
};

static void complete(void *user) {
  struct continuation *c = (struct continuation *)user;
  JNIEnv *env = NULL;
  if ((*g_vm)->GetEnv(g_vm, (void **)&env, JNI_VERSION_1_8) != JNI_OK)
    (*g_vm)->AttachCurrentThreadAsDaemon(g_vm, (void **)&env, NULL);
  jclass cls = (*env)->GetObjectClass(env, c->runnable);
  jmethodID run = (*env)->GetMethodID(env, cls, "run", "()V");
  (*env)->CallVoidMethod(env, c->runnable, run);
  if ((*env)->ExceptionCheck(env)) (*env)->ExceptionClear(env);
  (*env)->DeleteGlobalRef(env, c->runnable);
  (*env)->DeleteLocalRef(env, cls);
  free(c);
}

JNIEXPORT jint J(eventOnComplete)(JNIEnv *env, jobject self, jlong event, jobject runnable) {
  (void)self;
  if (!g_vm) (*env)->GetJavaVM(env, &g_vm);
  struct continuation *c = (struct continuation *)malloc(sizeof *c);
  if (!c) return CS_OUT_OF_HOST_MEMORY;
  c->runnable = (*env)->NewGlobalRef(env, runnable);
  cs_status s = cs_event_on_complete(H(cs_event, event), complete, c);
  if (s != CS_SUCCESS) {
    (*env)->DeleteGlobalRef(env, c->runnable);
    free(c);
  }
  return s;
}

/* ------------------------------------------------------------------ buffers */
JNIEXPORT jint J(bufferAllocOn)(JNIEnv *env, jobject self, jlong stream, jlong bytes, jlongArray out) {
  (void)self;
  cs_buffer *b = NULL;
  cs_status s = cs_buffer_alloc_on(H(cs_stream, stream), (size_t)bytes, &b);
  put_long(env, out, b);
  return s;
}

JNIEXPORT jint J(bufferRetain)(JNIEnv *env, jobject self, jlong buffer) {
  (void)env, (void)self;
  return cs_buffer_retain(H(cs_buffer, buffer));
}

JNIEXPORT jint J(bufferRelease)(JNIEnv *env, jobject self, jlong buffer) {
  (void)env, (void)self;
  return cs_buffer_release(H(cs_buffer, buffer));
}

JNIEXPORT jint J(bufferDevicePtr)(JNIEnv *env, jobject self, jlong buffer, jlongArray out) {
  (void)self;
  void *p = NULL;
  cs_status s = cs_buffer_device_ptr(H(cs_buffer, buffer), &p);
  put_long(env, out, p);
  return s;
}

JNIEXPORT jint J(bufferWrite)(JNIEnv *env, jobject self, jlong stream, jlong buffer, jlong offset, jobject host,
                              jlong bytes, jboolean blocking) {
  (void)self;
  return cs_buffer_write(H(cs_stream, stream), H(cs_buffer, buffer), (size_t)offset, direct(env, host),
                         (size_t)bytes, blocking ? 1 : 0);
}

JNIEXPORT jint J(bufferRead)(JNIEnv *env, jobject self, jlong stream, jlong buffer, jlong offset, jobject host,
                             jlong bytes) {
  (void)self;
  return cs_buffer_read(H(cs_stream, stream), H(cs_buffer, buffer), (size_t)offset, direct(env, host),
                        (size_t)bytes);
}

/* ------------------------------------------------------------------ fused kernels */
#define MAX_OUTPUTS 1024
#define MAX_LEAVES 4096

static int read_ints(JNIEnv *env, jintArray a, int32_t *dst, jsize cap, jsize *n) {
  *n = a ? (*env)->GetArrayLength(env, a) : 0;
  if (*n > cap) return 0;
  if (*n) (*env)->GetIntArrayRegion(env, a, 0, *n, (jint *)dst);
  return 1;
}

/* options: {math_mode, vector_width, items_per_thread, block_size, nontemporal_store
 * [, launch_order]}; a missing launch_order is CS_ORDER_AUTO */
static int read_options(JNIEnv *env, jintArray options, cs_fused_options *o) {
  int32_t v[6] = {0, 0, 0, 0, 0, CS_ORDER_AUTO};
  jsize n;
  if (!read_ints(env, options, v, 6, &n) || (n != 5 && n != 6)) return 0;
  o->math_mode = v[0], o->vector_width = v[1], o->items_per_thread = v[2], o->block_size = v[3];
  o->nontemporal_store = v[4];
  o->launch_order = v[5];
  return 1;
}

typedef cs_status (*compile_fn)(cs_device *, const cs_node *, int, const int32_t *, int, int, const int32_t *,
                                const cs_fused_options *, cs_fused **);

static jint compile_with(JNIEnv *env, compile_fn fn, jlong device, jobject nodes, jint nnodes, jintArray outputs,
                         jintArray shape, jintArray options, jlongArray out) {
  int32_t outs[MAX_OUTPUTS], shp[CS_MAX_RANK];
  jsize no, rank;
  cs_fused_options o;
  if (!read_ints(env, outputs, outs, MAX_OUTPUTS, &no) || !read_ints(env, shape, shp, CS_MAX_RANK, &rank) ||
      !read_options(env, options, &o))
    return CS_INVALID_VALUE;
  cs_fused *f = NULL;
  cs_status s = fn(H(cs_device, device), (const cs_node *)direct(env, nodes), nnodes, outs, no, rank, shp, &o, &f);
  put_long(env, out, f);
  return s;
}

JNIEXPORT jint J(fusedCompile)(JNIEnv *env, jobject self, jlong device, jobject nodes, jint nnodes,
                               jintArray outputs, jintArray shape, jintArray options, jlongArray out) {
  (void)self;
  return compile_with(env, cs_fused_compile, device, nodes, nnodes, outputs, shape, options, out);
}

JNIEXPORT jint J(fusedCompileSum)(JNIEnv *env, jobject self, jlong device, jobject nodes, jint nnodes,
                                  jintArray outputs, jintArray shape, jintArray options, jlongArray out) {
  (void)self;
  return compile_with(env, cs_fused_compile_sum, device, nodes, nnodes, outputs, shape, options, out);
}

JNIEXPORT jint J(fusedLookup)(JNIEnv *env, jobject self, jlong device, jobject nodes, jint nnodes, jintArray outputs,
                              jintArray shape, jintArray options, jintArray hit) {
  (void)self;
  int32_t outs[MAX_OUTPUTS], shp[CS_MAX_RANK];
  jsize no, rank;
  cs_fused_options o;
  if (!read_ints(env, outputs, outs, MAX_OUTPUTS, &no) || !read_ints(env, shape, shp, CS_MAX_RANK, &rank) ||
      !read_options(env, options, &o))
    return CS_INVALID_VALUE;
  int h = 0;
  cs_status s = cs_fused_lookup(H(cs_device, device), (const cs_node *)direct(env, nodes), nnodes, outs, no, rank,
                                shp, &o, &h);
  jint v = h;
  (*env)->SetIntArrayRegion(env, hit, 0, 1, &v);
  return s;
}

JNIEXPORT jint J(fusedRelease)(JNIEnv *env, jobject self, jlong fused) {
  (void)env, (void)self;
  return cs_fused_release(H(cs_fused, fused));
}

JNIEXPORT jint J(fusedCacheClear)(JNIEnv *env, jobject self, jlong device) {
  (void)env, (void)self;
  return cs_fused_cache_clear(H(cs_device, device));
}

static int read_leaves(JNIEnv *env, jlongArray leaves, const void **dst, jsize *n) {
  jlong tmp[MAX_LEAVES];
  *n = leaves ? (*env)->GetArrayLength(env, leaves) : 0;
  if (*n > MAX_LEAVES) return 0;
  if (*n) (*env)->GetLongArrayRegion(env, leaves, 0, *n, tmp);
  for (jsize i = 0; i < *n; ++i) dst[i] = P(tmp[i]);
  return 1;
}

JNIEXPORT jint J(fusedLaunch)(JNIEnv *env, jobject self, jlong stream, jlong fused, jlongArray leaves, jlong out) {
  (void)self;
  const void *ptrs[MAX_LEAVES];
  jsize n;
  if (!read_leaves(env, leaves, ptrs, &n)) return CS_INVALID_KERNEL_ARGS;
  return cs_fused_launch(H(cs_stream, stream), H(cs_fused, fused), ptrs, n, P(out));
}

JNIEXPORT jint J(fusedSum)(JNIEnv *env, jobject self, jlong stream, jlong fused, jlongArray leaves, jlong out,
                           jlong out64) {
  (void)self;
  const void *ptrs[MAX_LEAVES];
  jsize n;
  if (!read_leaves(env, leaves, ptrs, &n)) return CS_INVALID_KERNEL_ARGS;
  return cs_fused_sum(H(cs_stream, stream), H(cs_fused, fused), ptrs, n, (float *)P(out), (double *)P(out64));
}

/* ------------------------------------------------------------------ batches and graphs */
JNIEXPORT jint J(batchCreate)(JNIEnv *env, jobject self, jlong device, jlongArray exprs, jint block,
                              jlongArray out) {
  (void)self;
  jsize n = (*env)->GetArrayLength(env, exprs);
  jlong *h = (jlong *)malloc(sizeof(jlong) * (size_t)(n ? n : 1));
  cs_fused **f = (cs_fused **)malloc(sizeof(cs_fused *) * (size_t)(n ? n : 1));
  if (!h || !f) {
    free(h), free(f);
    return CS_OUT_OF_HOST_MEMORY;
  }
  (*env)->GetLongArrayRegion(env, exprs, 0, n, h);
  for (jsize i = 0; i < n; ++i) f[i] = H(cs_fused, h[i]);
  cs_batch *b = NULL;
  cs_status s = cs_batch_create(H(cs_device, device), f, n, block, &b);
  put_long(env, out, b);
  free(h), free(f);
  return s;
}

JNIEXPORT jint J(batchBind)(JNIEnv *env, jobject self, jlong stream, jlong batch, jlongArray args) {
  (void)self;
  jsize n = (*env)->GetArrayLength(env, args);
  jlong *h = (jlong *)malloc(sizeof(jlong) * (size_t)(n ? n : 1));
  const void **p = (const void **)malloc(sizeof(void *) * (size_t)(n ? n : 1));
  if (!h || !p) {
    free(h), free((void *)p);
    return CS_OUT_OF_HOST_MEMORY;
  }
  (*env)->GetLongArrayRegion(env, args, 0, n, h);
  for (jsize i = 0; i < n; ++i) p[i] = P(h[i]);
  cs_status s = cs_batch_bind(H(cs_stream, stream), H(cs_batch, batch), p, n);
  free(h), free((void *)p);
  return s;
}

JNIEXPORT jint J(batchLaunch)(JNIEnv *env, jobject self, jlong stream, jlong batch) {
  (void)env, (void)self;
  return cs_batch_launch(H(cs_stream, stream), H(cs_batch, batch));
}

JNIEXPORT jint J(batchRelease)(JNIEnv *env, jobject self, jlong batch) {
  (void)env, (void)self;
  return cs_batch_release(H(cs_batch, batch));
}

JNIEXPORT jint J(graphBeginCapture)(JNIEnv *env, jobject self, jlong stream) {
  (void)env, (void)self;
  return cs_graph_begin_capture(H(cs_stream, stream));
}

JNIEXPORT jint J(graphEndCapture)(JNIEnv *env, jobject self, jlong stream, jlongArray out) {
  (void)self;
  cs_graph *g = NULL;
  cs_status s = cs_graph_end_capture(H(cs_stream, stream), &g);
  put_long(env, out, g);
  return s;
}

JNIEXPORT jint J(graphLaunch)(JNIEnv *env, jobject self, jlong stream, jlong graph) {
  (void)env, (void)self;
  return cs_graph_launch(H(cs_stream, stream), H(cs_graph, graph));
}

JNIEXPORT jint J(graphRelease)(JNIEnv *env, jobject self, jlong graph) {
  (void)env, (void)self;
  return cs_graph_release(H(cs_graph, graph));
}

/* ------------------------------------------------------------------ fixed programs */
JNIEXPORT jint J(reduceSum)(JNIEnv *env, jobject self, jlong stream, jlong in, jlong n, jlong out, jlong out64) {
  (void)env, (void)self;
  return cs_reduce_sum(H(cs_stream, stream), (const float *)P(in), n, (float *)P(out), (double *)P(out64));
}

JNIEXPORT jint J(reduceMax)(JNIEnv *env, jobject self, jlong stream, jlong in, jlong n, jlong out) {
  (void)env, (void)self;
  return cs_reduce_max(H(cs_stream, stream), (const float *)P(in), n, (float *)P(out));
}

JNIEXPORT jint J(reduceRows)(JNIEnv *env, jobject self, jlong stream, jlong in, jlong rows, jlong cols, jlong rowSum,
                             jlong rowMax, jlong total, jlong total64) {
  (void)env, (void)self;
  return cs_reduce_rows(H(cs_stream, stream), (const float *)P(in), rows, cols, (float *)P(rowSum),
                        (float *)P(rowMax), (float *)P(total), (double *)P(total64));
}

JNIEXPORT jint J(reduceCols)(JNIEnv *env, jobject self, jlong stream, jlong in, jlong rows, jlong cols,
                             jlong colSum) {
  (void)env, (void)self;
  return cs_reduce_cols(H(cs_stream, stream), (const float *)P(in), rows, cols, (float *)P(colSum));
}

JNIEXPORT jint J(random)(JNIEnv *env, jobject self, jlong stream, jlong out, jlong n, jint seed) {
  (void)env, (void)self;
  return cs_random(H(cs_stream, stream), (float *)P(out), n, (uint32_t)seed);
}

JNIEXPORT jint J(randomNormal)(JNIEnv *env, jobject self, jlong stream, jlong out, jlong n, jint seed) {
  (void)env, (void)self;
  return cs_random_normal(H(cs_stream, stream), (float *)P(out), n, (uint32_t)seed);
}

JNIEXPORT jint J(gemm)(JNIEnv *env, jobject self, jlong stream, jint variant, jint m, jint n, jint k, jlong a,
                       jint lda, jlong b, jint ldb, jlong c, jint ldc) {
  (void)env, (void)self;
  return cs_gemm(H(cs_stream, stream), variant, m, n, k, (const float *)P(a), lda, (const float *)P(b), ldb,
                 (float *)P(c), ldc);
}
