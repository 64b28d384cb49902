/*
 * csmi.h -- C ABI of the MI355X (gfx950) back-end for Compute.scala's Tensor API.
 *
 * This library replaces two reference modules:
 *   - `OpenCL` (OpenCL/src/main/scala/com/thoughtworks/compute/OpenCL.scala), the LWJGL
 *     binding that `trait Tensors extends OpenCL` (Tensors.scala:199) drives through its
 *     protected members;
 *   - `Memory` (Memory/src/main/scala/com/thoughtworks/compute/Memory.scala), the host
 *     buffer typeclass. Device memory comes from a deterministic pooled hipMalloc arena and
 *     host transfers go through pinned staging.
 * It also exports path-level entry points for the kernels that `Tensors` generates or
 * launches: the fused elementwise/view kernel (Tensors.scala:1291-1392 with
 * OpenCLKernelBuilder.scala:135-571), `sum` (Tensors.scala:673-771), `random` and
 * `randomNormal` (Tensors.scala:397-524), and `matrixMultiply` (README.md:301-362).
 *
 * Conventions
 *   - Every function returns a cs_status: 0 on success, a negative code on failure. The
 *     codes are the OpenCL status values that the reference maps to typed exceptions
 *     (OpenCL.scala:143-312), so a JNI shim can raise the same exception classes.
 *   - The message of the last failure on the calling thread is read with cs_last_error().
 *   - Handles are reference counted: *_retain / *_release, released exactly once per
 *     reference, as cl_mem and cl_event are (OpenCL.scala:565-642).
 *   - Device pointers are plain `void*` / `float*` into device memory of the device the
 *     stream belongs to. No C++ exception crosses this interface.
 *   - Every entry point is thread safe. Calls that share one stream execute in order.
 */
#ifndef CSMI_H
#define CSMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status codes */
typedef int cs_status;
#define CS_SUCCESS                       0
#define CS_DEVICE_NOT_FOUND             (-1)   /* CL_DEVICE_NOT_FOUND */
#define CS_MEM_OBJECT_ALLOCATION_FAILURE (-4)  /* CL_MEM_OBJECT_ALLOCATION_FAILURE */
#define CS_OUT_OF_RESOURCES             (-5)   /* CL_OUT_OF_RESOURCES */
#define CS_OUT_OF_HOST_MEMORY           (-6)   /* CL_OUT_OF_HOST_MEMORY */
#define CS_BUILD_PROGRAM_FAILURE        (-11)  /* CL_BUILD_PROGRAM_FAILURE */
#define CS_INVALID_VALUE                (-30)  /* CL_INVALID_VALUE */
#define CS_INVALID_DEVICE               (-33)  /* CL_INVALID_DEVICE */
#define CS_INVALID_MEM_OBJECT           (-38)  /* CL_INVALID_MEM_OBJECT */
#define CS_INVALID_PROGRAM              (-44)  /* CL_INVALID_PROGRAM */
#define CS_INVALID_KERNEL_NAME          (-46)  /* CL_INVALID_KERNEL_NAME */
#define CS_INVALID_KERNEL_ARGS          (-52)  /* CL_INVALID_KERNEL_ARGS */
#define CS_INVALID_WORK_GROUP_SIZE      (-54)  /* CL_INVALID_WORK_GROUP_SIZE */
#define CS_INVALID_EVENT                (-58)  /* CL_INVALID_EVENT */
#define CS_INVALID_OPERATION            (-59)  /* CL_INVALID_OPERATION */
#define CS_INVALID_BUFFER_SIZE          (-61)  /* CL_INVALID_BUFFER_SIZE */

/* Copies the last failure message of the calling thread into buf (always NUL terminated)
 * and returns its full length. Replaces the exception messages of OpenCL.scala:143-312. */
size_t cs_last_error(char *buf, size_t buf_size);

/* The name of a status code ("CS_BUILD_PROGRAM_FAILURE", ...; "CS_UNKNOWN_STATUS" for any
 * other value), a static string. The JNI shim maps codes to exception classes by value
 * (OpenCL.scala:143-312); this name is for messages. */
const char *cs_status_string(cs_status status);

/* Library and build identification ("csmi <version> gfx950"). */
const char *cs_version(void);

/* ---------------------------------------------------------------- opaque handles */
typedef struct cs_device  cs_device;
typedef struct cs_stream  cs_stream;
typedef struct cs_event   cs_event;
typedef struct cs_buffer  cs_buffer;
typedef struct cs_program cs_program;
typedef struct cs_kernel  cs_kernel;
typedef struct cs_graph   cs_graph;
typedef struct cs_fused   cs_fused;

/* ---------------------------------------------------------------- devices
 * Replaces platform/device selection and the `DeviceId` queries (OpenCL.scala:340-374,
 * 450-527) that Tensors reads: deviceType (Tensors.scala:682), maxComputeUnits
 * (Tensors.scala:701) and the kernel work-group size (Tensors.scala:700, 738). */
cs_status cs_device_count(int *count);
cs_status cs_device_open(int ordinal, cs_device **device);
cs_status cs_device_close(cs_device *device);

typedef enum cs_info_key {
  CS_INFO_DEVICE_TYPE = 1,        /* CL_DEVICE_TYPE_GPU (4) for every device here */
  CS_INFO_COMPUTE_UNITS = 2,      /* CUs (256 on MI355X) */
  CS_INFO_MAX_WORK_GROUP_SIZE = 3,
  CS_INFO_WAVEFRONT_SIZE = 4,
  CS_INFO_LDS_BYTES = 5,          /* LDS per workgroup */
  CS_INFO_GLOBAL_MEM_BYTES = 6,
  CS_INFO_L2_BYTES = 7,
  CS_INFO_CLOCK_KHZ = 8,
  CS_INFO_ARENA_RESERVED_BYTES = 9,  /* bytes the arena holds from hipMalloc */
  CS_INFO_ARENA_IN_USE_BYTES = 10,   /* bytes handed out to live buffers */
  CS_INFO_ARENA_SLABS = 11,          /* number of hipMalloc slabs */
  CS_INFO_FUSED_CACHE_ENTRIES = 12   /* compiled fused kernels in the cache */
} cs_info_key;
cs_status cs_device_info(cs_device *device, int key, int64_t *value);
/* Device name string, e.g. "gfx950". Owned by the device. */
cs_status cs_device_name(cs_device *device, const char **name);

/* ---------------------------------------------------------------- streams and events
 * A stream stands in for one in-order command queue of `CommandQueuePool`
 * (OpenCL.scala:423-448). Events replace cl_event (OpenCL.scala:565-612, 1246-1263). */
cs_status cs_stream_create(cs_device *device, cs_stream **stream);
cs_status cs_stream_destroy(cs_stream *stream);
cs_status cs_stream_synchronize(cs_stream *stream);
cs_status cs_stream_device(cs_stream *stream, cs_device **device);
/* The native hipStream_t, for callers that launch their own HIP work on it. */
cs_status cs_stream_native(cs_stream *stream, void **hip_stream);

cs_status cs_event_record(cs_stream *stream, cs_event **event);      /* refcount = 1 */
cs_status cs_event_retain(cs_event *event);
cs_status cs_event_release(cs_event *event);
cs_status cs_event_wait(cs_event *event);                            /* host blocks */
cs_status cs_event_query(cs_event *event, int *complete);
cs_status cs_stream_wait_event(cs_stream *stream, cs_event *event);  /* device-side wait */
cs_status cs_event_elapsed_ms(cs_event *start, cs_event *end, float *ms);
/* Runs `callback(user)` on a library thread once `event` completes. The callback must
 * not call back into HIP (the same rule as OpenCL.scala:376-412 for AMD/Intel). */
cs_status cs_event_on_complete(cs_event *event, void (*callback)(void *user), void *user);

/* ---------------------------------------------------------------- buffers
 * Replaces `allocateBuffer`, `allocateBufferFrom` (OpenCL.scala:1397-1431), `DeviceBuffer`
 * retain/release/toHostBuffer (OpenCL.scala:614-717) and the Memory typeclass. Buffers are
 * carved from a pooled arena: the same sequence of alloc/release calls on one stream gives
 * the same device addresses. Blocks are 256-byte aligned. */
cs_status cs_buffer_alloc(cs_device *device, size_t bytes, cs_buffer **buffer);   /* refcount 1 */
/* Stream-ordered allocation (refcount 1): the same deterministic addresses as cs_buffer_alloc,
 * but the host never waits for earlier releases; `stream` waits on the device for the work
 * that other streams had queued when a reused block was released. The buffer is ready for
 * work queued on `stream` (or ordered after it). Replaces the per-materialisation
 * clCreateBuffer of allocateBuffer (OpenCL.scala:1397-1431) without its host round trip. */
cs_status cs_buffer_alloc_on(cs_stream *stream, size_t bytes, cs_buffer **buffer);
cs_status cs_buffer_retain(cs_buffer *buffer);
cs_status cs_buffer_release(cs_buffer *buffer);
cs_status cs_buffer_device_ptr(cs_buffer *buffer, void **device_ptr);
cs_status cs_buffer_size(cs_buffer *buffer, size_t *bytes);
/* Host -> device through pinned staging. Blocking when `blocking` != 0; otherwise the host
 * memory may be reused as soon as the call returns (it is staged). */
cs_status cs_buffer_write(cs_stream *stream, cs_buffer *buffer, size_t offset,
                          const void *host, size_t bytes, int blocking);
/* Device -> host through pinned staging. Always completes before returning. */
cs_status cs_buffer_read(cs_stream *stream, cs_buffer *buffer, size_t offset,
                         void *host, size_t bytes);
/* Raw-pointer variants, for memory the caller owns (e.g. a torch tensor). */
cs_status cs_memcpy_htod(cs_stream *stream, void *device_dst, const void *host_src, size_t bytes);
cs_status cs_memcpy_dtoh(cs_stream *stream, void *host_dst, const void *device_src, size_t bytes);
cs_status cs_memcpy_dtod(cs_stream *stream, void *device_dst, const void *device_src, size_t bytes);
/* Releases cached arena slabs that hold no live buffer back to HIP. */
cs_status cs_arena_trim(cs_device *device);

/* ---------------------------------------------------------------- programs and kernels
 * Replaces createProgramWithSource + Program.build (OpenCL.scala:1266-1294, 917-932),
 * createKernel (OpenCL.scala:1164-1173) and Kernel.enqueue (OpenCL.scala:787-844). Source
 * is HIP C++ compiled by hiprtc for gfx950; `options` are extra hiprtc options. On a build
 * failure the status is CS_BUILD_PROGRAM_FAILURE and *log (if log != NULL) receives the
 * compiler log, to be freed with cs_free (the BuildProgramFailure of OpenCL.scala:172-181). */
cs_status cs_program_build(cs_device *device, const char *source, const char *options,
                           cs_program **program, char **log);
cs_status cs_program_release(cs_program *program);
cs_status cs_program_get_kernel(cs_program *program, const char *name, cs_kernel **kernel);
cs_status cs_kernel_release(cs_kernel *kernel);
/* The largest workgroup size `kernel` can be launched with (its launch bounds and register
 * use), the `kernel.workGroupSize(deviceId)` query of Tensors.scala:700, 738
 * (CL_KERNEL_WORK_GROUP_SIZE, OpenCL.scala:739-749). */
cs_status cs_kernel_max_block(cs_kernel *kernel, int *max_block);
/* Launches `kernel` with hipModuleLaunchKernel semantics: args[i] points at the value of
 * argument i. grid is in workgroups. */
cs_status cs_launch(cs_stream *stream, cs_kernel *kernel, const uint32_t grid[3],
                    const uint32_t block[3], uint32_t dynamic_lds_bytes, void **args, int nargs);
void cs_free(void *p);

/* ---------------------------------------------------------------- graphs
 * Stream capture into a hipGraph, replayed with one launch; used for batches of
 * launch-bound work (no reference counterpart: the reference enqueues every kernel). */
cs_status cs_graph_begin_capture(cs_stream *stream);
cs_status cs_graph_end_capture(cs_stream *stream, cs_graph **graph);
cs_status cs_graph_launch(cs_stream *stream, cs_graph *graph);
cs_status cs_graph_release(cs_graph *graph);

/* ---------------------------------------------------------------- fused expression kernels
 * The serialised form of a Trees DAG (Trees.scala:303-1050) as built by `Tensor` closures.
 * Nodes reference earlier nodes by index. Leaves are numbered in first-visit depth-first
 * order, which is the argument order of Tensors.scala:230-251 / 1346-1376. */
#define CS_MAX_RANK 8

typedef enum cs_op {
  CS_OP_LITERAL = 1,    /* scalar float `value` (Trees FloatLiteral; Expressions float.literal) */
  CS_OP_PARAMETER = 2,  /* ArrayParameter: leaf `param`, padding `value`, shape[rank] */
  CS_OP_TRANSFORM = 3,  /* view of array node `a`: output shape[rank], matrix of
                           rank_src x (rank+1) doubles (NDimensionalAffineTransform) */
  CS_OP_FILL = 4,       /* array whose every element is float node `a` */
  CS_OP_EXTRACT = 5,    /* float at the current output coordinate of array node `a` */
  CS_OP_NEG = 10, CS_OP_ABS = 11, CS_OP_SQRT = 12, CS_OP_TANH = 13, CS_OP_EXP = 14,
  CS_OP_LOG = 15,
  CS_OP_ADD = 20, CS_OP_SUB = 21, CS_OP_MUL = 22, CS_OP_DIV = 23, CS_OP_MOD = 24,
  CS_OP_MIN = 25, CS_OP_MAX = 26
} cs_op;

typedef struct cs_node {
  int32_t op;                     /* cs_op */
  int32_t a, b;                   /* operand node indices (-1 when unused) */
  int32_t param;                  /* CS_OP_PARAMETER: leaf slot */
  float value;                    /* CS_OP_LITERAL value, CS_OP_PARAMETER padding */
  int32_t rank;                   /* PARAMETER: source rank; TRANSFORM: output rank */
  int32_t shape[CS_MAX_RANK];     /* PARAMETER: source shape; TRANSFORM: output shape */
  double matrix[CS_MAX_RANK * (CS_MAX_RANK + 1)]; /* TRANSFORM only, row-major */
} cs_node;

typedef enum cs_math_mode {
  CS_MATH_PRECISE = 0,  /* no FMA contraction, correctly rounded + - * / sqrt; exp, tanh and
                           log from the device math library (each within 2 ulp) */
  CS_MATH_FAST = 1,     /* as -cl-unsafe-math-optimizations (OpenCL.scala:1131-1135):
                           contraction, reassociation, reciprocals, no signed zeros, and the
                           hardware-rate exp and log; infinities and NaN keep their meaning */
  CS_MATH_ACCURATE = 2  /* as PRECISE, with exp and tanh within 1 ulp (compensated
                           polynomial evaluation; about twice the instructions) */
} cs_math_mode;

typedef struct cs_fused_options {
  int32_t math_mode;          /* cs_math_mode */
  int32_t vector_width;       /* consecutive innermost outputs per thread: 1,2,4,8; 0 = auto */
  int32_t items_per_thread;   /* vectors per thread: 1..8; 0 = auto */
  int32_t block_size;         /* threads per workgroup: 64..1024; 0 = auto */
  int32_t nontemporal_store;  /* 1 = stream the output past the caches */
  int32_t launch_order;       /* cs_launch_order: which outputs a workgroup computes together */
} cs_fused_options;

/* The automatic launch shape: vector_width 4 (halved until it divides the innermost extent),
 * items_per_thread 2 from 2^22 output positions and 1 below, 256 threads. An expression of
 * more than 256 canonical nodes (such as the convolution idiom) holds hundreds of values per
 * element: it gets vector_width 1 and items_per_thread 1, and is compiled for its own
 * workgroup size (a larger launch of it is CS_INVALID_WORK_GROUP_SIZE). */

/* CS_ORDER_AUTO: when a leaf is read through a view whose source rows run along an outer
 * output dimension (a transpose), workgroups take 2-D tiles across that dimension, so the
 * leaf's rows are read in whole cache lines; otherwise, and for fused sums, the linear order.
 * CS_ORDER_LINEAR: consecutive vectors of the output, always. Results are the same bits. */
enum { CS_ORDER_AUTO = 0, CS_ORDER_LINEAR = 1 };

/* Compiles (or fetches from the structural cache) the kernel that evaluates
 * outputs[0..noutputs) at every coordinate of out_shape. noutputs > 1 writes an
 * array-of-structs of noutputs floats per coordinate (`Tensor.join`, Tensors.scala:577-598).
 * Equivalent of enqueueClosure's compile step (Tensors.scala:1293-1386). */
cs_status cs_fused_compile(cs_device *device, const cs_node *nodes, int nnodes,
                           const int32_t *outputs, int noutputs,
                           int out_rank, const int32_t *out_shape,
                           const cs_fused_options *options, cs_fused **fused);
cs_status cs_fused_retain(cs_fused *fused);
cs_status cs_fused_release(cs_fused *fused);
/* Generated HIP source text (owned by `fused`). */
cs_status cs_fused_source(cs_fused *fused, const char **source);
/* Number of leaves the kernel takes and bytes of its output. */
cs_status cs_fused_signature(cs_fused *fused, int *nleaves, int64_t *out_elements);
/* 64-bit structural key (canonical, parameter ids numbered by first visit). */
cs_status cs_fused_key(cs_fused *fused, uint64_t *key);
/* Launches the kernel: leaves[i] is the device pointer bound to leaf slot i
 * (compiledKernel.run, Tensors.scala:1331-1391). */
cs_status cs_fused_launch(cs_stream *stream, cs_fused *fused, const void *const *leaves,
                          int nleaves, void *out);
/* Structural cache probe: 1 in *hit if an equal DAG was compiled on this device
 * (kernelCache.getIfPresent, Tensors.scala:1293). */
cs_status cs_fused_lookup(cs_device *device, const cs_node *nodes, int nnodes,
                          const int32_t *outputs, int noutputs, int out_rank,
                          const int32_t *out_shape, const cs_fused_options *options, int *hit);
/* Drops every cached fused kernel (clearCache, Tensors.scala:1282-1289). */
cs_status cs_fused_cache_clear(cs_device *device);
/* Generates the HIP source of the fused kernel without a device (the text that
 * generateKernelSourceCode returns, OpenCLKernelBuilder.scala:135-221). *source is
 * malloc'ed, to be freed with cs_free; *key (nullable) receives the structural key. */
cs_status cs_fused_generate_source(const cs_node *nodes, int nnodes, const int32_t *outputs,
                                   int noutputs, int out_rank, const int32_t *out_shape,
                                   const cs_fused_options *options, char **source, uint64_t *key);
/* Many expressions at once: one hiprtc module holds every kernel not already cached (one
 * hiprtc call instead of n), and out[i] receives the handle of descs[i]. */
typedef struct cs_fused_desc {
  const cs_node *nodes;
  int nnodes;
  const int32_t *outputs;
  int noutputs;
  int out_rank;
  const int32_t *out_shape;
  cs_fused_options options;
} cs_fused_desc;
cs_status cs_fused_compile_many(cs_device *device, const cs_fused_desc *descs, int n, cs_fused **out);
/* A reduction fused into its producer (the fusion barrier a `sum` puts after an inline
 * tensor, Tensors.scala:673-771 over 1291-1392, removed): cs_fused_compile_sum builds the
 * kernel of a single-output DAG that adds its elements (compensated FP32 per thread, FP64
 * across the block) into FP64 block partials instead of storing them, and cs_fused_sum runs
 * it and combines the partials in a fixed order: *out (FP32, nullable) and *out64 (FP64,
 * nullable) get the sum of the values the element-wise kernel would have written, within
 * 4 eps32 sum|x|, the same on every run. Such a handle is rejected by the element-wise
 * launches and by batches. */
cs_status cs_fused_compile_sum(cs_device *device, const cs_node *nodes, int nnodes, const int32_t *outputs,
                               int noutputs, int out_rank, const int32_t *out_shape,
                               const cs_fused_options *options, cs_fused **fused);
cs_status cs_fused_sum(cs_stream *stream, cs_fused *fused, const void *const *leaves, int nleaves, float *out,
                       double *out64);
/* cs_fused_launch that also returns two new events (refcount 1) stamped with the kernel
 * dispatch's own start and end, as OpenCL's profiling info gives for a command
 * (CL_PROFILING_COMMAND_START/END; the reference maps CL_PROFILING_INFO_NOT_AVAILABLE at
 * OpenCL.scala:259): cs_event_elapsed_ms(*start, *stop) is the kernel's duration, without
 * the launch gap before it. Not for use under graph capture. */
cs_status cs_fused_launch_timed(cs_stream *stream, cs_fused *fused, const void *const *leaves, int nleaves,
                                void *out, cs_event **start, cs_event **stop);
/* cs_fused_launch with the workgroup size chosen at launch (0 = the compiled default):
 * the generated code reads the workgroup size at run time, so one compile serves every
 * size the autotuner tries. A large expression (more than 256 nodes after merging equal
 * subtrees) is compiled for its own workgroup size only, so that a thread may hold more
 * registers: a larger block is rejected with CS_INVALID_WORK_GROUP_SIZE. */
cs_status cs_fused_launch_block(cs_stream *stream, cs_fused *fused, const void *const *leaves, int nleaves,
                                void *out, int block);
/* Compiles the fused kernel into the on-disk code-object cache without a device, so that a
 * later cs_fused_compile of the same expression and options only loads it. The cache lives
 * in $CSMI_KERNEL_CACHE or in `kcache` beside the library's directory. */
cs_status cs_fused_precompile(const cs_node *nodes, int nnodes, const int32_t *outputs, int noutputs,
                              int out_rank, const int32_t *out_shape, const cs_fused_options *options);
/* Compiles `source` for gfx950 with hiprtc without loading it (no device needed); on
 * failure returns CS_BUILD_PROGRAM_FAILURE with the log in *log (free with cs_free). */
cs_status cs_program_check(const char *source, const char *options, char **log);

/* ---------------------------------------------------------------- batches
 * Many fused expressions in one persistent kernel. The reference launches one kernel per
 * expression (Tensors.scala:1291-1392, at most 5 in flight per device, OpenCL.scala:1313);
 * here every expression's body becomes a device function (with its own vector width and
 * vectors per thread), and a fixed grid drains a queue of (expression, tile) work items
 * through an atomic counter. cs_batch_launch enqueues a counter reset and the persistent
 * kernel, so it can be captured in a graph and replayed. */
typedef struct cs_batch cs_batch;
cs_status cs_batch_create(cs_device *device, cs_fused *const *exprs, int n, int block, cs_batch **batch);
/* args: for each expression in order, its leaf pointers then its output pointer. Copies the
 * table to the device and waits (call outside graph capture). */
cs_status cs_batch_bind(cs_stream *stream, cs_batch *batch, const void *const *args, int nargs);
cs_status cs_batch_launch(cs_stream *stream, cs_batch *batch);
cs_status cs_batch_info(cs_batch *batch, int64_t *items, int *grid, int *block);
cs_status cs_batch_release(cs_batch *batch);
/* Generates the batch kernel's source without a device (the text cs_batch_create builds):
 * one device function per expression of `descs` and the persistent kernel. *source is
 * malloc'ed, to be freed with cs_free; *items (nullable) receives the work-item count. */
cs_status cs_batch_generate_source(const cs_fused_desc *descs, int n, int block, char **source, int64_t *items);

/* ---------------------------------------------------------------- cost-model network
 * A fully connected network widths[0] -> ... -> widths[nw-1] with ReLU after every layer but
 * the last, trained and evaluated on the device in FP32 (the learned cost model that
 * predicts fused-kernel times; the reference has no tuning). Parameters are one flat array
 * [W1 (w0 x w1, row-major [in][out]), b1, W2, b2, ...], initialised He-uniform from `seed`
 * with zero biases. X, y, target, loss and grads are device pointers; every reduction runs
 * in a fixed order, so results are deterministic. */
typedef struct cs_mlp cs_mlp;
cs_status cs_mlp_create(cs_device *device, const int32_t *widths, int nw, uint32_t seed, cs_mlp **mlp);
cs_status cs_mlp_param_count(cs_mlp *mlp, int64_t *count);
cs_status cs_mlp_get_params(cs_stream *stream, cs_mlp *mlp, float *host, int64_t count);
cs_status cs_mlp_set_params(cs_stream *stream, cs_mlp *mlp, const float *host, int64_t count);
/* y[n x widths[nw-1]] = network(X[n x widths[0]]) */
cs_status cs_mlp_forward(cs_stream *stream, cs_mlp *mlp, const float *X, int64_t n, float *y);
/* Mean squared error of network(X) (one output) against target[n]: *loss (nullable) and the
 * gradient of every parameter in grads[param_count]. */
cs_status cs_mlp_backward(cs_stream *stream, cs_mlp *mlp, const float *X, const float *target, int64_t n,
                          float *loss, float *grads);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8); t counts steps from 1. */
cs_status cs_mlp_adam(cs_stream *stream, cs_mlp *mlp, const float *grads, float lr, int t);
cs_status cs_mlp_release(cs_mlp *mlp);

/* ---------------------------------------------------------------- reductions
 * Replaces `sum` (Tensors.scala:673-771) with a correct tree (the reference's parallel
 * program drops elements for non power-of-two local sizes, Tensors.scala:375-387).
 * Per-thread and per-wave sums are FP32; block partials are combined in FP64. *out gets
 * the FP32 rounding of the FP64 total; *out64 (nullable) the FP64 total. */
cs_status cs_reduce_sum(cs_stream *stream, const float *in, int64_t n, float *out, double *out64);
/* Maximum of n floats (fmax semantics: NaN operands are ignored). */
cs_status cs_reduce_max(cs_stream *stream, const float *in, int64_t n, float *out);
/* One pass over a row-major [rows, cols] matrix. Each output pointer is nullable:
 * row_sum[rows], row_max[rows], total (FP32) and total64 (FP64) of all elements.
 * The reference expresses these only as `split(1).reduce(_ + _)` / `reduce(max)` chains
 * (README.md:342, TensorsSpec.scala:477). */
cs_status cs_reduce_rows(cs_stream *stream, const float *in, int64_t rows, int64_t cols,
                         float *row_sum, float *row_max, float *total, double *total64);
/* cs_reduce_rows without allocating: no counterpart in the reference. For callers that
 * cannot allocate device memory (code under graph capture, serving loops, hosts that forbid
 * hipMalloc). The same kernels and results as cs_reduce_rows; the FP64 row partials of the
 * total live in a fixed device array of the library (8 MiB per GPU, held until the process
 * ends), so a total (total or total64 non-null) is limited to 1048576 rows
 * (CS_INVALID_BUFFER_SIZE beyond). Calls on one GPU share that array: calls that overlap in
 * time on that GPU (on different streams) race and must not be made. */
cs_status cs_reduce_rows_noalloc(cs_stream *stream, const float *in, int64_t rows, int64_t cols,
                                 float *row_sum, float *row_max, float *total, double *total64);
/* Column sums of a row-major [rows, cols] matrix (the `split(0).reduce(_ + _)` idiom). */
cs_status cs_reduce_cols(cs_stream *stream, const float *in, int64_t rows, int64_t cols,
                         float *col_sum);

/* ---------------------------------------------------------------- random numbers
 * The reference's Wang-hash generators (Tensors.scala:106-117, 397-443, 479-524).
 * cs_random: out[i] = hash(i ^ seed) / 2^32 for i < n (bit exact with the reference).
 * cs_random_normal: Box-Muller pairs, writes 2*ceil(n/2) floats (odd n padded, as
 * Tensors.scala:509-513 does), so `out` must hold n + (n & 1) floats. */
cs_status cs_random(cs_stream *stream, float *out, int64_t n, uint32_t seed);
cs_status cs_random_normal(cs_stream *stream, float *out, int64_t n, uint32_t seed);
/* Synthetic-input generators built on the same hash: out[i] = lo + (hi-lo)*u_i and
 * N(0,1) from Box-Muller, each computed in FP64 and rounded once to FP32. */
cs_status cs_fill_uniform(cs_stream *stream, float *out, int64_t n, uint32_t seed,
                          double lo, double hi);
cs_status cs_fill_normal(cs_stream *stream, float *out, int64_t n, uint32_t seed);

/* ---------------------------------------------------------------- matrix multiplication
 * C[M,N] = A[M,K] . B[K,N], row-major with leading dimensions lda/ldb/ldc in elements.
 * The reference builds this from split/broadcast/join (README.md:301-362,
 * TensorsSpec.scala:468-528); every variant here accumulates in FP32. */
typedef enum cs_gemm_variant {
  CS_GEMM_FP32 = 0,     /* FP32 operands, FP32 products and sums (reference semantics) */
  CS_GEMM_BF16 = 1,     /* operands rounded to BF16, BF16 MFMA, FP32 accumulate */
  CS_GEMM_BF16X3 = 2,   /* split-BF16x3: A=Ah+Al, B=Bh+Bl, AhBh + AhBl + AlBh, FP32 accumulate */
  CS_GEMM_FP8 = 3       /* FP8 e4m3 with one scale per 128-wide K block of each row of A
                           and each column of B, FP32 accumulate */
} cs_gemm_variant;
cs_status cs_gemm(cs_stream *stream, int variant, int M, int N, int K,
                  const float *A, int lda, const float *B, int ldb, float *C, int ldc);

/* ---------------------------------------------------------------- accuracy check
 * For every FP32 bit pattern in [start, start + count), compares the elementwise function
 * `op` (CS_OP_EXP, CS_OP_TANH, CS_OP_LOG, CS_OP_SQRT or CS_CHECK_EXP_BOUNDED, as math mode `mode` computes it in
 * generated kernels) with its FP64 counterpart rounded to FP32. stats_device points at 3 zero-initialised
 * device uint64 values that receive max(ulp distance), the count of inputs over 1 ulp and
 * the bits of the worst input. The reference has no counterpart: its OpenCL runtime's
 * math library is not testable from its API. */
/* For cs_math_check only: exp as the precise mode generates it where its argument cannot
 * exceed 88 (the negation of a square or of an absolute value); meaningful for x <= 0. */
#define CS_CHECK_EXP_BOUNDED 114
cs_status cs_math_check(cs_stream *stream, int op, int mode, uint32_t start, uint32_t count,
                        unsigned long long *stats_device);

/* ---------------------------------------------------------------- HBM probe (measurement)
 * A streaming pass with no arithmetic over n floats (n a multiple of 4, 16-byte aligned
 * buffers): read src (kind READ), write dst (WRITE) or copy src to dst (COPY), with
 * non-temporal loads and the given store form, items_per_thread (1, 2, 4 or 8) 16-byte
 * vectors per thread in 256-thread workgroups, one vector of a wave contiguous per item (the
 * generated kernels' layout). Its time is the ceiling a memory-bound kernel of the same size
 * and shape can reach. No counterpart in the reference; the Tensor path never calls it. */
#define CS_PROBE_READ 0
#define CS_PROBE_WRITE 1
#define CS_PROBE_COPY 2
#define CS_PROBE_STORE_PLAIN 0
#define CS_PROBE_STORE_NT 1
#define CS_PROBE_STORE_NT_WRITE_THROUGH 2
cs_status cs_hbm_probe(cs_stream *stream, int kind, int store, const float *src, float *dst, int64_t n,
                       int items_per_thread);

#ifdef __cplusplus
}
#endif

#endif /* CSMI_H */
