// Reductions for gfx950: full sum and max, row-wise sum/max/total in one pass, column sums.
//
// Replaces the reference's `sum` programs (Tensors.scala:308-393, 673-771). Layout and
// arithmetic:
//   * 256-thread workgroups (4 wave64s); each thread reads 16-byte vectors, several in
//     flight, so a row of 4096 floats is 4 loads per thread;
//   * per-thread sums are FP32 with a TwoSum compensation term, combined across the wave
//     with xor shuffles and across the 4 waves through LDS, still as (sum, compensation)
//     pairs, so the in-block result is accurate to about one FP32 rounding;
//   * block results are combined in FP64: every block writes an FP64 partial and a second,
//     one-workgroup kernel adds the partials in a fixed order, so results are deterministic
//     run to run. A last-workgroup ticket in the first kernel was built and measured slower:
//     each workgroup's write-through of its partial, the wait for it and the atomic on one
//     shared counter cost more than the second launch (C3 49.7 us against 44.0, a 16 MiB sum
//     15.9 us against 5.8). One wave per row with eight rows per workgroup (so one partial
//     per eight rows) made the combine 0.45 us shorter and the row pass 0.8 us longer;
//   * max uses fmaxf, which returns the non-NaN operand (OpenCLKernelBuilder.scala:279-285).
// The reference's parallel program loses elements when the local size is not a power of
// two (Tensors.scala:375-387); these kernels sum every element.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "internal.h"

#pragma clang fp contract(off)

namespace csmi {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e) {
  s = a + b;
  float bp = s - a;
  e = (a - (s - bp)) + (b - bp);
}

// Once the running sum s is infinite or NaN it stays so (inf + finite = inf, inf - inf and
// NaN + x = NaN), and only then can a TwoSum error term be NaN. So c is used only while s is
// finite: the result is s itself otherwise, as a plain FP32 sum of the same values gives
// (inf for [1, inf, 2] or [3e38, 3e38], NaN for [inf, -inf]).
struct Acc {
  float s = 0.0f, c = 0.0f;
  __device__ __forceinline__ float result() const { return __builtin_isfinite(s) ? s + c : s; }
  __device__ __forceinline__ double result64() const {
    return __builtin_isfinite(s) ? double(s) + double(c) : double(s);
  }
  __device__ __forceinline__ void add(float x) {
    float t, e;
    two_sum(s, x, t, e);
    s = t;
    c += e;
  }
  __device__ __forceinline__ void merge(float s2, float c2) {
    float t, e;
    two_sum(s, s2, t, e);
    s = t;
    c += c2 + e;
  }
};

__device__ __forceinline__ float nanmax(float a, float b) { return fmaxf(a, b); }

// Sum pair and max across the 64 lanes of a wave.
__device__ __forceinline__ void wave_reduce(Acc &a, float &m) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    float s2 = __shfl_xor(a.s, off, 64);
    float c2 = __shfl_xor(a.c, off, 64);
    float m2 = __shfl_xor(m, off, 64);
    a.merge(s2, c2);
    m = nanmax(m, m2);
  }
}

// Reduces (a, m) over the block; the result is valid in thread 0.
__device__ __forceinline__ void block_reduce(Acc &a, float &m) {
  __shared__ float ls[kWaves], lc[kWaves], lm[kWaves];
  wave_reduce(a, m);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    ls[wave] = a.s;
    lc[wave] = a.c;
    lm[wave] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    Acc t;
    t.s = ls[0];
    t.c = lc[0];
    float mm = lm[0];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) {
      t.merge(ls[w], lc[w]);
      mm = nanmax(mm, lm[w]);
    }
    a = t;
    m = mm;
  }
}

// One workgroup of kFinish threads adds partials[0..n) in FP64 in a fixed order: thread t
// adds pairs 2t, 2t + 2 kFinish, ... (16-byte loads, eight in flight), then a fixed xor-shuffle
// tree inside each wave and the 16 waves in order.
constexpr int kFinish = 1024;
typedef double d2 __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(kFinish) void finish_sum_kernel(const double *__restrict__ partials, int64_t n,
                                                             float *out, double *out64) {
  __shared__ double lw[kFinish / 64];
  double acc = 0.0;
  const int64_t pairs = n / 2;
  const d2 *p2 = reinterpret_cast<const d2 *>(partials);  // partials are 16-byte aligned
  int64_t i = threadIdx.x;
  // eight 16-byte loads in flight per thread: one memory round trip for up to 16384 partials
  for (; i + 7 * kFinish < pairs; i += 8 * kFinish) {
    d2 v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = p2[i + k * kFinish];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += v[k].x + v[k].y;
  }
  for (; i < pairs; i += kFinish) {
    const d2 v = p2[i];
    acc += v.x + v.y;
  }
  if (threadIdx.x == 0 && (n & 1)) acc += partials[n - 1];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) lw[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = lw[0];
#pragma unroll
    for (int w = 1; w < kFinish / 64; ++w) t += lw[w];
    if (out) *out = float(t);
    if (out64) *out64 = t;
  }
}

__global__ __launch_bounds__(kFinish) void finish_max_kernel(const float *__restrict__ partials, int64_t n,
                                                             float *out) {
  __shared__ float lw[kFinish / 64];
  float m = __builtin_nanf("");
  for (int64_t i = threadIdx.x; i < n; i += kFinish) m = nanmax(m, partials[i]);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) m = nanmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) lw[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = lw[0];
#pragma unroll
    for (int w = 1; w < kFinish / 64; ++w) t = nanmax(t, lw[w]);
    *out = t;
  }
}

// ----------------------------------------------------------------------------- full sum/max
template <bool kMax>
__global__ __launch_bounds__(kBlock) void reduce_all_kernel(const float *__restrict__ in, int64_t n,
                                                            double *partial64, float *partialmax) {
  Acc a;
  float m = __builtin_nanf("");
  const int64_t nvec = ((reinterpret_cast<uintptr_t>(in) & 15) == 0) ? n / 4 : 0;
  const f4 *in4 = reinterpret_cast<const f4 *>(in);
  // Block-strided chunks of kBlock vectors: consecutive threads read consecutive vectors.
  const int64_t stride = int64_t(gridDim.x) * kBlock;
  int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  for (; i + 3 * stride < nvec; i += 4 * stride) {
    f4 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = __builtin_nontemporal_load(in4 + i + k * stride);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (kMax) {
        m = nanmax(m, nanmax(nanmax(v[k].x, v[k].y), nanmax(v[k].z, v[k].w)));
      } else {
        a.add(v[k].x); a.add(v[k].y); a.add(v[k].z); a.add(v[k].w);
      }
    }
  }
  for (; i < nvec; i += stride) {
    f4 v = __builtin_nontemporal_load(in4 + i);
    if (kMax) {
      m = nanmax(m, nanmax(nanmax(v.x, v.y), nanmax(v.z, v.w)));
    } else {
      a.add(v.x); a.add(v.y); a.add(v.z); a.add(v.w);
    }
  }
  for (int64_t j = nvec * 4 + int64_t(blockIdx.x) * kBlock + threadIdx.x; j < n; j += stride) {
    float x = in[j];
    if (kMax) m = nanmax(m, x); else a.add(x);
  }
  block_reduce(a, m);
  if (threadIdx.x == 0) {
    if (kMax) partialmax[blockIdx.x] = m;
    else partial64[blockIdx.x] = a.result64();
  }
}

// ----------------------------------------------------------------------------- rows
// One workgroup per row: row sum (compensated FP32, rounded once), row max, and the FP64
// partial of the row for the total.
template <bool kVec>
__global__ __launch_bounds__(kBlock) void reduce_rows_kernel(const float *__restrict__ in, int64_t rows,
                                                             int64_t cols, float *row_sum, float *row_max,
                                                             double *partial64) {
  const int64_t row = blockIdx.x;
  const float *p = in + row * cols;
  Acc a;
  float m = __builtin_nanf("");
  if (kVec) {
    const f4 *p4 = reinterpret_cast<const f4 *>(p);
    const int64_t nv = cols / 4;
    int64_t i = threadIdx.x;
    for (; i + 3 * kBlock < nv; i += 4 * kBlock) {
      f4 v[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) v[k] = __builtin_nontemporal_load(p4 + i + k * kBlock);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        // pairwise inside the vector, then one compensated add
        float s01, e01, s23, e23, s, e;
        two_sum(v[k].x, v[k].y, s01, e01);
        two_sum(v[k].z, v[k].w, s23, e23);
        two_sum(s01, s23, s, e);
        a.merge(s, e + e01 + e23);
        m = nanmax(m, nanmax(nanmax(v[k].x, v[k].y), nanmax(v[k].z, v[k].w)));
      }
    }
    for (; i < nv; i += kBlock) {
      f4 v = __builtin_nontemporal_load(p4 + i);
      a.add(v.x); a.add(v.y); a.add(v.z); a.add(v.w);
      m = nanmax(m, nanmax(nanmax(v.x, v.y), nanmax(v.z, v.w)));
    }
  } else {
    for (int64_t i = threadIdx.x; i < cols; i += kBlock) {
      float x = p[i];
      a.add(x);
      m = nanmax(m, x);
    }
  }
  block_reduce(a, m);
  if (threadIdx.x == 0) {
    if (row_sum) row_sum[row] = a.result();
    if (row_max) row_max[row] = m;
    if (partial64) partial64[row] = a.result64();
  }
}

// ----------------------------------------------------------------------------- columns
// Column sums: thread = column, blockIdx.y = a slice of rows; FP64 partials per slice, then
// one thread per column adds its slices in order. Rows are read eight at a time, all eight
// loads in flight before the first addition (the additions stay in row order).
__global__ __launch_bounds__(kBlock) void reduce_cols_kernel(const float *__restrict__ in, int64_t rows,
                                                             int64_t cols, int64_t rows_per_slice,
                                                             double *partial64) {
  const int64_t col = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (col >= cols) return;
  const int64_t r0 = int64_t(blockIdx.y) * rows_per_slice;
  const int64_t r1 = r0 + rows_per_slice < rows ? r0 + rows_per_slice : rows;
  Acc a;
  const float *p = in + r0 * cols + col;
  int64_t r = r0;
  for (; r + 8 <= r1; r += 8, p += 8 * cols) {
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = __builtin_nontemporal_load(p + k * cols);
#pragma unroll
    for (int k = 0; k < 8; ++k) a.add(v[k]);
  }
  for (; r < r1; ++r, p += cols) a.add(__builtin_nontemporal_load(p));
  partial64[int64_t(blockIdx.y) * cols + col] = a.result64();
}

// Four adjacent columns per thread with 16-byte loads, when cols is a multiple of 4 and the
// input is 16-byte aligned; the same row order per column as above. One wave per workgroup:
// many small workgroups spread evenly over the CUs.
constexpr int kColsBlock = 64;
__global__ __launch_bounds__(kColsBlock) void reduce_cols4_kernel(const float *__restrict__ in, int64_t rows,
                                                                  int64_t cols, int64_t rows_per_slice,
                                                                  double *partial64) {
  const int64_t c = (int64_t(blockIdx.x) * kColsBlock + threadIdx.x) * 4;
  if (c >= cols) return;
  const int64_t r0 = int64_t(blockIdx.y) * rows_per_slice;
  const int64_t r1 = r0 + rows_per_slice < rows ? r0 + rows_per_slice : rows;
  Acc a0, a1, a2, a3;
  const int64_t step = cols / 4;  // one row, in 16-byte vectors
  const f4 *p = reinterpret_cast<const f4 *>(in + r0 * cols + c);
  int64_t r = r0;
  for (; r + 8 <= r1; r += 8, p += 8 * step) {
    f4 v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = __builtin_nontemporal_load(p + k * step);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      a0.add(v[k].x);
      a1.add(v[k].y);
      a2.add(v[k].z);
      a3.add(v[k].w);
    }
  }
  for (; r < r1; ++r, p += step) {
    const f4 v = __builtin_nontemporal_load(p);
    a0.add(v.x);
    a1.add(v.y);
    a2.add(v.z);
    a3.add(v.w);
  }
  double *o = partial64 + int64_t(blockIdx.y) * cols + c;
  o[0] = a0.result64();
  o[1] = a1.result64();
  o[2] = a2.result64();
  o[3] = a3.result64();
}

// One thread per column adds its slices in order, eight loads in flight.
__global__ __launch_bounds__(kBlock) void finish_cols_kernel(const double *__restrict__ partial64, int64_t cols,
                                                             int64_t slices, float *col_sum) {
  const int64_t col = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (col >= cols) return;
  double s = 0.0;
  const double *p = partial64 + col;
  int64_t y = 0;
  for (; y + 8 <= slices; y += 8, p += 8 * cols) {
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = p[k * cols];
#pragma unroll
    for (int k = 0; k < 8; ++k) s += v[k];
  }
  for (; y < slices; ++y, p += cols) s += *p;
  col_sum[col] = float(s);
}

}  // namespace

// Per-stream scratch holding the FP64 (or FP32 max) partials.
struct ReduceScratch {
  double *partial64;
  float *partialmax;
};

}  // namespace csmi

using namespace csmi;

static ReduceScratch scratch_for(cs_stream *stream, size_t partial_doubles) {
  char *base = static_cast<char *>(stream_scratch(stream, partial_doubles * 8 + 4096));
  ReduceScratch r;
  r.partial64 = reinterpret_cast<double *>(base);
  r.partialmax = reinterpret_cast<float *>(base);
  return r;
}

namespace csmi {
double *sum_partials(cs_stream *stream, size_t n) { return scratch_for(stream, n).partial64; }

void launch_finish_sum(cs_stream *stream, const double *partials, int64_t n, float *out, double *out64) {
  hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(kFinish), 0, stream->stream, partials, n, out, out64);
}
}  // namespace csmi

// ----------------------------------------------------------------------------- no allocation
// cs_reduce_rows_noalloc keeps its row partials in this file-scope device array (one copy per
// GPU, held until the process ends). Host code reaches it only through the address
// hipGetSymbolAddress returns on each device.
constexpr int64_t kNoallocMaxRows = int64_t(1) << 20;
static __device__ __align__(256) double scratch[kNoallocMaxRows];

namespace {
std::mutex g_noalloc_mu;
std::map<int, double *> g_noalloc_scratch;  // device ordinal -> address of `scratch` there

// The device address of `scratch` on `device`, looked up once per device with that device
// current.
double *noalloc_scratch(cs_device *device) {
  std::lock_guard<std::mutex> l(g_noalloc_mu);
  auto it = g_noalloc_scratch.find(device->ordinal);
  if (it != g_noalloc_scratch.end()) return it->second;
  set_device(device);
  void *p = nullptr;
  CS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) fail(CS_OUT_OF_RESOURCES, "no device address for the reduction scratch");
  g_noalloc_scratch[device->ordinal] = static_cast<double *>(p);
  return static_cast<double *>(p);
}
}  // namespace

static int reduce_grid(cs_stream *stream, int64_t n) {
  int cus = stream->device->props.multiProcessorCount;
  int64_t want = (n / 4 + kBlock * 4 - 1) / (kBlock * 4);  // >= 4 vectors per thread
  int64_t cap = int64_t(cus) * 8;
  if (want > cap) want = cap;
  if (want < 1) want = 1;
  return static_cast<int>(want);
}

extern "C" {

cs_status cs_reduce_sum(cs_stream *stream, const float *in, int64_t n, float *out, double *out64) {
  return guard([&] {
    if (!stream) fail(CS_INVALID_VALUE, "stream is null");
    if (n < 0) fail(CS_INVALID_VALUE, "negative length");
    if (n > 0 && !in) fail(CS_INVALID_VALUE, "input is null");
    if (!out && !out64) fail(CS_INVALID_VALUE, "no output");
    set_device(stream->device);
    int grid = reduce_grid(stream, n);
    ReduceScratch s = scratch_for(stream, static_cast<size_t>(grid));
    hipLaunchKernelGGL(reduce_all_kernel<false>, dim3(grid), dim3(kBlock), 0, stream->stream, in, n,
                       s.partial64, s.partialmax);
    hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(kFinish), 0, stream->stream, s.partial64,
                       int64_t(grid), out, out64);
    CS_HIP(hipGetLastError());
  });
}

cs_status cs_reduce_max(cs_stream *stream, const float *in, int64_t n, float *out) {
  return guard([&] {
    if (!stream) fail(CS_INVALID_VALUE, "stream is null");
    if (n < 0) fail(CS_INVALID_VALUE, "negative length");
    if (n > 0 && !in) fail(CS_INVALID_VALUE, "input is null");
    if (!out) fail(CS_INVALID_VALUE, "output is null");
    set_device(stream->device);
    int grid = reduce_grid(stream, n);
    ReduceScratch s = scratch_for(stream, static_cast<size_t>(grid));
    hipLaunchKernelGGL(reduce_all_kernel<true>, dim3(grid), dim3(kBlock), 0, stream->stream, in, n,
                       s.partial64, s.partialmax);
    hipLaunchKernelGGL(finish_max_kernel, dim3(1), dim3(kFinish), 0, stream->stream, s.partialmax,
                       int64_t(grid), out);
    CS_HIP(hipGetLastError());
  });
}

cs_status cs_reduce_rows(cs_stream *stream, const float *in, int64_t rows, int64_t cols, float *row_sum,
                         float *row_max, float *total, double *total64) {
  return guard([&] {
    if (!stream) fail(CS_INVALID_VALUE, "stream is null");
    if (rows < 0 || cols < 0) fail(CS_INVALID_VALUE, "negative shape");
    if (rows > 0x7fffffff) fail(CS_INVALID_VALUE, "more than 2^31-1 rows");
    if (rows > 0 && cols > 0 && !in) fail(CS_INVALID_VALUE, "input is null");
    set_device(stream->device);
    bool want_total = total || total64;
    if (rows == 0) {
      // no rows: the total is an empty sum, written by the combine over zero partials
      if (want_total)
        hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(kFinish), 0, stream->stream, nullptr, int64_t(0), total,
                           total64);
      CS_HIP(hipGetLastError());
      return;
    }
    ReduceScratch s = scratch_for(stream, want_total ? static_cast<size_t>(rows) : 0);
    bool vec = cols % 4 == 0 && (reinterpret_cast<uintptr_t>(in) & 15) == 0;
    auto k = vec ? reduce_rows_kernel<true> : reduce_rows_kernel<false>;
    hipLaunchKernelGGL(k, dim3(static_cast<unsigned>(rows)), dim3(kBlock), 0, stream->stream, in, rows, cols,
                       row_sum, row_max, want_total ? s.partial64 : nullptr);
    if (want_total)
      hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(kFinish), 0, stream->stream, s.partial64, rows, total,
                         total64);
    CS_HIP(hipGetLastError());
  });
}

cs_status cs_reduce_rows_noalloc(cs_stream *stream, const float *in, int64_t rows, int64_t cols, float *row_sum,
                                 float *row_max, float *total, double *total64) {
  return guard([&] {
    if (!stream) fail(CS_INVALID_VALUE, "stream is null");
    if (rows < 0 || cols < 0) fail(CS_INVALID_VALUE, "negative shape");
    if (rows > 0x7fffffff) fail(CS_INVALID_VALUE, "more than 2^31-1 rows");
    if (rows > 0 && cols > 0 && !in) fail(CS_INVALID_VALUE, "input is null");
    const bool want_total = total || total64;
    if (want_total && rows > kNoallocMaxRows)
      fail(CS_INVALID_BUFFER_SIZE, "cs_reduce_rows_noalloc takes at most " + std::to_string(kNoallocMaxRows) +
                                       " rows when a total is asked for, got " + std::to_string(rows));
    set_device(stream->device);
    if (rows == 0) {
      // no rows: the total is an empty sum, written by the combine over zero partials
      if (want_total)
        hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(kFinish), 0, stream->stream, nullptr, int64_t(0), total,
                           total64);
      CS_HIP(hipGetLastError());
      return;
    }
    // every partial the combine reads is written by this call's row kernel first
    double *partial64 = want_total ? noalloc_scratch(stream->device) : nullptr;
    const bool vec = cols % 4 == 0 && (reinterpret_cast<uintptr_t>(in) & 15) == 0;
    auto k = vec ? reduce_rows_kernel<true> : reduce_rows_kernel<false>;
    hipLaunchKernelGGL(k, dim3(static_cast<unsigned>(rows)), dim3(kBlock), 0, stream->stream, in, rows, cols,
                       row_sum, row_max, partial64);
    if (want_total)
      hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(kFinish), 0, stream->stream, partial64, rows, total,
                         total64);
    CS_HIP(hipGetLastError());
  });
}

cs_status cs_reduce_cols(cs_stream *stream, const float *in, int64_t rows, int64_t cols, float *col_sum) {
  return guard([&] {
    if (!stream) fail(CS_INVALID_VALUE, "stream is null");
    if (rows < 0 || cols < 0) fail(CS_INVALID_VALUE, "negative shape");
    if (cols == 0) return;
    if (!col_sum || (rows > 0 && !in)) fail(CS_INVALID_VALUE, "null pointer");
    set_device(stream->device);
    const bool vec = cols % 4 == 0 && (reinterpret_cast<uintptr_t>(in) & 15) == 0;
    const int64_t per_block = vec ? int64_t(kColsBlock) * 4 : kBlock;  // columns per workgroup
    int64_t groups = (cols + per_block - 1) / per_block;
    if (groups > 0x7fffffff) fail(CS_INVALID_VALUE, "too many columns");
    int cus = stream->device->props.multiProcessorCount;
    // about 8 waves per CU, and at least 128 rows per slice so that the partials stay few
    int64_t slices = (int64_t(cus) * (vec ? 32 : 8) + groups - 1) / groups;
    int64_t max_slices = (rows + 127) / 128;
    if (slices > max_slices) slices = max_slices;
    if (slices < 1) slices = 1;
    if (slices > 65535) slices = 65535;
    int64_t per = (rows + slices - 1) / slices;
    if (per < 1) per = 1;
    slices = rows > 0 ? (rows + per - 1) / per : 1;
    ReduceScratch s = scratch_for(stream, static_cast<size_t>(slices * cols));
    const dim3 grid(static_cast<unsigned>(groups), static_cast<unsigned>(slices));
    if (vec)
      hipLaunchKernelGGL(reduce_cols4_kernel, grid, dim3(kColsBlock), 0, stream->stream, in, rows, cols, per,
                         s.partial64);
    else
      hipLaunchKernelGGL(reduce_cols_kernel, grid, dim3(kBlock), 0, stream->stream, in, rows, cols, per, s.partial64);
    hipLaunchKernelGGL(finish_cols_kernel, dim3(static_cast<unsigned>((cols + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       stream->stream, s.partial64, cols, slices, col_sum);
    CS_HIP(hipGetLastError());
  });
}

}  // extern "C"
