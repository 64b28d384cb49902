// Random number kernels: the reference's Wang-hash `random` and Box-Muller `randomNormal`
// (Tensors.scala:106-117, 397-443), plus FP64-exact synthetic-input generators on the same
// hash.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.h"

#pragma clang fp contract(off)

namespace csmi {
namespace {

// Thomas Wang's 32-bit integer hash, in the variant the reference emits
// (Tensors.scala:107-115).
__device__ __forceinline__ uint32_t wang_hash(uint32_t v) {
  v = (v ^ 61u) ^ (v >> 16);
  v *= 9u;
  v ^= v << 4;
  v *= 0x27d4eb2du;
  v ^= v >> 15;
  return v;
}

// Marsaglia xorshift step used by random_normal (Tensors.scala:403-407).
__device__ __forceinline__ uint32_t xorshift32(uint32_t s) {
  uint32_t t1 = s ^ (s << 13);
  uint32_t t2 = t1 ^ (t1 >> 17);
  return t2 ^ (t2 << 5);
}

constexpr int kBlock = 256;

// buffer[i] = hash(i ^ seed) / 2^32, converted to float then scaled (Tensors.scala:436-439).
__global__ __launch_bounds__(kBlock) void random_kernel(float *out, int64_t n, uint32_t seed) {
  int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t h = wang_hash(static_cast<uint32_t>(i) ^ seed);
  out[i] = static_cast<float>(h) / 4294967296.0f;
}

// Pair i writes buffer[2i], buffer[2i+1] (Tensors.scala:409-425). Every FP32 operation is
// correctly rounded: log, cos and sin are evaluated in FP64 and rounded once, as a
// correctly rounded libm does on the host.
__global__ __launch_bounds__(kBlock) void random_normal_kernel(float *out, int64_t pairs, uint32_t seed) {
  int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= pairs) return;
  uint32_t r1 = wang_hash(static_cast<uint32_t>(i) ^ seed);
  uint32_t r2 = xorshift32(r1);
  float u1 = static_cast<float>(r1) / 4294967296.0f;
  float u2 = static_cast<float>(r2) / 4294967296.0f;
  float r = sqrtf(-2.0f * static_cast<float>(log(static_cast<double>(u1))));
  float theta = 2.0f * 3.14159274101257324219f * u2;  // 2 * M_PI_F * u2, left to right
  out[2 * i] = r * static_cast<float>(cos(static_cast<double>(theta)));
  out[2 * i + 1] = r * static_cast<float>(sin(static_cast<double>(theta)));
}

// Synthetic inputs: u = hash(i ^ seed) / 2^32 exactly in FP64; x = lo + (hi - lo) * u,
// rounded once to FP32.
__global__ __launch_bounds__(kBlock) void fill_uniform_kernel(float *out, int64_t n, uint32_t seed,
                                                              double lo, double hi) {
  int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  double u = static_cast<double>(wang_hash(static_cast<uint32_t>(i) ^ seed)) * (1.0 / 4294967296.0);
  double w = hi - lo;
  double t = w * u;
  out[i] = static_cast<float>(lo + t);
}

// N(0,1): u1 = (hash + 1) / 2^32 in (0, 1], u2 = xorshift(hash) / 2^32;
// z = sqrt(-2 ln u1) * cos(2 pi u2) in FP64, rounded once to FP32.
__global__ __launch_bounds__(kBlock) void fill_normal_kernel(float *out, int64_t n, uint32_t seed) {
  int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t h = wang_hash(static_cast<uint32_t>(i) ^ seed);
  double u1 = (static_cast<double>(h) + 1.0) * (1.0 / 4294967296.0);
  double u2 = static_cast<double>(xorshift32(h)) * (1.0 / 4294967296.0);
  double r = sqrt(-2.0 * log(u1));
  double th = 6.283185307179586 * u2;
  out[i] = static_cast<float>(r * cos(th));
}

unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

}  // namespace
}  // namespace csmi

using namespace csmi;

extern "C" {

cs_status cs_random(cs_stream *stream, float *out, int64_t n, uint32_t seed) {
  return guard([&] {
    if (!stream) fail(CS_INVALID_VALUE, "stream is null");
    if (n < 0 || n > 0xffffffffll) fail(CS_INVALID_VALUE, "length out of range");
    if (n == 0) return;
    if (!out) fail(CS_INVALID_VALUE, "out is null");
    set_device(stream->device);
    hipLaunchKernelGGL(random_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, stream->stream, out, n, seed);
    CS_HIP(hipGetLastError());
  });
}

cs_status cs_random_normal(cs_stream *stream, float *out, int64_t n, uint32_t seed) {
  return guard([&] {
    if (!stream) fail(CS_INVALID_VALUE, "stream is null");
    if (n < 0 || n > 0xffffffffll) fail(CS_INVALID_VALUE, "length out of range");
    if (n == 0) return;
    if (!out) fail(CS_INVALID_VALUE, "out is null");
    set_device(stream->device);
    int64_t pairs = (n + 1) / 2;
    hipLaunchKernelGGL(random_normal_kernel, dim3(blocks_for(pairs)), dim3(kBlock), 0, stream->stream, out,
                       pairs, seed);
    CS_HIP(hipGetLastError());
  });
}

cs_status cs_fill_uniform(cs_stream *stream, float *out, int64_t n, uint32_t seed, double lo, double hi) {
  return guard([&] {
    if (!stream) fail(CS_INVALID_VALUE, "stream is null");
    if (n < 0 || n > 0xffffffffll) fail(CS_INVALID_VALUE, "length out of range");
    if (n == 0) return;
    if (!out) fail(CS_INVALID_VALUE, "out is null");
    set_device(stream->device);
    hipLaunchKernelGGL(fill_uniform_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, stream->stream, out, n, seed,
                       lo, hi);
    CS_HIP(hipGetLastError());
  });
}

cs_status cs_fill_normal(cs_stream *stream, float *out, int64_t n, uint32_t seed) {
  return guard([&] {
    if (!stream) fail(CS_INVALID_VALUE, "stream is null");
    if (n < 0 || n > 0xffffffffll) fail(CS_INVALID_VALUE, "length out of range");
    if (n == 0) return;
    if (!out) fail(CS_INVALID_VALUE, "out is null");
    set_device(stream->device);
    hipLaunchKernelGGL(fill_normal_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, stream->stream, out, n, seed);
    CS_HIP(hipGetLastError());
  });
}

}  // extern "C"
