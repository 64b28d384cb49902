// On-device accuracy check of the elementwise math against FP64.
//
// For every FP32 bit pattern in [start, start + count) the kernel evaluates one function
// in FP32 and its FP64 counterpart rounded to FP32, and accumulates the largest distance in
// units in the last place, how many inputs exceed 1 ulp, and the worst input. FP64 is used
// here only as the oracle.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "internal.h"
#include "precise_math.h"

namespace csmi {
namespace {

__device__ __forceinline__ int64_t ordered(float f) {
  int32_t i = __float_as_int(f);
  return i < 0 ? int64_t(INT32_MIN) - i : int64_t(i);
}

// The functions each math mode emits (fused.cpp); the fast mode's are compiled here
// without fast-math flags, so they check the library functions the mode calls.
__device__ float eval32(int op, int mode, float x) {
  switch (op) {
    case CS_OP_EXP: return mode == CS_MATH_ACCURATE ? csmi_expf(x) : mode == CS_MATH_FAST ? __expf(x) : expf(x);
    case CS_CHECK_EXP_BOUNDED: return csmi_expf_np(x);
    case CS_OP_TANH: return mode == CS_MATH_ACCURATE ? csmi_tanhf(x) : csmi_tanhf_p(x);
    case CS_OP_LOG: return mode == CS_MATH_FAST ? __logf(x) : logf(x);
    case CS_OP_SQRT: return sqrtf(x);
    default: return __builtin_nanf("");
  }
}

__device__ double eval64(int op, double x) {
  switch (op) {
    case CS_OP_EXP:
    case CS_CHECK_EXP_BOUNDED: return exp(x);
    case CS_OP_TANH: return tanh(x);
    case CS_OP_LOG: return log(x);
    case CS_OP_SQRT: return sqrt(x);
    default: return __builtin_nan("");
  }
}

// stats[0] = max ulp, stats[1] = inputs over 1 ulp, stats[2] = worst input bits
__global__ __launch_bounds__(256) void math_check_kernel(int op, int mode, uint32_t start, uint32_t count,
                                                         unsigned long long *stats) {
  uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
  uint64_t stride = uint64_t(gridDim.x) * 256;
  unsigned long long worst = 0, worst_bits = 0, over = 0;
  for (; i < count; i += stride) {
    const uint32_t bits = start + static_cast<uint32_t>(i);
    const float x = __int_as_float(static_cast<int32_t>(bits));
    const float got = eval32(op, mode, x);
    const float want = static_cast<float>(eval64(op, static_cast<double>(x)));
    unsigned long long d;
    if (want != want || got != got) {
      d = (want != want && got != got) ? 0ull : (1ull << 40);
    } else {
      int64_t a = ordered(got), b = ordered(want);
      d = static_cast<unsigned long long>(a > b ? a - b : b - a);
    }
    if (d > 1) ++over;
    if (d > worst) {
      worst = d;
      worst_bits = bits;
    }
  }
  if (worst) {
    unsigned long long prev = atomicMax(&stats[0], worst);
    if (worst > prev) atomicExch(&stats[2], worst_bits);
  }
  if (over) atomicAdd(&stats[1], over);
}

}  // namespace
}  // namespace csmi

using namespace csmi;

extern "C" cs_status cs_math_check(cs_stream *stream, int op, int mode, uint32_t start, uint32_t count,
                                   unsigned long long *stats_device) {
  return guard([&] {
    if (!stream || !stats_device) fail(CS_INVALID_VALUE, "null argument");
    if (op != CS_OP_EXP && op != CS_OP_TANH && op != CS_OP_LOG && op != CS_OP_SQRT && op != CS_CHECK_EXP_BOUNDED)
      fail(CS_INVALID_VALUE, "unsupported op");
    if (mode != CS_MATH_PRECISE && mode != CS_MATH_FAST && mode != CS_MATH_ACCURATE)
      fail(CS_INVALID_VALUE, "unknown math mode");
    if (count == 0) return;
    set_device(stream->device);
    unsigned grid = static_cast<unsigned>(std::min<uint64_t>((uint64_t(count) + 255) / 256, 8192));
    hipLaunchKernelGGL(math_check_kernel, dim3(grid), dim3(256), 0, stream->stream, op, mode, start, count,
                       stats_device);
    CS_HIP(hipGetLastError());
  });
}
