// Matrix multiplication C = A . B (row-major FP32 in and out).
//
// The reference has no GEMM kernel: matrixMultiply is a user idiom of split, broadcast,
// product and join that unrolls into one elementwise kernel (README.md:301-362,
// TensorsSpec.scala:468-528). This file is the MFMA replacement for that idiom.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.h"

namespace csmi {
namespace {

// FP32 SIMT tile: 64x64 outputs per 256-thread workgroup, 4x4 per thread, K in steps of 16
// staged through LDS. Products and sums stay FP32 (fused multiply-add).
constexpr int kT = 64, kK = 16;

__global__ __launch_bounds__(256) void gemm_fp32_simt(int M, int N, int K, const float *__restrict__ A, int lda,
                                                      const float *__restrict__ B, int ldb, float *__restrict__ C,
                                                      int ldc) {
  __shared__ float As[kK][kT + 4];
  __shared__ float Bs[kK][kT + 4];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int m0 = blockIdx.y * kT, n0 = blockIdx.x * kT;
  float acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += kK) {
    for (int e = threadIdx.x; e < kT * kK; e += 256) {
      int r = e / kK, c = e % kK;  // A tile: 64 rows x 16 k
      int gm = m0 + r, gk = k0 + c;
      As[c][r] = (gm < M && gk < K) ? A[int64_t(gm) * lda + gk] : 0.0f;
      int rb = e / kT, cb = e % kT;  // B tile: 16 k x 64 cols
      int gkb = k0 + rb, gn = n0 + cb;
      Bs[rb][cb] = (gkb < K && gn < N) ? B[int64_t(gkb) * ldb + gn] : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kK; ++k) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[k][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[k][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      int gm = m0 + ty * 4 + i, gn = n0 + tx * 4 + j;
      if (gm < M && gn < N) C[int64_t(gm) * ldc + gn] = acc[i][j];
    }
}

}  // namespace
}  // namespace csmi

using namespace csmi;

extern "C" cs_status cs_gemm(cs_stream *stream, int variant, int M, int N, int K, const float *A, int lda,
                             const float *B, int ldb, float *C, int ldc) {
  return guard([&] {
    if (!stream) fail(CS_INVALID_VALUE, "stream is null");
    if (M < 0 || N < 0 || K < 0) fail(CS_INVALID_VALUE, "negative dimension");
    if (lda < K || ldb < N || ldc < N) fail(CS_INVALID_VALUE, "leading dimension too small");
    if (M == 0 || N == 0) return;
    if (!A || !B || !C) fail(CS_INVALID_VALUE, "null matrix pointer");
    set_device(stream->device);
    switch (variant) {
      case CS_GEMM_FP32: {
        dim3 grid((N + kT - 1) / kT, (M + kT - 1) / kT);
        hipLaunchKernelGGL(gemm_fp32_simt, grid, dim3(256), 0, stream->stream, M, N, K, A, lda, B, ldb, C, ldc);
        CS_HIP(hipGetLastError());
        break;
      }
      default:
        fail(CS_INVALID_VALUE, "unknown GEMM variant " + std::to_string(variant));
    }
  });
}
