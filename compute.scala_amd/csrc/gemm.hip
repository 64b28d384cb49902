// Matrix multiplication C = A . B (row-major FP32 in and out) on gfx950 MFMA.
//
// The reference has no GEMM kernel: matrixMultiply is a user idiom of split, broadcast,
// product and join that unrolls into one elementwise kernel (README.md:301-362,
// TensorsSpec.scala:468-528), so every output is a sequential FP32 dot product. Here:
//
//   CS_GEMM_FP32    FP32 operands on v_mfma_f32_32x32x2_f32 (FP32 products and sums)
//   CS_GEMM_BF16    operands rounded to BF16, v_mfma_f32_32x32x16_bf16, FP32 accumulate
//   CS_GEMM_BF16X3  split-BF16x3 (Henry, Tang & Heinecke 2019): A = Ah + Al, B = Bh + Bl
//                   with Ah = bf16(A), Al = bf16(A - Ah); C = Al.Bh + Ah.Bl + Ah.Bh, all
//                   three products on the BF16 MFMA into one FP32 accumulator
//   CS_GEMM_FP8     OCP e4m3 operands with one power-of-two scale per 128-wide K block of
//                   each row of A and each column of B (amax / 448 rounded up to 2^e),
//                   applied by v_mfma_scale_f32_32x32x64_f8f6f4, FP32 accumulate
//
// Structure: a pre-pass converts A to a K-contiguous operand X[Mp][Kp] and B to its
// transpose Y[Np][Kp] (zero padded to the tile sizes, split or quantised as the variant
// needs) in the stream's scratch; the main kernel computes C[m][n] = sum_k X[m][k] Y[n][k].
// It runs 256 threads (4 wave64s in 2x2) per 256x256 output tile; each wave owns 128x128
// as 4x4 MFMA tiles of 32x32 (256 FP32 accumulators per lane). K advances in stages staged
// through LDS (rows padded by 16 bytes against bank conflicts); the next stage's global
// loads are in flight while the MFMAs of the current one run.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.h"

namespace csmi {
namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

enum Kind { kF32 = 0, kBF16 = 1, kBF16X3 = 2, kFP8 = 3 };

template <int KIND>
struct Traits;
template <>
struct Traits<kF32> {
  static constexpr int kElem = 4, kStageK = 16, kInstrK = 2, kParts = 1;
};
template <>
struct Traits<kBF16> {
  static constexpr int kElem = 2, kStageK = 32, kInstrK = 16, kParts = 1;
};
template <>
struct Traits<kBF16X3> {
  static constexpr int kElem = 2, kStageK = 32, kInstrK = 16, kParts = 2;
};
template <>
struct Traits<kFP8> {
  static constexpr int kElem = 1, kStageK = 128, kInstrK = 64, kParts = 1;
};

constexpr int kTile = 256;     // output tile edge (rows and columns)
constexpr int kThreads = 256;  // 4 waves
constexpr int kWaveTiles = 4;  // 4x4 MFMA tiles of 32x32 per wave

// ---------------------------------------------------------------------------- pre-pass
__device__ __forceinline__ uint16_t bf16_rne(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return static_cast<uint16_t>((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// Converts rows of a row-major matrix (rows x cols, leading dimension ld) into a padded
// [rows_p][cols_p] operand. kTranspose reads the source as its transpose (B -> Y).
// parts == 2 writes the BF16 high part to out0 and the low part to out1.
template <int KIND, bool kTranspose>
__global__ __launch_bounds__(256) void convert_kernel(const float *__restrict__ src, int rows, int cols, int ld,
                                                      int rows_p, int cols_p, void *out0, void *out1) {
  __shared__ float tile[32][33];
  const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 32 x 8
  for (int i = ty; i < 32; i += 8) {
    int r = r0 + i, c = c0 + tx;
    float v = 0.0f;
    if (kTranspose) {  // output row r = source column r; output col c = source row c
      if (r < rows && c < cols) v = src[int64_t(c) * ld + r];
    } else {
      if (r < rows && c < cols) v = src[int64_t(r) * ld + c];
    }
    tile[i][tx] = v;
  }
  __syncthreads();
  for (int i = ty; i < 32; i += 8) {
    int r = r0 + i, c = c0 + tx;
    if (r >= rows_p || c >= cols_p) continue;
    float v = tile[i][tx];
    int64_t o = int64_t(r) * cols_p + c;
    if (KIND == kF32) {
      static_cast<float *>(out0)[o] = v;
    } else {
      uint16_t h = bf16_rne(v);
      static_cast<uint16_t *>(out0)[o] = h;
      if (KIND == kBF16X3) static_cast<uint16_t *>(out1)[o] = bf16_rne(v - bf16_to_f32(h));
    }
  }
}

// FP8 quantisation of a [rows][cols] FP32 matrix (already K-contiguous, zero padded) in
// blocks of 128 along each row: one wave per block, scale 2^e with 2^e >= amax / 448.
__global__ __launch_bounds__(256) void quantize_fp8_kernel(const float *__restrict__ src, int rows, int cols,
                                                           uint8_t *__restrict__ q, uint8_t *__restrict__ scales) {
  const int wave = (blockIdx.x * 256 + threadIdx.x) / 64;
  const int lane = threadIdx.x % 64;
  const int blocks_per_row = cols / 128;
  if (wave >= rows * blocks_per_row) return;
  const int r = wave / blocks_per_row, kb = wave % blocks_per_row;
  const float *p = src + int64_t(r) * cols + kb * 128;
  float x0 = p[lane], x1 = p[lane + 64];
  float m = fmaxf(fabsf(x0), fabsf(x1));
  for (int off = 32; off >= 1; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  // e = smallest integer with 448 * 2^e >= m (so |x| / 2^e <= 448); m = 0 -> e = 0
  int e = 0;
  if (m > 0.0f) {
    int ex;
    frexpf(m / 448.0f, &ex);  // m/448 = f * 2^ex, f in [0.5, 1)
    e = ex;
    if (ldexpf(1.0f, e - 1) >= m / 448.0f) e -= 1;
  }
  e = e < -127 ? -127 : (e > 127 ? 127 : e);
  const float inv = ldexpf(1.0f, -e);
  uint8_t *qp = q + int64_t(r) * cols + kb * 128;
  int packed0 = __builtin_amdgcn_cvt_pk_fp8_f32(x0 * inv, 0.0f, 0, false);
  int packed1 = __builtin_amdgcn_cvt_pk_fp8_f32(x1 * inv, 0.0f, 0, false);
  qp[lane] = static_cast<uint8_t>(packed0 & 0xff);
  qp[lane + 64] = static_cast<uint8_t>(packed1 & 0xff);
  if (lane == 0) scales[int64_t(r) * blocks_per_row + kb] = static_cast<uint8_t>(e + 127);
}

// ---------------------------------------------------------------------------- main kernel
template <int KIND>
struct Frag;
template <>
struct Frag<kF32> {
  float v;
};
template <>
struct Frag<kBF16> {
  bf16x8 v;
};
template <>
struct Frag<kBF16X3> {
  bf16x8 v;
};
template <>
struct Frag<kFP8> {
  i32x8 v;
};

template <int KIND>
__device__ __forceinline__ void mfma(const Frag<KIND> &a, const Frag<KIND> &b, f32x16 &c, int sa, int sb);
template <>
__device__ __forceinline__ void mfma<kF32>(const Frag<kF32> &a, const Frag<kF32> &b, f32x16 &c, int, int) {
  c = __builtin_amdgcn_mfma_f32_32x32x2f32(a.v, b.v, c, 0, 0, 0);
}
template <>
__device__ __forceinline__ void mfma<kBF16>(const Frag<kBF16> &a, const Frag<kBF16> &b, f32x16 &c, int, int) {
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.v, b.v, c, 0, 0, 0);
}
template <>
__device__ __forceinline__ void mfma<kBF16X3>(const Frag<kBF16X3> &a, const Frag<kBF16X3> &b, f32x16 &c, int,
                                              int) {
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.v, b.v, c, 0, 0, 0);
}
template <>
__device__ __forceinline__ void mfma<kFP8>(const Frag<kFP8> &a, const Frag<kFP8> &b, f32x16 &c, int sa, int sb) {
  // cbsz = blgp = 0: both operands OCP e4m3; one E8M0 scale per lane (byte 0)
  c = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a.v, b.v, c, 0, 0, 0, sa, 0, sb);
}

template <int KIND>
__device__ __forceinline__ Frag<KIND> lds_frag(const char *p);
template <>
__device__ __forceinline__ Frag<kF32> lds_frag<kF32>(const char *p) {
  return Frag<kF32>{*reinterpret_cast<const float *>(p)};
}
template <>
__device__ __forceinline__ Frag<kBF16> lds_frag<kBF16>(const char *p) {
  return Frag<kBF16>{__builtin_bit_cast(bf16x8, *reinterpret_cast<const i32x4 *>(p))};
}
template <>
__device__ __forceinline__ Frag<kBF16X3> lds_frag<kBF16X3>(const char *p) {
  return Frag<kBF16X3>{__builtin_bit_cast(bf16x8, *reinterpret_cast<const i32x4 *>(p))};
}
template <>
__device__ __forceinline__ Frag<kFP8> lds_frag<kFP8>(const char *p) {
  return Frag<kFP8>{*reinterpret_cast<const i32x8 *>(p)};
}

// X: [Mp][Kp] operand(s), Y: [Np][Kp] operand(s); parts == 2 adds the low halves X1, Y1.
// sx / sy: FP8 scales [Mp][Kp/128] and [Np][Kp/128].
template <int KIND>
__global__ __launch_bounds__(kThreads, 1) void gemm_nt_kernel(int M, int N, int Kp, const char *__restrict__ X0,
                                                             const char *__restrict__ X1, const char *__restrict__ Y0,
                                                             const char *__restrict__ Y1, const uint8_t *__restrict__ sx,
                                                             const uint8_t *__restrict__ sy, float *__restrict__ C,
                                                             int ldc) {
  using T = Traits<KIND>;
  constexpr int kRowBytes = T::kStageK * T::kElem;  // bytes of one operand row per stage
  constexpr int kStride = kRowBytes + 16;           // padded LDS row
  constexpr int kInstr = T::kStageK / T::kInstrK;   // MFMAs per tile per stage
  constexpr int kLaneBytes = kRowBytes / 2 / kInstr;
  constexpr int kChunks = kTile * kRowBytes / 16 / kThreads;  // 16-byte loads per thread per operand
  constexpr int kParts = T::kParts;
  extern __shared__ __attribute__((aligned(16))) char lds[];
  char *xs = lds;                                // [kParts][kTile][kStride]
  char *ys = lds + kParts * kTile * kStride;      // [kParts][kTile][kStride]

  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int wm = wave / 2, wn = wave % 2;
  const int half = lane / 32, l32 = lane % 32;
  // Tile order: consecutive blocks walk down a column of tiles so that blocks on the same
  // XCD (blockIdx % 8) share B panels in their L2.
  const int tiles_m = (M + kTile - 1) / kTile;
  const int m0 = (blockIdx.x % tiles_m) * kTile;
  const int n0 = (blockIdx.x / tiles_m) * kTile;
  const int64_t kpb = int64_t(Kp) * T::kElem;  // operand row bytes in global memory

  i32x4 rx[kParts][kChunks], ry[kParts][kChunks];
  auto load_stage = [&](int ks) {
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
      const int idx = threadIdx.x + c * kThreads;
      const int row = idx / (kRowBytes / 16), col = idx % (kRowBytes / 16);
      const int64_t go = int64_t(ks) * kRowBytes + col * 16;
      rx[0][c] = *reinterpret_cast<const i32x4 *>(X0 + (m0 + row) * kpb + go);
      ry[0][c] = *reinterpret_cast<const i32x4 *>(Y0 + (n0 + row) * kpb + go);
      if (kParts == 2) {
        rx[kParts - 1][c] = *reinterpret_cast<const i32x4 *>(X1 + (m0 + row) * kpb + go);
        ry[kParts - 1][c] = *reinterpret_cast<const i32x4 *>(Y1 + (n0 + row) * kpb + go);
      }
    }
  };
  auto store_stage = [&]() {
#pragma unroll
    for (int p = 0; p < kParts; ++p)
#pragma unroll
      for (int c = 0; c < kChunks; ++c) {
        const int idx = threadIdx.x + c * kThreads;
        const int row = idx / (kRowBytes / 16), col = idx % (kRowBytes / 16);
        *reinterpret_cast<i32x4 *>(xs + (p * kTile + row) * kStride + col * 16) = rx[p][c];
        *reinterpret_cast<i32x4 *>(ys + (p * kTile + row) * kStride + col * 16) = ry[p][c];
      }
  };

  f32x16 acc[kWaveTiles][kWaveTiles];
#pragma unroll
  for (int i = 0; i < kWaveTiles; ++i)
#pragma unroll
    for (int j = 0; j < kWaveTiles; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;

  const int stages = Kp / T::kStageK;
  load_stage(0);
  for (int ks = 0; ks < stages; ++ks) {
    store_stage();
    __syncthreads();
    if (ks + 1 < stages) load_stage(ks + 1);
    int sa[kWaveTiles], sb[kWaveTiles];
#pragma unroll
    for (int i = 0; i < kWaveTiles; ++i) {
      sa[i] = sb[i] = 127;
      if (KIND == kFP8) {
        const int kb = ks, kbs = Kp / 128;
        sa[i] = sx[int64_t(m0 + wm * 128 + i * 32 + l32) * kbs + kb];
        sb[i] = sy[int64_t(n0 + wn * 128 + i * 32 + l32) * kbs + kb];
      }
    }
#pragma unroll
    for (int s = 0; s < kInstr; ++s) {
      const int off = half * (kRowBytes / 2) + s * kLaneBytes;
      Frag<KIND> fa[kParts][kWaveTiles], fb[kParts][kWaveTiles];
#pragma unroll
      for (int p = 0; p < kParts; ++p)
#pragma unroll
        for (int i = 0; i < kWaveTiles; ++i) {
          fa[p][i] = lds_frag<KIND>(xs + (p * kTile + wm * 128 + i * 32 + l32) * kStride + off);
          fb[p][i] = lds_frag<KIND>(ys + (p * kTile + wn * 128 + i * 32 + l32) * kStride + off);
        }
#pragma unroll
      for (int i = 0; i < kWaveTiles; ++i)
#pragma unroll
        for (int j = 0; j < kWaveTiles; ++j) {
          if (kParts == 2) {  // small products first: Al.Bh + Ah.Bl, then Ah.Bh
            mfma<KIND>(fa[1][i], fb[0][j], acc[i][j], 0, 0);
            mfma<KIND>(fa[0][i], fb[1][j], acc[i][j], 0, 0);
          }
          mfma<KIND>(fa[0][i], fb[0][j], acc[i][j], sa[i], sb[j]);
        }
    }
    __syncthreads();
  }
  // acc[i][j][e]: row = 8*(e/4) + 4*half + e%4, col = l32 of the 32x32 tile (i, j)
#pragma unroll
  for (int i = 0; i < kWaveTiles; ++i)
#pragma unroll
    for (int j = 0; j < kWaveTiles; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int r = m0 + wm * 128 + i * 32 + 8 * (e / 4) + 4 * half + (e % 4);
        const int c = n0 + wn * 128 + j * 32 + l32;
        if (r < M && c < N) C[int64_t(r) * ldc + c] = acc[i][j][e];
      }
}

// LDS of one stage: the X and Y panels (kParts each) of kTile padded rows.
template <int KIND>
size_t lds_bytes() {
  using T = Traits<KIND>;
  return size_t(2) * T::kParts * kTile * (T::kStageK * T::kElem + 16);
}

int64_t round_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

template <int KIND>
void run_gemm(cs_stream *stream, int M, int N, int K, const float *A, int lda, const float *B, int ldb, float *C,
              int ldc) {
  using T = Traits<KIND>;
  const int64_t Mp = round_up(M, kTile), Np = round_up(N, kTile), Kp = round_up(K, T::kStageK);
  const int64_t xbytes = Mp * Kp * T::kElem, ybytes = Np * Kp * T::kElem;
  const int64_t parts = T::kParts;
  int64_t total = parts * (xbytes + ybytes) + 4096;
  int64_t f32_stage = 0;
  if (KIND == kFP8) {
    f32_stage = (Mp > Np ? Mp : Np) * Kp * 4;  // FP32 staging for the padded, K-contiguous copy
    total += f32_stage + (Mp + Np) * (Kp / 128) + 4096;
  }
  char *base = static_cast<char *>(stream_scratch(stream, static_cast<size_t>(total)));
  char *x0 = base, *y0 = x0 + xbytes;
  char *x1 = parts == 2 ? y0 + ybytes : nullptr;
  char *y1 = parts == 2 ? x1 + xbytes : nullptr;
  uint8_t *sx = nullptr, *sy = nullptr;
  hipStream_t s = stream->stream;
  if (KIND == kFP8) {
    float *stage = reinterpret_cast<float *>(base + xbytes + ybytes);
    sx = reinterpret_cast<uint8_t *>(base + xbytes + ybytes + f32_stage);
    sy = sx + Mp * (Kp / 128);
    dim3 ga((Kp + 31) / 32, (Mp + 31) / 32);
    hipLaunchKernelGGL((convert_kernel<kF32, false>), ga, dim3(256), 0, s, A, M, K, lda, int(Mp), int(Kp), stage,
                       nullptr);
    hipLaunchKernelGGL(quantize_fp8_kernel, dim3(unsigned((Mp * (Kp / 128) + 3) / 4)), dim3(256), 0, s, stage,
                       int(Mp), int(Kp), reinterpret_cast<uint8_t *>(x0), sx);
    dim3 gb((Kp + 31) / 32, (Np + 31) / 32);
    hipLaunchKernelGGL((convert_kernel<kF32, true>), gb, dim3(256), 0, s, B, N, K, ldb, int(Np), int(Kp), stage,
                       nullptr);
    hipLaunchKernelGGL(quantize_fp8_kernel, dim3(unsigned((Np * (Kp / 128) + 3) / 4)), dim3(256), 0, s, stage,
                       int(Np), int(Kp), reinterpret_cast<uint8_t *>(y0), sy);
  } else {
    dim3 ga((Kp + 31) / 32, (Mp + 31) / 32);
    hipLaunchKernelGGL((convert_kernel<KIND, false>), ga, dim3(256), 0, s, A, M, K, lda, int(Mp), int(Kp), x0, x1);
    dim3 gb((Kp + 31) / 32, (Np + 31) / 32);
    hipLaunchKernelGGL((convert_kernel<KIND, true>), gb, dim3(256), 0, s, B, N, K, ldb, int(Np), int(Kp), y0, y1);
  }
  const size_t lds = lds_bytes<KIND>();
  static bool configured = false;
  if (!configured) {
    CS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&gemm_nt_kernel<KIND>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    configured = true;
  }
  const unsigned grid = static_cast<unsigned>((Mp / kTile) * (Np / kTile));
  hipLaunchKernelGGL(gemm_nt_kernel<KIND>, dim3(grid), dim3(kThreads), lds, s, M, N, int(Kp), x0, x1, y0, y1, sx,
                     sy, C, ldc);
  CS_HIP(hipGetLastError());
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
    if (variant < CS_GEMM_FP32 || variant > CS_GEMM_FP8)
      fail(CS_INVALID_VALUE, "unknown GEMM variant " + std::to_string(variant));
    if (M == 0 || N == 0) return;
    if (!A || !B || !C) fail(CS_INVALID_VALUE, "null matrix pointer");
    if (int64_t(M) * K >= (int64_t(1) << 31) || int64_t(N) * K >= (int64_t(1) << 31))
      fail(CS_INVALID_VALUE, "operands are limited to 2^31 elements");
    set_device(stream->device);
    const int Ke = K == 0 ? 1 : K;  // K = 0: a zero-padded stage gives C = 0
    switch (variant) {
      case CS_GEMM_FP32: run_gemm<kF32>(stream, M, N, Ke, A, lda, B, ldb, C, ldc); break;
      case CS_GEMM_BF16: run_gemm<kBF16>(stream, M, N, Ke, A, lda, B, ldb, C, ldc); break;
      case CS_GEMM_BF16X3: run_gemm<kBF16X3>(stream, M, N, Ke, A, lda, B, ldb, C, ldc); break;
      case CS_GEMM_FP8: run_gemm<kFP8>(stream, M, N, Ke, A, lda, B, ldb, C, ldc); break;
    }
  });
}
