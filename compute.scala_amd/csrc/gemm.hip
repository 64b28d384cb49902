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
// A pre-pass converts A to a K-contiguous operand X[Mp][.] and B to its transpose Y[Np][.]
// (zero padded to the tile sizes; split, interleaved or quantised as the variant needs) in
// the stream's scratch; the main kernel computes C[m][n] = sum_k X[m][k] Y[n][k] with 256
// threads (4 wave64s in 2x2) per 256x256 output tile, each wave owning 128x128 as 4x4 MFMA
// tiles of 32x32 (256 FP32 accumulators per lane).
//
// All variants share one pipeline. A K stage is 128 bytes of every operand row (32 FP32
// values, 64 BF16 values, 32 hi/lo pairs or 128 e4m3 values); two stages live in LDS (128 KiB).
// The next stage streams global -> LDS directly (buffer_load ... lds, no registers) while
// the MFMAs of the current one run; inside a stage the fragments of the next MFMA step are
// read from LDS while the current step's MFMAs issue. Each 128-byte LDS row R stores its
// 16-byte piece p at slot p ^ ((R >> 1) & 7), so the fragment reads of 16 consecutive rows
// (the lanes the LDS serves in one cycle) hit 16 different bank groups.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "internal.h"

namespace csmi {
namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void;

enum Kind { kF32 = 0, kBF16 = 1, kBF16X3 = 2, kFP8 = 3 };

constexpr int kTile = 256;     // output tile edge (rows and columns)
constexpr int kThreads = 256;  // 4 waves
constexpr int kWaveTiles = 4;  // 4x4 MFMA tiles of 32x32 per wave

// Blocks are dispatched round-robin over the 8 XCDs, each with its own L2. The tile of
// block b: the XCDs get contiguous ranges of tiles, walked in groups of 4 tile rows (column
// by column inside a group), so one XCD's blocks share 4 A panels and a run of B panels.
__device__ __forceinline__ void tile_coords(int b, int tiles_m, int tiles_n, int &tm, int &tn) {
  const int total = tiles_m * tiles_n;
  int t = b;
  if (total % 8 == 0) t = (b % 8) * (total / 8) + b / 8;
  constexpr int G = 4;
  const int per_group = G * tiles_n;
  const int g = t / per_group, local = t % per_group;
  const int rows = min(G, tiles_m - g * G);
  tm = g * G + local % rows;
  tn = local / rows;
}

// ---------------------------------------------------------------------------- pre-pass

// FP32 operands for the FP32 kernel, zero padded to [rows_p][kp]. A (K contiguous): one
// thread per 4 consecutive K values of a row, 16-byte loads when the source is 16-byte
// aligned (`vec`) and 16-byte stores.
__device__ __forceinline__ void pack_rows_f32(unsigned bx, const float *__restrict__ src, int rows, int cols, int ld,
                                              int rows_p, int kp, int vec, float *__restrict__ out) {
  const int chunks = kp / 4;
  const int64_t idx = int64_t(bx) * 256 + threadIdx.x;
  if (idx >= int64_t(rows_p) * chunks) return;
  const int r = static_cast<int>(idx / chunks), k = static_cast<int>(idx % chunks) * 4;
  const float *p = src + int64_t(r) * ld + k;
  float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (vec && r < rows && k + 4 <= cols) {
    v = *reinterpret_cast<const float4 *>(p);
  } else if (r < rows) {
    v.x = k < cols ? p[0] : 0.0f;
    v.y = k + 1 < cols ? p[1] : 0.0f;
    v.z = k + 2 < cols ? p[2] : 0.0f;
    v.w = k + 3 < cols ? p[3] : 0.0f;
  }
  *reinterpret_cast<float4 *>(out + int64_t(r) * kp + k) = v;
}

// A TK x TN tile of B (K x N, N contiguous): rows k0.., columns n0.., zero outside the source.
// TN / 4 threads per row, one 16-byte load each (`vec`), so a row's 4 TN bytes are read
// contiguously. Column n of row k sits at (n + k) mod TN: the reads of one column down K
// (lane l at row 4l + j or 8l + e) then spread over the 64 banks without padding.
constexpr int kPackTN = 64;  // B-tile width of the FP32 and BF16 packs
template <int TN>
__device__ __forceinline__ int tile_col(int k, int n) { return (n + k) & (TN - 1); }

template <int TK, int TN, int kThreadsT>
__device__ __forceinline__ void load_b_tile(float (*tile)[TN], const float *__restrict__ src, int n_src, int k_src,
                                            int ld, int n0, int k0, int vec) {
  constexpr int kRowsPerPass = kThreadsT / (TN / 4), kPasses = TK / kRowsPerPass;
  const int t = threadIdx.x, n = (t % (TN / 4)) * 4;
  if (vec && n0 + TN <= n_src && k0 + TK <= k_src) {
    // the tile lies inside the source: all loads in flight before the first LDS write
    float4 a[kPasses];
#pragma unroll
    for (int i = 0; i < kPasses; ++i)
      a[i] = *reinterpret_cast<const float4 *>(src + int64_t(k0 + t / (TN / 4) + kRowsPerPass * i) * ld + (n0 + n));
#pragma unroll
    for (int i = 0; i < kPasses; ++i) {
      const int k = t / (TN / 4) + kRowsPerPass * i;
      tile[k][tile_col<TN>(k, n)] = a[i].x;
      tile[k][tile_col<TN>(k, n + 1)] = a[i].y;
      tile[k][tile_col<TN>(k, n + 2)] = a[i].z;
      tile[k][tile_col<TN>(k, n + 3)] = a[i].w;
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < kPasses; ++i) {
    const int k = t / (TN / 4) + kRowsPerPass * i;
    const float *p = src + int64_t(k0 + k) * ld + (n0 + n);
    float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (k0 + k < k_src) {
      if (vec && n0 + n + 4 <= n_src) {
        a = *reinterpret_cast<const float4 *>(p);
      } else {
        a.x = n0 + n < n_src ? p[0] : 0.0f;
        a.y = n0 + n + 1 < n_src ? p[1] : 0.0f;
        a.z = n0 + n + 2 < n_src ? p[2] : 0.0f;
        a.w = n0 + n + 3 < n_src ? p[3] : 0.0f;
      }
    }
    tile[k][tile_col<TN>(k, n)] = a.x;
    tile[k][tile_col<TN>(k, n + 1)] = a.y;
    tile[k][tile_col<TN>(k, n + 2)] = a.z;
    tile[k][tile_col<TN>(k, n + 3)] = a.w;
  }
}

// B (K x N, N contiguous) -> Y[rows_p][kp], row n = column n of B: 64 K x TN N tiles through
// LDS, 16-byte stores along K (each row of Y gets 256 contiguous bytes per tile).
template <int TN>
__device__ __forceinline__ void pack_cols_f32(unsigned bx, unsigned by, const float *__restrict__ src, int n_src,
                                              int k_src, int ld, int rows_p, int kp, int vec,
                                              float *__restrict__ out) {
  __shared__ float tile[64][TN];  // [k][rotated n]
  const int n0 = int(bx) * TN, k0 = int(by) * 64, t = threadIdx.x;
  load_b_tile<64, TN, 256>(tile, src, n_src, k_src, ld, n0, k0, vec);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < TN / 16; ++i) {
    const int n = t / 16 + 16 * i, c = t % 16, k = k0 + 4 * c;
    if (n0 + n >= rows_p || k >= kp) continue;
    const float4 v =
        make_float4(tile[4 * c][tile_col<TN>(4 * c, n)], tile[4 * c + 1][tile_col<TN>(4 * c + 1, n)],
                    tile[4 * c + 2][tile_col<TN>(4 * c + 2, n)], tile[4 * c + 3][tile_col<TN>(4 * c + 3, n)]);
    *reinterpret_cast<float4 *>(out + int64_t(n0 + n) * kp + k) = v;
  }
}

// Operand rows for the LDS pipeline. BF16: Kp values per row. Split (BF16X3): per 32-wide K
// block, 32 BF16 high parts then 32 BF16 low parts (128 bytes), so that one K stage of a row
// is 128 contiguous bytes in every variant.
// Two floats rounded to BF16 (nearest even) by v_cvt_pk_bf16_f32: a in the low half.
__device__ __forceinline__ uint32_t bf16x2(float a, float b) {
  typedef float f32x2 __attribute__((ext_vector_type(2)));
  typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, bf16x2_t));
}

__device__ __forceinline__ void store_bf16x8(const float (&v)[8], char *row, int k, bool split) {
  i32x4 hi, lo;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const uint32_t h = bf16x2(v[2 * e], v[2 * e + 1]);
    hi[e] = int(h);
    // the low parts: each value minus its high part (exact), rounded once more
    lo[e] = int(bf16x2(v[2 * e] - __uint_as_float(h << 16), v[2 * e + 1] - __uint_as_float(h & 0xffff0000u)));
  }
  if (split) {
    char *p = row + (k / 32) * 128 + (k % 32) * 2;
    *reinterpret_cast<i32x4 *>(p) = hi;
    *reinterpret_cast<i32x4 *>(p + 64) = lo;
  } else {
    *reinterpret_cast<i32x4 *>(row + int64_t(k) * 2) = hi;
  }
}

// A (rows x cols, K contiguous) -> X: one thread per 8 consecutive K values of a row, two
// 16-byte loads when the source is 16-byte aligned (`vec`), and one (split: two) 16-byte
// stores. Rows and K values beyond the source are zero.
template <bool kSplit>
__device__ __forceinline__ void pack_rows_bf16(unsigned bx, const float *__restrict__ src, int rows, int cols, int ld,
                                               int rows_p, int kp, int vec, char *__restrict__ out) {
  const int chunks = kp / 8;
  const int64_t idx = int64_t(bx) * 256 + threadIdx.x;
  if (idx >= int64_t(rows_p) * chunks) return;
  const int r = static_cast<int>(idx / chunks), k = static_cast<int>(idx % chunks) * 8;
  float v[8];
  const float *p = src + int64_t(r) * ld + k;
  if (vec && r < rows && k + 8 <= cols) {
    const float4 a = *reinterpret_cast<const float4 *>(p), b = *reinterpret_cast<const float4 *>(p + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = (r < rows && k + e < cols) ? p[e] : 0.0f;
  }
  store_bf16x8(v, out + int64_t(r) * (int64_t(kp) * (kSplit ? 4 : 2)), k, kSplit);
}

// B (K x N, N contiguous) -> Y (row n holds column n of B along K): 64 K x TN N tiles
// through LDS; stores are 16 bytes along K (split: 16 high and 16 low bytes).
template <bool kSplit, int TN>
__device__ __forceinline__ void pack_cols_bf16(unsigned bx, unsigned by, const float *__restrict__ src, int n_src,
                                               int k_src, int ld, int rows_p, int kp, int vec,
                                               char *__restrict__ out) {
  __shared__ float tile[64][TN];  // [k][rotated n]
  const int n0 = int(bx) * TN, k0 = int(by) * 64, t = threadIdx.x;
  load_b_tile<64, TN, 256>(tile, src, n_src, k_src, ld, n0, k0, vec);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < TN / 32; ++i) {
    const int n = t / 8 + 32 * i, c = t % 8, k = k0 + 8 * c;
    if (n0 + n >= rows_p || k >= kp) continue;
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = tile[8 * c + e][tile_col<TN>(8 * c + e, n)];
    store_bf16x8(v, out + int64_t(n0 + n) * (int64_t(kp) * (kSplit ? 4 : 2)), k, kSplit);
  }
}

// FP8 quantisation of 128 values: the scale is 2^e with e the smallest integer such that
// 448 * 2^e >= amax (|x| / 2^e <= 448, the e4m3 maximum); amax = 0 gives e = 0. Returns the
// E8M0 scale byte e + 127.
__device__ __forceinline__ int fp8_scale_exponent(float amax) {
  int e = 0;
  if (amax > 0.0f) {
    int ex;
    frexpf(amax / 448.0f, &ex);  // amax/448 = f * 2^ex, f in [0.5, 1)
    e = ex;
    if (ldexpf(1.0f, e - 1) >= amax / 448.0f) e -= 1;
  }
  return e < -127 ? -127 : (e > 127 ? 127 : e);
}

// Four e4m3 bytes of x[0..3] * inv, packed into one word (byte i = value i).
__device__ __forceinline__ uint32_t fp8x4(float x0, float x1, float x2, float x3, float inv) {
  int w = __builtin_amdgcn_cvt_pk_fp8_f32(x0 * inv, x1 * inv, 0, false);
  w = __builtin_amdgcn_cvt_pk_fp8_f32(x2 * inv, x3 * inv, w, true);
  return static_cast<uint32_t>(w);
}

// A (rows x cols, K contiguous) -> X[rows_p][kp] e4m3 plus scales[kb * rows_p + r]: 16 lanes
// per 128-wide block of a row, 8 consecutive values per lane (two 16-byte loads when the
// source is 16-byte aligned, `vec`), one 8-byte store per lane.
__device__ __forceinline__ void quantize_rows_fp8(unsigned bx, const float *__restrict__ src, int rows, int cols,
                                                  int ld, int rows_p, int kp, int vec, uint8_t *__restrict__ q,
                                                  uint8_t *__restrict__ scales) {
  const int64_t group = (int64_t(bx) * 256 + threadIdx.x) / 16;
  const int l16 = threadIdx.x % 16;
  const int blocks = kp / 128;
  if (group >= int64_t(rows_p) * blocks) return;  // whole groups of 16 lanes leave together
  const int r = static_cast<int>(group / blocks), kb = static_cast<int>(group % blocks);
  const int k = kb * 128 + 8 * l16;
  const float *p = src + int64_t(r) * ld + k;
  float v[8];
  if (vec && r < rows && k + 8 <= cols) {
    const float4 a = *reinterpret_cast<const float4 *>(p), b = *reinterpret_cast<const float4 *>(p + 4);
    v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = (r < rows && k + e < cols) ? p[e] : 0.0f;
  }
  float m = 0.0f;
#pragma unroll
  for (int e = 0; e < 8; ++e) m = fmaxf(m, fabsf(v[e]));
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 16));
  const int e = fp8_scale_exponent(m);
  const float inv = ldexpf(1.0f, -e);
  const uint2 w = make_uint2(fp8x4(v[0], v[1], v[2], v[3], inv), fp8x4(v[4], v[5], v[6], v[7], inv));
  *reinterpret_cast<uint2 *>(q + int64_t(r) * kp + k) = w;
  if (l16 == 0) scales[int64_t(kb) * rows_p + r] = static_cast<uint8_t>(e + 127);
}

// B (K x N, N contiguous) -> Y[rows_p][kp] e4m3 (row n = column n of B) plus
// scales[kb * rows_p + n]: a block stages 128 K x 64 N in LDS (16-byte loads along N when
// `vec`, all in flight when the tile lies inside the source), then 4 lanes quantise one
// column's 128-wide block, 32 consecutive K values per lane and two 16-byte stores each (the
// column's 128 bytes are contiguous in Y).
__device__ __forceinline__ void quantize_cols_fp8(unsigned bx, unsigned by, const float *__restrict__ src, int n_src,
                                                  int k_src, int ld, int rows_p, int kp, int vec,
                                                  uint8_t *__restrict__ q_out, uint8_t *__restrict__ scales) {
  __shared__ float tile[128][65];  // [k][n]; row stride 65 words: read j of lane (n, q) hits bank 32q + j + n
  const int n0 = int(bx) * 64, kb = int(by), k0 = kb * 128, t = threadIdx.x;
  const int c4 = (t % 16) * 4;  // first of the four columns this thread loads
  if (vec && n0 + 64 <= n_src && k0 + 128 <= k_src) {
    float4 a[8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
      a[i] = *reinterpret_cast<const float4 *>(src + int64_t(k0 + t / 16 + 16 * i) * ld + (n0 + c4));
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int k = t / 16 + 16 * i;
      tile[k][c4] = a[i].x;
      tile[k][c4 + 1] = a[i].y;
      tile[k][c4 + 2] = a[i].z;
      tile[k][c4 + 3] = a[i].w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int k = t / 16 + 16 * i;
      const float *p = src + int64_t(k0 + k) * ld + (n0 + c4);
      float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (k0 + k < k_src) {
        if (vec && n0 + c4 + 4 <= n_src) {
          a = *reinterpret_cast<const float4 *>(p);
        } else {
          a.x = n0 + c4 < n_src ? p[0] : 0.0f;
          a.y = n0 + c4 + 1 < n_src ? p[1] : 0.0f;
          a.z = n0 + c4 + 2 < n_src ? p[2] : 0.0f;
          a.w = n0 + c4 + 3 < n_src ? p[3] : 0.0f;
        }
      }
      tile[k][c4] = a.x;
      tile[k][c4 + 1] = a.y;
      tile[k][c4 + 2] = a.z;
      tile[k][c4 + 3] = a.w;
    }
  }
  __syncthreads();
  // 4 lanes per column, 32 consecutive K values each: lane q of column n quantises K rows
  // 32q..32q+31 of the tile and stores them as 32 contiguous bytes of Y's row n.
  const int n = t / 4, q = t % 4;
  if (n0 + n < rows_p) {  // the 4 lanes of a column agree
    float x[32];
    float m = 0.0f;
#pragma unroll
    for (int j = 0; j < 32; ++j) {
      x[j] = tile[32 * q + j][n];
      m = fmaxf(m, fabsf(x[j]));
    }
    m = fmaxf(m, __shfl_xor(m, 1, 4));
    m = fmaxf(m, __shfl_xor(m, 2, 4));
    const int e = fp8_scale_exponent(m);
    const float inv = ldexpf(1.0f, -e);
    uint32_t w[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) w[j] = fp8x4(x[4 * j], x[4 * j + 1], x[4 * j + 2], x[4 * j + 3], inv);
    uint4 *dst = reinterpret_cast<uint4 *>(q_out + int64_t(n0 + n) * kp + k0 + 32 * q);
    dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
    dst[1] = make_uint4(w[4], w[5], w[6], w[7]);
    if (q == 0) scales[int64_t(kb) * rows_p + n0 + n] = static_cast<uint8_t>(e + 127);
  }
}

// Both operands in one launch: workgroups [0, row_blocks) pack A's rows, the rest B's 64-column
// tiles (N fastest), so the tail of one operand's pass overlaps the start of the other's and
// the call pays one dispatch instead of two.
struct PackArgs {
  const float *a;
  int m, k, lda, mp, kp, vec_a;
  void *x;
  unsigned row_blocks;
  const float *b;
  int n, ldb, np, vec_b;
  void *y;
  unsigned cols_gx;
  uint8_t *sx, *sy;  // FP8: the scales of A's rows and B's columns
};

template <int KIND>
__global__ __launch_bounds__(256) void pack_operands_kernel(PackArgs p) {
  const unsigned b = blockIdx.x;
  if (b < p.row_blocks) {
    if constexpr (KIND == kFP8)
      quantize_rows_fp8(b, p.a, p.m, p.k, p.lda, p.mp, p.kp, p.vec_a, static_cast<uint8_t *>(p.x), p.sx);
    else if constexpr (KIND == kF32)
      pack_rows_f32(b, p.a, p.m, p.k, p.lda, p.mp, p.kp, p.vec_a, static_cast<float *>(p.x));
    else
      pack_rows_bf16<KIND == kBF16X3>(b, p.a, p.m, p.k, p.lda, p.mp, p.kp, p.vec_a, static_cast<char *>(p.x));
    return;
  }
  const unsigned c = b - p.row_blocks;
  if constexpr (KIND == kFP8)
    quantize_cols_fp8(c % p.cols_gx, c / p.cols_gx, p.b, p.n, p.k, p.ldb, p.np, p.kp, p.vec_b,
                      static_cast<uint8_t *>(p.y), p.sy);
  else if constexpr (KIND == kF32)
    pack_cols_f32<kPackTN>(c % p.cols_gx, c / p.cols_gx, p.b, p.n, p.k, p.ldb, p.np, p.kp, p.vec_b,
                           static_cast<float *>(p.y));
  else
    pack_cols_bf16<KIND == kBF16X3, kPackTN>(c % p.cols_gx, c / p.cols_gx, p.b, p.n, p.k, p.ldb, p.np, p.kp,
                                              p.vec_b, static_cast<char *>(p.y));
}

// ------------------------------------------------------------------------ main kernel
constexpr int kRowB = 128;                          // bytes of one operand row per K stage
constexpr int kPanelB = kTile * kRowB;              // 32 KiB: one operand's rows of a stage
constexpr int kStageB = 2 * kPanelB;                // X and Y panels
constexpr size_t kLdsBytes = size_t(2) * kStageB;   // two stages in flight: 128 KiB

template <int KIND>
struct Pipe;
template <>
struct Pipe<kF32> {
  // 4 fragment steps of 8 K values; each fragment (4 FP32 values per lane) feeds 4 MFMAs of K 2
  static constexpr int kBytesPerK = 4, kK = 32, kSteps = 4;
  typedef i32x4 Frag;
};
template <>
struct Pipe<kBF16> {
  static constexpr int kBytesPerK = 2, kK = 64, kSteps = 4;  // 4 MFMA K steps of 16
  typedef bf16x8 Frag;
};
template <>
struct Pipe<kBF16X3> {
  static constexpr int kBytesPerK = 4, kK = 32, kSteps = 2;  // hi and lo per K value
  typedef bf16x8 Frag;
};
template <>
struct Pipe<kFP8> {
  static constexpr int kBytesPerK = 1, kK = 128, kSteps = 2;  // 2 MFMA K steps of 64
  typedef i32x8 Frag;
};

#define CSMI_SCHED_FENCE() __builtin_amdgcn_sched_barrier(0)

template <int KIND>
__device__ __forceinline__ void mma(const typename Pipe<KIND>::Frag &a, const typename Pipe<KIND>::Frag &b,
                                    f32x16 &c, int sa, int sb) {
  if constexpr (KIND == kFP8) {
    // cbsz = blgp = 0: both operands OCP e4m3; one E8M0 scale per lane (byte 0)
    c = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 0, 0, 0, sa, 0, sb);
  } else {
    (void)sa;
    (void)sb;
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
}

// Runs stage(ks, more) for ks = 0 .. stages - 1 with `more` (a next stage exists) as a
// compile-time constant.
template <class Stage>
__device__ __forceinline__ void run_stages(int stages, Stage &stage) {
  for (int ks = 0; ks + 1 < stages; ++ks) stage(ks, std::true_type());
  stage(stages - 1, std::false_type());
}

// X: [Mp] rows of Kp * kBytesPerK bytes, Y: [Np] rows likewise. sx / sy: FP8 scales
// [Kp/128][Mp] and [Kp/128][Np]. Host guarantees 256 * row bytes < 2^31.
template <int KIND>
__global__ __launch_bounds__(kThreads, 1) void gemm_lds_kernel(int M, int N, int Kp, int Mp, int Np,
                                                               const char *__restrict__ X,
                                                               const char *__restrict__ Y,
                                                               const uint8_t *__restrict__ sx,
                                                               const uint8_t *__restrict__ sy,
                                                               float *__restrict__ C, int ldc) {
  using P = Pipe<KIND>;
  using F = typename P::Frag;
  extern __shared__ __attribute__((aligned(1024))) char lds[];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int wm = wave >> 1, wn = wave & 1, half = lane >> 5, l32 = lane & 31;
  int tm, tn;
  tile_coords(blockIdx.x, Mp / kTile, Np / kTile, tm, tn);
  const int m0 = tm * kTile, n0 = tn * kTile;
  const int rowb = Kp * P::kBytesPerK;
  const int stages = Kp / P::kK;

  // Global -> LDS. One DMA instruction of a wave fills a 1 KiB chunk = 8 rows of a panel,
  // lane l writing LDS bytes 16 l .. 16 l + 15: slot l & 7 of row l >> 3 of the chunk. Row R
  // of a panel stores its piece p at slot p ^ ((R >> 1) & 7). DMA d (0..15) of a stage:
  // operand d & 1, chunk c = 4 (d >> 1) + wave, i.e. rows R = 8 c + (l >> 3), whose key
  // (R >> 1) & 7 = 4 (wave & 1) + (l >> 4) is the same for all 8 DMAs of a lane.
  const int drow = lane >> 3;
  const int dkey = 4 * (wave & 1) + (drow >> 1);
  const unsigned voff = unsigned((drow + 8 * wave) * rowb + (((lane & 7) ^ dkey) << 4));
  const __amdgpu_buffer_rsrc_t rx =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<char *>(X) + int64_t(m0) * rowb, 0, kTile * rowb, 0x00020000);
  const __amdgpu_buffer_rsrc_t ry =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<char *>(Y) + int64_t(n0) * rowb, 0, kTile * rowb, 0x00020000);
  auto dma = [&](int d, int ks, int buf) {
    const int j = d >> 1;
    const unsigned soff = unsigned(32 * j * rowb + ks * kRowB);
    char *l = lds + buf * kStageB + (d & 1) * kPanelB + (4 * j + wave) * 1024;
    __builtin_amdgcn_raw_ptr_buffer_load_lds((d & 1) ? ry : rx, (lds_void *)(l), 16, voff, soff, 0, 0);
  };

  // Fragment reads: lane (half, l32) of tile i reads row R = wm*128 + 32 i + l32 (A) or
  // wn*128 + 32 i + l32 (B); piece p sits at slot p ^ ((R >> 1) & 7) = p ^ ((l32 >> 1) & 7).
  // The LDS serves 16 lanes per cycle over 64 banks (two 128-byte rows): lanes l32 .. l32 + 15
  // have 16 different (R & 1, (R >> 1) & 7), so they hit 16 different bank groups.
  const int sw = (l32 >> 1) & 7;
  const int arow = (wm * 128 + l32) * kRowB, brow = kPanelB + (wn * 128 + l32) * kRowB;
  auto piece = [&](int buf, int row, int p) -> i32x4 {
    return *reinterpret_cast<const i32x4 *>(lds + buf * kStageB + row + ((p ^ sw) << 4));
  };
  // BF16 step s uses piece 2 s + half (K 16 s + 8 half .. + 7, the MFMA's own K order);
  // BF16X3 step s: high part piece 2 s + half, low part 4 + 2 s + half; FP8 step s: pieces
  // 4 s + 2 half and 4 s + 2 half + 1 (K 64 s + 32 half .. + 31). FP32 step s: piece
  // 2 s + half, K 8 s + 4 half + t for t = 0 .. 3; MFMA t of the step takes value t, so it
  // sums K 8 s + t and 8 s + 4 + t (a K order that A and B share: the sum is unchanged).
  auto frag = [&](int buf, int row, int s, int part) -> F {
    if constexpr (KIND == kFP8) {
      const int p0 = 4 * s + 2 * half;
      const i32x4 x = piece(buf, row, p0), y = piece(buf, row, p0 + 1);
      return i32x8{x[0], x[1], x[2], x[3], y[0], y[1], y[2], y[3]};
    } else if constexpr (KIND == kF32) {
      return piece(buf, row, 2 * s + half);
    } else {
      return __builtin_bit_cast(bf16x8, piece(buf, row, 4 * part + 2 * s + half));
    }
  };

  f32x16 acc[kWaveTiles][kWaveTiles];
#pragma unroll
  for (int i = 0; i < kWaveTiles; ++i)
#pragma unroll
    for (int j = 0; j < kWaveTiles; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;

  // FP8 scales of the current stage (set 0) and of the next one (set 1, loaded one stage
  // ahead and moved to set 0 at the stage boundary).
  int sa[2][kWaveTiles], sb[2][kWaveTiles];
  auto scales = [&](int ks, int set) {
#pragma unroll
    for (int i = 0; i < kWaveTiles; ++i) {
      if constexpr (KIND == kFP8) {
        sa[set][i] = sx[int64_t(ks) * Mp + m0 + wm * 128 + i * 32 + l32];
        sb[set][i] = sy[int64_t(ks) * Np + n0 + wn * 128 + i * 32 + l32];
      } else {
        sa[set][i] = sb[set][i] = 0;
      }
    }
  };
  // DMAs of the next stage ride in the first MFMA groups of a stage so that they have most of
  // the stage to land: 4 per group in 4 groups, or 8 per group in 2 for FP8, whose MFMAs take
  // twice the cycles (a group of 4 spans 256 of the stage's 2048).
  auto dma_slot = [&](int group, int ks_next, int buf) {
    constexpr int kPer = KIND == kFP8 ? 8 : 4;
    if (group < 16 / kPer) {
#pragma unroll
      for (int e = 0; e < kPer; ++e) dma(kPer * group + e, ks_next, buf);
    }
  };

#pragma unroll
  for (int d = 0; d < 16; ++d) dma(d, 0, 0);
  scales(0, 0);
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();

  if constexpr (KIND == kBF16X3) {
    // Fragments: al, ah, bl single; bh double (it is used by the first and the last MFMA
    // group of a step). Order per step: A = Al.Bh, B = Ah.Bl, C = Ah.Bh. Ah, Bl of step s
    // are read at its start (used from group B on); Al, Bh of step s + 1 right after group A.
    F al[kWaveTiles], ah[kWaveTiles], bl[kWaveTiles], bh[2][kWaveTiles];
#pragma unroll
    for (int i = 0; i < kWaveTiles; ++i) {
      al[i] = frag(0, arow + i * 32 * kRowB, 0, 1);
      bh[0][i] = frag(0, brow + i * 32 * kRowB, 0, 0);
    }
    auto stage = [&](int ks, auto more_tag) {
      constexpr bool kMore = decltype(more_tag)::value;
      const int buf = ks & 1;
#pragma unroll
      for (int s = 0; s < P::kSteps; ++s) {
#pragma unroll
        for (int i = 0; i < kWaveTiles; ++i) {
          ah[i] = frag(buf, arow + i * 32 * kRowB, s, 0);
          bl[i] = frag(buf, brow + i * 32 * kRowB, s, 1);
        }
        CSMI_SCHED_FENCE();
#pragma unroll
        for (int g = 0; g < kWaveTiles; ++g) {
#pragma unroll
          for (int j = 0; j < kWaveTiles; ++j) mma<KIND>(al[g], bh[s & 1][j], acc[g][j], 0, 0);
          if (kMore && s == 0) dma_slot(g, ks + 1, buf ^ 1);
          CSMI_SCHED_FENCE();
        }
        if (s + 1 < P::kSteps) {
#pragma unroll
          for (int i = 0; i < kWaveTiles; ++i) {
            al[i] = frag(buf, arow + i * 32 * kRowB, s + 1, 1);
            bh[(s + 1) & 1][i] = frag(buf, brow + i * 32 * kRowB, s + 1, 0);
          }
        } else {
          // Every read of this stage's buffer has been issued (Ah, Bl above): publish the next
          // stage (own DMAs landed, then the barrier) and read its first fragments while the
          // remaining two groups run.
          asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
          __syncthreads();
          if (kMore) {
#pragma unroll
            for (int i = 0; i < kWaveTiles; ++i) {
              al[i] = frag(buf ^ 1, arow + i * 32 * kRowB, 0, 1);
              bh[0][i] = frag(buf ^ 1, brow + i * 32 * kRowB, 0, 0);
            }
          }
        }
        CSMI_SCHED_FENCE();
#pragma unroll
        for (int g = 0; g < kWaveTiles; ++g) {
#pragma unroll
          for (int j = 0; j < kWaveTiles; ++j) mma<KIND>(ah[g], bl[j], acc[g][j], 0, 0);
          CSMI_SCHED_FENCE();
        }
#pragma unroll
        for (int g = 0; g < kWaveTiles; ++g) {
#pragma unroll
          for (int j = 0; j < kWaveTiles; ++j) mma<KIND>(ah[g], bh[s & 1][j], acc[g][j], 0, 0);
          CSMI_SCHED_FENCE();
        }
      }
    };
    run_stages(stages, stage);
  } else {
    // Fragments double-buffered across the K steps of a stage (kSteps is even, so step 0
    // of every stage uses set 0).
    F fa[2][kWaveTiles], fb[2][kWaveTiles];
#pragma unroll
    for (int i = 0; i < kWaveTiles; ++i) {
      fa[0][i] = frag(0, arow + i * 32 * kRowB, 0, 0);
      fb[0][i] = frag(0, brow + i * 32 * kRowB, 0, 0);
    }
    auto stage = [&](int ks, auto more_tag) {
      constexpr bool kMore = decltype(more_tag)::value;
      const int buf = ks & 1;
      if (kMore) scales(ks + 1, 1);
#pragma unroll
      for (int s = 0; s < P::kSteps; ++s) {
        if (s + 1 < P::kSteps) {
#pragma unroll
          for (int i = 0; i < kWaveTiles; ++i) {
            fa[(s + 1) & 1][i] = frag(buf, arow + i * 32 * kRowB, s + 1, 0);
            fb[(s + 1) & 1][i] = frag(buf, brow + i * 32 * kRowB, s + 1, 0);
          }
        } else {
          // The last step's fragments were read during the previous step, so this stage's
          // buffer is no longer read: publish the next stage (own DMAs and scale loads landed,
          // then the barrier) and read its first fragments while this step's MFMAs run.
          asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
          __syncthreads();
          if (kMore) {
#pragma unroll
            for (int i = 0; i < kWaveTiles; ++i) {
              fa[0][i] = frag(buf ^ 1, arow + i * 32 * kRowB, 0, 0);
              fb[0][i] = frag(buf ^ 1, brow + i * 32 * kRowB, 0, 0);
            }
          }
        }
        CSMI_SCHED_FENCE();
        if constexpr (KIND == kF32) {
          // 4 MFMAs per tile, value t of every fragment in round t: 16 independent
          // accumulators between two uses of one. (__int_as_float, not __builtin_bit_cast: the
          // latter applied to a vector element reads element 0 with this compiler.)
#pragma unroll
          for (int t = 0; t < 4; ++t) {
#pragma unroll
            for (int g = 0; g < kWaveTiles; ++g) {
#pragma unroll
              for (int j = 0; j < kWaveTiles; ++j)
                acc[g][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(__int_as_float(fa[s & 1][g][t]),
                                                                 __int_as_float(fb[s & 1][j][t]),
                                                                 acc[g][j], 0, 0, 0);
              if (kMore && t == 0) dma_slot(s * kWaveTiles + g, ks + 1, buf ^ 1);
              CSMI_SCHED_FENCE();
            }
          }
        } else {
#pragma unroll
          for (int g = 0; g < kWaveTiles; ++g) {
#pragma unroll
            for (int j = 0; j < kWaveTiles; ++j)
              mma<KIND>(fa[s & 1][g], fb[s & 1][j], acc[g][j], sa[0][g], sb[0][j]);
            if (kMore) dma_slot(s * kWaveTiles + g, ks + 1, buf ^ 1);
            CSMI_SCHED_FENCE();
          }
        }
      }
      if (kMore) {
#pragma unroll
        for (int i = 0; i < kWaveTiles; ++i) {
          sa[0][i] = sa[1][i];
          sb[0][i] = sb[1][i];
        }
      }
    };
    run_stages(stages, stage);
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

int64_t round_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

__global__ __launch_bounds__(256) void zero_matrix_kernel(float *__restrict__ C, int M, int N, int ldc) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= int64_t(M) * N) return;
  C[(i / N) * ldc + i % N] = 0.0f;
}

template <int KIND>
void run_lds(cs_stream *stream, int M, int N, int K, const float *A, int lda, const float *B, int ldb, float *C,
             int ldc) {
  using P = Pipe<KIND>;
  const int64_t Mp = round_up(M, kTile), Np = round_up(N, kTile), Kp = round_up(K, P::kK);
  const int64_t rowb = Kp * P::kBytesPerK;
  if (rowb * kTile >= (int64_t(1) << 31))
    fail(CS_INVALID_VALUE, "K = " + std::to_string(K) + " is too large for this GEMM variant (K * bytes per "
                           "value must stay below 8 MiB)");
  const int64_t xbytes = Mp * rowb, ybytes = Np * rowb;
  int64_t total = xbytes + ybytes + 4096;
  if (KIND == kFP8) total += (Mp + Np) * (Kp / 128) + 4096;
  char *base = static_cast<char *>(stream_scratch(stream, static_cast<size_t>(total)));
  char *x = base, *y = base + xbytes;
  uint8_t *sx = nullptr, *sy = nullptr;
  hipStream_t s = stream->stream;
  const int vec_a = (lda % 4 == 0 && reinterpret_cast<uintptr_t>(A) % 16 == 0) ? 1 : 0;
  const int vec_b = (ldb % 4 == 0 && reinterpret_cast<uintptr_t>(B) % 16 == 0) ? 1 : 0;
  if constexpr (KIND == kFP8) {
    sx = reinterpret_cast<uint8_t *>(base + xbytes + ybytes);
    sy = sx + Mp * (Kp / 128);
  }
  {
    PackArgs pa;
    pa.a = A, pa.m = M, pa.k = K, pa.lda = lda, pa.mp = int(Mp), pa.kp = int(Kp), pa.vec_a = vec_a, pa.x = x;
    pa.b = B, pa.n = N, pa.ldb = ldb, pa.np = int(Np), pa.vec_b = vec_b, pa.y = y;
    pa.sx = sx, pa.sy = sy;
    int64_t col_tiles_k;
    if constexpr (KIND == kFP8) {  // 16 lanes per 128-wide block of a row; 128 K x 64 N column tiles
      pa.row_blocks = unsigned((Mp * (Kp / 128) * 16 + 255) / 256);
      pa.cols_gx = unsigned(Np / 64);
      col_tiles_k = Kp / 128;
    } else {  // one thread per 4 (FP32) or 8 K values of a row; 64 K x kPackTN N column tiles
      pa.row_blocks = unsigned((Mp * (Kp / (KIND == kF32 ? 4 : 8)) + 255) / 256);
      pa.cols_gx = unsigned(Np / kPackTN);
      col_tiles_k = (Kp + 63) / 64;
    }
    const int64_t blocks = int64_t(pa.row_blocks) + int64_t(pa.cols_gx) * col_tiles_k;
    hipLaunchKernelGGL(pack_operands_kernel<KIND>, dim3(unsigned(blocks)), dim3(256), 0, s, pa);
  }
  static const bool configured = [] {
    CS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&gemm_lds_kernel<KIND>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kLdsBytes)));
    return true;
  }();
  (void)configured;
  const unsigned grid = static_cast<unsigned>((Mp / kTile) * (Np / kTile));
  hipLaunchKernelGGL(gemm_lds_kernel<KIND>, dim3(grid), dim3(kThreads), kLdsBytes, s, M, N, int(Kp), int(Mp),
                     int(Np), x, y, sx, sy, C, ldc);
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
    if (!C) fail(CS_INVALID_VALUE, "null matrix pointer");
    set_device(stream->device);
    if (K == 0) {  // empty sums: C = 0, and A and B are not read
      const int64_t n = int64_t(M) * N;
      hipLaunchKernelGGL(zero_matrix_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream->stream, C, M, N,
                         ldc);
      CS_HIP(hipGetLastError());
      return;
    }
    if (!A || !B) fail(CS_INVALID_VALUE, "null matrix pointer");
    if (int64_t(M) * K >= (int64_t(1) << 31) || int64_t(N) * K >= (int64_t(1) << 31))
      fail(CS_INVALID_VALUE, "operands are limited to 2^31 elements");
    switch (variant) {
      case CS_GEMM_FP32: run_lds<kF32>(stream, M, N, K, A, lda, B, ldb, C, ldc); break;
      case CS_GEMM_BF16: run_lds<kBF16>(stream, M, N, K, A, lda, B, ldb, C, ldc); break;
      case CS_GEMM_BF16X3: run_lds<kBF16X3>(stream, M, N, K, A, lda, B, ldb, C, ldc); break;
      case CS_GEMM_FP8: run_lds<kFP8>(stream, M, N, K, A, lda, B, ldb, C, ldc); break;
    }
  });
}
