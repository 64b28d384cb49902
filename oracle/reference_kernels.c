/*
 * Host-C restatement of the reference's fixed OpenCL programs and of the CPU baselines
 * (test infrastructure: only tests, smoke checks and the CPU-baseline leg of the
 * benchmark load it).
 *
 *   refcl_sequential_sum   -- `reduce` with reduce16 + reduce_rest, launched with
 *                             globalWorkSize = localWorkSize = 1 on CPU devices
 *                             (Tensors.scala:312-351, 682-696)
 *   refcl_parallel_sum     -- the two-stage work-group tree of the GPU path, including its
 *                             `offset > 1` loop (Tensors.scala:357-392, 697-753), simulated
 *                             group by group
 *   refcl_random           -- `random` (Tensors.scala:431-443)
 *   refcl_random_normal    -- `random_normal` (Tensors.scala:397-429), float libm
 *   base_row_fold          -- row sums / row maxima as the left fold of
 *                             split(1).reduce(_ + _) / reduce(max) (README.md:342)
 *   base_gemm_seqk         -- C[i,k] = ((a_i0 b_0k + a_i1 b_1k) + ...) in FP32, the
 *                             matrixMultiply1 kernel's sum order (README.md:312-325)
 *   fp64_row_stats         -- FP64 row sums, row maxima and total
 *
 * Build: oracle/build.py (gcc -O2 -fopenmp, -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

/* ------------------------------------------------------------------ K2 sequential */
float refcl_sequential_sum(const float *buffer, uint32_t length) {
  /* One work-item: get_global_id(0) = 0, get_global_size(0) = 1. */
  const uint32_t vector_length = length / 16;
  float result;
  if (vector_length >= 1) {
    float acc[16];
    memcpy(acc, buffer, sizeof acc);                       /* vload16(0, buffer) */
    for (uint32_t gi = 1; gi < vector_length; ++gi) {     /* accumulator + vload16(gi) */
      const float *v = buffer + 16 * (size_t)gi;
      for (int l = 0; l < 16; ++l) acc[l] = acc[l] + v[l];
    }
    float f8[8], f4[4], f2[2];
    for (int l = 0; l < 8; ++l) f8[l] = acc[8 + l] + acc[l]; /* accumulator.hi + accumulator.lo */
    for (int l = 0; l < 4; ++l) f4[l] = f8[4 + l] + f8[l];
    for (int l = 0; l < 2; ++l) f2[l] = f4[2 + l] + f4[l];
    float accumulator = f2[0] + f2[1];
    const uint32_t stage1 = 16 * vector_length;
    for (uint32_t i = 0; i < length - stage1; ++i)           /* reduce_rest */
      accumulator = accumulator + buffer[stage1 + i];
    result = accumulator;
  } else {
    float accumulator = 0.0f;
    for (uint32_t i = 0; i < length; ++i) accumulator = accumulator + buffer[i];
    result = accumulator;
  }
  return result;
}

/* ------------------------------------------------------------------ K3 parallel */
static void parallel_stage(const float *buffer, int length, int local_size, int groups, float *result) {
  const int global_size = local_size * groups;
  float scratch[1024];
  for (int g = 0; g < groups; ++g) {
    for (int l = 0; l < local_size; ++l) {
      int gi = g * local_size + l;
      float acc = 0.0f;
      while (gi < length) {
        acc = acc + buffer[gi];
        gi += global_size;
      }
      scratch[l] = acc;
    }
    for (unsigned offset = (unsigned)local_size / 2; offset > 1; offset = offset / 2)
      for (unsigned l = 0; l < offset; ++l) scratch[l] = scratch[l] + scratch[l + offset];
    result[g] = scratch[0] + scratch[1];
  }
}

/* work_group_size: CL_KERNEL_WORK_GROUP_SIZE; compute_units: maxComputeUnits. */
float refcl_parallel_sum(const float *buffer, int length, int work_group_size, int compute_units,
                         float *scratch_groups) {
  int local1 = length < work_group_size ? length : work_group_size;
  int groups = length % local1 == 0 ? length / local1 : length / local1 + 1;
  if (groups > compute_units) groups = compute_units;
  float out;
  if (groups == 1) {
    parallel_stage(buffer, length, local1, 1, &out);
    return out;
  }
  parallel_stage(buffer, length, local1, groups, scratch_groups);
  int local2 = groups < work_group_size ? groups : work_group_size;
  parallel_stage(scratch_groups, groups, local2, 1, &out);
  return out;
}

/* ------------------------------------------------------------------ K4 / K5 */
static inline uint32_t wang_hash(uint32_t v) {
  v = (v ^ 61u) ^ (v >> 16);
  v *= 9u;
  v ^= v << 4;
  v *= 0x27d4eb2du;
  v ^= v >> 15;
  return v;
}

static inline uint32_t xorshift32(uint32_t s) {
  const uint32_t t1 = s ^ (s << 13);
  const uint32_t t2 = t1 ^ (t1 >> 17);
  return t2 ^ (t2 << 5);
}

void refcl_random(float *buffer, int64_t size, uint32_t seed) {
  for (int64_t i = 0; i < size; ++i) buffer[i] = (float)wang_hash((uint32_t)i ^ seed) / 4294967296.0f;
}

/* Writes 2 * ceil(size / 2) floats. */
void refcl_random_normal(float *buffer, int64_t size, uint32_t seed) {
  const int64_t pairs = (size + 1) / 2;
  for (int64_t i = 0; i < pairs; ++i) {
    const uint32_t r1 = wang_hash((uint32_t)i ^ seed);
    const uint32_t r2 = xorshift32(r1);
    const float u1 = (float)r1 / 4294967296.0f;
    const float u2 = (float)r2 / 4294967296.0f;
    const float r = sqrtf(-2.0f * logf(u1));
    const float theta = 2.0f * 3.14159274101257324219f * u2;
    buffer[2 * i] = r * cosf(theta);
    buffer[2 * i + 1] = r * sinf(theta);
  }
}

/* ------------------------------------------------------------------ baselines */
void base_row_fold(const float *a, int64_t rows, int64_t cols, float *row_sum, float *row_max) {
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < rows; ++r) {
    const float *p = a + r * cols;
    float s = p[0], m = p[0];
    for (int64_t c = 1; c < cols; ++c) {
      s = s + p[c];
      m = fmaxf(m, p[c]);
    }
    if (row_sum) row_sum[r] = s;
    if (row_max) row_max[r] = m;
  }
}

void base_gemm_seqk(const float *A, const float *B, float *C, int M, int N, int K) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < M; ++i) {
    float *c = C + (int64_t)i * N;
    for (int j = 0; j < N; ++j) c[j] = A[(int64_t)i * K] * B[j];
    for (int k = 1; k < K; ++k) {
      const float a = A[(int64_t)i * K + k];
      const float *b = B + (int64_t)k * N;
      for (int j = 0; j < N; ++j) c[j] = c[j] + a * b[j];
    }
  }
}

void fp64_row_stats(const float *a, int64_t rows, int64_t cols, double *row_sum, float *row_max,
                    double *row_abs, double *total) {
  double t = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : t)
  for (int64_t r = 0; r < rows; ++r) {
    const float *p = a + r * cols;
    double s = 0.0, sa = 0.0;
    float m = p[0];
    for (int64_t c = 0; c < cols; ++c) {
      s += (double)p[c];
      sa += fabs((double)p[c]);
      m = fmaxf(m, p[c]);
    }
    row_sum[r] = s;
    row_abs[r] = sa;
    row_max[r] = m;
    t += s;
  }
  *total = t;
}
