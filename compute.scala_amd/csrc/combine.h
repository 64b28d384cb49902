// Cross-workgroup combine of a reduction inside the kernel that produced the partials.
//
// Every workgroup writes its partial with an agent-scope store (`sc1`: it writes through the
// XCD's L2, so a workgroup on another XCD reads the new value), waits for that store to
// complete and then takes a ticket from an arrival counter. The workgroup that draws the last
// ticket invalidates its view of the caches (an agent-scope acquire), adds every partial in an
// order that depends only on their count, writes the result and puts the counter back to zero
// for the next launch on the stream. So the result is the same bits whichever workgroup
// arrives last, and no second kernel (with the launch gap before it) is needed.
//
// Only the partials are published: the release that a C++ fetch_add would carry also writes
// back every dirty line of the XCD's L2 (`buffer_wbl2`), once per workgroup; a store that is
// already coherent at agent scope needs only to have completed (`s_waitcnt vmcnt(0)`).
//
// Used by the reduction kernels of the library and, as source text, by generated sum kernels.
#pragma once

// Thread 0 of each workgroup: publishes `v` as partial `i` of `n`; true in the workgroup that
// arrives last. The counter must be zero at launch (it is left at zero by the last workgroup).
__device__ __forceinline__ bool csmi_publish_f64(double *partials, unsigned i, double v, unsigned *ticket,
                                                 unsigned n) {
  __hip_atomic_store(partials + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // the partial is in memory before the ticket
  return __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == n - 1;
}

__device__ __forceinline__ bool csmi_publish_f32(float *partials, unsigned i, float v, unsigned *ticket,
                                                 unsigned n) {
  __hip_atomic_store(partials + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  return __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == n - 1;
}

// All threads of the last workgroup, after a barrier: later loads see every published partial.
__device__ __forceinline__ void csmi_acquire_partials() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

// Thread 0 of the last workgroup, after the result is written: the counter back to zero.
__device__ __forceinline__ void csmi_reset_ticket(unsigned *ticket) {
  __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of p[0..n) in FP64 by the calling workgroup of `threads` threads (a multiple of 64, at
// most 1024), valid in thread 0. The order depends on n and `threads` only: thread t adds
// the pairs t, t + threads, ... in index order (16-byte loads, sixteen in flight), then a
// fixed xor-shuffle tree inside each wave and the waves in order. p is 16-byte aligned.
__device__ __forceinline__ double csmi_fixed_sum(const double *p, long long n, int threads) {
  typedef double csmi_d2 __attribute__((ext_vector_type(2)));
  __shared__ double csmi_lw[16];
  double acc = 0.0;
  const long long pairs = n / 2;
  const csmi_d2 *p2 = reinterpret_cast<const csmi_d2 *>(p);
  long long i = threadIdx.x;
  for (; i + 15LL * threads < pairs; i += 16LL * threads) {
    csmi_d2 v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = p2[i + k * threads];
#pragma unroll
    for (int k = 0; k < 16; ++k) acc += v[k].x + v[k].y;
  }
  for (; i < pairs; i += threads) {
    const csmi_d2 v = p2[i];
    acc += v.x + v.y;
  }
  if (threadIdx.x == 0 && (n & 1)) acc += p[n - 1];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) csmi_lw[threadIdx.x >> 6] = acc;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
    t = csmi_lw[0];
    for (int w = 1; w < threads / 64; ++w) t += csmi_lw[w];
  }
  return t;
}

// Maximum of p[0..n) (fmaxf: NaN only if every value is NaN), valid in thread 0.
__device__ __forceinline__ float csmi_fixed_max(const float *p, long long n, int threads) {
  __shared__ float csmi_lm[16];
  float m = __builtin_nanf("");
  for (long long i = threadIdx.x; i < n; i += threads) m = fmaxf(m, p[i]);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) csmi_lm[threadIdx.x >> 6] = m;
  __syncthreads();
  float t = m;
  if (threadIdx.x == 0) {
    t = csmi_lm[0];
    for (int w = 1; w < threads / 64; ++w) t = fmaxf(t, csmi_lm[w]);
  }
  return t;
}
