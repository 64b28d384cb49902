// Streaming kernels that measure what HBM delivers to a kernel of this library: read-only,
// write-only and copy passes over 16-byte vectors with the load and store forms the generated
// kernels use (non-temporal loads; plain, non-temporal, or non-temporal write-through
// stores). They hold no arithmetic, so their time at a given size and launch shape is the
// ceiling a memory-bound kernel of that size and shape can reach; the benchmark states the
// fused kernels against it. Measurement only: nothing in the Tensor path calls them.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.h"

namespace csmi {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

template <int kStore>
__device__ __forceinline__ void store4(f4 *p, f4 v) {
  if constexpr (kStore == CS_PROBE_STORE_PLAIN) {
    *p = v;
  } else if constexpr (kStore == CS_PROBE_STORE_NT) {
    __builtin_nontemporal_store(v, p);
  } else {
    // non-temporal and written through at device scope (a vector store), as the generated
    // kernels' whole-line stores
    asm volatile("global_store_dwordx4 %0, %1, off sc1 nt" ::"v"(p), "v"(v) : "memory");
  }
}

// Thread t of block b handles vectors (b * U + u) * 256 + t for u < U (a wave's accesses are
// contiguous, as in the generated kernels); all U loads are issued before the first use.
template <int kKind, int kStore, int U>
__global__ __launch_bounds__(256) void probe_kernel(const f4 *__restrict__ src, f4 *__restrict__ dst, int64_t n4,
                                                    float *__restrict__ sink) {
  const int64_t base = int64_t(blockIdx.x) * U * 256 + threadIdx.x;
  f4 v[U];
  if constexpr (kKind != CS_PROBE_WRITE) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = base + int64_t(u) * 256;
      v[u] = i < n4 ? __builtin_nontemporal_load(src + i) : f4{0.0f, 0.0f, 0.0f, 0.0f};
    }
  }
  if constexpr (kKind == CS_PROBE_READ) {
    float s = 0.0f;
#pragma unroll
    for (int u = 0; u < U; ++u) s += (v[u].x + v[u].y) + (v[u].z + v[u].w);
    if (s == -1.0f) sink[0] = s;  // keeps the loads; never true for the inputs the benchmark uses
  } else {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = base + int64_t(u) * 256;
      if (i < n4) store4<kStore>(dst + i, kKind == CS_PROBE_WRITE ? f4{1.0f, 2.0f, 3.0f, 4.0f} : v[u]);
    }
  }
}

template <int kKind, int kStore>
void launch_probe(hipStream_t s, unsigned grid, int items, const f4 *src, f4 *dst, int64_t n4, float *sink) {
  switch (items) {
    case 1: hipLaunchKernelGGL((probe_kernel<kKind, kStore, 1>), dim3(grid), dim3(256), 0, s, src, dst, n4, sink); break;
    case 2: hipLaunchKernelGGL((probe_kernel<kKind, kStore, 2>), dim3(grid), dim3(256), 0, s, src, dst, n4, sink); break;
    case 4: hipLaunchKernelGGL((probe_kernel<kKind, kStore, 4>), dim3(grid), dim3(256), 0, s, src, dst, n4, sink); break;
    default: hipLaunchKernelGGL((probe_kernel<kKind, kStore, 8>), dim3(grid), dim3(256), 0, s, src, dst, n4, sink); break;
  }
}

template <int kKind>
void launch_kind(hipStream_t s, int store, unsigned grid, int items, const f4 *src, f4 *dst, int64_t n4, float *sink) {
  if (store == CS_PROBE_STORE_PLAIN) launch_probe<kKind, CS_PROBE_STORE_PLAIN>(s, grid, items, src, dst, n4, sink);
  else if (store == CS_PROBE_STORE_NT) launch_probe<kKind, CS_PROBE_STORE_NT>(s, grid, items, src, dst, n4, sink);
  else launch_probe<kKind, CS_PROBE_STORE_NT_WRITE_THROUGH>(s, grid, items, src, dst, n4, sink);
}

}  // namespace
}  // namespace csmi

using namespace csmi;

extern "C" cs_status cs_hbm_probe(cs_stream *stream, int kind, int store, const float *src, float *dst, int64_t n,
                                  int items_per_thread) {
  return guard([&] {
    if (!stream) fail(CS_INVALID_VALUE, "stream is null");
    if (kind != CS_PROBE_READ && kind != CS_PROBE_WRITE && kind != CS_PROBE_COPY) fail(CS_INVALID_VALUE, "unknown kind");
    if (store != CS_PROBE_STORE_PLAIN && store != CS_PROBE_STORE_NT && store != CS_PROBE_STORE_NT_WRITE_THROUGH)
      fail(CS_INVALID_VALUE, "unknown store form");
    if (items_per_thread != 1 && items_per_thread != 2 && items_per_thread != 4 && items_per_thread != 8)
      fail(CS_INVALID_VALUE, "items_per_thread is 1, 2, 4 or 8");
    if (n < 0 || n % 4) fail(CS_INVALID_VALUE, "n must be a non-negative multiple of 4");
    if (n == 0) return;
    if ((kind != CS_PROBE_WRITE && !src) || (kind != CS_PROBE_READ && !dst) || !dst)
      fail(CS_INVALID_VALUE, "null buffer (dst also receives the read pass's unused sink)");
    if ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15)
      fail(CS_INVALID_VALUE, "buffers must be 16-byte aligned");
    const int64_t n4 = n / 4;
    const int64_t per_block = int64_t(items_per_thread) * 256;
    const int64_t grid = (n4 + per_block - 1) / per_block;
    if (grid > 0x7fffffff) fail(CS_INVALID_VALUE, "too many vectors");
    set_device(stream->device);
    const f4 *s4 = reinterpret_cast<const f4 *>(src);
    f4 *d4 = reinterpret_cast<f4 *>(dst);
    hipStream_t s = stream->stream;
    const unsigned g = static_cast<unsigned>(grid);
    if (kind == CS_PROBE_READ) launch_kind<CS_PROBE_READ>(s, store, g, items_per_thread, s4, d4, n4, dst);
    else if (kind == CS_PROBE_WRITE) launch_kind<CS_PROBE_WRITE>(s, store, g, items_per_thread, s4, d4, n4, dst);
    else launch_kind<CS_PROBE_COPY>(s, store, g, items_per_thread, s4, d4, n4, dst);
    CS_HIP(hipGetLastError());
  });
}
