// A small fully connected network (the cost model's predictor), trained and evaluated on
// the device: forward, mean-squared-error gradient, backward and Adam, all in FP32.
//
// The reference has no counterpart (it has no tuning). Layout: parameters are one flat
// FP32 array [W1 (d0 x d1), b1 (d1), W2 (d1 x d2), b2, ...]; W is row-major [in][out]; ReLU
// follows every layer but the last. Every reduction runs in a fixed order (one thread per
// output sum), so results are deterministic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "internal.h"

struct cs_mlp {
  cs_device *device = nullptr;
  std::vector<int> widths;
  std::vector<int64_t> w_off, b_off;  // offsets of each layer's W and b in the flat arrays
  int64_t count = 0;
  float *params = nullptr, *m = nullptr, *v = nullptr;
  float *acts = nullptr;  // activations of every layer for the last forward pass
  int64_t acts_capacity = 0;
  float *deltas = nullptr;  // two ping-pong gradient buffers of n x max width
  int64_t deltas_capacity = 0;
};

namespace csmi {
namespace {

constexpr int kB = 256;

__global__ __launch_bounds__(kB) void linear_fwd(const float *__restrict__ X, int64_t n, int din,
                                                 const float *__restrict__ W, const float *__restrict__ b, int dout,
                                                 int relu, float *__restrict__ Y) {
  int64_t t = int64_t(blockIdx.x) * kB + threadIdx.x;
  if (t >= n * dout) return;
  int64_t s = t / dout;
  int j = static_cast<int>(t % dout);
  const float *x = X + s * din;
  float acc = b[j];
  for (int i = 0; i < din; ++i) acc = fmaf(x[i], W[int64_t(i) * dout + j], acc);
  Y[t] = relu ? fmaxf(acc, 0.0f) : acc;
}

// dpred = 2 (pred - target) / n; loss = sum (pred - target)^2 / n (one workgroup, fixed order).
__global__ __launch_bounds__(1024) void mse_grad(const float *__restrict__ pred, const float *__restrict__ target,
                                                 int64_t n, float *__restrict__ dpred, float *loss) {
  __shared__ double part[1024];
  double acc = 0.0;
  for (int64_t s = threadIdx.x; s < n; s += 1024) {
    float d = pred[s] - target[s];
    dpred[s] = 2.0f * d / float(n);
    acc += double(d) * double(d);
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if (int(threadIdx.x) < off) part[threadIdx.x] += part[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0 && loss) *loss = float(part[0] / double(n));
}

// g = dY * relu'(Y) is formed on the fly. dX[s][i] = sum_j g[s][j] W[i][j].
__global__ __launch_bounds__(kB) void linear_bwd_dx(const float *__restrict__ dY, const float *__restrict__ Y,
                                                    int relu, const float *__restrict__ W, int64_t n, int din,
                                                    int dout, float *__restrict__ dX) {
  int64_t t = int64_t(blockIdx.x) * kB + threadIdx.x;
  if (t >= n * din) return;
  int64_t s = t / din;
  int i = static_cast<int>(t % din);
  float acc = 0.0f;
  for (int j = 0; j < dout; ++j) {
    float g = dY[s * dout + j];
    if (relu && Y[s * dout + j] <= 0.0f) g = 0.0f;
    acc = fmaf(g, W[int64_t(i) * dout + j], acc);
  }
  dX[t] = acc;
}

// dW[i][j] = sum_s X[s][i] g[s][j] for i < din; row i == din holds db[j] = sum_s g[s][j].
__global__ __launch_bounds__(kB) void linear_bwd_dw(const float *__restrict__ X, const float *__restrict__ dY,
                                                    const float *__restrict__ Y, int relu, int64_t n, int din,
                                                    int dout, float *__restrict__ dW, float *__restrict__ db) {
  int64_t t = int64_t(blockIdx.x) * kB + threadIdx.x;
  if (t >= int64_t(din + 1) * dout) return;
  int i = static_cast<int>(t / dout), j = static_cast<int>(t % dout);
  float acc = 0.0f;
  for (int64_t s = 0; s < n; ++s) {
    float g = dY[s * dout + j];
    if (relu && Y[s * dout + j] <= 0.0f) g = 0.0f;
    acc = fmaf(i < din ? X[s * din + i] : 1.0f, g, acc);
  }
  if (i < din) dW[int64_t(i) * dout + j] = acc;
  else db[j] = acc;
}

__global__ __launch_bounds__(kB) void adam_kernel(float *__restrict__ p, const float *__restrict__ g,
                                                  float *__restrict__ m, float *__restrict__ v, int64_t count,
                                                  float lr, float b1, float b2, float eps, float c1, float c2) {
  int64_t t = int64_t(blockIdx.x) * kB + threadIdx.x;
  if (t >= count) return;
  float gi = g[t];
  float mi = b1 * m[t] + (1.0f - b1) * gi;
  float vi = b2 * v[t] + (1.0f - b2) * gi * gi;
  m[t] = mi;
  v[t] = vi;
  p[t] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

unsigned blocks(int64_t n) { return static_cast<unsigned>((n + kB - 1) / kB); }

uint32_t hash32(uint32_t v) {
  v = (v ^ 61u) ^ (v >> 16);
  v *= 9u;
  v ^= v << 4;
  v *= 0x27d4eb2du;
  v ^= v >> 15;
  return v;
}

void ensure(float **buf, int64_t *cap, int64_t need) {
  if (*cap >= need) return;
  if (*buf) CS_HIP(hipFree(*buf));
  *buf = nullptr;
  CS_HIP(hipMalloc(reinterpret_cast<void **>(buf), static_cast<size_t>(need) * 4));
  *cap = need;
}

int64_t act_offset(const cs_mlp *net, int layer, int64_t n) {  // layer 1..L outputs
  int64_t off = 0;
  for (int l = 1; l < layer; ++l) off += n * net->widths[l];
  return off;
}

void forward(cs_stream *stream, cs_mlp *net, const float *X, int64_t n) {
  const int L = static_cast<int>(net->widths.size()) - 1;
  int64_t total = 0;
  for (int l = 1; l <= L; ++l) total += n * net->widths[l];
  ensure(&net->acts, &net->acts_capacity, total);
  const float *in = X;
  for (int l = 1; l <= L; ++l) {
    float *out = net->acts + act_offset(net, l, n);
    int din = net->widths[l - 1], dout = net->widths[l];
    hipLaunchKernelGGL(linear_fwd, dim3(blocks(n * dout)), dim3(kB), 0, stream->stream, in, n, din,
                       net->params + net->w_off[l - 1], net->params + net->b_off[l - 1], dout, l < L ? 1 : 0, out);
    in = out;
  }
  CS_HIP(hipGetLastError());
}

}  // namespace
}  // namespace csmi

using namespace csmi;

extern "C" {

cs_status cs_mlp_create(cs_device *device, const int32_t *widths, int nw, uint32_t seed, cs_mlp **mlp) {
  return guard([&] {
    if (!device || !widths || !mlp) fail(CS_INVALID_VALUE, "null argument");
    if (nw < 2 || nw > 16) fail(CS_INVALID_VALUE, "a network has 2 to 16 widths");
    for (int i = 0; i < nw; ++i)
      if (widths[i] < 1 || widths[i] > 4096) fail(CS_INVALID_VALUE, "layer width out of range");
    // freed with its device arrays if a later step fails
    std::unique_ptr<cs_mlp, void (*)(cs_mlp *)> net(new cs_mlp(), [](cs_mlp *x) {
      for (float *p : {x->params, x->m, x->v})
        if (p) (void)hipFree(p);
      delete x;
    });
    net->device = device;
    for (int i = 0; i < nw; ++i) net->widths.push_back(widths[i]);
    std::vector<float> host;
    for (int l = 1; l < nw; ++l) {
      int din = widths[l - 1], dout = widths[l];
      net->w_off.push_back(static_cast<int64_t>(host.size()));
      // He-uniform: U(-sqrt(6/din), sqrt(6/din)) from the Wang hash of (seed, index)
      double a = std::sqrt(6.0 / din);
      for (int64_t k = 0; k < int64_t(din) * dout; ++k) {
        uint32_t h = hash32(static_cast<uint32_t>(host.size()) ^ hash32(seed + 0x9e3779b9u * l));
        host.push_back(static_cast<float>(-a + 2.0 * a * (h * (1.0 / 4294967296.0))));
      }
      net->b_off.push_back(static_cast<int64_t>(host.size()));
      for (int j = 0; j < dout; ++j) host.push_back(0.0f);
    }
    net->count = static_cast<int64_t>(host.size());
    set_device(device);
    CS_HIP(hipMalloc(reinterpret_cast<void **>(&net->params), host.size() * 4));
    CS_HIP(hipMalloc(reinterpret_cast<void **>(&net->m), host.size() * 4));
    CS_HIP(hipMalloc(reinterpret_cast<void **>(&net->v), host.size() * 4));
    CS_HIP(hipMemcpy(net->params, host.data(), host.size() * 4, hipMemcpyHostToDevice));
    CS_HIP(hipMemset(net->m, 0, host.size() * 4));
    CS_HIP(hipMemset(net->v, 0, host.size() * 4));
    *mlp = net.release();
  });
}

cs_status cs_mlp_param_count(cs_mlp *mlp, int64_t *count) {
  return guard([&] {
    if (!mlp || !count) fail(CS_INVALID_VALUE, "null argument");
    *count = mlp->count;
  });
}

cs_status cs_mlp_get_params(cs_stream *stream, cs_mlp *mlp, float *host, int64_t count) {
  return guard([&] {
    if (!stream || !mlp || !host) fail(CS_INVALID_VALUE, "null argument");
    if (count != mlp->count) fail(CS_INVALID_VALUE, "parameter count mismatch");
    set_device(stream->device);
    CS_HIP(hipMemcpyAsync(host, mlp->params, size_t(count) * 4, hipMemcpyDeviceToHost, stream->stream));
    CS_HIP(hipStreamSynchronize(stream->stream));
  });
}

cs_status cs_mlp_set_params(cs_stream *stream, cs_mlp *mlp, const float *host, int64_t count) {
  return guard([&] {
    if (!stream || !mlp || !host) fail(CS_INVALID_VALUE, "null argument");
    if (count != mlp->count) fail(CS_INVALID_VALUE, "parameter count mismatch");
    set_device(stream->device);
    CS_HIP(hipMemcpyAsync(mlp->params, host, size_t(count) * 4, hipMemcpyHostToDevice, stream->stream));
    CS_HIP(hipStreamSynchronize(stream->stream));
  });
}

cs_status cs_mlp_forward(cs_stream *stream, cs_mlp *mlp, const float *X, int64_t n, float *y) {
  return guard([&] {
    if (!stream || !mlp || !X || !y) fail(CS_INVALID_VALUE, "null argument");
    if (n <= 0) fail(CS_INVALID_VALUE, "no samples");
    set_device(stream->device);
    forward(stream, mlp, X, n);
    const int L = static_cast<int>(mlp->widths.size()) - 1;
    CS_HIP(hipMemcpyAsync(y, mlp->acts + act_offset(mlp, L, n), size_t(n) * mlp->widths[L] * 4,
                          hipMemcpyDeviceToDevice, stream->stream));
  });
}

cs_status cs_mlp_backward(cs_stream *stream, cs_mlp *mlp, const float *X, const float *target, int64_t n,
                          float *loss, float *grads) {
  return guard([&] {
    if (!stream || !mlp || !X || !target || !grads) fail(CS_INVALID_VALUE, "null argument");
    if (n <= 0) fail(CS_INVALID_VALUE, "no samples");
    const int L = static_cast<int>(mlp->widths.size()) - 1;
    if (mlp->widths[L] != 1) fail(CS_INVALID_VALUE, "the loss needs a single output");
    set_device(stream->device);
    forward(stream, mlp, X, n);
    int wmax = 1;
    for (int w : mlp->widths) wmax = std::max(wmax, w);
    ensure(&mlp->deltas, &mlp->deltas_capacity, 2 * n * wmax);
    float *d_cur = mlp->deltas, *d_next = mlp->deltas + n * wmax;
    const float *pred = mlp->acts + act_offset(mlp, L, n);
    hipLaunchKernelGGL(mse_grad, dim3(1), dim3(1024), 0, stream->stream, pred, target, n, d_cur, loss);
    for (int l = L; l >= 1; --l) {
      int din = mlp->widths[l - 1], dout = mlp->widths[l];
      const float *Y = mlp->acts + act_offset(mlp, l, n);
      const float *Xin = l == 1 ? X : mlp->acts + act_offset(mlp, l - 1, n);
      int relu = l < L ? 1 : 0;
      hipLaunchKernelGGL(linear_bwd_dw, dim3(blocks(int64_t(din + 1) * dout)), dim3(kB), 0, stream->stream, Xin,
                         d_cur, Y, relu, n, din, dout, grads + mlp->w_off[l - 1], grads + mlp->b_off[l - 1]);
      if (l > 1) {
        hipLaunchKernelGGL(linear_bwd_dx, dim3(blocks(n * din)), dim3(kB), 0, stream->stream, d_cur, Y, relu,
                           mlp->params + mlp->w_off[l - 1], n, din, dout, d_next);
        std::swap(d_cur, d_next);
      }
    }
    CS_HIP(hipGetLastError());
  });
}

cs_status cs_mlp_adam(cs_stream *stream, cs_mlp *mlp, const float *grads, float lr, int t) {
  return guard([&] {
    if (!stream || !mlp || !grads) fail(CS_INVALID_VALUE, "null argument");
    if (t < 1) fail(CS_INVALID_VALUE, "Adam steps count from 1");
    const float b1 = 0.9f, b2 = 0.999f;
    float c1 = 1.0f - std::pow(b1, float(t)), c2 = 1.0f - std::pow(b2, float(t));
    set_device(stream->device);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks(mlp->count)), dim3(kB), 0, stream->stream, mlp->params, grads, mlp->m,
                       mlp->v, mlp->count, lr, b1, b2, 1e-8f, c1, c2);
    CS_HIP(hipGetLastError());
  });
}

cs_status cs_mlp_release(cs_mlp *mlp) {
  return guard([&] {
    if (!mlp) fail(CS_INVALID_VALUE, "mlp is null");
    set_device(mlp->device);
    for (float *p : {mlp->params, mlp->m, mlp->v, mlp->acts, mlp->deltas})
      if (p) (void)hipFree(p);
    delete mlp;
  });
}

}  // extern "C"
