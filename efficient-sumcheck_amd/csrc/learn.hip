// A small fp32 network trained on the GPU: the serving scheduler's model of a
// request's service time (its admission priority).  Proof bits never depend
// on it; it only orders admissions.
//
// Network: x (D inputs) -> tanh(W1 x + b1) (H) -> tanh(W2 h1 + b2) (H) ->
// y = w3 . h2 + b3.  Loss: mean squared error over the batch.
// Parameters are one flat fp32 array: W1 [H][D], b1 [H], W2 [H][H], b2 [H],
// w3 [H], b3.
//
// Kernels (VALU; the layers are far too small for matrix cores):
//   k_mlp_forward   one thread per sample, parameters staged in LDS
//   k_mlp_backward  one workgroup per sample (a thread per hidden unit):
//                   forward, then the per-sample gradient of (y - t)^2 / n
//                   into row s of a scratch matrix
//   k_mlp_reduce    one thread per parameter sums its column in sample order
//                   (deterministic), plus the loss
//   k_mlp_adam      one thread per parameter
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "effsc_hip.h"

namespace {

constexpr int kMlpBlock = 256;
constexpr int kMaxIn = 16;
constexpr int kMaxHidden = 64;

struct Shape {
  int D, H;
  __host__ __device__ int w1() const { return 0; }
  __host__ __device__ int b1() const { return H * D; }
  __host__ __device__ int w2() const { return H * D + H; }
  __host__ __device__ int b2() const { return H * D + H + H * H; }
  __host__ __device__ int w3() const { return H * D + 2 * H + H * H; }
  __host__ __device__ int b3() const { return H * D + 3 * H + H * H; }
  __host__ __device__ int count() const { return H * D + 3 * H + H * H + 1; }
};

__device__ __forceinline__ void stage_params(const float* p, float* sp, int count) {
  for (int i = threadIdx.x; i < count; i += blockDim.x) sp[i] = p[i];
  __syncthreads();
}

// Forward of one sample per thread.  The input and the first hidden layer sit
// in LDS after the parameters, column k of thread t at k * blockDim.x + t (a
// private array indexed by a run-time width would live in scratch); the second
// layer is folded into the output as it is computed, in the same order.
constexpr int kFwdBlock = 64;

__global__ __launch_bounds__(kFwdBlock) void k_mlp_forward(Shape S, const float* params, const float* X, size_t n,
                                                           float* Y) {
  extern __shared__ float sp[];
  stage_params(params, sp, S.count());
  const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int t = threadIdx.x, B = blockDim.x;
  float* x = sp + S.count();
  float* h1 = x + S.D * B;
  for (int k = 0; k < S.D; k++) x[k * B + t] = X[s * S.D + k];
  for (int i = 0; i < S.H; i++) {
    float a = sp[S.b1() + i];
    for (int k = 0; k < S.D; k++) a = fmaf(sp[S.w1() + i * S.D + k], x[k * B + t], a);
    h1[i * B + t] = tanhf(a);
  }
  float y = sp[S.b3()];
  for (int i = 0; i < S.H; i++) {
    float a = sp[S.b2() + i];
    for (int k = 0; k < S.H; k++) a = fmaf(sp[S.w2() + i * S.H + k], h1[k * B + t], a);
    y = fmaf(sp[S.w3() + i], tanhf(a), y);
  }
  Y[s] = y;
}

// Row s of G: gradient of (y_s - t_s)^2 / n; E[s] = (y_s - t_s)^2 / n.
// One workgroup per sample: thread i < H owns hidden unit i, activations and
// deltas live in LDS next to the parameters.
__global__ __launch_bounds__(kMlpBlock) void k_mlp_backward(Shape S, const float* params, const float* X,
                                                            const float* T, size_t n, float* G, float* E) {
  extern __shared__ float sp[];
  const int P = S.count(), D = S.D, H = S.H;
  float* xs = sp + P;
  float* h1 = xs + kMaxIn;
  float* h2 = h1 + kMaxHidden;
  float* d2 = h2 + kMaxHidden;
  float* ys = d2 + kMaxHidden;
  const size_t s = blockIdx.x;
  const int i = threadIdx.x;
  for (int k = i; k < P; k += blockDim.x) sp[k] = params[k];
  if (i < D) xs[i] = X[s * D + i];
  __syncthreads();
  if (i < H) {
    float a = sp[S.b1() + i];
    for (int k = 0; k < D; k++) a = fmaf(sp[S.w1() + i * D + k], xs[k], a);
    h1[i] = tanhf(a);
  }
  __syncthreads();
  if (i < H) {
    float a = sp[S.b2() + i];
    for (int k = 0; k < H; k++) a = fmaf(sp[S.w2() + i * H + k], h1[k], a);
    h2[i] = tanhf(a);
  }
  __syncthreads();
  if (i == 0) {
    float y = sp[S.b3()];
    for (int k = 0; k < H; k++) y = fmaf(sp[S.w3() + k], h2[k], y);
    ys[0] = y;
  }
  __syncthreads();
  const float r = ys[0] - T[s];
  const float dy = 2.0f * r / (float)n;
  float* g = G + s * (size_t)P;
  if (i == 0) {
    E[s] = r * r / (float)n;
    g[S.b3()] = dy;
  }
  if (i < H) {
    g[S.w3() + i] = dy * h2[i];
    d2[i] = dy * sp[S.w3() + i] * (1.0f - h2[i] * h2[i]);  // through tanh
  }
  __syncthreads();
  // second layer: dW2[a][b] = d2[a] h1[b], db2 = d2
  for (int e = i; e < H * H; e += blockDim.x) g[S.w2() + e] = d2[e / H] * h1[e % H];
  if (i < H) {
    g[S.b2() + i] = d2[i];
    // first layer: d1[i] = sum_a d2[a] W2[a][i], through tanh
    float a = 0.0f;
    for (int k = 0; k < H; k++) a = fmaf(d2[k], sp[S.w2() + k * H + i], a);
    const float di = a * (1.0f - h1[i] * h1[i]);
    g[S.b1() + i] = di;
    for (int k = 0; k < D; k++) g[S.w1() + i * D + k] = di * xs[k];
  }
}

// grad[p] = sum_s G[s][p] in sample order; grad[count] = sum_s E[s].
__global__ __launch_bounds__(kMlpBlock) void k_mlp_reduce(Shape S, const float* G, const float* E, size_t n,
                                                          float* grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int P = S.count();
  if (p > P) return;
  float a = 0.0f;
  if (p < P) {
    for (size_t s = 0; s < n; s++) a += G[s * (size_t)P + p];
  } else {
    for (size_t s = 0; s < n; s++) a += E[s];
  }
  grad[p] = a;
}

// Adam (Kingma & Ba 2015) with bias correction.
__global__ __launch_bounds__(kMlpBlock) void k_mlp_adam(int P, float* params, const float* grad, float* m, float* v,
                                                        float lr, float b1, float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const float g = grad[p];
  const float mp = b1 * m[p] + (1.0f - b1) * g;
  const float vp = b2 * v[p] + (1.0f - b2) * g * g;
  m[p] = mp;
  v[p] = vp;
  params[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

}  // namespace

struct effsc_hip_mlp {
  int dev = 0;
  Shape S{0, 0};
  hipStream_t stream = nullptr;
  float* d_params = nullptr;
  float* d_m = nullptr;
  float* d_v = nullptr;
  float* d_grad = nullptr;  // count + 1 (loss)
  float* d_x = nullptr;
  float* d_t = nullptr;
  float* d_y = nullptr;
  float* d_G = nullptr;
  float* d_E = nullptr;
  size_t cap = 0;  // samples the data buffers hold
  long adam_t = 0;
};

namespace {

using effsc_hip::set_error;

int mlp_fail(int code, const std::string& m) { return set_error(code, m); }

void mlp_destroy(effsc_hip_mlp* m) {
  if (!m) return;
  (void)hipSetDevice(m->dev);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (float* p : {m->d_params, m->d_m, m->d_v, m->d_grad, m->d_x, m->d_t, m->d_y, m->d_G, m->d_E})
    if (p) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

int ensure_cap(effsc_hip_mlp* m, size_t n) {
  if (n <= m->cap) return EFFSC_HIP_OK;
  HIPCHK(hipStreamSynchronize(m->stream));
  for (float** p : {&m->d_x, &m->d_t, &m->d_y, &m->d_G, &m->d_E}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  const size_t P = m->S.count();
  HIPCHK(hipMalloc((void**)&m->d_x, n * m->S.D * 4));
  HIPCHK(hipMalloc((void**)&m->d_t, n * 4));
  HIPCHK(hipMalloc((void**)&m->d_y, n * 4));
  HIPCHK(hipMalloc((void**)&m->d_G, n * P * 4));
  HIPCHK(hipMalloc((void**)&m->d_E, n * 4));
  m->cap = n;
  return EFFSC_HIP_OK;
}

int upload(effsc_hip_mlp* m, const float* x, const float* t, size_t n) {
  TRY(ensure_cap(m, n));
  HIPCHK(hipMemcpyAsync(m->d_x, x, n * m->S.D * 4, hipMemcpyHostToDevice, m->stream));
  if (t) HIPCHK(hipMemcpyAsync(m->d_t, t, n * 4, hipMemcpyHostToDevice, m->stream));
  return EFFSC_HIP_OK;
}

int launch_grad(effsc_hip_mlp* m, size_t n) {
  const int P = m->S.count();
  // parameters, x, h1, h2, d2 and y in LDS
  const size_t lds = ((size_t)P + kMaxIn + 3 * kMaxHidden + 4) * 4;
  hipLaunchKernelGGL(k_mlp_backward, dim3((unsigned)n), dim3(kMlpBlock), lds, m->stream, m->S, m->d_params, m->d_x,
                     m->d_t, n, m->d_G, m->d_E);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_mlp_reduce, dim3((P + 1 + kMlpBlock - 1) / kMlpBlock), dim3(kMlpBlock), 0, m->stream, m->S,
                     m->d_G, m->d_E, n, m->d_grad);
  HIPCHK(hipGetLastError());
  return EFFSC_HIP_OK;
}

}  // namespace

extern "C" {

int effsc_hip_mlp_new(int device, int inputs, int hidden, uint64_t seed, effsc_hip_mlp** out) {
  if (!out) return mlp_fail(EFFSC_HIP_EINVAL, "null out");
  *out = nullptr;
  if (inputs < 1 || inputs > kMaxIn || hidden < 1 || hidden > kMaxHidden)
    return mlp_fail(EFFSC_HIP_EINVAL, "inputs must be in [1, 16] and hidden in [1, 64]");
  HIPCHK(hipSetDevice(device));
  auto* m = new effsc_hip_mlp();
  m->dev = device;
  m->S = Shape{inputs, hidden};
  const int P = m->S.count();
  // Glorot-uniform weights from a splitmix64 stream, zero biases
  std::vector<float> p(P, 0.0f);
  uint64_t z = seed;
  auto uni = [&]() {
    z += 0x9E3779B97F4A7C15ull;
    uint64_t x = z;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x ^= x >> 31;
    return (float)((double)(x >> 11) * (1.0 / 9007199254740992.0)) * 2.0f - 1.0f;
  };
  const Shape S = m->S;
  const float a1 = std::sqrt(6.0f / (S.D + S.H)), a2 = std::sqrt(6.0f / (2 * S.H)), a3 = std::sqrt(6.0f / (S.H + 1));
  for (int i = 0; i < S.H * S.D; i++) p[S.w1() + i] = a1 * uni();
  for (int i = 0; i < S.H * S.H; i++) p[S.w2() + i] = a2 * uni();
  for (int i = 0; i < S.H; i++) p[S.w3() + i] = a3 * uni();
  hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
  for (float** q : {&m->d_params, &m->d_m, &m->d_v})
    if (e == hipSuccess) e = hipMalloc((void**)q, P * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&m->d_grad, (P + 1) * 4);
  if (e == hipSuccess) e = hipMemcpy(m->d_params, p.data(), P * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(m->d_m, 0, P * 4);
  if (e == hipSuccess) e = hipMemset(m->d_v, 0, P * 4);
  if (e != hipSuccess) {
    int rc = mlp_fail(EFFSC_HIP_ENOMEM, std::string("mlp: ") + hipGetErrorString(e));
    mlp_destroy(m);
    return rc;
  }
  *out = m;
  return EFFSC_HIP_OK;
}

int effsc_hip_mlp_param_count(const effsc_hip_mlp* m, size_t* out) {
  if (!m || !out) return mlp_fail(EFFSC_HIP_EINVAL, "null argument");
  *out = (size_t)m->S.count();
  return EFFSC_HIP_OK;
}

int effsc_hip_mlp_get_params(effsc_hip_mlp* m, float* out) {
  if (!m || !out) return mlp_fail(EFFSC_HIP_EINVAL, "null argument");
  HIPCHK(hipSetDevice(m->dev));
  HIPCHK(hipMemcpyAsync(out, m->d_params, m->S.count() * 4, hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return EFFSC_HIP_OK;
}

int effsc_hip_mlp_set_params(effsc_hip_mlp* m, const float* in) {
  if (!m || !in) return mlp_fail(EFFSC_HIP_EINVAL, "null argument");
  HIPCHK(hipSetDevice(m->dev));
  HIPCHK(hipMemcpyAsync(m->d_params, in, m->S.count() * 4, hipMemcpyHostToDevice, m->stream));
  HIPCHK(hipMemsetAsync(m->d_m, 0, m->S.count() * 4, m->stream));
  HIPCHK(hipMemsetAsync(m->d_v, 0, m->S.count() * 4, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  m->adam_t = 0;
  return EFFSC_HIP_OK;
}

int effsc_hip_mlp_forward(effsc_hip_mlp* m, const float* x, size_t n, float* y) {
  if (!m || (n && (!x || !y))) return mlp_fail(EFFSC_HIP_EINVAL, "null argument");
  if (!n) return EFFSC_HIP_OK;
  HIPCHK(hipSetDevice(m->dev));
  TRY(upload(m, x, nullptr, n));
  const unsigned g = (unsigned)((n + kFwdBlock - 1) / kFwdBlock);
  const size_t lds = ((size_t)m->S.count() + (size_t)(m->S.D + m->S.H) * kFwdBlock) * 4;  // at most 41 KiB
  hipLaunchKernelGGL(k_mlp_forward, dim3(g), dim3(kFwdBlock), lds, m->stream, m->S, m->d_params, m->d_x, n,
                     m->d_y);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(y, m->d_y, n * 4, hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return EFFSC_HIP_OK;
}

int effsc_hip_mlp_loss_grad(effsc_hip_mlp* m, const float* x, const float* t, size_t n, float* loss, float* grad) {
  if (!m || !x || !t || n == 0) return mlp_fail(EFFSC_HIP_EINVAL, "bad argument");
  HIPCHK(hipSetDevice(m->dev));
  TRY(upload(m, x, t, n));
  TRY(launch_grad(m, n));
  const int P = m->S.count();
  std::vector<float> h(P + 1);
  HIPCHK(hipMemcpyAsync(h.data(), m->d_grad, (P + 1) * 4, hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  if (grad) std::memcpy(grad, h.data(), P * 4);
  if (loss) *loss = h[P];
  return EFFSC_HIP_OK;
}

int effsc_hip_mlp_train(effsc_hip_mlp* m, const float* x, const float* t, size_t n, int steps, float lr,
                        float* final_loss) {
  if (!m || !x || !t || n == 0 || steps < 0) return mlp_fail(EFFSC_HIP_EINVAL, "bad argument");
  HIPCHK(hipSetDevice(m->dev));
  TRY(upload(m, x, t, n));
  const int P = m->S.count();
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int it = 0; it < steps; it++) {
    TRY(launch_grad(m, n));
    m->adam_t++;
    const float c1 = 1.0f - std::pow(b1, (float)m->adam_t), c2 = 1.0f - std::pow(b2, (float)m->adam_t);
    hipLaunchKernelGGL(k_mlp_adam, dim3((P + kMlpBlock - 1) / kMlpBlock), dim3(kMlpBlock), 0, m->stream, P,
                       m->d_params, m->d_grad, m->d_m, m->d_v, lr, b1, b2, eps, c1, c2);
    HIPCHK(hipGetLastError());
  }
  // loss at the trained parameters
  TRY(launch_grad(m, n));
  float l = 0;
  HIPCHK(hipMemcpyAsync(&l, m->d_grad + P, 4, hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  if (final_loss) *final_loss = l;
  return EFFSC_HIP_OK;
}

void effsc_hip_mlp_free(effsc_hip_mlp* m) { mlp_destroy(m); }

}  // extern "C"
