// rk_mlp.hip -- the rebuild-interval proposal network: an MLP
// 16 -> 64 (tanh) -> 64 (tanh) -> 1 in FP32, with its forward, backward and
// Adam kernels.  It has no counterpart in the reference: it only chooses how
// many iterations the host enqueues before it reads the device displacement
// guard (rk_engine.cpp run_guarded), so it never changes a result.
//
// Parameters, one float array of kMlpParams:
//   W1[64][16] b1[64] W2[64][64] b2[64] W3[64] b3[1]
// Training: one thread per sample computes the activations and the
// back-propagated deltas (k_mlp_fwd_bwd); one thread per parameter then sums
// its gradient over the samples in sample order (k_mlp_grad), so gradients are
// deterministic; k_mlp_adam applies Adam.  Loss = mean of (y - t)^2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/rockable_amd.h"

namespace {

constexpr int kIn = 16, kH = 64;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH, oB3 = oW3 + kH;
constexpr int kMlpParams = oB3 + 1;  // 5313
constexpr int kBlock = 256;

// forward of one sample; h1, h2 receive the activations
__device__ __forceinline__ float mlp_forward(const float* __restrict__ P, const float* x, float* h1, float* h2) {
  for (int j = 0; j < kH; j++) {
    float a = P[oB1 + j];
    for (int i = 0; i < kIn; i++) a += P[oW1 + j * kIn + i] * x[i];
    h1[j] = tanhf(a);
  }
  for (int j = 0; j < kH; j++) {
    float a = P[oB2 + j];
    for (int i = 0; i < kH; i++) a += P[oW2 + j * kH + i] * h1[i];
    h2[j] = tanhf(a);
  }
  float y = P[oB3];
  for (int i = 0; i < kH; i++) y += P[oW3 + i] * h2[i];
  return y;
}

__global__ void __launch_bounds__(kBlock) k_mlp_forward(const float* __restrict__ P, const float* __restrict__ X, int n,
                                                        float* __restrict__ Y) {
  __shared__ float sp[kMlpParams];
  for (int q = threadIdx.x; q < kMlpParams; q += blockDim.x) sp[q] = P[q];
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[kIn], h1[kH], h2[kH];
  for (int i = 0; i < kIn; i++) x[i] = X[(size_t)s * kIn + i];
  Y[s] = mlp_forward(sp, x, h1, h2);
}

// per sample: activations h1, h2 and deltas d1, d2 (dL/d pre-activation),
// e = dL/dy = 2 (y - t) / n, and the squared error
__global__ void __launch_bounds__(kBlock) k_mlp_fwd_bwd(const float* __restrict__ P, const float* __restrict__ X,
                                                        const float* __restrict__ T, int n, float* __restrict__ H1,
                                                        float* __restrict__ H2, float* __restrict__ D1,
                                                        float* __restrict__ D2, float* __restrict__ E,
                                                        float* __restrict__ SQ) {
  __shared__ float sp[kMlpParams];
  for (int q = threadIdx.x; q < kMlpParams; q += blockDim.x) sp[q] = P[q];
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[kIn], h1[kH], h2[kH], d2[kH];
  for (int i = 0; i < kIn; i++) x[i] = X[(size_t)s * kIn + i];
  const float y = mlp_forward(sp, x, h1, h2);
  const float r = y - T[s];
  const float e = 2.0f * r / (float)n;
  for (int j = 0; j < kH; j++) {
    d2[j] = e * sp[oW3 + j] * (1.0f - h2[j] * h2[j]);
    H1[(size_t)s * kH + j] = h1[j];
    H2[(size_t)s * kH + j] = h2[j];
    D2[(size_t)s * kH + j] = d2[j];
  }
  for (int i = 0; i < kH; i++) {
    float a = 0.0f;
    for (int j = 0; j < kH; j++) a += sp[oW2 + j * kH + i] * d2[j];
    D1[(size_t)s * kH + i] = a * (1.0f - h1[i] * h1[i]);
  }
  E[s] = e;
  SQ[s] = r * r;
}

// one thread per parameter: its gradient summed over the samples in order;
// thread kMlpParams sums the loss
__global__ void __launch_bounds__(kBlock) k_mlp_grad(const float* __restrict__ X, int n, const float* __restrict__ H1,
                                                     const float* __restrict__ H2, const float* __restrict__ D1,
                                                     const float* __restrict__ D2, const float* __restrict__ E,
                                                     const float* __restrict__ SQ, float* __restrict__ G,
                                                     float* __restrict__ loss) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > kMlpParams) return;
  float g = 0.0f;
  if (q == kMlpParams) {
    for (int s = 0; s < n; s++) g += SQ[s];
    *loss = g / (float)n;
    return;
  }
  if (q < oB1) {
    const int j = q / kIn, i = q % kIn;
    for (int s = 0; s < n; s++) g += D1[(size_t)s * kH + j] * X[(size_t)s * kIn + i];
  } else if (q < oW2) {
    const int j = q - oB1;
    for (int s = 0; s < n; s++) g += D1[(size_t)s * kH + j];
  } else if (q < oB2) {
    const int j = (q - oW2) / kH, i = (q - oW2) % kH;
    for (int s = 0; s < n; s++) g += D2[(size_t)s * kH + j] * H1[(size_t)s * kH + i];
  } else if (q < oW3) {
    const int j = q - oB2;
    for (int s = 0; s < n; s++) g += D2[(size_t)s * kH + j];
  } else if (q < oB3) {
    const int i = q - oW3;
    for (int s = 0; s < n; s++) g += E[s] * H2[(size_t)s * kH + i];
  } else {
    for (int s = 0; s < n; s++) g += E[s];
  }
  G[q] = g;
}

__global__ void __launch_bounds__(kBlock) k_mlp_adam(float* __restrict__ P, const float* __restrict__ G,
                                                     float* __restrict__ M, float* __restrict__ V, float lr,
                                                     float b1t, float b2t) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= kMlpParams) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float g = G[q];
  const float m = b1 * M[q] + (1.0f - b1) * g;
  const float v = b2 * V[q] + (1.0f - b2) * g * g;
  M[q] = m;
  V[q] = v;
  P[q] -= lr * (m / (1.0f - b1t)) / (sqrtf(v / (1.0f - b2t)) + eps);
}

inline int nblk(long n) { return (int)((n + kBlock - 1) / kBlock); }

}  // namespace

struct rk_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  float* P = nullptr;  // parameters
  float *M = nullptr, *V = nullptr, *G = nullptr, *loss = nullptr;
  int64_t adam_t = 0;
  // sample buffers
  size_t cap = 0;
  float *X = nullptr, *T = nullptr, *Y = nullptr, *H1 = nullptr, *H2 = nullptr, *D1 = nullptr, *D2 = nullptr,
        *E = nullptr, *SQ = nullptr;
};

namespace {

int mfail(rk_mlp* m, const std::string& s) {
  if (m) m->err = s;
  return -1;
}
#define MCHK(x)                                                                          \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) return mfail(mlp, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

int reserve(rk_mlp* mlp, size_t n) {
  if (n <= mlp->cap) return 0;
  float** bufs[] = {&mlp->X, &mlp->T, &mlp->Y, &mlp->H1, &mlp->H2, &mlp->D1, &mlp->D2, &mlp->E, &mlp->SQ};
  for (float** b : bufs)
    if (*b) (void)hipFree(*b);
  const size_t c = n + n / 2 + 64;
  const size_t sz[] = {kIn, 1, 1, kH, kH, kH, kH, 1, 1};
  for (int k = 0; k < 9; k++) MCHK(hipMalloc(bufs[k], c * sz[k] * sizeof(float)));
  mlp->cap = c;
  return 0;
}

// one Adam step on the n samples already in X, T
int adam_step(rk_mlp* mlp, int n, float lr) {
  hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(nblk(n)), dim3(kBlock), 0, mlp->stream, mlp->P, mlp->X, mlp->T, n, mlp->H1,
                     mlp->H2, mlp->D1, mlp->D2, mlp->E, mlp->SQ);
  hipLaunchKernelGGL(k_mlp_grad, dim3(nblk(kMlpParams + 1)), dim3(kBlock), 0, mlp->stream, mlp->X, n, mlp->H1,
                     mlp->H2, mlp->D1, mlp->D2, mlp->E, mlp->SQ, mlp->G, mlp->loss);
  mlp->adam_t++;
  const float b1t = (float)std::pow(0.9, (double)mlp->adam_t), b2t = (float)std::pow(0.999, (double)mlp->adam_t);
  hipLaunchKernelGGL(k_mlp_adam, dim3(nblk(kMlpParams)), dim3(kBlock), 0, mlp->stream, mlp->P, mlp->G, mlp->M,
                     mlp->V, lr, b1t, b2t);
  MCHK(hipGetLastError());
  return 0;
}

}  // namespace

// device entry used by the engine: y[0] = MLP(x) for one feature vector
// already on the device, on the given stream
int rk_mlp_forward_device(rk_mlp* mlp, const float* x_dev, float* y_dev, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_forward, dim3(1), dim3(64), 0, s, mlp->P, x_dev, 1, y_dev);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" {

int rk_mlp_create(rk_mlp** out, int device, uint64_t seed) {
  if (!out) return -1;
  *out = nullptr;
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || device < 0 || device >= nd) return -2;
  if (hipSetDevice(device) != hipSuccess) return -2;
  rk_mlp* mlp = new rk_mlp();
  mlp->device = device;
  if (hipStreamCreate(&mlp->stream) != hipSuccess) {
    delete mlp;
    return -2;
  }
  // Glorot-uniform weights, zero biases
  std::vector<float> p(kMlpParams, 0.0f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int off, int fan_out, int fan_in) {
    const float a = std::sqrt(6.0f / (float)(fan_in + fan_out));
    std::uniform_real_distribution<float> u(-a, a);
    for (int k = 0; k < fan_out * fan_in; k++) p[off + k] = u(rng);
  };
  fill(oW1, kH, kIn);
  fill(oW2, kH, kH);
  fill(oW3, 1, kH);
  std::vector<float> zero(kMlpParams, 0.0f);
  if (hipMalloc(&mlp->P, kMlpParams * sizeof(float)) != hipSuccess ||
      hipMalloc(&mlp->M, kMlpParams * sizeof(float)) != hipSuccess ||
      hipMalloc(&mlp->V, kMlpParams * sizeof(float)) != hipSuccess ||
      hipMalloc(&mlp->G, kMlpParams * sizeof(float)) != hipSuccess ||
      hipMalloc(&mlp->loss, sizeof(float)) != hipSuccess ||
      hipMemcpy(mlp->P, p.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(mlp->M, zero.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(mlp->V, zero.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
    rk_mlp_destroy(mlp);
    return -2;
  }
  *out = mlp;
  return 0;
}

void rk_mlp_destroy(rk_mlp* mlp) {
  if (!mlp) return;
  float* bufs[] = {mlp->P, mlp->M, mlp->V, mlp->G, mlp->loss, mlp->X, mlp->T, mlp->Y,
                   mlp->H1, mlp->H2, mlp->D1, mlp->D2, mlp->E, mlp->SQ};
  for (float* b : bufs)
    if (b) (void)hipFree(b);
  if (mlp->stream) (void)hipStreamDestroy(mlp->stream);
  delete mlp;
}

const char* rk_mlp_last_error(const rk_mlp* mlp) { return mlp ? mlp->err.c_str() : "null model"; }

int rk_mlp_nparams(void) { return kMlpParams; }

int rk_mlp_get_params(rk_mlp* mlp, float* out, size_t n) {
  if (!mlp || !out || n < (size_t)kMlpParams) return mfail(mlp, "rk_mlp_get_params: buffer too small");
  MCHK(hipStreamSynchronize(mlp->stream));
  MCHK(hipMemcpy(out, mlp->P, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int rk_mlp_set_params(rk_mlp* mlp, const float* in, size_t n) {
  if (!mlp || !in || n < (size_t)kMlpParams) return mfail(mlp, "rk_mlp_set_params: buffer too small");
  MCHK(hipStreamSynchronize(mlp->stream));
  MCHK(hipMemcpy(mlp->P, in, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
  std::vector<float> zero(kMlpParams, 0.0f);
  MCHK(hipMemcpy(mlp->M, zero.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
  MCHK(hipMemcpy(mlp->V, zero.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
  mlp->adam_t = 0;
  return 0;
}

int rk_mlp_forward(rk_mlp* mlp, const float* X, size_t n, float* y) {
  if (!mlp || (!X && n) || (!y && n)) return -1;
  if (n == 0) return 0;
  (void)hipSetDevice(mlp->device);
  if (reserve(mlp, n)) return -1;
  MCHK(hipMemcpyAsync(mlp->X, X, n * kIn * sizeof(float), hipMemcpyHostToDevice, mlp->stream));
  hipLaunchKernelGGL(k_mlp_forward, dim3(nblk((long)n)), dim3(kBlock), 0, mlp->stream, mlp->P, mlp->X, (int)n, mlp->Y);
  MCHK(hipGetLastError());
  MCHK(hipMemcpyAsync(y, mlp->Y, n * sizeof(float), hipMemcpyDeviceToHost, mlp->stream));
  MCHK(hipStreamSynchronize(mlp->stream));
  return 0;
}

int rk_mlp_loss_grad(rk_mlp* mlp, const float* X, const float* t, size_t n, float* loss, float* grad) {
  if (!mlp || !X || !t || n == 0) return mfail(mlp, "rk_mlp_loss_grad: empty input");
  (void)hipSetDevice(mlp->device);
  if (reserve(mlp, n)) return -1;
  MCHK(hipMemcpyAsync(mlp->X, X, n * kIn * sizeof(float), hipMemcpyHostToDevice, mlp->stream));
  MCHK(hipMemcpyAsync(mlp->T, t, n * sizeof(float), hipMemcpyHostToDevice, mlp->stream));
  hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(nblk((long)n)), dim3(kBlock), 0, mlp->stream, mlp->P, mlp->X, mlp->T, (int)n,
                     mlp->H1, mlp->H2, mlp->D1, mlp->D2, mlp->E, mlp->SQ);
  hipLaunchKernelGGL(k_mlp_grad, dim3(nblk(kMlpParams + 1)), dim3(kBlock), 0, mlp->stream, mlp->X, (int)n, mlp->H1,
                     mlp->H2, mlp->D1, mlp->D2, mlp->E, mlp->SQ, mlp->G, mlp->loss);
  MCHK(hipGetLastError());
  if (grad) MCHK(hipMemcpyAsync(grad, mlp->G, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, mlp->stream));
  if (loss) MCHK(hipMemcpyAsync(loss, mlp->loss, sizeof(float), hipMemcpyDeviceToHost, mlp->stream));
  MCHK(hipStreamSynchronize(mlp->stream));
  return 0;
}

int rk_mlp_train(rk_mlp* mlp, const float* X, const float* t, size_t n, int epochs, int batch, float lr,
                 uint64_t seed, float* final_loss) {
  if (!mlp || !X || !t || n == 0 || epochs < 1 || batch < 1) return mfail(mlp, "rk_mlp_train: bad arguments");
  (void)hipSetDevice(mlp->device);
  const size_t b = std::min<size_t>((size_t)batch, n);
  if (reserve(mlp, std::max(n, b))) return -1;
  std::vector<size_t> order(n);
  for (size_t k = 0; k < n; k++) order[k] = k;
  std::mt19937_64 rng(seed);
  std::vector<float> bx(b * kIn), bt(b);
  for (int ep = 0; ep < epochs; ep++) {
    std::shuffle(order.begin(), order.end(), rng);
    for (size_t s0 = 0; s0 + b <= n; s0 += b) {
      for (size_t k = 0; k < b; k++) {
        memcpy(&bx[k * kIn], X + order[s0 + k] * kIn, kIn * sizeof(float));
        bt[k] = t[order[s0 + k]];
      }
      MCHK(hipMemcpyAsync(mlp->X, bx.data(), b * kIn * sizeof(float), hipMemcpyHostToDevice, mlp->stream));
      MCHK(hipMemcpyAsync(mlp->T, bt.data(), b * sizeof(float), hipMemcpyHostToDevice, mlp->stream));
      if (adam_step(mlp, (int)b, lr)) return -1;
      MCHK(hipStreamSynchronize(mlp->stream));  // bx, bt are reused
    }
  }
  if (final_loss) {
    float l = 0.0f;
    if (rk_mlp_loss_grad(mlp, X, t, n, &l, nullptr)) return -1;
    *final_loss = l;
  }
  return 0;
}

}  // extern "C"
