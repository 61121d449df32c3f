// Parameter-selection network: a 4 -> 64 -> 64 -> 4 multilayer perceptron
// with tanh hidden layers, trained and evaluated by these kernels (FP32).
//
// Parameter vector layout (NPARAM = 4*64 + 64 + 64*64 + 64 + 64*4 + 4):
//   W1[64][4], b1[64], W2[64][64], b2[64], W3[4][64], b3[4]   (row = output)
//   h1 = tanh(W1 x + b1), h2 = tanh(W2 h1 + b2), y = W3 h2 + b3.
// Training minimises the mean squared error (1/B) sum_b |y_b - t_b|^2 with
// Adam; gradients are reduced over the batch per parameter, in a fixed
// order (no atomics), so a training run is bitwise reproducible.
#include <hip/hip_runtime.h>

#include "spme_mlp.h"

namespace spme {

constexpr int NI = 4, NH = 64, NO = 4;
constexpr int OW1 = 0, OB1 = OW1 + NH * NI, OW2 = OB1 + NH, OB2 = OW2 + NH * NH,
              OW3 = OB2 + NH, OB3 = OW3 + NO * NH, NPARAM = OB3 + NO;
static_assert(NPARAM == MLP_NPARAM, "parameter count");

__device__ __forceinline__ void load_params(float *s, const float *__restrict__ w) {
  for (int i = threadIdx.x; i < NPARAM; i += blockDim.x) s[i] = w[i];
  __syncthreads();
}

// One thread per sample.  Stores h1, h2 (post-activation) when non-null.
__global__ __launch_bounds__(256) void k_mlp_forward(const float *__restrict__ w,
                                                     const float *__restrict__ x, int B,
                                                     float *__restrict__ y,
                                                     float *__restrict__ h1o,
                                                     float *__restrict__ h2o) {
  __shared__ float s[NPARAM];
  load_params(s, w);
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float xi[NI];
#pragma unroll
  for (int k = 0; k < NI; ++k) xi[k] = x[b * NI + k];
  float h1[NH];
#pragma unroll
  for (int i = 0; i < NH; ++i) {
    float a = s[OB1 + i];
#pragma unroll
    for (int k = 0; k < NI; ++k) a += s[OW1 + i * NI + k] * xi[k];
    h1[i] = tanhf(a);
  }
  float out[NO];
#pragma unroll
  for (int o = 0; o < NO; ++o) out[o] = s[OB3 + o];
  for (int j = 0; j < NH; ++j) {
    float a = s[OB2 + j];
#pragma unroll
    for (int i = 0; i < NH; ++i) a += s[OW2 + j * NH + i] * h1[i];
    const float h2 = tanhf(a);
    if (h2o) h2o[(size_t)b * NH + j] = h2;
#pragma unroll
    for (int o = 0; o < NO; ++o) out[o] += s[OW3 + o * NH + j] * h2;
  }
  if (h1o) {
#pragma unroll
    for (int i = 0; i < NH; ++i) h1o[(size_t)b * NH + i] = h1[i];
  }
#pragma unroll
  for (int o = 0; o < NO; ++o) y[b * NO + o] = out[o];
}

// Per sample: dy = 2 (y - t) m / B (m: per-output loss weight),
// d2 = (W3^T dy) (1 - h2^2), d1 = (W2^T d2) (1 - h1^2).
__global__ __launch_bounds__(256) void k_mlp_delta(const float *__restrict__ w,
                                                   const float *__restrict__ y,
                                                   const float *__restrict__ t,
                                                   const float *__restrict__ m, int B,
                                                   const float *__restrict__ h1,
                                                   const float *__restrict__ h2,
                                                   float *__restrict__ dy,
                                                   float *__restrict__ d1,
                                                   float *__restrict__ d2) {
  __shared__ float s[NPARAM];
  load_params(s, w);
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float g[NO];
#pragma unroll
  for (int o = 0; o < NO; ++o) {
    g[o] = 2.0f * (y[b * NO + o] - t[b * NO + o]) * m[o] / float(B);
    dy[b * NO + o] = g[o];
  }
  float e2[NH];
#pragma unroll
  for (int j = 0; j < NH; ++j) {
    float a = 0.f;
#pragma unroll
    for (int o = 0; o < NO; ++o) a += s[OW3 + o * NH + j] * g[o];
    const float h = h2[(size_t)b * NH + j];
    e2[j] = a * (1.f - h * h);
    d2[(size_t)b * NH + j] = e2[j];
  }
  for (int i = 0; i < NH; ++i) {
    float a = 0.f;
#pragma unroll
    for (int j = 0; j < NH; ++j) a += s[OW2 + j * NH + i] * e2[j];
    const float h = h1[(size_t)b * NH + i];
    d1[(size_t)b * NH + i] = a * (1.f - h * h);
  }
}

// One thread per parameter: the gradient summed over the batch in sample
// order.
__global__ __launch_bounds__(256) void k_mlp_grad(const float *__restrict__ x, int B,
                                                  const float *__restrict__ h1,
                                                  const float *__restrict__ h2,
                                                  const float *__restrict__ dy,
                                                  const float *__restrict__ d1,
                                                  const float *__restrict__ d2,
                                                  float *__restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  float acc = 0.f;
  if (p < OB1) {  // W1[i][k]
    const int i = (p - OW1) / NI, k = (p - OW1) % NI;
    for (int b = 0; b < B; ++b) acc += d1[(size_t)b * NH + i] * x[b * NI + k];
  } else if (p < OW2) {  // b1[i]
    const int i = p - OB1;
    for (int b = 0; b < B; ++b) acc += d1[(size_t)b * NH + i];
  } else if (p < OB2) {  // W2[j][i]
    const int j = (p - OW2) / NH, i = (p - OW2) % NH;
    for (int b = 0; b < B; ++b) acc += d2[(size_t)b * NH + j] * h1[(size_t)b * NH + i];
  } else if (p < OW3) {  // b2[j]
    const int j = p - OB2;
    for (int b = 0; b < B; ++b) acc += d2[(size_t)b * NH + j];
  } else if (p < OB3) {  // W3[o][j]
    const int o = (p - OW3) / NH, j = (p - OW3) % NH;
    for (int b = 0; b < B; ++b) acc += dy[b * NO + o] * h2[(size_t)b * NH + j];
  } else {  // b3[o]
    const int o = p - OB3;
    for (int b = 0; b < B; ++b) acc += dy[b * NO + o];
  }
  grad[p] = acc;
}

// Adam (Kingma & Ba 2015) with bias correction, step t >= 1.
__global__ __launch_bounds__(256) void k_mlp_adam(float *__restrict__ w,
                                                  const float *__restrict__ g,
                                                  float *__restrict__ m1,
                                                  float *__restrict__ m2, float lr, float b1,
                                                  float b2, float eps, int t) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  const float gp = g[p];
  const float a = b1 * m1[p] + (1.f - b1) * gp;
  const float v = b2 * m2[p] + (1.f - b2) * gp * gp;
  m1[p] = a;
  m2[p] = v;
  const float ah = a / (1.f - powf(b1, float(t)));
  const float vh = v / (1.f - powf(b2, float(t)));
  w[p] -= lr * ah / (sqrtf(vh) + eps);
}

}  // namespace spme

using namespace spme;

namespace {
int grid1(int n, int b) { return (n + b - 1) / b; }
}

extern "C" {

int spme_mlp_forward(const float *w, const float *x, int B, float *y, float *h1, float *h2,
                     void *stream) {
  if (!w || !x || !y || B < 0) return -1;
  if (B == 0) return 0;
  (void)hipGetLastError();  // clear an error left by an earlier, unrelated call
  hipLaunchKernelGGL(k_mlp_forward, dim3(grid1(B, 256)), dim3(256), 0, (hipStream_t)stream, w,
                     x, B, y, h1, h2);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : -100 - int(e);
}

int spme_mlp_backward(const float *w, const float *x, const float *y, const float *t,
                      const float *mask, int B, const float *h1, const float *h2,
                      float *work, float *grad, void *stream) {
  if (!w || !x || !y || !t || !mask || !h1 || !h2 || !work || !grad || B <= 0) return -1;
  hipStream_t st = (hipStream_t)stream;
  (void)hipGetLastError();
  float *dy = work, *d1 = work + (size_t)B * NO, *d2 = d1 + (size_t)B * NH;
  hipLaunchKernelGGL(k_mlp_delta, dim3(grid1(B, 256)), dim3(256), 0, st, w, y, t, mask, B, h1,
                     h2, dy, d1, d2);
  hipLaunchKernelGGL(k_mlp_grad, dim3(grid1(NPARAM, 256)), dim3(256), 0, st, x, B, h1, h2, dy,
                     d1, d2, grad);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : -100 - int(e);
}

int spme_mlp_adam(float *w, const float *g, float *m1, float *m2, float lr, float beta1,
                  float beta2, float eps, int step, void *stream) {
  if (!w || !g || !m1 || !m2 || step < 1) return -1;
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_mlp_adam, dim3(grid1(NPARAM, 256)), dim3(256), 0, (hipStream_t)stream,
                     w, g, m1, m2, lr, beta1, beta2, eps, step);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : -100 - int(e);
}

int spme_mlp_nparam(void) { return NPARAM; }

}  // extern "C"
