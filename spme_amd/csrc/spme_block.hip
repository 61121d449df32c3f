// Tall-skinny block operations for block Krylov methods on batches of
// right-hand sides (the block-Lanczos M^{1/2} z of Brownian dynamics): blocks
// are row-major [s][ld] like the engine's batched inputs, small matrices are
// FP64 row-major on the device.  No counterpart in the reference.
//
//   gram:   C[a][b] = sum_i X[a][i] Y[b][i]        (FP64 accumulation)
//   update: Y[b][i] = beta Y[b][i] + sum_a X[a][i] A[a][b]
//
// The gram is deterministic: a fixed number of blocks (a function of n
// only) each sum a fixed set of chunks in a fixed order into a workspace of
// partial sums, then one block per entry reduces them in a fixed tree.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>

#include "../../include/spme_ext.h"
#include "spme_launch.h"

namespace spme {
namespace {
int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  return set_last_error(code, buf);
}

constexpr int BS = 16;      // largest block width
constexpr int GCH = 256;    // vector entries per staged chunk
constexpr int GMAXB = 512;  // most partial-sum blocks

int gram_blocks(long n) { return int(std::min<long>((n + GCH - 1) / GCH, GMAXB)); }

// Block g sums chunks g, g + nb, g + 2 nb, ... of the vectors; thread
// (a, b) keeps its entry's partial sum in a register across chunks.
template <typename T>
__global__ __launch_bounds__(256) void k_block_gram_part(long n, int sx, int sy,
                                                         const T *__restrict__ X, long ldx,
                                                         const T *__restrict__ Y, long ldy,
                                                         double *__restrict__ part) {
  __shared__ T xs[BS][GCH + 1], ys[BS][GCH + 1];
  const int a = threadIdx.x / BS, b = threadIdx.x % BS;
  const long nch = (n + GCH - 1) / GCH;
  double s = 0.0;
  for (long c = blockIdx.x; c < nch; c += gridDim.x) {
    const long i0 = c * GCH;
    const int m = int(n - i0 < GCH ? n - i0 : GCH);
    __syncthreads();  // the previous chunk has been read
    for (int t = threadIdx.x; t < BS * GCH; t += blockDim.x) {
      const int r = t / GCH, k = t - r * GCH;
      xs[r][k] = (r < sx && k < m) ? X[r * ldx + i0 + k] : T(0);
      ys[r][k] = (r < sy && k < m) ? Y[r * ldy + i0 + k] : T(0);
    }
    __syncthreads();
    if (a < sx && b < sy)
      for (int k = 0; k < m; ++k) s = fma(double(xs[a][k]), double(ys[b][k]), s);
  }
  if (a < sx && b < sy) part[(size_t)blockIdx.x * (sx * sy) + a * sy + b] = s;
}

// Entry e = blockIdx.x: lane l sums partials l, l + 64, ... in order, then a
// fixed butterfly over the wave.
__global__ __launch_bounds__(64) void k_block_gram_reduce(int nb, int ss,
                                                          const double *__restrict__ part,
                                                          double *__restrict__ C) {
  const int e = blockIdx.x, l = threadIdx.x;
  double s = 0.0;
  for (int k = l; k < nb; k += 64) s += part[(size_t)k * ss + e];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (l == 0) C[e] = s;
}

template <typename T>
__global__ __launch_bounds__(256) void k_block_update(long n, int sx, int sy,
                                                      const T *__restrict__ X, long ldx,
                                                      const double *__restrict__ A,
                                                      double beta, T *__restrict__ Y,
                                                      long ldy) {
  __shared__ double sA[BS * BS];
  for (int t = threadIdx.x; t < sx * sy; t += blockDim.x) sA[t] = A[t];
  __syncthreads();
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x[BS];
#pragma unroll
  for (int a = 0; a < BS; ++a) x[a] = a < sx ? double(X[a * ldx + i]) : 0.0;
  for (int b = 0; b < sy; ++b) {
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < BS; ++a)
      if (a < sx) s = fma(x[a], sA[a * sy + b], s);
    T *y = Y + b * ldy + i;
    *y = T((beta == 0.0 ? 0.0 : beta * double(*y)) + s);
  }
}

int check_block(int precision, long n, int sx, int sy, long ldx, long ldy) {
  if (precision != SPME_FP32 && precision != SPME_FP64 && precision != SPME_FP64_PLAIN)
    return fail(-1, "precision must be SPME_FP32, SPME_FP64 or SPME_FP64_PLAIN");
  if (n < 1) return fail(-1, "n must be positive");
  if (sx < 1 || sx > BS || sy < 1 || sy > BS)
    return fail(-1, "block widths must be in [1, %d]", BS);
  if (ldx < n || ldy < n) return fail(-1, "leading dimensions must be >= n");
  return 0;
}
}  // namespace
}  // namespace spme

using namespace spme;

extern "C" {

long spme_block_gram_worksize(long n, int sx, int sy) {
  return (long)gram_blocks(n) * sx * sy;
}

int spme_block_gram(int precision, long n, int sx, int sy, const void *X, long ldx,
                    const void *Y, long ldy, double *C, double *work, long work_len,
                    void *stream) {
  if (int rc = check_block(precision, n, sx, sy, ldx, ldy)) return rc;
  if (!X || !Y || !C || !work) return fail(-1, "null pointer");
  const long need = spme_block_gram_worksize(n, sx, sy);
  if (work_len < need) return fail(-1, "workspace too small (%ld < %ld doubles)", work_len, need);
  (void)hipGetLastError();
  hipStream_t st = (hipStream_t)stream;
  const int nb = gram_blocks(n);
  if (precision == SPME_FP32)
    hipLaunchKernelGGL(k_block_gram_part<float>, dim3(nb), dim3(256), 0, st, n, sx, sy,
                       (const float *)X, ldx, (const float *)Y, ldy, work);
  else
    hipLaunchKernelGGL(k_block_gram_part<double>, dim3(nb), dim3(256), 0, st, n, sx, sy,
                       (const double *)X, ldx, (const double *)Y, ldy, work);
  hipLaunchKernelGGL(k_block_gram_reduce, dim3(sx * sy), dim3(64), 0, st, nb, sx * sy, work, C);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : fail(-2, "block gram launch: %s", hipGetErrorString(e));
}

int spme_block_update(int precision, long n, int sx, int sy, const void *X, long ldx,
                      const double *A, double beta, void *Y, long ldy, void *stream) {
  if (int rc = check_block(precision, n, sx, sy, ldx, ldy)) return rc;
  if (!X || !A || !Y) return fail(-1, "null pointer");
  if (X == Y) return fail(-1, "X and Y must not alias");
  (void)hipGetLastError();
  hipStream_t st = (hipStream_t)stream;
  const int nb = int((n + 255) / 256);
  if (precision == SPME_FP32)
    hipLaunchKernelGGL(k_block_update<float>, dim3(nb), dim3(256), 0, st, n, sx, sy,
                       (const float *)X, ldx, A, beta, (float *)Y, ldy);
  else
    hipLaunchKernelGGL(k_block_update<double>, dim3(nb), dim3(256), 0, st, n, sx, sy,
                       (const double *)X, ldx, A, beta, (double *)Y, ldy);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : fail(-2, "block update launch: %s", hipGetErrorString(e));
}

}  // extern "C"
