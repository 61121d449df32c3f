// Tall-skinny block operations for block Krylov methods on batches of
// right-hand sides (the block-Lanczos M^{1/2} z of Brownian dynamics): blocks
// are row-major [s][ld] like the engine's batched inputs, small matrices are
// FP64 row-major on the device.  No counterpart in the reference.
//
//   gram:   C[a][b] = sum_i X[a][i] Y[b][i]        (FP64 accumulation)
//   update: Y[b][i] = beta Y[b][i] + sum_a X[a][i] A[a][b]
//
// Y has at most 16 rows; X up to 4096 (the tall form: all earlier blocks
// of a Krylov basis stacked, one pass over them instead of one call each).
// The gram is deterministic: a fixed number of blocks (a function of n and
// sx only) each sum a fixed set of chunks in a fixed order into a workspace
// of partial sums, then one block per entry reduces them in a fixed tree.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <mutex>

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

constexpr int BS = 16;      // largest block width (Y; X up to 16: square form)
constexpr int BX = 4096;    // most rows of X (tall form)
constexpr int GCH = 256;    // vector entries per staged chunk
constexpr int GMAXB = 512;  // most partial-sum blocks
constexpr int TA = 64;      // tall form: rows of X per block
constexpr int TCH = 64;     // tall form: vector entries per staged chunk
constexpr int TGRID = 2048; // tall form: blocks per launch (about)

int gram_blocks(long n) { return int(std::min<long>((n + GCH - 1) / GCH, GMAXB)); }
// tall form: blocks along n per group of TA rows; each block leaves one
// partial sum per wave (4)
int gram_blocks_tall(long n, int sx) {
  const int nab = (sx + TA - 1) / TA;
  return int(std::min<long>((n + TCH - 1) / TCH, std::max(8, TGRID / nab)));
}
int gram_parts(long n, int sx) {
  return sx <= BS ? gram_blocks(n) : 4 * gram_blocks_tall(n, sx);
}

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

// Tall form, block (g, ab): rows ab*TA .. ab*TA + TA - 1 of X against all
// rows of Y over chunks g, g + gridDim.x, ...  Wave w sums entries
// w*TCH/4 .. (w+1)*TCH/4 - 1 of each chunk; its lane (rg, cg) keeps the 4 x 4
// tile rows 4rg.., columns 4cg.. in registers (4 + 4 LDS reads per 16 FMAs,
// Y staged transposed in FP64) and leaves it as partial sum 4g + w.
template <typename T>
__global__ __launch_bounds__(256) void k_block_gram_tall(long n, int sx, int sy,
                                                         const T *__restrict__ X, long ldx,
                                                         const T *__restrict__ Y, long ldy,
                                                         double *__restrict__ part) {
  static_assert(TA == 64 && BS == 16 && TA * TCH % 256 == 0 && BS * TCH % 256 == 0,
                "lane tiles: 16 x 4 of 4 x 4; 256 threads stage whole tiles");
  constexpr int KQ = TCH / 4;  // entries per wave per chunk
  __shared__ T xs[TA][TCH + 1];
  __shared__ __attribute__((aligned(16))) double yt[TCH][BS];
  const int a0 = blockIdx.y * TA, na = min(TA, sx - a0);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;  // 256 threads: w < 4
  const int rg = lane & 15, cg = lane >> 4;                 // rows 4rg.., columns 4cg..
  const long nch = (n + TCH - 1) / TCH;
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  // the next chunk's entries, loaded into registers while the current one
  // is summed: thread t holds entries t + 256 q of the X tile and of Y
  constexpr int NXR = TA * TCH / 256, NYR = BS * TCH / 256;
  T xr[NXR], yr[NYR];
  auto fetch = [&](long c) {
    const long i0 = c * TCH;
    const int m = int(n - i0 < TCH ? n - i0 : TCH);
#pragma unroll
    for (int q = 0; q < NXR; ++q) {
      const int t = threadIdx.x + 256 * q, r = t / TCH, k = t - r * TCH;
      xr[q] = (r < na && k < m) ? X[(long)(a0 + r) * ldx + i0 + k] : T(0);
    }
#pragma unroll
    for (int q = 0; q < NYR; ++q) {
      const int t = threadIdx.x + 256 * q, r = t / TCH, k = t - r * TCH;
      yr[q] = (r < sy && k < m) ? Y[(long)r * ldy + i0 + k] : T(0);
    }
  };
  if (blockIdx.x < nch) fetch(blockIdx.x);
  for (long c = blockIdx.x; c < nch; c += gridDim.x) {
    __syncthreads();  // the previous chunk has been read
#pragma unroll
    for (int q = 0; q < NXR; ++q) {
      const int t = threadIdx.x + 256 * q, r = t / TCH, k = t - r * TCH;
      xs[r][k] = xr[q];
    }
#pragma unroll
    for (int q = 0; q < NYR; ++q) {
      const int t = threadIdx.x + 256 * q, r = t / TCH, k = t - r * TCH;
      yt[k][r] = double(yr[q]);
    }
    __syncthreads();
    if (c + gridDim.x < nch) fetch(c + gridDim.x);
#pragma unroll 4
    for (int k = w * KQ; k < w * KQ + KQ; ++k) {  // entries past m are zero
      double x[4], y[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) x[i] = double(xs[4 * rg + i][k]);
#pragma unroll
      for (int j = 0; j < 4; ++j) y[j] = yt[k][4 * cg + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fma(x[i], y[j], acc[i][j]);
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int a = 4 * rg + i;
    if (a < na) {
      double *p = part + ((size_t)(4 * blockIdx.x + w) * sx + a0 + a) * sy;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (4 * cg + j < sy) p[4 * cg + j] = acc[i][j];
    }
  }
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

// Tall form: thread i sums over all rows a of X in order (TA rows of A at
// a time in LDS), 16 FP64 partial sums in registers.
template <typename T>
__global__ __launch_bounds__(256) void k_block_update_tall(long n, int sx, int sy,
                                                           const T *__restrict__ X, long ldx,
                                                           const double *__restrict__ A,
                                                           double beta, T *__restrict__ Y,
                                                           long ldy) {
  __shared__ __attribute__((aligned(16))) double sA[TA][BS];
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool on = i < n;
  double acc[BS];
#pragma unroll
  for (int b = 0; b < BS; ++b) acc[b] = 0.0;
  for (int a0 = 0; a0 < sx; a0 += TA) {
    const int na = min(TA, sx - a0);
    __syncthreads();  // the previous rows of A have been read
    for (int t = threadIdx.x; t < TA * BS; t += blockDim.x) {
      const int r = t / BS, b = t - r * BS;
      sA[r][b] = (r < na && b < sy) ? A[(size_t)(a0 + r) * sy + b] : 0.0;
    }
    __syncthreads();
    if (on) {
      int r = 0;
      for (; r + 8 <= na; r += 8) {  // eight rows' loads in flight
        double x[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) x[q] = double(X[(long)(a0 + r + q) * ldx + i]);
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
          for (int b = 0; b < BS; ++b) acc[b] = fma(x[q], sA[r + q][b], acc[b]);
      }
      for (; r < na; ++r) {
        const double x = double(X[(long)(a0 + r) * ldx + i]);
#pragma unroll
        for (int b = 0; b < BS; ++b) acc[b] = fma(x, sA[r][b], acc[b]);
      }
    }
  }
  if (on) {
#pragma unroll
    for (int b = 0; b < BS; ++b)
      if (b < sy) {
        T *y = Y + b * ldy + i;
        *y = T((beta == 0.0 ? 0.0 : beta * double(*y)) + acc[b]);
      }
  }
}

// Launches the partial-sum kernel over nbx blocks along n (tall form: nbx
// blocks per group of TA rows, 4 partial sums each) and the fixed-order
// reduction of the partial sums in work into C.
int gram_launch(int precision, long n, int sx, int sy, const void *X, long ldx, const void *Y,
                long ldy, double *C, double *work, int nbx, hipStream_t st) {
  int nb = nbx;
  if (sx <= BS) {
    if (precision == SPME_FP32)
      hipLaunchKernelGGL(k_block_gram_part<float>, dim3(nbx), dim3(256), 0, st, n, sx, sy,
                         (const float *)X, ldx, (const float *)Y, ldy, work);
    else
      hipLaunchKernelGGL(k_block_gram_part<double>, dim3(nbx), dim3(256), 0, st, n, sx, sy,
                         (const double *)X, ldx, (const double *)Y, ldy, work);
  } else {
    nb = 4 * nbx;
    const dim3 grid(nbx, (sx + TA - 1) / TA);
    if (precision == SPME_FP32)
      hipLaunchKernelGGL(k_block_gram_tall<float>, grid, dim3(256), 0, st, n, sx, sy,
                         (const float *)X, ldx, (const float *)Y, ldy, work);
    else
      hipLaunchKernelGGL(k_block_gram_tall<double>, grid, dim3(256), 0, st, n, sx, sy,
                         (const double *)X, ldx, (const double *)Y, ldy, work);
  }
  hipLaunchKernelGGL(k_block_gram_reduce, dim3(sx * sy), dim3(64), 0, st, nb, sx * sy, work, C);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : fail(-2, "block gram launch: %s", hipGetErrorString(e));
}

// Blocks along n of the workspace form: gram_parts partial sums.
int gram_nbx(long n, int sx) { return sx <= BS ? gram_blocks(n) : gram_blocks_tall(n, sx); }

int check_block(int precision, long n, int sx, int sy, long ldx, long ldy) {
  if (precision != SPME_FP32 && precision != SPME_FP64 && precision != SPME_FP64_PLAIN)
    return fail(-1, "precision must be SPME_FP32, SPME_FP64 or SPME_FP64_PLAIN");
  if (n < 1) return fail(-1, "n must be positive");
  if (sx < 1 || sx > BX || sy < 1 || sy > BS)
    return fail(-1, "block widths must be sx in [1, %d], sy in [1, %d]", BX, BS);
  if (ldx < n || ldy < n) return fail(-1, "leading dimensions must be >= n");
  return 0;
}
// ---------------------------------------------------------------------------
// The same product without a workspace argument, for callers that cannot
// allocate: the partial sums live in this file-scope device array, one copy
// per GPU, held until the process ends.  Host code must not pass the
// array's name to a kernel (on the host it is only a stand-in for the
// symbol): its device address is looked up once per device with
// hipGetSymbolAddress, on that device's first call, and used from then on.
// Every partial sum a launch reduces is written by that launch, so nothing
// needs clearing between calls.  All calls on one GPU share the array:
// calls that may overlap on one GPU (on different streams, or from several
// threads without ordering) race.
constexpr long GRAM_SCRATCH = 1L << 20;  // doubles (8 MiB)
static __device__ __align__(256) double scratch[GRAM_SCRATCH];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[MAX_DEVICES] = {};

// Device address of this device's scratch array (nullptr and an error message
// on failure).
double *scratch_on_current_device() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
    fail(-2, "no usable current device for the no-alloc gram");
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      fail(-2, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
           hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = static_cast<double *>(ptr);
  }
  return g_scratch_ptr[dev];
}

}  // namespace
}  // namespace spme

using namespace spme;

extern "C" {

long spme_block_gram_worksize(long n, int sx, int sy) {
  if (n < 1 || sx < 1 || sx > BX || sy < 1 || sy > BS) return 0;
  return (long)gram_parts(n, sx) * sx * sy;
}

int spme_block_gram(int precision, long n, int sx, int sy, const void *X, long ldx,
                    const void *Y, long ldy, double *C, double *work, long work_len,
                    void *stream) {
  if (int rc = check_block(precision, n, sx, sy, ldx, ldy)) return rc;
  if (!X || !Y || !C || !work) return fail(-1, "null pointer");
  const long need = spme_block_gram_worksize(n, sx, sy);
  if (work_len < need) return fail(-1, "workspace too small (%ld < %ld doubles)", work_len, need);
  (void)hipGetLastError();
  return gram_launch(precision, n, sx, sy, X, ldx, Y, ldy, C, work, gram_nbx(n, sx),
                     (hipStream_t)stream);
}


long spme_block_gram_noalloc_max(void) { return GRAM_SCRATCH; }

int spme_block_gram_noalloc(int precision, long n, int sx, int sy, const void *X, long ldx,
                            const void *Y, long ldy, double *C, void *stream) {
  if (int rc = check_block(precision, n, sx, sy, ldx, ldy)) return rc;
  if (!X || !Y || !C) return fail(-1, "null pointer");
  // the workspace form's partition when its partial sums fit, so the result
  // is bitwise that of spme_block_gram; otherwise fewer blocks along n (the
  // tall form only: a partition of its own, fixed by n and sx, so the result
  // stays deterministic)
  int nbx = gram_nbx(n, sx);
  const long per = sx <= BS ? (long)sx * sy : 4L * sx * sy;  // doubles per block along n
  if ((long)nbx * per > GRAM_SCRATCH) nbx = int(GRAM_SCRATCH / per);
  if (nbx < 1)
    return fail(-1, "no-alloc gram: %d x %d partial sums exceed the %ld-double scratch array",
                sx, sy, GRAM_SCRATCH);
  double *work = scratch_on_current_device();
  if (!work) return -2;
  (void)hipGetLastError();
  return gram_launch(precision, n, sx, sy, X, ldx, Y, ldy, C, work, nbx, (hipStream_t)stream);
}

int spme_block_update(int precision, long n, int sx, int sy, const void *X, long ldx,
                      const double *A, double beta, void *Y, long ldy, void *stream) {
  if (int rc = check_block(precision, n, sx, sy, ldx, ldy)) return rc;
  if (!X || !A || !Y) return fail(-1, "null pointer");
  {  // the spans [first, last element] of X and Y must not overlap
    const size_t tsz = precision == SPME_FP32 ? sizeof(float) : sizeof(double);
    const char *x0 = (const char *)X, *y0 = (const char *)Y;
    const char *x1 = x0 + ((size_t)(sx - 1) * ldx + n) * tsz;
    const char *y1 = y0 + ((size_t)(sy - 1) * ldy + n) * tsz;
    if (x0 < y1 && y0 < x1) return fail(-1, "X and Y must not alias (overlapping spans)");
  }
  (void)hipGetLastError();
  hipStream_t st = (hipStream_t)stream;
  const int nb = int((n + 255) / 256);
  if (sx <= BS) {
    if (precision == SPME_FP32)
      hipLaunchKernelGGL(k_block_update<float>, dim3(nb), dim3(256), 0, st, n, sx, sy,
                         (const float *)X, ldx, A, beta, (float *)Y, ldy);
    else
      hipLaunchKernelGGL(k_block_update<double>, dim3(nb), dim3(256), 0, st, n, sx, sy,
                         (const double *)X, ldx, A, beta, (double *)Y, ldy);
  } else {
    if (precision == SPME_FP32)
      hipLaunchKernelGGL(k_block_update_tall<float>, dim3(nb), dim3(256), 0, st, n, sx, sy,
                         (const float *)X, ldx, A, beta, (float *)Y, ldy);
    else
      hipLaunchKernelGGL(k_block_update_tall<double>, dim3(nb), dim3(256), 0, st, n, sx, sy,
                         (const double *)X, ldx, A, beta, (double *)Y, ldy);
  }
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : fail(-2, "block update launch: %s", hipGetErrorString(e));
}

}  // extern "C"
