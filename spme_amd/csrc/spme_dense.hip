// Plain FP64 reciprocal path with the reference's data structures and call
// pattern (old-version/spme.c:158-324,560-726; old-version/spme_kernel.c):
// per particle a dense row of p^3 FP64 weights and p^3 int32 mesh offsets,
// spreading by atomic adds into a zeroed mesh, one right-hand side at a time,
// the influence and the 1/K^3 scale as passes of their own, interpolation by
// gathering through the offsets with alpha' = alpha K^3 / L^3.
//
// This is the yardstick the compact and batched paths are measured against:
// the reference algorithm written competently for the MI355X (one wave per
// particle, coalesced weight/offset rows, hardware FP64 atomics, the same
// FFT kernels as the compact path) rather than translated.
#include <hip/hip_runtime.h>

#include "spme_kernels.h"
#include "spme_launch.h"

namespace spme {

namespace {
constexpr int DW = 4;  // particles (waves) per 256-thread block

// Wave-wide sum (64 lanes).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
}  // namespace

// compute_P (old-version/spme.c:281-295): row i holds
//   Pm[i*ldP + (z*p + y)*p + x]  = q1[x] q2[y] q3[z]
//   ind[i*ldind + (z*p + y)*p + x] = (zz*K + yy)*RS + xx,
// zz = (z + base_z) mod K etc., from the separable weights w[i][3][p] and
// the packed base (bx | by << 10 | bz << 20) of the P build.
template <int P>
__global__ __launch_bounds__(256) void k_dense_P(const int *__restrict__ base,
                                                 const double *__restrict__ w, int N, int K,
                                                 int RS, int ldP, int ldind,
                                                 double *__restrict__ Pm, int *__restrict__ ind) {
  const int i = blockIdx.x * DW + threadIdx.x / 64, lane = threadIdx.x & 63;
  if (i >= N) return;
  const int b = base[i];
  const int b0 = b & 1023, b1 = (b >> 10) & 1023, b2 = (b >> 20) & 1023;
  const double *wi = w + (size_t)i * 3 * P;
  for (int j = lane; j < P * P * P; j += 64) {
    const int x = j % P, y = (j / P) % P, z = j / (P * P);
    int xx = b0 + x, yy = b1 + y, zz = b2 + z;
    if (xx >= K) xx -= K;
    if (yy >= K) yy -= K;
    if (zz >= K) zz -= K;
    Pm[(size_t)i * ldP + j] = wi[x] * wi[P + y] * wi[2 * P + z];
    ind[(size_t)i * ldind + j] = (zz * K + yy) * RS + xx;
  }
}

// spread (old-version/spme_kernel.c:231-312): grid[c][ind] += P * f[3i + c]
// for the 3 components, by hardware FP64 atomic adds (the reference's GPU
// variant adds with compare-and-swap loops).  The mesh is zeroed before.
template <int P>
__global__ __launch_bounds__(256) void k_dense_spread(const double *__restrict__ Pm,
                                                      const int *__restrict__ ind, int N,
                                                      int ldP, int ldind,
                                                      const double *__restrict__ f,
                                                      double *__restrict__ grid, long cs) {
  const int i = blockIdx.x * DW + threadIdx.x / 64, lane = threadIdx.x & 63;
  if (i >= N) return;
  const double f0 = f[3 * i], f1 = f[3 * i + 1], f2 = f[3 * i + 2];
  for (int j = lane; j < P * P * P; j += 64) {
    const double w = Pm[(size_t)i * ldP + j];
    const int o = ind[(size_t)i * ldind + j];
    unsafeAtomicAdd(grid + o, w * f0);
    unsafeAtomicAdd(grid + cs + o, w * f1);
    unsafeAtomicAdd(grid + 2 * cs + o, w * f2);
  }
}

// apply_influence (old-version/spme_kernel.c:73-228): for every half-spectrum
// point, g_c <- sum_d lm2 (delta_cd - k_c k_d / k^2) g_d; g(0) = 0.
// lm2[(z*K + y)*Khp + x] = psi(|k|) |b(k)|^2 (zero at k = 0 and for x > K/2).
__global__ __launch_bounds__(256) void k_dense_influence(double *__restrict__ grid, long cs,
                                                         int K, int Khp,
                                                         const double *__restrict__ lm2,
                                                         double dk) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)K * K * Khp) return;
  const int x = int(t % Khp);
  if (x > K / 2) return;
  const long zy = t / Khp;
  const int y = int(zy % K), z = int(zy / K);
  const double kx = x * dk, ky = (y <= K / 2 ? y : y - K) * dk;
  const double kz = (z <= K / 2 ? z : z - K) * dk;
  const double k2 = kx * kx + ky * ky + kz * kz;
  const double m = lm2[t];
  const double ik = k2 > 0.0 ? 1.0 / k2 : 0.0;
  const double B0 = m - m * kx * kx * ik, B1 = -m * kx * ky * ik, B2 = -m * kx * kz * ik;
  const double B3 = m - m * ky * ky * ik, B4 = -m * ky * kz * ik, B5 = m - m * kz * kz * ik;
  double2 *g = reinterpret_cast<double2 *>(grid);
  const long h = cs / 2;
  const double2 va = g[t], vb = g[h + t], vc = g[2 * h + t];
  g[t] = {B0 * va.x + B1 * vb.x + B2 * vc.x, B0 * va.y + B1 * vb.y + B2 * vc.y};
  g[h + t] = {B1 * va.x + B3 * vb.x + B4 * vc.x, B1 * va.y + B3 * vb.y + B4 * vc.y};
  g[2 * h + t] = {B2 * va.x + B4 * vb.x + B5 * vc.x, B2 * va.y + B4 * vb.y + B5 * vc.y};
}

// The backward transform's 1/K^3 (the reference's scaling pass), over the
// whole 3-component mesh.
__global__ __launch_bounds__(256) void k_dense_scale(double *__restrict__ grid, long n,
                                                     double s) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) grid[t] *= s;
}

// interpolate (old-version/spme_kernel.c:8-31): u[3i + c] = beta u[3i + c]
// + alpha' sum_j P[i][j] grid[c][ind[i][j]] (beta == 0 does not read u).
template <int P>
__global__ __launch_bounds__(256) void k_dense_interp(const double *__restrict__ Pm,
                                                      const int *__restrict__ ind, int N,
                                                      int ldP, int ldind,
                                                      const double *__restrict__ grid, long cs,
                                                      double alpha, double beta,
                                                      double *__restrict__ u) {
  const int i = blockIdx.x * DW + threadIdx.x / 64, lane = threadIdx.x & 63;
  if (i >= N) return;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int j = lane; j < P * P * P; j += 64) {
    const double w = Pm[(size_t)i * ldP + j];
    const int o = ind[(size_t)i * ldind + j];
    a0 += grid[o] * w;
    a1 += grid[cs + o] * w;
    a2 += grid[2 * cs + o] * w;
  }
  a0 = wave_sum(a0);
  a1 = wave_sum(a1);
  a2 = wave_sum(a2);
  if (lane < 3) {
    const double a = lane == 0 ? a0 : (lane == 1 ? a1 : a2);
    double *ui = u + 3L * i + lane;
    *ui = (beta == 0.0 ? 0.0 : beta * *ui) + alpha * a;
  }
}

// ---------------------------------------------------------------------------
static int dense_blocks(int N) { return (N + DW - 1) / DW; }

int launch_dense_P(const int *base, const double *w, int N, int K, int RS, int P, int ldP,
                   int ldind, double *Pm, int *ind, hipStream_t st) {
  switch (P) {
#define X(PP)                                                                                \
  case PP:                                                                                   \
    hipLaunchKernelGGL(k_dense_P<PP>, dim3(dense_blocks(N)), dim3(256), 0, st, base, w, N, K, \
                       RS, ldP, ldind, Pm, ind);                                             \
    return 0;
    X(4) X(6) X(8)
#undef X
  }
  return -1;
}

int launch_dense_spread(const double *Pm, const int *ind, int N, int P, int ldP, int ldind,
                        const double *f, double *grid, long cs, hipStream_t st) {
  switch (P) {
#define X(PP)                                                                         \
  case PP:                                                                            \
    hipLaunchKernelGGL(k_dense_spread<PP>, dim3(dense_blocks(N)), dim3(256), 0, st, Pm, \
                       ind, N, ldP, ldind, f, grid, cs);                              \
    return 0;
    X(4) X(6) X(8)
#undef X
  }
  return -1;
}

void launch_dense_influence(double *grid, long cs, int K, int Khp, const double *lm2,
                            double dk, hipStream_t st) {
  const long n = (long)K * K * Khp;
  hipLaunchKernelGGL(k_dense_influence, dim3(int((n + 255) / 256)), dim3(256), 0, st, grid, cs,
                     K, Khp, lm2, dk);
}

void launch_dense_scale(double *grid, long n, double s, hipStream_t st) {
  hipLaunchKernelGGL(k_dense_scale, dim3(int((n + 255) / 256)), dim3(256), 0, st, grid, n, s);
}

int launch_dense_interp(const double *Pm, const int *ind, int N, int P, int ldP, int ldind,
                        const double *grid, long cs, double alpha, double beta, double *u,
                        hipStream_t st) {
  switch (P) {
#define X(PP)                                                                         \
  case PP:                                                                            \
    hipLaunchKernelGGL(k_dense_interp<PP>, dim3(dense_blocks(N)), dim3(256), 0, st, Pm, \
                       ind, N, ldP, ldind, grid, cs, alpha, beta, u);                 \
    return 0;
    X(4) X(6) X(8)
#undef X
  }
  return -1;
}

}  // namespace spme
