// Hand-written HIP kernels of the SPME reciprocal-space path for gfx950.
//
// Mesh layout (per component, components contiguous, comp_stride reals):
//   real view     r[(z*K + y)*RS + x],  x < K
//   complex view  c[(z*K + y)*Khp + x], x < Kh = K/2+1, Khp = RS/2
// The 3-D r2c / c2r transforms run in place in this layout.
//
// Particles are sorted into buckets keyed by (xy tile of the stencil base,
// z of the stencil base); bucket k holds sorted slots [boff[k], boff[k+1]).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "spme_bodies.h"
#include "spme_kernels.h"
#include "spme_launch.h"

namespace spme {

// ---------------------------------------------------------------------------
// P build (old-version/spme.c:232-297, compute_P): per particle, the wrapped
// stencil base index and the separable centred B-spline weights
// q_d[x] = sum_y W[x][y] g_d^y with g = c - floor(c) - 1/2, c = pos/L*K.
// Weights are FP64 here (cast to the mesh precision when gathered).
__global__ __launch_bounds__(256) void k_pbuild(const double *__restrict__ pos, int N,
                                                double L, int K, int P, WTable W,
                                                int TX, int TY, int *__restrict__ base_tmp,
                                                double *__restrict__ w_tmp,
                                                int *__restrict__ key_tmp,
                                                int *__restrict__ counts, int *__restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int b[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double c = pos[3 * i + d] / L * K;
    if (!(c >= -double(K) && c < 2.0 * K)) {  // also catches NaN
      atomicOr(err, 1);
      c = 0.0;
    }
    const double fl = floor(c);
    int i0 = int(fl) - P / 2 + 1;
    i0 %= K;
    if (i0 < 0) i0 += K;
    b[d] = i0;
    const double g = c - fl - 0.5;
    double pw[8];
    pw[0] = 1.0;
    for (int y = 1; y < P; ++y) pw[y] = g * pw[y - 1];
    for (int x = 0; x < P; ++x) {
      double s = 0.0;
      for (int y = 0; y < P; ++y) s += W.w[x * P + y] * pw[y];
      w_tmp[(size_t)i * 3 * P + d * P + x] = s;
    }
  }
  base_tmp[i] = b[0] | (b[1] << 10) | (b[2] << 20);
  const int key = ((b[1] / TY) * (K / TX) + b[0] / TX) * K + b[2];
  key_tmp[i] = key;
  atomicAdd(&counts[key], 1);
}

// Exclusive scan of n ints into out[0..n] (out[n] = total) in three passes
// over tiles of SCAN_TILE values (1024 threads, 4 consecutive values each):
// tile totals, a one-block scan of the totals, then each tile scanned again
// from its offset.  n <= 1024 * SCAN_TILE.
constexpr int SCAN_TILE = 4096;

// Inclusive scan over the block of one value per thread (1024 threads);
// returns the thread's inclusive prefix, and the block total in *tot.
__device__ __forceinline__ int block_scan_1024(int v, int *s_w, int *tot) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o, 64);
    v += lane >= o ? u : 0;
  }
  if (lane == 63) s_w[w] = v;
  __syncthreads();
  if (threadIdx.x < 16) {
    int x = s_w[threadIdx.x];
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
      const int u = __shfl_up(x, o, 16);
      x += (threadIdx.x & 15) >= o ? u : 0;
    }
    s_w[16 + threadIdx.x] = x;
  }
  __syncthreads();
  *tot = s_w[31];
  return v + (w > 0 ? s_w[16 + w - 1] : 0);
}

__device__ __forceinline__ void scan_load4(const int *__restrict__ in, int n, long i0, int (&v)[4]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = i0 + k < n ? in[i0 + k] : 0;
}

__global__ __launch_bounds__(1024) void k_scan_tiles(const int *__restrict__ in, int n,
                                                     int *__restrict__ sums) {
  __shared__ int s_w[32];
  int v[4];
  scan_load4(in, n, (long)blockIdx.x * SCAN_TILE + 4 * threadIdx.x, v);
  int tot;
  block_scan_1024(v[0] + v[1] + v[2] + v[3], s_w, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// sums[0..nt) -> exclusive offsets in place (nt <= 1024); out[n] = total
__global__ __launch_bounds__(1024) void k_scan_top(int *__restrict__ sums, int nt,
                                                   int *__restrict__ out, int n) {
  __shared__ int s_w[32];
  const int t = threadIdx.x;
  const int x = t < nt ? sums[t] : 0;
  int tot;
  const int inc = block_scan_1024(x, s_w, &tot);
  if (t < nt) sums[t] = inc - x;
  if (t == 0) out[n] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_out(const int *__restrict__ in, int n,
                                                   const int *__restrict__ sums,
                                                   int *__restrict__ out) {
  __shared__ int s_w[32];
  const long i0 = (long)blockIdx.x * SCAN_TILE + 4 * threadIdx.x;
  int v[4];
  scan_load4(in, n, i0, v);
  const int own = v[0] + v[1] + v[2] + v[3];
  int tot;
  int run = block_scan_1024(own, s_w, &tot) - own + sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (i0 + k < n) out[i0 + k] = run;
    run += v[k];
  }
}

__global__ __launch_bounds__(256) void k_pscatter(const int *__restrict__ key,
                                                  const int *__restrict__ boff,
                                                  int *__restrict__ cursor,
                                                  int *__restrict__ perm, int N) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int k = key[i];
  perm[boff[k] + atomicAdd(&cursor[k], 1)] = i;
}

// Order each bucket by original particle index, so the sorted layout (and
// with it every summation order downstream) does not depend on atomics.
// Particle i of a bucket of at most SORT_MAX particles is placed at its
// rank, the number of the bucket's particles with a smaller index (one
// thread per particle, independent loads); a larger bucket (clustered
// input) is queued in big[] for k_bucket_big, so that the cost stays
// bounded however the particles cluster.  Both write out[], a second array.
__global__ __launch_bounds__(256) void k_bucket_rank(const int *__restrict__ key, int N,
                                                     const int *__restrict__ boff,
                                                     const int *__restrict__ perm,
                                                     int *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int k = key[i], a = boff[k], e = boff[k + 1];
  if (e - a > SORT_MAX) return;
  int r = 0;
  for (int m = a; m < e; ++m) r += perm[m] < i;
  out[a + r] = i;
}

__global__ __launch_bounds__(256) void k_bucket_queue(const int *__restrict__ boff, int nkeys,
                                                      int *__restrict__ big,
                                                      int *__restrict__ nbig) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nkeys && boff[k + 1] - boff[k] > SORT_MAX) big[atomicAdd(nbig, 1)] = k;
}

// The queued buckets: block b rebuilds buckets big[b], big[b + gridDim.x],
// ... by one stable pass over all keys (a block-wide prefix count of the
// matches in each 256-index chunk), which lists each bucket's particles in
// increasing original index.  At most N / (SORT_MAX + 1) buckets qualify,
// so the pass costs at most N^2 / (SORT_MAX + 1) key reads.
__global__ __launch_bounds__(256) void k_bucket_big(const int *__restrict__ key, int N,
                                                    const int *__restrict__ boff,
                                                    const int *__restrict__ big,
                                                    const int *__restrict__ nbig,
                                                    int *__restrict__ out) {
  __shared__ int wsum[4];
  const int nb = *nbig;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int q = blockIdx.x; q < nb; q += gridDim.x) {
    const int k = big[q];
    int base = boff[k];
    for (int i0 = 0; i0 < N; i0 += 256) {
      const int i = i0 + threadIdx.x;
      const bool hit = i < N && key[i] == k;
      const unsigned long long m = __ballot(hit);
      if (lane == 0) wsum[w] = __popcll(m);
      __syncthreads();
      int off = 0;
      for (int v = 0; v < w; ++v) off += wsum[v];
      const int tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
      if (hit) out[base + off + __popcll(m & ((1ull << lane) - 1))] = i;
      base += tot;
      __syncthreads();  // wsum is rewritten by the next chunk
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_pgather(const int *__restrict__ perm,
                                                 const int *__restrict__ base_tmp,
                                                 const double *__restrict__ w_tmp,
                                                 int *__restrict__ pbase, T *__restrict__ pw,
                                                 int N, int P3) {
  // thread t: weight j = t % P3 of sorted particle s = t / P3, so that a
  // particle's row of weights is read and written by consecutive lanes
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)N * P3) return;
  const int s = int(t / P3), j = int(t - (long)s * P3);
  const int i = perm[s];
  if (j == 0) pbase[s] = base_tmp[i];
  pw[t] = T(w_tmp[(size_t)i * P3 + j]);
}

// Forces into sorted order, one or two orders in one pass: for spreading,
// particle-major fa[ia[i]*C + 3r + d] = in[r*ldin + 3i + d]; for the
// real-space SpMV (cell order, the right-hand sides of one component
// together) fb[ib[i]*C + d*(C/3) + r].  Either may be null; C <= 48.
// With radii (rad != null, unequal-radius engines, C <= 24) fa takes six
// components per right-hand side, f and a_i^2 f:
// fa[ia[i]*2C + 6r + d] = f, fa[ia[i]*2C + 6r + 3 + d] = a_i^2 f.
// A block stages PB consecutive original particles in LDS (each input row
// read once, coalesced), then writes each particle's C values as one
// contiguous record per order.
// 64 particles per block keep many blocks in flight (at N = 200k, 16 RHS,
// 256 per block took 0.100 ms, 64 take 0.042 ms).
template <typename T> struct PermuteShape {
  static constexpr int PB = 64;      // particles per block
  static constexpr int CP = 48 + 1;  // LDS pitch per particle
};

template <typename T>
__global__ __launch_bounds__(256) void k_permute_in(const int *__restrict__ ia,
                                                    T *__restrict__ fa,
                                                    const int *__restrict__ ib,
                                                    T *__restrict__ fb,
                                                    const T *__restrict__ in, long ldin, int N,
                                                    int C, const double *__restrict__ rad) {
  constexpr int PB = PermuteShape<T>::PB, CP = PermuteShape<T>::CP;
  __shared__ T s[PB * CP];
  const int i0 = blockIdx.x * PB, n = min(PB, N - i0), tid = threadIdx.x;
  for (int r = 0; r < C / 3; ++r) {
    const T *src = in + r * ldin + 3L * i0;
    for (int e = tid; e < 3 * n; e += 256) {
      const int i = e / 3;
      s[i * CP + 3 * r + (e - 3 * i)] = src[e];
    }
  }
  __syncthreads();
  const int R = C / 3;
  for (int e = tid; e < n * C; e += 256) {
    const int i = e / C, c = e - i * C;
    const T v = s[i * CP + c];
    const int r = c / 3, d = c - 3 * r;
    if (fa) {
      if (rad) {
        const double a = rad[i0 + i];
        T *o = fa + (size_t)ia[i0 + i] * (2 * C) + 6 * r + d;
        o[0] = v;
        o[3] = T(a * a) * v;
      } else {
        fa[(size_t)ia[i0 + i] * C + c] = v;
      }
    }
    if (fb) fb[(size_t)ib[i0 + i] * C + d * R + r] = v;
  }
}

// inv[perm[s]] = s
__global__ __launch_bounds__(256) void k_invert(const int *__restrict__ perm,
                                                int *__restrict__ inv, int N) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < N) inv[perm[s]] = s;
}

// Spread: single-system kernel (logical blocks XCD-remapped).
template <typename T, int P, int CC, int TX, int TY>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 3 : 1))) void k_spread(const int *__restrict__ boff,
                                               const int *__restrict__ pbase,
                                               const T *__restrict__ pw,
                                               const T *__restrict__ fs, int C,
                                               T *__restrict__ grid, long comp_stride, int RS,
                                               int K, int ZC) {
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  spread_block<T, P, CC, TX, TY>(boff, pbase, pw, fs, C, grid, comp_stride, RS, K, ZC, lb);
}

// Interpolation: single-system kernel.
template <typename T, int P, int CC, int TXY, int NT, typename TO>
__global__ __launch_bounds__(NT) void k_interp(const int *__restrict__ boff,
                                                const int *__restrict__ pbase,
                                                const T *__restrict__ pw,
                                                const int *__restrict__ perm,
                                                const T *__restrict__ grid, long comp_stride,
                                                int RS, int K, int C, int ZC, TO alpha, TO beta,
                                                TO *__restrict__ out, long ldout) {
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  interp_block<T, P, CC, TXY, NT, TO>(boff, pbase, pw, perm, grid, comp_stride, RS, K, C, ZC,
                                      alpha, beta, out, ldout, lb);
}


// Single-system FFT kernels: one logical block per hardware block.
template <typename T, int K, int NL, int NT>
__global__ __launch_bounds__(NT) void k_fft_x_r2c(T *__restrict__ grid, int RS, long npairs,
                                                  const cpx<T> *__restrict__ tw) {
  __shared__ cpx<T> buf[NL * (K + 1)];
  __shared__ cpx<T> s_tw[K];
  fft_x_r2c_block<T, K, NL, NT>(grid, RS, npairs, tw, blockIdx.x, buf, s_tw);
}

template <typename T, int K, int NL, int NT>
__global__ __launch_bounds__(NT) void k_fft_x_c2r(T *__restrict__ grid, int RS, long npairs,
                                                  const cpx<T> *__restrict__ tw) {
  __shared__ cpx<T> buf[NL * (K + 1)];
  __shared__ cpx<T> s_tw[K];
  fft_x_c2r_block<T, K, NL, NT>(grid, RS, npairs, tw, blockIdx.x, buf, s_tw);
}

template <typename T, int K, int XW, int S, int NT>
__global__ __launch_bounds__(NT) void k_fft_y(T *__restrict__ grid, long comp_stride, int Khp,
                                              int nxc, const cpx<T> *__restrict__ tw) {
  __shared__ cpx<T> buf[XW * (K + 1)];
  __shared__ cpx<T> s_tw[K];
  fft_y_block<T, K, XW, S, NT>(grid, comp_stride, Khp, nxc, tw, xcd_remap(blockIdx.x, gridDim.x),
                               buf, s_tw);
}

// FFT along z alone (the plain path applies the influence in a pass of its own)
template <typename T, int K, int XW, int S, int NT>
__global__ __launch_bounds__(NT) void k_fft_z(T *__restrict__ grid, long comp_stride, int Khp,
                                              int nxc, const cpx<T> *__restrict__ tw) {
  __shared__ cpx<T> buf[XW * (K + 1)];
  __shared__ cpx<T> s_tw[K];
  fft_y_block<T, K, XW, S, NT, true>(grid, comp_stride, Khp, nxc, tw,
                                     xcd_remap(blockIdx.x, gridDim.x), buf, s_tw);
}

// 512-thread blocks are held to 128 VGPRs, so that two blocks fit a CU
template <typename T, int K, int XW, int NT, bool POLY = false>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT == 512 ? 4 : 1)))
void k_fft_z_infl(T *__restrict__ grid, long comp_stride, int Khp, const T *__restrict__ tab,
                  T dk, const cpx<T> *__restrict__ tw, cpx<T> *__restrict__ nyq) {
  __shared__ cpx<T> buf[(POLY ? 6 : 3) * XW * (K + 1)];
  __shared__ cpx<T> s_tw[K];
  fft_z_infl_block<T, K, XW, NT, POLY>(grid, comp_stride, Khp, tab, dk, tw,
                                       xcd_remap(blockIdx.x, gridDim.x), buf, s_tw, nyq);
}

// ---------------------------------------------------------------------------
// Fused x + y transforms (bodies in spme_bodies.h): single-system kernels,
// each block striding over the (component, z) planes.
template <typename T, int K, int NT, bool NYQ>
__global__ __launch_bounds__(NT) void k_fft_xy_fwd(T *__restrict__ grid, int RS, long comp_stride,
                                                   int nplanes, const cpx<T> *__restrict__ tw,
                                                   cpx<T> *__restrict__ nyq) {
  __shared__ cpx<T> buf[XYShape<T, K>::NCPX];
  __shared__ cpx<T> s_tw[K];
  xy_fwd_block<T, K, NT, NYQ>(grid, RS, comp_stride, nplanes, tw, nyq, blockIdx.x, gridDim.x,
                              buf, s_tw);
}

template <typename T, int K, int NT, bool NYQ>
__global__ __launch_bounds__(NT) void k_fft_xy_inv(T *__restrict__ grid, int RS, long comp_stride,
                                                   int nplanes, const cpx<T> *__restrict__ tw,
                                                   const cpx<T> *__restrict__ nyq) {
  __shared__ cpx<T> buf[XYShape<T, K>::NCPX];
  __shared__ cpx<T> s_tw[K];
  xy_inv_block<T, K, NT, NYQ>(grid, RS, comp_stride, nplanes, tw, nyq, blockIdx.x, gridDim.x,
                              buf, s_tw);
}

// ---------------------------------------------------------------------------
// Influence table (old-version/spme.c:128-150), FP64:
//   tab[(z*K + y)*Khp + x] = psi(|k|) |b_x|^2 |b_y|^2 |b_z|^2 / V, x <= K/2,
// psi(k) = 6 pi e^{-k^2/4xi^2} (1 - k^2/3)(1 + k^2/4xi^2 + k^4/8xi^4)/k^2
// (old-version/spme.c:43-60 with unit radii); 0 at k = 0 and in padding.
template <typename T>
__global__ __launch_bounds__(256) void k_influence_table(const double *__restrict__ bsq,
                                                         int K, int Khp, double dk, double xi,
                                                         double inv_v, T *__restrict__ tab) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)K * K * Khp) return;
  const int x = int(t % Khp);
  const long zy = t / Khp;
  const int y = int(zy % K), z = int(zy / K);
  double v = 0.0;
  if (x <= K / 2) {
    const double kx = x * dk, ky = (y <= K / 2 ? y : y - K) * dk;
    const double kz = (z <= K / 2 ? z : z - K) * dk;
    const double k2 = kx * kx + ky * ky + kz * kz;
    if (k2 > 0.0) {
      const double q = k2 / (xi * xi);
      const double psi = 6.0 * M_PI * exp(-q / 4.0) * (1.0 - k2 / 3.0) / k2 *
                         (1.0 + q * (0.25 + q / 8.0));
      v = psi * bsq[x] * bsq[y] * bsq[z] * inv_v;
    }
  }
  tab[t] = T(v);
}

// Factors of the same table for the z pass, FP64-built, stored in mesh
// precision: fac[n] = psi(dk sqrt(n)) inv_v for n = nx^2 + ny^2 + nz^2 <=
// 3 (K/2)^2 (the wave numbers are integer multiples of dk), then |b(m)|^2
// for m < K; the z pass forms psi |b_x|^2 |b_y|^2 |b_z|^2 / V from them.
// radius_free: A(k), psi without its radius term (1 - k^2 (a_i^2 + a_j^2)/6),
// for unequal radii, which the z pass applies per pair of spectra.
template <typename T>
__global__ __launch_bounds__(256) void k_influence_factors(const double *__restrict__ bsq,
                                                           int K, double dk, double xi,
                                                           double inv_v, int radius_free,
                                                           T *__restrict__ fac) {
  const int npsi = influence_npsi(K);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < npsi) {
    const double k2 = double(t) * dk * dk;
    double v = 0.0;
    if (t > 0) {
      const double q = k2 / (xi * xi);
      const double rterm = radius_free ? 1.0 : (1.0 - k2 / 3.0);
      v = 6.0 * M_PI * exp(-q / 4.0) * rterm / k2 * (1.0 + q * (0.25 + q / 8.0)) * inv_v;
    }
    fac[t] = T(v);
  } else if (t < npsi + K) {
    fac[t] = T(bsq[t - npsi]);
  }
}

// ---------------------------------------------------------------------------
// Host launchers.

template <typename T>
void launch_influence_table(const double *bsq, int K, int Khp, double dk, double xi,
                            double inv_v, T *tab, hipStream_t st) {
  const long n = (long)K * K * Khp;
  hipLaunchKernelGGL(k_influence_table<T>, dim3(int((n + 255) / 256)), dim3(256), 0, st, bsq, K,
                     Khp, dk, xi, inv_v, tab);
}
template <typename T>
void launch_influence_factors(const double *bsq, int K, double dk, double xi, double inv_v,
                              int radius_free, T *fac, hipStream_t st) {
  const int n = influence_npsi(K) + K;
  hipLaunchKernelGGL(k_influence_factors<T>, dim3((n + 255) / 256), dim3(256), 0, st, bsq, K, dk,
                     xi, inv_v, radius_free, fac);
}
template void launch_influence_factors<float>(const double *, int, double, double, double, int,
                                              float *, hipStream_t);
template void launch_influence_factors<double>(const double *, int, double, double, double, int,
                                               double *, hipStream_t);
template void launch_influence_table<float>(const double *, int, int, double, double, double,
                                            float *, hipStream_t);
template void launch_influence_table<double>(const double *, int, int, double, double, double,
                                             double *, hipStream_t);

static int grid1(long n, int b) { return int((n + b - 1) / b); }

void launch_pbuild(const double *pos, int N, double L, int K, int P, const WTable &W, int TX,
                   int TY, int *base_tmp, double *w_tmp, int *key_tmp, int *counts, int *err,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_pbuild, dim3(grid1(N, 256)), dim3(256), 0, st, pos, N, L, K, P, W, TX,
                     TY, base_tmp, w_tmp, key_tmp, counts, err);
}

int scan_tiles(int n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

// sums: scan_tiles(n) <= 1024 ints of scratch; -1 if n is too large
int launch_scan(const int *in, int *out, int n, int *sums, hipStream_t st) {
  const int nt = scan_tiles(n);
  if (nt > 1024) return -1;
  if (nt > 0)
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(1024), 0, st, in, n, sums);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, sums, nt, out, n);
  if (nt > 0)
    hipLaunchKernelGGL(k_scan_out, dim3(nt), dim3(1024), 0, st, in, n, sums, out);
  return 0;
}

void launch_pscatter(const int *key, const int *boff, int *cursor, int *perm, int N,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_pscatter, dim3(grid1(N, 256)), dim3(256), 0, st, key, boff, cursor,
                     perm, N);
}

int launch_bucket_sort(const int *key, int N, const int *boff, int *perm, int nkeys, int *bigq,
                       int *tmp, hipStream_t st) {
  // bigq: the queue count, then room for N / (SORT_MAX + 1) + 1 buckets;
  // tmp: N ints, the ordered buckets before they are copied back to perm
  if (hipMemsetAsync(bigq, 0, sizeof(int), st) != hipSuccess) return -1;
  hipLaunchKernelGGL(k_bucket_queue, dim3(grid1(nkeys, 256)), dim3(256), 0, st, boff, nkeys,
                     bigq + 1, bigq);
  hipLaunchKernelGGL(k_bucket_rank, dim3(grid1(N, 256)), dim3(256), 0, st, key, N, boff, perm,
                     tmp);
  hipLaunchKernelGGL(k_bucket_big, dim3(256), dim3(256), 0, st, key, N, boff, bigq + 1, bigq,
                     tmp);
  if (hipMemcpyAsync(perm, tmp, sizeof(int) * N, hipMemcpyDeviceToDevice, st) != hipSuccess)
    return -1;
  return 0;
}

template <typename T>
void launch_pgather(const int *perm, const int *base_tmp, const double *w_tmp, int *pbase,
                    T *pw, int N, int P3, hipStream_t st) {
  hipLaunchKernelGGL(k_pgather<T>, dim3(grid1((long)N * P3, 256)), dim3(256), 0, st, perm,
                     base_tmp, w_tmp, pbase, pw, N, P3);
}

template <typename T>
void launch_permute_in(const int *ia, T *fa, const int *ib, T *fb, const T *in, long ldin, int N,
                       int C, hipStream_t st, const double *rad) {
  hipLaunchKernelGGL(k_permute_in<T>, dim3(grid1(N, PermuteShape<T>::PB)), dim3(256), 0, st, ia,
                     fa, ib, fb, in, ldin, N, C, rad);
}

// Unequal radii: out[r*ldout + 3i + d] = beta*out + alpha*(g[2r][3i + d] -
// a_i^2 g[2r + 1][3i + d]), the interpolated out_f and out_a of the
// influence (each a vector of 3N values at stride 3N).
template <typename T, typename TO>
__global__ __launch_bounds__(256) void k_combine_radii(const T *__restrict__ g,
                                                       const double *__restrict__ rad, int N,
                                                       int nrhs, TO alpha, TO beta,
                                                       TO *__restrict__ out, long ldout) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n3 = 3L * N;
  if (t >= n3 * nrhs) return;
  const int r = int(t / n3);
  const long e = t - r * n3;
  const double a = rad[e / 3];
  const TO v = TO(g[2 * r * n3 + e]) - TO(a * a) * TO(g[(2 * r + 1) * n3 + e]);
  TO *u = out + r * ldout + e;
  *u = beta == TO(0) ? alpha * v : beta * *u + alpha * v;
}

template <typename T, typename TO>
void launch_combine_radii(const T *g, const double *rad, int N, int nrhs, double alpha,
                          double beta, TO *out, long ldout, hipStream_t st) {
  hipLaunchKernelGGL((k_combine_radii<T, TO>), dim3(grid1(3L * N * nrhs, 256)), dim3(256), 0, st,
                     g, rad, N, nrhs, TO(alpha), TO(beta), out, ldout);
}
template void launch_combine_radii<float, float>(const float *, const double *, int, int,
                                                 double, double, float *, long, hipStream_t);
template void launch_combine_radii<float, double>(const float *, const double *, int, int,
                                                  double, double, double *, long, hipStream_t);
template void launch_combine_radii<double, double>(const double *, const double *, int, int,
                                                   double, double, double *, long, hipStream_t);

void launch_invert(const int *perm, int *inv, int N, hipStream_t st) {
  hipLaunchKernelGGL(k_invert, dim3(grid1(N, 256)), dim3(256), 0, st, perm, inv, N);
}

template <typename T, int P, int CC>
static void spread_cc(const SpreadArgs<T> &a, hipStream_t st) {
  constexpr int TX = 8, TY = 8;
  const int nblk = spread_blocks(a.K, a.ZC, a.C, CC);
  hipLaunchKernelGGL((k_spread<T, P, CC, TX, TY>), dim3(nblk), dim3(64), 0, st, a.boff, a.pbase,
                     a.pw, a.fs, a.C, a.grid, a.comp_stride, a.RS, a.K, a.ZC);
}

// Components per pass: the register window holds P*CC values per lane.
template <typename T, int P>
static int spread_p(const SpreadArgs<T> &a, hipStream_t st) {
  if (a.CC == 3) spread_cc<T, P, 3>(a, st);
  else if (a.CC == 6) spread_cc<T, P, 6>(a, st);
  else if (a.CC == 12) spread_cc<T, P, 12>(a, st);
  else return -1;
  return 0;
}

template <typename T>
int launch_spread(const SpreadArgs<T> &a, hipStream_t st) {
  if (a.C % a.CC || a.ZC < 1 || a.ZC > SPREAD_MAXZC) return -1;
  switch (a.P) {
    case 4: return spread_p<T, 4>(a, st);
    case 6: return spread_p<T, 6>(a, st);
    case 8: return spread_p<T, 8>(a, st);
  }
  return -1;
}

// Interpolation shape: 16 bytes of components (4 FP32, 2 FP64) per block
// over a 16 x 16 tile (8 x 8 when K is not a multiple of 16).
template <typename T, int P, int CC, int TXY, typename TO>
static void interp_cc(const InterpArgs<T, TO> &a, hipStream_t st) {
  constexpr int NT = InterpShape<T>::NT;
  const int nblk = interp_blocks(a.K, a.ZC, a.C, CC, TXY);
  hipLaunchKernelGGL((k_interp<T, P, CC, TXY, NT, TO>), dim3(nblk), dim3(NT), 0, st, a.boff,
                     a.pbase, a.pw, a.perm, a.grid, a.comp_stride, a.RS, a.K, a.C, a.ZC, a.alpha,
                     a.beta, a.out, a.ldout);
}

template <typename T, int P, typename TO>
static int interp_p(const InterpArgs<T, TO> &a, hipStream_t st) {
  constexpr int CC = InterpShape<T>::CC;
  if (a.ZC < 1 || a.ZC > InterpShape<T>::MAXZC) return -1;
  if (a.K % 16 == 0) interp_cc<T, P, CC, 16>(a, st);
  else interp_cc<T, P, CC, 8>(a, st);
  return 0;
}

template <typename T, typename TO>
int launch_interp(const InterpArgs<T, TO> &a, hipStream_t st) {
  switch (a.P) {
    case 4: return interp_p<T, 4>(a, st);
    case 6: return interp_p<T, 6>(a, st);
    case 8: return interp_p<T, 8>(a, st);
  }
  return -1;
}

// Fused x + y passes wherever a plane fits in LDS (every K up to 192 in
// FP32, up to 128 in FP64); otherwise separate x and y passes.  The grouped
// kernels run the same plane bodies for K <= 96, so grouped and
// single-system applies of those meshes give the same bits.
template <typename T, int K> constexpr bool xy_fits() { return xy_fused<T>(K); }
// Threads per plane block: enough registers per lane (168 at 768 threads)
// for the next plane's prefetch beside the transform's register tiles.
template <typename T, int K> constexpr int xy_threads() {
  return K >= 192 ? 768 : (K >= 128 ? (sizeof(T) == 8 ? 768 : 512) : 256);
}

template <typename T> bool fft_fused(int K) { return fft_size_supported(K) && xy_fused<T>(K); }
template bool fft_fused<float>(int);
template bool fft_fused<double>(int);

template <typename T, int K>
static void fft3d_k(const FftArgs<T> &a, hipStream_t st, FftMarks *mk) {
  using S = FftShape<T>;
  constexpr int NT = 256;
  const long rows = (long)a.ncomp * K * K;
  const long npairs = rows / 2;
  const int Khp = a.RS / 2;
  const int nxc = (K / 2 + 1 + S::XW - 1) / S::XW;
  using ZS = ZShape<T, K>;
  const int nzb = ZLayout<ZS::XW>::per_rhs(K) * (a.ncomp / 3);
  // unequal radii: blocks of the 6 components of a right-hand side, in the
  // narrow shape (twice the components per block)
  constexpr int XWP = FftShape<T>::XWZ, NTP = 256;
  const int nzbp = ZLayout<XWP>::per_rhs(K) * (a.ncomp / 6);
  auto launch_z = [&](cpx<T> *nyq) {
    if (a.poly)
      hipLaunchKernelGGL((k_fft_z_infl<T, K, XWP, NTP, true>), dim3(nzbp), dim3(NTP), 0, st,
                         a.grid, a.comp_stride, Khp, a.tab, a.dk, a.tw, nyq);
    else
      hipLaunchKernelGGL((k_fft_z_infl<T, K, ZS::XW, ZS::NT>), dim3(nzb), dim3(ZS::NT), 0, st,
                         a.grid, a.comp_stride, Khp, a.tab, a.dk, a.tw, nyq);
  };
  auto mark = [&](int pass) {
    if (mk) {
      (void)hipEventRecord(mk->ev[mk->n], st);
      mk->pass[mk->n++] = pass;
    }
  };
  if constexpr (xy_fits<T, K>()) {
    constexpr int NTXY = xy_threads<T, K>();
    const int nplanes = a.ncomp * K;
    const int per_cu =
        std::max(1, std::min(K >= 128 ? 2 : 4, int((160 * 1024) / XYShape<T, K>::LDS)));
    const int grid = std::min(nplanes, std::max(1, a.ncu) * per_cu);
    if (a.nyq)
      hipLaunchKernelGGL((k_fft_xy_fwd<T, K, NTXY, true>), dim3(grid), dim3(NTXY), 0, st, a.grid,
                         a.RS, a.comp_stride, nplanes, a.tw, a.nyq);
    else
      hipLaunchKernelGGL((k_fft_xy_fwd<T, K, NTXY, false>), dim3(grid), dim3(NTXY), 0, st,
                         a.grid, a.RS, a.comp_stride, nplanes, a.tw, a.nyq);
    mark(0);
    launch_z(a.nyq);
    mark(2);
    if (a.nyq)
      hipLaunchKernelGGL((k_fft_xy_inv<T, K, NTXY, true>), dim3(grid), dim3(NTXY), 0, st, a.grid,
                         a.RS, a.comp_stride, nplanes, a.tw, a.nyq);
    else
      hipLaunchKernelGGL((k_fft_xy_inv<T, K, NTXY, false>), dim3(grid), dim3(NTXY), 0, st,
                         a.grid, a.RS, a.comp_stride, nplanes, a.tw, a.nyq);
    mark(4);
  } else {
    hipLaunchKernelGGL((k_fft_x_r2c<T, K, S::NLX, NT>), dim3(grid1(npairs, S::NLX)), dim3(NT),
                       0, st, a.grid, a.RS, npairs, a.tw);
    mark(0);
    hipLaunchKernelGGL((k_fft_y<T, K, S::XW, -1, NT>), dim3(nxc * K * a.ncomp), dim3(NT), 0,
                       st, a.grid, a.comp_stride, Khp, nxc, a.tw);
    mark(1);
    launch_z(nullptr);
    mark(2);
    hipLaunchKernelGGL((k_fft_y<T, K, S::XW, +1, NT>), dim3(nxc * K * a.ncomp), dim3(NT), 0,
                       st, a.grid, a.comp_stride, Khp, nxc, a.tw);
    mark(3);
    hipLaunchKernelGGL((k_fft_x_c2r<T, K, S::NLX, NT>), dim3(grid1(npairs, S::NLX)), dim3(NT),
                       0, st, a.grid, a.RS, npairs, a.tw);
    mark(4);
  }
}

template <typename T>
int launch_fft_influence(const FftArgs<T> &a, hipStream_t st, FftMarks *mk) {
  switch (a.K) {
    case 16: fft3d_k<T, 16>(a, st, mk); return 0;
    case 24: fft3d_k<T, 24>(a, st, mk); return 0;
    case 32: fft3d_k<T, 32>(a, st, mk); return 0;
    case 48: fft3d_k<T, 48>(a, st, mk); return 0;
    case 64: fft3d_k<T, 64>(a, st, mk); return 0;
    case 96: fft3d_k<T, 96>(a, st, mk); return 0;
    case 128: fft3d_k<T, 128>(a, st, mk); return 0;
    case 192: fft3d_k<T, 192>(a, st, mk); return 0;
    case 256: fft3d_k<T, 256>(a, st, mk); return 0;
  }
  return -1;
}

// Forward (dir -1: x r2c, y, z) or backward (dir +1: z, y, x c2r) transform
// of a.ncomp components, unnormalised, no influence.
template <typename T, int K>
static void fft3d_plain_k(const FftArgs<T> &a, int dir, hipStream_t st) {
  using S = FftShape<T>;
  constexpr int NT = 256;
  const long npairs = (long)a.ncomp * K * K / 2;
  const int Khp = a.RS / 2;
  const int nxc = (K / 2 + 1 + S::XW - 1) / S::XW;
  const dim3 gyz(nxc * K * a.ncomp);
  if (dir < 0) {
    hipLaunchKernelGGL((k_fft_x_r2c<T, K, S::NLX, NT>), dim3(grid1(npairs, S::NLX)), dim3(NT),
                       0, st, a.grid, a.RS, npairs, a.tw);
    hipLaunchKernelGGL((k_fft_y<T, K, S::XW, -1, NT>), gyz, dim3(NT), 0, st, a.grid,
                       a.comp_stride, Khp, nxc, a.tw);
    hipLaunchKernelGGL((k_fft_z<T, K, S::XW, -1, NT>), gyz, dim3(NT), 0, st, a.grid,
                       a.comp_stride, Khp, nxc, a.tw);
  } else {
    hipLaunchKernelGGL((k_fft_z<T, K, S::XW, +1, NT>), gyz, dim3(NT), 0, st, a.grid,
                       a.comp_stride, Khp, nxc, a.tw);
    hipLaunchKernelGGL((k_fft_y<T, K, S::XW, +1, NT>), gyz, dim3(NT), 0, st, a.grid,
                       a.comp_stride, Khp, nxc, a.tw);
    hipLaunchKernelGGL((k_fft_x_c2r<T, K, S::NLX, NT>), dim3(grid1(npairs, S::NLX)), dim3(NT),
                       0, st, a.grid, a.RS, npairs, a.tw);
  }
}

template <typename T>
int launch_fft_plain(const FftArgs<T> &a, int dir, hipStream_t st) {
  switch (a.K) {
    case 16: fft3d_plain_k<T, 16>(a, dir, st); return 0;
    case 24: fft3d_plain_k<T, 24>(a, dir, st); return 0;
    case 32: fft3d_plain_k<T, 32>(a, dir, st); return 0;
    case 48: fft3d_plain_k<T, 48>(a, dir, st); return 0;
    case 64: fft3d_plain_k<T, 64>(a, dir, st); return 0;
    case 96: fft3d_plain_k<T, 96>(a, dir, st); return 0;
    case 128: fft3d_plain_k<T, 128>(a, dir, st); return 0;
    case 192: fft3d_plain_k<T, 192>(a, dir, st); return 0;
    case 256: fft3d_plain_k<T, 256>(a, dir, st); return 0;
  }
  return -1;
}
template int launch_fft_plain<double>(const FftArgs<double> &, int, hipStream_t);

bool fft_size_supported(int K) {
  return K == 16 || K == 24 || K == 32 || K == 48 || K == 64 || K == 96 || K == 128 ||
         K == 192 || K == 256;
}

template void launch_pgather<float>(const int *, const int *, const double *, int *, float *,
                                    int, int, hipStream_t);
template void launch_pgather<double>(const int *, const int *, const double *, int *, double *,
                                     int, int, hipStream_t);
template void launch_permute_in<float>(const int *, float *, const int *, float *, const float *,
                                       long, int, int, hipStream_t, const double *);
template void launch_permute_in<double>(const int *, double *, const int *, double *,
                                        const double *, long, int, int, hipStream_t,
                                        const double *);
template int launch_spread<float>(const SpreadArgs<float> &, hipStream_t);
template int launch_spread<double>(const SpreadArgs<double> &, hipStream_t);
template int launch_interp<float, float>(const InterpArgs<float, float> &, hipStream_t);
template int launch_interp<float, double>(const InterpArgs<float, double> &, hipStream_t);
template int launch_interp<double, double>(const InterpArgs<double, double> &, hipStream_t);
template int launch_fft_influence<float>(const FftArgs<float> &, hipStream_t, FftMarks *);
template int launch_fft_influence<double>(const FftArgs<double> &, hipStream_t, FftMarks *);

}  // namespace spme
