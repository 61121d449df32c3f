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

// Exclusive scan of n ints into out[0..n] (out[n] = total), one block.
__global__ __launch_bounds__(1024) void k_scan(const int *__restrict__ in, int *__restrict__ out,
                                               int n) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int chunk = (n + 1023) / 1024;
  const int a = min(n, t * chunk), e = min(n, a + chunk);
  int s = 0;
  for (int j = a; j < e; ++j) s += in[j];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - s;
  for (int j = a; j < e; ++j) {
    out[j] = run;
    run += in[j];
  }
  if (t == 1023) out[n] = part[1023];
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
__global__ __launch_bounds__(256) void k_bucket_sort(const int *__restrict__ boff,
                                                     int *__restrict__ perm, int nkeys) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nkeys) return;
  const int a = boff[k], e = boff[k + 1];
  for (int j = a + 1; j < e; ++j) {
    const int v = perm[j];
    int m = j - 1;
    while (m >= a && perm[m] > v) {
      perm[m + 1] = perm[m];
      --m;
    }
    perm[m + 1] = v;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_pgather(const int *__restrict__ perm,
                                                 const int *__restrict__ base_tmp,
                                                 const double *__restrict__ w_tmp,
                                                 int *__restrict__ pbase, T *__restrict__ pw,
                                                 int N, int P3) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const int i = perm[s];
  pbase[s] = base_tmp[i];
  for (int j = 0; j < P3; ++j) pw[(size_t)s * P3 + j] = T(w_tmp[(size_t)i * P3 + j]);
}

// fs[s*C + 3r + d] = in[r*ldin + 3*perm[s] + d]  (sorted, particle-major).
template <typename T>
__global__ __launch_bounds__(256) void k_permute_in(const int *__restrict__ perm,
                                                    const T *__restrict__ in, long ldin,
                                                    T *__restrict__ fs, int N, int C) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)N * C) return;
  const int s = int(t / C), c = int(t - (long)s * C);
  const int r = c / 3, d = c - 3 * r;
  fs[t] = in[r * ldin + 3L * perm[s] + d];
}

// ---------------------------------------------------------------------------
// Spread (old-version/spme_kernel.c:231-312): grid[c] = sum_i q_i (x) f_ic.
// One wave owns a TX x TY column of cells over a z range and keeps a window
// of P planes per component in registers; particles are visited in
// increasing base-z, and a plane is final (and stored) once the sweep has
// passed every base that reaches it.  Every cell is written exactly once, so
// the mesh needs no clearing and no atomics are used.
template <typename T, int P, int CC, int TX, int TY>
__global__ __launch_bounds__(64) void k_spread(const int *__restrict__ boff,
                                               const int *__restrict__ pbase,
                                               const T *__restrict__ pw,
                                               const T *__restrict__ fs, int C,
                                               T *__restrict__ grid, long comp_stride, int RS,
                                               int K, int ZC) {
  static_assert(TX * TY == 64, "one wave per tile");
  constexpr int NXR = 1 + (P + TX - 2) / TX;
  constexpr int NYR = 1 + (P + TY - 2) / TY;
  const int ntx = K / TX, nty = K / TY;
  const int tx = blockIdx.x % ntx, ty = blockIdx.x / ntx;
  const int lane = threadIdx.x;
  const int x = tx * TX + lane % TX, y = ty * TY + lane / TX;
  const int z0 = blockIdx.y * ZC, zend = min(z0 + ZC, K);
  const int cc0 = blockIdx.z * CC;
  const int zs = z0 - (P - 1);
  const int nsteps = zend - zs;

  T acc[P][CC];
#pragma unroll
  for (int j = 0; j < P; ++j)
#pragma unroll
    for (int c = 0; c < CC; ++c) acc[j][c] = T(0);

  for (int s0 = 0; s0 < nsteps; s0 += P) {
#pragma unroll
    for (int ph = 0; ph < P; ++ph) {
      const int s = s0 + ph;
      if (s < nsteps) {
        int bz = zs + s;
        if (bz < 0) bz += K;
#pragma unroll
        for (int a = 0; a < NYR; ++a) {
          int tyy = ty - a;
          if (tyy < 0) tyy += nty;
#pragma unroll
          for (int b = 0; b < NXR; ++b) {
            int txx = tx - b;
            if (txx < 0) txx += ntx;
            const int key = (tyy * ntx + txx) * K + bz;
            const int i0 = boff[key], i1 = boff[key + 1];
            for (int i = i0; i < i1; ++i) {
              const int pb = pbase[i];
              int dx = x - (pb & 1023);
              dx += dx < 0 ? K : 0;
              int dy = y - ((pb >> 10) & 1023);
              dy += dy < 0 ? K : 0;
              if (dx < P && dy < P) {
                const T *w = pw + (size_t)i * 3 * P;
                T wx = w[0], wy = w[P];
#pragma unroll
                for (int j = 1; j < P; ++j) {
                  wx = dx == j ? w[j] : wx;
                  wy = dy == j ? w[P + j] : wy;
                }
                const T wxy = wx * wy;
                T wf[CC];
#pragma unroll
                for (int c = 0; c < CC; ++c) wf[c] = wxy * fs[(size_t)i * C + cc0 + c];
#pragma unroll
                for (int j = 0; j < P; ++j) {
                  const T wz = w[2 * P + j];
#pragma unroll
                  for (int c = 0; c < CC; ++c) acc[(ph + j) % P][c] += wz * wf[c];
                }
              }
            }
          }
        }
        if (s >= P - 1) {
          T *g = grid + ((size_t)(zs + s) * K + y) * RS + x + (size_t)cc0 * comp_stride;
#pragma unroll
          for (int c = 0; c < CC; ++c) g[(size_t)c * comp_stride] = acc[ph][c];
        }
#pragma unroll
        for (int c = 0; c < CC; ++c) acc[ph][c] = T(0);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Interpolation (old-version/spme_kernel.c:8-70):
//   out[r*ldout + 3*i + d] = beta*out + alpha * sum_j q_i(j) grid[3r+d](j).
// One wave per particle; lane (dx, dy) walks the P planes of its stencil
// column, then the wave reduces.
template <typename T, int P, int CC>
__global__ __launch_bounds__(256) void k_interp(const int *__restrict__ pbase,
                                                const T *__restrict__ pw,
                                                const int *__restrict__ perm,
                                                const T *__restrict__ grid, long comp_stride,
                                                int RS, int K, int N, int C, T alpha, T beta,
                                                T *__restrict__ out, long ldout) {
  const int lane = threadIdx.x & 63;
  const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (s >= N) return;
  const int pb = __builtin_amdgcn_readfirstlane(pbase[s]);
  const int bx = pb & 1023, by = (pb >> 10) & 1023, bz = pb >> 20;
  const int dx = lane % P, dy = lane / P;
  const bool act = lane < P * P;
  int x = bx + dx;
  if (x >= K) x -= K;
  int y = by + dy;
  if (y >= K) y -= K;
  const T *w = pw + (size_t)s * 3 * P;
  const T wxy = act ? w[dx] * w[P + dy] : T(0);
  const int o = __builtin_amdgcn_readfirstlane(perm[s]);
  for (int c0 = 0; c0 < C; c0 += CC) {
    T acc[CC];
#pragma unroll
    for (int c = 0; c < CC; ++c) acc[c] = T(0);
    if (act) {
#pragma unroll
      for (int j = 0; j < P; ++j) {
        int z = bz + j;
        if (z >= K) z -= K;
        const T wz = w[2 * P + j];
        const T *g = grid + ((size_t)z * K + y) * RS + x + (size_t)c0 * comp_stride;
#pragma unroll
        for (int c = 0; c < CC; ++c) acc[c] += wz * g[(size_t)c * comp_stride];
      }
    }
#pragma unroll
    for (int c = 0; c < CC; ++c) {
      const T v = wave_sum(acc[c] * wxy);
      if (lane == 0) {
        const int comp = c0 + c, r = comp / 3, d = comp - 3 * r;
        T *u = out + r * ldout + 3L * o + d;
        *u = beta == T(0) ? alpha * v : beta * *u + alpha * v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// FFT passes.  tw[t] = exp(-2 pi i t/K), t < K.

template <typename T, int K>
__device__ __forceinline__ void load_tw(cpx<T> *s_tw, const cpx<T> *tw) {
  for (int t = threadIdx.x; t < K; t += blockDim.x) s_tw[t] = tw[t];
}

// Forward r2c along x: two real rows per complex line (z = a + i b).
template <typename T, int K, int NL, int NT>
__global__ __launch_bounds__(NT) void k_fft_x_r2c(T *__restrict__ grid, int RS, long npairs,
                                                  const cpx<T> *__restrict__ tw) {
  constexpr int Kh = K / 2 + 1;
  __shared__ cpx<T> buf[NL * K];
  __shared__ cpx<T> s_tw[K];
  load_tw<T, K>(s_tw, tw);
  const long q0 = (long)blockIdx.x * NL;
  for (int t = threadIdx.x; t < NL * K; t += NT) {
    const int l = t / K, n = t - l * K;
    const long q = q0 + l;
    cpx<T> v = {T(0), T(0)};
    if (q < npairs) {
      const T *ra = grid + (2 * q) * RS;
      v = {ra[n], ra[RS + n]};
    }
    buf[t] = v;
  }
  __syncthreads();
  lds_fft<T, K, -1, NL, K, NT>(buf, s_tw);
  for (int t = threadIdx.x; t < NL * Kh; t += NT) {
    const int l = t / Kh, k = t - l * Kh;
    const long q = q0 + l;
    if (q < npairs) {
      const cpx<T> Z = buf[l * K + k];
      const cpx<T> Zm = cconj(buf[l * K + (k == 0 ? 0 : K - k)]);
      const cpx<T> A = {T(0.5) * (Z.x + Zm.x), T(0.5) * (Z.y + Zm.y)};
      const cpx<T> D = csub(Z, Zm);
      const cpx<T> B = {T(0.5) * D.y, T(-0.5) * D.x};
      cpx<T> *ca = reinterpret_cast<cpx<T> *>(grid + (2 * q) * RS);
      ca[k] = A;
      ca[RS / 2 + k] = B;
    }
  }
}

// Backward c2r along x.  The imaginary parts of the k = 0 and k = K/2
// entries are dropped, so the result is the real part of the inverse DFT.
template <typename T, int K, int NL, int NT>
__global__ __launch_bounds__(NT) void k_fft_x_c2r(T *__restrict__ grid, int RS, long npairs,
                                                  const cpx<T> *__restrict__ tw) {
  constexpr int Kh = K / 2 + 1;
  __shared__ cpx<T> buf[NL * K];
  __shared__ cpx<T> s_tw[K];
  load_tw<T, K>(s_tw, tw);
  const long q0 = (long)blockIdx.x * NL;
  for (int t = threadIdx.x; t < NL * Kh; t += NT) {
    const int l = t / Kh, k = t - l * Kh;
    const long q = q0 + l;
    cpx<T> A = {T(0), T(0)}, B = {T(0), T(0)};
    if (q < npairs) {
      const cpx<T> *ca = reinterpret_cast<const cpx<T> *>(grid + (2 * q) * RS);
      A = ca[k];
      B = ca[RS / 2 + k];
    }
    if (k == 0 || k == K / 2) {
      A.y = T(0);
      B.y = T(0);
    }
    buf[l * K + k] = {A.x - B.y, A.y + B.x};
    if (k > 0 && k < K / 2) buf[l * K + K - k] = {A.x + B.y, B.x - A.y};
  }
  __syncthreads();
  lds_fft<T, K, +1, NL, K, NT>(buf, s_tw);
  for (int t = threadIdx.x; t < NL * K; t += NT) {
    const int l = t / K, n = t - l * K;
    const long q = q0 + l;
    if (q < npairs) {
      T *ra = grid + (2 * q) * RS;
      const cpx<T> v = buf[t];
      ra[n] = v.x;
      ra[RS + n] = v.y;
    }
  }
}

// Complex FFT along y, in place.  Block = (x chunk of XW, plane z, comp).
template <typename T, int K, int XW, int S, int NT>
__global__ __launch_bounds__(NT) void k_fft_y(T *__restrict__ grid, long comp_stride, int Khp,
                                              int nxc, const cpx<T> *__restrict__ tw) {
  constexpr int Kh = K / 2 + 1;
  constexpr int LS = K + 1;
  __shared__ cpx<T> buf[XW * LS];
  __shared__ cpx<T> s_tw[K];
  load_tw<T, K>(s_tw, tw);
  const int xc = blockIdx.x % nxc;
  const int zc = blockIdx.x / nxc;  // z + K*comp
  const int comp = zc / K, z = zc - comp * K;
  const int x0 = xc * XW;
  cpx<T> *plane = reinterpret_cast<cpx<T> *>(grid + comp * comp_stride) + (size_t)z * K * Khp;
  for (int t = threadIdx.x; t < XW * K; t += NT) {
    const int y = t / XW, xo = t - y * XW;
    const int x = x0 + xo;
    buf[xo * LS + y] = x < Kh ? plane[(size_t)y * Khp + x] : cpx<T>{T(0), T(0)};
  }
  __syncthreads();
  lds_fft<T, K, S, XW, LS, NT>(buf, s_tw);
  for (int t = threadIdx.x; t < XW * K; t += NT) {
    const int y = t / XW, xo = t - y * XW;
    const int x = x0 + xo;
    if (x < Kh) plane[(size_t)y * Khp + x] = buf[xo * LS + y];
  }
}

// Forward FFT along z, RPY influence (old-version/spme_kernel.c:73-228),
// backward FFT along z, for the 3 components of one right-hand side.
// Block = (x chunk, y) x rhs.  tab[(z*K + y)*Khp + x] = psi(|k|)|b(k)|^2/V.
// Wave numbers follow old-version/spme.c:95-107 (index K/2 -> +pi K/L).
template <typename T, int K, int XW, int NT>
__global__ __launch_bounds__(NT) void k_fft_z_infl(T *__restrict__ grid, long comp_stride,
                                                   int Khp, int nxc, const T *__restrict__ tab,
                                                   T dk, const cpx<T> *__restrict__ tw) {
  constexpr int Kh = K / 2 + 1;
  constexpr int LS = K + 1;
  constexpr int NLN = 3 * XW;
  __shared__ cpx<T> buf[NLN * LS];
  __shared__ cpx<T> s_tw[K];
  __shared__ T s_tab[XW * K];
  load_tw<T, K>(s_tw, tw);
  const int xc = blockIdx.x % nxc, y = blockIdx.x / nxc;
  const int x0 = xc * XW;
  const int rhs = blockIdx.y;
  for (int t = threadIdx.x; t < XW * K; t += NT) {
    const int z = t / XW, xo = t - z * XW;
    const int x = x0 + xo;
    s_tab[xo * K + z] = x < Kh ? tab[((size_t)z * K + y) * Khp + x] : T(0);
  }
  cpx<T> *g0 = reinterpret_cast<cpx<T> *>(grid + (3L * rhs) * comp_stride);
  const long cs = comp_stride / 2;  // complex elements
  for (int t = threadIdx.x; t < 3 * XW * K; t += NT) {
    const int c = t / (XW * K), r = t - c * XW * K;
    const int z = r / XW, xo = r - z * XW;
    const int x = x0 + xo;
    buf[(c * XW + xo) * LS + z] =
        x < Kh ? g0[c * cs + ((size_t)z * K + y) * Khp + x] : cpx<T>{T(0), T(0)};
  }
  __syncthreads();
  lds_fft<T, K, -1, NLN, LS, NT>(buf, s_tw);
  const T ky = T(y <= K / 2 ? y : y - K) * dk;
  for (int t = threadIdx.x; t < XW * K; t += NT) {
    const int xo = t / K, z = t - xo * K;
    const int x = x0 + xo;
    const T kx = T(x) * dk;  // x <= K/2 in the half spectrum
    const T kz = T(z <= K / 2 ? z : z - K) * dk;
    const T k2 = kx * kx + ky * ky + kz * kz;
    cpx<T> *a = buf + (0 * XW + xo) * LS + z;
    cpx<T> *b = buf + (1 * XW + xo) * LS + z;
    cpx<T> *c = buf + (2 * XW + xo) * LS + z;
    if (k2 == T(0) || x >= Kh) {
      *a = {T(0), T(0)};
      *b = {T(0), T(0)};
      *c = {T(0), T(0)};
      continue;
    }
    const T m2 = s_tab[xo * K + z];
    const T ik = T(1) / k2;
    const T B0 = m2 - m2 * kx * kx * ik, B1 = -m2 * kx * ky * ik, B2 = -m2 * kx * kz * ik;
    const T B3 = m2 - m2 * ky * ky * ik, B4 = -m2 * ky * kz * ik, B5 = m2 - m2 * kz * kz * ik;
    const cpx<T> va = *a, vb = *b, vc = *c;
    *a = {B0 * va.x + B1 * vb.x + B2 * vc.x, B0 * va.y + B1 * vb.y + B2 * vc.y};
    *b = {B1 * va.x + B3 * vb.x + B4 * vc.x, B1 * va.y + B3 * vb.y + B4 * vc.y};
    *c = {B2 * va.x + B4 * vb.x + B5 * vc.x, B2 * va.y + B4 * vb.y + B5 * vc.y};
  }
  __syncthreads();
  lds_fft<T, K, +1, NLN, LS, NT>(buf, s_tw);
  for (int t = threadIdx.x; t < 3 * XW * K; t += NT) {
    const int c = t / (XW * K), r = t - c * XW * K;
    const int z = r / XW, xo = r - z * XW;
    const int x = x0 + xo;
    if (x < Kh) g0[c * cs + ((size_t)z * K + y) * Khp + x] = buf[(c * XW + xo) * LS + z];
  }
}

// ---------------------------------------------------------------------------
// Host launchers.

static int grid1(long n, int b) { return int((n + b - 1) / b); }

void launch_pbuild(const double *pos, int N, double L, int K, int P, const WTable &W, int TX,
                   int TY, int *base_tmp, double *w_tmp, int *key_tmp, int *counts, int *err,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_pbuild, dim3(grid1(N, 256)), dim3(256), 0, st, pos, N, L, K, P, W, TX,
                     TY, base_tmp, w_tmp, key_tmp, counts, err);
}

void launch_scan(const int *in, int *out, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, in, out, n);
}

void launch_pscatter(const int *key, const int *boff, int *cursor, int *perm, int N,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_pscatter, dim3(grid1(N, 256)), dim3(256), 0, st, key, boff, cursor,
                     perm, N);
}

void launch_bucket_sort(const int *boff, int *perm, int nkeys, hipStream_t st) {
  hipLaunchKernelGGL(k_bucket_sort, dim3(grid1(nkeys, 256)), dim3(256), 0, st, boff, perm,
                     nkeys);
}

template <typename T>
void launch_pgather(const int *perm, const int *base_tmp, const double *w_tmp, int *pbase,
                    T *pw, int N, int P3, hipStream_t st) {
  hipLaunchKernelGGL(k_pgather<T>, dim3(grid1(N, 256)), dim3(256), 0, st, perm, base_tmp, w_tmp,
                     pbase, pw, N, P3);
}

template <typename T>
void launch_permute_in(const int *perm, const T *in, long ldin, T *fs, int N, int C,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_permute_in<T>, dim3(grid1((long)N * C, 256)), dim3(256), 0, st, perm, in,
                     ldin, fs, N, C);
}

template <typename T, int P, int CC>
static void spread_cc(const SpreadArgs<T> &a, hipStream_t st) {
  constexpr int TX = 8, TY = 8;
  const int ntiles = (a.K / TX) * (a.K / TY);
  const int nz = (a.K + a.ZC - 1) / a.ZC;
  hipLaunchKernelGGL((k_spread<T, P, CC, TX, TY>), dim3(ntiles, nz, a.C / CC), dim3(64), 0, st,
                     a.boff, a.pbase, a.pw, a.fs, a.C, a.grid, a.comp_stride, a.RS, a.K, a.ZC);
}

// Components per pass: the register window holds P*CC values per lane;
// P = 8 with 12 components no longer unrolls into registers (it spills).
template <typename T, int P>
static int spread_p(const SpreadArgs<T> &a, hipStream_t st) {
  if (a.CC == 3) spread_cc<T, P, 3>(a, st);
  else if (a.CC == 6) spread_cc<T, P, 6>(a, st);
  else if constexpr (P < 8) {
    if (a.CC == 12) spread_cc<T, P, 12>(a, st);
    else return -1;
  } else return -1;
  return 0;
}

template <typename T>
int launch_spread(const SpreadArgs<T> &a, hipStream_t st) {
  if (a.C % a.CC) return -1;
  switch (a.P) {
    case 4: return spread_p<T, 4>(a, st);
    case 6: return spread_p<T, 6>(a, st);
    case 8: return spread_p<T, 8>(a, st);
  }
  return -1;
}

template <typename T, int P, int CC>
static void interp_cc(const InterpArgs<T> &a, hipStream_t st) {
  hipLaunchKernelGGL((k_interp<T, P, CC>), dim3(grid1(a.N, 4)), dim3(256), 0, st, a.pbase, a.pw,
                     a.perm, a.grid, a.comp_stride, a.RS, a.K, a.N, a.C, a.alpha, a.beta,
                     a.out, a.ldout);
}

template <typename T, int P>
static int interp_p(const InterpArgs<T> &a, hipStream_t st) {
  if (a.C % 12 == 0) interp_cc<T, P, 12>(a, st);
  else if (a.C % 6 == 0) interp_cc<T, P, 6>(a, st);
  else if (a.C % 3 == 0) interp_cc<T, P, 3>(a, st);
  else return -1;
  return 0;
}

template <typename T>
int launch_interp(const InterpArgs<T> &a, hipStream_t st) {
  switch (a.P) {
    case 4: return interp_p<T, 4>(a, st);
    case 6: return interp_p<T, 6>(a, st);
    case 8: return interp_p<T, 8>(a, st);
  }
  return -1;
}

// Line batch sizes: keep ~24 KB of LDS per FFT block.
template <typename T> struct FftShape {
  static constexpr int NLX = sizeof(T) == 8 ? 8 : 16;  // complex lines per x block
  static constexpr int XW = sizeof(T) == 8 ? 8 : 16;   // y-pass columns
  static constexpr int XWZ = sizeof(T) == 8 ? 4 : 8;   // z-pass columns (x3 comps)
};

template <typename T, int K>
static void fft3d_k(const FftArgs<T> &a, hipStream_t st, hipEvent_t *ev) {
  using S = FftShape<T>;
  constexpr int NT = 256;
  const long rows = (long)a.ncomp * K * K;
  const long npairs = rows / 2;
  const int Khp = a.RS / 2;
  const int nxc = (K / 2 + 1 + S::XW - 1) / S::XW;
  const int nxcz = (K / 2 + 1 + S::XWZ - 1) / S::XWZ;
  auto mark = [&](int i) {
    if (ev) (void)hipEventRecord(ev[i], st);
  };
  hipLaunchKernelGGL((k_fft_x_r2c<T, K, S::NLX, NT>), dim3(grid1(npairs, S::NLX)), dim3(NT), 0,
                     st, a.grid, a.RS, npairs, a.tw);
  mark(0);
  hipLaunchKernelGGL((k_fft_y<T, K, S::XW, -1, NT>), dim3(nxc * K * a.ncomp), dim3(NT), 0, st,
                     a.grid, a.comp_stride, Khp, nxc, a.tw);
  mark(1);
  hipLaunchKernelGGL((k_fft_z_infl<T, K, S::XWZ, NT>), dim3(nxcz * K, a.ncomp / 3), dim3(NT), 0,
                     st, a.grid, a.comp_stride, Khp, nxcz, a.tab, a.dk, a.tw);
  mark(2);
  hipLaunchKernelGGL((k_fft_y<T, K, S::XW, +1, NT>), dim3(nxc * K * a.ncomp), dim3(NT), 0, st,
                     a.grid, a.comp_stride, Khp, nxc, a.tw);
  mark(3);
  hipLaunchKernelGGL((k_fft_x_c2r<T, K, S::NLX, NT>), dim3(grid1(npairs, S::NLX)), dim3(NT), 0,
                     st, a.grid, a.RS, npairs, a.tw);
  mark(4);
}

template <typename T>
int launch_fft_influence(const FftArgs<T> &a, hipStream_t st, hipEvent_t *ev) {
  switch (a.K) {
    case 16: fft3d_k<T, 16>(a, st, ev); return 0;
    case 24: fft3d_k<T, 24>(a, st, ev); return 0;
    case 32: fft3d_k<T, 32>(a, st, ev); return 0;
    case 48: fft3d_k<T, 48>(a, st, ev); return 0;
    case 64: fft3d_k<T, 64>(a, st, ev); return 0;
    case 96: fft3d_k<T, 96>(a, st, ev); return 0;
    case 128: fft3d_k<T, 128>(a, st, ev); return 0;
    case 192: fft3d_k<T, 192>(a, st, ev); return 0;
    case 256: fft3d_k<T, 256>(a, st, ev); return 0;
  }
  return -1;
}

bool fft_size_supported(int K) {
  return K == 16 || K == 24 || K == 32 || K == 48 || K == 64 || K == 96 || K == 128 ||
         K == 192 || K == 256;
}

template void launch_pgather<float>(const int *, const int *, const double *, int *, float *,
                                    int, int, hipStream_t);
template void launch_pgather<double>(const int *, const int *, const double *, int *, double *,
                                     int, int, hipStream_t);
template void launch_permute_in<float>(const int *, const float *, long, float *, int, int,
                                       hipStream_t);
template void launch_permute_in<double>(const int *, const double *, long, double *, int, int,
                                        hipStream_t);
template int launch_spread<float>(const SpreadArgs<float> &, hipStream_t);
template int launch_spread<double>(const SpreadArgs<double> &, hipStream_t);
template int launch_interp<float>(const InterpArgs<float> &, hipStream_t);
template int launch_interp<double>(const InterpArgs<double> &, hipStream_t);
template int launch_fft_influence<float>(const FftArgs<float> &, hipStream_t, hipEvent_t *);
template int launch_fft_influence<double>(const FftArgs<double> &, hipStream_t, hipEvent_t *);

}  // namespace spme
