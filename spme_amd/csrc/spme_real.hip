// Real-space RPY part of the Beenakker-Ewald mobility (absent from the
// reference, which computes only the wave-space part): a matrix-free 3x3
// block SpMV over a cell list, with the Rotne-Prager overlap correction and
// the self term.  Radius a = 1, mobility in units of 1/(6 pi eta a).
//
//   M_real(r) = F(r) I + G(r) r r^T / r^2                (Beenakker 1986)
//   F = erfc(xi r)(3/(4r) + 1/(2r^3))
//       + e^{-xi^2 r^2}/sqrt(pi) (4xi^7 r^4 + 3xi^3 r^2 - 20xi^5 r^2 - 9xi/2 + 14xi^3 + xi/r^2)
//   G = erfc(xi r)(3/(4r) - 3/(2r^3))
//       + e^{-xi^2 r^2}/sqrt(pi) (-4xi^7 r^4 - 3xi^3 r^2 + 16xi^5 r^2 + 3xi/2 - 2xi^3 - 3xi/r^2)
//   r < 2: add (1 - 9r/32) I + (3r/32) rr^T/r^2 - [3/(4r)(1 + 2/(3r^2)) I + 3/(4r)(1 - 2/r^2) rr^T/r^2]
//   self:  (1 - 6xi/sqrt(pi) + 40xi^3/(3 sqrt(pi))) I
//
// Pair coefficients are evaluated in FP64 on every path: at overlapping
// pairs the 1/r^3 terms of M_real and of the far-field subtraction cancel.
#include <hip/hip_runtime.h>

#include "spme_real.h"

namespace spme {

__device__ __forceinline__ void pair_coeffs(double r2, double xi, double rc2, double &F,
                                            double &G) {
  const double r = sqrt(r2);
  F = 0.0;
  G = 0.0;
  if (r2 < rc2) {
    const double xr = xi * r;
    const double e = exp(-xr * xr) * 0.56418958354775628695;  // 1/sqrt(pi)
    const double ec = erfc(xr);
    const double x2 = xi * xi, x3 = x2 * xi, x5 = x3 * x2, x7 = x5 * x2;
    const double ir = 1.0 / r, ir2 = ir * ir, ir3 = ir2 * ir;
    F = ec * (0.75 * ir + 0.5 * ir3) +
        e * (4.0 * x7 * r2 * r2 + 3.0 * x3 * r2 - 20.0 * x5 * r2 - 4.5 * xi + 14.0 * x3 +
             xi * ir2);
    G = ec * (0.75 * ir - 1.5 * ir3) +
        e * (-4.0 * x7 * r2 * r2 - 3.0 * x3 * r2 + 16.0 * x5 * r2 + 1.5 * xi - 2.0 * x3 -
             3.0 * xi * ir2);
  }
  if (r2 < 4.0) {
    const double ir = 1.0 / r;
    F += (1.0 - 0.28125 * r) - 0.75 * ir * (1.0 + 2.0 / (3.0 * r2));
    G += 0.09375 * r - 0.75 * ir * (1.0 - 2.0 / r2);
  }
}

// Cell key of a position (cells of side L/nc, wrapped).
__global__ __launch_bounds__(256) void k_cell_key(const double *__restrict__ pos, int N,
                                                  double L, int nc, int *__restrict__ key,
                                                  int *__restrict__ counts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double s = pos[3 * i + d] / L;
    s -= floor(s);
    int v = int(s * nc);
    c[d] = v >= nc ? nc - 1 : (v < 0 ? 0 : v);
  }
  const int k = (c[2] * nc + c[1]) * nc + c[0];
  key[i] = k;
  atomicAdd(&counts[k], 1);
}

__global__ __launch_bounds__(256) void k_cell_gather(const int *__restrict__ perm,
                                                     const double *__restrict__ pos,
                                                     double *__restrict__ spos, int N) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const int i = perm[s];
  spos[3 * s + 0] = pos[3 * i + 0];
  spos[3 * s + 1] = pos[3 * i + 1];
  spos[3 * s + 2] = pos[3 * i + 2];
}

// One thread per particle (cell order).  fs: forces in cell order [N][C].
// out[r*ldout + 3*perm[s] + d] = (accum ? out : beta*out) + alpha*(self f + sum M f).
template <typename T, int C>
__global__ __launch_bounds__(128) void k_real(const double *__restrict__ spos,
                                              const int *__restrict__ perm,
                                              const int *__restrict__ coff, int nc, double L,
                                              double xi, double rc2, double self,
                                              const T *__restrict__ fs, int N, double alpha,
                                              double beta, int accum, T *__restrict__ out,
                                              long ldout) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const double px = spos[3 * s], py = spos[3 * s + 1], pz = spos[3 * s + 2];
  double acc[C];
#pragma unroll
  for (int c = 0; c < C; ++c) acc[c] = self * double(fs[(size_t)s * C + c]);
  int cc[3];
  {
    double q[3] = {px / L, py / L, pz / L};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double t = q[d] - floor(q[d]);
      int v = int(t * nc);
      cc[d] = v >= nc ? nc - 1 : (v < 0 ? 0 : v);
    }
  }
  const double iL = 1.0 / L;
  for (int dz = -1; dz <= 1; ++dz) {
    int cz = cc[2] + dz;
    cz += cz < 0 ? nc : (cz >= nc ? -nc : 0);
    for (int dy = -1; dy <= 1; ++dy) {
      int cy = cc[1] + dy;
      cy += cy < 0 ? nc : (cy >= nc ? -nc : 0);
      for (int dx = -1; dx <= 1; ++dx) {
        int cx = cc[0] + dx;
        cx += cx < 0 ? nc : (cx >= nc ? -nc : 0);
        const int cell = (cz * nc + cy) * nc + cx;
        const int j0 = coff[cell], j1 = coff[cell + 1];
        for (int j = j0; j < j1; ++j) {
          if (j == s) continue;
          double ex = px - spos[3 * j], ey = py - spos[3 * j + 1], ez = pz - spos[3 * j + 2];
          ex -= L * rint(ex * iL);
          ey -= L * rint(ey * iL);
          ez -= L * rint(ez * iL);
          const double r2 = ex * ex + ey * ey + ez * ez;
          if (r2 >= rc2 && r2 >= 4.0) continue;
          if (r2 == 0.0) continue;
          double F, G;
          pair_coeffs(r2, xi, rc2, F, G);
          const double g = G / r2;
          const T *fj = fs + (size_t)j * C;
#pragma unroll
          for (int r = 0; r < C / 3; ++r) {
            const double f0 = fj[3 * r], f1 = fj[3 * r + 1], f2 = fj[3 * r + 2];
            const double dot = g * (ex * f0 + ey * f1 + ez * f2);
            acc[3 * r + 0] += F * f0 + ex * dot;
            acc[3 * r + 1] += F * f1 + ey * dot;
            acc[3 * r + 2] += F * f2 + ez * dot;
          }
        }
      }
    }
  }
  const int o = perm[s];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int r = c / 3, d = c - 3 * r;
    T *u = out + r * ldout + 3L * o + d;
    const double prev = accum ? double(*u) : (beta == 0.0 ? 0.0 : beta * double(*u));
    *u = T(prev + alpha * acc[c]);
  }
}

static int grid1(long n, int b) { return int((n + b - 1) / b); }

void launch_cell_key(const double *pos, int N, double L, int nc, int *key, int *counts,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_cell_key, dim3(grid1(N, 256)), dim3(256), 0, st, pos, N, L, nc, key,
                     counts);
}

void launch_cell_gather(const int *perm, const double *pos, double *spos, int N,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_cell_gather, dim3(grid1(N, 256)), dim3(256), 0, st, perm, pos, spos, N);
}

template <typename T, int C>
static void real_c(const RealArgs<T> &a, hipStream_t st) {
  hipLaunchKernelGGL((k_real<T, C>), dim3(grid1(a.N, 128)), dim3(128), 0, st, a.spos, a.perm,
                     a.coff, a.nc, a.L, a.xi, a.rc2, a.self, a.fs, a.N, a.alpha, a.beta,
                     a.accum, a.out, a.ldout);
}

template <typename T>
int launch_real(const RealArgs<T> &a, hipStream_t st) {
  switch (a.C) {
    case 3: real_c<T, 3>(a, st); return 0;
    case 6: real_c<T, 6>(a, st); return 0;
    case 12: real_c<T, 12>(a, st); return 0;
    case 24: real_c<T, 24>(a, st); return 0;
    case 48: real_c<T, 48>(a, st); return 0;
  }
  return -1;
}

template int launch_real<float>(const RealArgs<float> &, hipStream_t);
template int launch_real<double>(const RealArgs<double> &, hipStream_t);

}  // namespace spme
