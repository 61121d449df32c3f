// Real-space RPY part of the Beenakker-Ewald mobility (absent from the
// reference, which computes only the wave-space part): a 3x3 block-sparse
// SpMV whose blocks are built once per position set over a cell list, with
// the Rotne-Prager overlap correction and the self term.  Radius a = 1, mobility in units of 1/(6 pi eta a).
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

#include "spme_bodies.h"
#include "spme_kernels.h"
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

// Unequal radii a_i, a_j (units of the reference radius): the far field is
// linear in s2 = (a_i^2 + a_j^2)/2, M_far = 3/(4r)(I + rr^T/r^2) + s2/(2r^3)
// (I - 3 rr^T/r^2), so Beenakker's real-space terms split by dimension into
// an Oseen part and an s2 part (s2 = 1 gives pair_coeffs):
//   F = erfc(xi r) (3/(4r) + s2/(2r^3))
//       + e (3xi^3 r^2 - 9xi/2 + s2 (4xi^7 r^4 - 20xi^5 r^2 + 14xi^3 + xi/r^2))
//   G = erfc(xi r) (3/(4r) - 3 s2/(2r^3))
//       + e (-3xi^3 r^2 + 3xi/2 + s2 (-4xi^7 r^4 + 16xi^5 r^2 - 2xi^3 - 3xi/r^2))
// with e = e^{-xi^2 r^2}/sqrt(pi).  Overlap, r < a_i + a_j (Zuk, Wajnryb,
// Mizerski, Szymczak 2014): add M_ov - M_far with
//   |a_i - a_j| < r: M_ov = [16 r^3 (a_i + a_j) - ((a_i - a_j)^2 + 3r^2)^2]
//                           / (32 r^3 a_i a_j) I
//                          + 3 ((a_i - a_j)^2 - r^2)^2 / (32 r^3 a_i a_j) rr^T/r^2
//   r <= |a_i - a_j|: M_ov = I / max(a_i, a_j).
__device__ __forceinline__ void pair_coeffs_radii(double r2, double xi, double rc2, double ai,
                                                  double aj, double &F, double &G) {
  const double r = sqrt(r2), s2 = 0.5 * (ai * ai + aj * aj);
  F = 0.0;
  G = 0.0;
  if (r2 < rc2) {
    const double xr = xi * r;
    const double e = exp(-xr * xr) * 0.56418958354775628695;  // 1/sqrt(pi)
    const double ec = erfc(xr);
    const double x2 = xi * xi, x3 = x2 * xi, x5 = x3 * x2, x7 = x5 * x2;
    const double ir = 1.0 / r, ir2 = ir * ir, ir3 = ir2 * ir;
    F = ec * (0.75 * ir + 0.5 * s2 * ir3) +
        e * (3.0 * x3 * r2 - 4.5 * xi +
             s2 * (4.0 * x7 * r2 * r2 - 20.0 * x5 * r2 + 14.0 * x3 + xi * ir2));
    G = ec * (0.75 * ir - 1.5 * s2 * ir3) +
        e * (-3.0 * x3 * r2 + 1.5 * xi +
             s2 * (-4.0 * x7 * r2 * r2 + 16.0 * x5 * r2 - 2.0 * x3 - 3.0 * xi * ir2));
  }
  const double sa = ai + aj;
  if (r < sa) {
    const double d = ai - aj, r3 = r2 * r;
    double Fov, Gov;
    if (r <= fabs(d)) {
      Fov = 1.0 / fmax(ai, aj);
      Gov = 0.0;
    } else {
      const double q = d * d + 3.0 * r2, w = d * d - r2;
      Fov = (16.0 * r3 * sa - q * q) / (32.0 * r3 * ai * aj);
      Gov = 3.0 * w * w / (32.0 * r3 * ai * aj);
    }
    F += Fov - (0.75 / r + s2 / (2.0 * r3));
    G += Gov - (0.75 / r - 1.5 * s2 / r3);
  }
}

// Self coefficient of each row (cell order) with unequal radii:
// 1/a - 6 xi/sqrt(pi) + 40 xi^3 a^2/(3 sqrt(pi)).
template <typename T>
__global__ __launch_bounds__(256) void k_self_radii(const double *__restrict__ rrad, int N,
                                                    double xi, T *__restrict__ selfv) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const double a = rrad[s], isp = 0.56418958354775628695;
  selfv[s] = T(1.0 / a - 6.0 * xi * isp + 40.0 * xi * xi * xi * a * a * isp / 3.0);
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

// FP32 engines' real-space positions (cell order): the offset of each
// particle from the corner of its cell (the cell of cell_of, side h = L/nc),
// formed in FP64 and rounded to FP32, and the cell's coordinates packed as
// cx | cy << 10 | cz << 20 (nc <= 128).
__global__ __launch_bounds__(256) void k_cell_rel(const double *__restrict__ spos, int N,
                                                  double L, int nc, float4 *__restrict__ cpos) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const double h = L / nc;
  float o[3];
  int key = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double t = spos[3L * s + d] / L;
    t -= floor(t);
    int v = int(t * nc);
    v = v >= nc ? nc - 1 : (v < 0 ? 0 : v);
    o[d] = float(t * L - v * h);
    key |= v << (10 * d);
  }
  cpos[s] = make_float4(o[0], o[1], o[2], __int_as_float(key));
}

// positions (and radii, when rad != null) into cell order
__global__ __launch_bounds__(256) void k_cell_gather(const int *__restrict__ perm,
                                                     const double *__restrict__ pos,
                                                     double *__restrict__ spos,
                                                     const double *__restrict__ rad,
                                                     double *__restrict__ srad, int N) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const int i = perm[s];
  spos[3 * s + 0] = pos[3 * i + 0];
  spos[3 * s + 1] = pos[3 * i + 1];
  spos[3 * s + 2] = pos[3 * i + 2];
  if (rad) srad[s] = rad[i];
}

// ---------------------------------------------------------------------------
// Pair blocks, built once per position set.  Rows are particles in cell
// order, grouped in slices of 64 rows; slice q holds width[q] entries per
// row, stored entry-major (entry k of row 64q + l at soff[q] + 64k + l) so
// that a wave reads each entry array coalesced.  Each entry is a neighbour
// j (cell order) and the two coefficients F and G/r^2 of its symmetric 3x3
// block M_real + overlap = F I + (G/r^2) e e^T (the SpMV forms e from the
// positions); past the end of a row, the row's own index with zero
// coefficients.

__device__ __forceinline__ int cell_of(double p, double L, int nc) {
  double t = p / L;
  t -= floor(t);
  int v = int(t * nc);
  return v >= nc ? nc - 1 : (v < 0 ? 0 : v);
}

// fn(j, ex, ey, ez, r2) for every j != s within the cutoff, over the 27
// cells around s's in a fixed order (dz, dy, dx ascending, j ascending
// within a cell).  The distance test is evaluated without contraction, so
// the count and column passes (separate kernels) decide alike.
template <typename F>
__device__ __forceinline__ void for_neighbours(const double *__restrict__ spos,
                                               const int *__restrict__ coff, int nc, double L,
                                               int s, double cut2, F &&fn) {
  const double px = spos[3 * s], py = spos[3 * s + 1], pz = spos[3 * s + 2];
  const int cx0 = cell_of(px, L, nc), cy0 = cell_of(py, L, nc), cz0 = cell_of(pz, L, nc);
  const double iL = 1.0 / L;
  for (int dz = -1; dz <= 1; ++dz) {
    int cz = cz0 + dz;
    cz += cz < 0 ? nc : (cz >= nc ? -nc : 0);
    for (int dy = -1; dy <= 1; ++dy) {
      int cy = cy0 + dy;
      cy += cy < 0 ? nc : (cy >= nc ? -nc : 0);
      for (int dx = -1; dx <= 1; ++dx) {
        int cx = cx0 + dx;
        cx += cx < 0 ? nc : (cx >= nc ? -nc : 0);
        const int cell = (cz * nc + cy) * nc + cx;
        for (int j = coff[cell], j1 = coff[cell + 1]; j < j1; ++j) {
#pragma clang fp contract(off)
          if (j == s) continue;
          double ex = px - spos[3 * j], ey = py - spos[3 * j + 1], ez = pz - spos[3 * j + 2];
          ex -= L * rint(ex * iL);
          ey -= L * rint(ey * iL);
          ez -= L * rint(ez * iL);
          const double r2 = ex * ex + ey * ey + ez * ez;
          if (r2 >= cut2 || r2 == 0.0) continue;
          fn(j, ex, ey, ez, r2);
        }
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_pair_count(const double *__restrict__ spos,
                                                    const int *__restrict__ coff, int nc,
                                                    double L, double cut2, int N,
                                                    int *__restrict__ cnt) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  int n = 0;
  for_neighbours(spos, coff, nc, L, s, cut2,
                 [&](int, double, double, double, double) { ++n; });
  cnt[s] = n;
}

// slice q: size[q] = 64 * max count over its rows; width[q] = that max.
__global__ __launch_bounds__(64) void k_slice_width(const int *__restrict__ cnt, int N,
                                                    int *__restrict__ width,
                                                    int *__restrict__ size) {
  const int q = blockIdx.x, s = q * 64 + threadIdx.x;
  int v = s < N ? cnt[s] : 0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  if (threadIdx.x == 0) {
    width[q] = v;
    size[q] = 64 * v;
  }
}

// Columns of each row's entries in neighbour order, padded with the row's
// own index.  The coefficients follow in a second pass over the stored
// entries only: computed here, the FP64 erfc/exp of a hit would hold up the
// whole wave on every candidate that one lane accepts (about one in seven).
// The slice widths come from k_pair_count's decisions; a row that finds
// more neighbours here than its slice holds (the two passes disagreeing on
// a distance test) stores none past the width and sets bit 2 of err.
__global__ __launch_bounds__(256) void k_pair_cols(const double *__restrict__ spos,
                                                   const int *__restrict__ coff, int nc,
                                                   double L, double cut2, int N,
                                                   const int *__restrict__ width,
                                                   const int *__restrict__ soff,
                                                   int *__restrict__ col, int *__restrict__ err) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const int q = s >> 6, l = s & 63;
  const long base = soff[q] + l;
  const int w = width[q];
  int k = 0;
  bool over = false;
  for_neighbours(spos, coff, nc, L, s, cut2, [&](int j, double, double, double, double) {
    if (k < w) col[base + 64L * k] = j;
    else over = true;
    ++k;
  });
  if (over) atomicOr(err, 2);
  for (; k < w; ++k) col[base + 64L * k] = s;
}

// F and G/r^2 of every stored entry (zero for the padding on the row itself);
// with srad (radii in cell order) those of the pair's radii.
template <typename T>
__global__ __launch_bounds__(256) void k_pair_coeffs(const double *__restrict__ spos, int N,
                                                     double L, double xi, double rc2,
                                                     const int *__restrict__ width,
                                                     const int *__restrict__ soff,
                                                     const int *__restrict__ col,
                                                     const double *__restrict__ srad,
                                                     T *__restrict__ blk) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const int q = s >> 6, l = s & 63;
  const long base = soff[q] + l;
  const double px = spos[3L * s], py = spos[3L * s + 1], pz = spos[3L * s + 2];
  const double iL = 1.0 / L;
  const int w = width[q];
  // four entries at a time: their columns, then their positions, are loaded
  // together
  for (int k0 = 0; k0 < w; k0 += 4) {
    int jj[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) jj[u] = col[base + 64L * min(k0 + u, w - 1)];
    double qx[4], qy[4], qz[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      qx[u] = spos[3L * jj[u]];
      qy[u] = spos[3L * jj[u] + 1];
      qz[u] = spos[3L * jj[u] + 2];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (k0 + u >= w) break;
      const long e = base + 64L * (k0 + u);
      const int j = jj[u];
      double F = 0.0, G = 0.0, r2 = 1.0;
      if (j != s) {
        double ex = px - qx[u], ey = py - qy[u], ez = pz - qz[u];
        ex -= L * rint(ex * iL);
        ey -= L * rint(ey * iL);
        ez -= L * rint(ez * iL);
        r2 = ex * ex + ey * ey + ez * ez;
        if (srad)
          pair_coeffs_radii(r2, xi, rc2, srad[s], srad[j], F, G);
        else
          pair_coeffs(r2, xi, rc2, F, G);
      }
      blk[2 * e] = T(F);
      blk[2 * e + 1] = T(G / r2);
    }
  }
}

// Lanes = (row in cell order, right-hand side), right-hand side fastest.
template <typename T, int C, typename TO>
__global__ __launch_bounds__(256) void k_real_spmv(const int *__restrict__ perm,
                                                   const int *__restrict__ width,
                                                   const int *__restrict__ soff,
                                                   const int *__restrict__ col,
                                                   const T *__restrict__ blk,
                                                   const double *__restrict__ spos,
                                                   const float4 *__restrict__ cpos, int nc,
                                                   float h, double L,
                                                   T self, const T *__restrict__ selfv,
                                                   const T *__restrict__ fs, int N,
                                                   TO alpha, TO beta, int accum,
                                                   TO *__restrict__ out, long ldout) {
  constexpr int LPR = RealShape<C>::LPR;
  const long g = (long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int s = int(g / LPR), l = int(g % LPR);
  if (s < N)
    real_lane<T, C, TO>(perm, width, soff, col, blk, spos, cpos, nc, h, L, self, selfv, fs,
                        alpha, beta, accum, out, ldout, s, l);
}

static int grid1(long n, int b) { return int((n + b - 1) / b); }

void launch_cell_key(const double *pos, int N, double L, int nc, int *key, int *counts,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_cell_key, dim3(grid1(N, 256)), dim3(256), 0, st, pos, N, L, nc, key,
                     counts);
}

void launch_cell_gather(const int *perm, const double *pos, double *spos, const double *rad,
                        double *srad, int N, hipStream_t st) {
  hipLaunchKernelGGL(k_cell_gather, dim3(grid1(N, 256)), dim3(256), 0, st, perm, pos, spos, rad,
                     srad, N);
}

template <typename T>
void launch_self_radii(const double *srad, int N, double xi, T *selfv, hipStream_t st) {
  hipLaunchKernelGGL(k_self_radii<T>, dim3(grid1(N, 256)), dim3(256), 0, st, srad, N, xi, selfv);
}
template void launch_self_radii<float>(const double *, int, double, float *, hipStream_t);
template void launch_self_radii<double>(const double *, int, double, double *, hipStream_t);

void launch_cell_rel(const double *spos, int N, double L, int nc, float4 *cpos, hipStream_t st) {
  hipLaunchKernelGGL(k_cell_rel, dim3(grid1(N, 256)), dim3(256), 0, st, spos, N, L, nc, cpos);
}

void launch_pair_count(const double *spos, const int *coff, int nc, double L, double cut2,
                       int N, int *cnt, hipStream_t st) {
  hipLaunchKernelGGL(k_pair_count, dim3(grid1(N, 256)), dim3(256), 0, st, spos, coff, nc, L,
                     cut2, N, cnt);
}

void launch_slice_width(const int *cnt, int N, int *width, int *size, hipStream_t st) {
  hipLaunchKernelGGL(k_slice_width, dim3(grid1(N, 64)), dim3(64), 0, st, cnt, N, width, size);
}

template <typename T>
void launch_pair_fill(const double *spos, const int *coff, int nc, double L, double xi,
                      double rc2, double cut2, int N, const int *width, const int *soff,
                      int *col, const double *srad, T *blk, int *err, hipStream_t st) {
  hipLaunchKernelGGL(k_pair_cols, dim3(grid1(N, 256)), dim3(256), 0, st, spos, coff, nc, L,
                     cut2, N, width, soff, col, err);
  hipLaunchKernelGGL(k_pair_coeffs<T>, dim3(grid1(N, 256)), dim3(256), 0, st, spos, N, L, xi,
                     rc2, width, soff, col, srad, blk);
}

template <typename T, int C, typename TO>
static void real_c(const RealArgs<T, TO> &a, hipStream_t st) {
  hipLaunchKernelGGL((k_real_spmv<T, C, TO>), dim3(grid1((long)a.N * RealShape<C>::LPR, 256)),
                     dim3(256), 0, st, a.perm, a.width, a.soff, a.col, a.blk, a.spos, a.cpos,
                     a.nc, a.h, a.L, T(a.self), a.selfv, a.fs, a.N, TO(a.alpha), TO(a.beta),
                     a.accum, a.out, a.ldout);
}

template <typename T, typename TO>
int launch_real(const RealArgs<T, TO> &a, hipStream_t st) {
  switch (a.C) {  // every batch width 1..16
#define X(CC) \
  case CC: real_c<T, CC, TO>(a, st); return 0;
    X(3) X(6) X(9) X(12) X(15) X(18) X(21) X(24) X(27) X(30) X(33) X(36) X(39) X(42) X(45)
    X(48)
#undef X
  }
  return -1;
}

template void launch_pair_fill<float>(const double *, const int *, int, double, double, double,
                                      double, int, const int *, const int *, int *,
                                      const double *, float *, int *, hipStream_t);
template void launch_pair_fill<double>(const double *, const int *, int, double, double, double,
                                       double, int, const int *, const int *, int *,
                                       const double *, double *, int *, hipStream_t);
template int launch_real<float, float>(const RealArgs<float, float> &, hipStream_t);
template int launch_real<float, double>(const RealArgs<float, double> &, hipStream_t);
template int launch_real<double, double>(const RealArgs<double, double> &, hipStream_t);

}  // namespace spme
