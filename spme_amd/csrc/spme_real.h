// Host-side launch interface of the real-space kernels (spme_real.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace spme {

// TO: type of the output vectors and of the final sum (see InterpArgs).
template <typename T, typename TO = T> struct RealArgs {
  const int *perm;   // cell order -> original index
  const int *width;  // entries per row, per slice of 64 rows
  const int *soff;   // slice offsets (entries)
  const int *col;    // neighbour (cell order)
  const T *blk;      // pair coefficients (F, G / r^2) of each entry, interleaved
  const double *spos;  // FP64 engines: positions in cell order [N][3]
  const float4 *cpos;  // FP32 engines: offsets in the cell and cell coordinates
  int nc;              // cells per side
  float h;             // cell side
  double L;
  double self;
  const T *selfv;  // per-row self coefficients (cell order), or null: self
  const T *fs;  // forces in cell order, fs[s*C + d*C/3 + r]
  int N, C;
  double alpha, beta;
  int accum;
  TO *out;
  long ldout;
};

void launch_cell_key(const double *pos, int N, double L, int nc, int *key, int *counts,
                     hipStream_t st);
void launch_cell_gather(const int *perm, const double *pos, double *spos, const double *rad,
                        double *srad, int N, hipStream_t st);
void launch_cell_rel(const double *spos, int N, double L, int nc, float4 *cpos, hipStream_t st);
template <typename T>
void launch_self_radii(const double *srad, int N, double xi, T *selfv, hipStream_t st);
void launch_pair_count(const double *spos, const int *coff, int nc, double L, double cut2,
                       int N, int *cnt, hipStream_t st);
void launch_slice_width(const int *cnt, int N, int *width, int *size, hipStream_t st);
template <typename T>
void launch_pair_fill(const double *spos, const int *coff, int nc, double L, double xi,
                      double rc2, double cut2, int N, const int *width, const int *soff,
                      int *col, const double *srad, T *blk, int *err, hipStream_t st);
template <typename T, typename TO> int launch_real(const RealArgs<T, TO> &a, hipStream_t st);

}  // namespace spme
