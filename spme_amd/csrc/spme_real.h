// Host-side launch interface of the real-space kernels (spme_real.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace spme {

template <typename T> struct RealArgs {
  const double *spos;  // positions in cell order [N][3]
  const int *perm;     // cell order -> original index
  const int *coff;     // cell offsets [nc^3 + 1]
  int nc;
  double L, xi, rc2, self;
  const T *fs;  // forces in cell order [N][C]
  int N, C;
  double alpha, beta;
  int accum;
  T *out;
  long ldout;
};

void launch_cell_key(const double *pos, int N, double L, int nc, int *key, int *counts,
                     hipStream_t st);
void launch_cell_gather(const int *perm, const double *pos, double *spos, int N,
                        hipStream_t st);
template <typename T> int launch_real(const RealArgs<T> &a, hipStream_t st);

}  // namespace spme
