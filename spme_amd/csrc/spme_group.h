// Host launch interface of the grouped (multi-system) kernels, spme_group.hip.
//
// A group stage runs the logical blocks of many independent systems in one
// launch.  Logical block g of the group (after the XCD remap) belongs to
// system sys[g], as its local block g - pre[sys[g]]; args[s] holds system
// s's kernel arguments.  Launchers take a hardware grid size: the whole
// logical range runs one logical block per hardware block; a smaller grid
// selects the persistent form, which loops over the logical blocks with
// stride = grid (then a multiple of 8, so that every XCD keeps one
// contiguous range of logical blocks).  The FFT passes are never persistent.
#pragma once
#include <hip/hip_runtime.h>

#include "spme_launch.h"
#include "spme_real.h"

namespace spme {

template <class A> struct GroupStage {
  const A *args;
  const int *pre;  // first logical block of each system
  const int *sys;  // system of each logical block
  int total;       // logical blocks
};

// Force permutation of one system for both orderings: fs (P buckets) and,
// when rperm is set, rfs (real-space cells).
template <typename T> struct GPermArgs {
  const int *perm;
  const int *rperm;
  const T *in;
  long ldin;
  T *fs;
  T *rfs;
  int N, C;
};

constexpr int GROUP_ELEMS = 256;  // elements (rows) per logical block

inline int group_perm_blocks(int N, int C) {
  return int(((long)N * C + GROUP_ELEMS - 1) / GROUP_ELEMS);
}
// real space: real_lpr(C) lanes per row
inline int group_real_blocks(int N, int C) {
  return int(((long)N * real_lpr(C) + GROUP_ELEMS - 1) / GROUP_ELEMS);
}
// Meshes whose grouped x passes are the fused x + y plane passes of the
// single-system path (256 threads per block, as their y passes).
template <typename T> constexpr bool group_xy(int K) { return K <= 96 && xy_fused<T>(K); }

// Logical blocks of FFT pass i (0 x fwd, 1 y fwd, 2 z + influence,
// 3 y inv, 4 x inv) of a K^3 mesh with ncomp components; with fused x + y
// planes, passes 0 and 4 take one block per plane and passes 1 and 3 none.
template <typename T> inline int fft_pass_blocks(int K, int ncomp, int pass) {
  using S = FftShape<T>;
  if (group_xy<T>(K)) {
    if (pass == 0 || pass == 4) return ncomp * K;
    if (pass == 1 || pass == 3) return 0;
  }
  if (pass == 0 || pass == 4) {
    const long npairs = (long)ncomp * K * K / 2;
    return int((npairs + S::NLX - 1) / S::NLX);
  }
  if (pass == 1 || pass == 3) return ((K / 2 + 1 + S::XW - 1) / S::XW) * K * ncomp;
  return ZLayout<S::XWZ>::per_rhs(K) * (ncomp / 3);
}
// Largest mesh the grouped FFT kernels are instantiated for: 96 or 256.
inline int group_kmax(int K) { return K <= 96 ? 96 : 256; }

template <typename T>
void launch_group_permute(const GroupStage<GPermArgs<T>> &t, int grid, hipStream_t st);
template <typename T>
int launch_group_spread(const GroupStage<SpreadArgs<T>> &t, int P, int CC, int grid,
                        hipStream_t st);
template <typename T>
int launch_group_fft(const GroupStage<FftArgs<T>> &t, int pass, int kmax, hipStream_t st);
template <typename T>
int launch_group_interp(const GroupStage<InterpArgs<T>> &t, int P, int txy, int grid,
                        hipStream_t st);
template <typename T>
int launch_group_real(const GroupStage<RealArgs<T>> &t, int C, int grid, hipStream_t st);

}  // namespace spme
