// Grouped kernels: one launch per apply stage for many independent systems
// (uneven N, box and mesh), over the logical blocks of every system.  The per-block work is the same device body as the
// single-system kernels (spme_bodies.h), so a grouped apply gives the same
// bits as the systems' separate applies.
#include <hip/hip_runtime.h>

#include "spme_bodies.h"
#include "spme_group.h"
#include "spme_kernels.h"

namespace spme {

// Logical block g -> (system, local block).  remap: XCD-aware order (valid
// when the grid is the whole range or a multiple of 8).
template <class A>
__device__ __forceinline__ int group_locate(const GroupStage<A> &t, int g, bool remap,
                                            int &lb) {
  const int lg = remap ? xcd_remap(g, t.total) : g;
  const int s = t.sys[lg];
  lb = lg - t.pre[s];
  return s;
}

// Runs f(g) for the logical blocks of this hardware block: one (grid = all
// logical blocks) or, PERSIST, a grid-stride loop with a barrier between
// blocks so that LDS is free for the next one.  The loop is a separate
// instantiation: hoisting work out of it costs registers, so it is used
// only where measurement favours it.
template <bool PERSIST, class F>
__device__ __forceinline__ void group_blocks(int total, F &&f) {
  if constexpr (PERSIST) {
    for (int g = blockIdx.x; g < total; g += gridDim.x) {
      f(g);
      __syncthreads();
    }
  } else {
    if ((int)blockIdx.x < total) f((int)blockIdx.x);
  }
}

template <typename T, bool PERSIST>
__global__ __launch_bounds__(GROUP_ELEMS) void kg_permute(GroupStage<GPermArgs<T>> t) {
  group_blocks<PERSIST>(t.total, [&](int g) {
    int lb;
    const int s = group_locate(t, g, false, lb);
    const GPermArgs<T> &a = t.args[s];
    const long e = (long)lb * GROUP_ELEMS + threadIdx.x;
    if (e < (long)a.N * a.C) {
      if (a.perm) permute_elem<T>(a.perm, a.in, a.ldin, a.fs, a.C, e);
      if (a.rperm) permute_elem_cell<T>(a.rperm, a.in, a.ldin, a.rfs, a.C, e);
    }
  });
}

template <typename T, int P, int CC, bool PERSIST>
__global__ __launch_bounds__(64) void kg_spread(GroupStage<SpreadArgs<T>> t) {
  group_blocks<PERSIST>(t.total, [&](int g) {
    int lb;
    const int s = group_locate(t, g, true, lb);
    const SpreadArgs<T> &a = t.args[s];
    spread_block<T, P, CC, 8, 8>(a.boff, a.pbase, a.pw, a.fs, a.C, a.grid, a.comp_stride, a.RS,
                                 a.K, a.ZC, lb);
  });
}

template <typename T, int P, int CC, int TXY, int NT, bool PERSIST>
__global__ __launch_bounds__(NT) void kg_interp(GroupStage<InterpArgs<T>> t) {
  group_blocks<PERSIST>(t.total, [&](int g) {
    int lb;
    const int s = group_locate(t, g, true, lb);
    const InterpArgs<T> &a = t.args[s];
    interp_block<T, P, CC, TXY, NT>(a.boff, a.pbase, a.pw, a.perm, a.grid, a.comp_stride, a.RS,
                                    a.K, a.C, a.ZC, a.alpha, a.beta, a.out, a.ldout, lb);
  });
}

template <typename T, int C, bool PERSIST>
__global__ __launch_bounds__(GROUP_ELEMS) void kg_real(GroupStage<RealArgs<T>> t) {
  group_blocks<PERSIST>(t.total, [&](int g) {
    int lb;
    const int s = group_locate(t, g, true, lb);
    const RealArgs<T> &a = t.args[s];
    constexpr int LPR = RealShape<C>::LPR;
    const long gi = (long)lb * GROUP_ELEMS + threadIdx.x;
    const int row = int(gi / LPR), l = int(gi % LPR);
    if (row < a.N)
      real_lane<T, C>(a.perm, a.width, a.soff, a.col, a.blk, a.spos, a.cpos, a.nc, a.h, a.L,
                      T(a.self), a.selfv, a.fs, T(a.alpha), T(a.beta), a.accum, a.out, a.ldout,
                      row, l);
  });
}

// FFT passes with the mesh size taken from each system's arguments; LDS is
// sized for KMAX.  K is uniform over a block, so the switch does not
// diverge and every barrier inside a body is reached by the whole block.
// One logical block per hardware block (a persistent loop over these bodies
// doubles their register use).
#define SPME_FOR_EACH_K(X) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256)

// Pass 0 (x forward) and pass 4 (x inverse).  Meshes with K <= 96 run the
// fused x + y plane bodies of the single-system path (one (component, z)
// plane per logical block; their y passes have no blocks), larger ones the
// separate x pass.  LDS holds the larger of the two buffers.
template <typename T, int KMAX>
constexpr int group_x_buf() {
  constexpr int NL = FftShape<T>::NLX;
  constexpr int KS = KMAX < 96 ? KMAX : 96;
  return NL * (KMAX + 1) > XYShape<T, KS>::NCPX ? NL * (KMAX + 1) : XYShape<T, KS>::NCPX;
}

template <typename T, int KMAX, int NT>
__global__ __launch_bounds__(NT) void kg_fft_x_r2c(GroupStage<FftArgs<T>> t) {
  constexpr int NL = FftShape<T>::NLX;
  __shared__ cpx<T> buf[group_x_buf<T, KMAX>()];
  __shared__ cpx<T> s_tw[KMAX];
  {
    const int g = blockIdx.x;
    if (g >= t.total) return;
    int lb;
    const int s = group_locate(t, g, false, lb);
    const FftArgs<T> &a = t.args[s];
    const long npairs = (long)a.ncomp * a.K * a.K / 2;
    switch (a.K) {
#define X(KK)                                                                          \
  case KK:                                                                             \
    if constexpr (KK <= KMAX) {                                                        \
      if constexpr (group_xy<T>(KK))                                                   \
        xy_fwd_block<T, KK, NT, false>(a.grid, a.RS, a.comp_stride, a.ncomp * KK, a.tw, \
                                       nullptr, lb, a.ncomp * KK, buf, s_tw);          \
      else                                                                             \
        fft_x_r2c_block<T, KK, NL, NT>(a.grid, a.RS, npairs, a.tw, lb, buf, s_tw);     \
    }                                                                                  \
    break;
      SPME_FOR_EACH_K(X)
#undef X
    }
  }
}

template <typename T, int KMAX, int NT>
__global__ __launch_bounds__(NT) void kg_fft_x_c2r(GroupStage<FftArgs<T>> t) {
  constexpr int NL = FftShape<T>::NLX;
  __shared__ cpx<T> buf[group_x_buf<T, KMAX>()];
  __shared__ cpx<T> s_tw[KMAX];
  {
    const int g = blockIdx.x;
    if (g >= t.total) return;
    int lb;
    const int s = group_locate(t, g, false, lb);
    const FftArgs<T> &a = t.args[s];
    const long npairs = (long)a.ncomp * a.K * a.K / 2;
    switch (a.K) {
#define X(KK)                                                                          \
  case KK:                                                                             \
    if constexpr (KK <= KMAX) {                                                        \
      if constexpr (group_xy<T>(KK))                                                   \
        xy_inv_block<T, KK, NT, false>(a.grid, a.RS, a.comp_stride, a.ncomp * KK, a.tw, \
                                       nullptr, lb, a.ncomp * KK, buf, s_tw);          \
      else                                                                             \
        fft_x_c2r_block<T, KK, NL, NT>(a.grid, a.RS, npairs, a.tw, lb, buf, s_tw);     \
    }                                                                                  \
    break;
      SPME_FOR_EACH_K(X)
#undef X
    }
  }
}

template <typename T, int KMAX, int S, int NT>
__global__ __launch_bounds__(NT) void kg_fft_y(GroupStage<FftArgs<T>> t) {
  constexpr int XW = FftShape<T>::XW;
  __shared__ cpx<T> buf[XW * (KMAX + 1)];
  __shared__ cpx<T> s_tw[KMAX];
  {
    const int g = blockIdx.x;
    if (g >= t.total) return;
    int lb;
    const int s = group_locate(t, g, true, lb);
    const FftArgs<T> &a = t.args[s];
    const int Khp = a.RS / 2;
    const int nxc = (a.K / 2 + 1 + XW - 1) / XW;
    switch (a.K) {
#define X(KK)                                                                          \
  case KK:                                                                             \
    if constexpr (KK <= KMAX)                                                          \
      fft_y_block<T, KK, XW, S, NT>(a.grid, a.comp_stride, Khp, nxc, a.tw, lb, buf, s_tw); \
    break;
      SPME_FOR_EACH_K(X)
#undef X
    }
  }
}

template <typename T, int KMAX, int NT>
__global__ __launch_bounds__(NT) void kg_fft_z_infl(GroupStage<FftArgs<T>> t) {
  constexpr int XW = FftShape<T>::XWZ;
  __shared__ cpx<T> buf[3 * XW * (KMAX + 1)];
  __shared__ cpx<T> s_tw[KMAX];
  {
    const int g = blockIdx.x;
    if (g >= t.total) return;
    int lb;
    const int s = group_locate(t, g, true, lb);
    const FftArgs<T> &a = t.args[s];
    const int Khp = a.RS / 2;
    switch (a.K) {
#define X(KK)                                                                                \
  case KK:                                                                                   \
    if constexpr (KK <= KMAX)                                                                \
      fft_z_infl_block<T, KK, XW, NT>(a.grid, a.comp_stride, Khp, a.tab, a.dk, a.tw, lb, buf, \
                                      s_tw, nullptr);                                        \
    break;
      SPME_FOR_EACH_K(X)
#undef X
    }
  }
}

// ---------------------------------------------------------------------------
// Launchers.  grid: hardware blocks (the caller clamps it to total).

// A grid smaller than the logical block count selects the persistent form.
template <typename T>
void launch_group_permute(const GroupStage<GPermArgs<T>> &t, int grid, hipStream_t st) {
  if (grid < t.total)
    hipLaunchKernelGGL((kg_permute<T, true>), dim3(grid), dim3(GROUP_ELEMS), 0, st, t);
  else
    hipLaunchKernelGGL((kg_permute<T, false>), dim3(t.total), dim3(GROUP_ELEMS), 0, st, t);
}

template <typename T, int P, int CC>
static void group_spread_cc(const GroupStage<SpreadArgs<T>> &t, int grid, hipStream_t st) {
  if (grid < t.total)
    hipLaunchKernelGGL((kg_spread<T, P, CC, true>), dim3(grid), dim3(64), 0, st, t);
  else
    hipLaunchKernelGGL((kg_spread<T, P, CC, false>), dim3(t.total), dim3(64), 0, st, t);
}

template <typename T, int P>
static int group_spread_p(const GroupStage<SpreadArgs<T>> &t, int CC, int grid,
                          hipStream_t st) {
  if (CC == 3) group_spread_cc<T, P, 3>(t, grid, st);
  else if (CC == 6) group_spread_cc<T, P, 6>(t, grid, st);
  else if (CC == 12) group_spread_cc<T, P, 12>(t, grid, st);
  else return -1;
  return 0;
}

template <typename T>
int launch_group_spread(const GroupStage<SpreadArgs<T>> &t, int P, int CC, int grid,
                        hipStream_t st) {
  switch (P) {
    case 4: return group_spread_p<T, 4>(t, CC, grid, st);
    case 6: return group_spread_p<T, 6>(t, CC, grid, st);
    case 8: return group_spread_p<T, 8>(t, CC, grid, st);
  }
  return -1;
}

template <typename T, int KMAX>
static int group_fft_k(const GroupStage<FftArgs<T>> &t, int pass, hipStream_t st) {
  constexpr int NT = 256;
  const int grid = t.total;
  if (grid == 0) return 0;  // the y passes of groups whose meshes all run fused
  switch (pass) {
    case 0: hipLaunchKernelGGL((kg_fft_x_r2c<T, KMAX, NT>), dim3(grid), dim3(NT), 0, st, t); break;
    case 1: hipLaunchKernelGGL((kg_fft_y<T, KMAX, -1, NT>), dim3(grid), dim3(NT), 0, st, t); break;
    case 2: hipLaunchKernelGGL((kg_fft_z_infl<T, KMAX, NT>), dim3(grid), dim3(NT), 0, st, t); break;
    case 3: hipLaunchKernelGGL((kg_fft_y<T, KMAX, +1, NT>), dim3(grid), dim3(NT), 0, st, t); break;
    case 4: hipLaunchKernelGGL((kg_fft_x_c2r<T, KMAX, NT>), dim3(grid), dim3(NT), 0, st, t); break;
    default: return -1;
  }
  return 0;
}

template <typename T>
int launch_group_fft(const GroupStage<FftArgs<T>> &t, int pass, int kmax, hipStream_t st) {
  if (kmax == 96) return group_fft_k<T, 96>(t, pass, st);
  if (kmax == 256) return group_fft_k<T, 256>(t, pass, st);
  return -1;
}

template <typename T, int P, int TXY>
static void group_interp_t(const GroupStage<InterpArgs<T>> &t, int grid, hipStream_t st) {
  constexpr int NT = InterpShape<T>::NT, CC = InterpShape<T>::CC;
  if (grid < t.total)
    hipLaunchKernelGGL((kg_interp<T, P, CC, TXY, NT, true>), dim3(grid), dim3(NT), 0, st, t);
  else
    hipLaunchKernelGGL((kg_interp<T, P, CC, TXY, NT, false>), dim3(t.total), dim3(NT), 0, st, t);
}

template <typename T, int P>
static int group_interp_p(const GroupStage<InterpArgs<T>> &t, int txy, int grid,
                          hipStream_t st) {
  if (txy == 16) group_interp_t<T, P, 16>(t, grid, st);
  else if (txy == 8) group_interp_t<T, P, 8>(t, grid, st);
  else return -1;
  return 0;
}

template <typename T>
int launch_group_interp(const GroupStage<InterpArgs<T>> &t, int P, int txy, int grid,
                        hipStream_t st) {
  switch (P) {
    case 4: return group_interp_p<T, 4>(t, txy, grid, st);
    case 6: return group_interp_p<T, 6>(t, txy, grid, st);
    case 8: return group_interp_p<T, 8>(t, txy, grid, st);
  }
  return -1;
}

template <typename T>
int launch_group_real(const GroupStage<RealArgs<T>> &t, int C, int grid, hipStream_t st) {
  switch (C) {
#define X(CC)                                                                            \
  case CC:                                                                               \
    if (grid < t.total)                                                                  \
      hipLaunchKernelGGL((kg_real<T, CC, true>), dim3(grid), dim3(GROUP_ELEMS), 0, st, t); \
    else                                                                                 \
      hipLaunchKernelGGL((kg_real<T, CC, false>), dim3(t.total), dim3(GROUP_ELEMS), 0, st, \
                         t);                                                             \
    return 0;
    X(3) X(6) X(9) X(12) X(15) X(18) X(21) X(24) X(27) X(30) X(33) X(36) X(39) X(42) X(45)
    X(48)
#undef X
  }
  return -1;
}

#define SPME_INST(T)                                                                        \
  template void launch_group_permute<T>(const GroupStage<GPermArgs<T>> &, int, hipStream_t); \
  template int launch_group_spread<T>(const GroupStage<SpreadArgs<T>> &, int, int, int,     \
                                      hipStream_t);                                         \
  template int launch_group_fft<T>(const GroupStage<FftArgs<T>> &, int, int, hipStream_t);  \
  template int launch_group_interp<T>(const GroupStage<InterpArgs<T>> &, int, int, int,     \
                                      hipStream_t);                                         \
  template int launch_group_real<T>(const GroupStage<RealArgs<T>> &, int, int, hipStream_t);
SPME_INST(float)
SPME_INST(double)
#undef SPME_INST

}  // namespace spme
