// Host-side launch interface of the SPME kernels (spme_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace spme {

template <typename T> struct cpx;

// Sets the calling thread's message for spme_last_error(); returns code.
int set_last_error(int code, const char *msg);

struct WTable {
  double w[64];  // W[x*P + y], P <= 8
};

template <typename T> struct SpreadArgs {
  const int *boff;
  const int *pbase;
  const T *pw;
  const T *fs;
  int C, CC, P;
  T *grid;
  long comp_stride;
  int RS, K, ZC;
};

// TO: type of the output vectors and of the final beta*out + alpha*sum
// (T, or double when an FP32 mesh path forms its final sum in FP64).
template <typename T, typename TO = T> struct InterpArgs {
  const int *boff;
  const int *pbase;
  const T *pw;
  const int *perm;
  const T *grid;
  long comp_stride;
  int RS, K, N, C, P, ZC;
  TO alpha, beta;
  TO *out;
  long ldout;
};

// Spread: one wave per 8 x 8 column tile, z chunk of ZC planes, CC
// components per pass; SPREAD_MAXZC bounds ZC (LDS bucket table).
constexpr int SPREAD_MAXZC = 256;
inline int spread_blocks(int K, int ZC, int C, int CC) {
  return (K / 8) * (K / 8) * ((K + ZC - 1) / ZC) * (C / CC);
}
// Interpolation: a chunk of CC components (16 bytes, the last chunk may be
// partial) per block over a TXY x TXY tile and a z chunk of ZC planes.
template <typename T> struct InterpShape {
  static constexpr int CC = 16 / sizeof(T);
  static constexpr int MAXZC = 256;
  static constexpr int NT = 1024;
};
inline int interp_blocks(int K, int ZC, int C, int CC, int TXY) {
  return (K / TXY) * (K / TXY) * ((K + ZC - 1) / ZC) * ((C + CC - 1) / CC);
}
// FFT line batches: keep ~24 KB of LDS per block.
template <typename T> struct FftShape {
  static constexpr int NLX = sizeof(T) == 8 ? 8 : 16;  // complex lines per x block
  static constexpr int XW = 16;                        // y-pass columns
  static constexpr int XWZ = sizeof(T) == 8 ? 4 : 8;   // z-pass columns (x3 comps)
};

// Real space: lanes per row for C = 3R components, each lane taking 4, 2 or 1
// consecutive right-hand sides.
__host__ __device__ constexpr int real_rpl(int C) {
  return (C / 3) % 4 == 0 ? 4 : ((C / 3) % 2 == 0 ? 2 : 1);
}
__host__ __device__ constexpr int real_lpr(int C) { return C / 3 / real_rpl(C); }

// Blocks of the z pass for one right-hand side (3 components): chunks of XW
// x columns for every y.  When K/2 is a multiple of XW the chunks cover
// x < K/2 and the Nyquist column x = K/2 gets K/XW narrow blocks of XW rows
// of y each, instead of a chunk that is one column wide.
template <int XW> struct ZLayout {
  __host__ __device__ static constexpr bool nyq(int K) { return (K / 2) % XW == 0; }
  __host__ __device__ static constexpr int chunks(int K) {
    return nyq(K) ? K / 2 / XW : (K / 2 + XW) / XW;
  }
  __host__ __device__ static constexpr int per_rhs(int K) {
    return chunks(K) * K + (nyq(K) ? K / XW : 0);
  }
};

// Single-system z pass shape: FP32 meshes of 128 and 192 points use chunks
// of whole 128-byte lines (16 complex) with 512 threads; the rest (and the
// grouped kernels) use FftShape::XWZ columns with 256 threads.
template <typename T, int K> struct ZShape {
  static constexpr bool wide = sizeof(T) == 4 && (K == 128 || K == 192);
  static constexpr int XW = wide ? 16 : FftShape<T>::XWZ;
  static constexpr int NT = wide ? 512 : 256;
};

// Entries of the psi factor table: n = nx^2 + ny^2 + nz^2 <= 3 (K/2)^2.
__host__ __device__ constexpr int influence_npsi(int K) { return 3 * (K / 2) * (K / 2) + 1; }

template <typename T> struct FftArgs {
  T *grid;
  long comp_stride;
  int RS, K, ncomp;
  const T *tab;  // influence factors (psi over n^2, then |b|^2), see the z pass
  T dk;
  const cpx<T> *tw;
  int ncu;  // compute units (grid of the fused plane passes)
  // Nyquist column x = K/2 of the half spectrum as [component][z][y]
  // complex, used by the fused plane passes and the z pass; null: the column
  // stays in the spectrum rows (separate passes, grouped applies)
  cpx<T> *nyq;
  // unequal radii: ncomp = 6 per right-hand side (S_f, S_{a^2 f}) and tab
  // holds the radius-free factors A(k); see the z pass
  int poly;
};
// True when the single-system apply of this mesh uses the fused x + y plane
// passes (and then a separate Nyquist column).
// LDS bytes of one fused x + y plane: K/2 line pairs of 2(K/2 + 1) complex
// values, and the twiddles.
template <typename T> constexpr size_t xy_lds(int K) {
  return size_t((K / 2) * 2 * (K / 2 + 1) + K) * 2 * sizeof(T);
}
template <typename T> constexpr bool xy_fused(int K) { return xy_lds<T>(K) <= 160 * 1024; }
template <typename T> bool fft_fused(int K);

// Events recorded after each FFT launch and the pass each one ends
// (0 x or fused x+y forward, 1 y forward, 2 z + influence, 3 y inverse,
// 4 x or fused x+y inverse).
struct FftMarks {
  hipEvent_t ev[5];
  int pass[5];
  int n;
};

void launch_pbuild(const double *pos, int N, double L, int K, int P, const WTable &W, int TX,
                   int TY, int *base_tmp, double *w_tmp, int *key_tmp, int *counts, int *err,
                   hipStream_t st);
// Exclusive scan of n <= 2^22 ints, out[n] = total; sums: 1024 ints of scratch.
int launch_scan(const int *in, int *out, int n, int *sums, hipStream_t st);
void launch_pscatter(const int *key, const int *boff, int *cursor, int *perm, int N,
                     hipStream_t st);
// buckets above SORT_MAX particles are rebuilt by a stable pass over the keys
constexpr int SORT_MAX = 256;
inline long bucket_queue_ints(int N) { return N / (SORT_MAX + 1) + 2; }
int launch_bucket_sort(const int *key, int N, const int *boff, int *perm, int nkeys, int *bigq,
                       int *tmp, hipStream_t st);
template <typename T>
void launch_pgather(const int *perm, const int *base_tmp, const double *w_tmp, int *pbase,
                    T *pw, int N, int P3, hipStream_t st);
template <typename T>
void launch_permute_in(const int *ia, T *fa, const int *ib, T *fb, const T *in, long ldin, int N,
                       int C, hipStream_t st, const double *rad = nullptr);
// Unequal radii: out_r = beta out_r + alpha (g[2r] - a_i^2 g[2r+1]).
template <typename T, typename TO>
void launch_combine_radii(const T *g, const double *rad, int N, int nrhs, double alpha,
                          double beta, TO *out, long ldout, hipStream_t st);
void launch_invert(const int *perm, int *inv, int N, hipStream_t st);
template <typename T> int launch_spread(const SpreadArgs<T> &a, hipStream_t st);
template <typename T, typename TO> int launch_interp(const InterpArgs<T, TO> &a, hipStream_t st);
// mk (optional): events to record after each launch (mk->n = 0 on entry).
template <typename T>
int launch_fft_influence(const FftArgs<T> &a, hipStream_t st, FftMarks *mk = nullptr);
// Unnormalised 3-D transform without influence: dir -1 forward (x r2c, y, z),
// dir +1 backward (z, y, x c2r).
template <typename T> int launch_fft_plain(const FftArgs<T> &a, int dir, hipStream_t st);

// Plain FP64 path with the reference's data structures (spme_dense.hip).
int launch_dense_P(const int *base, const double *w, int N, int K, int RS, int P, int ldP,
                   int ldind, double *Pm, int *ind, hipStream_t st);
int launch_dense_spread(const double *Pm, const int *ind, int N, int P, int ldP, int ldind,
                        const double *f, double *grid, long cs, hipStream_t st);
void launch_dense_influence(double *grid, long cs, int K, int Khp, const double *lm2,
                            double dk, hipStream_t st);
void launch_dense_scale(double *grid, long n, double s, hipStream_t st);
int launch_dense_interp(const double *Pm, const int *ind, int N, int P, int ldP, int ldind,
                        const double *grid, long cs, double alpha, double beta, double *u,
                        hipStream_t st);
bool fft_size_supported(int K);
template <typename T>
void launch_influence_table(const double *bsq, int K, int Khp, double dk, double xi,
                            double inv_v, T *tab, hipStream_t st);
template <typename T>
void launch_influence_factors(const double *bsq, int K, double dk, double xi, double inv_v,
                              int radius_free, T *fac, hipStream_t st);

}  // namespace spme
