// Host engine and C ABI of the MI355X SPME mobility product.
//
// Drop-in surface: include/spme.h (old-version/spme.h:85-101).
// Extensions:      include/spme_ext.h.
//
// One engine owns one HIP stream and all device buffers; calls on one engine
// are not re-entrant (the mesh is shared scratch), separate engines are
// independent.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/spme.h"
#include "../../include/spme_ext.h"
#include "spme_group.h"
#include "spme_kernels.h"
#include "spme_launch.h"
#include "spme_real.h"

using namespace spme;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

}  // namespace

// The error message of the calling thread, for the other translation units
// of the library (spme_last_error() returns it).
int spme::set_last_error(int code, const char *msg) {
  g_err = msg;
  return code;
}

namespace {

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(-2, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                  __LINE__);                                                             \
  } while (0)

// ---------------------------------------------------------------------------
// B-spline tables from the closed form (the reference stores them as
// literals for p = 4, 6, old-version/spme.c:14-40).  M_p is the cardinal
// B-spline on [0, p]; W[x][y] is the coefficient of g^y in
// M_p(g + 1/2 + p - 1 - x), g in [-1/2, 1/2).  Every intermediate is an
// integer multiple of 2^-(p-1) below 2^53, so the sums are exact and each
// entry is the correctly rounded quotient.

double binom(int n, int k) {
  double r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return std::round(r);
}

double factorial(int n) {
  double r = 1;
  for (int i = 2; i <= n; ++i) r *= i;
  return r;
}

// M_p(k) for integer k, times (p-1)!.
double euler_num(int p, int k) {
  double s = 0;
  for (int j = 0; j <= k; ++j) {
    if (k - j <= 0) continue;
    s += ((j & 1) ? -1.0 : 1.0) * binom(p, j) * std::pow(double(k - j), p - 1);
  }
  return s;
}

void make_W(int p, double *W) {
  const double fac = factorial(p - 1);
  for (int x = 0; x < p; ++x) {
    double coef[8] = {0};
    for (int j = 0; j <= p - 1 - x; ++j) {
      const double a = p - 0.5 - x - j;  // (g + a)^(p-1)
      const double sj = ((j & 1) ? -1.0 : 1.0) * binom(p, j);
      for (int y = 0; y < p; ++y)
        coef[y] += sj * binom(p - 1, y) * std::pow(a, p - 1 - y);
    }
    for (int y = 0; y < p; ++y) W[x * p + y] = coef[y] / fac;
  }
}

// |b(m)|^2 = 1/|sum_{i<p-1} M_p(i+1) e^{i phi i}|^2, phi = 2 pi m/K with m in
// FFT order (old-version/spme.c:109-126; only the modulus enters lm2).
void make_bsq(int K, int p, std::vector<double> &bsq) {
  const double fac = factorial(p - 1);
  bsq.resize(K);
  for (int x = 0; x < K; ++x) {
    const int m = x <= K / 2 ? x : x - K;
    const double phi = 2.0 * M_PI * m / K;
    double dr = 0, di = 0;
    for (int i = 0; i < p - 1; ++i) {
      const double s = euler_num(p, i + 1) / fac;
      dr += s * std::cos(i * phi);
      di += s * std::sin(i * phi);
    }
    bsq[x] = 1.0 / (dr * dr + di * di);
  }
}

}  // namespace

// ---------------------------------------------------------------------------

struct StageEvents {
  hipEvent_t start;
  std::vector<std::pair<int, hipEvent_t>> marks;
};

struct _spme_t {
  spme_params_t prm;
  int K, P, N, prec, max_rhs, Kh, Khp, RS;
  bool dense;  // SPME_FP64_PLAIN: the reference's data structures
  long comp_stride;  // reals
  size_t tsz;        // bytes per real
  hipStream_t stream;
  int device;
  bool have_pos;
  // mesh and tables
  void *grid;  // 3*max_rhs components
  void *nyq;   // Nyquist column of the fused FFT path, [3*max_rhs][K][K] complex (or null)
  void *tab;   // plain path: influence table [K][K][Khp]; otherwise its factors
  void *tw;    // twiddles exp(-2 pi i t/K), K complex
  // P data
  int nkeys;
  double *pos;  // device positions (N x 3)
  int *base_tmp, *key_tmp, *counts, *cursor, *boff, *perm, *pbase, *err;
  int *bigq;  // queue of buckets too large for the per-thread sort (count first)
  double *w_tmp;
  void *pw;  // sorted weights [N][3][P]
  void *fs;  // sorted forces [N][3*max_rhs]
  // plain path: dense p^3 weights and int32 mesh offsets per particle
  int ldP, ldind;
  double *dP;
  int *dind;
  // real space
  int nc;
  int *r_key, *r_counts, *r_cursor, *r_off, *r_perm;
  int *iperm, *r_iperm;  // inverses of perm and r_perm (original -> sorted)
  double *r_pos;
  float4 *r_cpos;  // FP32 engines: cell-relative positions for the SpMV
  int *scan_sums;  // scratch of the device scans (1024 tile totals)
  void *r_fs;
  // real-space pair blocks (sliced ELL, 64 rows per slice)
  int nslices;
  int *r_cnt, *r_width, *r_size, *r_soff, *r_col;
  void *r_blk;
  long r_E, r_cap, r_used;
  // drop-in staging (FP64 host API), h_rhs right-hand sides of 3N values
  double *h_in, *h_out;
  int h_rhs;
  // per-kernel timing: when enabled, events are recorded on the apply
  // stream after every launch; a stage's time is the gap to the event before
  int timing;
  std::vector<StageEvents> pending;
  std::vector<hipEvent_t> pool;
  double stage_ms[SPME_NSTAGES];
  int stage_n[SPME_NSTAGES];
  // requested tile shapes (0 = default) and a counter of position sets, so
  // that groups holding this engine notice a rebuild of its buffers
  spme_tiles_t tiles;
  unsigned pos_version;
  int ncu;  // compute units of the device
  // unequal radii (prm.radii != 0): the mesh carries Cm = 6 components per
  // right-hand side (f and a^2 f), rad the radii (original order, 1 by
  // default), r_rad in cell order, selfv the per-row self coefficients,
  // tmp the interpolated influence outputs [2 nrhs][3N] before combining
  int poly, Cm;
  double *rad, *r_rad;
  void *selfv, *tmp;
  double cut;    // pair search radius: max(rc, 2 a_max), at least 2
  int nc_alloc;  // cells per side the cell arrays were sized for
};

namespace {

template <typename T>
int upload_tables(_spme_t *e) {
  const int K = e->K;
  std::vector<double> bsq;
  make_bsq(K, e->P, bsq);
  const double L = e->prm.box_size, V = L * L * L, xi = e->prm.xi;
  double *d_bsq = nullptr;
  HIP_TRY(hipMalloc(&d_bsq, sizeof(double) * K));
  HIP_TRY(hipMemcpy(d_bsq, bsq.data(), sizeof(double) * K, hipMemcpyHostToDevice));
  if (e->dense) {
    // the plain path keeps the reference's lm2 = psi |b|^2 table and scales
    // by 1/K^3 and alpha K^3/V in passes of their own
    HIP_TRY(hipMalloc(&e->tab, sizeof(T) * (size_t)K * K * e->Khp));
    launch_influence_table<T>(d_bsq, K, e->Khp, 2.0 * M_PI / L, xi, 1.0, (T *)e->tab,
                              e->stream);
  } else {
    // the z pass forms psi |b|^2 / V from its factors
    HIP_TRY(hipMalloc(&e->tab, sizeof(T) * (size_t)(influence_npsi(K) + K)));
    // unequal radii: the factors of A(k), psi without its radius term
    launch_influence_factors<T>(d_bsq, K, 2.0 * M_PI / L, xi, 1.0 / V, e->poly, (T *)e->tab,
                                e->stream);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(e->stream));
  HIP_TRY(hipFree(d_bsq));
  std::vector<T> tw(2 * K);
  for (int t = 0; t < K; ++t) {
    const double a = -2.0 * M_PI * t / K;
    tw[2 * t] = T(std::cos(a));
    tw[2 * t + 1] = T(std::sin(a));
  }
  HIP_TRY(hipMalloc(&e->tw, tw.size() * sizeof(T)));
  HIP_TRY(hipMemcpy(e->tw, tw.data(), tw.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

int validate(const spme_params_t *p) {
  if (!p) return fail(-1, "null parameters");
  if (!(p->porder == 4 || p->porder == 6 || p->porder == 8))
    return fail(-1, "porder must be 4, 6 or 8 (got %d)", p->porder);
  if (!fft_size_supported(p->dim))
    return fail(-1, "dim %d not supported (smooth sizes 16..256 of the form 2^a 3^b, "
                    "multiple of 8)", p->dim);
  if (4 * p->porder >= p->dim)
    return fail(-1, "dim must be larger than 4 * porder (old-version/main.c:205)");
  if (p->np <= 0) return fail(-1, "np must be positive");
  if (!(p->box_size > 0.0) || !std::isfinite(p->box_size))
    return fail(-1, "box_size must be positive");
  if (!(p->xi > 0.0) || !std::isfinite(p->xi)) return fail(-1, "xi must be positive");
  if (p->precision != SPME_FP64 && p->precision != SPME_FP32 &&
      p->precision != SPME_FP64_PLAIN && p->precision != SPME_FP32_SUM64)
    return fail(-1, "precision must be SPME_FP64, SPME_FP32, SPME_FP64_PLAIN or "
                    "SPME_FP32_SUM64");
  if (p->max_rhs < 1 || p->max_rhs > 16) return fail(-1, "max_rhs must be in [1, 16]");
  if (p->radii) {
    if (p->precision == SPME_FP64_PLAIN)
      return fail(-1, "unequal radii are not available on the plain path");
    if (p->max_rhs > 8)
      return fail(-1, "with unequal radii max_rhs must be in [1, 8] (six mesh components per "
                      "right-hand side)");
  }
  // the spreading stages particle data by 32-bit dword offsets into the
  // sorted force and weight arrays (at most N * 48 values of 2 dwords)
  if (p->np > (1 << 23)) return fail(-1, "np must be at most %d (got %d)", 1 << 23, p->np);
  return 0;
}

void free_all(_spme_t *e) {
  void *ptrs[] = {e->grid, e->nyq, e->tab, e->tw, e->pos, e->base_tmp, e->key_tmp, e->counts,
                  e->cursor, e->boff, e->perm, e->pbase, e->err, e->w_tmp, e->pw, e->fs,
                  e->r_key, e->r_counts, e->r_cursor, e->r_off, e->r_perm, e->r_pos,
                  e->r_cpos, e->r_fs, e->r_cnt, e->r_width, e->r_size, e->r_soff, e->r_col, e->r_blk,
                  e->h_in, e->h_out, e->dP, e->dind, e->iperm, e->r_iperm, e->bigq,
                  e->rad, e->r_rad, e->selfv, e->tmp, e->scan_sums};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (auto &r : e->pending) {
    e->pool.push_back(r.start);
    for (auto &m : r.marks) e->pool.push_back(m.second);
  }
  for (hipEvent_t ev : e->pool) (void)hipEventDestroy(ev);
  if (e->stream) (void)hipStreamDestroy(e->stream);
}

int engine_create(const spme_params_t *prm, _spme_t **out) {
  if (int rc = validate(prm)) return rc;
  _spme_t *e = new _spme_t();  // value-initialised: every pointer starts null
  e->prm = *prm;
  e->K = prm->dim;
  e->P = prm->porder;
  e->N = prm->np;
  e->prec = prm->precision;
  e->dense = prm->precision == SPME_FP64_PLAIN;
  e->max_rhs = prm->max_rhs;
  e->Kh = e->K / 2 + 1;
  // rows padded to whole 128-byte lines in FP32 (16 complex) so that the
  // column chunks of the y and z passes read whole lines
  e->Khp = (e->Kh + 15) / 16 * 16;
  e->RS = 2 * e->Khp;
  e->comp_stride = (long)e->K * e->K * e->RS;
  // the mesh path runs in FP32 for SPME_FP32 and SPME_FP32_SUM64
  e->tsz = (e->prec == SPME_FP32 || e->prec == SPME_FP32_SUM64) ? 4 : 8;
  e->poly = prm->radii ? 1 : 0;
  e->Cm = e->poly ? 6 : 3;
  // C: mesh components; the real-space forces keep 3 per right-hand side
  const int K = e->K, N = e->N, C = e->Cm * e->max_rhs;
  auto bail = [&](int rc) {
    free_all(e);
    delete e;
    return rc;
  };
#define E_TRY(expr)                                                                      \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return bail(fail(-2, "%s failed: %s", #expr, hipGetErrorString(e_)));              \
  } while (0)
  E_TRY(hipGetDevice(&e->device));
  E_TRY(hipDeviceGetAttribute(&e->ncu, hipDeviceAttributeMultiprocessorCount, e->device));
  E_TRY(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  // the plain path transforms one right-hand side (3 components) at a time
  E_TRY(hipMalloc(&e->grid, (size_t)(e->dense ? 3 : C) * e->comp_stride * e->tsz));
  const bool fused = e->tsz == 4 ? fft_fused<float>(K) : fft_fused<double>(K);
  // a separate Nyquist column array for the large fused meshes only (the
  // grouped kernels, which serve K <= 96, keep it in the spectrum rows)
  if (!e->dense && fused && K >= 128)
    E_TRY(hipMalloc(&e->nyq, (size_t)C * K * K * 2 * e->tsz));
  int rc = e->tsz == 4 ? upload_tables<float>(e) : upload_tables<double>(e);
  if (rc) return bail(rc);
  e->nkeys = (K / 8) * (K / 8) * K;
  E_TRY(hipMalloc(&e->pos, sizeof(double) * 3 * N));
  E_TRY(hipMalloc(&e->base_tmp, sizeof(int) * N));
  E_TRY(hipMalloc(&e->key_tmp, sizeof(int) * N));
  E_TRY(hipMalloc(&e->counts, sizeof(int) * e->nkeys));
  E_TRY(hipMalloc(&e->cursor, sizeof(int) * e->nkeys));
  E_TRY(hipMalloc(&e->bigq, sizeof(int) * bucket_queue_ints(N)));
  E_TRY(hipMalloc(&e->boff, sizeof(int) * (e->nkeys + 1)));
  E_TRY(hipMalloc(&e->scan_sums, sizeof(int) * 1024));
  E_TRY(hipMalloc(&e->perm, sizeof(int) * N));
  E_TRY(hipMalloc(&e->iperm, sizeof(int) * N));
  E_TRY(hipMalloc(&e->pbase, sizeof(int) * N));
  E_TRY(hipMalloc(&e->err, sizeof(int)));
  E_TRY(hipMalloc(&e->w_tmp, sizeof(double) * 3 * e->P * N));
  if (e->dense) {
    // PAD_LEN(p^3, 8) and PAD_LEN(p^3, 4) (old-version/spme.c:457-460)
    const int p3 = e->P * e->P * e->P;
    e->ldP = (p3 + 7) / 8 * 8;
    e->ldind = (p3 + 15) / 16 * 16;
    E_TRY(hipMalloc(&e->dP, sizeof(double) * (size_t)e->ldP * N));
    E_TRY(hipMalloc(&e->dind, sizeof(int) * (size_t)e->ldind * N));
  } else {
    E_TRY(hipMalloc(&e->pw, e->tsz * 3 * e->P * N));
    E_TRY(hipMalloc(&e->fs, e->tsz * (size_t)C * N));
  }
  const double L = prm->box_size;
  // Real-space cell list only when a cutoff is given (the drop-in surface is
  // reciprocal-only).  Cells are at least max(rc, 2) wide so that every pair
  // within the cutoff or overlapping lies in adjacent cells.
  e->cut = prm->rc > 2.0 ? prm->rc : 2.0;
  e->nc = prm->rc > 0.0 ? int(std::floor(L / e->cut)) : 0;
  if (e->nc > 128) e->nc = 128;
  e->nc_alloc = e->nc;
  if (e->poly) {
    std::vector<double> ones(N, 1.0);
    E_TRY(hipMalloc(&e->rad, sizeof(double) * N));
    E_TRY(hipMemcpy(e->rad, ones.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    E_TRY(hipMalloc(&e->tmp, e->tsz * (size_t)C * N));
  }
  if (e->nc >= 3) {
    const int nc3 = e->nc * e->nc * e->nc;
    E_TRY(hipMalloc(&e->r_key, sizeof(int) * N));
    E_TRY(hipMalloc(&e->r_counts, sizeof(int) * nc3));
    E_TRY(hipMalloc(&e->r_cursor, sizeof(int) * nc3));
    E_TRY(hipMalloc(&e->r_off, sizeof(int) * (nc3 + 1)));
    E_TRY(hipMalloc(&e->r_perm, sizeof(int) * N));
    E_TRY(hipMalloc(&e->r_iperm, sizeof(int) * N));
    E_TRY(hipMalloc(&e->r_pos, sizeof(double) * 3 * N));
    if (e->tsz == 4) E_TRY(hipMalloc(&e->r_cpos, sizeof(float4) * N));
    E_TRY(hipMalloc(&e->r_fs, e->tsz * (size_t)3 * e->max_rhs * N));
    e->nslices = (N + 63) / 64;
    E_TRY(hipMalloc(&e->r_cnt, sizeof(int) * N));
    E_TRY(hipMalloc(&e->r_width, sizeof(int) * e->nslices));
    E_TRY(hipMalloc(&e->r_size, sizeof(int) * e->nslices));
    E_TRY(hipMalloc(&e->r_soff, sizeof(int) * (e->nslices + 1)));
    if (e->poly) {
      E_TRY(hipMalloc(&e->r_rad, sizeof(double) * N));
      E_TRY(hipMalloc(&e->selfv, e->tsz * N));
    }
  }
#undef E_TRY
  *out = e;
  return 0;
}

template <typename T>
int build_P(_spme_t *e, hipStream_t st) {
  const int N = e->N, K = e->K, P = e->P;
  WTable W;
  make_W(P, W.w);
  HIP_TRY(hipMemsetAsync(e->counts, 0, sizeof(int) * e->nkeys, st));
  HIP_TRY(hipMemsetAsync(e->cursor, 0, sizeof(int) * e->nkeys, st));
  HIP_TRY(hipMemsetAsync(e->err, 0, sizeof(int), st));
  launch_pbuild(e->pos, N, e->prm.box_size, K, P, W, 8, 8, e->base_tmp, e->w_tmp, e->key_tmp,
                e->counts, e->err, st);
  if (e->dense) {
    if (launch_dense_P(e->base_tmp, e->w_tmp, N, K, e->RS, P, e->ldP, e->ldind, e->dP, e->dind,
                       st))
      return fail(-1, "dense P: unsupported porder");
  } else {
    if (launch_scan(e->counts, e->boff, e->nkeys, e->scan_sums, st))
      return fail(-1, "scan of more than 2^22 values");
    launch_pscatter(e->key_tmp, e->boff, e->cursor, e->perm, N, st);
    if (launch_bucket_sort(e->key_tmp, N, e->boff, e->perm, e->nkeys, e->bigq, e->iperm, st))
      return fail(-2, "bucket sort launch failed");
    launch_invert(e->perm, e->iperm, N, st);
    launch_pgather<T>(e->perm, e->base_tmp, e->w_tmp, e->pbase, (T *)e->pw, N, 3 * P, st);
  }
  if (e->nc >= 3) {
    const int nc3 = e->nc * e->nc * e->nc;
    HIP_TRY(hipMemsetAsync(e->r_counts, 0, sizeof(int) * nc3, st));
    HIP_TRY(hipMemsetAsync(e->r_cursor, 0, sizeof(int) * nc3, st));
    launch_cell_key(e->pos, N, e->prm.box_size, e->nc, e->r_key, e->r_counts, st);
    if (launch_scan(e->r_counts, e->r_off, nc3, e->scan_sums, st))
      return fail(-1, "scan of more than 2^22 values");
    launch_pscatter(e->r_key, e->r_off, e->r_cursor, e->r_perm, N, st);
    if (launch_bucket_sort(e->r_key, N, e->r_off, e->r_perm, nc3, e->bigq, e->r_iperm, st))
      return fail(-2, "cell sort launch failed");
    launch_invert(e->r_perm, e->r_iperm, N, st);
    launch_cell_gather(e->r_perm, e->pos, e->r_pos, e->poly ? e->rad : nullptr, e->r_rad, N, st);
    if (e->r_cpos) launch_cell_rel(e->r_pos, N, e->prm.box_size, e->nc, e->r_cpos, st);
    if (e->poly) launch_self_radii<T>(e->r_rad, N, e->prm.xi, (T *)e->selfv, st);
    const double L = e->prm.box_size, rc = e->prm.rc;
    const double cut = e->cut;
    launch_pair_count(e->r_pos, e->r_off, e->nc, L, cut * cut, N, e->r_cnt, st);
    launch_slice_width(e->r_cnt, N, e->r_width, e->r_size, st);
    // the entry offsets are ints: sum the slice sizes in 64 bits first and
    // refuse a block list that would not fit
    std::vector<int> hw(e->nslices);
    HIP_TRY(hipMemcpyAsync(hw.data(), e->r_width, sizeof(int) * e->nslices,
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    long total = 0;
    for (int w : hw) total += 64L * w;
    if (total > INT_MAX)
      return fail(-1,
                  "real-space block list needs %ld entries (at most %d): the particles are "
                  "too clustered for rc = %g",
                  total, INT_MAX, rc);
    if (launch_scan(e->r_size, e->r_soff, e->nslices, e->scan_sums, st))
      return fail(-1, "scan of more than 2^22 values");
    if (total > e->r_cap) {
      (void)hipFree(e->r_col);
      (void)hipFree(e->r_blk);
      e->r_col = nullptr;
      e->r_blk = nullptr;
      e->r_cap = 0;
      const long cap = total + total / 8 + 64;
      HIP_TRY(hipMalloc(&e->r_col, sizeof(int) * cap));
      HIP_TRY(hipMalloc(&e->r_blk, sizeof(T) * 2 * cap));
      e->r_cap = cap;
    }
    e->r_E = e->r_cap;
    e->r_used = total;
    launch_pair_fill<T>(e->r_pos, e->r_off, e->nc, L, e->prm.xi, rc * rc, cut * cut, N,
                        e->r_width, e->r_soff, e->r_col, e->poly ? e->r_rad : nullptr,
                        (T *)e->r_blk, e->err, st);
  }
  HIP_TRY(hipGetLastError());
  return 0;
}

int set_positions(_spme_t *e, const double *pos, int on_device, hipStream_t st) {
  if (!pos) return fail(-1, "positions are NULL");
  // from here on the P data and the block list are being rebuilt: until
  // every step and check has passed the engine holds no usable positions,
  // and groups holding it see a new version
  e->have_pos = false;
  ++e->pos_version;
  (void)hipGetLastError();
  HIP_TRY(hipMemcpyAsync(e->pos, pos, sizeof(double) * 3 * e->N,
                         on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
  int rc = e->tsz == 4 ? build_P<float>(e, st) : build_P<double>(e, st);
  if (rc) return rc;
  int herr = 0;
  HIP_TRY(hipMemcpyAsync(&herr, e->err, sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (herr & 1) return fail(-1, "positions must be finite and within [-L, 2L)");
  if (herr & 2)
    return fail(-2, "real-space block list: a row found more neighbours than its count pass");
  e->have_pos = true;
  return 0;
}

// Radii (N values, FP64, > 0) for an engine created with radii: kept in
// original order; the pair search radius becomes max(rc, 2, 2 a_max), so
// the cells may grow (never beyond the arrays sized at create).  The P data
// and block list depend on the radii: positions must be set again.
int set_radii(_spme_t *e, const double *rad, int on_device, hipStream_t st) {
  if (!e->poly) return fail(-1, "the engine was created without radii (spme_params_t.radii)");
  if (!rad) return fail(-1, "radii are NULL");
  e->have_pos = false;
  ++e->pos_version;
  std::vector<double> h(e->N);
  if (on_device) {
    HIP_TRY(hipMemcpyAsync(h.data(), rad, sizeof(double) * e->N, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
  } else {
    std::copy(rad, rad + e->N, h.begin());
  }
  double amax = 0.0;
  for (double a : h) {
    if (!(a > 0.0) || !std::isfinite(a)) return fail(-1, "radii must be positive and finite");
    amax = std::max(amax, a);
  }
  const double rc = e->prm.rc;
  const double cut = std::max(std::max(rc, 2.0), 2.0 * amax);
  if (e->nc_alloc >= 3) {
    int nc = int(std::floor(e->prm.box_size / cut));
    nc = std::min(nc, e->nc_alloc);
    if (nc < 3)
      return fail(-1, "radii up to %g need box_size >= 3 * max(rc, 2 a_max) = %g for the "
                      "real-space part", amax, 3.0 * cut);
    e->nc = nc;
  }
  e->cut = cut;
  HIP_TRY(hipMemcpyAsync(e->rad, h.data(), sizeof(double) * e->N, hipMemcpyHostToDevice, st));
  HIP_TRY(hipStreamSynchronize(st));
  return 0;
}

// Room for batches of up to R right-hand sides: the mesh, Nyquist column,
// sorted forces and (with radii) interpolation buffers are reallocated; the
// P data and block list do not depend on the batch width and stay.  Groups
// holding the engine see a new position version and rebuild their tables.
int engine_grow(_spme_t *e, int R) {
  if (R <= e->max_rhs) return 0;
  if (R > 16 || (e->poly && R > 8)) return fail(-1, "batch of %d right-hand sides too large", R);
  HIP_TRY(hipStreamSynchronize(e->stream));
  const size_t C = (size_t)e->Cm * R;
  void **bufs[] = {&e->grid, &e->nyq, (void **)&e->fs, &e->r_fs, &e->tmp};
  const bool had[] = {e->grid != nullptr, e->nyq != nullptr, e->fs != nullptr,
                      e->r_fs != nullptr, e->tmp != nullptr};
  const size_t bytes[] = {(e->dense ? 3 : C) * e->comp_stride * e->tsz,
                          C * e->K * e->K * 2 * e->tsz, e->tsz * C * e->N,
                          e->tsz * (size_t)3 * R * e->N, e->tsz * C * e->N};
  for (int b = 0; b < 5; ++b) {
    if (!had[b]) continue;
    (void)hipFree(*bufs[b]);
    *bufs[b] = nullptr;
  }
  e->max_rhs = e->prm.max_rhs = R;
  ++e->pos_version;
  for (int b = 0; b < 5; ++b)
    if (had[b]) HIP_TRY(hipMalloc(bufs[b], bytes[b]));
  return 0;
}

hipEvent_t take_event(_spme_t *e) {
  if (!e->pool.empty()) {
    hipEvent_t ev = e->pool.back();
    e->pool.pop_back();
    return ev;
  }
  hipEvent_t ev = nullptr;
  (void)hipEventCreate(&ev);
  return ev;
}

int collect_timing(_spme_t *e) {
  for (auto &r : e->pending) {
    hipEvent_t prev = r.start;
    for (auto &m : r.marks) {
      HIP_TRY(hipEventSynchronize(m.second));
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, prev, m.second));
      e->stage_ms[m.first] += ms;
      e->stage_n[m.first] += 1;
      prev = m.second;
    }
    e->pool.push_back(r.start);
    for (auto &m : r.marks) e->pool.push_back(m.second);
  }
  e->pending.clear();
  return 0;
}

int pick_cc(int C, size_t tsz) {
  if (tsz == 4 && C % 12 == 0) return 12;
  return C % 6 == 0 ? 6 : 3;
}

// Tile shapes of one system for C components, defaults resolved.
struct Tiles {
  int spread_cc, spread_zc, interp_zc;
};

Tiles resolve_tiles(const _spme_t *e, const spme_tiles_t &t, int C) {
  const int K = e->K;
  Tiles r;
  r.spread_cc = (t.spread_cc > 0 && C % t.spread_cc == 0) ? t.spread_cc : pick_cc(C, e->tsz);
  r.spread_zc = t.spread_zc > 0 ? std::min(t.spread_zc, K) : (K >= 64 ? K / 4 : K / 2);
  r.interp_zc = t.interp_zc > 0 ? std::min(t.interp_zc, K) : (K >= 128 ? K / 2 : K);
  return r;
}

int check_tiles(const spme_tiles_t &t) {
  if (!(t.spread_cc == 0 || t.spread_cc == 3 || t.spread_cc == 6 || t.spread_cc == 12))
    return fail(-1, "spread_cc must be 0, 3, 6 or 12 (got %d)", t.spread_cc);
  if (t.spread_zc < 0 || t.spread_zc > SPREAD_MAXZC)
    return fail(-1, "spread_zc must be in [0, %d] (got %d)", SPREAD_MAXZC, t.spread_zc);
  if (t.interp_zc < 0 || t.interp_zc > InterpShape<float>::MAXZC)
    return fail(-1, "interp_zc must be in [0, %d] (got %d)", InterpShape<float>::MAXZC,
                t.interp_zc);
  for (int v : {t.spread_persist, t.interp_persist, t.real_persist})
    if (v < 0 || v > 64) return fail(-1, "persist values must be in [0, 64] (got %d)", v);
  return 0;
}

// T: mesh and input type; TO: output type (double for SPME_FP32_SUM64)
template <typename T, typename TO>
int apply_t(_spme_t *e, int nrhs, double alpha, const T *in, long ldin, double beta, TO *out,
            long ldout, int parts, hipStream_t st, int stop_after) {
  const int N = e->N, C = 3 * nrhs;
  if ((parts & SPME_SELF) && !(parts & SPME_REAL))
    return fail(-1, "SPME_SELF without SPME_REAL is not supported");
  if ((parts & SPME_REAL) && e->nc < 3)
    return fail(-1, "real-space part needs rc > 0 and box_size >= 3*max(rc, 2)");
  StageEvents rec;
  const bool timed = e->timing && stop_after == 0;
  if (timed) {
    rec.start = take_event(e);
    HIP_TRY(hipEventRecord(rec.start, st));
  }
  auto mark = [&](int stage) -> int {
    if (!timed) return 0;
    hipEvent_t ev = take_event(e);
    HIP_TRY(hipEventRecord(ev, st));
    rec.marks.push_back({stage, ev});
    return 0;
  };
  if ((parts & SPME_RECIP) && e->dense) {
    if (stop_after) return fail(-1, "debug stages are not available on the plain path");
    // one right-hand side at a time (old-version/spme.c:631-702)
    const int K = e->K;
    const long cs = e->comp_stride;
    const double L = e->prm.box_size;
    const double *fin = reinterpret_cast<const double *>(in);
    double *uout = reinterpret_cast<double *>(out);
    double *grid = (double *)e->grid;
    FftArgs<double> fa{grid, cs, e->RS, K, 3, (const double *)e->tab, 2.0 * M_PI / L,
                       (const cpx<double> *)e->tw, e->ncu, nullptr};
    const double k3 = double(K) * K * K;
    for (int r = 0; r < nrhs; ++r) {
      HIP_TRY(hipMemsetAsync(grid, 0, sizeof(double) * 3 * cs, st));
      if (launch_dense_spread(e->dP, e->dind, N, e->P, e->ldP, e->ldind, fin + r * ldin, grid,
                              cs, st))
        return fail(-1, "spread: unsupported porder");
      mark(SPME_ST_SPREAD);
      if (launch_fft_plain<double>(fa, -1, st)) return fail(-1, "fft: unsupported size");
      mark(SPME_ST_FFT_X_FWD);
      launch_dense_influence(grid, cs, K, e->Khp, (const double *)e->tab, 2.0 * M_PI / L, st);
      mark(SPME_ST_FFT_Z_INFL);
      launch_fft_plain<double>(fa, +1, st);
      launch_dense_scale(grid, 3 * cs, 1.0 / k3, st);
      mark(SPME_ST_FFT_X_INV);
      launch_dense_interp(e->dP, e->dind, N, e->P, e->ldP, e->ldind, grid, cs,
                          alpha * k3 / (L * L * L), beta, uout + r * ldout, st);
      mark(SPME_ST_INTERP);
    }
  }
  // mesh components: 3 per right-hand side, 6 with unequal radii (f, a^2 f)
  const int Cv = e->Cm * nrhs;
  const Tiles tl = resolve_tiles(e, e->tiles, Cv);
  // one pass writes the forces in both sorted orders when both are needed
  const bool both = (parts & SPME_RECIP) && !e->dense && (parts & SPME_REAL);
  if ((parts & SPME_RECIP) && !e->dense) {
    launch_permute_in<T>(e->iperm, (T *)e->fs, both ? e->r_iperm : nullptr,
                         both ? (T *)e->r_fs : nullptr, in, ldin, N, C, st,
                         e->poly ? e->rad : nullptr);
    mark(SPME_ST_PERMUTE);
    SpreadArgs<T> sa{e->boff, e->pbase, (const T *)e->pw, (const T *)e->fs, Cv, tl.spread_cc,
                     e->P, (T *)e->grid, e->comp_stride, e->RS, e->K, tl.spread_zc};
    if (launch_spread<T>(sa, st)) return fail(-1, "spread: unsupported configuration");
    mark(SPME_ST_SPREAD);
    if (stop_after == 1) return 0;
    FftArgs<T> fa{(T *)e->grid, e->comp_stride, e->RS, e->K, Cv, (const T *)e->tab,
                  T(2.0 * M_PI / e->prm.box_size), (const cpx<T> *)e->tw, e->ncu,
                  (cpx<T> *)e->nyq, e->poly};
    FftMarks fm{};
    if (timed)
      for (int i = 0; i < 5; ++i) fm.ev[i] = take_event(e);
    if (launch_fft_influence<T>(fa, st, timed ? &fm : nullptr))
      return fail(-1, "fft: unsupported size");
    if (timed) {
      for (int i = 0; i < fm.n; ++i) rec.marks.push_back({SPME_ST_FFT_X_FWD + fm.pass[i], fm.ev[i]});
      for (int i = fm.n; i < 5; ++i) e->pool.push_back(fm.ev[i]);
    }
    if (stop_after == 2) return 0;
    if (e->poly) {
      // out_f and out_a of every right-hand side into tmp, then
      // out = beta out + alpha (I[out_f] - a_i^2 I[out_a])
      InterpArgs<T, T> ia{e->boff, e->pbase, (const T *)e->pw, e->perm, (const T *)e->grid,
                          e->comp_stride, e->RS, e->K, N, Cv, e->P, tl.interp_zc, T(1), T(0),
                          (T *)e->tmp, 3L * N};
      if (launch_interp(ia, st)) return fail(-1, "interpolate: unsupported configuration");
      launch_combine_radii<T, TO>((const T *)e->tmp, e->rad, N, nrhs, alpha, beta, out, ldout,
                                  st);
    } else {
      InterpArgs<T, TO> ia{e->boff, e->pbase, (const T *)e->pw, e->perm, (const T *)e->grid,
                           e->comp_stride, e->RS, e->K, N, C, e->P, tl.interp_zc, TO(alpha),
                           TO(beta), out, ldout};
      if (launch_interp(ia, st)) return fail(-1, "interpolate: unsupported configuration");
    }
    mark(SPME_ST_INTERP);
  }
  if (parts & SPME_REAL) {
    if (!both) {
      launch_permute_in<T>(nullptr, nullptr, e->r_iperm, (T *)e->r_fs, in, ldin, N, C, st);
      mark(SPME_ST_PERMUTE_REAL);
    }
    const double xi = e->prm.xi;
    const double self = (parts & SPME_SELF)
                            ? 1.0 - 6.0 * xi / std::sqrt(M_PI) +
                                  40.0 * xi * xi * xi / (3.0 * std::sqrt(M_PI))
                            : 0.0;
    RealArgs<T, TO> ra{e->r_perm, e->r_width, e->r_soff, e->r_col, (const T *)e->r_blk,
                       e->r_pos, e->r_cpos, e->nc, float(e->prm.box_size / e->nc),
                       e->prm.box_size, self, e->poly ? (const T *)e->selfv : nullptr,
                       (const T *)e->r_fs, N, C, alpha, beta,
                       (parts & SPME_RECIP) ? 1 : 0, out, ldout};
    if (launch_real(ra, st)) return fail(-1, "real space: unsupported batch width");
    mark(SPME_ST_REAL);
  }
  HIP_TRY(hipGetLastError());
  if (timed) e->pending.push_back(std::move(rec));
  return 0;
}

int apply(_spme_t *e, int nrhs, double alpha, const void *in, long ldin, double beta,
          void *out, long ldout, int parts, hipStream_t st, int stop_after = 0) {
  // hipGetLastError is sticky per thread: clear what earlier, unrelated
  // calls left so that the check after the launches sees only ours
  (void)hipGetLastError();
  if (!e->have_pos) return fail(-1, "positions have not been set");
  if (nrhs < 1 || nrhs > e->max_rhs)
    return fail(-1, "nrhs must be in [1, %d] (got %d)", e->max_rhs, nrhs);
  if (ldin < 3L * e->N || ldout < 3L * e->N) return fail(-1, "ldin/ldout must be >= 3*np");
  if (!in || !out) return fail(-1, "null vector");
  if (parts == 0 || (parts & ~SPME_FULL)) return fail(-1, "bad parts mask %d", parts);
  if (e->prec == SPME_FP32_SUM64)
    return apply_t<float, double>(e, nrhs, alpha, (const float *)in, ldin, beta, (double *)out,
                                  ldout, parts, st, stop_after);
  if (e->prec == SPME_FP32)
    return apply_t<float, float>(e, nrhs, alpha, (const float *)in, ldin, beta, (float *)out, ldout,
                          parts, st, stop_after);
  return apply_t<double, double>(e, nrhs, alpha, (const double *)in, ldin, beta, (double *)out, ldout,
                         parts, st, stop_after);
}

// ---------------------------------------------------------------------------
// Autotuning: candidate shapes are timed with HIP events on the caller's
// stream.  Every candidate computes the same bits, so the choice is purely
// by time.

struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  ~EventPair() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  int init() {
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    return 0;
  }
};

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// Mean milliseconds per call of launch() (which returns 0 or an error code)
// over reps timed calls after one untimed call.
template <class F>
int time_launch(const EventPair &ev, hipStream_t st, int reps, F &&launch, float &ms) {
  if (launch()) return fail(-1, "autotune: launch rejected");
  HIP_TRY(hipEventRecord(ev.a, st));
  for (int r = 0; r < reps; ++r)
    if (launch()) return fail(-1, "autotune: launch rejected");
  HIP_TRY(hipEventRecord(ev.b, st));
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventSynchronize(ev.b));
  HIP_TRY(hipEventElapsedTime(&ms, ev.a, ev.b));
  ms /= float(reps);
  return 0;
}

// z-chunk candidates: K/d for small d, within [lo, hi].
std::vector<int> zc_candidates(int K, int lo, int hi) {
  std::vector<int> v;
  for (int d : {1, 2, 3, 4, 6, 8}) {
    const int z = (K + d - 1) / d;
    if (z >= lo && z <= hi && std::find(v.begin(), v.end(), z) == v.end()) v.push_back(z);
  }
  return v;
}

template <typename T, typename TO>
int autotune_t(_spme_t *e, int nrhs, int reps, hipStream_t st) {
  const int C = e->Cm * nrhs, K = e->K, N = e->N;  // mesh components
  EventPair ev;
  if (int rc = ev.init()) return rc;
  DevBuf out;
  HIP_TRY(hipMalloc(&out.p, sizeof(TO) * (size_t)C * N));
  HIP_TRY(hipMemsetAsync(e->fs, 0, e->tsz * (size_t)C * N, st));
  spme_tiles_t best = e->tiles;
  float best_ms = 1e30f;
  for (int cc : {3, 6, 12}) {
    if (C % cc) continue;
    for (int zc : zc_candidates(K, 8, SPREAD_MAXZC)) {
      SpreadArgs<T> sa{e->boff, e->pbase, (const T *)e->pw, (const T *)e->fs, C, cc, e->P,
                       (T *)e->grid, e->comp_stride, e->RS, K, zc};
      float ms = 0.f;
      if (int rc = time_launch(ev, st, reps, [&] { return launch_spread<T>(sa, st); }, ms))
        return rc;
      if (ms < best_ms) {
        best_ms = ms;
        best.spread_cc = cc;
        best.spread_zc = zc;
      }
    }
  }
  best_ms = 1e30f;
  for (int zc : zc_candidates(K, 8, std::min(K, (int)InterpShape<T>::MAXZC))) {
    InterpArgs<T, TO> ia{e->boff, e->pbase, (const T *)e->pw, e->perm, (const T *)e->grid,
                         e->comp_stride, e->RS, K, N, C, e->P, zc, TO(1), TO(0), (TO *)out.p,
                         3L * N};
    float ms = 0.f;
    if (int rc = time_launch(ev, st, reps, [&] { return launch_interp(ia, st); }, ms))
      return rc;
    if (ms < best_ms) {
      best_ms = ms;
      best.interp_zc = zc;
    }
  }
  HIP_TRY(hipStreamSynchronize(st));
  e->tiles = best;
  return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// Groups of independent systems.  Stage order of a grouped apply: force
// permutation (both orderings), spread, the five FFT passes, interpolation,
// real space; one launch each.  Per stage, the logical blocks of all
// systems are numbered system after system; maps[stage] holds pre (nsys+1
// first blocks) followed by the system of every block.

enum { GS_PERM = 0, GS_SPREAD, GS_FFT0, GS_INTERP = GS_FFT0 + 5, GS_REAL, GS_N };

struct _spme_group_t {
  std::vector<_spme_t *> sys;
  std::vector<unsigned> ver;  // position version of each engine at the last map build
  int nrhs = 0, prec = 0, P = 0, kmax = 96, txy = 16, ncu = 256;
  size_t tsz = 4;
  spme_tiles_t tiles{};
  // block maps (device)
  int *d_maps = nullptr;
  size_t maps_cap = 0;
  int total[GS_N] = {};
  const int *pre[GS_N] = {};
  const int *smap[GS_N] = {};
  // argument tables: device copy and pinned host staging
  char *d_args = nullptr, *h_args = nullptr;
  size_t args_cap = 0;
  size_t a_off[5] = {};  // perm, spread, fft, interp, real
  hipEvent_t up_ev = nullptr;
  hipStream_t last = nullptr;
  bool used = false, maps_ok = false, args_ok = false;
  // the call the uploaded arguments describe
  std::vector<const void *> k_in;
  std::vector<void *> k_out;
  std::vector<long> k_ldin, k_ldout;
  double k_alpha = 0, k_beta = 0;
  int k_parts = 0;
  // captured apply
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_parts = -1;
};

namespace {

void group_drop_graph(_spme_group_t *g) {
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  g->exec = nullptr;
  g->graph = nullptr;
  g->graph_parts = -1;
}

// Waits for every earlier use of the group's tables.
int group_quiesce(_spme_group_t *g) {
  if (g->used) HIP_TRY(hipStreamSynchronize(g->last));
  if (g->up_ev) HIP_TRY(hipEventSynchronize(g->up_ev));
  return 0;
}

int group_stage_blocks(const _spme_group_t *g, const _spme_t *e, int stage) {
  const int C = 3 * g->nrhs;
  const Tiles tl = resolve_tiles(e, g->tiles, C);
  switch (stage) {
    case GS_PERM: return group_perm_blocks(e->N, C);
    case GS_SPREAD: return spread_blocks(e->K, tl.spread_zc, C, tl.spread_cc);
    case GS_INTERP:
      return interp_blocks(e->K, tl.interp_zc, C,
                           g->tsz == 4 ? InterpShape<float>::CC : InterpShape<double>::CC, g->txy);
    case GS_REAL: return group_real_blocks(e->N, C);
    default:
      return g->tsz == 4 ? fft_pass_blocks<float>(e->K, C, stage - GS_FFT0)
                         : fft_pass_blocks<double>(e->K, C, stage - GS_FFT0);
  }
}

int group_build_maps(_spme_group_t *g) {
  const int n = (int)g->sys.size();
  std::vector<int> h;
  size_t off[GS_N];
  for (int stage = 0; stage < GS_N; ++stage) {
    off[stage] = h.size();
    long acc = 0;
    std::vector<int> cnt(n);
    for (int s = 0; s < n; ++s) {
      cnt[s] = group_stage_blocks(g, g->sys[s], stage);
      h.push_back(int(acc));
      acc += cnt[s];
      if (acc > (1L << 30)) return fail(-1, "group too large: stage %d has %ld blocks", stage, acc);
    }
    h.push_back(int(acc));
    for (int s = 0; s < n; ++s) h.insert(h.end(), cnt[s], s);
    g->total[stage] = int(acc);
  }
  if (int rc = group_quiesce(g)) return rc;
  group_drop_graph(g);
  if (h.size() > g->maps_cap) {
    if (g->d_maps) (void)hipFree(g->d_maps);
    g->d_maps = nullptr;
    g->maps_cap = 0;
    HIP_TRY(hipMalloc(&g->d_maps, sizeof(int) * h.size()));
    g->maps_cap = h.size();
  }
  HIP_TRY(hipMemcpy(g->d_maps, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
  for (int stage = 0; stage < GS_N; ++stage) {
    g->pre[stage] = g->d_maps + off[stage];
    g->smap[stage] = g->d_maps + off[stage] + n + 1;
  }
  for (int s = 0; s < n; ++s) g->ver[s] = g->sys[s]->pos_version;
  g->maps_ok = true;
  g->args_ok = false;
  return 0;
}

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

template <typename T>
int group_build_args(_spme_group_t *g, double alpha, const void *const *in, const long *ldin,
                     double beta, void *const *out, const long *ldout, int parts,
                     hipStream_t st) {
  const int n = (int)g->sys.size(), C = 3 * g->nrhs;
  const size_t sz[5] = {sizeof(GPermArgs<T>), sizeof(SpreadArgs<T>), sizeof(FftArgs<T>),
                        sizeof(InterpArgs<T>), sizeof(RealArgs<T>)};
  size_t bytes = 0;
  for (int k = 0; k < 5; ++k) {
    g->a_off[k] = bytes;
    bytes += align256(sz[k] * n);
  }
  if (int rc = group_quiesce(g)) return rc;
  if (bytes > g->args_cap) {
    group_drop_graph(g);
    if (g->d_args) (void)hipFree(g->d_args);
    if (g->h_args) (void)hipHostFree(g->h_args);
    g->d_args = g->h_args = nullptr;
    g->args_cap = 0;
    HIP_TRY(hipMalloc(&g->d_args, bytes));
    HIP_TRY(hipHostMalloc(&g->h_args, bytes));
    g->args_cap = bytes;
  }
  auto *pa = reinterpret_cast<GPermArgs<T> *>(g->h_args + g->a_off[0]);
  auto *sa = reinterpret_cast<SpreadArgs<T> *>(g->h_args + g->a_off[1]);
  auto *fa = reinterpret_cast<FftArgs<T> *>(g->h_args + g->a_off[2]);
  auto *ia = reinterpret_cast<InterpArgs<T> *>(g->h_args + g->a_off[3]);
  auto *ra = reinterpret_cast<RealArgs<T> *>(g->h_args + g->a_off[4]);
  for (int s = 0; s < n; ++s) {
    _spme_t *e = g->sys[s];
    const Tiles tl = resolve_tiles(e, g->tiles, C);
    const bool rec = parts & SPME_RECIP, real = parts & SPME_REAL;
    pa[s] = GPermArgs<T>{rec ? e->perm : nullptr, real ? e->r_perm : nullptr,
                         (const T *)in[s], ldin[s], (T *)e->fs, (T *)e->r_fs, e->N, C};
    sa[s] = SpreadArgs<T>{e->boff, e->pbase, (const T *)e->pw, (const T *)e->fs, C,
                          tl.spread_cc, e->P, (T *)e->grid, e->comp_stride, e->RS, e->K,
                          tl.spread_zc};
    // grouped applies run the separate passes: the Nyquist column stays in the rows
    fa[s] = FftArgs<T>{(T *)e->grid, e->comp_stride, e->RS, e->K, C, (const T *)e->tab,
                       T(2.0 * M_PI / e->prm.box_size), (const cpx<T> *)e->tw, e->ncu, nullptr};
    ia[s] = InterpArgs<T>{e->boff, e->pbase, (const T *)e->pw, e->perm, (const T *)e->grid,
                          e->comp_stride, e->RS, e->K, e->N, C, e->P, tl.interp_zc, T(alpha),
                          T(beta), (T *)out[s], ldout[s]};
    const double xi = e->prm.xi;
    const double self = (parts & SPME_SELF) ? 1.0 - 6.0 * xi / std::sqrt(M_PI) +
                                                  40.0 * xi * xi * xi / (3.0 * std::sqrt(M_PI))
                                            : 0.0;
    ra[s] = RealArgs<T>{e->r_perm, e->r_width, e->r_soff, e->r_col, (const T *)e->r_blk,
                        e->r_pos, e->r_cpos, e->nc, float(e->prm.box_size / e->nc),
                        e->prm.box_size, self, nullptr, (const T *)e->r_fs, e->N, C,
                        alpha, beta, rec ? 1 : 0, (T *)out[s], ldout[s]};
  }
  HIP_TRY(hipMemcpyAsync(g->d_args, g->h_args, bytes, hipMemcpyHostToDevice, st));
  HIP_TRY(hipEventRecord(g->up_ev, st));
  g->last = st;
  g->used = true;
  g->k_in.assign(in, in + n);
  g->k_out.assign(out, out + n);
  g->k_ldin.assign(ldin, ldin + n);
  g->k_ldout.assign(ldout, ldout + n);
  g->k_alpha = alpha;
  g->k_beta = beta;
  g->k_parts = parts;
  g->args_ok = true;
  return 0;
}

// Hardware grid of a stage for a persistence setting (blocks per CU).
int group_grid(const _spme_group_t *g, int stage, int persist) {
  const int total = g->total[stage];
  if (persist <= 0) return total;
  const long want = ((long)persist * g->ncu + 7) / 8 * 8;
  return want < total ? int(want) : total;
}

template <typename T>
int group_launch(_spme_group_t *g, int stage, hipStream_t st) {
  const spme_tiles_t &tl = g->tiles;
  const char *A = g->d_args;
  if (g->total[stage] == 0) return 0;
  switch (stage) {
    case GS_PERM: {
      GroupStage<GPermArgs<T>> t{(const GPermArgs<T> *)(A + g->a_off[0]), g->pre[stage],
                                 g->smap[stage], g->total[stage]};
      launch_group_permute<T>(t, group_grid(g, stage, tl.real_persist), st);
      return 0;
    }
    case GS_SPREAD: {
      GroupStage<SpreadArgs<T>> t{(const SpreadArgs<T> *)(A + g->a_off[1]), g->pre[stage],
                                  g->smap[stage], g->total[stage]};
      const int cc = resolve_tiles(g->sys[0], tl, 3 * g->nrhs).spread_cc;
      return launch_group_spread<T>(t, g->P, cc, group_grid(g, stage, tl.spread_persist), st);
    }
    case GS_INTERP: {
      GroupStage<InterpArgs<T>> t{(const InterpArgs<T> *)(A + g->a_off[3]), g->pre[stage],
                                  g->smap[stage], g->total[stage]};
      return launch_group_interp<T>(t, g->P, g->txy, group_grid(g, stage, tl.interp_persist),
                                    st);
    }
    case GS_REAL: {
      GroupStage<RealArgs<T>> t{(const RealArgs<T> *)(A + g->a_off[4]), g->pre[stage],
                                g->smap[stage], g->total[stage]};
      return launch_group_real<T>(t, 3 * g->nrhs, group_grid(g, stage, tl.real_persist), st);
    }
    default: {
      GroupStage<FftArgs<T>> t{(const FftArgs<T> *)(A + g->a_off[2]), g->pre[stage],
                               g->smap[stage], g->total[stage]};
      return launch_group_fft<T>(t, stage - GS_FFT0, g->kmax, st);
    }
  }
}

template <typename T>
int group_launch_all(_spme_group_t *g, int parts, hipStream_t st) {
  if (int rc = group_launch<T>(g, GS_PERM, st)) return rc;
  if (parts & SPME_RECIP)
    for (int stage = GS_SPREAD; stage <= GS_INTERP; ++stage)
      if (int rc = group_launch<T>(g, stage, st)) return rc;
  if (parts & SPME_REAL)
    if (int rc = group_launch<T>(g, GS_REAL, st)) return rc;
  return 0;
}

int group_check(_spme_group_t *g, int parts) {
  if (parts == 0 || (parts & ~SPME_FULL)) return fail(-1, "bad parts mask %d", parts);
  if ((parts & SPME_SELF) && !(parts & SPME_REAL))
    return fail(-1, "SPME_SELF without SPME_REAL is not supported");
  for (size_t s = 0; s < g->sys.size(); ++s) {
    _spme_t *e = g->sys[s];
    if (!e->have_pos) return fail(-1, "system %zu: positions have not been set", s);
    if ((parts & SPME_REAL) && e->nc < 3)
      return fail(-1, "system %zu: real-space part needs rc > 0 and box_size >= 3*max(rc, 2)",
                  s);
    if (e->pos_version != g->ver[s]) g->maps_ok = false;
  }
  return 0;
}

template <typename T>
int group_apply_t(_spme_group_t *g, double alpha, const void *const *in, const long *ldin,
                  double beta, void *const *out, const long *ldout, int parts, int graph,
                  hipStream_t st) {
  const int n = (int)g->sys.size();
  for (int s = 0; s < n; ++s) {
    if (!in[s] || !out[s]) return fail(-1, "system %d: null vector", s);
    if (ldin[s] < 3L * g->sys[s]->N || ldout[s] < 3L * g->sys[s]->N)
      return fail(-1, "system %d: ldin/ldout must be >= 3*np", s);
  }
  if (int rc = group_check(g, parts)) return rc;
  if (!g->maps_ok)
    if (int rc = group_build_maps(g)) return rc;
  const bool same = g->args_ok && g->k_alpha == alpha && g->k_beta == beta &&
                    g->k_parts == parts && std::equal(in, in + n, g->k_in.begin()) &&
                    std::equal(out, out + n, g->k_out.begin()) &&
                    std::equal(ldin, ldin + n, g->k_ldin.begin()) &&
                    std::equal(ldout, ldout + n, g->k_ldout.begin());
  if (g->used && g->last != st) HIP_TRY(hipStreamSynchronize(g->last));
  if (!same)
    if (int rc = group_build_args<T>(g, alpha, in, ldin, beta, out, ldout, parts, st))
      return rc;
  g->last = st;
  g->used = true;
  if (!graph) {
    if (int rc = group_launch_all<T>(g, parts, st)) return rc;
    HIP_TRY(hipGetLastError());
    return 0;
  }
  if (!g->exec || g->graph_parts != parts) {
    group_drop_graph(g);
    HIP_TRY(hipStreamBeginCapture(g->cap, hipStreamCaptureModeThreadLocal));
    const int rc = group_launch_all<T>(g, parts, g->cap);
    hipGraph_t gr = nullptr;
    const hipError_t ec = hipStreamEndCapture(g->cap, &gr);
    if (rc || ec != hipSuccess) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc ? rc : fail(-2, "graph capture failed: %s", hipGetErrorString(ec));
    }
    g->graph = gr;
    HIP_TRY(hipGraphInstantiate(&g->exec, gr, nullptr, nullptr, 0));
    g->graph_parts = parts;
  }
  HIP_TRY(hipGraphLaunch(g->exec, st));
  return 0;
}

template <typename T>
int group_autotune_t(_spme_group_t *g, int parts, int reps, hipStream_t st) {
  const int n = (int)g->sys.size(), C = 3 * g->nrhs;
  if (int rc = group_check(g, parts)) return rc;
  // scratch inputs (zero) and outputs, one slice per system
  std::vector<long> ld(n), off(n);
  long tot = 0;
  for (int s = 0; s < n; ++s) {
    ld[s] = 3L * g->sys[s]->N;
    off[s] = tot;
    tot += ld[s] * g->nrhs;
  }
  DevBuf din, dout;
  HIP_TRY(hipMalloc(&din.p, sizeof(T) * tot));
  HIP_TRY(hipMalloc(&dout.p, sizeof(T) * tot));
  HIP_TRY(hipMemsetAsync(din.p, 0, sizeof(T) * tot, st));
  std::vector<const void *> vin(n);
  std::vector<void *> vout(n);
  for (int s = 0; s < n; ++s) {
    vin[s] = (const T *)din.p + off[s];
    vout[s] = (T *)dout.p + off[s];
  }
  EventPair ev;
  if (int rc = ev.init()) return rc;
  g->last = st;
  g->used = true;
  auto run = [&](const spme_tiles_t &cand, int stage, float &ms) -> int {
    g->tiles = cand;
    if (int rc = group_build_maps(g)) return rc;
    if (int rc = group_build_args<T>(g, 1.0, vin.data(), ld.data(), 0.0, vout.data(),
                                     ld.data(), parts, st))
      return rc;
    if (int rc = group_launch<T>(g, GS_PERM, st)) return rc;
    return time_launch(ev, st, reps, [&] { return group_launch<T>(g, stage, st); }, ms);
  };
  spme_tiles_t best = g->tiles;
  float best_ms = 1e30f, ms = 0.f;
  // spreading: components per pass and z chunk (0 = each system's default),
  // then persistence
  for (int cc : {3, 6, 12}) {
    if (C % cc) continue;
    for (int zc : {0, 16, 24, 32, 48, 64, 96}) {
      spme_tiles_t c = best;
      c.spread_cc = cc;
      c.spread_zc = zc;
      c.spread_persist = 0;
      if (int rc = run(c, GS_SPREAD, ms)) return rc;
      if (ms < best_ms) {
        best_ms = ms;
        best = c;
      }
    }
  }
  for (int ps : {1, 2, 4, 8}) {
    spme_tiles_t c = best;
    c.spread_persist = ps;
    if (int rc = run(c, GS_SPREAD, ms)) return rc;
    if (ms < best_ms) {
      best_ms = ms;
      best = c;
    }
  }
  best_ms = 1e30f;
  for (int zc : {0, 16, 32, 48, 64, 96}) {
    for (int ps : {0, 1, 2}) {
      spme_tiles_t c = best;
      c.interp_zc = zc;
      c.interp_persist = ps;
      if (int rc = run(c, GS_INTERP, ms)) return rc;
      if (ms < best_ms) {
        best_ms = ms;
        best = c;
      }
    }
  }
  if (parts & SPME_REAL) {
    best_ms = 1e30f;
    for (int ps : {0, 1, 2, 4, 8}) {
      spme_tiles_t c = best;
      c.real_persist = ps;
      if (int rc = run(c, GS_REAL, ms)) return rc;
      if (ms < best_ms) {
        best_ms = ms;
        best = c;
      }
    }
  }
  HIP_TRY(hipStreamSynchronize(st));
  g->tiles = best;
  g->maps_ok = false;  // rebuilt (and arguments re-uploaded) by the next apply
  g->args_ok = false;
  return 0;
}

void die(const char *where) {
  fprintf(stderr, "%s: %s\n", where, g_err.c_str());
  abort();
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

const char *spme_last_error(void) { return g_err.c_str(); }

int spme_engine_create(const spme_params_t *prm, spme_t **out) {
  if (!out) return fail(-1, "null output pointer");
  g_err.clear();
  return engine_create(prm, out);
}

int spme_engine_destroy(spme_t *spme) {
  if (!spme) return 0;
  (void)hipStreamSynchronize(spme->stream);
  free_all(spme);
  delete spme;
  return 0;
}

int spme_engine_set_positions(spme_t *spme, const double *pos, int on_device, void *stream) {
  if (!spme) return fail(-1, "null engine");
  return set_positions(spme, pos, on_device, (hipStream_t)stream);
}

int spme_engine_set_radii(spme_t *spme, const double *radii, int on_device, void *stream) {
  if (!spme) return fail(-1, "null engine");
  return set_radii(spme, radii, on_device, (hipStream_t)stream);
}

int spme_engine_apply(spme_t *spme, int nrhs, double alpha, const void *in, long ldin,
                      double beta, void *out, long ldout, int parts, void *stream) {
  if (!spme) return fail(-1, "null engine");
  return apply(spme, nrhs, alpha, in, ldin, beta, out, ldout, parts, (hipStream_t)stream);
}

int spme_engine_set_timing(spme_t *e, int enable) {
  if (!e) return fail(-1, "null engine");
  if (int rc = collect_timing(e)) return rc;
  e->timing = enable ? 1 : 0;
  for (int i = 0; i < SPME_NSTAGES; ++i) {
    e->stage_ms[i] = 0.0;
    e->stage_n[i] = 0;
  }
  return 0;
}

int spme_engine_get_timing(spme_t *e, double *ms_total, int *launches) {
  if (!e || !ms_total || !launches) return fail(-1, "null argument");
  if (int rc = collect_timing(e)) return rc;
  for (int i = 0; i < SPME_NSTAGES; ++i) {
    ms_total[i] = e->stage_ms[i];
    launches[i] = e->stage_n[i];
  }
  return 0;
}

int spme_engine_info(const spme_t *e, spme_info_t *info) {
  if (!e || !info) return fail(-1, "null argument");
  info->K = e->K;
  info->P = e->P;
  info->N = e->N;
  info->precision = e->prec;
  info->max_rhs = e->max_rhs;
  info->RS = e->RS;
  info->comp_stride = e->comp_stride;
  info->L = e->prm.box_size;
  info->xi = e->prm.xi;
  info->rc = e->prm.rc;
  info->mesh_bytes = (long)(3 * e->max_rhs) * e->comp_stride * (long)e->tsz;
  info->stream = e->stream;
  info->real_entries = e->have_pos && e->nc >= 3 ? e->r_used : 0;
  return 0;
}

// stage 1: mesh after spreading; stage 2: after FFT, influence, inverse FFT.
int spme_engine_debug_stage(spme_t *e, int stage, int nrhs, const void *in, long ldin,
                            void *host_mesh) {
  if (!e) return fail(-1, "null engine");
  if (stage != 1 && stage != 2) return fail(-1, "stage must be 1 or 2");
  if (e->dense) return fail(-1, "debug stages are not available on the plain path");
  if (e->poly) return fail(-1, "debug stages are not available with unequal radii");
  int rc = apply(e, nrhs, 1.0, in, ldin, 0.0, (void *)in, ldin, SPME_RECIP, e->stream, stage);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(host_mesh, e->grid, (size_t)3 * nrhs * e->comp_stride * e->tsz,
                         hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  return 0;
}

// --- drop-in (include/spme.h) ---------------------------------------------

void cutoff2xi(double rmax, double ewald_eps, double box_size, double *_xi) {
  if (!(rmax > 0.0) || !(ewald_eps > 0.0) || !(ewald_eps < 1.0) || !_xi) {
    fail(-1, "cutoff2xi: need rmax > 0, 0 < ewald_eps < 1");
    die("cutoff2xi");
  }
  (void)box_size;
  // Largest |eigenvalue| of M_real(rmax) (F and F+G) decreases with xi once
  // xi*rmax > ~1; bisect on xi for the value ewald_eps.
  auto tail = [&](double xi) {
    const double r = rmax, r2 = r * r, xr = xi * r;
    const double e = std::exp(-xr * xr) / std::sqrt(M_PI), ec = std::erfc(xr);
    const double x3 = xi * xi * xi, x5 = x3 * xi * xi, x7 = x5 * xi * xi;
    const double F = ec * (0.75 / r + 0.5 / (r2 * r)) +
                     e * (4 * x7 * r2 * r2 + 3 * x3 * r2 - 20 * x5 * r2 - 4.5 * xi +
                          14 * x3 + xi / r2);
    const double G = ec * (0.75 / r - 1.5 / (r2 * r)) +
                     e * (-4 * x7 * r2 * r2 - 3 * x3 * r2 + 16 * x5 * r2 + 1.5 * xi -
                          2 * x3 - 3 * xi / r2);
    return std::fmax(std::fabs(F), std::fabs(F + G));
  };
  double lo = 1.0 / rmax, hi = 50.0 / rmax;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (tail(mid) > ewald_eps) lo = mid;
    else hi = mid;
  }
  *_xi = hi;
}

int spme_engine_set_tiles(spme_t *spme, const spme_tiles_t *tiles) {
  if (!spme) return fail(-1, "null engine");
  if (spme->dense) return fail(-1, "the plain path has no tile shapes");
  const spme_tiles_t t = tiles ? *tiles : spme_tiles_t{};
  if (int rc = check_tiles(t)) return rc;
  spme->tiles = t;
  return 0;
}

int spme_engine_get_tiles(const spme_t *spme, int nrhs, spme_tiles_t *tiles) {
  if (!spme || !tiles) return fail(-1, "null argument");
  if (spme->dense) return fail(-1, "the plain path has no tile shapes");
  if (nrhs < 1 || nrhs > spme->max_rhs)
    return fail(-1, "nrhs must be in [1, %d] (got %d)", spme->max_rhs, nrhs);
  const Tiles r = resolve_tiles(spme, spme->tiles, 3 * nrhs);
  *tiles = spme->tiles;
  tiles->spread_cc = r.spread_cc;
  tiles->spread_zc = r.spread_zc;
  tiles->interp_zc = r.interp_zc;
  return 0;
}

int spme_engine_autotune(spme_t *spme, int nrhs, int reps, spme_tiles_t *best, void *stream) {
  if (!spme) return fail(-1, "null engine");
  if (spme->dense) return fail(-1, "the plain path has no tile shapes");
  (void)hipGetLastError();
  if (!spme->have_pos) return fail(-1, "positions have not been set");
  if (nrhs < 1 || nrhs > spme->max_rhs)
    return fail(-1, "nrhs must be in [1, %d] (got %d)", spme->max_rhs, nrhs);
  if (reps < 1) return fail(-1, "reps must be positive");
  hipStream_t st = (hipStream_t)stream;
  const int rc = spme->prec == SPME_FP32 ? autotune_t<float, float>(spme, nrhs, reps, st)
                 : spme->prec == SPME_FP32_SUM64 ? autotune_t<float, double>(spme, nrhs, reps, st)
                                                 : autotune_t<double, double>(spme, nrhs, reps, st);
  if (rc) return rc;
  return best ? spme_engine_get_tiles(spme, nrhs, best) : 0;
}

int spme_group_create(int nsys, spme_t *const *engines, int nrhs, spme_group_t **out) {
  if (!out || !engines) return fail(-1, "null argument");
  *out = nullptr;
  if (nsys < 1) return fail(-1, "nsys must be positive");
  for (int s = 0; s < nsys; ++s) {
    const spme_t *e = engines[s];
    if (!e) return fail(-1, "engine %d is null", s);
    if (e->dense) return fail(-1, "engine %d: plain-path engines cannot be grouped", s);
    if (e->prec == SPME_FP32_SUM64)
      return fail(-1, "engine %d: groups take SPME_FP32 or SPME_FP64 engines", s);
    if (e->poly) return fail(-1, "engine %d: engines with unequal radii cannot be grouped", s);
    if (e->prec != engines[0]->prec || e->P != engines[0]->P)
      return fail(-1, "engine %d: all engines need one precision and one porder", s);
    if (nrhs < 1 || nrhs > e->max_rhs)
      return fail(-1, "engine %d: nrhs must be in [1, max_rhs = %d]", s, e->max_rhs);
    for (int r = 0; r < s; ++r)
      if (engines[r] == e) return fail(-1, "engine %d appears twice", s);
  }
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  for (int s = 0; s < nsys; ++s)
    if (engines[s]->device != dev) return fail(-1, "engine %d lives on another device", s);
  _spme_group_t *g = new _spme_group_t();
  g->sys.assign(engines, engines + nsys);
  g->ver.assign(nsys, ~0u);
  g->nrhs = nrhs;
  g->prec = engines[0]->prec;
  g->P = engines[0]->P;
  g->tsz = engines[0]->tsz;
  int kmax = 0;
  bool all16 = true;
  for (auto *e : g->sys) {
    kmax = std::max(kmax, e->K);
    all16 = all16 && e->K % 16 == 0;
  }
  g->kmax = group_kmax(kmax);
  g->txy = all16 ? 16 : 8;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipEventCreateWithFlags(&g->up_ev, hipEventDisableTiming) != hipSuccess ||
      hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking) != hipSuccess) {
    spme_group_destroy(g);
    return fail(-2, "group set-up failed");
  }
  g->ncu = ncu > 0 ? ncu : 256;
  *out = g;
  return 0;
}

int spme_group_destroy(spme_group_t *g) {
  if (!g) return 0;
  (void)group_quiesce(g);
  group_drop_graph(g);
  if (g->d_maps) (void)hipFree(g->d_maps);
  if (g->d_args) (void)hipFree(g->d_args);
  if (g->h_args) (void)hipHostFree(g->h_args);
  if (g->up_ev) (void)hipEventDestroy(g->up_ev);
  if (g->cap) (void)hipStreamDestroy(g->cap);
  delete g;
  return 0;
}

int spme_group_set_tiles(spme_group_t *g, const spme_tiles_t *tiles) {
  if (!g) return fail(-1, "null group");
  const spme_tiles_t t = tiles ? *tiles : spme_tiles_t{};
  if (int rc = check_tiles(t)) return rc;
  g->tiles = t;
  g->maps_ok = false;
  g->args_ok = false;
  return 0;
}

int spme_group_get_tiles(const spme_group_t *g, spme_tiles_t *tiles) {
  if (!g || !tiles) return fail(-1, "null argument");
  *tiles = g->tiles;
  return 0;
}

int spme_group_autotune(spme_group_t *g, int parts, int reps, spme_tiles_t *best,
                        void *stream) {
  if (!g) return fail(-1, "null group");
  (void)hipGetLastError();
  if (reps < 1) return fail(-1, "reps must be positive");
  hipStream_t st = (hipStream_t)stream;
  const int rc = g->prec == SPME_FP32 ? group_autotune_t<float>(g, parts, reps, st)
                                      : group_autotune_t<double>(g, parts, reps, st);
  if (rc) return rc;
  if (best) *best = g->tiles;
  return 0;
}

int spme_group_apply(spme_group_t *g, double alpha, const void *const *in, const long *ldin,
                     double beta, void *const *out, const long *ldout, int parts, int graph,
                     void *stream) {
  if (!g) return fail(-1, "null group");
  if (!in || !out || !ldin || !ldout) return fail(-1, "null argument");
  (void)hipGetLastError();
  hipStream_t st = (hipStream_t)stream;
  return g->prec == SPME_FP32
             ? group_apply_t<float>(g, alpha, in, ldin, beta, out, ldout, parts, graph, st)
             : group_apply_t<double>(g, alpha, in, ldin, beta, out, ldout, parts, graph, st);
}

void create_spme_engine(double xi, int dim, int porder, int np, double box_size, int nthreads,
                        spme_t **_spme) {
  (void)nthreads;  // host work here is not threaded; the GPU does the rest
  spme_params_t p{};
  p.xi = xi;
  p.dim = dim;
  p.porder = porder;
  p.np = np;
  p.box_size = box_size;
  p.rc = 0.0;
  p.precision = SPME_FP64;
  p.max_rhs = 1;
  if (!_spme || spme_engine_create(&p, _spme)) die("create_spme_engine");
  if (hipMalloc(&(*_spme)->h_in, sizeof(double) * 3 * np) != hipSuccess ||
      hipMalloc(&(*_spme)->h_out, sizeof(double) * 3 * np) != hipSuccess) {
    fail(-2, "hipMalloc of staging vectors failed");
    die("create_spme_engine");
  }
  (*_spme)->h_rhs = 1;
}

void destroy_spme_engine(spme_t *spme) { spme_engine_destroy(spme); }

void compute_spme(spme_t *spme, double *pos, int nrhs, double alpha, double *vec_in, int ldin,
                  double beta, double *vec_out, int ldout) {
  if (!spme) {
    fail(-1, "null engine");
    die("compute_spme");
  }
  hipStream_t st = spme->stream;
  if (pos && set_positions(spme, pos, 0, st)) die("compute_spme");
  if (!spme->have_pos) {
    fail(-1, "the first call must pass positions");
    die("compute_spme");
  }
  if (nrhs <= 0) return;
  if (ldin < 3 * spme->N || ldout < 3 * spme->N) {
    fail(-1, "ldin/ldout must be >= 3*np");
    die("compute_spme");
  }
  // right-hand sides go through the batched apply (the reference loops
  // them one by one, old-version/spme.c:631-702; FP64 batches give bitwise
  // the single results), in batches whose mesh stays within 4 GiB
  const size_t n3 = 3 * (size_t)spme->N;
  const double per_rhs = 3.0 * spme->comp_stride * spme->tsz;
  const int cap = int(std::max(1.0, std::min(16.0, std::floor(4294967296.0 / per_rhs))));
  const int R = std::min(nrhs, cap);
  if (engine_grow(spme, R)) die("compute_spme");
  if (spme->h_rhs < R) {
    (void)hipFree(spme->h_in);
    (void)hipFree(spme->h_out);
    spme->h_in = spme->h_out = nullptr;
    spme->h_rhs = 0;
    if (hipMalloc(&spme->h_in, sizeof(double) * n3 * R) != hipSuccess ||
        hipMalloc(&spme->h_out, sizeof(double) * n3 * R) != hipSuccess) {
      fail(-2, "hipMalloc of staging vectors failed");
      die("compute_spme");
    }
    spme->h_rhs = R;
  }
  const size_t w = n3 * sizeof(double);
  for (int nn = 0; nn < nrhs; nn += R) {
    const int nb = std::min(R, nrhs - nn);
    const double *vin = vec_in + (size_t)nn * ldin;
    double *vout = vec_out + (size_t)nn * ldout;
    if (hipMemcpy2DAsync(spme->h_in, w, vin, sizeof(double) * ldin, w, nb,
                         hipMemcpyHostToDevice, st) != hipSuccess ||
        (beta != 0.0 && hipMemcpy2DAsync(spme->h_out, w, vout, sizeof(double) * ldout, w, nb,
                                          hipMemcpyHostToDevice, st) != hipSuccess)) {
      fail(-2, "host-to-device copy failed");
      die("compute_spme");
    }
    if (apply(spme, nb, alpha, spme->h_in, (long)n3, beta, spme->h_out, (long)n3, SPME_RECIP,
              st))
      die("compute_spme");
    if (hipMemcpy2DAsync(vout, sizeof(double) * ldout, spme->h_out, w, w, nb,
                         hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      fail(-2, "device-to-host copy failed");
      die("compute_spme");
    }
  }
}

}  // extern "C"
