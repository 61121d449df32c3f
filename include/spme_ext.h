/*
 * spme_ext.h -- extension entry points of the MI355X SPME engine.
 *
 * The reference has no counterpart for these (its only surface is spme.h);
 * they expose what the drop-in keeps internal:
 *   - precision (FP64 or FP32 mesh path) and a batch width of right-hand
 *     sides that share one P build (the reference loops RHS serially,
 *     old-version/spme.c:631-702, re-spreading each);
 *   - device-resident inputs/outputs on a caller's HIP stream (hipStream_t
 *     passed as void*), so no host copies sit in the timed path;
 *   - the real-space, overlap and self parts, giving the full RPY M.f.
 * All functions return 0 on success and a negative code on error; the
 * message is available from spme_last_error().  Plain C types only.
 */
#ifndef SPME_EXT_H_
#define SPME_EXT_H_

#include "spme.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Precision and data layout of the reciprocal path:
 *   SPME_FP64        FP64, compact separable weights, sorted particles,
 *                    batched right-hand sides, influence fused in the FFT;
 *   SPME_FP32        the same in FP32 (tables and setup in FP64);
 *   SPME_FP64_PLAIN  FP64 with the reference's data structures and call
 *                    pattern (old-version/spme.c:158-324,560-726): a dense
 *                    p^3 weight row and int32 mesh offsets per particle,
 *                    spreading by atomic adds into a zeroed mesh, one
 *                    right-hand side at a time, influence and 1/K^3 scale as
 *                    passes of their own.  The yardstick for the others.
 *                    No tiles, groups or debug stages.
 *   SPME_FP32_SUM64  the SPME_FP32 mesh path (spreading, FFTs, interpolation
 *                    and the real-space pair sums in FP32; tables and setup
 *                    in FP64) whose final beta*out + alpha*(recip + real)
 *                    sum is formed in FP64: inputs are float, outputs
 *                    double (the reference keeps that sum in FP64,
 *                    old-version/spme_kernel.c:28-30).  Not for groups. */
enum { SPME_FP64 = 0, SPME_FP32 = 1, SPME_FP64_PLAIN = 2, SPME_FP32_SUM64 = 3 };

/* Parts of M to apply (bit mask). */
enum { SPME_RECIP = 1, SPME_REAL = 2, SPME_SELF = 4, SPME_FULL = 7 };

typedef struct spme_params {
  double xi;       /* Ewald splitting parameter */
  int dim;         /* mesh points per side, K */
  int porder;      /* B-spline order p: 4, 6 or 8 */
  int np;          /* number of particles N, at most 2^23 */
  double box_size; /* cube side L */
  double rc;       /* real-space cutoff (used by SPME_REAL) */
  int precision;   /* SPME_FP64, SPME_FP32, SPME_FP64_PLAIN or SPME_FP32_SUM64 */
  int max_rhs;     /* largest batch of right-hand sides per apply */
  int radii;       /* nonzero: per-particle radii (spme_engine_set_radii, 1 until
                      set); the mesh then carries f and a^2 f, six components per
                      right-hand side, so max_rhs <= 8; not for the plain path,
                      groups or debug stages */
} spme_params_t;

/* Message of the last failed call on this thread ("" if none). */
const char *spme_last_error(void);

int spme_engine_create(const spme_params_t *prm, spme_t **out);
int spme_engine_destroy(spme_t *spme);

/* Build the interpolation (P) data from positions (N x 3 FP64).  pos is a
 * device pointer if on_device != 0, else host memory.  stream is a
 * hipStream_t (NULL: the default stream); the call waits for the build. */
int spme_engine_set_positions(spme_t *spme, const double *pos, int on_device, void *stream);

/* Per-particle sphere radii (N FP64 values, in units of the reference
 * radius 1; device pointer if on_device) for an engine created with
 * radii != 0.  The reference's psi(k) takes a pair's radii aa, ab
 * (old-version/spme.c:43-60) but is only called with 1, 1 (spme.c:141);
 * here psi_ij uses (a_i^2 + a_j^2)/2, the real-space part its Ewald split,
 * the overlap correction the unequal-sphere RPY tensor of Zuk et al. (2014)
 * and the self term 1/a_i.  Positions must be set (again) after radii. */
int spme_engine_set_radii(spme_t *spme, const double *radii, int on_device, void *stream);

/* out[r*ldout + 3i + c] <- beta*out + alpha*(M f_r)[3i + c], r < nrhs,
 * for the parts in `parts`.  in/out are device pointers of the engine's
 * precision (float for SPME_FP32, double for SPME_FP64 and SPME_FP64_PLAIN;
 * for SPME_FP32_SUM64 in is float and out double).  Asynchronous on
 * stream, a hipStream_t (NULL: the default stream). */
int spme_engine_apply(spme_t *spme, int nrhs, double alpha, const void *in, long ldin,
                      double beta, void *out, long ldout, int parts, void *stream);

/* Per-kernel timing.  With timing on, every apply records HIP events on its
 * stream after each kernel launch; get_timing waits for them and returns,
 * per stage, the summed milliseconds and the number of launches since
 * timing was (re)enabled.  Stages, in launch order: */
enum {
  SPME_ST_PERMUTE = 0,   /* forces into sorted order (reciprocal part; with the real
                            part too, both orders in this one launch) */
  SPME_ST_SPREAD,        /* B-spline spreading onto the mesh */
  SPME_ST_FFT_X_FWD,     /* r2c along x */
  SPME_ST_FFT_Y_FWD,     /* complex FFT along y */
  SPME_ST_FFT_Z_INFL,    /* FFT along z, influence, inverse FFT along z */
  SPME_ST_FFT_Y_INV,     /* inverse FFT along y */
  SPME_ST_FFT_X_INV,     /* c2r along x */
  SPME_ST_INTERP,        /* interpolation back to particles */
  SPME_ST_PERMUTE_REAL,  /* forces into cell order (real-space part without the
                            reciprocal part) */
  SPME_ST_REAL,          /* real-space pairs, overlap and self */
  SPME_NSTAGES
};
int spme_engine_set_timing(spme_t *spme, int enable);
int spme_engine_get_timing(spme_t *spme, double *ms_total, int *launches);

/* Tile shapes of the spreading, interpolation and real-space kernels (the
 * reference has no counterpart; its kernels have fixed shapes).  A field of
 * 0 selects the default.  No setting changes the arithmetic: every valid
 * choice gives bitwise the same result, so tuning is safe at any time.
 *   spread_cc   components per spreading pass: 3, 6 or 12; ignored when it
 *               does not divide 3*nrhs
 *   spread_zc   z planes per spreading block, 1..256 (clamped to K)
 *   interp_zc   z planes per interpolation block, 1..256 (clamped to K)
 *   *_persist   grouped launches only (spme_group_*): resident blocks per
 *               compute unit of a persistent launch of that stage, 1..64;
 *               0 launches one block per tile. */
typedef struct spme_tiles {
  int spread_cc;
  int spread_zc;
  int interp_zc;
  int spread_persist;
  int interp_persist;
  int real_persist;
} spme_tiles_t;

/* NULL restores the defaults. */
int spme_engine_set_tiles(spme_t *spme, const spme_tiles_t *tiles);
/* The shapes an apply of nrhs right-hand sides uses, defaults resolved. */
int spme_engine_get_tiles(const spme_t *spme, int nrhs, spme_tiles_t *tiles);
/* Autotuner: times every candidate shape of the spreading and interpolation
 * kernels on the engine's current positions for batches of nrhs (reps timed
 * launches each), keeps the fastest (as set_tiles) and returns it in best
 * (may be NULL).  Synchronous on stream; overwrites the engine's mesh. */
int spme_engine_autotune(spme_t *spme, int nrhs, int reps, spme_tiles_t *best, void *stream);

/* Groups of independent systems (uneven N, box and mesh) applied together:
 * every stage of the apply is one launch over the tiles of all systems, and
 * with graph != 0 the whole apply is one captured hipGraph, replayed on
 * later calls.  The engines must have positions set, one precision and one
 * B-spline order, and max_rhs >= nrhs; they stay owned by the caller and
 * must outlive the group.  For members with K <= 96 a group gives bitwise
 * the result of the separate spme_engine_apply calls; larger members use
 * separate x and y FFT passes where a single apply fuses them, so their
 * results agree to rounding.  It uses the engines' meshes and buffers: do not
 * apply a member engine on another stream while a group apply may run.
 * Changing an engine's positions is allowed (the group notices). */
typedef struct _spme_group_t spme_group_t;
int spme_group_create(int nsys, spme_t *const *engines, int nrhs, spme_group_t **out);
int spme_group_destroy(spme_group_t *group);
/* Shapes for every member (zc values clamp per system); NULL: defaults. */
int spme_group_set_tiles(spme_group_t *group, const spme_tiles_t *tiles);
int spme_group_get_tiles(const spme_group_t *group, spme_tiles_t *tiles);
/* Autotuner for the grouped kernels (as spme_engine_autotune, over the
 * group's systems; the real-space stage is tuned when parts has it). */
int spme_group_autotune(spme_group_t *group, int parts, int reps, spme_tiles_t *best,
                        void *stream);
/* For s < nsys: out[s][r*ldout[s] + 3i + c] <- beta*out + alpha*(M_s f_{s,r})
 * for r < nrhs, in[s] and out[s] device pointers of the engines' precision.
 * Asynchronous on stream (NULL: the default stream). */
int spme_group_apply(spme_group_t *group, double alpha, const void *const *in,
                     const long *ldin, double beta, void *const *out, const long *ldout,
                     int parts, int graph, void *stream);

/* Parameter-selection network (the reference has none; its cutoff2xi is
 * declared only).  A 4 -> 64 -> 64 -> 4 tanh MLP in FP32; w holds
 * W1[64][4], b1[64], W2[64][64], b2[64], W3[4][64], b3[4] (spme_mlp_nparam()
 * floats).  All pointers are device memory; calls are asynchronous on stream
 * (NULL: the default stream).
 *   forward:  y[B][4] = net(x[B][4]); h1, h2 ([B][64] each, may be NULL)
 *             receive the hidden activations for a backward pass.
 *   backward: gradient of (1/B) sum_b sum_o mask[o] (y - t)^2 into
 *             grad[nparam]; work holds B*(4 + 2*64) floats.
 *   adam:     one Adam step (step counts from 1). */
int spme_mlp_nparam(void);
int spme_mlp_forward(const float *w, const float *x, int B, float *y, float *h1, float *h2,
                     void *stream);
int spme_mlp_backward(const float *w, const float *x, const float *y, const float *t,
                      const float *mask, int B, const float *h1, const float *h2,
                      float *work, float *grad, void *stream);
int spme_mlp_adam(float *w, const float *g, float *m1, float *m2, float lr, float beta1,
                  float beta2, float eps, int step, void *stream);

/* Tall-skinny block operations for block Krylov methods over batches of
 * right-hand sides (block Lanczos for M^{1/2} z in Brownian dynamics).  No
 * counterpart in the reference.  Blocks are row-major [s][ld] device arrays
 * of the precision's type (float for SPME_FP32, double otherwise) with
 * ld >= n, 1 <= sy <= 16 rows of Y and 1 <= sx <= 4096 rows of X (X may
 * stack every earlier block of a Krylov basis: one call instead of one per
 * block); small matrices are FP64 row-major device arrays.
 *   gram:   C[a][b] = sum_{i<n} X[a][i] Y[b][i]   (a < sx, b < sy; FP64
 *           accumulation, deterministic; work: spme_block_gram_worksize
 *           doubles of device memory, 0 for sizes out of range)
 *   update: Y[b][i] = beta Y[b][i] + sum_a X[a][i] A[a][b]  (beta == 0 does
 *           not read Y; the spans of X and Y must not overlap)
 * Asynchronous on stream (a hipStream_t, NULL: the default stream). */
long spme_block_gram_worksize(long n, int sx, int sy);
int spme_block_gram(int precision, long n, int sx, int sy, const void *X, long ldx,
                    const void *Y, long ldy, double *C, double *work, long work_len,
                    void *stream);
int spme_block_update(int precision, long n, int sx, int sy, const void *X, long ldx,
                      const double *A, double beta, void *Y, long ldy, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): spme_block_gram without the workspace argument, calling
 * no allocator.  The partial sums live in a static device array of
 * spme_block_gram_noalloc_max() doubles (8 MiB) that each GPU holds until
 * the process ends.  Where spme_block_gram's partial sums fit in it the
 * partition, and so the result, is bitwise the same; larger tall products
 * (sx > 16) use fewer blocks along n, a deterministic partition of their
 * own.  Sizes beyond the block limits above are refused with a message.
 * All calls on one GPU share the array: calls that may overlap on one GPU
 * (other streams, other threads) race. */
long spme_block_gram_noalloc_max(void);
int spme_block_gram_noalloc(int precision, long n, int sx, int sy, const void *X, long ldx,
                            const void *Y, long ldy, double *C, void *stream);

/* No counterpart in the reference: a measurement aid.  Streaming probes give
 * the HBM rates that plain hand-written streaming kernels reach on the
 * current GPU, the practical ceilings beside the published peak.  mode 0
 * reads a (b[k] = sum of what block k read, for k < blocks), 1 fills b with
 * value, 2 copies a to b, 3 scales b in place by value.  bytes per buffer (a
 * positive multiple of 16), device buffers 16-byte aligned and not
 * overlapping; blocks of 256 threads in a grid-stride loop (0: 16 per
 * compute unit); nontemporal != 0 uses non-temporal loads and stores.  One
 * untimed launch, then reps timed ones on stream (NULL: the default stream);
 * *ms = the mean time of one launch.  Synchronous. */
int spme_stream_probe(int mode, int nontemporal, const void *a, void *b, long bytes,
                      float value, int blocks, int reps, void *stream, float *ms);

/* Layout and sizes, for tests and benchmarks. */
typedef struct spme_info {
  int K, P, N, precision, max_rhs;
  int RS;              /* reals per mesh row (padded) */
  long comp_stride;    /* reals per mesh component */
  double L, xi, rc;
  long mesh_bytes;     /* device bytes of the mesh */
  void *stream;        /* the engine's own stream */
  long real_entries;   /* stored real-space pair entries (rows padded to
                          their 64-row slice's width); 0 before positions */
} spme_info_t;
int spme_engine_info(const spme_t *spme, spme_info_t *info);

/* Copy the current mesh (max_rhs*3 components, real view, RS-padded rows)
 * to host memory; for tests of the intermediate stages. */
int spme_engine_debug_stage(spme_t *spme, int stage, int nrhs, const void *in, long ldin,
                            void *host_mesh);

#ifdef __cplusplus
}
#endif

#endif /* SPME_EXT_H_ */
