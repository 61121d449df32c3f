/*
 * spme.h -- drop-in C interface of the MI355X SPME engine.
 *
 * Same entry points, argument meaning and padding macros as the reference's
 * old-version/spme.h; each declaration cites the line it replaces.  The
 * engine struct is opaque here (the reference exposes MKL/cuFFT members,
 * old-version/spme.h:7-82, but its callers only hold pointers,
 * old-version/main.c:136,215-216).
 *
 * Semantics (old-version/spme.c:560-726):
 *   vec_out[nn*ldout + 3i + c] <- beta*vec_out[..] + alpha*(M_recip f_nn)[3i + c]
 * for nn < nrhs, with M_recip the k != 0 Beenakker wave-space RPY operator
 * (unit spheres, mobility in units of 1/(6 pi eta a)) evaluated by SPME.
 * pos is N x 3 (AoS) in [0, box_size); pos == NULL reuses the previous
 * positions (the first call must pass them).  All pointers are host memory;
 * the call returns after vec_out is written.  With beta == 0, vec_out is not
 * read.  Errors: invalid arguments abort with a message on stderr, as the
 * reference's asserts do; spme_last_error() (spme_ext.h) reports the cause.
 */
#ifndef SPME_H_
#define SPME_H_

#ifdef __cplusplus
extern "C" {
#endif

typedef struct _spme_t spme_t;

/* old-version/spme.h:85-86 (declared there, never defined): choose the
 * Ewald splitting parameter xi so that the real-space RPY tensor of a pair
 * at distance rmax is ewald_eps relative to the self mobility. */
void cutoff2xi(double rmax, double ewald_eps, double box_size, double *_xi);

/* old-version/spme.h:89-91 */
void create_spme_engine(double xi, int dim, int porder, int np, double box_size,
                        int nthreads, spme_t **_spme);

/* old-version/spme.h:93 */
void destroy_spme_engine(spme_t *spme);

/* old-version/spme.h:95-97 */
void compute_spme(spme_t *spme, double *pos, int nrhs, double alpha, double *vec_in,
                  int ldin, double beta, double *vec_out, int ldout);

/* old-version/spme.h:100-101 */
#define PAD_LEN(N, size) (((N + (64 / size) - 1) / (64 / size)) * (64 / size))
#define PAD_FFT_LEN(N, size) ((((N + (64 / size) - 1) / (64 / size)) | 1) * (64 / size))

#ifdef __cplusplus
}
#endif

#endif /* SPME_H_ */
