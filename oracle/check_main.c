/*
 * check_main.c -- drives every function of the CPU oracle (spme_oracle.c)
 * once at small sizes, for the AddressSanitizer / UndefinedBehaviorSanitizer
 * build of the oracle (oracle/Makefile, target sanitize).  Test
 * infrastructure only: it checks memory and arithmetic safety of the
 * checker, not its values (tests/test_oracle.py pins those).
 * Exit status 0, and a sanitizer report with a nonzero status otherwise.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

int orc_tab_splines(int p, double *out);
int orc_tab_W(int p, double *out);
void orc_compute_P(int K, int p, const double *pos, int N, double L, double *P, int *ind,
                   int *base, double *q);
void orc_spread(int K, int p, int N, const double *P, const int *ind, const double *f,
                double *grid);
void orc_interpolate(int K, int p, int N, const double *P, const int *ind, const double *grid,
                     double alpha, double beta, double *u);
void orc_fft1d(int n, double *data, int sign);
void orc_fft3d_r2c(int K, const double *r, double *c);
void orc_fft3d_c2r(int K, double *c, double *r);
void orc_compute_spme(double xi, int K, int p, int N, double L, const double *pos, int nrhs,
                      double alpha, const double *vec_in, int ldin, double beta,
                      double *vec_out, int ldout);
void orc_spme_full(double xi, int K, int p, double rc, int N, double L, const double *pos,
                   int nrhs, double alpha, const double *vec_in, int ldin, double beta,
                   double *vec_out, int ldout);
void orc_real_space(int N, const double *pos, double L, double xi, double rc, const double *f,
                    int nrows, const int *rows, double *u);
void orc_ewald_direct(int N, const double *pos, double L, double xi, double rc, double kmax,
                      const double *f, int nrows, const int *rows, double *u);
void orc_ewald_direct_radii(int N, const double *pos, const double *rad, double L, double xi,
                            double rc, double kmax, const double *f, int nrows,
                            const int *rows, double *u);
void orc_spme_full_radii(double xi, int K, int p, double rc, int N, double L,
                         const double *pos, const double *rad, int nrhs, double alpha,
                         const double *vec_in, int ldin, double beta, double *vec_out,
                         int ldout);

static double urand(unsigned *s)
{
    *s = *s * 1664525u + 1013904223u;
    return (*s >> 8) * (1.0 / 16777216.0);
}

int main(void)
{
    unsigned seed = 12345u;
    const int N = 40, NR = 2;
    double L = 12.0, sum = 0.0, tab[64];
    double *pos = malloc(sizeof(double) * 3 * N), *rad = malloc(sizeof(double) * N);
    double *f = malloc(sizeof(double) * 3 * N * NR), *u = malloc(sizeof(double) * 3 * N * NR);
    int rows[5] = {0, 3, 7, 11, 39}, i, p, K;
    for (i = 0; i < 3 * N; i++) pos[i] = L * urand(&seed);
    for (i = 0; i < N; i++) rad[i] = 0.5 + urand(&seed);
    for (i = 0; i < 3 * N * NR; i++) f[i] = urand(&seed) - 0.5;
    for (p = 4; p <= 8; p += 2) {
        orc_tab_splines(p, tab);
        orc_tab_W(p, tab);
        sum += tab[0];
    }
    {   /* 1-D transforms of smooth and prime-factor lengths */
        int n;
        for (n = 2; n <= 48; n++) {
            double *d = malloc(sizeof(double) * 2 * n);
            for (i = 0; i < 2 * n; i++) d[i] = urand(&seed);
            orc_fft1d(n, d, -1);
            orc_fft1d(n, d, +1);
            sum += d[0];
            free(d);
        }
    }
    for (K = 16; K <= 24; K += 8) {
        size_t K3 = (size_t)K * K * K, Kc = (size_t)K * K * (K / 2 + 1);
        double *r = malloc(sizeof(double) * K3), *c = malloc(sizeof(double) * 2 * Kc);
        for (i = 0; i < (int)K3; i++) r[i] = urand(&seed);
        orc_fft3d_r2c(K, r, c);
        orc_fft3d_c2r(K, c, r);
        sum += r[1];
        for (p = 4; p <= 6; p += 2) {
            int p3 = p * p * p;
            double *P = malloc(sizeof(double) * N * p3), *q = malloc(sizeof(double) * N * 3 * p);
            int *ind = malloc(sizeof(int) * N * p3), *base = malloc(sizeof(int) * 3 * N);
            double *grid = malloc(sizeof(double) * 3 * K3);
            orc_compute_P(K, p, pos, N, L, P, ind, base, q);
            orc_spread(K, p, N, P, ind, f, grid);
            orc_interpolate(K, p, N, P, ind, grid, 1.0, 0.5, u);
            orc_compute_spme(0.6, K, p, N, L, pos, NR, 1.0, f, 3 * N, 0.0, u, 3 * N);
            orc_spme_full(0.6, K, p, 3.9, N, L, pos, NR, 1.0, f, 3 * N, 0.0, u, 3 * N);
            orc_spme_full_radii(0.6, K, p, 3.9, N, L, pos, rad, NR, 1.0, f, 3 * N, 0.0, u,
                                3 * N);
            sum += u[0] + grid[5];
            free(P); free(q); free(ind); free(base); free(grid);
        }
        free(r);
        free(c);
    }
    orc_real_space(N, pos, L, 0.6, 3.9, f, N, NULL, u);   /* cell list (L/rc >= 3) */
    orc_real_space(N, pos, L, 0.6, 5.0, f, 5, rows, u);   /* brute force over images */
    orc_ewald_direct(N, pos, L, 0.5, 6.6 / 0.5, 2 * 0.5 * 6.6, f, 5, rows, u);
    orc_ewald_direct_radii(N, pos, rad, L, 0.5, 6.6 / 0.5, 2 * 0.5 * 6.6, f, 5, rows, u);
    sum += u[0];
    free(pos); free(rad); free(f); free(u);
    if (!isfinite(sum)) {
        fprintf(stderr, "non-finite result\n");
        return 1;
    }
    printf("oracle sanitizer run: ok (%.6g)\n", sum);
    return 0;
}
