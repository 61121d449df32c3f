/*
 * spme_oracle.c -- CPU restatement of the reference's reciprocal-space RPY
 * SPME (billy-inn/SPME, old-version/spme.c and old-version/spme_kernel.c),
 * plus the FP64 direct Beenakker-Ewald sum and the real-space RPY pieces the
 * reference does not have.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (spme_amd/) links, loads
 * or calls this file.  It is the checker used by tests/, by the smoke test of
 * the package entry point, and by the CPU-baseline leg of bench.py.
 *
 * Everything is FP64, single-threaded unless OpenMP is enabled at build time
 * (the build recipe in oracle/Makefile turns it on).
 *
 * Layout (differs from the reference's padded ld1/ld2/ld3; the values are
 * the same):
 *   real mesh     r[z*K*K + y*K + x]                 (K^3 doubles / component)
 *   half spectrum c[(z*K + y)*(K/2+1) + x], x<=K/2   (complex, re/im interleaved)
 *   P             P[i*p^3 + z*p^2 + y*p + x]         (as the reference)
 *   ind           ind[i*p^3 + ...] = zz*K*K + yy*K + xx
 *
 * Pins (tests/test_oracle.py): the B-spline tables against their exact
 * rational closed form; the FFT against numpy.fft and a brute-force DFT; the
 * dense-P brute force for spread/interpolate; the direct Ewald sum against
 * itself at two splitting parameters (xi-independence) and the Hasimoto
 * single-sphere mobility; the real-space tensor against Ewald splitting.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* B-spline tables.  spme.c:14-22 (tab_splines) and spme.c:25-40 (tab_W) for
 * p = 4 and 6, given there as fractions; p = 8 is not in the reference and
 * is the same closed form: W[x][y] = coefficient of g^y in
 * M_p(g + 1/2 + p - 1 - x), denominators 48, 3840, 645120.                */

static const double euler4[3] = {1, 4, 1};
static const double euler6[5] = {1, 26, 66, 26, 1};
static const double euler8[7] = {1, 120, 1191, 2416, 1191, 120, 1};

static const double W4[16] = {1, -6, 12, -8, 23, -30, -12, 24,
                              23, 30, -12, -24, 1, 6, 12, 8};
static const double W6[36] = {
    1, -10, 40, -80, 80, -32,        237, -750, 840, -240, -240, 160,
    1682, -1540, -880, 1120, 160, -320, 1682, 1540, -880, -1120, 160, 320,
    237, 750, 840, 240, -240, -160,  1, 10, 40, 80, 80, 32};
static const double W8[64] = {
    1, -14, 84, -280, 560, -672, 448, -128,
    2179, -10094, 19740, -20440, 10640, -672, -2240, 896,
    60657, -137494, 101556, -1400, -35280, 12768, 4032, -2688,
    259723, -182070, -121380, 108360, 24080, -30240, -2240, 4480,
    259723, 182070, -121380, -108360, 24080, 30240, -2240, -4480,
    60657, 137494, 101556, 1400, -35280, -12768, 4032, 2688,
    2179, 10094, 19740, 20440, 10640, 672, -2240, -896,
    1, 14, 84, 280, 560, 672, 448, 128};

/* Euler-spline values M_p(1..p-1); returns 0 on unsupported p. */
int orc_tab_splines(int p, double *out)
{
    const double *e;
    double den;
    int i;
    if (p == 4) { e = euler4; den = 6.0; }
    else if (p == 6) { e = euler6; den = 120.0; }
    else if (p == 8) { e = euler8; den = 5040.0; }
    else return 0;
    for (i = 0; i < p - 1; i++) out[i] = e[i] / den;
    return p - 1;
}

/* Polynomial coefficients W[x*p + y]; returns 0 on unsupported p. */
int orc_tab_W(int p, double *out)
{
    const double *w;
    double den;
    int i;
    if (p == 4) { w = W4; den = 48.0; }
    else if (p == 6) { w = W6; den = 3840.0; }
    else if (p == 8) { w = W8; den = 645120.0; }
    else return 0;
    for (i = 0; i < p * p; i++) out[i] = w[i] / den;
    return p * p;
}

/* ------------------------------------------------------------------------ */
/* spme.c:43-60 scalar_recip: psi(k) for radii aa, ab.  k == 0 divides by
 * zero in the reference; callers never use that value (it is zeroed).      */
double orc_scalar_recip(double k, double xi, double aa, double ab)
{
    double kx = k / xi;
    double kx2 = kx * kx;
    double k2 = k * k;
    return 6.0 * M_PI * exp(-kx2 / 4.0) * (1.0 - k2 * (aa * aa + ab * ab) / 6.0)
           / k2 * (1.0 + kx2 * (1.0 / 4.0 + kx2 / 8.0));
}

/* spme.c:95-107: FFT-ordered wave numbers, index K/2 carries +pi*K/L. */
void orc_compute_map(int K, double L, double *map)
{
    int x = 0, i;
    for (i = 0; i <= K / 2; i++) map[x++] = i * 2.0 * M_PI / L;
    for (i = -(K / 2 - 1); i <= -1; i++) map[x++] = i * 2.0 * M_PI / L;
}

/* spme.c:109-126: Euler exponential-spline factor b(m), complex (re,im). */
void orc_compute_b(int K, int p, double L, const double *map, double *b)
{
    double spl[8];
    int x, i;
    orc_tab_splines(p, spl);
    for (x = 0; x < K; x++) {
        double kx = map[x], dr = 0.0, di = 0.0, phi, tr, ti, m2;
        for (i = 0; i < p - 1; i++) {
            phi = i * kx * L / K;
            dr += spl[i] * cos(phi);
            di += spl[i] * sin(phi);
        }
        phi = L * kx * (p - 1) / K;
        tr = cos(phi);
        ti = sin(phi);
        m2 = dr * dr + di * di;
        b[2 * x + 0] = (tr * dr + ti * di) / m2;
        b[2 * x + 1] = (ti * dr - tr * di) / m2;
    }
}

/* spme.c:128-150: lm2[(z*K+y)*(K/2+1)+x] = psi(|k|)*|b_x b_y b_z|^2.
 * The k = 0 entry is set to 0 here (the reference stores inf and the
 * influence step zeroes the k = 0 mode afterwards, spme_kernel.c:211-219). */
void orc_compute_lm2(double xi, int K, int p, double L, double *lm2)
{
    int Kh = K / 2 + 1;
    double *map = (double *)malloc(sizeof(double) * K);
    double *b = (double *)malloc(sizeof(double) * 2 * K);
    int j;
    orc_compute_map(K, L, map);
    orc_compute_b(K, p, L, map, b);
#pragma omp parallel for schedule(static)
    for (j = 0; j < K * K; j++) {
        int z = j / K, y = j % K, x;
        double kz = map[z], ky = map[y];
        for (x = 0; x <= K / 2; x++) {
            double kx = map[x];
            double kk = sqrt(kx * kx + ky * ky + kz * kz);
            double m2, tr, ti, br, bi;
            if (kk == 0.0) { lm2[(size_t)j * Kh + x] = 0.0; continue; }
            m2 = orc_scalar_recip(kk, xi, 1.0, 1.0);
            tr = b[2 * x] * b[2 * y] - b[2 * x + 1] * b[2 * y + 1];
            ti = b[2 * x] * b[2 * y + 1] + b[2 * x + 1] * b[2 * y];
            br = tr * b[2 * z] - ti * b[2 * z + 1];
            bi = tr * b[2 * z + 1] + ti * b[2 * z];
            lm2[(size_t)j * Kh + x] = m2 * (br * br + bi * bi);
        }
    }
    free(map);
    free(b);
}

/* ------------------------------------------------------------------------ */
/* spme.c:206-297 compute_P, per particle: centred B-spline weights and the
 * wrapped mesh index of every stencil point.  Also returns the base index
 * (ind0 per dimension) and the 1-D weights q_d[x] (separable form) when the
 * optional outputs are non-NULL.                                            */
void orc_compute_P(int K, int p, const double *pos, int N, double L,
                   double *P, int *ind, int *base, double *q)
{
    double W[64];
    int i;
    orc_tab_W(p, W);
#pragma omp parallel for schedule(static)
    for (i = 0; i < N; i++) {
        double c[3], g[3], pw[3][8], qq[3][8];
        int i0[3], d, x, y, z;
        for (d = 0; d < 3; d++) {
            double fl;
            c[d] = pos[3 * i + d] / L * K;
            fl = floor(c[d]);
            i0[d] = (int)(fl - p / 2 + 1);
            i0[d] = ((i0[d] % K) + K) % K;
            g[d] = c[d] - fl - 0.5;
            pw[d][0] = 1.0;
            for (x = 1; x < p; x++) pw[d][x] = g[d] * pw[d][x - 1];
            for (x = 0; x < p; x++) {
                double s = 0.0;
                for (y = 0; y < p; y++) s += W[x * p + y] * pw[d][y];
                qq[d][x] = s;
            }
        }
        if (base) for (d = 0; d < 3; d++) base[3 * i + d] = i0[d];
        if (q) for (d = 0; d < 3; d++) for (x = 0; x < p; x++)
            q[(size_t)i * 3 * p + d * p + x] = qq[d][x];
        if (P || ind) {
            int p2 = p * p, p3 = p2 * p;
            for (z = 0; z < p; z++) {
                int zz = (z + i0[2]) % K;
                for (y = 0; y < p; y++) {
                    int yy = (y + i0[1]) % K;
                    for (x = 0; x < p; x++) {
                        int xx = (x + i0[0]) % K;
                        size_t o = (size_t)i * p3 + z * p2 + y * p + x;
                        if (P) P[o] = qq[0][x] * qq[1][y] * qq[2][z];
                        if (ind) ind[o] = zz * K * K + yy * K + xx;
                    }
                }
            }
        }
    }
}

/* spme_kernel.c:231-312 spread, in the reference's coloured parallel form:
 * each axis is cut into nb blocks [floor(b K/nb), floor((b+1) K/nb)), all at
 * least p wide, every particle belongs to the block of its stencil's first
 * point, and the blocks of one of the 8 parity colours are spread in
 * parallel: a stencil spans at most two consecutive blocks per axis
 * (periodically: block nb-1 is followed by block 0, of the other parity
 * since nb is even), so same-colour blocks never write the same point.  The
 * reference takes nb = ceil(K/p) blocks of side p, which races and drops
 * blocks when nb is odd (spme_kernel.c:268,280) and leaves a short last
 * block when p does not divide K; here nb = 2 floor(K/2p) and the blocks
 * tile the axis exactly.  Each mesh point receives its additions in a fixed
 * order (colour, block, particle index), whatever the thread count.
 * grid: 3 components of K^3, zeroed here.  f: AoS 3N.                      */
void orc_spread(int K, int p, int N, const double *P, const int *ind,
                const double *f, double *grid)
{
    size_t K3 = (size_t)K * K * K;
    int p3 = p * p * p, i, col;
    int nb = 2 * (K / (2 * p)), nb3, b;
    int *cnt, *start, *order, *bo;
    if (nb < 2) nb = 2;
    nb3 = nb * nb * nb;
    bo = (int *)malloc(sizeof(int) * K);  /* block of each mesh index */
    for (b = 0; b < nb; b++) {
        int x;
        for (x = (int)((long)b * K / nb); x < (int)((long)(b + 1) * K / nb); x++) bo[x] = b;
    }
    memset(grid, 0, sizeof(double) * 3 * K3);
    cnt = (int *)calloc((size_t)nb3 + 1, sizeof(int));
    start = (int *)malloc(sizeof(int) * ((size_t)nb3 + 1));
    order = (int *)malloc(sizeof(int) * (size_t)(N > 0 ? N : 1));
    /* block of each particle from its stencil's first point (x, y, z = 0) */
    for (i = 0; i < N; i++) {
        int i0 = ind[(size_t)i * p3], xx = i0 % K, yy = (i0 / K) % K, zz = i0 / (K * K);
        cnt[(bo[zz] * nb + bo[yy]) * nb + bo[xx]]++;
    }
    start[0] = 0;
    for (b = 0; b < nb3; b++) start[b + 1] = start[b] + cnt[b];
    for (b = 0; b < nb3; b++) cnt[b] = start[b];
    for (i = 0; i < N; i++) {
        int i0 = ind[(size_t)i * p3], xx = i0 % K, yy = (i0 / K) % K, zz = i0 / (K * K);
        order[cnt[(bo[zz] * nb + bo[yy]) * nb + bo[xx]]++] = i;
    }
    for (col = 0; col < 8; col++) {
        int h = nb / 2, nbc = h * h * h, t;
#pragma omp parallel for schedule(dynamic, 1)
        for (t = 0; t < nbc; t++) {
            int bx = 2 * (t % h) + (col & 1), by = 2 * ((t / h) % h) + ((col >> 1) & 1);
            int bz = 2 * (t / (h * h)) + ((col >> 2) & 1);
            int blk = (bz * nb + by) * nb + bx, q, j;
            for (q = start[blk]; q < start[blk + 1]; q++) {
                int ii = order[q];
                for (j = 0; j < p3; j++) {
                    size_t o = (size_t)ii * p3 + j;
                    int idx = ind[o];
                    double pv = P[o];
                    grid[0 * K3 + idx] += pv * f[3 * ii + 0];
                    grid[1 * K3 + idx] += pv * f[3 * ii + 1];
                    grid[2 * K3 + idx] += pv * f[3 * ii + 2];
                }
            }
        }
    }
    free(cnt);
    free(start);
    free(order);
    free(bo);
}

/* spme_kernel.c:8-70 interpolate: u = beta*u + alpha*P^T grid. */
void orc_interpolate(int K, int p, int N, const double *P, const int *ind,
                     const double *grid, double alpha, double beta, double *u)
{
    size_t K3 = (size_t)K * K * K;
    int p3 = p * p * p, i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < N; i++) {
        double t0 = 0.0, t1 = 0.0, t2 = 0.0;
        int j;
        for (j = 0; j < p3; j++) {
            size_t o = (size_t)i * p3 + j;
            int idx = ind[o];
            double pv = P[o];
            t0 += grid[0 * K3 + idx] * pv;
            t1 += grid[1 * K3 + idx] * pv;
            t2 += grid[2 * K3 + idx] * pv;
        }
        u[3 * i + 0] = beta * u[3 * i + 0] + alpha * t0;
        u[3 * i + 1] = beta * u[3 * i + 1] + alpha * t1;
        u[3 * i + 2] = beta * u[3 * i + 2] + alpha * t2;
    }
}

/* ------------------------------------------------------------------------ */
/* Project FP64 FFT (the reference calls MKL DFTI, spme.c:370-412, which is
 * absent here).  Recursive mixed radix, decimation in time, any length;
 * prime factors other than 2 and 3 use a direct DFT butterfly.
 * sign = -1: forward e^{-2 pi i jk/n}; sign = +1: backward; unnormalised.  */

static int smallest_factor(int n)
{
    int f;
    if (n % 4 == 0 && n != 4) return 4;
    for (f = 2; f * f <= n; f++) if (n % f == 0) return f;
    return n;
}

/* w[2t], w[2t+1] = cos, sin(2 pi t / n0), t < n0: the twiddles of every
 * sub-transform (length n divides n0: W_n^e = w[e n0/n]).                 */
static double *twiddles(int n0)
{
    double *w = (double *)malloc(sizeof(double) * 2 * (size_t)n0);
    int t;
    for (t = 0; t < n0; t++) {
        w[2 * t] = cos(2.0 * M_PI * t / n0);
        w[2 * t + 1] = sin(2.0 * M_PI * t / n0);
    }
    return w;
}

static void fft_rec(int n, int stride, const double *in, double *out,
                    int sign, double *tmp, const double *w, int n0)
{
    int r, m, q, s, k, ws;
    if (n == 1) { out[0] = in[0]; out[1] = in[1]; return; }
    r = smallest_factor(n);
    m = n / r;
    ws = n0 / n;
    /* sub-transforms of the r decimated sequences into out[q*m .. ] */
    for (q = 0; q < r; q++)
        fft_rec(m, stride * r, in + 2 * q * stride, out + 2 * q * m, sign, tmp, w, n0);
    /* butterflies: X[k + s*m] = sum_q W_n^{q(k+s m)} Y_q[k] */
    for (k = 0; k < m; k++) {
        for (s = 0; s < r; s++) {
            double sr = 0.0, si = 0.0;
            for (q = 0; q < r; q++) {
                long e = (long)q * (k + s * m) % n;
                double wr = w[2 * e * ws], wi = sign * w[2 * e * ws + 1];
                double yr = out[2 * (q * m + k)], yi = out[2 * (q * m + k) + 1];
                sr += wr * yr - wi * yi;
                si += wr * yi + wi * yr;
            }
            tmp[2 * (k + s * m)] = sr;
            tmp[2 * (k + s * m) + 1] = si;
        }
    }
    memcpy(out, tmp, sizeof(double) * 2 * n);
}

/* 1-D transform of n points in place with a twiddle table of n points and
 * scratch a, t (2n doubles each). */
static void fft_line(int n, double *data, int sign, const double *w, double *a, double *t)
{
    fft_rec(n, 1, data, a, sign, t, w, n);
    memcpy(data, a, sizeof(double) * 2 * n);
}

/* 1-D complex FFT of length n, in place on data[2n]. */
void orc_fft1d(int n, double *data, int sign)
{
    double *a = (double *)malloc(sizeof(double) * 2 * n);
    double *t = (double *)malloc(sizeof(double) * 2 * n);
    double *w = twiddles(n);
    fft_line(n, data, sign, w, a, t);
    free(a);
    free(t);
    free(w);
}

/* Along y then z of a half spectrum c[(z*K+y)*Kh + x]; one twiddle table,
 * scratch per thread. */
static void fft_yz(int K, double *c, int sign)
{
    int Kh = K / 2 + 1;
    double *w = twiddles(K);
#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * 6 * K), *a = buf + 2 * K, *tt = buf + 4 * K;
        int t;
#pragma omp for schedule(static)
        for (t = 0; t < K * Kh; t++) {       /* y lines: fixed (z, x) */
            int z = t / Kh, x = t % Kh, y;
            for (y = 0; y < K; y++) {
                size_t o = ((size_t)(z * K + y) * Kh + x) * 2;
                buf[2 * y] = c[o]; buf[2 * y + 1] = c[o + 1];
            }
            fft_line(K, buf, sign, w, a, tt);
            for (y = 0; y < K; y++) {
                size_t o = ((size_t)(z * K + y) * Kh + x) * 2;
                c[o] = buf[2 * y]; c[o + 1] = buf[2 * y + 1];
            }
        }
#pragma omp for schedule(static)
        for (t = 0; t < K * Kh; t++) {       /* z lines: fixed (y, x) */
            int y = t / Kh, x = t % Kh, z;
            for (z = 0; z < K; z++) {
                size_t o = ((size_t)(z * K + y) * Kh + x) * 2;
                buf[2 * z] = c[o]; buf[2 * z + 1] = c[o + 1];
            }
            fft_line(K, buf, sign, w, a, tt);
            for (z = 0; z < K; z++) {
                size_t o = ((size_t)(z * K + y) * Kh + x) * 2;
                c[o] = buf[2 * z]; c[o + 1] = buf[2 * z + 1];
            }
        }
        free(buf);
    }
    free(w);
}

/* 3-D real-to-complex forward transform, unnormalised.  r: K^3 reals;
 * c: K*K*(K/2+1) complex.  Same transform as spme.c:654-656.               */
void orc_fft3d_r2c(int K, const double *r, double *c)
{
    int Kh = K / 2 + 1;
    double *w = twiddles(K);
#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * 6 * K), *a = buf + 2 * K, *tt = buf + 4 * K;
        int t;
#pragma omp for schedule(static)
        for (t = 0; t < K * K; t++) {
            int x;
            for (x = 0; x < K; x++) { buf[2 * x] = r[(size_t)t * K + x]; buf[2 * x + 1] = 0.0; }
            fft_line(K, buf, -1, w, a, tt);
            for (x = 0; x < Kh; x++) {
                c[((size_t)t * Kh + x) * 2] = buf[2 * x];
                c[((size_t)t * Kh + x) * 2 + 1] = buf[2 * x + 1];
            }
        }
        free(buf);
    }
    free(w);
    fft_yz(K, c, -1);
}

/* 3-D complex-to-real backward transform, unnormalised, destroying c.
 * Convention for a half spectrum that is not Hermitian in its x = 0 and
 * x = K/2 planes (the reference's influence makes it so, because index K/2
 * maps to +pi*K/L): inverse along z and y first, then along x with the
 * imaginary parts of the x = 0 and x = K/2 entries dropped.  The result is
 * the real part of the inverse DFT of the stored spectrum.                 */
void orc_fft3d_c2r(int K, double *c, double *r)
{
    int Kh = K / 2 + 1;
    double *w;
    fft_yz(K, c, +1);
    w = twiddles(K);
#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * 6 * K), *a = buf + 2 * K, *tt = buf + 4 * K;
        int t;
#pragma omp for schedule(static)
        for (t = 0; t < K * K; t++) {
            int x;
            for (x = 0; x < Kh; x++) {
                buf[2 * x] = c[((size_t)t * Kh + x) * 2];
                buf[2 * x + 1] = c[((size_t)t * Kh + x) * 2 + 1];
            }
            buf[1] = 0.0;
            buf[2 * (K / 2) + 1] = 0.0;
            for (x = Kh; x < K; x++) {
                buf[2 * x] = buf[2 * (K - x)];
                buf[2 * x + 1] = -buf[2 * (K - x) + 1];
            }
            fft_line(K, buf, +1, w, a, tt);
            for (x = 0; x < K; x++) r[(size_t)t * K + x] = buf[2 * x];
        }
        free(buf);
    }
    free(w);
}

/* spme_kernel.c:114-228 apply_influence on 3 half spectra (re/im pairs):
 * g_c(k) <- sum_d lm2(k) (delta_cd - k_c k_d/|k|^2) g_d(k), then g(0) = 0. */
void orc_apply_influence(int K, double L, const double *lm2, double *c0,
                         double *c1, double *c2)
{
    int Kh = K / 2 + 1, j;
    double *map = (double *)malloc(sizeof(double) * K);
    orc_compute_map(K, L, map);
#pragma omp parallel for schedule(static)
    for (j = 0; j < K * K; j++) {
        int z = j / K, y = j % K, x, h;
        double kz = map[z], ky = map[y], kyz = ky * ky + kz * kz;
        for (x = 0; x <= K / 2; x++) {
            size_t o = (size_t)j * Kh + x;
            double kx = map[x], m2 = lm2[o], kk, B0, B1, B2, B3, B4, B5;
            if (kx * kx + kyz == 0.0) continue;
            kk = 1.0 / (kx * kx + kyz);
            B0 = m2 - m2 * kx * kx * kk;
            B1 = -m2 * kx * ky * kk;
            B2 = -m2 * kx * kz * kk;
            B3 = m2 - m2 * ky * ky * kk;
            B4 = -m2 * ky * kz * kk;
            B5 = m2 - m2 * kz * kz * kk;
            for (h = 0; h < 2; h++) {
                double a = c0[2 * o + h], b = c1[2 * o + h], c = c2[2 * o + h];
                c0[2 * o + h] = B0 * a + B1 * b + B2 * c;
                c1[2 * o + h] = B1 * a + B3 * b + B4 * c;
                c2[2 * o + h] = B2 * a + B4 * b + B5 * c;
            }
        }
    }
    c0[0] = c0[1] = c1[0] = c1[1] = c2[0] = c2[1] = 0.0;
    free(map);
}

/* spme.c:560-726 compute_spme for one engine configuration, all RHS:
 * vec_out[nn*ldout + 3i+c] = beta*vec_out[..] + alpha*(M_recip f_nn)[3i+c].
 * The P matrix is rebuilt from pos on every call (the reference reuses it
 * when pos == NULL; the result is the same).                               */
void orc_compute_spme(double xi, int K, int p, int N, double L,
                      const double *pos, int nrhs, double alpha,
                      const double *vec_in, int ldin, double beta,
                      double *vec_out, int ldout)
{
    size_t K3 = (size_t)K * K * K, Kc = (size_t)K * K * (K / 2 + 1);
    int p3 = p * p * p, nn, c;
    size_t n;
    double *P = (double *)malloc(sizeof(double) * (size_t)N * p3);
    int *ind = (int *)malloc(sizeof(int) * (size_t)N * p3);
    double *grid = (double *)malloc(sizeof(double) * 3 * K3);
    double *spec = (double *)malloc(sizeof(double) * 3 * 2 * Kc);
    double *lm2 = (double *)malloc(sizeof(double) * Kc);
    double a2 = alpha * ((double)K * K * K) / (L * L * L);
    orc_compute_lm2(xi, K, p, L, lm2);
    orc_compute_P(K, p, pos, N, L, P, ind, NULL, NULL);
    for (nn = 0; nn < nrhs; nn++) {
        orc_spread(K, p, N, P, ind, vec_in + (size_t)nn * ldin, grid);
        for (c = 0; c < 3; c++) orc_fft3d_r2c(K, grid + c * K3, spec + c * 2 * Kc);
        orc_apply_influence(K, L, lm2, spec, spec + 2 * Kc, spec + 4 * Kc);
        for (c = 0; c < 3; c++) {
            orc_fft3d_c2r(K, spec + c * 2 * Kc, grid + c * K3);
            for (n = 0; n < K3; n++) grid[c * K3 + n] *= 1.0 / ((double)K * K * K);
        }
        orc_interpolate(K, p, N, P, ind, grid, a2, beta, vec_out + (size_t)nn * ldout);
    }
    free(P); free(ind); free(grid); free(spec); free(lm2);
}

/* ------------------------------------------------------------------------ */
/* Real-space RPY (not in the reference).  Beenakker, J. Chem. Phys. 85, 1581
 * (1986), radius a = 1, mobility in units of 1/(6 pi eta a):
 *   M_real(r) = F(r) I + G(r) rr^T/r^2
 *   F = erfc(xi r)(3/(4r) + 1/(2r^3))
 *       + e^{-xi^2 r^2}/sqrt(pi) (4xi^7 r^4 + 3xi^3 r^2 - 20xi^5 r^2 - 9xi/2
 *                                 + 14xi^3 + xi/r^2)
 *   G = erfc(xi r)(3/(4r) - 3/(2r^3))
 *       + e^{-xi^2 r^2}/sqrt(pi) (-4xi^7 r^4 - 3xi^3 r^2 + 16xi^5 r^2 + 3xi/2
 *                                 - 2xi^3 - 3xi/r^2)
 * Self: 1 - 6xi/sqrt(pi) + 40xi^3/(3 sqrt(pi)).
 * Overlap (Rotne-Prager 1969), r < 2: add M_ov - M_far with
 *   M_ov  = (1 - 9r/32) I + (3r/32) rr^T/r^2
 *   M_far = 3/(4r)(1 + 2/(3r^2)) I + 3/(4r)(1 - 2/r^2) rr^T/r^2.        */

void orc_real_coeffs(double r, double xi, double *F, double *G)
{
    double r2 = r * r, xr = xi * r, e = exp(-xr * xr) / sqrt(M_PI);
    double ec = erfc(xr);
    double x3 = xi * xi * xi, x5 = x3 * xi * xi, x7 = x5 * xi * xi;
    *F = ec * (0.75 / r + 0.5 / (r2 * r))
         + e * (4.0 * x7 * r2 * r2 + 3.0 * x3 * r2 - 20.0 * x5 * r2 - 4.5 * xi
                + 14.0 * x3 + xi / r2);
    *G = ec * (0.75 / r - 1.5 / (r2 * r))
         + e * (-4.0 * x7 * r2 * r2 - 3.0 * x3 * r2 + 16.0 * x5 * r2 + 1.5 * xi
                - 2.0 * x3 - 3.0 * xi / r2);
}

void orc_overlap_coeffs(double r, double *F, double *G)
{
    if (r >= 2.0) { *F = 0.0; *G = 0.0; return; }
    *F = (1.0 - 9.0 * r / 32.0) - 0.75 / r * (1.0 + 2.0 / (3.0 * r * r));
    *G = (3.0 * r / 32.0) - 0.75 / r * (1.0 - 2.0 / (r * r));
}

double orc_self_coeff(double xi)
{
    return 1.0 - 6.0 * xi / sqrt(M_PI) + 40.0 * xi * xi * xi / (3.0 * sqrt(M_PI));
}

static inline void add_block(double F, double G, const double *d, double r2,
                             const double *f, double *u)
{
    double dot = (d[0] * f[0] + d[1] * f[1] + d[2] * f[2]) / r2;
    u[0] += F * f[0] + G * d[0] * dot;
    u[1] += F * f[1] + G * d[1] * dot;
    u[2] += F * f[2] + G * d[2] * dot;
}

/* Real-space sum for rows[0..nrows): u_i = sum over j and periodic images
 * with 0 < |r| < rc of M_real(r) f_j, plus the overlap correction on the
 * minimum-image pair when |r| < 2 (i != j).  Adds into u (3*nrows).
 * rows == NULL means all N particles.  When L/rc >= 3 a cell list is used
 * (only the minimum image can then lie within rc); otherwise every image
 * whose shift can reach rc is visited.                                     */
static void real_pair(const double *d0, double L, double xi, double rc,
                      int self, const double *fj, double *acc)
{
    int lo[3], hi[3], a, b, c, m;
    for (m = 0; m < 3; m++) {
        lo[m] = (int)ceil((-rc - d0[m]) / L);
        hi[m] = (int)floor((rc - d0[m]) / L);
    }
    for (a = lo[0]; a <= hi[0]; a++)
    for (b = lo[1]; b <= hi[1]; b++)
    for (c = lo[2]; c <= hi[2]; c++) {
        double d[3] = {d0[0] + a * L, d0[1] + b * L, d0[2] + c * L};
        double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2], F, G;
        if (r2 == 0.0 || r2 >= rc * rc) continue;
        orc_real_coeffs(sqrt(r2), xi, &F, &G);
        add_block(F, G, d, r2, fj, acc);
    }
    if (!self) {
        double r2 = d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2], F, G;
        if (r2 < 4.0 && r2 > 0.0) {
            orc_overlap_coeffs(sqrt(r2), &F, &G);
            add_block(F, G, d0, r2, fj, acc);
        }
    }
}

static inline void min_image(const double *pi, const double *pj, double L,
                             double *d)
{
    int m;
    for (m = 0; m < 3; m++) {
        double t = pi[m] - pj[m];
        d[m] = t - L * round(t / L);
    }
}

void orc_real_space(int N, const double *pos, double L, double xi, double rc,
                    const double *f, int nrows, const int *rows, double *u)
{
    int nc = (int)floor(L / rc), t;
    if (nc >= 3) {
        int nc3 = nc * nc * nc, j;
        int *head = (int *)malloc(sizeof(int) * nc3);
        int *next = (int *)malloc(sizeof(int) * N);
        int *cell = (int *)malloc(sizeof(int) * N);
        for (j = 0; j < nc3; j++) head[j] = -1;
        for (j = N - 1; j >= 0; j--) {
            int cc[3], m;
            for (m = 0; m < 3; m++) {
                double s = pos[3 * j + m] / L;
                s -= floor(s);
                cc[m] = (int)(s * nc);
                if (cc[m] >= nc) cc[m] = nc - 1;
            }
            cell[j] = (cc[2] * nc + cc[1]) * nc + cc[0];
            next[j] = head[cell[j]];
            head[cell[j]] = j;
        }
#pragma omp parallel for schedule(dynamic, 16)
        for (t = 0; t < nrows; t++) {
            int i = rows ? rows[t] : t, ci = cell[i], jj;
            int cx = ci % nc, cy = (ci / nc) % nc, cz = ci / (nc * nc), a, b, c;
            double acc[3] = {0.0, 0.0, 0.0};
            for (c = -1; c <= 1; c++)
            for (b = -1; b <= 1; b++)
            for (a = -1; a <= 1; a++) {
                int nb = ((((cz + c) % nc + nc) % nc) * nc + (((cy + b) % nc + nc) % nc)) * nc
                         + (((cx + a) % nc + nc) % nc);
                for (jj = head[nb]; jj >= 0; jj = next[jj]) {
                    double d[3], r2, F, G;
                    if (jj == i) continue;
                    min_image(pos + 3 * i, pos + 3 * jj, L, d);
                    r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                    if (r2 == 0.0) continue;
                    if (r2 < rc * rc) {
                        orc_real_coeffs(sqrt(r2), xi, &F, &G);
                        add_block(F, G, d, r2, f + 3 * jj, acc);
                    }
                    if (r2 < 4.0) {
                        orc_overlap_coeffs(sqrt(r2), &F, &G);
                        add_block(F, G, d, r2, f + 3 * jj, acc);
                    }
                }
            }
            u[3 * t + 0] += acc[0];
            u[3 * t + 1] += acc[1];
            u[3 * t + 2] += acc[2];
        }
        free(head); free(next); free(cell);
        return;
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (t = 0; t < nrows; t++) {
        int i = rows ? rows[t] : t, j;
        double acc[3] = {0.0, 0.0, 0.0};
        for (j = 0; j < N; j++) {
            double d0[3];
            min_image(pos + 3 * i, pos + 3 * j, L, d0);
            real_pair(d0, L, xi, rc, j == i, f + 3 * j, acc);
        }
        u[3 * t + 0] += acc[0];
        u[3 * t + 1] += acc[1];
        u[3 * t + 2] += acc[2];
    }
}

/* Exact Ewald wave-space sum (no mesh) for rows[0..nrows):
 *   u_i += (1/V) sum_{k != 0, |k| < kmax} psi(k)(I - kk^T/k^2) Re[e^{ik.r_i} S(k)]
 *   S(k) = sum_j f_j e^{-ik.r_j}.
 * Sum over the half space (nz > 0, or nz = 0 and ny > 0, or nz = ny = 0 and
 * nx > 0), each term doubled.                                              */
void orc_recip_exact(int N, const double *pos, double L, double xi, double kmax,
                     const double *f, int nrows, const int *rows, double *u)
{
    double dk = 2.0 * M_PI / L, V = L * L * L;
    int nmax = (int)floor(kmax / dk), W = 2 * nmax + 1;
    /* per-particle phase tables e^{-i dk n x} for n in [-nmax, nmax] */
    double *ex = (double *)malloc(sizeof(double) * 2 * (size_t)N * W * 3);
    double *ur = (double *)calloc((size_t)3 * nrows, sizeof(double));
    int nz, j;
#pragma omp parallel for schedule(static)
    for (j = 0; j < N; j++) {
        int d, n;
        for (d = 0; d < 3; d++) {
            double th = dk * pos[3 * j + d];
            for (n = -nmax; n <= nmax; n++) {
                size_t o = (((size_t)j * 3 + d) * W + (n + nmax)) * 2;
                ex[o] = cos(n * th);
                ex[o + 1] = -sin(n * th);
            }
        }
    }
#pragma omp parallel
    {
        double *loc = (double *)calloc((size_t)3 * nrows, sizeof(double));
        double *A = (double *)malloc(sizeof(double) * 6 * (size_t)N);
#pragma omp for schedule(dynamic, 1)
        for (nz = 0; nz <= nmax; nz++) {
            int ny;
            for (ny = -nmax; ny <= nmax; ny++) {
                int nx;
                double kz = nz * dk, ky = ny * dk;
                if (ky * ky + kz * kz >= kmax * kmax) continue;
                if (nz == 0 && ny < 0) continue;
                /* A_j = f_j e^{-i(ky y_j + kz z_j)} */
                for (j = 0; j < N; j++) {
                    const double *ey = ex + (((size_t)j * 3 + 1) * W + (ny + nmax)) * 2;
                    const double *ez = ex + (((size_t)j * 3 + 2) * W + (nz + nmax)) * 2;
                    double pr = ey[0] * ez[0] - ey[1] * ez[1];
                    double pi = ey[0] * ez[1] + ey[1] * ez[0];
                    int m;
                    for (m = 0; m < 3; m++) {
                        A[6 * j + 2 * m] = pr * f[3 * j + m];
                        A[6 * j + 2 * m + 1] = pi * f[3 * j + m];
                    }
                }
                for (nx = -nmax; nx <= nmax; nx++) {
                    double kx = nx * dk, k2 = kx * kx + ky * ky + kz * kz, ps;
                    double S[6] = {0, 0, 0, 0, 0, 0}, kv[3] = {kx, ky, kz};
                    int t, m;
                    if (k2 >= kmax * kmax || k2 == 0.0) continue;
                    if (nz == 0 && ny == 0 && nx < 0) continue;
                    for (j = 0; j < N; j++) {
                        const double *e = ex + (((size_t)j * 3 + 0) * W + (nx + nmax)) * 2;
                        for (m = 0; m < 3; m++) {
                            S[2 * m] += A[6 * j + 2 * m] * e[0] - A[6 * j + 2 * m + 1] * e[1];
                            S[2 * m + 1] += A[6 * j + 2 * m] * e[1] + A[6 * j + 2 * m + 1] * e[0];
                        }
                    }
                    ps = 2.0 * orc_scalar_recip(sqrt(k2), xi, 1.0, 1.0) / V;
                    /* project S onto the plane normal to k */
                    {
                        double sr = (S[0] * kx + S[2] * ky + S[4] * kz) / k2;
                        double si = (S[1] * kx + S[3] * ky + S[5] * kz) / k2;
                        for (m = 0; m < 3; m++) {
                            S[2 * m] -= sr * kv[m];
                            S[2 * m + 1] -= si * kv[m];
                        }
                    }
                    for (t = 0; t < nrows; t++) {
                        int i = rows ? rows[t] : t;
                        double th = kx * pos[3 * i] + ky * pos[3 * i + 1] + kz * pos[3 * i + 2];
                        double c = cos(th), s = sin(th);
                        for (m = 0; m < 3; m++)
                            loc[3 * t + m] += ps * (c * S[2 * m] - s * S[2 * m + 1]);
                    }
                }
            }
        }
#pragma omp critical
        {
            int t;
            for (t = 0; t < 3 * nrows; t++) ur[t] += loc[t];
        }
        free(loc);
        free(A);
    }
    for (j = 0; j < 3 * nrows; j++) u[j] += ur[j];
    free(ur);
    free(ex);
}

/* FP64 direct Beenakker-Ewald mobility product for rows[0..nrows):
 * u = self + real (to rc) + overlap + exact wave-space sum (to kmax).
 * Overwrites u (3*nrows).                                                  */
void orc_ewald_direct(int N, const double *pos, double L, double xi, double rc,
                      double kmax, const double *f, int nrows, const int *rows,
                      double *u)
{
    double s = orc_self_coeff(xi);
    int t, m;
    for (t = 0; t < nrows; t++) {
        int i = rows ? rows[t] : t;
        for (m = 0; m < 3; m++) u[3 * t + m] = s * f[3 * i + m];
    }
    orc_real_space(N, pos, L, xi, rc, f, nrows, rows, u);
    orc_recip_exact(N, pos, L, xi, kmax, f, nrows, rows, u);
}

/* Full SPME mobility product restated on the CPU (reciprocal part as the
 * reference computes it, plus real space, self and overlap), all rows:
 * u = beta*u + alpha*M f for nrhs vectors.                                 */
void orc_spme_full(double xi, int K, int p, double rc, int N, double L,
                   const double *pos, int nrhs, double alpha,
                   const double *vec_in, int ldin, double beta,
                   double *vec_out, int ldout)
{
    double *t = (double *)malloc(sizeof(double) * 3 * (size_t)N);
    double s = orc_self_coeff(xi);
    int nn, n;
    orc_compute_spme(xi, K, p, N, L, pos, nrhs, alpha, vec_in, ldin, beta,
                     vec_out, ldout);
    for (nn = 0; nn < nrhs; nn++) {
        const double *f = vec_in + (size_t)nn * ldin;
        double *u = vec_out + (size_t)nn * ldout;
        for (n = 0; n < 3 * N; n++) t[n] = s * f[n];
        orc_real_space(N, pos, L, xi, rc, f, N, NULL, t);
        for (n = 0; n < 3 * N; n++) u[n] += alpha * t[n];
    }
    free(t);
}

/* ------------------------------------------------------------------------ */
/* Unequal radii (not built by the reference: its psi takes the pair's radii
 * aa, ab, spme.c:43-60, but is only called with 1, 1, spme.c:141).  Units:
 * 1/(6 pi eta) with lengths in units of the reference radius 1.
 *
 * The RPY far field of spheres a_i, a_j is linear in s2 = (a_i^2 + a_j^2)/2:
 *   M_far = 3/(4r) (I + rr^T/r^2) + s2/(2r^3) (I - 3 rr^T/r^2),
 * so every Ewald part splits into an Oseen part and an s2 part.  Real space:
 * Beenakker's terms (orc_real_coeffs, s2 = 1) split by dimension,
 *   F = F_J + s2 F_s,  G = G_J + s2 G_s,
 *   F_J = erfc(xi r) 3/(4r) + e (3 xi^3 r^2 - 9 xi/2)
 *   F_s = erfc(xi r)/(2r^3) + e (4 xi^7 r^4 - 20 xi^5 r^2 + 14 xi^3 + xi/r^2)
 *   G_J = erfc(xi r) 3/(4r) + e (-3 xi^3 r^2 + 3 xi/2)
 *   G_s = -erfc(xi r) 3/(2r^3) + e (-4 xi^7 r^4 + 16 xi^5 r^2 - 2 xi^3 - 3 xi/r^2)
 * with e = exp(-xi^2 r^2)/sqrt(pi).  Wave space: psi_ij(k) = A(k)(1 - k^2
 * (a_i^2 + a_j^2)/6) with A = scalar_recip(k, xi, 0, 0), so
 *   u_i = I[A P (S_f - k^2/6 S_{a^2 f})](r_i) - a_i^2 I[(k^2/6) A P S_f](r_i)
 * (S the structure factor, P = I - kk^T/k^2).  Self: 1/a_i - 6 xi/sqrt(pi)
 * + 40 xi^3 a_i^2/(3 sqrt(pi)).  Overlap (Zuk, Wajnryb, Mizerski and
 * Szymczak, J. Fluid Mech. 741 (2014) R5), r < a_i + a_j: M_ov - M_far with
 *   |a_i - a_j| < r:  M_ov = [16 r^3 (a_i + a_j) - ((a_i - a_j)^2 + 3r^2)^2]
 *                           / (32 r^3 a_i a_j) I
 *                          + 3((a_i - a_j)^2 - r^2)^2 / (32 r^3 a_i a_j) rr^T/r^2
 *   r <= |a_i - a_j|: M_ov = I / max(a_i, a_j).                             */

void orc_real_coeffs_split(double r, double xi, double *FJ, double *GJ, double *Fs,
                           double *Gs)
{
    double r2 = r * r, xr = xi * r, e = exp(-xr * xr) / sqrt(M_PI);
    double ec = erfc(xr);
    double x3 = xi * xi * xi, x5 = x3 * xi * xi, x7 = x5 * xi * xi;
    *FJ = ec * (0.75 / r) + e * (3.0 * x3 * r2 - 4.5 * xi);
    *Fs = ec * (0.5 / (r2 * r)) + e * (4.0 * x7 * r2 * r2 - 20.0 * x5 * r2 + 14.0 * x3 + xi / r2);
    *GJ = ec * (0.75 / r) + e * (-3.0 * x3 * r2 + 1.5 * xi);
    *Gs = -ec * (1.5 / (r2 * r)) + e * (-4.0 * x7 * r2 * r2 + 16.0 * x5 * r2 - 2.0 * x3
                                        - 3.0 * xi / r2);
}

void orc_overlap_coeffs_radii(double r, double ai, double aj, double *F, double *G)
{
    double s2 = 0.5 * (ai * ai + aj * aj), d = ai - aj, r3 = r * r * r;
    double Fov, Gov;
    if (r >= ai + aj) { *F = 0.0; *G = 0.0; return; }
    if (r <= fabs(d)) {
        Fov = 1.0 / (ai > aj ? ai : aj);
        Gov = 0.0;
    } else {
        double q = d * d + 3.0 * r * r, w = d * d - r * r;
        Fov = (16.0 * r3 * (ai + aj) - q * q) / (32.0 * r3 * ai * aj);
        Gov = 3.0 * w * w / (32.0 * r3 * ai * aj);
    }
    *F = Fov - (0.75 / r + s2 / (2.0 * r3));
    *G = Gov - (0.75 / r - 1.5 * s2 / r3);
}

double orc_self_coeff_radius(double xi, double a)
{
    return 1.0 / a - 6.0 * xi / sqrt(M_PI) + 40.0 * xi * xi * xi * a * a / (3.0 * sqrt(M_PI));
}

/* Real-space sum with radii for rows[0..nrows), added into u: pairs within
 * rc (every periodic image) and the overlap correction on the minimum image
 * (r < a_i + a_j).  All pairs are visited (brute force, O(nrows N)).       */
void orc_real_space_radii(int N, const double *pos, const double *rad, double L, double xi,
                          double rc, const double *f, int nrows, const int *rows, double *u)
{
    int t;
#pragma omp parallel for schedule(dynamic, 4)
    for (t = 0; t < nrows; t++) {
        int i = rows ? rows[t] : t, j, m;
        double acc[3] = {0.0, 0.0, 0.0};
        for (j = 0; j < N; j++) {
            double d0[3], s2 = 0.5 * (rad[i] * rad[i] + rad[j] * rad[j]);
            int lo[3], hi[3], a, b, c;
            min_image(pos + 3 * i, pos + 3 * j, L, d0);
            for (m = 0; m < 3; m++) {
                lo[m] = (int)ceil((-rc - d0[m]) / L);
                hi[m] = (int)floor((rc - d0[m]) / L);
            }
            for (a = lo[0]; a <= hi[0]; a++)
            for (b = lo[1]; b <= hi[1]; b++)
            for (c = lo[2]; c <= hi[2]; c++) {
                double d[3] = {d0[0] + a * L, d0[1] + b * L, d0[2] + c * L};
                double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2], FJ, GJ, Fs, Gs;
                if (r2 == 0.0 || r2 >= rc * rc) continue;
                orc_real_coeffs_split(sqrt(r2), xi, &FJ, &GJ, &Fs, &Gs);
                add_block(FJ + s2 * Fs, GJ + s2 * Gs, d, r2, f + 3 * j, acc);
            }
            if (j != i) {
                double r2 = d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2], F, G;
                double sa = rad[i] + rad[j];
                if (r2 > 0.0 && r2 < sa * sa) {
                    orc_overlap_coeffs_radii(sqrt(r2), rad[i], rad[j], &F, &G);
                    add_block(F, G, d0, r2, f + 3 * j, acc);
                }
            }
        }
        u[3 * t + 0] += acc[0];
        u[3 * t + 1] += acc[1];
        u[3 * t + 2] += acc[2];
    }
}

/* Exact wave-space sum with radii (see orc_recip_exact): the structure
 * factors of f and of a^2 f, u_i += (1/V) sum_k A(k) P Re[e^{ik.r_i}
 * (S_f - k^2/6 (a_i^2 S_f + S_{a^2 f}))], over the half space, doubled.    */
void orc_recip_exact_radii(int N, const double *pos, const double *rad, double L, double xi,
                           double kmax, const double *f, int nrows, const int *rows,
                           double *u)
{
    double dk = 2.0 * M_PI / L, V = L * L * L;
    int nmax = (int)floor(kmax / dk), nz;
    double *ur = (double *)calloc((size_t)3 * nrows, sizeof(double));
#pragma omp parallel
    {
        double *loc = (double *)calloc((size_t)3 * nrows, sizeof(double));
#pragma omp for schedule(dynamic, 1)
        for (nz = 0; nz <= nmax; nz++) {
            int ny, nx, j, t, m;
            for (ny = -nmax; ny <= nmax; ny++)
            for (nx = -nmax; nx <= nmax; nx++) {
                double kx = nx * dk, ky = ny * dk, kz = nz * dk;
                double k2 = kx * kx + ky * ky + kz * kz, kv[3] = {kx, ky, kz}, A;
                double S[6] = {0, 0, 0, 0, 0, 0}, Sa[6] = {0, 0, 0, 0, 0, 0};
                if (k2 >= kmax * kmax || k2 == 0.0) continue;
                if (nz == 0 && (ny < 0 || (ny == 0 && nx < 0))) continue;
                for (j = 0; j < N; j++) {
                    double th = kx * pos[3 * j] + ky * pos[3 * j + 1] + kz * pos[3 * j + 2];
                    double c = cos(th), sn = -sin(th), a2 = rad[j] * rad[j];
                    for (m = 0; m < 3; m++) {
                        S[2 * m] += c * f[3 * j + m];
                        S[2 * m + 1] += sn * f[3 * j + m];
                        Sa[2 * m] += a2 * c * f[3 * j + m];
                        Sa[2 * m + 1] += a2 * sn * f[3 * j + m];
                    }
                }
                /* project both onto the plane normal to k */
                {
                    double sr = (S[0] * kx + S[2] * ky + S[4] * kz) / k2;
                    double si = (S[1] * kx + S[3] * ky + S[5] * kz) / k2;
                    double ar = (Sa[0] * kx + Sa[2] * ky + Sa[4] * kz) / k2;
                    double ai = (Sa[1] * kx + Sa[3] * ky + Sa[5] * kz) / k2;
                    for (m = 0; m < 3; m++) {
                        S[2 * m] -= sr * kv[m];
                        S[2 * m + 1] -= si * kv[m];
                        Sa[2 * m] -= ar * kv[m];
                        Sa[2 * m + 1] -= ai * kv[m];
                    }
                }
                A = 2.0 * orc_scalar_recip(sqrt(k2), xi, 0.0, 0.0) / V;
                for (t = 0; t < nrows; t++) {
                    int i = rows ? rows[t] : t;
                    double th = kx * pos[3 * i] + ky * pos[3 * i + 1] + kz * pos[3 * i + 2];
                    double c = cos(th), sn = sin(th), w = 1.0 - k2 * rad[i] * rad[i] / 6.0;
                    for (m = 0; m < 3; m++) {
                        double gr = w * S[2 * m] - k2 / 6.0 * Sa[2 * m];
                        double gi = w * S[2 * m + 1] - k2 / 6.0 * Sa[2 * m + 1];
                        loc[3 * t + m] += A * (c * gr - sn * gi);
                    }
                }
            }
        }
#pragma omp critical
        {
            int t;
            for (t = 0; t < 3 * nrows; t++) ur[t] += loc[t];
        }
        free(loc);
    }
    {
        int t;
        for (t = 0; t < 3 * nrows; t++) u[t] += ur[t];
    }
    free(ur);
}

/* FP64 direct Ewald mobility product with radii (overwrites u). */
void orc_ewald_direct_radii(int N, const double *pos, const double *rad, double L, double xi,
                            double rc, double kmax, const double *f, int nrows,
                            const int *rows, double *u)
{
    int t, m;
    for (t = 0; t < nrows; t++) {
        int i = rows ? rows[t] : t;
        double s = orc_self_coeff_radius(xi, rad[i]);
        for (m = 0; m < 3; m++) u[3 * t + m] = s * f[3 * i + m];
    }
    orc_real_space_radii(N, pos, rad, L, xi, rc, f, nrows, rows, u);
    orc_recip_exact_radii(N, pos, rad, L, xi, kmax, f, nrows, rows, u);
}

/* Full SPME mobility product with radii: the reference's mesh steps (P,
 * spread, FFT, influence, inverse FFT, interpolation) applied to f and to
 * a^2 f, the influence of the pair (A |b|^2 from scalar_recip with zero
 * radii), u = beta u + alpha (I[G1] - a_i^2 I[G2] + self + real).           */
void orc_spme_full_radii(double xi, int K, int p, double rc, int N, double L,
                         const double *pos, const double *rad, int nrhs, double alpha,
                         const double *vec_in, int ldin, double beta, double *vec_out,
                         int ldout)
{
    size_t K3 = (size_t)K * K * K, Kc = (size_t)K * K * (K / 2 + 1);
    int p3 = p * p * p, Kh = K / 2 + 1, nn, c, j;
    size_t n;
    double *P = (double *)malloc(sizeof(double) * (size_t)N * p3);
    int *ind = (int *)malloc(sizeof(int) * (size_t)N * p3);
    double *grid = (double *)malloc(sizeof(double) * 6 * K3);
    double *spec = (double *)malloc(sizeof(double) * 6 * 2 * Kc);
    double *lmA = (double *)malloc(sizeof(double) * Kc);
    double *f2 = (double *)malloc(sizeof(double) * 3 * (size_t)N);
    double *t2 = (double *)malloc(sizeof(double) * 3 * (size_t)N);
    double *map = (double *)malloc(sizeof(double) * K);
    double *b = (double *)malloc(sizeof(double) * 2 * K);
    double a2 = alpha * ((double)K * K * K) / (L * L * L);
    orc_compute_map(K, L, map);
    orc_compute_b(K, p, L, map, b);
    /* A |b_x b_y b_z|^2 (lm2 of compute_spline with zero radii) */
#pragma omp parallel for schedule(static)
    for (j = 0; j < K * K; j++) {
        int z = j / K, y = j % K, x;
        for (x = 0; x <= K / 2; x++) {
            double kx = map[x], ky = map[y], kz = map[z];
            double kk = sqrt(kx * kx + ky * ky + kz * kz), tr, ti, br, bi;
            if (kk == 0.0) { lmA[(size_t)j * Kh + x] = 0.0; continue; }
            tr = b[2 * x] * b[2 * y] - b[2 * x + 1] * b[2 * y + 1];
            ti = b[2 * x] * b[2 * y + 1] + b[2 * x + 1] * b[2 * y];
            br = tr * b[2 * z] - ti * b[2 * z + 1];
            bi = tr * b[2 * z + 1] + ti * b[2 * z];
            lmA[(size_t)j * Kh + x] = orc_scalar_recip(kk, xi, 0.0, 0.0) * (br * br + bi * bi);
        }
    }
    orc_compute_P(K, p, pos, N, L, P, ind, NULL, NULL);
    for (nn = 0; nn < nrhs; nn++) {
        const double *f = vec_in + (size_t)nn * ldin;
        double *u = vec_out + (size_t)nn * ldout;
        for (n = 0; n < (size_t)N; n++)
            for (c = 0; c < 3; c++) f2[3 * n + c] = rad[n] * rad[n] * f[3 * n + c];
        orc_spread(K, p, N, P, ind, f, grid);
        orc_spread(K, p, N, P, ind, f2, grid + 3 * K3);
        for (c = 0; c < 6; c++) orc_fft3d_r2c(K, grid + c * K3, spec + c * 2 * Kc);
        /* G1 = A P (S_f - k^2/6 S_a), G2 = (k^2/6) A P S_f, in place */
#pragma omp parallel for schedule(static)
        for (j = 0; j < K * K; j++) {
            int z = j / K, y = j % K, x, h, m;
            double kz = map[z], ky = map[y];
            for (x = 0; x <= K / 2; x++) {
                size_t o = (size_t)j * Kh + x;
                double kx = map[x], k2 = kx * kx + ky * ky + kz * kz, kv[3] = {kx, ky, kz};
                if (k2 == 0.0) {
                    for (c = 0; c < 6; c++) spec[c * 2 * Kc + 2 * o] = spec[c * 2 * Kc + 2 * o + 1] = 0.0;
                    continue;
                }
                for (h = 0; h < 2; h++) {
                    double g1[3], g2[3], d1 = 0.0, d2 = 0.0;
                    for (m = 0; m < 3; m++) {
                        double sf = spec[m * 2 * Kc + 2 * o + h];
                        double sa = spec[(3 + m) * 2 * Kc + 2 * o + h];
                        g1[m] = sf - k2 / 6.0 * sa;
                        g2[m] = sf;
                        d1 += g1[m] * kv[m];
                        d2 += g2[m] * kv[m];
                    }
                    for (m = 0; m < 3; m++) {
                        spec[m * 2 * Kc + 2 * o + h] = lmA[o] * (g1[m] - d1 * kv[m] / k2);
                        spec[(3 + m) * 2 * Kc + 2 * o + h] =
                            lmA[o] * k2 / 6.0 * (g2[m] - d2 * kv[m] / k2);
                    }
                }
            }
        }
        for (c = 0; c < 6; c++) {
            orc_fft3d_c2r(K, spec + c * 2 * Kc, grid + c * K3);
            for (n = 0; n < K3; n++) grid[c * K3 + n] *= 1.0 / ((double)K * K * K);
        }
        orc_interpolate(K, p, N, P, ind, grid, a2, beta, u);
        orc_interpolate(K, p, N, P, ind, grid + 3 * K3, a2, 0.0, t2);
        for (n = 0; n < (size_t)N; n++)
            for (c = 0; c < 3; c++) u[3 * n + c] -= rad[n] * rad[n] * t2[3 * n + c];
        for (n = 0; n < (size_t)N; n++) {
            double sc = orc_self_coeff_radius(xi, rad[n]);
            for (c = 0; c < 3; c++) t2[3 * n + c] = sc * f[3 * n + c];
        }
        orc_real_space_radii(N, pos, rad, L, xi, rc, f, N, NULL, t2);
        for (n = 0; n < 3 * (size_t)N; n++) u[n] += alpha * t2[n];
    }
    free(P); free(ind); free(grid); free(spec); free(lmA); free(f2); free(t2); free(map);
    free(b);
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
