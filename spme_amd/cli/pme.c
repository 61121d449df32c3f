/*
 * pme -- command-line driver with the reference driver's interface
 * (old-version/main.c): read an xyz configuration, build an SPME engine
 * (xi = 0.5, --dim, --porder), fill forces and velocities as the reference
 * does (srand(1234); f = rand()*3.0, v = rand()*4.0, five vectors of
 * ldm = PAD_LEN(3N, 8) doubles) and apply v <- 2.2 v + 2.1 M_recip f once,
 * through include/spme.h on the MI355X library.
 *
 * Options (old-version/main.c:16-38,166-194):
 *   -h, --help       usage
 *   -v, --version    version
 *   --xyz FILE       particle configuration (default input.xyz)
 *   --dim K          mesh points per side (default 64)
 *   --porder P       B-spline order (default 4)
 * Extensions (no reference counterpart):
 *   --nrhs R         apply to the first R of the five vectors (default 1,
 *                    as the reference's loop `for (i = 0; i < 1; i++)`)
 *   --out PREFIX     write PREFIX.f, PREFIX.v0 and PREFIX.v: the R force
 *                    vectors, the velocities before and after, each R x 3N
 *                    little-endian doubles
 *   --check-xyz      read the file, print N, L and the wrapped positions
 *                    (%.17g), and exit without touching the GPU
 */
#include <getopt.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "spme.h"

static const char *version_info = "0.1.0";

static void usage(const char *call) {
  fprintf(stderr, "Usage: %s [OPTIONS]\n", call);
  fprintf(stderr, "Options:\n");
  fprintf(stderr, "\t-h or --help         Display this information\n");
  fprintf(stderr, "\t-v or --version      Display version information\n");
  fprintf(stderr, "\t--dim                nfft grids\n");
  fprintf(stderr, "\t--porder             interpolation order\n");
  fprintf(stderr, "\t--xyz                xyz file\n");
  fprintf(stderr, "\t--nrhs               right-hand sides to apply (1..5)\n");
  fprintf(stderr, "\t--out                prefix of the binary dumps\n");
  fprintf(stderr, "\t--check-xyz          print the configuration read and exit\n");
}

/* The xyz format of old-version/main.c:45-127: the first line that is not
 * blank and does not start with '#' holds "N L"; every later such line holds
 * "name x y z a" (only x, y, z are used).  Positions are wrapped into
 * [0, L) with fmod.  A count different from N is reported, as the reference
 * does; unlike the reference, rows beyond N are not stored past the array
 * and missing rows are an error. */
static int read_xyz(const char *path, double **pos_out, int *np_out, double *L_out) {
  char line[1024], name[128];
  FILE *fp = fopen(path, "r");
  if (!fp) {
    fprintf(stderr, "Can not open file %s\n", path);
    return -1;
  }
  int np = 0;
  double L = 0.0;
  while (fgets(line, sizeof line, fp)) {
    if (sscanf(line, "%127s", name) != 1 || name[0] == '#') continue;
    if (sscanf(line, "%d %lf", &np, &L) != 2) np = 0;
    break;
  }
  if (np <= 0) {
    fprintf(stderr, "Invalid number of particles\n");
    fclose(fp);
    return -1;
  }
  if (!(L > 0.0)) {
    fprintf(stderr, "Invalid box size\n");
    fclose(fp);
    return -1;
  }
  double *pos = (double *)malloc(sizeof(double) * 3 * (size_t)np);
  if (!pos) {
    fclose(fp);
    return -1;
  }
  int count = 0;
  while (fgets(line, sizeof line, fp)) {
    double x, y, z, a;
    const int n = sscanf(line, "%127s %lf %lf %lf %lf", name, &x, &y, &z, &a);
    if (n < 1 || name[0] == '#') continue;
    if (n < 4) {
      fprintf(stderr, "Malformed line in %s: %s", path, line);
      free(pos);
      fclose(fp);
      return -1;
    }
    if (count < np) {
      pos[3 * count] = x;
      pos[3 * count + 1] = y;
      pos[3 * count + 2] = z;
    }
    ++count;
  }
  fclose(fp);
  if (count != np) {
    fprintf(stderr, "%d particles are found in %s, but num-particles = %d\n", count, path, np);
    if (count < np) {
      free(pos);
      return -1;
    }
  }
  for (int i = 0; i < 3 * np; ++i) {
    const double w = fmod(pos[i], L);
    pos[i] = w >= 0.0 ? w : L + w;
  }
  *pos_out = pos;
  *np_out = np;
  *L_out = L;
  return 0;
}

static int dump(const char *prefix, const char *ext, const double *v, int nrhs, int ldm,
                int n3) {
  char path[1100];
  snprintf(path, sizeof path, "%s.%s", prefix, ext);
  FILE *fp = fopen(path, "wb");
  if (!fp) {
    fprintf(stderr, "Can not write %s\n", path);
    return -1;
  }
  for (int r = 0; r < nrhs; ++r)
    if (fwrite(v + (size_t)r * ldm, sizeof(double), (size_t)n3, fp) != (size_t)n3) {
      fclose(fp);
      return -1;
    }
  return fclose(fp);
}

int main(int argc, char **argv) {
  static const struct option long_options[] = {
      {"help", 0, NULL, 'h'},   {"version", 0, NULL, 'v'}, {"xyz", 1, NULL, 12},
      {"dim", 1, NULL, 21},     {"porder", 1, NULL, 22},   {"nrhs", 1, NULL, 31},
      {"out", 1, NULL, 32},     {"check-xyz", 0, NULL, 33}, {NULL, 0, NULL, 0}};
  char xyz_file[1024] = "input.xyz";
  const char *out = NULL;
  int dim = 64, porder = 4, nrhs = 1, check = 0, c;
  while ((c = getopt_long(argc, argv, ":hv", long_options, NULL)) != -1) {
    switch (c) {
      case 'h': usage(argv[0]); return 0;
      case 'v': fprintf(stdout, "%s version %s\n", argv[0], version_info); return 0;
      case 12:
        strncpy(xyz_file, optarg, sizeof xyz_file - 1);
        break;
      case 21: dim = atoi(optarg); break;
      case 22: porder = atoi(optarg); break;
      case 31: nrhs = atoi(optarg); break;
      case 32: out = optarg; break;
      case 33: check = 1; break;
      case ':': fprintf(stderr, "Option -%c requires an argument.\n", optopt); return -1;
      case '?': fprintf(stderr, "Unknown option `-%c'.\n", optopt); return -1;
      default: usage(argv[0]); return -1;
    }
  }
  if (dim <= 0 || porder <= 0) {
    fprintf(stderr, "dim and porder must be positive\n");
    return -1;
  }
  if (nrhs < 1 || nrhs > 5) {
    fprintf(stderr, "nrhs must be in [1, 5]\n");
    return -1;
  }
  double *pos = NULL, L = 0.0;
  int np = 0;
  if (read_xyz(xyz_file, &pos, &np, &L) != 0) return -1;
  if (check) {
    printf("%d %.17g\n", np, L);
    for (int i = 0; i < np; ++i)
      printf("%.17g %.17g %.17g\n", pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    free(pos);
    return 0;
  }
  printf("imported %d particles, L = %lf\n", np, L);
  const int ldm = PAD_LEN(np * 3, sizeof(double));
  if (4 * porder >= dim) {
    fprintf(stderr, "dim must be larger than 4 * porder\n");
    free(pos);
    return -1;
  }
  const double xi = 0.5;
  printf("SPME: xi = %lf, dim = %d, porder = %d\n", xi, dim, porder);
  spme_t *spme = NULL;
  create_spme_engine(xi, dim, porder, np, L, 1, &spme);
  const double alpha = 2.1, beta = 2.2;
  double *f = (double *)malloc(sizeof(double) * (size_t)ldm * 5);
  double *v = (double *)malloc(sizeof(double) * (size_t)ldm * 5);
  if (!f || !v) return -1;
  srand(1234);
  for (int i = 0; i < ldm * 5; ++i) {  /* old-version/main.c:229-234 */
    f[i] = rand() * 3.0;
    v[i] = rand() * 4.0;
  }
  if (out && (dump(out, "f", f, nrhs, ldm, 3 * np) || dump(out, "v0", v, nrhs, ldm, 3 * np)))
    return -1;
  struct timeval t0, t1;
  gettimeofday(&t0, NULL);
  for (int i = 0; i < nrhs; ++i)  /* old-version/main.c:238-251: pos passed every call */
    compute_spme(spme, pos, 1, alpha, f + (size_t)i * ldm, ldm, beta, v + (size_t)i * ldm, ldm);
  gettimeofday(&t1, NULL);
  printf("compute_spme: %d call(s) in %.3le secs\n", nrhs,
         (t1.tv_sec - t0.tv_sec) + (t1.tv_usec - t0.tv_usec) / 1e6);
  if (out && dump(out, "v", v, nrhs, ldm, 3 * np)) return -1;
  destroy_spme_engine(spme);
  free(pos);
  free(f);
  free(v);
  return 0;
}
