// Hand-written HIP kernels of the SPME reciprocal-space path for gfx950.
//
// Mesh layout (per component, components contiguous, comp_stride reals):
//   real view     r[(z*K + y)*RS + x],  x < K
//   complex view  c[(z*K + y)*Khp + x], x < Kh = K/2+1, Khp = RS/2
// The 3-D r2c / c2r transforms run in place in this layout.
//
// Particles are sorted into buckets keyed by (xy tile of the stencil base,
// z of the stencil base); bucket k holds sorted slots [boff[k], boff[k+1]).
#include <hip/hip_runtime.h>

#include "spme_kernels.h"
#include "spme_launch.h"

namespace spme {

// ---------------------------------------------------------------------------
// P build (old-version/spme.c:232-297, compute_P): per particle, the wrapped
// stencil base index and the separable centred B-spline weights
// q_d[x] = sum_y W[x][y] g_d^y with g = c - floor(c) - 1/2, c = pos/L*K.
// Weights are FP64 here (cast to the mesh precision when gathered).
__global__ __launch_bounds__(256) void k_pbuild(const double *__restrict__ pos, int N,
                                                double L, int K, int P, WTable W,
                                                int TX, int TY, int *__restrict__ base_tmp,
                                                double *__restrict__ w_tmp,
                                                int *__restrict__ key_tmp,
                                                int *__restrict__ counts, int *__restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int b[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double c = pos[3 * i + d] / L * K;
    if (!(c >= -double(K) && c < 2.0 * K)) {  // also catches NaN
      atomicOr(err, 1);
      c = 0.0;
    }
    const double fl = floor(c);
    int i0 = int(fl) - P / 2 + 1;
    i0 %= K;
    if (i0 < 0) i0 += K;
    b[d] = i0;
    const double g = c - fl - 0.5;
    double pw[8];
    pw[0] = 1.0;
    for (int y = 1; y < P; ++y) pw[y] = g * pw[y - 1];
    for (int x = 0; x < P; ++x) {
      double s = 0.0;
      for (int y = 0; y < P; ++y) s += W.w[x * P + y] * pw[y];
      w_tmp[(size_t)i * 3 * P + d * P + x] = s;
    }
  }
  base_tmp[i] = b[0] | (b[1] << 10) | (b[2] << 20);
  const int key = ((b[1] / TY) * (K / TX) + b[0] / TX) * K + b[2];
  key_tmp[i] = key;
  atomicAdd(&counts[key], 1);
}

// Exclusive scan of n ints into out[0..n] (out[n] = total), one block.
__global__ __launch_bounds__(1024) void k_scan(const int *__restrict__ in, int *__restrict__ out,
                                               int n) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int chunk = (n + 1023) / 1024;
  const int a = min(n, t * chunk), e = min(n, a + chunk);
  int s = 0;
  for (int j = a; j < e; ++j) s += in[j];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - s;
  for (int j = a; j < e; ++j) {
    out[j] = run;
    run += in[j];
  }
  if (t == 1023) out[n] = part[1023];
}

__global__ __launch_bounds__(256) void k_pscatter(const int *__restrict__ key,
                                                  const int *__restrict__ boff,
                                                  int *__restrict__ cursor,
                                                  int *__restrict__ perm, int N) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int k = key[i];
  perm[boff[k] + atomicAdd(&cursor[k], 1)] = i;
}

// Order each bucket by original particle index, so the sorted layout (and
// with it every summation order downstream) does not depend on atomics.
__global__ __launch_bounds__(256) void k_bucket_sort(const int *__restrict__ boff,
                                                     int *__restrict__ perm, int nkeys) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nkeys) return;
  const int a = boff[k], e = boff[k + 1];
  for (int j = a + 1; j < e; ++j) {
    const int v = perm[j];
    int m = j - 1;
    while (m >= a && perm[m] > v) {
      perm[m + 1] = perm[m];
      --m;
    }
    perm[m + 1] = v;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_pgather(const int *__restrict__ perm,
                                                 const int *__restrict__ base_tmp,
                                                 const double *__restrict__ w_tmp,
                                                 int *__restrict__ pbase, T *__restrict__ pw,
                                                 int N, int P3) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const int i = perm[s];
  pbase[s] = base_tmp[i];
  for (int j = 0; j < P3; ++j) pw[(size_t)s * P3 + j] = T(w_tmp[(size_t)i * P3 + j]);
}

// fs[s*C + 3r + d] = in[r*ldin + 3*perm[s] + d]  (sorted, particle-major).
template <typename T>
__global__ __launch_bounds__(256) void k_permute_in(const int *__restrict__ perm,
                                                    const T *__restrict__ in, long ldin,
                                                    T *__restrict__ fs, int N, int C) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)N * C) return;
  const int s = int(t / C), c = int(t - (long)s * C);
  const int r = c / 3, d = c - 3 * r;
  fs[t] = in[r * ldin + 3L * perm[s] + d];
}

// ---------------------------------------------------------------------------
// Spread (old-version/spme_kernel.c:231-312): grid[c] = sum_i q_i (x) f_ic.
// One wave owns a TX x TY column of cells over a z range and keeps a window
// of P planes per component in registers; particles are visited in
// increasing base z, and the lowest plane of the window is final (and
// stored) once the sweep has passed every base that reaches it; the window
// then shifts by one plane.  Every cell is written exactly once, so the
// mesh needs no clearing and no atomics are used.
// Latency: the bucket offsets of the whole sweep are read into LDS up front,
// and each step's particles (base, weights, forces) are staged into LDS with
// one batch of independent loads.
template <typename T, int P, int CC, int TX, int TY>
__global__ __launch_bounds__(64) void k_spread(const int *__restrict__ boff,
                                               const int *__restrict__ pbase,
                                               const T *__restrict__ pw,
                                               const T *__restrict__ fs, int C,
                                               T *__restrict__ grid, long comp_stride, int RS,
                                               int K, int ZC) {
  static_assert(TX * TY == 64, "one wave per tile");
  constexpr int NXR = 1 + (P + TX - 2) / TX;
  constexpr int NYR = 1 + (P + TY - 2) / TY;
  constexpr int NTL = NXR * NYR;
  constexpr int MAXST = 256 + P;  // >= ZC + P - 1
  constexpr int MAXB = 16;        // particles staged per batch
  constexpr int NW = 3 * P;
  __shared__ int s_lo[NTL * MAXST];
  __shared__ int s_n[NTL * MAXST];
  __shared__ int s_pb[MAXB];
  __shared__ T s_w[MAXB * NW];
  __shared__ T s_f[MAXB * CC];
  const int ntx = K / TX, nty = K / TY;
  const int lb = xcd_remap(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z),
                           gridDim.x * gridDim.y * gridDim.z);
  const int tile = lb % gridDim.x, rest = lb / gridDim.x;
  const int tx = tile % ntx, ty = tile / ntx;
  const int lane = threadIdx.x;
  const int x = tx * TX + lane % TX, y = ty * TY + lane / TX;
  const int z0 = (rest % gridDim.y) * ZC, zend = min(z0 + ZC, K);
  const int cc0 = (rest / gridDim.y) * CC;
  const int zs = z0 - (P - 1);
  const int nsteps = zend - zs;

  for (int q = lane; q < NTL * nsteps; q += 64) {
    const int t = q / nsteps, st = q - t * nsteps;
    int tyy = ty - t / NXR, txx = tx - t % NXR;
    tyy += tyy < 0 ? nty : 0;
    txx += txx < 0 ? ntx : 0;
    int bz = zs + st;
    bz += bz < 0 ? K : 0;
    const int key = (tyy * ntx + txx) * K + bz;
    const int lo = boff[key];
    s_lo[q] = lo;
    s_n[q] = boff[key + 1] - lo;
  }
  __syncthreads();

  T acc[P][CC];
#pragma unroll
  for (int j = 0; j < P; ++j)
#pragma unroll
    for (int c = 0; c < CC; ++c) acc[j][c] = T(0);

  for (int st = 0; st < nsteps; ++st) {
    int cnt[NTL], lo[NTL];
    int n = 0;
#pragma unroll
    for (int t = 0; t < NTL; ++t) {
      cnt[t] = s_n[t * nsteps + st];
      lo[t] = s_lo[t * nsteps + st];
      n += cnt[t];
    }
    // sorted index of the k-th particle of this step
    auto locate = [&](int k) {
      int i = lo[0] + k;
#pragma unroll
      for (int t = 0; t < NTL - 1; ++t) {
        if (k >= cnt[t]) {
          k -= cnt[t];
          i = lo[t + 1] + k;
        } else {
          break;
        }
      }
      return i;
    };
    for (int b0 = 0; b0 < n; b0 += MAXB) {
      const int nb = min(MAXB, n - b0);
      // stage the batch: all loads first, then the LDS stores
      constexpr int RW = (MAXB * NW + 63) / 64, RF = (MAXB * CC + 63) / 64;
      int vpb = 0;
      T vw[RW], vf[RF];
      if (lane < nb) vpb = pbase[locate(b0 + lane)];
#pragma unroll
      for (int r = 0; r < RW; ++r) {
        const int g = lane + 64 * r;
        vw[r] = g < nb * NW ? pw[(size_t)locate(b0 + g / NW) * NW + g % NW] : T(0);
      }
#pragma unroll
      for (int r = 0; r < RF; ++r) {
        const int g = lane + 64 * r;
        vf[r] = g < nb * CC ? fs[(size_t)locate(b0 + g / CC) * C + cc0 + g % CC] : T(0);
      }
      if (lane < nb) s_pb[lane] = vpb;
#pragma unroll
      for (int r = 0; r < RW; ++r)
        if (lane + 64 * r < MAXB * NW) s_w[lane + 64 * r] = vw[r];
#pragma unroll
      for (int r = 0; r < RF; ++r)
        if (lane + 64 * r < MAXB * CC) s_f[lane + 64 * r] = vf[r];
      __syncthreads();
      for (int k = 0; k < nb; ++k) {
        const int pb = s_pb[k];
        int dx = x - (pb & 1023);
        dx += dx < 0 ? K : 0;
        int dy = y - ((pb >> 10) & 1023);
        dy += dy < 0 ? K : 0;
        if (dx < P && dy < P) {
          const T *w = s_w + k * NW;
          const T wxy = w[dx] * w[P + dy];
          T wf[CC];
#pragma unroll
          for (int c = 0; c < CC; ++c) wf[c] = wxy * s_f[k * CC + c];
#pragma unroll
          for (int j = 0; j < P; ++j) {
            const T wz = w[2 * P + j];
#pragma unroll
            for (int c = 0; c < CC; ++c) acc[j][c] += wz * wf[c];
          }
        }
      }
      __syncthreads();
    }
    if (st >= P - 1) {
      T *g = grid + ((size_t)(zs + st) * K + y) * RS + x + (size_t)cc0 * comp_stride;
#pragma unroll
      for (int c = 0; c < CC; ++c) g[(size_t)c * comp_stride] = acc[0][c];
    }
#pragma unroll
    for (int j = 0; j < P - 1; ++j)
#pragma unroll
      for (int c = 0; c < CC; ++c) acc[j][c] = acc[j + 1][c];
#pragma unroll
    for (int c = 0; c < CC; ++c) acc[P - 1][c] = T(0);
  }
}

// ---------------------------------------------------------------------------
// Interpolation (old-version/spme_kernel.c:8-70):
//   out[r*ldout + 3*i + d] = beta*out + alpha * sum_j q_i(j) grid[3r+d](j).
// A block owns a TXY x TXY tile of stencil bases over a z range and a chunk
// of CC components.  It streams the tile plus its (P-1)-wide halo through
// LDS one z plane at a time (P+1 slots: the P planes the current particles
// need and the one being filled); each particle is handled by exactly one
// wave (lanes = the P x P stencil columns) and written without atomics.
// The loads of step s+1 (its new plane and its particles) are issued before
// step s is processed, so each step costs one barrier and no exposed
// global-memory latency.
template <typename T, int P, int CC, int TXY, int NT>
__global__ __launch_bounds__(NT) void k_interp(const int *__restrict__ boff,
                                                const int *__restrict__ pbase,
                                                const T *__restrict__ pw,
                                                const int *__restrict__ perm,
                                                const T *__restrict__ grid, long comp_stride,
                                                int RS, int K, int C, int ZC, T alpha, T beta,
                                                T *__restrict__ out, long ldout) {
  constexpr int E = TXY + P - 1;
  constexpr int EX = E | 1;  // odd row pitch
  constexpr int PLANE = CC * E * EX;
  constexpr int NSLOT = P + 1;
  constexpr int NSUB = TXY / 8;  // sort tiles are 8 x 8
  constexpr int NS2 = NSUB * NSUB;
  constexpr int MAXST = 64;  // >= ZC
  constexpr int NWV = NT / 64;
  constexpr int MAXB = 32;  // particles staged per batch
  constexpr int NW = 3 * P;
  constexpr int NPL = (CC * E * E + NT - 1) / NT;  // plane values per thread
  constexpr int RW = (MAXB * NW + NT - 1) / NT;    // weights per thread
  __shared__ T win[NSLOT * PLANE];
  __shared__ int s_lo[NS2 * MAXST];
  __shared__ int s_n[NS2 * MAXST];
  __shared__ int s_pb[2][MAXB], s_perm[2][MAXB];
  __shared__ T s_w[2][MAXB * NW];
  const int nt = K / TXY;
  // logical block: tiles fastest, so each XCD gets a band of adjacent tiles
  const int lb = xcd_remap(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z),
                           gridDim.x * gridDim.y * gridDim.z);
  const int tile = lb % (nt * nt), rest = lb / (nt * nt);
  const int tx = tile % nt, ty = tile / nt;
  const int x0 = tx * TXY, y0 = ty * TXY;
  const int z0 = (rest % gridDim.y) * ZC, zend = min(z0 + ZC, K);
  const int nst = zend - z0;
  const int cc0 = (rest / gridDim.y) * CC;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const T *g0 = grid + (size_t)cc0 * comp_stride;

  for (int q = tid; q < NS2 * nst; q += NT) {
    const int sq = q / nst, st = q - sq * nst;
    const int sx = tx * NSUB + sq % NSUB, sy = ty * NSUB + sq / NSUB;
    const int key = (sy * (K / 8) + sx) * K + z0 + st;
    const int lo = boff[key];
    s_lo[q] = lo;
    s_n[q] = boff[key + 1] - lo;
  }
  // per-thread plane offsets, fixed over the sweep: source offset within a
  // mesh z plane (components of the chunk included) and LDS offset within a
  // window slot; -1 past the plane
  long src_off[NPL];
  int dst_off[NPL];
#pragma unroll
  for (int r = 0; r < NPL; ++r) {
    const int q = tid + NT * r;
    const int c = q / (E * E), rr = q - c * (E * E);
    const int ey = rr / E, ex = rr - ey * E;
    int yy = y0 + ey, xx = x0 + ex;
    yy -= yy >= K ? K : 0;
    xx -= xx >= K ? K : 0;
    src_off[r] = (long)c * comp_stride + (long)yy * RS + xx;
    dst_off[r] = q < CC * E * E ? (c * E + ey) * EX + ex : -1;
  }
  const long zstride = (long)K * RS;
  for (int j = 0; j < P - 1; ++j) {
    const T *gz = g0 + ((z0 + j) % K) * zstride;
    T *wz = win + ((z0 + j) % NSLOT) * PLANE;
#pragma unroll
    for (int r = 0; r < NPL; ++r)
      if (dst_off[r] >= 0) wz[dst_off[r]] = gz[src_off[r]];
  }
  __syncthreads();

  // registers carrying the next step's loads
  T vp[NPL], vw[RW];
  int vpb = 0, vpm = 0, nb_next = 0;
  // issue the loads of step st, particles [b0, b0 + MAXB) of it
  auto fetch = [&](int st, int b0, bool with_plane) {
    int cnt[NS2], lo[NS2];
    int n = 0;
#pragma unroll
    for (int q = 0; q < NS2; ++q) {
      cnt[q] = s_n[q * nst + st];
      lo[q] = s_lo[q * nst + st];
      n += cnt[q];
    }
    auto locate = [&](int k) {
      int i = lo[0] + k;
#pragma unroll
      for (int q = 0; q < NS2 - 1; ++q) {
        if (k >= cnt[q]) {
          k -= cnt[q];
          i = lo[q + 1] + k;
        } else {
          break;
        }
      }
      return i;
    };
    const int nb = max(0, min(MAXB, n - b0));
    if (with_plane) {
      const T *gz = g0 + ((z0 + st + P - 1) % K) * zstride;
#pragma unroll
      for (int r = 0; r < NPL; ++r) vp[r] = dst_off[r] >= 0 ? gz[src_off[r]] : T(0);
    }
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const int g = tid + NT * r;
      vw[r] = g < nb * NW ? pw[(size_t)locate(b0 + g / NW) * NW + g % NW] : T(0);
    }
    if (tid < nb) {
      const int i = locate(b0 + tid);
      vpb = pbase[i];
      vpm = perm[i];
    }
    nb_next = nb;
    return n;
  };
  auto commit = [&](int st, int buf, bool with_plane) {
    if (with_plane) {
      T *wz = win + ((z0 + st + P - 1) % NSLOT) * PLANE;
#pragma unroll
      for (int r = 0; r < NPL; ++r)
        if (dst_off[r] >= 0) wz[dst_off[r]] = vp[r];
    }
#pragma unroll
    for (int r = 0; r < RW; ++r)
      if (tid + NT * r < MAXB * NW) s_w[buf][tid + NT * r] = vw[r];
    if (tid < MAXB) {
      s_pb[buf][tid] = vpb;
      s_perm[buf][tid] = vpm;
    }
  };

  const int dx = lane % P, dy = lane / P;
  const bool act = lane < P * P;
  int n_cur = fetch(0, 0, true);
  int b_cur = 0, st = 0, buf = 0;
  while (st < nst) {
    const int bz = z0 + st;
    const int slot0 = bz % NSLOT;
    const int nb = nb_next;
    commit(st, buf, b_cur == 0);
    __syncthreads();
    // next unit of work: the rest of this step, or the next step
    const int b_nxt = b_cur + MAXB < n_cur ? b_cur + MAXB : 0;
    const int st_nxt = b_nxt ? st : st + 1;
    int n_nxt = n_cur;
    if (st_nxt < nst) n_nxt = fetch(st_nxt, b_nxt, b_nxt == 0);
    for (int k = wave; k < nb; k += NWV) {
      const int pb = s_pb[buf][k];
      const int rx = (pb & 1023) - x0, ry = ((pb >> 10) & 1023) - y0;
      const T *w = s_w[buf] + k * NW;
      const T wxy = act ? w[dx] * w[P + dy] : T(0);
      T acc[CC];
#pragma unroll
      for (int c = 0; c < CC; ++c) acc[c] = T(0);
      if (act) {
        const int off = (ry + dy) * EX + rx + dx;
#pragma unroll
        for (int j = 0; j < P; ++j) {
          const T wz = w[2 * P + j];
          const int sl = slot0 + j >= NSLOT ? slot0 + j - NSLOT : slot0 + j;
          const T *pl = win + sl * PLANE + off;
#pragma unroll
          for (int c = 0; c < CC; ++c) acc[c] += wz * pl[c * E * EX];
        }
      }
      const int o = s_perm[buf][k];
#pragma unroll
      for (int c = 0; c < CC; ++c) {
        const T v = wave_sum(acc[c] * wxy);
        if (lane == 0) {
          const int comp = cc0 + c, r = comp / 3, d = comp - 3 * r;
          T *u = out + r * ldout + 3L * o + d;
          *u = beta == T(0) ? alpha * v : beta * *u + alpha * v;
        }
      }
    }
    b_cur = b_nxt;
    st = st_nxt;
    n_cur = n_nxt;
    buf ^= 1;
  }
}

// ---------------------------------------------------------------------------
// FFT passes: four-step transforms of length K = N1 * N2 (Split<K>).  Step A:
// per n2, an N1-point FFT over x[n1 N2 + n2] in registers, times
// W_K^{n2 k1}, stored at k1 N2 + n2.  Step B: per k1, an N2-point FFT over
// that contiguous block in registers gives X[k1 + N1 k2].  One LDS transpose
// per transform.  tw[t] = exp(-2 pi i t/K), t < K.

template <typename T, int K>
__device__ __forceinline__ void load_tw(cpx<T> *s_tw, const cpx<T> *tw) {
  for (int t = threadIdx.x; t < K; t += blockDim.x) s_tw[t] = tw[t];
}

// Forward r2c along x: two real rows per complex line (z = a + i b).
template <typename T, int K, int NL, int NT>
__global__ __launch_bounds__(NT) void k_fft_x_r2c(T *__restrict__ grid, int RS, long npairs,
                                                  const cpx<T> *__restrict__ tw) {
  constexpr int N1 = Split<K>::N1, N2 = Split<K>::N2;
  constexpr int Kh = K / 2 + 1;
  constexpr int LS = K + 1;
  __shared__ cpx<T> buf[NL * LS];
  __shared__ cpx<T> s_tw[K];
  load_tw<T, K>(s_tw, tw);
  const long q0 = (long)blockIdx.x * NL;
  for (int t = threadIdx.x; t < NL * K; t += NT) {
    const int l = t / K, n = t - l * K;
    const long q = q0 + l;
    cpx<T> v = {T(0), T(0)};
    if (q < npairs) {
      const T *ra = grid + (2 * q) * RS;
      v = {ra[n], ra[RS + n]};
    }
    buf[l * LS + n] = v;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < NL * N2; t += NT) {  // step A, in place per thread
    const int l = t / N2, n2 = t - l * N2;
    cpx<T> a[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) a[n1] = buf[l * LS + n1 * N2 + n2];
    reg_fft<N1, -1>(a);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) a[k1] = cmul(a[k1], twiddle<-1>(s_tw, n2 * k1));
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) buf[l * LS + k1 * N2 + n2] = a[k1];
  }
  __syncthreads();
  constexpr int PB = (NL * N1 + NT - 1) / NT;
  cpx<T> b[PB][N2];
#pragma unroll
  for (int u = 0; u < PB; ++u) {  // step B into registers
    const int t = threadIdx.x + u * NT;
    if (t < NL * N1) {
      const int l = t / N1, k1 = t - l * N1;
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) b[u][n2] = buf[l * LS + k1 * N2 + n2];
      reg_fft<N2, -1>(b[u]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < PB; ++u) {  // natural order back into LDS
    const int t = threadIdx.x + u * NT;
    if (t < NL * N1) {
      const int l = t / N1, k1 = t - l * N1;
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) buf[l * LS + k1 + N1 * k2] = b[u][k2];
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < NL * Kh; t += NT) {
    const int l = t / Kh, k = t - l * Kh;
    const long q = q0 + l;
    if (q < npairs) {
      const cpx<T> Z = buf[l * LS + k];
      const cpx<T> Zm = cconj(buf[l * LS + (k == 0 ? 0 : K - k)]);
      const cpx<T> A = {T(0.5) * (Z.x + Zm.x), T(0.5) * (Z.y + Zm.y)};
      const cpx<T> D = csub(Z, Zm);
      const cpx<T> B = {T(0.5) * D.y, T(-0.5) * D.x};
      cpx<T> *ca = reinterpret_cast<cpx<T> *>(grid + (2 * q) * RS);
      ca[k] = A;
      ca[RS / 2 + k] = B;
    }
  }
}

// Backward c2r along x.  The imaginary parts of the k = 0 and k = K/2
// entries are dropped, so the result is the real part of the inverse DFT.
template <typename T, int K, int NL, int NT>
__global__ __launch_bounds__(NT) void k_fft_x_c2r(T *__restrict__ grid, int RS, long npairs,
                                                  const cpx<T> *__restrict__ tw) {
  constexpr int N1 = Split<K>::N1, N2 = Split<K>::N2;
  constexpr int Kh = K / 2 + 1;
  constexpr int LS = K + 1;
  __shared__ cpx<T> buf[NL * LS];
  __shared__ cpx<T> s_tw[K];
  load_tw<T, K>(s_tw, tw);
  const long q0 = (long)blockIdx.x * NL;
  for (int t = threadIdx.x; t < NL * Kh; t += NT) {
    const int l = t / Kh, k = t - l * Kh;
    const long q = q0 + l;
    cpx<T> A = {T(0), T(0)}, B = {T(0), T(0)};
    if (q < npairs) {
      const cpx<T> *ca = reinterpret_cast<const cpx<T> *>(grid + (2 * q) * RS);
      A = ca[k];
      B = ca[RS / 2 + k];
    }
    if (k == 0 || k == K / 2) {
      A.y = T(0);
      B.y = T(0);
    }
    buf[l * LS + k] = {A.x - B.y, A.y + B.x};
    if (k > 0 && k < K / 2) buf[l * LS + K - k] = {A.x + B.y, B.x - A.y};
  }
  __syncthreads();
  for (int t = threadIdx.x; t < NL * N2; t += NT) {  // step A, in place per thread
    const int l = t / N2, n2 = t - l * N2;
    cpx<T> a[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) a[n1] = buf[l * LS + n1 * N2 + n2];
    reg_fft<N1, +1>(a);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) a[k1] = cmul(a[k1], twiddle<+1>(s_tw, n2 * k1));
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) buf[l * LS + k1 * N2 + n2] = a[k1];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < NL * N1; t += NT) {  // step B, straight to the rows
    const int l = t / N1, k1 = t - l * N1;
    const long q = q0 + l;
    cpx<T> b[N2];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) b[n2] = buf[l * LS + k1 * N2 + n2];
    reg_fft<N2, +1>(b);
    if (q < npairs) {
      T *ra = grid + (2 * q) * RS;
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        ra[k1 + N1 * k2] = b[k2].x;
        ra[RS + k1 + N1 * k2] = b[k2].y;
      }
    }
  }
}

// Complex FFT along y, in place.  Block = (x chunk of XW, plane z, comp).
template <typename T, int K, int XW, int S, int NT>
__global__ __launch_bounds__(NT) void k_fft_y(T *__restrict__ grid, long comp_stride, int Khp,
                                              int nxc, const cpx<T> *__restrict__ tw) {
  constexpr int N1 = Split<K>::N1, N2 = Split<K>::N2;
  constexpr int Kh = K / 2 + 1;
  constexpr int LS = K + 1;
  __shared__ cpx<T> buf[XW * LS];
  __shared__ cpx<T> s_tw[K];
  load_tw<T, K>(s_tw, tw);
  __syncthreads();
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const int xc = lb % nxc;
  const int zc = lb / nxc;  // z + K*comp
  const int comp = zc / K, z = zc - comp * K;
  const int x0 = xc * XW;
  cpx<T> *plane = reinterpret_cast<cpx<T> *>(grid + comp * comp_stride) + (size_t)z * K * Khp;
  for (int t = threadIdx.x; t < XW * N2; t += NT) {  // step A from global
    const int xo = t % XW, n2 = t / XW;
    const int x = x0 + xo;
    cpx<T> a[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1)
      a[n1] = x < Kh ? plane[(size_t)(n1 * N2 + n2) * Khp + x] : cpx<T>{T(0), T(0)};
    reg_fft<N1, S>(a);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) a[k1] = cmul(a[k1], twiddle<S>(s_tw, n2 * k1));
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) buf[xo * LS + k1 * N2 + n2] = a[k1];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < XW * N1; t += NT) {  // step B to global
    const int xo = t % XW, k1 = t / XW;
    const int x = x0 + xo;
    cpx<T> b[N2];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) b[n2] = buf[xo * LS + k1 * N2 + n2];
    reg_fft<N2, S>(b);
    if (x < Kh) {
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) plane[(size_t)(k1 + N1 * k2) * Khp + x] = b[k2];
    }
  }
}

// Forward FFT along z, RPY influence (old-version/spme_kernel.c:73-228),
// backward FFT along z, for the 3 components of one right-hand side.
// Block = (x chunk, y) x rhs.  tab[(z*K + y)*Khp + x] = psi(|k|)|b(k)|^2/V.
// Wave numbers follow old-version/spme.c:95-107 (index K/2 -> +pi K/L).
// The spectrum of a (column, k1) stays in registers between the forward
// N2-point FFT, the 3x3 influence and the backward N2-point FFT.
template <typename T, int K, int XW, int NT>
__global__ __launch_bounds__(NT) void k_fft_z_infl(T *__restrict__ grid, long comp_stride,
                                                   int Khp, int nxc, const T *__restrict__ tab,
                                                   T dk, const cpx<T> *__restrict__ tw) {
  constexpr int N1 = Split<K>::N1, N2 = Split<K>::N2;
  constexpr int Kh = K / 2 + 1;
  constexpr int LS = K + 1;
  __shared__ cpx<T> buf[3 * XW * LS];
  __shared__ cpx<T> s_tw[K];
  __shared__ T s_tab[XW * K];
  load_tw<T, K>(s_tw, tw);
  // adjacent x chunks (halves of one cache line) on the same XCD
  const int lb = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const int xc = lb % nxc, y = (lb / nxc) % K;
  const int x0 = xc * XW;
  const int rhs = lb / (nxc * K);
  for (int t = threadIdx.x; t < XW * K; t += NT) {
    const int z = t / XW, xo = t - z * XW;
    const int x = x0 + xo;
    s_tab[xo * K + z] = x < Kh ? tab[((size_t)z * K + y) * Khp + x] : T(0);
  }
  __syncthreads();
  cpx<T> *g0 = reinterpret_cast<cpx<T> *>(grid + (3L * rhs) * comp_stride);
  const long cs = comp_stride / 2;  // complex elements
  for (int t = threadIdx.x; t < 3 * XW * N2; t += NT) {  // step A (forward) from global
    const int xo = t % XW, r = t / XW;
    const int c = r % 3, n2 = r / 3;
    const int x = x0 + xo;
    cpx<T> a[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1)
      a[n1] = x < Kh ? g0[c * cs + ((size_t)(n1 * N2 + n2) * K + y) * Khp + x]
                     : cpx<T>{T(0), T(0)};
    reg_fft<N1, -1>(a);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) a[k1] = cmul(a[k1], twiddle<-1>(s_tw, n2 * k1));
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) buf[(c * XW + xo) * LS + k1 * N2 + n2] = a[k1];
  }
  __syncthreads();
  const T ky = T(y <= K / 2 ? y : y - K) * dk;
  for (int t = threadIdx.x; t < XW * N1; t += NT) {  // step B, influence, inverse step B
    const int xo = t % XW, k1 = t / XW;
    const int x = x0 + xo;
    cpx<T> b[3][N2];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) b[c][n2] = buf[(c * XW + xo) * LS + k1 * N2 + n2];
      reg_fft<N2, -1>(b[c]);
    }
    const T kx = T(x) * dk;  // x <= K/2 in the half spectrum
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) {
      const int z = k1 + N1 * k2;
      const T kz = T(z <= K / 2 ? z : z - K) * dk;
      const T k2n = kx * kx + ky * ky + kz * kz;
      const T m2 = s_tab[xo * K + z];
      // k = 0 and padding columns carry m2 = 0; guard the division there
      const T ik = k2n > T(0) ? T(1) / k2n : T(0);
      const T B0 = m2 - m2 * kx * kx * ik, B1 = -m2 * kx * ky * ik, B2 = -m2 * kx * kz * ik;
      const T B3 = m2 - m2 * ky * ky * ik, B4 = -m2 * ky * kz * ik, B5 = m2 - m2 * kz * kz * ik;
      const cpx<T> va = b[0][k2], vb = b[1][k2], vc = b[2][k2];
      b[0][k2] = {B0 * va.x + B1 * vb.x + B2 * vc.x, B0 * va.y + B1 * vb.y + B2 * vc.y};
      b[1][k2] = {B1 * va.x + B3 * vb.x + B4 * vc.x, B1 * va.y + B3 * vb.y + B4 * vc.y};
      b[2][k2] = {B2 * va.x + B4 * vb.x + B5 * vc.x, B2 * va.y + B4 * vb.y + B5 * vc.y};
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      reg_fft<N2, +1>(b[c]);
#pragma unroll
      for (int n2 = 1; n2 < N2; ++n2) b[c][n2] = cmul(b[c][n2], twiddle<+1>(s_tw, n2 * k1));
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) buf[(c * XW + xo) * LS + k1 * N2 + n2] = b[c][n2];
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 3 * XW * N2; t += NT) {  // inverse step A to global
    const int xo = t % XW, r = t / XW;
    const int c = r % 3, n2 = r / 3;
    const int x = x0 + xo;
    cpx<T> a[N1];
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) a[k1] = buf[(c * XW + xo) * LS + k1 * N2 + n2];
    reg_fft<N1, +1>(a);
    if (x < Kh) {
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1)
        g0[c * cs + ((size_t)(n2 + N2 * n1) * K + y) * Khp + x] = a[n1];
    }
  }
}

// ---------------------------------------------------------------------------
// Influence table (old-version/spme.c:128-150), FP64:
//   tab[(z*K + y)*Khp + x] = psi(|k|) |b_x|^2 |b_y|^2 |b_z|^2 / V, x <= K/2,
// psi(k) = 6 pi e^{-k^2/4xi^2} (1 - k^2/3)(1 + k^2/4xi^2 + k^4/8xi^4)/k^2
// (old-version/spme.c:43-60 with unit radii); 0 at k = 0 and in padding.
template <typename T>
__global__ __launch_bounds__(256) void k_influence_table(const double *__restrict__ bsq,
                                                         int K, int Khp, double dk, double xi,
                                                         double inv_v, T *__restrict__ tab) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)K * K * Khp) return;
  const int x = int(t % Khp);
  const long zy = t / Khp;
  const int y = int(zy % K), z = int(zy / K);
  double v = 0.0;
  if (x <= K / 2) {
    const double kx = x * dk, ky = (y <= K / 2 ? y : y - K) * dk;
    const double kz = (z <= K / 2 ? z : z - K) * dk;
    const double k2 = kx * kx + ky * ky + kz * kz;
    if (k2 > 0.0) {
      const double q = k2 / (xi * xi);
      const double psi = 6.0 * M_PI * exp(-q / 4.0) * (1.0 - k2 / 3.0) / k2 *
                         (1.0 + q * (0.25 + q / 8.0));
      v = psi * bsq[x] * bsq[y] * bsq[z] * inv_v;
    }
  }
  tab[t] = T(v);
}

// ---------------------------------------------------------------------------
// Host launchers.

template <typename T>
void launch_influence_table(const double *bsq, int K, int Khp, double dk, double xi,
                            double inv_v, T *tab, hipStream_t st) {
  const long n = (long)K * K * Khp;
  hipLaunchKernelGGL(k_influence_table<T>, dim3(int((n + 255) / 256)), dim3(256), 0, st, bsq, K,
                     Khp, dk, xi, inv_v, tab);
}
template void launch_influence_table<float>(const double *, int, int, double, double, double,
                                            float *, hipStream_t);
template void launch_influence_table<double>(const double *, int, int, double, double, double,
                                             double *, hipStream_t);

static int grid1(long n, int b) { return int((n + b - 1) / b); }

void launch_pbuild(const double *pos, int N, double L, int K, int P, const WTable &W, int TX,
                   int TY, int *base_tmp, double *w_tmp, int *key_tmp, int *counts, int *err,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_pbuild, dim3(grid1(N, 256)), dim3(256), 0, st, pos, N, L, K, P, W, TX,
                     TY, base_tmp, w_tmp, key_tmp, counts, err);
}

void launch_scan(const int *in, int *out, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, in, out, n);
}

void launch_pscatter(const int *key, const int *boff, int *cursor, int *perm, int N,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_pscatter, dim3(grid1(N, 256)), dim3(256), 0, st, key, boff, cursor,
                     perm, N);
}

void launch_bucket_sort(const int *boff, int *perm, int nkeys, hipStream_t st) {
  hipLaunchKernelGGL(k_bucket_sort, dim3(grid1(nkeys, 256)), dim3(256), 0, st, boff, perm,
                     nkeys);
}

template <typename T>
void launch_pgather(const int *perm, const int *base_tmp, const double *w_tmp, int *pbase,
                    T *pw, int N, int P3, hipStream_t st) {
  hipLaunchKernelGGL(k_pgather<T>, dim3(grid1(N, 256)), dim3(256), 0, st, perm, base_tmp, w_tmp,
                     pbase, pw, N, P3);
}

template <typename T>
void launch_permute_in(const int *perm, const T *in, long ldin, T *fs, int N, int C,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_permute_in<T>, dim3(grid1((long)N * C, 256)), dim3(256), 0, st, perm, in,
                     ldin, fs, N, C);
}

template <typename T, int P, int CC>
static void spread_cc(const SpreadArgs<T> &a, hipStream_t st) {
  constexpr int TX = 8, TY = 8;
  const int ntiles = (a.K / TX) * (a.K / TY);
  const int nz = (a.K + a.ZC - 1) / a.ZC;
  hipLaunchKernelGGL((k_spread<T, P, CC, TX, TY>), dim3(ntiles, nz, a.C / CC), dim3(64), 0, st,
                     a.boff, a.pbase, a.pw, a.fs, a.C, a.grid, a.comp_stride, a.RS, a.K, a.ZC);
}

// Components per pass: the register window holds P*CC values per lane.
template <typename T, int P>
static int spread_p(const SpreadArgs<T> &a, hipStream_t st) {
  if (a.CC == 3) spread_cc<T, P, 3>(a, st);
  else if (a.CC == 6) spread_cc<T, P, 6>(a, st);
  else if (a.CC == 12) spread_cc<T, P, 12>(a, st);
  else return -1;
  return 0;
}

template <typename T>
int launch_spread(const SpreadArgs<T> &a, hipStream_t st) {
  if (a.C % a.CC) return -1;
  switch (a.P) {
    case 4: return spread_p<T, 4>(a, st);
    case 6: return spread_p<T, 6>(a, st);
    case 8: return spread_p<T, 8>(a, st);
  }
  return -1;
}

// Interpolation shape: FP32 takes 3 components per block over a 16 x 16
// tile; FP64 takes 1 (LDS: P planes x CC x (TXY+P-1)^2 values).
template <typename T, int P, int CC, int TXY>
static void interp_cc(const InterpArgs<T> &a, hipStream_t st) {
  constexpr int NT = 512;
  const int nt = a.K / TXY;
  const int zc = a.K >= 64 ? a.K / 4 : a.K;  // <= 64
  hipLaunchKernelGGL((k_interp<T, P, CC, TXY, NT>), dim3(nt * nt, (a.K + zc - 1) / zc, a.C / CC),
                     dim3(NT), 0, st, a.boff, a.pbase, a.pw, a.perm, a.grid, a.comp_stride,
                     a.RS, a.K, a.C, zc, a.alpha, a.beta, a.out, a.ldout);
}

template <typename T, int P>
static int interp_p(const InterpArgs<T> &a, hipStream_t st) {
  constexpr int CC = sizeof(T) == 4 ? 3 : 1;
  if (a.C % CC) return -1;
  if (a.K % 16 == 0) interp_cc<T, P, CC, 16>(a, st);
  else interp_cc<T, P, CC, 8>(a, st);
  return 0;
}

template <typename T>
int launch_interp(const InterpArgs<T> &a, hipStream_t st) {
  switch (a.P) {
    case 4: return interp_p<T, 4>(a, st);
    case 6: return interp_p<T, 6>(a, st);
    case 8: return interp_p<T, 8>(a, st);
  }
  return -1;
}

// Line batch sizes: keep ~24 KB of LDS per FFT block.
template <typename T> struct FftShape {
  static constexpr int NLX = sizeof(T) == 8 ? 8 : 16;  // complex lines per x block
  static constexpr int XW = 16;                        // y-pass columns
  static constexpr int XWZ = sizeof(T) == 8 ? 4 : 8;   // z-pass columns (x3 comps)
};

template <typename T, int K>
static void fft3d_k(const FftArgs<T> &a, hipStream_t st, hipEvent_t *ev) {
  using S = FftShape<T>;
  constexpr int NT = 256;
  const long rows = (long)a.ncomp * K * K;
  const long npairs = rows / 2;
  const int Khp = a.RS / 2;
  const int nxc = (K / 2 + 1 + S::XW - 1) / S::XW;
  const int nxcz = (K / 2 + 1 + S::XWZ - 1) / S::XWZ;
  auto mark = [&](int i) {
    if (ev) (void)hipEventRecord(ev[i], st);
  };
  hipLaunchKernelGGL((k_fft_x_r2c<T, K, S::NLX, NT>), dim3(grid1(npairs, S::NLX)), dim3(NT), 0,
                     st, a.grid, a.RS, npairs, a.tw);
  mark(0);
  hipLaunchKernelGGL((k_fft_y<T, K, S::XW, -1, NT>), dim3(nxc * K * a.ncomp), dim3(NT), 0, st,
                     a.grid, a.comp_stride, Khp, nxc, a.tw);
  mark(1);
  hipLaunchKernelGGL((k_fft_z_infl<T, K, S::XWZ, NT>), dim3(nxcz * K, a.ncomp / 3), dim3(NT), 0,
                     st, a.grid, a.comp_stride, Khp, nxcz, a.tab, a.dk, a.tw);
  mark(2);
  hipLaunchKernelGGL((k_fft_y<T, K, S::XW, +1, NT>), dim3(nxc * K * a.ncomp), dim3(NT), 0, st,
                     a.grid, a.comp_stride, Khp, nxc, a.tw);
  mark(3);
  hipLaunchKernelGGL((k_fft_x_c2r<T, K, S::NLX, NT>), dim3(grid1(npairs, S::NLX)), dim3(NT), 0,
                     st, a.grid, a.RS, npairs, a.tw);
  mark(4);
}

template <typename T>
int launch_fft_influence(const FftArgs<T> &a, hipStream_t st, hipEvent_t *ev) {
  switch (a.K) {
    case 16: fft3d_k<T, 16>(a, st, ev); return 0;
    case 24: fft3d_k<T, 24>(a, st, ev); return 0;
    case 32: fft3d_k<T, 32>(a, st, ev); return 0;
    case 48: fft3d_k<T, 48>(a, st, ev); return 0;
    case 64: fft3d_k<T, 64>(a, st, ev); return 0;
    case 96: fft3d_k<T, 96>(a, st, ev); return 0;
    case 128: fft3d_k<T, 128>(a, st, ev); return 0;
    case 192: fft3d_k<T, 192>(a, st, ev); return 0;
    case 256: fft3d_k<T, 256>(a, st, ev); return 0;
  }
  return -1;
}

bool fft_size_supported(int K) {
  return K == 16 || K == 24 || K == 32 || K == 48 || K == 64 || K == 96 || K == 128 ||
         K == 192 || K == 256;
}

template void launch_pgather<float>(const int *, const int *, const double *, int *, float *,
                                    int, int, hipStream_t);
template void launch_pgather<double>(const int *, const int *, const double *, int *, double *,
                                     int, int, hipStream_t);
template void launch_permute_in<float>(const int *, const float *, long, float *, int, int,
                                       hipStream_t);
template void launch_permute_in<double>(const int *, const double *, long, double *, int, int,
                                        hipStream_t);
template int launch_spread<float>(const SpreadArgs<float> &, hipStream_t);
template int launch_spread<double>(const SpreadArgs<double> &, hipStream_t);
template int launch_interp<float>(const InterpArgs<float> &, hipStream_t);
template int launch_interp<double>(const InterpArgs<double> &, hipStream_t);
template int launch_fft_influence<float>(const FftArgs<float> &, hipStream_t, hipEvent_t *);
template int launch_fft_influence<double>(const FftArgs<double> &, hipStream_t, hipEvent_t *);

}  // namespace spme
