// Device bodies of the apply-path kernels.  Each body processes one logical
// block (one row or element for the elementwise ones) given explicitly, so
// that a kernel can run it once per hardware block for one system
// (spme_kernels.hip) or loop it over the blocks of many systems
// (spme_group.hip).  Included by those sources only.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "spme_kernels.h"
#include "spme_launch.h"

namespace spme {

// fs[s*C + 3r + d] = in[r*ldin + 3*perm[s] + d]  (sorted, particle-major);
// element t of the N*C outputs.
template <typename T>
__device__ __forceinline__ void permute_elem(const int *__restrict__ perm,
                                             const T *__restrict__ in, long ldin,
                                             T *__restrict__ fs, int C, long t) {
  const int s = int(t / C), c = int(t - (long)s * C);
  const int r = c / 3, d = c - 3 * r;
  fs[t] = in[r * ldin + 3L * perm[s] + d];
}

// The same into the real-space layout (cell order, the right-hand sides of
// one component together): fs[s*C + d*R + r], R = C/3; element t of the
// N*C outputs.
template <typename T>
__device__ __forceinline__ void permute_elem_cell(const int *__restrict__ perm,
                                                  const T *__restrict__ in, long ldin,
                                                  T *__restrict__ fs, int C, long t) {
  const int s = int(t / C), c = int(t - (long)s * C);
  const int R = C / 3, d = c / R, r = c - d * R;
  fs[t] = in[r * ldin + 3L * perm[s] + d];
}

// Non-temporal store of a 4-, 8- or 16-byte value.  The forward plane pass
// and the z pass write the half spectrum this way: the next pass reads it
// back from HBM in any case (the mesh is 1.6 GB at 16 RHS and K = 192), and
// the hint keeps the written lines from displacing what the pass still
// reads (measured at K = 192, FP32: 0.676 -> 0.636 ms for the z pass,
// 0.677 -> 0.666 ms for the forward plane pass).  The inverse plane pass
// and spreading keep ordinary stores (there the hint cost 0.04 and 1.2 ms).
template <typename V>
__device__ __forceinline__ void st_stream(V *p, const V &v) {
  typedef float v4 __attribute__((ext_vector_type(4)));
  typedef float v2 __attribute__((ext_vector_type(2)));
  if constexpr (sizeof(V) == 16) {
    v4 t;
    __builtin_memcpy(&t, &v, 16);
    __builtin_nontemporal_store(t, reinterpret_cast<v4 *>(p));
  } else if constexpr (sizeof(V) == 8) {
    v2 t;
    __builtin_memcpy(&t, &v, 8);
    __builtin_nontemporal_store(t, reinterpret_cast<v2 *>(p));
  } else {
    __builtin_nontemporal_store(v, p);
  }
}

// ---------------------------------------------------------------------------
// Spread (old-version/spme_kernel.c:231-312): grid[c] = sum_i q_i (x) f_ic.
// One wave owns a TX x TY column of cells over a z range and keeps a window
// of P planes per component in registers; particles are visited in
// increasing base z, and the lowest plane of the window is final (and
// stored) once the sweep has passed every base that reaches it; the window
// then shifts by one plane.  Every cell is written exactly once, so the
// mesh needs no clearing and no atomics are used.
// Latency: the bucket offsets of a sweep (at most 96 planes; longer z
// chunks run several sweeps, which changes no sum) are read into LDS up
// front.  The particles (base, weights, forces) of G consecutive steps are
// one staging group (bucket ranges of consecutive bases are contiguous in
// the sorted order): they are copied into LDS by direct global->LDS loads
// (no registers), double-buffered, so the next group's copy is in flight
// while the current group is spread.  A group that wraps through z = 0 or
// exceeds a batch is staged step by step instead.  Each step adds its
// particles in ascending sorted order per tile either way, so the result
// does not depend on the staging.  The block is one wave: no barriers (LDS
// accesses of one wave execute in order; the compiler waits for a direct
// load before the LDS reads it feeds).
// f(st0 + U, integral_constant<U>) for U < P while st0 + U < n: a loop of P
// steps unrolled so that U is a constant in each copy of the body.
template <int U, int P, class F>
__device__ __forceinline__ void unroll_steps(int st0, int n, F &f) {
  if constexpr (U < P) {
    if (st0 + U < n) {
      f(st0 + U, std::integral_constant<int, U>{});
      unroll_steps<U + 1, P>(st0, n, f);
    }
  }
}

// Steps per staging group (an even number of groups per P unrolled steps,
// so the two buffers alternate) and particles per batch (one bit each of a
// 64-bit mask; about twice a group's expected count at 0.03-0.05 particles
// per cell).
template <int P> struct SpreadStage {
  static constexpr int G = P % 4 == 0 ? P / 4 : P / 2;
  static constexpr int MAXB = P % 4 == 0 ? 32 : 64;
};

// s_waitcnt vmcnt(0): every vector memory access of the wave, direct LDS
// loads included, has completed.  The empty asm keeps the compiler from
// moving memory accesses across the wait (the wait itself touches no memory).
__device__ __forceinline__ void wait_vm_all() {
  __builtin_amdgcn_s_waitcnt(0x0F70);
  asm volatile("" ::: "memory");
}

template <typename T, int P, int CC, int TX, int TY>
__device__ __forceinline__ void spread_block(const int *__restrict__ boff,
                                             const int *__restrict__ pbase,
                                             const T *__restrict__ pw, const T *__restrict__ fs,
                                             int C, T *__restrict__ grid, long comp_stride,
                                             int RS, int K, int ZC, int lb) {
  static_assert(TX * TY == 64, "one wave per tile");
  constexpr int NXR = 1 + (P + TX - 2) / TX;
  constexpr int NYR = 1 + (P + TY - 2) / TY;
  constexpr int NTL = NXR * NYR;
  constexpr int MAXZ = 96;            // planes per sweep
  constexpr int MAXST = MAXZ + P - 1;
  constexpr int G = SpreadStage<P>::G, MAXB = SpreadStage<P>::MAXB;
  static_assert(P % (2 * G) == 0 && MAXB <= 64, "buffers alternate per group");
  using Mask = std::conditional_t<(MAXB > 32), unsigned long long, unsigned>;
  constexpr Mask ONE = 1;
  constexpr int NW = 3 * P;
  constexpr int Q = sizeof(T) / 4;                 // dwords per value
  constexpr int RWD = (MAXB * NW * Q + 63) / 64;   // weight dwords per lane
  constexpr int RFD = (MAXB * CC * Q + 63) / 64;   // force dwords per lane
  __shared__ int s_lo[NTL * MAXST];
  __shared__ int s_n[NTL * MAXST];
  __shared__ int s_pb0[MAXB], s_pb1[MAXB];
  __shared__ __attribute__((aligned(16))) T s_w0[MAXB * NW], s_w1[MAXB * NW];
  __shared__ __attribute__((aligned(16))) T s_f0[MAXB * CC], s_f1[MAXB * CC];
  const int ntx = K / TX, nty = K / TY, nz = (K + ZC - 1) / ZC;
  const int tile = lb % (ntx * nty), rest = lb / (ntx * nty);
  const int tx = tile % ntx, ty = tile / ntx;
  const int lane = threadIdx.x;
  const int x = tx * TX + lane % TX, y = ty * TY + lane / TX;
  const int z0 = (rest % nz) * ZC, zend = min(z0 + ZC, K);
  const int cc0 = (rest / nz) * CC;
  const unsigned *pw_d = reinterpret_cast<const unsigned *>(pw);
  const unsigned *fs_d = reinterpret_cast<const unsigned *>(fs);
  // 16-byte staging: weight records of whole 16-byte pieces; force slices
  // too when the rows (C values) keep them 16-byte aligned
  constexpr bool W16 = NW * Q % 4 == 0, F16 = CC * Q % 4 == 0;
  const bool f16ok = C * Q % 4 == 0;

  // copy particles [b0, b0 + nb) of the concatenated sorted ranges
  // (lo[t], cnt[t]), t < NTL, into buffer B: base, NW weights and the CC
  // forces of this pass, 16 bytes or one dword per lane per load (32-bit
  // dword offsets: the engine bounds N so that N * C values fit)
  auto stage = [&](auto bc, const int (&lo)[NTL], const int (&cnt)[NTL], int b0, int nb) {
    constexpr int B = decltype(bc)::value;
    int *spb = B ? s_pb1 : s_pb0;
    T *sw = B ? s_w1 : s_w0;
    T *sf = B ? s_f1 : s_f0;
    auto locate = [&](int k) {
      int i = lo[0] + k, rem = k - cnt[0];
#pragma unroll
      for (int t = 1; t < NTL; ++t) {
        i = rem >= 0 ? lo[t] + rem : i;
        rem -= cnt[t];
      }
      return i;
    };
    if (lane < nb) __builtin_amdgcn_global_load_lds(pbase + locate(b0 + lane), spb, 4, 0, 0);
    // a particle's weights (and, when its rows are 16-byte aligned, its
    // forces) in 16-byte pieces: a quarter of the loads and index searches
    // of one dword per lane
    if constexpr (W16) {
      constexpr int PW = NW * Q / 4, RW16 = (MAXB * PW + 63) / 64;
#pragma unroll
      for (int r = 0; r < RW16; ++r) {
        const int e = lane + 64 * r;
        if (e < nb * PW) {
          const int k = e / PW;
          const unsigned off =
              unsigned(locate(b0 + k)) * unsigned(NW * Q) + unsigned(e - k * PW) * 4u;
          __builtin_amdgcn_global_load_lds(pw_d + off, reinterpret_cast<char *>(sw) + 1024 * r, 16, 0, 0);
        }
      }
    } else {
#pragma unroll
      for (int r = 0; r < RWD; ++r) {
        const int e = lane + 64 * r;
        if (e < nb * NW * Q) {
          const int k = e / (NW * Q);
          const unsigned off = unsigned(locate(b0 + k)) * unsigned(NW * Q) + unsigned(e - k * NW * Q);
          __builtin_amdgcn_global_load_lds(pw_d + off, reinterpret_cast<char *>(sw) + 256 * r, 4, 0, 0);
        }
      }
    }
    if (F16 && f16ok) {
      constexpr int PF = CC * Q / 4, RF16 = (MAXB * PF + 63) / 64;
#pragma unroll
      for (int r = 0; r < RF16; ++r) {
        const int e = lane + 64 * r;
        if (e < nb * PF) {
          const int k = e / PF;
          const unsigned off =
              (unsigned(locate(b0 + k)) * unsigned(C) + unsigned(cc0)) * unsigned(Q) +
              unsigned(e - k * PF) * 4u;
          __builtin_amdgcn_global_load_lds(fs_d + off, reinterpret_cast<char *>(sf) + 1024 * r, 16, 0, 0);
        }
      }
    } else {
#pragma unroll
      for (int r = 0; r < RFD; ++r) {
        const int e = lane + 64 * r;
        if (e < nb * CC * Q) {
          const int k = e / (CC * Q);
          const unsigned off =
              (unsigned(locate(b0 + k)) * unsigned(C) + unsigned(cc0)) * unsigned(Q) +
              unsigned(e - k * CC * Q);
          __builtin_amdgcn_global_load_lds(fs_d + off, reinterpret_cast<char *>(sf) + 256 * r, 4, 0, 0);
        }
      }
    }
  };
  // staged particles k < nb of buffer B whose stencil covers this lane's column
  auto covers = [&](auto bc, int nb) {
    constexpr int B = decltype(bc)::value;
    const int *spb = B ? s_pb1 : s_pb0;
    Mask m = 0;
    for (int k = 0; k < nb; ++k) {
      const int pb = spb[k];
      int dx = x - (pb & 1023);
      dx += dx < 0 ? K : 0;
      int dy = y - ((pb >> 10) & 1023);
      dy += dy < 0 ? K : 0;
      m |= (dx < P && dy < P) ? ONE << k : Mask(0);
    }
    return m;
  };

  for (int zb = z0; zb < zend; zb += MAXZ) {
    const int zs = zb - (P - 1);
    const int nsteps = min(zb + MAXZ, zend) - zs;
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

    // bucket offsets are wave-uniform: keep them in scalar registers
    auto lo_of = [&](int l) { return __builtin_amdgcn_readfirstlane(s_lo[l]); };
    auto n_of = [&](int l) { return __builtin_amdgcn_readfirstlane(s_n[l]); };
    // group starting at step st: its ranges, and, per tile, the offset of
    // the tile's particles in the staging order minus its first sorted
    // index; true if it is one batch
    auto plan = [&](int st, int (&glo)[NTL], int (&gcnt)[NTL], int (&gdel)[NTL], int &tot) {
      tot = 0;
      if (st >= nsteps) return false;
      const int gn = min(G, nsteps - st), bz0 = zs + st;
#pragma unroll
      for (int t = 0; t < NTL; ++t) {
        const int l = t * nsteps + st;
        glo[t] = lo_of(l);
        gcnt[t] = lo_of(l + gn - 1) + n_of(l + gn - 1) - glo[t];
        gdel[t] = tot - glo[t];
        tot += gcnt[t];
      }
      return !(bz0 < 0 && bz0 + gn > 0) && tot <= MAXB;
    };
    // the group being copied (n*) and the current one (c*)
    int nlo[NTL], ncnt[NTL], ndel[NTL], ntot;
    bool nok = plan(0, nlo, ncnt, ndel, ntot);
    if (nok) stage(std::integral_constant<int, 0>{}, nlo, ncnt, 0, ntot);
    bool cok = false;
    int cdel[NTL];
    Mask cmask = 0;

    T acc[P][CC];
#pragma unroll
    for (int j = 0; j < P; ++j)
#pragma unroll
      for (int c = 0; c < CC; ++c) acc[j][c] = T(0);

    // Step st adds the particles whose stencil starts at plane zs + st to the
    // planes zs + st + j, j < P, held in acc[(st + j) % P]; after the step,
    // plane zs + st (slot st % P) is complete: it is written and its slot
    // cleared for plane zs + st + P.  The steps are unrolled by P so that
    // every slot index and staging buffer is a constant.  In the first P
    // steps of a sweep (PRO) the planes below the sweep, which another block
    // or sweep stores, are not accumulated: step u adds to planes j >= P-1-u.
    auto step_any = [&](int st, auto uc, auto pc) {
      constexpr int u = decltype(uc)::value;  // st % P
      constexpr int JMIN = decltype(pc)::value ? P - 1 - u : 0;
      constexpr int B = (u / G) & 1;          // this group's buffer
      const std::integral_constant<int, B> bc{};
      const int *spb = B ? s_pb1 : s_pb0;
      const T *sw = B ? s_w1 : s_w0;
      const T *sf = B ? s_f1 : s_f0;
      // the wave runs as many rounds as its busiest lane has particles, each
      // lane taking its own next covering particle (ascending)
      auto run = [&](Mask m) {
        auto take = [&]() {
          const int k = sizeof(Mask) == 8 ? __builtin_ctzll(m) : __builtin_ctz(unsigned(m));
          m &= m - 1;
          return k;
        };
        auto wxy_of = [&](int k) {
          const int pb = spb[k];
          int dx = x - (pb & 1023);
          dx += dx < 0 ? K : 0;
          int dy = y - ((pb >> 10) & 1023);
          dy += dy < 0 ? K : 0;
          const T *w = sw + k * NW;
          return w[dx] * w[P + dy];
        };
        if (!m) return;
        int k = take();
        T wxy = wxy_of(k);
        while (true) {
          const int kc = k;
          const T wc = wxy;
          const bool more = m != 0;
          if (more) {  // the next particle's base and xy weight, ahead of this one's sums
            k = take();
            wxy = wxy_of(k);
          }
          const T *w = sw + kc * NW;
          T wf[CC];
#pragma unroll
          for (int c = 0; c < CC; ++c) wf[c] = wc * sf[kc * CC + c];
#pragma unroll
          for (int j = JMIN; j < P; ++j) {
            const T wz = w[2 * P + j];
#pragma unroll
            for (int c = 0; c < CC; ++c) acc[(u + j) % P][c] += wz * wf[c];
          }
          if (!more) break;
        }
      };
      if constexpr (u % G == 0) {
        // the copied group becomes current (its copy, issued a group ago,
        // is waited for); copy the next one into the other buffer, which
        // the group before the current one has finished reading
        wait_vm_all();
        cok = nok;
        const int ctot = ntot;
#pragma unroll
        for (int t = 0; t < NTL; ++t) cdel[t] = ndel[t];
        nok = plan(st + G, nlo, ncnt, ndel, ntot);
        if (nok) stage(std::integral_constant<int, 1 - B>{}, nlo, ncnt, 0, ntot);
        if (cok) cmask = covers(bc, ctot);
      }
      int lo[NTL], cnt[NTL], n = 0;
      Mask sb = 0;  // this step's particles in the staged group
#pragma unroll
      for (int t = 0; t < NTL; ++t) {
        const int l = t * nsteps + st;
        cnt[t] = n_of(l);
        lo[t] = lo_of(l);
        n += cnt[t];
        if (cok && cnt[t] > 0) {  // offset + count <= MAXB
          const Mask ones = cnt[t] >= MAXB ? ~Mask(0) : (ONE << cnt[t]) - 1;
          sb |= ones << (cdel[t] + lo[t]);
        }
      }
      const int nbatch = cok ? 1 : (n + MAXB - 1) / MAXB;
      for (int b = 0; b < nbatch; ++b) {
        Mask m;
        if (cok) {
          m = cmask & sb;
        } else {
          const int nb = min(MAXB, n - b * MAXB);
          stage(bc, lo, cnt, b * MAXB, nb);
          wait_vm_all();
          m = covers(bc, nb);
        }
        run(m);
      }
      if (st >= P - 1) {
        T *g = grid + ((size_t)(zs + st) * K + y) * RS + x + (size_t)cc0 * comp_stride;
#pragma unroll
        for (int c = 0; c < CC; ++c) g[(size_t)c * comp_stride] = acc[u][c];
      }
#pragma unroll
      for (int c = 0; c < CC; ++c) acc[u][c] = T(0);
    };
    auto step = [&](int st, auto uc) { step_any(st, uc, std::false_type{}); };
    auto step_pro = [&](int st, auto uc) { step_any(st, uc, std::true_type{}); };
    unroll_steps<0, P>(0, nsteps, step_pro);
    for (int st0 = P; st0 < nsteps; st0 += P) unroll_steps<0, P>(st0, nsteps, step);
  }
  wait_vm_all();
}

// ---------------------------------------------------------------------------
// Interpolation (old-version/spme_kernel.c:8-70):
//   out[r*ldout + 3*i + d] = beta*out + alpha * sum_j q_i(j) grid[3r+d](j).
// A block owns a TXY x TXY tile of stencil bases, a z range and a chunk of
// CC components (CC*sizeof(T) = 16 bytes: 4 FP32 or 2 FP64).  The tile and
// its (P-1)-wide halo are staged in LDS with the CC components of a point
// side by side, so one 16-byte LDS read gives a lane all its components.
// The z range is swept in phases of S steps: a ring of S + P - 1 = 16
// planes holds what the particles of one phase need, and the S new planes and the
// particles of the next phase are loaded into registers while the current
// phase is processed (two barriers per phase).  Each particle is handled by
// one 16-lane row of a wave (four particles per wave), a lane summing one
// stencil column over P/2 stencil rows; the row's partial sums are reduced
// for all CC components at once and written without atomics.  Particles of a phase are the contiguous slot ranges of the
// 8 x 8 sort tiles inside the block tile.

// Wave-wide sums of CC values per lane (DPP lane moves and two cross-row
// shuffles).  Returns, in every lane, the total of component (lane % CC).
// dpp_ctrl: quad_perm [1,0,3,2] = 0xB1, [2,3,0,1] = 0x4E, row_ror:4 =
// 0x124, row_ror:8 = 0x128.
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, int(b), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}

template <typename T>
__device__ __forceinline__ T sum_same_residue4(T c) {
  // lanes i, i^4, i^8, ... (all lanes with equal i % 4)
  c += dpp_mov<0x124>(c);
  c += dpp_mov<0x128>(c);
  c += __shfl_xor(c, 16);
  c += __shfl_xor(c, 32);
  return c;
}

template <typename T, int CC>
__device__ __forceinline__ T reduce_components(const T (&v)[CC], int lane) {
  static_assert(CC == 4 || CC == 2, "4 FP32 or 2 FP64 components");
  const bool o = lane & 1;
  if constexpr (CC == 4) {
    // round 1 (partner lane ^ 1): even lanes keep components 0, 2, odd 1, 3
    const T a = (o ? v[1] : v[0]) + dpp_mov<0xB1>(o ? v[0] : v[1]);
    const T b = (o ? v[3] : v[2]) + dpp_mov<0xB1>(o ? v[2] : v[3]);
    // round 2 (partner lane ^ 2): component lane % 4
    const bool h = lane & 2;
    const T c = (h ? b : a) + dpp_mov<0x4E>(h ? a : b);
    return sum_same_residue4(c);
  } else {
    T c = (o ? v[1] : v[0]) + dpp_mov<0xB1>(o ? v[0] : v[1]);
    c += dpp_mov<0x4E>(c);  // lanes with equal parity of the quad
    return sum_same_residue4(c);
  }
}

// The same over the 16 lanes of a DPP row only: every lane gets the total
// of component (lane % 4) over its own row.
template <typename T, int CC>
__device__ __forceinline__ T reduce_components_row(const T (&v)[CC], int lane) {
  static_assert(CC == 4 || CC == 2, "4 FP32 or 2 FP64 components");
  const bool o = lane & 1;
  T c;
  if constexpr (CC == 4) {
    const T a = (o ? v[1] : v[0]) + dpp_mov<0xB1>(o ? v[0] : v[1]);
    const T b = (o ? v[3] : v[2]) + dpp_mov<0xB1>(o ? v[2] : v[3]);
    const bool h = lane & 2;
    c = (h ? b : a) + dpp_mov<0x4E>(h ? a : b);
  } else {
    c = (o ? v[1] : v[0]) + dpp_mov<0xB1>(o ? v[0] : v[1]);
    c += dpp_mov<0x4E>(c);
  }
  c += dpp_mov<0x124>(c);
  c += dpp_mov<0x128>(c);
  return c;
}

template <typename T> struct alignas(16) Vec16 {
  static constexpr int N = 16 / sizeof(T);
  T v[N];
};

// Ring of planes (a power of two, so a plane's slot is z & (NSLOT-1)),
// steps per phase (the ring holds a phase's S + P - 1 planes) and particles
// staged per batch.
template <int P> struct InterpPhase {
  static constexpr int NSLOT = 16;
  static constexpr int S = NSLOT - P + 1;
  static constexpr int MAXB = 128;
};

// TO is the output type: the stencil sum is formed in T, the final
// beta*out + alpha*sum in TO (double for the FP32 form with an FP64 sum).
template <typename T, int P, int CC, int TXY, int NT, typename TO = T>
__device__ __forceinline__ void interp_block(const int *__restrict__ boff,
                                             const int *__restrict__ pbase,
                                             const T *__restrict__ pw,
                                             const int *__restrict__ perm,
                                             const T *__restrict__ grid, long comp_stride,
                                             int RS, int K, int C, int ZC, TO alpha, TO beta,
                                             TO *__restrict__ out, long ldout, int lb) {
  static_assert(CC * sizeof(T) == 16, "16-byte component groups");
  using V = Vec16<T>;
  constexpr int S = InterpPhase<P>::S, MAXB = InterpPhase<P>::MAXB;
  constexpr int NSLOT = InterpPhase<P>::NSLOT;
  static_assert((NSLOT & (NSLOT - 1)) == 0 && S + P - 1 == NSLOT, "ring of planes");
  constexpr int E = TXY + P - 1;        // tile + halo points per side
  constexpr int NSUB = TXY / 8;         // sort tiles are 8 x 8
  constexpr int NS2 = NSUB * NSUB;
  constexpr int NWV = NT / 64;
  constexpr int NW = 3 * P;
  // Window row pitch, in 16-byte points: a wave's 16-lane phase of a b128
  // read covers two stencil rows, which fall in disjoint halves of the 64
  // banks when the pitch is 8 mod 16.  Padded where it still fits in LDS.
  constexpr int EP_PAD = E + ((8 - E) % 16 + 16) % 16;
  constexpr bool PAD_FITS =
      (size_t)NSLOT * E * EP_PAD * 16 + (size_t)MAXB * NW * sizeof(T) + 2 * MAXB * 4 <=
      160 * 1024;
  constexpr int EP = PAD_FITS ? EP_PAD : E;
  constexpr int NPTS = E * E;           // points of a plane
  constexpr int PLANE = E * EP;         // window slot stride
  constexpr int NPT = (S * NPTS + NT - 1) / NT;    // plane points per thread per phase
  constexpr int RW = (MAXB * NW + NT - 1) / NT;    // weights per thread per batch
  __shared__ V win[NSLOT * PLANE];
  __shared__ int s_pb[MAXB], s_perm[MAXB];
  __shared__ __attribute__((aligned(16))) T s_w[MAXB * NW];

  const int nt = K / TXY, nz = (K + ZC - 1) / ZC;
  const int tile = lb % (nt * nt), rest = lb / (nt * nt);
  const int tx = tile % nt, ty = tile / nt;
  const int x0 = tx * TXY, y0 = ty * TXY;
  const int z0 = (rest % nz) * ZC, zend = min(z0 + ZC, K);
  const int cc0 = (rest / nz) * CC;
  const int ncc = min(CC, C - cc0);  // valid components of this chunk
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const T *g0 = grid + (size_t)cc0 * comp_stride;
  const long zstride = (long)K * RS;
  // key of sort tile q at base z: ((sy*(K/8) + sx)*K + z)
  int key0[NS2];
#pragma unroll
  for (int q = 0; q < NS2; ++q) {
    const int sx = tx * NSUB + q % NSUB, sy = ty * NSUB + q / NSUB;
    key0[q] = (sy * (K / 8) + sx) * K;
  }

  // plane point p of the tile (p < NPTS): source offset in a mesh plane
  // and position in a window slot
  auto src_of = [&](int p) {
    const int ey = p / E, ex = p - ey * E;
    int yy = y0 + ey, xx = x0 + ex;
    yy -= yy >= K ? K : 0;
    xx -= xx >= K ? K : 0;
    return (long)yy * RS + xx;
  };
  auto win_of = [&](int p) {
    const int ey = p / E;
    return ey * EP + (p - ey * E);
  };
  auto load_point = [&](int z, int p) {  // z unwrapped
    V v;
    const T *g = g0 + (z >= K ? z - K : z) * zstride + src_of(p);
#pragma unroll
    for (int c = 0; c < CC; ++c) v.v[c] = c < ncc ? g[(size_t)c * comp_stride] : T(0);
    return v;
  };
  // prologue: planes z0 .. z0+P-2
  for (int q = tid; q < (P - 1) * NPTS; q += NT) {
    const int j = q / NPTS, p = q - j * NPTS;
    win[((z0 + j) & (NSLOT - 1)) * PLANE + win_of(p)] = load_point(z0 + j, p);
  }
  // the phase loads of this thread: plane j_of[r] of the phase, point with
  // in-plane offset src_off[r] (mesh) and pt_of[r] (window); fixed over the
  // sweep, so computed once
  long src_off[NPT];
  int pt_of[NPT], j_of[NPT];
#pragma unroll
  for (int r = 0; r < NPT; ++r) {
    const int q = tid + NT * r;
    const int j = q / NPTS, p = q - j * NPTS;
    j_of[r] = q < S * NPTS ? j : S;  // S: no load
    pt_of[r] = win_of(p);
    src_off[r] = src_of(p);
  }

  // registers carrying the next phase's loads
  V vp[NPT];
  T vw[RW];
  int vpb = 0, vpm = 0;
  int lo[NS2], cnt[NS2];
  int n_all = 0;
  // phase starting at step zs: the new planes zs+P-1 .. zs+P-1+S-1 (those
  // below zend+P-1) and the particle ranges; loads of batch b0
  auto ranges = [&](int zs) {
    const int ze = min(zs + S, zend);
    n_all = 0;
#pragma unroll
    for (int q = 0; q < NS2; ++q) {
      lo[q] = boff[key0[q] + zs];
      cnt[q] = boff[key0[q] + ze] - lo[q];
      n_all += cnt[q];
    }
  };
  auto locate = [&](int k) {  // k-th particle of the phase -> sorted slot
    int i = lo[0] + k, rem = k - cnt[0];
#pragma unroll
    for (int q = 1; q < NS2; ++q) {
      i = rem >= 0 ? lo[q] + rem : i;
      rem -= cnt[q];
    }
    return i;
  };
  auto fetch_planes = [&](int zs) {
    const int zfirst = zs + P - 1, zlast = min(zs + S, zend) + P - 1;  // [zfirst, zlast)
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
      const int z = zfirst + j_of[r];
      if (z < zlast) {
        const T *g = g0 + (z >= K ? z - K : z) * zstride + src_off[r];
#pragma unroll
        for (int c = 0; c < CC; ++c) vp[r].v[c] = c < ncc ? g[(size_t)c * comp_stride] : T(0);
      }
    }
  };
  auto commit_planes = [&](int zs) {
    const int zfirst = zs + P - 1, zlast = min(zs + S, zend) + P - 1;
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
      const int z = zfirst + j_of[r];
      if (z < zlast) win[(z & (NSLOT - 1)) * PLANE + pt_of[r]] = vp[r];
    }
  };
  auto fetch_particles = [&](int b0) {
    const int nb = max(0, min(MAXB, n_all - b0));
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
    return nb;
  };
  auto commit_particles = [&]() {
#pragma unroll
    for (int r = 0; r < RW; ++r)
      if (tid + NT * r < MAXB * NW) s_w[tid + NT * r] = vw[r];
    if (tid < MAXB) {
      s_pb[tid] = vpb;
      s_perm[tid] = vpm;
    }
  };

  // Four particles per wave, one per 16-lane DPP row: lane = (row g, h,
  // dx); the lane sums stencil column dx of stencil rows 2i + h, i < P/2
  // (its two rows of a 16-lane LDS phase are adjacent: disjoint bank
  // halves), and the row's 16 lanes are reduced with DPP moves only.
  constexpr int NR = P / 2;
  const int g = lane >> 4, h = (lane >> 3) & 1, dx = lane & 7;
  const bool act = dx < P;
  const int dxc = min(dx, P - 1);
  auto process = [&](int nb) {
    for (int k0 = 4 * wave; k0 < nb; k0 += 4 * NWV) {
      const int k = k0 + g;
      const bool kon = k < nb;
      const int kk = kon ? k : k0;  // a staged particle (loads stay in range)
      const int pb = s_pb[kk];
      const int rx = (pb & 1023) - x0, ry = ((pb >> 10) & 1023) - y0;
      const int bz = pb >> 20;  // in [z0, zend)
      const T *w = s_w + kk * NW;
      const T wx = act ? w[dxc] : T(0);
      // window points of plane j of this lane's column (first row h)
      const int sl0 = bz & (NSLOT - 1);
      const V *pt = win + (ry + h) * EP + rx + dxc;
      const V *pj[P];
#pragma unroll
      for (int j = 0; j < P; ++j) pj[j] = pt + ((sl0 + j) & (NSLOT - 1)) * PLANE;
      T acc[CC];
#pragma unroll
      for (int c = 0; c < CC; ++c) acc[c] = T(0);
#pragma unroll
      for (int i = 0; i < NR; ++i) {
        T ar[CC];
#pragma unroll
        for (int c = 0; c < CC; ++c) ar[c] = T(0);
#pragma unroll
        for (int j = 0; j < P; ++j) {
          const T wz = w[2 * P + j];
          const V v = pj[j][2 * i * EP];
#pragma unroll
          for (int c = 0; c < CC; ++c) ar[c] += wz * v.v[c];
        }
        const T wy = w[P + 2 * i + h];
#pragma unroll
        for (int c = 0; c < CC; ++c) acc[c] += wy * ar[c];
      }
#pragma unroll
      for (int c = 0; c < CC; ++c) acc[c] *= wx;
      const T tot = reduce_components_row<T, CC>(acc, lane);
      const int cl = lane & 15;
      if (kon && cl < ncc) {
        const int comp = cc0 + cl, r = comp / 3, d = comp - 3 * r;
        TO *u = out + r * ldout + 3L * s_perm[k] + d;
        *u = beta == TO(0) ? alpha * TO(tot) : beta * *u + alpha * TO(tot);
      }
    }
  };

  ranges(z0);
  fetch_planes(z0);
  int nb = fetch_particles(0);
  for (int zs = z0; zs < zend; zs += S) {
    const int n_phase = n_all;
    commit_planes(zs);
    commit_particles();
    __syncthreads();
    const int zn = zs + S;
    int nb_next = 0;
    if (n_phase <= MAXB && zn < zend) {
      ranges(zn);
      fetch_planes(zn);
      nb_next = fetch_particles(0);
    }
    process(nb);
    if (n_phase > MAXB) {  // rare: the phase holds more particles than a batch
      for (int b0 = MAXB; b0 < n_phase; b0 += MAXB) {
        __syncthreads();
        const int nbb = fetch_particles(b0);
        commit_particles();
        __syncthreads();
        process(nbb);
      }
      if (zn < zend) {
        ranges(zn);
        fetch_planes(zn);
        nb_next = fetch_particles(0);
      }
    }
    __syncthreads();
    nb = nb_next;
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
__device__ __forceinline__ void fft_x_r2c_block(T *__restrict__ grid, int RS, long npairs,
                                                const cpx<T> *__restrict__ tw, int lb,
                                                cpx<T> *buf, cpx<T> *s_tw) {
  constexpr int N1 = Split<K>::N1, N2 = Split<K>::N2;
  constexpr int Kh = K / 2 + 1;
  constexpr int LS = K + 1;
  load_tw<T, K>(s_tw, tw);
  const long q0 = (long)lb * NL;
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
__device__ __forceinline__ void fft_x_c2r_block(T *__restrict__ grid, int RS, long npairs,
                                                const cpx<T> *__restrict__ tw, int lb,
                                                cpx<T> *buf, cpx<T> *s_tw) {
  constexpr int N1 = Split<K>::N1, N2 = Split<K>::N2;
  constexpr int Kh = K / 2 + 1;
  constexpr int LS = K + 1;
  load_tw<T, K>(s_tw, tw);
  const long q0 = (long)lb * NL;
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

// Complex FFT along y (ZAX false: block = (x chunk of XW, plane z, comp)) or
// along z (ZAX true: block = (x chunk of XW, row y, comp)), in place.
template <typename T, int K, int XW, int S, int NT, bool ZAX = false>
__device__ __forceinline__ void fft_y_block(T *__restrict__ grid, long comp_stride, int Khp,
                                            int nxc, const cpx<T> *__restrict__ tw, int lb,
                                            cpx<T> *buf, cpx<T> *s_tw) {
  constexpr int N1 = Split<K>::N1, N2 = Split<K>::N2;
  constexpr int Kh = K / 2 + 1;
  constexpr int LS = K + 1;
  load_tw<T, K>(s_tw, tw);
  __syncthreads();
  const int xc = lb % nxc;
  const int zc = lb / nxc;  // (z, or y along z) + K*comp
  const int comp = zc / K, z = zc - comp * K;
  const int x0 = xc * XW;
  cpx<T> *plane = reinterpret_cast<cpx<T> *>(grid + comp * comp_stride) +
                  (ZAX ? (size_t)z * Khp : (size_t)z * K * Khp);
  const size_t cstr = ZAX ? (size_t)K * Khp : (size_t)Khp;  // stride of the transformed index
  for (int t = threadIdx.x; t < XW * N2; t += NT) {  // step A from global
    const int xo = t % XW, n2 = t / XW;
    const int x = x0 + xo;
    cpx<T> a[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1)
      a[n1] = x < Kh ? plane[(size_t)(n1 * N2 + n2) * cstr + x] : cpx<T>{T(0), T(0)};
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
      for (int k2 = 0; k2 < N2; ++k2) plane[(size_t)(k1 + N1 * k2) * cstr + x] = b[k2];
    }
  }
}

// Forward FFT along z, RPY influence (old-version/spme_kernel.c:73-228),
// backward FFT along z, for the 3 components of one right-hand side.
// Blocks follow ZLayout: per right-hand side, (x chunk, y) blocks of XW
// adjacent columns, then (when K/2 is a multiple of XW) Nyquist blocks whose
// XW lanes take column x = K/2 of XW consecutive rows y.
// m2(k) = psi(|k|)|b(k)|^2/V is formed from FP64-built factors:
// tab[n] = psi(dk sqrt(n))/V for n = nx^2 + ny^2 + nz^2 (integer wave
// numbers) and tab[NPSI + m] = |b(m)|^2, so the z pass reads a table of
// O(K^2) entries that stays in cache instead of one of K^2 (K/2 + 1).
// Wave numbers follow old-version/spme.c:95-107 (index K/2 -> +pi K/L).
// The spectrum of a (column, k1) stays in registers between the forward
// N2-point FFT, the 3x3 influence and the backward N2-point FFT.
// nyq: the Nyquist column as [component][z][y] (its blocks then read 16
// consecutive y per z), or null when it sits in the spectrum rows.
// POLY (unequal radii): a block takes the 6 components of a right-hand side,
// the spectra S_f (components 0-2) and S_{a^2 f} (3-5), and tab holds the
// factors of A(k) = psi(k) without its radius term, so that
//   out_f = A|b|^2 P (S_f - k^2/6 S_{a^2 f}),   out_a = (k^2/6) A|b|^2 P S_f
// (P = I - kk^T/k^2); interpolation then forms I[out_f] - a_i^2 I[out_a].
template <typename T, int K, int XW, int NT, bool POLY = false>
__device__ __forceinline__ void fft_z_infl_block(T *__restrict__ grid, long comp_stride,
                                                 int Khp, const T *__restrict__ tab, T dk,
                                                 const cpx<T> *__restrict__ tw, int lb,
                                                 cpx<T> *buf, cpx<T> *s_tw,
                                                 cpx<T> *__restrict__ nyq) {
  constexpr int N1 = Split<K>::N1, N2 = Split<K>::N2;
  constexpr int Kh = K / 2 + 1;
  constexpr int LS = K + 1;
  constexpr int NCP = POLY ? 6 : 3;  // components per block
  using Z = ZLayout<XW>;
  static_assert(XW * N1 <= NT, "one step-B column per thread");
  constexpr int NXC = Z::chunks(K), PER = Z::per_rhs(K);
  // with narrow Nyquist blocks every lane's column has x <= K/2 < Kh
  constexpr bool ALL_X = Z::nyq(K);
  load_tw<T, K>(s_tw, tw);
  __syncthreads();  // step A reads twiddles that other threads loaded
  const int rhs = lb / PER, j = lb - rhs * PER;
  // a lane's column (x, y) is (xa + xo*dx, ya + xo*dy)
  int xa, ya, dx, dy;
  if (j < NXC * K) {  // adjacent x chunks (halves of one cache line) on the same XCD
    xa = (j % NXC) * XW, ya = j / NXC, dx = 1, dy = 0;
  } else {
    xa = K / 2, ya = (j - NXC * K) * XW, dx = 0, dy = 1;
  }
  cpx<T> *g0 = reinterpret_cast<cpx<T> *>(grid + ((long)NCP * rhs) * comp_stride);
  const long cs = comp_stride / 2;  // complex elements
  // this block's column values in the separate Nyquist array (block-uniform)
  cpx<T> *nq = (nyq && j >= NXC * K) ? nyq + (long)NCP * rhs * K * K : nullptr;
  constexpr long NQC = (long)K * K;  // Nyquist values per component
  for (int t = threadIdx.x; t < NCP * XW * N2; t += NT) {  // step A (forward) from global
    const int xo = t % XW, r = t / XW;
    const int c = r % NCP, n2 = r / NCP;
    const int x = xa + xo * dx, y = ya + xo * dy;
    cpx<T> a[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1)
      a[n1] = nq ? nq[c * NQC + (long)(n1 * N2 + n2) * K + y]
                 : (ALL_X || x < Kh ? g0[c * cs + ((size_t)(n1 * N2 + n2) * K + y) * Khp + x]
                                    : cpx<T>{T(0), T(0)});
    reg_fft<N1, -1>(a);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) a[k1] = cmul(a[k1], twiddle<-1>(s_tw, n2 * k1));
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) buf[(c * XW + xo) * LS + k1 * N2 + n2] = a[k1];
  }
  __syncthreads();
  if (threadIdx.x < XW * N1) {  // step B, influence, inverse step B
    const int xo = threadIdx.x % XW, k1 = threadIdx.x / XW;
    const int x = xa + xo * dx, y = ya + xo * dy;
    // k = 0 (tab[0] = 0) and padding columns (x > K/2) carry m2 = 0
    constexpr int NPSI = influence_npsi(K);
    const T *bsq = tab + NPSI;
    const int nx = min(x, K / 2), ny = y <= K / 2 ? y : K - y;
    const int nxy = nx * nx + ny * ny;
    const T bxy = x < Kh ? bsq[nx] * bsq[y] : T(0);
    T m2[N2];
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) {
      const int z = k1 + N1 * k2, nz = z <= K / 2 ? z : K - z;
      m2[k2] = tab[nxy + nz * nz] * (bxy * bsq[z]);
    }
    cpx<T> b[NCP][N2];
#pragma unroll
    for (int c = 0; c < NCP; ++c) {
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) b[c][n2] = buf[(c * XW + xo) * LS + k1 * N2 + n2];
      reg_fft<N2, -1>(b[c]);
    }
    const T kx = T(x) * dk;  // x <= K/2 in the half spectrum
    const T ky = T(y <= K / 2 ? y : y - K) * dk;
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) {
      // every fused multiply-add is written out, so the single-system and
      // grouped kernels round alike whatever their surrounding code
#pragma clang fp contract(off)
      const int z = k1 + N1 * k2;
      const T kz = T(z <= K / 2 ? z : z - K) * dk;
      const T k2n = fma(kz, kz, fma(ky, ky, kx * kx));
      // guard the division at k = 0
      const T mi = k2n > T(0) ? m2[k2] / k2n : T(0);
      const T ax = mi * kx, ay = mi * ky, az = mi * kz;
      const T B0 = fma(-ax, kx, m2[k2]), B1 = -ax * ky, B2 = -ax * kz;
      const T B3 = fma(-ay, ky, m2[k2]), B4 = -ay * kz, B5 = fma(-az, kz, m2[k2]);
      cpx<T> va = b[0][k2], vb = b[1][k2], vc = b[2][k2];  // S_f
      if constexpr (POLY) {
        // out_a = (k^2/6) B S_f into components 3-5; out_f = B (S_f - k^2/6
        // S_{a^2 f}) below, from the combination formed in va, vb, vc
        const T q = k2n * T(1.0 / 6.0);
        const cpx<T> sa = b[3][k2], sb = b[4][k2], sc = b[5][k2];  // S_{a^2 f}
        b[3][k2] = {q * fma(B2, vc.x, fma(B1, vb.x, B0 * va.x)),
                    q * fma(B2, vc.y, fma(B1, vb.y, B0 * va.y))};
        b[4][k2] = {q * fma(B4, vc.x, fma(B3, vb.x, B1 * va.x)),
                    q * fma(B4, vc.y, fma(B3, vb.y, B1 * va.y))};
        b[5][k2] = {q * fma(B5, vc.x, fma(B4, vb.x, B2 * va.x)),
                    q * fma(B5, vc.y, fma(B4, vb.y, B2 * va.y))};
        va = {fma(-q, sa.x, va.x), fma(-q, sa.y, va.y)};
        vb = {fma(-q, sb.x, vb.x), fma(-q, sb.y, vb.y)};
        vc = {fma(-q, sc.x, vc.x), fma(-q, sc.y, vc.y)};
      }
      b[0][k2] = {fma(B2, vc.x, fma(B1, vb.x, B0 * va.x)),
                  fma(B2, vc.y, fma(B1, vb.y, B0 * va.y))};
      b[1][k2] = {fma(B4, vc.x, fma(B3, vb.x, B1 * va.x)),
                  fma(B4, vc.y, fma(B3, vb.y, B1 * va.y))};
      b[2][k2] = {fma(B5, vc.x, fma(B4, vb.x, B2 * va.x)),
                  fma(B5, vc.y, fma(B4, vb.y, B2 * va.y))};
    }
#pragma unroll
    for (int c = 0; c < NCP; ++c) {
      reg_fft<N2, +1>(b[c]);
#pragma unroll
      for (int n2 = 1; n2 < N2; ++n2) b[c][n2] = cmul(b[c][n2], twiddle<+1>(s_tw, n2 * k1));
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) buf[(c * XW + xo) * LS + k1 * N2 + n2] = b[c][n2];
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < NCP * XW * N2; t += NT) {  // inverse step A to global
    const int xo = t % XW, r = t / XW;
    const int c = r % NCP, n2 = r / NCP;
    const int x = xa + xo * dx, y = ya + xo * dy;
    cpx<T> a[N1];
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) a[k1] = buf[(c * XW + xo) * LS + k1 * N2 + n2];
    reg_fft<N1, +1>(a);
    if (nq) {
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1)
        st_stream(nq + c * NQC + (long)(n2 + N2 * n1) * K + y, a[n1]);
    } else if (ALL_X || x < Kh) {
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1)
        st_stream(g0 + c * cs + ((size_t)(n2 + N2 * n1) * K + y) * Khp + x, a[n1]);
    }
  }
}

// Real-space SpMV: a row (cell order) is handled by LPR lanes, each with RPL
// consecutive right-hand sides, so the pair's coefficients and column index
// are loaded once per RPL right-hand sides.  fs: forces in cell order with
// the right-hand sides of one component together, fs[s*C + d*R + r] (R =
// C/3), so that for each d the LPR lanes of a row read one contiguous
// segment of R values.  For right-hand side r and d < 3:
//   out[r*ldout + 3*perm[s] + d] = (accum ? out : beta*out)
//                                  + alpha*(self f + sum_j B_sj f_j),
// B_sj = F I + (G/r^2) e e^T with e the minimum-image separation, formed
// from the stored FP64-built pair (F, G/r^2) (interleaved, one 2-value load
// per entry) and the positions.  FP64 engines take e from the FP64
// positions.  FP32 engines take it from cpos: per particle its offset from
// the corner of its cell, in FP32 (at most one cell side, so within about
// 1e-6 of the FP64 value), and the cell's coordinates; e = offset_j -
// offset_i + dc*h with dc the wrapped cell difference (-1, 0 or 1: a stored
// neighbour lies in one of the 27 cells around the row's, and there are at
// least 3 cells per side).  With unequal radii the stored F and G carry the
// pair's radii and the self coefficient is per row (selfv, cell order;
// null: the scalar self).
// Rows are packed in slices of 64 (entry-major within a slice); the rows of
// a wave lie in one slice, so every lane runs the same entry count.
template <int C> struct RealShape {
  static constexpr int RPL = real_rpl(C);  // right-hand sides per lane
  static constexpr int LPR = real_lpr(C);  // lanes per row
};

// N values from p into v; 16-byte (8-byte) loads when N values fill whole
// 16-byte (8-byte) units (p is then so aligned at every call site).
template <typename T, int N>
__device__ __forceinline__ void load_n(const T *__restrict__ p, T (&v)[N]) {
  if constexpr (N * sizeof(T) % 16 == 0) {
    constexpr int W = 16 / sizeof(T);
    struct alignas(16) U { T v[W]; };
#pragma unroll
    for (int i = 0; i < N / W; ++i) {
      const U t = reinterpret_cast<const U *>(p)[i];
#pragma unroll
      for (int k = 0; k < W; ++k) v[W * i + k] = t.v[k];
    }
  } else if constexpr (sizeof(T) == 4 && N % 2 == 0) {
    struct alignas(8) U { T v[2]; };
#pragma unroll
    for (int i = 0; i < N / 2; ++i) {
      const U t = reinterpret_cast<const U *>(p)[i];
      v[2 * i] = t.v[0];
      v[2 * i + 1] = t.v[1];
    }
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = p[i];
  }
}

// Cell coordinates packed into cpos[s].w: cx | cy << 10 | cz << 20.
struct CellPos {
  float4 v;
  __device__ __forceinline__ int c(int d) const {
    return (__float_as_int(v.w) >> (10 * d)) & 1023;
  }
};

// Wrapped difference of two cell coordinates of neighbouring cells.
__device__ __forceinline__ int cell_step(int cj, int ci, int nc) {
  const int d = cj - ci;
  return d > 1 ? d - nc : (d < -1 ? d + nc : d);
}

// TO: output type (see interp_block); the pair sums are formed in T.
template <typename T, int C, typename TO = T>
__device__ __forceinline__ void real_lane(const int *__restrict__ perm,
                                          const int *__restrict__ width,
                                          const int *__restrict__ soff,
                                          const int *__restrict__ col,
                                          const T *__restrict__ blk,
                                          const double *__restrict__ spos,
                                          const float4 *__restrict__ cpos, int nc, float h,
                                          double L, T self, const T *__restrict__ selfv,
                                          const T *__restrict__ fs, TO alpha, TO beta, int accum,
                                          TO *__restrict__ out, long ldout, int s, int l) {
  constexpr int RPL = RealShape<C>::RPL, R = C / 3;
  constexpr bool F32 = sizeof(T) == 4;
  const int q = s >> 6, ls = s & 63;
  const int r0 = l * RPL;
  T a[3][RPL], f[3][RPL];
#pragma unroll
  for (int d = 0; d < 3; ++d) load_n<T, RPL>(fs + (size_t)s * C + d * R + r0, f[d]);
  // self term: one coefficient, or one per row (unequal radii)
  const T sf = selfv ? selfv[s] : self;
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int r = 0; r < RPL; ++r) a[d][r] = sf * f[d][r];
  double px = 0.0, py = 0.0, pz = 0.0;
  CellPos ci{};
  if constexpr (F32) {
    ci.v = cpos[s];
  } else {
    px = spos[3L * s];
    py = spos[3L * s + 1];
    pz = spos[3L * s + 2];
  }
  const double iL = 1.0 / L;
  const long base = soff[q] + ls;
  const int w = width[q];
  struct alignas(2 * sizeof(T)) Pair { T F, g; };
  const Pair *pairs = reinterpret_cast<const Pair *>(blk) + base;
  const int *cols = col + base;
  // U neighbours per step.  The column indices and coefficients of the next
  // step are loaded while the current step's gathers are in flight, and all
  // U gathers are issued before any of them is used.  Entries past the row's
  // width repeat its last entry with zero coefficients (they add exact
  // zeros), so no load leaves the slice.
  constexpr int U = 6;
  int jn[U];
  Pair gn[U];
  auto fetch_entries = [&](int k0) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long k = min(k0 + u, w - 1);
      jn[u] = cols[64L * k];
      gn[u] = pairs[64L * k];
    }
  };
  if (w > 0) fetch_entries(0);
  for (int k0 = 0; k0 < w; k0 += U) {
    // every fused multiply-add is written out, so that the instantiations
    // for different batch widths (right-hand sides per lane) round alike
#pragma clang fp contract(off)
    int j[U];
    T F[U], g[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      j[u] = jn[u];
      const bool on = k0 + u < w;
      F[u] = on ? gn[u].F : T(0);
      g[u] = on ? gn[u].g : T(0);
    }
    T fv[U][3][RPL];
    T x[U], y[U], z[U];
    if constexpr (F32) {
      CellPos cj[U];
#pragma unroll
      for (int u = 0; u < U; ++u) cj[u].v = cpos[j[u]];
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int d = 0; d < 3; ++d) load_n<T, RPL>(fs + (size_t)j[u] * C + d * R + r0, fv[u][d]);
      if (k0 + U < w) fetch_entries(k0 + U);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        x[u] = (cj[u].v.x - ci.v.x) + float(cell_step(cj[u].c(0), ci.c(0), nc)) * h;
        y[u] = (cj[u].v.y - ci.v.y) + float(cell_step(cj[u].c(1), ci.c(1), nc)) * h;
        z[u] = (cj[u].v.z - ci.v.z) + float(cell_step(cj[u].c(2), ci.c(2), nc)) * h;
      }
    } else {
      double ex[U], ey[U], ez[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        ex[u] = spos[3L * j[u]];
        ey[u] = spos[3L * j[u] + 1];
        ez[u] = spos[3L * j[u] + 2];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int d = 0; d < 3; ++d) load_n<T, RPL>(fs + (size_t)j[u] * C + d * R + r0, fv[u][d]);
      if (k0 + U < w) fetch_entries(k0 + U);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        double dx = ex[u] - px, dy = ey[u] - py, dz = ez[u] - pz;
        dx = fma(-L, rint(dx * iL), dx);
        dy = fma(-L, rint(dy * iL), dy);
        dz = fma(-L, rint(dz * iL), dz);
        x[u] = T(dx);
        y[u] = T(dy);
        z[u] = T(dz);
      }
    }
    // each neighbour's product is formed apart from the running sums and
    // added last, in entry order
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const T gx = g[u] * x[u], gy = g[u] * y[u], gz = g[u] * z[u];
      const T b0 = fma(gx, x[u], F[u]), b1 = gx * y[u], b2 = gx * z[u];
      const T b3 = fma(gy, y[u], F[u]), b4 = gy * z[u], b5 = fma(gz, z[u], F[u]);
#pragma unroll
      for (int r = 0; r < RPL; ++r) {
        const T f0 = fv[u][0][r], f1 = fv[u][1][r], f2 = fv[u][2][r];
        a[0][r] += fma(b2, f2, fma(b1, f1, b0 * f0));
        a[1][r] += fma(b4, f2, fma(b3, f1, b1 * f0));
        a[2][r] += fma(b5, f2, fma(b4, f1, b2 * f0));
      }
    }
  }
  const long po = 3L * perm[s];
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    TO *u = out + (r0 + r) * ldout + po;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
#pragma clang fp contract(off)
      const TO prev = accum ? u[d] : (beta == TO(0) ? TO(0) : beta * u[d]);
      u[d] = fma(alpha, TO(a[d][r]), prev);
    }
  }
}

// ---------------------------------------------------------------------------
// Fused x + y transforms: one (component, z) plane at a time in LDS, so the
// real mesh is read once and the half spectrum written once per direction
// (the separate x and y passes move the spectrum through HBM twice more).
// LDS holds K/2 line pairs of pitch LP = 2*Kh complex: during the x
// transform line l is the complex row (row 2l) + i (row 2l+1); after the
// split, row y of the half spectrum sits at y*Kh (2l*Kh = l*LP and
// (2l+1)*Kh = l*LP + Kh), i.e. the plane is [y][x] with pitch Kh for the y
// transform.  Blocks loop over planes and load the next plane into
// registers while the second half of the current one is transformed and
// stored.
template <typename T, int K> struct XYShape {
  static constexpr int Kh = K / 2 + 1;
  static constexpr int LP = 2 * Kh;
  static constexpr int NCPX = (K / 2) * LP;
  static constexpr size_t LDS = (NCPX + K) * sizeof(cpx<T>);
  static_assert(LDS == xy_lds<T>(K), "the host's plane size");
};

// Split of packed lines into two half spectra, in place: for line l and
// k <= K/2, A_k = (Z_k + conj Z_{K-k})/2 (row 2l), B_k = (Z_k - conj
// Z_{K-k})/(2i) (row 2l+1).  Writing B_k (position Kh + k) clobbers Z at
// K - (K/2 - 1 - k), so k and K/2 - 1 - k go to one thread.
template <typename T, int K>
__device__ __forceinline__ void xy_split(cpx<T> *line, int k) {
  constexpr int Kh = K / 2 + 1;
  const cpx<T> Z = line[k], Zm = cconj(line[k == 0 ? 0 : K - k]);
  const cpx<T> A = {T(0.5) * (Z.x + Zm.x), T(0.5) * (Z.y + Zm.y)};
  const cpx<T> D = csub(Z, Zm);
  const cpx<T> B = {T(0.5) * D.y, T(-0.5) * D.x};
  line[k] = A;
  line[Kh + k] = B;
}

// Inverse of the split (packing two half spectra into one line of K),
// dropping the imaginary parts at k = 0 and K/2 (real part of the IDFT).
template <typename T, int K>
__device__ __forceinline__ void xy_pack(cpx<T> *line, int k) {
  constexpr int Kh = K / 2 + 1;
  cpx<T> A = line[k], B = line[Kh + k];
  if (k == 0 || k == K / 2) {
    A.y = T(0);
    B.y = T(0);
  }
  line[k] = {A.x - B.y, A.y + B.x};
  if (k > 0 && k < K / 2) line[K - k] = {A.x + B.y, B.x - A.y};
}

// NYQ: the Nyquist column x = K/2 goes to its own array (nyq), which the
// launcher passes exactly when NYQ is set.  The body transforms planes q0,
// q0 + qs, ... below nplanes (buf: S::NCPX values, s_tw: K, in LDS); a
// single-system kernel strides over the planes, a grouped one takes one.
template <typename T, int K, int NT, bool NYQ>
__device__ __forceinline__ void xy_fwd_block(T *__restrict__ grid, int RS, long comp_stride,
                                             int nplanes, const cpx<T> *__restrict__ tw,
                                             cpx<T> *__restrict__ nyq, int q0, int qs,
                                             cpx<T> *buf, cpx<T> *s_tw) {
  using S = XYShape<T, K>;
  constexpr int N1 = Split<K>::N1, N2 = Split<K>::N2;
  constexpr int Kh = S::Kh, LP = S::LP, L2 = K / 2;
  constexpr int G = 16 / sizeof(T);  // reals per 16-byte load
  constexpr int UPL = K / G, NUNIT = L2 * UPL, NU = (NUNIT + NT - 1) / NT;
  using V = Vec16<T>;
  load_tw<T, K>(s_tw, tw);
  int tid = threadIdx.x;
  const int Khp = RS / 2;
  V ra[NU], rb[NU];
  auto fetch = [&](int q) {
    const int comp = q / K, z = q - comp * K;
    const T *base = grid + comp * comp_stride + (size_t)z * K * RS;
#pragma unroll
    for (int r = 0; r < NU; ++r) {
      const bool full = NT * (r + 1) <= NUNIT;  // a constant once unrolled
      const int u = tid + NT * r;
      if (full || u < NUNIT) {
        const int l = u / UPL, g = u - l * UPL;
        ra[r] = *reinterpret_cast<const V *>(base + (size_t)(2 * l) * RS + G * g);
        rb[r] = *reinterpret_cast<const V *>(base + (size_t)(2 * l + 1) * RS + G * g);
      }
    }
  };
  auto commit = [&]() {
#pragma unroll
    for (int r = 0; r < NU; ++r) {
      const bool full = NT * (r + 1) <= NUNIT;
      const int u = tid + NT * r;
      if (full || u < NUNIT) {
        const int l = u / UPL, g = u - l * UPL;
#pragma unroll
        for (int i = 0; i < G; ++i) buf[l * LP + G * g + i] = {ra[r].v[i], rb[r].v[i]};
      }
    }
  };
  int q = q0;
  if (q < nplanes) fetch(q);
  for (; q < nplanes; q += qs) {
    // keep per-thread index arithmetic inside the loop: hoisted, it would
    // stay live across the transform and spill
    asm volatile("" : "+v"(tid));
    commit();
    __syncthreads();
    // the next plane's loads, in flight through this one
    if (q + qs < nplanes) fetch(q + qs);
    // x, step A: per line and n2, an N1-point FFT, twiddled, in place
    for (int t = tid; t < L2 * N2; t += NT) {
      const int l = t / N2, n2 = t - l * N2;
      cpx<T> *ln = buf + l * LP;
      cpx<T> a[N1];
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) a[n1] = ln[n1 * N2 + n2];
      reg_fft<N1, -1>(a);
#pragma unroll
      for (int k1 = 1; k1 < N1; ++k1) a[k1] = cmul(a[k1], twiddle<-1>(s_tw, n2 * k1));
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) ln[k1 * N2 + n2] = a[k1];
    }
    __syncthreads();
    // x, step B, natural order: in rounds of whole lines (a line's tasks
    // read and write the same positions), one register tile per thread
    constexpr int LPR = NT / N1;  // lines per round
    for (int l0 = 0; l0 < L2; l0 += LPR) {
      const int l = l0 + tid / N1, k1 = tid % N1;
      const bool on = tid < LPR * N1 && l < L2;
      cpx<T> b[N2];
      if (on) {
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) b[n2] = buf[l * LP + k1 * N2 + n2];
        reg_fft<N2, -1>(b);
      }
      __syncthreads();
      if (on) {
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) buf[l * LP + k1 + N1 * k2] = b[k2];
      }
      __syncthreads();
    }
    // split into half spectra: k and K/2-1-k per thread, K/2 alone
    for (int t = tid; t < L2 * (K / 4 + 1); t += NT) {
      const int l = t / (K / 4 + 1), m = t - l * (K / 4 + 1);
      cpx<T> *ln = buf + l * LP;
      if (m < K / 4) {
        // both reads before either write
        const int k0 = m, k1 = K / 2 - 1 - m;
        const cpx<T> Z0 = ln[k0], Z0m = cconj(ln[k0 == 0 ? 0 : K - k0]);
        const cpx<T> Z1 = ln[k1], Z1m = cconj(ln[K - k1]);
        const cpx<T> D0 = csub(Z0, Z0m), D1 = csub(Z1, Z1m);
        ln[k0] = {T(0.5) * (Z0.x + Z0m.x), T(0.5) * (Z0.y + Z0m.y)};
        ln[Kh + k0] = {T(0.5) * D0.y, T(-0.5) * D0.x};
        ln[k1] = {T(0.5) * (Z1.x + Z1m.x), T(0.5) * (Z1.y + Z1m.y)};
        ln[Kh + k1] = {T(0.5) * D1.y, T(-0.5) * D1.x};
      } else {
        xy_split<T, K>(ln, K / 2);
      }
    }
    __syncthreads();
    // y, step A over the Kh columns of the [y][x] plane (pitch Kh)
    for (int t = tid; t < Kh * N2; t += NT) {
      const int x = t % Kh, n2 = t / Kh;
      cpx<T> a[N1];
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) a[n1] = buf[(n1 * N2 + n2) * Kh + x];
      reg_fft<N1, -1>(a);
#pragma unroll
      for (int k1 = 1; k1 < N1; ++k1) a[k1] = cmul(a[k1], twiddle<-1>(s_tw, n2 * k1));
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) buf[(k1 * N2 + n2) * Kh + x] = a[k1];
    }
    __syncthreads();
    // y, step B straight to the spectrum rows: the columns that stay in
    // the rows (all Kh, or the K/2 below the Nyquist column), then, with a
    // separate Nyquist array, that column's N1 tasks in the first lanes
    {
      const int comp = q / K, z = q - comp * K;
      cpx<T> *spec = reinterpret_cast<cpx<T> *>(grid + comp * comp_stride) + (size_t)z * K * Khp;
      constexpr int XR = NYQ ? K / 2 : Kh;
      for (int t = tid; t < XR * N1; t += NT) {
        const int x = t % XR, k1 = t / XR;
        cpx<T> bb[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) bb[n2] = buf[(k1 * N2 + n2) * Kh + x];
        reg_fft<N2, -1>(bb);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2)
          st_stream(spec + (size_t)(k1 + N1 * k2) * Khp + x, bb[k2]);
      }
      if (NYQ && tid < N1) {
        cpx<T> *nq = nyq + ((size_t)comp * K + z) * K;
        const int k1 = tid;
        cpx<T> bb[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) bb[n2] = buf[(k1 * N2 + n2) * Kh + K / 2];
        reg_fft<N2, -1>(bb);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) st_stream(nq + k1 + N1 * k2, bb[k2]);
      }
    }
    __syncthreads();
  }
}


// NYQ: the Nyquist column x = K/2 comes from its own array (nyq), which
// the launcher passes exactly when NYQ is set.  Planes as in xy_fwd_block.
template <typename T, int K, int NT, bool NYQ>
__device__ __forceinline__ void xy_inv_block(T *__restrict__ grid, int RS, long comp_stride,
                                             int nplanes, const cpx<T> *__restrict__ tw,
                                             const cpx<T> *__restrict__ nyq, int q0, int qs,
                                             cpx<T> *buf, cpx<T> *s_tw) {
  using S = XYShape<T, K>;
  constexpr int N1 = Split<K>::N1, N2 = Split<K>::N2;
  constexpr int Kh = S::Kh, LP = S::LP, L2 = K / 2;
  constexpr int G = 16 / sizeof(cpx<T>);  // complex per 16-byte load
  // units of G complex per plane: whole rows x < Kh (UPR per row), or, with
  // a separate Nyquist column, rows x < K/2 (UPQ per row) and the column
  // (K/G units)
  constexpr int UPR = (Kh + G - 1) / G, UPQ = (K / 2) / G;
  static_assert((K / 2) % G == 0 && K % G == 0, "whole units");
  constexpr int NUNIT = NYQ ? K * UPQ + K / G : K * UPR;
  constexpr int NU = (NUNIT + NT - 1) / NT;
  struct alignas(16) C16 { cpx<T> v[G]; };
  load_tw<T, K>(s_tw, tw);
  int tid = threadIdx.x;
  const int Khp = RS / 2;
  C16 rv[NU];
  constexpr int upr = NYQ ? UPQ : UPR;  // row units (a constant: divisions by it are cheap)
  auto fetch = [&](int q) {
    const int comp = q / K, z = q - comp * K;
    const cpx<T> *spec =
        reinterpret_cast<const cpx<T> *>(grid + comp * comp_stride) + (size_t)z * K * Khp;
    const cpx<T> *nq = NYQ ? nyq + ((size_t)comp * K + z) * K : nullptr;
#pragma unroll
    for (int r = 0; r < NU; ++r) {
      // full(r): every thread's unit r exists; rows(r): all of them are row
      // units (constants once the loop is unrolled: no branch where true)
      const bool full = NT * (r + 1) <= NUNIT, rows = NT * (r + 1) <= K * upr;
      const int u = tid + NT * r;
      if (full || u < NUNIT) {
        const int y = u / upr, g = u - y * upr;
        if (!NYQ || rows || y < K)
          rv[r] = *reinterpret_cast<const C16 *>(spec + (size_t)y * Khp + G * g);
        else
          rv[r] = *reinterpret_cast<const C16 *>(nq + G * (u - K * upr));
      }
    }
  };
  auto commit = [&]() {
#pragma unroll
    for (int r = 0; r < NU; ++r) {
      const bool full = NT * (r + 1) <= NUNIT, rows = NT * (r + 1) <= K * upr;
      const int u = tid + NT * r;
      if (full || u < NUNIT) {
        const int y = u / upr, g = u - y * upr;
        if (!NYQ || rows || y < K) {
          // with the Nyquist column apart, row units cover x < K/2 < Kh
#pragma unroll
          for (int i = 0; i < G; ++i)
            if (NYQ || G * g + i < Kh) buf[y * Kh + G * g + i] = rv[r].v[i];
        } else {  // Nyquist units: rows G w .. G w + G - 1 of column K/2
          const int w = u - K * upr;
#pragma unroll
          for (int i = 0; i < G; ++i) buf[(G * w + i) * Kh + K / 2] = rv[r].v[i];
        }
      }
    }
  };
  int q = q0;
  if (q < nplanes) fetch(q);
  for (; q < nplanes; q += qs) {
    // keep per-thread index arithmetic inside the loop: hoisted, it would
    // stay live across the transform and spill
    asm volatile("" : "+v"(tid));
    commit();
    __syncthreads();
    // the next plane's loads, in flight through this one
    if (q + qs < nplanes) fetch(q + qs);
    // y inverse, step A in place
    for (int t = tid; t < Kh * N2; t += NT) {
      const int x = t % Kh, n2 = t / Kh;
      cpx<T> a[N1];
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) a[n1] = buf[(n1 * N2 + n2) * Kh + x];
      reg_fft<N1, +1>(a);
#pragma unroll
      for (int k1 = 1; k1 < N1; ++k1) a[k1] = cmul(a[k1], twiddle<+1>(s_tw, n2 * k1));
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) buf[(k1 * N2 + n2) * Kh + x] = a[k1];
    }
    __syncthreads();
    // y inverse, step B, natural order: in rounds of whole columns (a
    // column's tasks read and write the same positions), one register tile
    // per thread.  When the columns before the Nyquist column fill whole
    // rounds, the Nyquist column x = K/2 is done in the last round by the
    // first N1 lanes of wave 0 after their own column: its 16 tasks lie in
    // one wave, and each task's writes wait on its reads through the FFT,
    // so it needs no barrier of its own (a round of its own would cost two).
    constexpr int CPR = NT / N1;  // columns per round
    constexpr bool NYQ_IN_ROUND = (Kh - 1) % CPR == 0 && N1 <= 64;
    constexpr int KX = NYQ_IN_ROUND ? Kh - 1 : Kh;  // columns done in rounds
    for (int x0 = 0; x0 < KX; x0 += CPR) {
      const int x = x0 + tid % CPR, k1 = tid / CPR;  // consecutive lanes: consecutive x
      const bool on = tid < CPR * N1 && x < KX;
      cpx<T> c[N2];
      if (on) {
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) c[n2] = buf[(k1 * N2 + n2) * Kh + x];
        reg_fft<N2, +1>(c);
      }
      __syncthreads();
      if (on) {
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) buf[(k1 + N1 * k2) * Kh + x] = c[k2];
      }
      if (NYQ_IN_ROUND && x0 + CPR >= KX && tid < N1) {
        const int kn = tid;  // task k1 of column K/2
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) c[n2] = buf[(kn * N2 + n2) * Kh + K / 2];
        reg_fft<N2, +1>(c);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) buf[(kn + N1 * k2) * Kh + K / 2] = c[k2];
      }
      __syncthreads();
    }
    // pack row pairs into lines: k and K/2-1-k per thread, K/2 alone
    for (int t = tid; t < L2 * (K / 4 + 1); t += NT) {
      const int l = t / (K / 4 + 1), m = t - l * (K / 4 + 1);
      cpx<T> *ln = buf + l * LP;
      if (m < K / 4) {
        const int k0 = m, k1 = K / 2 - 1 - m;
        cpx<T> A0 = ln[k0], B0 = ln[Kh + k0], A1 = ln[k1], B1 = ln[Kh + k1];
        if (k0 == 0) {
          A0.y = T(0);
          B0.y = T(0);
        }
        ln[k0] = {A0.x - B0.y, A0.y + B0.x};
        if (k0 > 0) ln[K - k0] = {A0.x + B0.y, B0.x - A0.y};
        ln[k1] = {A1.x - B1.y, A1.y + B1.x};
        ln[K - k1] = {A1.x + B1.y, B1.x - A1.y};
      } else {
        xy_pack<T, K>(ln, K / 2);
      }
    }
    __syncthreads();
    // x inverse, step A in place
    for (int t = tid; t < L2 * N2; t += NT) {
      const int l = t / N2, n2 = t - l * N2;
      cpx<T> *ln = buf + l * LP;
      cpx<T> a[N1];
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) a[n1] = ln[n1 * N2 + n2];
      reg_fft<N1, +1>(a);
#pragma unroll
      for (int k1 = 1; k1 < N1; ++k1) a[k1] = cmul(a[k1], twiddle<+1>(s_tw, n2 * k1));
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) ln[k1 * N2 + n2] = a[k1];
    }
    __syncthreads();
    // x inverse, step B straight to the real rows 2l (real part), 2l+1 (imag)
    {
      const int comp = q / K, z = q - comp * K;
      T *base = grid + comp * comp_stride + (size_t)z * K * RS;
      for (int t = tid; t < L2 * N1; t += NT) {
        const int l = t / N1, k1 = t - l * N1;
        cpx<T> bb[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) bb[n2] = buf[l * LP + k1 * N2 + n2];
        reg_fft<N2, +1>(bb);
        T *ra = base + (size_t)(2 * l) * RS;
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) {
          ra[k1 + N1 * k2] = bb[k2].x;
          ra[RS + k1 + N1 * k2] = bb[k2].y;
        }
      }
    }
    __syncthreads();
  }
}

}  // namespace spme
