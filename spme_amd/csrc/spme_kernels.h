// Device-side building blocks shared by the SPME kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

namespace spme {

template <typename T> struct cpx { T x, y; };

template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cconj(cpx<T> a) { return {a.x, -a.y}; }
// multiply by s*i (s = +1 or -1)
template <int S, typename T>
__device__ __forceinline__ cpx<T> mul_si(cpx<T> a) {
  return S > 0 ? cpx<T>{-a.y, a.x} : cpx<T>{a.y, -a.x};
}

// In-register DFT of size R with exponent sign S (S = -1 forward).
template <int R, int S, typename T>
__device__ __forceinline__ void small_dft(cpx<T> *v) {
  if constexpr (R == 2) {
    cpx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    cpx<T> s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    cpx<T> s13 = cadd(v[1], v[3]), d13 = mul_si<S>(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
  } else if constexpr (R == 3) {
    const T h = T(0.86602540378443864676372317075294);  // sqrt(3)/2
    cpx<T> t = cadd(v[1], v[2]);
    cpx<T> u = csub(v[1], v[2]);
    cpx<T> m = {v[0].x - T(0.5) * t.x, v[0].y - T(0.5) * t.y};
    cpx<T> r = mul_si<S>(cpx<T>{h * u.x, h * u.y});
    v[0] = cadd(v[0], t);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  }
}

// Radix of the next Stockham stage: 4s first, then a 2, then 3s.
__host__ __device__ constexpr int next_radix(int rem) {
  return (rem % 4 == 0) ? 4 : ((rem % 2 == 0) ? 2 : 3);
}

// One Stockham autosort stage over NL lines of length N held in LDS
// (line l at buf + l*LS), NT threads.  tw[t] = exp(-2 pi i t / N).
// Reads of a stage all complete (barrier) before any of its writes, so the
// stage runs in place.
template <typename T, int N, int NS, int R, int S, int NL, int LS, int NT>
__device__ __forceinline__ void stockham_stage(cpx<T> *buf, const cpx<T> *tw) {
  constexpr int NR = N / R;
  constexpr int NB = NL * NR;
  constexpr int PER = (NB + NT - 1) / NT;
  constexpr int TWS = N / (NS * R);
  cpx<T> v[PER][R];
#pragma unroll
  for (int b = 0; b < PER; ++b) {
    const int g = threadIdx.x + b * NT;
    if (NB % NT == 0 || g < NB) {
      const int line = g / NR, j = g - line * NR;
      const cpx<T> *L = buf + line * LS;
      const int k = j % NS;
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = L[j + r * NR];
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cpx<T> w = tw[k * r * TWS];
          if (S > 0) w.y = -w.y;
          v[b][r] = cmul(v[b][r], w);
        }
      }
      small_dft<R, S>(v[b]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < PER; ++b) {
    const int g = threadIdx.x + b * NT;
    if (NB % NT == 0 || g < NB) {
      const int line = g / NR, j = g - line * NR;
      cpx<T> *L = buf + line * LS;
      const int k = j % NS;
      const int d = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) L[d + r * NS] = v[b][r];
    }
  }
  __syncthreads();
}

template <typename T, int N, int NS, int S, int NL, int LS, int NT>
__device__ __forceinline__ void stockham_all(cpx<T> *buf, const cpx<T> *tw) {
  if constexpr (NS < N) {
    constexpr int R = next_radix(N / NS);
    stockham_stage<T, N, NS, R, S, NL, LS, NT>(buf, tw);
    stockham_all<T, N, NS * R, S, NL, LS, NT>(buf, tw);
  }
}

// Full in-LDS FFT (unnormalised) of NL lines; caller synchronises before.
template <typename T, int N, int S, int NL, int LS, int NT>
__device__ __forceinline__ void lds_fft(cpx<T> *buf, const cpx<T> *tw) {
  stockham_all<T, N, 1, S, NL, LS, NT>(buf, tw);
}

// Wave64 sum.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

}  // namespace spme
