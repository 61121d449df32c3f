// Device-side building blocks shared by the SPME kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

namespace spme {

template <typename T> struct cpx { T x, y; };

// The FFT arithmetic states its fused multiply-adds (contraction off), so
// that every kernel that inlines it rounds alike, whatever the code around.
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
#pragma clang fp contract(off)
  return {fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x)};
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
#pragma clang fp contract(off)
    const T h = T(0.86602540378443864676372317075294);  // sqrt(3)/2
    cpx<T> t = cadd(v[1], v[2]);
    cpx<T> u = csub(v[1], v[2]);
    cpx<T> m = {fma(T(-0.5), t.x, v[0].x), fma(T(-0.5), t.y, v[0].y)};
    cpx<T> r = mul_si<S>(cpx<T>{h * u.x, h * u.y});
    v[0] = cadd(v[0], t);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  }
}

// cos/sin(2 pi t / M) for M = 16 and 12 (M = 8 and 6 use every other
// entry), so that the twiddles of in-register FFTs fold to constants.
__host__ __device__ constexpr double tw_c16(int t) {
  constexpr double v[16] = {1.0, 0.92387953251128675613, 0.70710678118654752440,
                            0.38268343236508977173, 0.0, -0.38268343236508977173,
                            -0.70710678118654752440, -0.92387953251128675613, -1.0,
                            -0.92387953251128675613, -0.70710678118654752440,
                            -0.38268343236508977173, 0.0, 0.38268343236508977173,
                            0.70710678118654752440, 0.92387953251128675613};
  return v[t & 15];
}
__host__ __device__ constexpr double tw_s16(int t) { return tw_c16(t + 12); }
__host__ __device__ constexpr double tw_c12(int t) {
  constexpr double v[12] = {1.0, 0.86602540378443864676, 0.5, 0.0, -0.5,
                            -0.86602540378443864676, -1.0, -0.86602540378443864676, -0.5,
                            0.0, 0.5, 0.86602540378443864676};
  return v[t % 12];
}
__host__ __device__ constexpr double tw_s12(int t) { return tw_c12(t + 9); }

template <int M> __host__ __device__ constexpr double twc(int t) {
  return M == 16 ? tw_c16(t) : M == 8 ? tw_c16(2 * t) : M == 12 ? tw_c12(t) : tw_c12(2 * t);
}
template <int M> __host__ __device__ constexpr double tws(int t) {
  return M == 16 ? tw_s16(t) : M == 8 ? tw_s16(2 * t) : M == 12 ? tw_s12(t) : tw_s12(2 * t);
}

// In-register FFT of size M in {2, 3, 4, 6, 8, 12, 16}, natural order in
// and out, exponent sign S: mixed-radix decimation in time,
// X[k + Q s] = sum_r W_R^{rs} (W_M^{rk} Y_r[k]) with Y_r the size-Q FFT of
// v[r + R q].
template <int M, int S, typename T>
__device__ __forceinline__ void reg_fft(cpx<T> *v) {
  if constexpr (M == 2 || M == 3 || M == 4) {
    small_dft<M, S>(v);
  } else {
    constexpr int R = (M % 4 == 0) ? 4 : ((M % 2 == 0) ? 2 : 3);
    constexpr int Q = M / R;
    cpx<T> sub[R][Q];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int q = 0; q < Q; ++q) sub[r][q] = v[r + R * q];
#pragma unroll
    for (int r = 0; r < R; ++r) reg_fft<Q, S>(sub[r]);
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      cpx<T> t[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (r * k == 0) {
          t[r] = sub[r][k];
        } else {
          const cpx<T> w = {T(twc<M>(r * k)), T(S * tws<M>(r * k))};
          t[r] = cmul(sub[r][k], w);
        }
      }
      small_dft<R, S>(t);
#pragma unroll
      for (int s = 0; s < R; ++s) v[k + Q * s] = t[s];
    }
  }
}

// Four-step split K = N1 * N2 (N1, N2 <= 16) used by the FFT passes.
template <int K> struct Split;
template <> struct Split<16> { static constexpr int N1 = 4, N2 = 4; };
template <> struct Split<24> { static constexpr int N1 = 6, N2 = 4; };
template <> struct Split<32> { static constexpr int N1 = 8, N2 = 4; };
template <> struct Split<48> { static constexpr int N1 = 8, N2 = 6; };
template <> struct Split<64> { static constexpr int N1 = 8, N2 = 8; };
template <> struct Split<96> { static constexpr int N1 = 12, N2 = 8; };
template <> struct Split<128> { static constexpr int N1 = 16, N2 = 8; };
template <> struct Split<192> { static constexpr int N1 = 16, N2 = 12; };
template <> struct Split<256> { static constexpr int N1 = 16, N2 = 16; };

// twiddle exp(S 2 pi i t / K) from the table tw[t] = exp(-2 pi i t / K)
template <int S, typename T>
__device__ __forceinline__ cpx<T> twiddle(const cpx<T> *tw, int t) {
  cpx<T> w = tw[t];
  if (S > 0) w.y = -w.y;
  return w;
}

// Blocks are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8), each with its own L2.  xcd_remap(b, n) gives the logical index of
// hardware block b such that every XCD processes one contiguous range of
// logical blocks, so neighbouring blocks (which share halo data) share an L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
  const int q = n >> 3, r = n & 7;
  const int x = b & 7, i = b >> 3;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// Wave64 sum through DPP lane moves (no LDS traffic): xor-1 and xor-2 quad
// permutes, half-row and row mirrors (each 16-lane row then holds its sum),
// then row_bcast:15 and row_bcast:31 fold the four rows into lane 63, which
// is broadcast.  dpp_ctrl encodings: quad_perm [1,0,3,2] = 0xB1,
// [2,3,0,1] = 0x4E, row_half_mirror = 0x141, row_mirror = 0x140,
// row_bcast:15 = 0x142, row_bcast:31 = 0x143.
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_add(float v) {
  const int m = __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false);
  return v + __int_as_float(m);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_add(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, int(b), CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, ROWS, 0xF, false);
  return v + __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  v = dpp_add<0xB1, 0xF>(v);
  v = dpp_add<0x4E, 0xF>(v);
  v = dpp_add<0x141, 0xF>(v);
  v = dpp_add<0x140, 0xF>(v);
  v = dpp_add<0x142, 0xA>(v);
  v = dpp_add<0x143, 0xC>(v);
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
  } else {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(int(b), 63);
    const int hi = __builtin_amdgcn_readlane(int(b >> 32), 63);
    return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
  }
}

}  // namespace spme
