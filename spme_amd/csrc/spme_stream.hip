// Streaming probes: the HBM rates that plain streaming kernels reach on the
// GPU, the practical ceilings beside which the mesh kernels' rates are read
// (the published 8 TB/s is the roofline's peak).  No counterpart in the
// reference.
//
// One kernel per access pattern: read (a sum, so the loads stay live),
// write (fill), copy, and in-place scale.  Each lane keeps U 16-byte
// accesses in flight per step of a grid-stride loop; the caller sweeps
// the grid and the non-temporal variant, since the best shape differs per
// pattern.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>

#include "../../include/spme_ext.h"
#include "spme_launch.h"

namespace spme {
namespace {

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  return set_last_error(code, buf);
}

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int NT = 256;
constexpr int U = 4;  // 16-byte accesses in flight per lane

template <bool NTMP>
__device__ __forceinline__ f4 ld(const f4 *p) {
  if constexpr (NTMP) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NTMP>
__device__ __forceinline__ void st(f4 *p, f4 v) {
  if constexpr (NTMP) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// MODE 0: read a, b[block] = the block's sum; 1: b = s; 2: b = a;
// 3: b = s * b.  n: 16-byte elements per buffer.
template <int MODE, bool NTMP>
__global__ __launch_bounds__(NT) void k_stream(const f4 *__restrict__ a, f4 *__restrict__ b,
                                               long n, float s) {
  const long step = (long)gridDim.x * NT * U;
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  for (long i0 = (long)blockIdx.x * NT * U + threadIdx.x; i0 < n; i0 += step) {
    f4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long i = i0 + (long)u * NT;
      if (i < n) {
        if constexpr (MODE == 0 || MODE == 2) v[u] = ld<NTMP>(a + i);
        else if constexpr (MODE == 3) v[u] = ld<NTMP>(b + i);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long i = i0 + (long)u * NT;
      if (i < n) {
        if constexpr (MODE == 0) acc += v[u];
        else if constexpr (MODE == 1) st<NTMP>(b + i, f4{s, s, s, s});
        else if constexpr (MODE == 2) st<NTMP>(b + i, v[u]);
        else st<NTMP>(b + i, s * v[u]);
      }
    }
  }
  if constexpr (MODE == 0) {
    __shared__ float wsum[NT / 64];
    float t = (acc.x + acc.y) + (acc.z + acc.w);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
      float tot = 0.f;
#pragma unroll
      for (int w = 0; w < NT / 64; ++w) tot += wsum[w];
      reinterpret_cast<float *>(b)[blockIdx.x] = tot;
    }
  }
}

template <int MODE, bool NTMP>
void launch(const f4 *a, f4 *b, long n, float s, int blocks, hipStream_t st) {
  hipLaunchKernelGGL((k_stream<MODE, NTMP>), dim3(blocks), dim3(NT), 0, st, a, b, n, s);
}

template <bool NTMP>
void launch_mode(int mode, const f4 *a, f4 *b, long n, float s, int blocks, hipStream_t st) {
  switch (mode) {
    case 0: launch<0, NTMP>(a, b, n, s, blocks, st); break;
    case 1: launch<1, NTMP>(a, b, n, s, blocks, st); break;
    case 2: launch<2, NTMP>(a, b, n, s, blocks, st); break;
    default: launch<3, NTMP>(a, b, n, s, blocks, st); break;
  }
}

}  // namespace
}  // namespace spme

using namespace spme;

extern "C" int spme_stream_probe(int mode, int nontemporal, const void *a, void *b, long bytes,
                                 float value, int blocks, int reps, void *stream, float *ms) {
  if (mode < 0 || mode > 3) return fail(-1, "stream probe: mode must be 0..3 (got %d)", mode);
  if (bytes <= 0 || bytes % 16) return fail(-1, "stream probe: bytes must be a positive multiple of 16");
  if (!b || ((mode == 0 || mode == 2) && !a)) return fail(-1, "stream probe: null buffer");
  if ((reinterpret_cast<size_t>(a) | reinterpret_cast<size_t>(b)) % 16)
    return fail(-1, "stream probe: buffers must be 16-byte aligned");
  if (blocks < 0 || reps < 1 || !ms) return fail(-1, "stream probe: bad blocks, reps or ms");
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return fail(-2, "stream probe: no device");
  const long n = bytes / 16;
  if (blocks == 0) blocks = 16 * ncu;
  if (mode == 0 && (long)blocks * 4 > bytes)
    return fail(-1, "stream probe: the read probe writes one float per block into b");
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return fail(-2, "stream probe: event");
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return fail(-2, "stream probe: event");
  }
  const f4 *pa = static_cast<const f4 *>(a);
  f4 *pb = static_cast<f4 *>(b);
  auto run = [&] {
    if (nontemporal) launch_mode<true>(mode, pa, pb, n, value, blocks, st);
    else launch_mode<false>(mode, pa, pb, n, value, blocks, st);
  };
  run();  // warm-up (code object load, first touch)
  hipEventRecord(e0, st);
  for (int r = 0; r < reps; ++r) run();
  hipEventRecord(e1, st);
  hipError_t err = hipEventSynchronize(e1);
  float t = 0.f;
  if (err == hipSuccess) err = hipEventElapsedTime(&t, e0, e1);
  if (err == hipSuccess) err = hipGetLastError();
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (err != hipSuccess) return fail(-2, "stream probe: %s", hipGetErrorString(err));
  *ms = t / reps;
  return 0;
}
