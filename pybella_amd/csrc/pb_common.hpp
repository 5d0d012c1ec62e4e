// Shared definitions of the pybella_amd HIP library (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/pybella_amd.h"

struct pb_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;   // stream all work is issued on
  std::string err;
  int n_cu = 0;
};

// Record a formatted error on the context and return code.
inline int pb_fail(pb_ctx* ctx, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  return code;
}

#define PB_HIP(ctx, call)                                                         \
  do {                                                                            \
    hipError_t _e = (call);                                                       \
    if (_e != hipSuccess)                                                         \
      return pb_fail((ctx), PB_ERR_HIP, "%s failed: %s (%s:%d)", #call,           \
                     hipGetErrorString(_e), __FILE__, __LINE__);                  \
  } while (0)

#define PB_LAUNCH_CHECK(ctx) PB_HIP(ctx, hipGetLastError())

constexpr int kWave = 64;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Block-wide sum of NV doubles; result valid in thread 0.  ``sh`` must hold
// NV * (blockDim.x / 64) doubles.  Deterministic order.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* sh) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  const int nw = (blockDim.x * blockDim.y) / kWave;
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int lane2 = tid & (kWave - 1);
  const int wid2 = tid / kWave;
  (void)lane;
  (void)wid;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = wave_sum(v[k]);
    if (lane2 == 0) sh[k * nw + wid2] = s;
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += sh[k * nw + w];
      v[k] = s;
    }
  }
}

inline int ceil_div(int a, int b) { return (a + b - 1) / b; }
