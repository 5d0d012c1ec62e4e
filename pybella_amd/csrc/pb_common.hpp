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
  // this device's copy of the library's static scratch array (device
  // address from hipGetSymbolAddress at context creation; null if that failed)
  double* scratch = nullptr;
  size_t scratch_bytes = 0;
  int scratch_err = 0;   // hipError_t of the address lookup
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

// v of the lane selected by the DPP control CTRL (both 32-bit halves moved).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum over the 64 lanes of a wave (every lane active), valid in lane 0.  The
// order is the xor butterfly's: lane 0 adds the partial sum at xor distance
// 32, 16, 8, 4, 2, 1 in turn.  The two cross-row steps go through
// ds_bpermute; the four row-local ones are DPP moves.  After the distance-8
// step lanes i and i^8 hold equal sums, so the row rotation by 4 delivers
// the distance-4 partner's value whichever way it rotates.
__device__ __forceinline__ double wave_sum(double v) {
  v += __shfl_xor(v, 32, kWave);
  v += __shfl_xor(v, 16, kWave);
  v += dpp_f64<0x128>(v);   // row_ror 8: lane i ^ 8
  v += dpp_f64<0x124>(v);   // row_ror 4: lane i ^ 4 or i ^ 12 (equal sums)
  v += dpp_f64<0x4E>(v);    // quad_perm [2,3,0,1]: lane i ^ 2
  v += dpp_f64<0xB1>(v);    // quad_perm [1,0,3,2]: lane i ^ 1
  return v;
}

// wave_sum of NV values at once, step by step over all values, so the
// cross-row exchanges of the NV values are in flight together instead of
// one round trip after another.  Same order of additions as wave_sum.
template <int NV>
__device__ __forceinline__ void wave_sum_n(double (&v)[NV]) {
  double t[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) t[k] = __shfl_xor(v[k], 32, kWave);
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] += t[k];
#pragma unroll
  for (int k = 0; k < NV; ++k) t[k] = __shfl_xor(v[k], 16, kWave);
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] += t[k];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] += dpp_f64<0x128>(v[k]);
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] += dpp_f64<0x124>(v[k]);
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] += dpp_f64<0x4E>(v[k]);
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] += dpp_f64<0xB1>(v[k]);
}

// Block-wide sum of NV doubles over a block of exactly NW waves; the result is
// valid in thread 0.  ``sh`` holds NV * NW doubles.  Deterministic order: wave
// sums (wave_sum_n), then thread 0 adds them in wave order.  With the wave
// count a constant, thread 0 reads the NV x NW wave sums in one batch.
template <int NV, int NW>
__device__ __forceinline__ void block_sum_w(double (&v)[NV], double* sh) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int lane = tid & (kWave - 1), wid = tid / kWave;
  double s[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) s[k] = v[k];
  wave_sum_n<NV>(s);
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    if (lane == 0) sh[k * NW + wid] = s[k];
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double w[NW];   // one value's wave sums at a time: NW registers, not NV x NW
#pragma unroll
      for (int i = 0; i < NW; ++i) w[i] = sh[k * NW + i];
      double a = 0.0;
#pragma unroll
      for (int i = 0; i < NW; ++i) a += w[i];
      v[k] = a;
    }
  }
}

inline int ceil_div(int a, int b) { return (a + b - 1) / b; }
