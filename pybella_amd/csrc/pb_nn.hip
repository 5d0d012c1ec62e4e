// Learned preconditioner network on gfx950: a 4-level convolutional
// encoder-decoder (3x3 convolutions, 16-32-64-64 channels, GELU) with
// hand-written BF16 MFMA kernels and FP32 accumulation, for forward, backward
// and training.  No reference counterpart (pyBELLA has no learned component).
//
// Tensors are NHWC bf16 with a channel stride ``ctot`` and offset ``coff`` so
// that concatenations are views.  Grid levels are node-based 2:1 coarsenings,
// H_{l+1} = (H_l + 1) / 2, matching 2^k + 1 grids (513 -> 257 -> 129 -> 65).
// Down-sampling is injection (coarse(i, j) = fine(2i, 2j)); up-sampling is
// linear interpolation (fine(2i) = coarse(i), fine(2i+1) = mean of neighbours).
//
// Convolution as implicit GEMM: out[p][co] = sum_{tap, ci} in[p + off(tap)][ci]
// W[tap][ci][co] (zero padding).  One wavefront computes a 16-pixel x 16-channel
// tile per v_mfma_f32_16x16x16bf16: lane l holds A[l%16][4(l/16)..+3] (pixel,
// 4 input channels), B[4(l/16)..+3][l%16] (4 input channels, output channel)
// and D[4(l/16)..+3][l%16] (4 pixels, output channel).  The input-gradient
// (dgrad) is the same kernel on the output gradient with taps flipped and the
// weights in [tap][ci][co] layout; the weight gradient reduces over pixels
// (the GEMM K dimension) in per-chunk partials summed in a fixed order.
#include <vector>

#include "pb_op2d.hpp"

using namespace pb2d;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
// LDS of a convolution block's weight tile: 16 NT rows x (cin + 4); the
// launches keep 16 NT cin <= 8192 (checked on the host)
constexpr int kConvWsh = 8192 + 64 * 4;
// grids of at most this many pixels (all images) take k_conv3x3_small
constexpr long kSmallConvPix = 32768;
typedef unsigned short bf16;

namespace {

__device__ __forceinline__ float bf2f(bf16 b) { return __uint_as_float(((unsigned)b) << 16); }
__device__ __forceinline__ bf16 f2bf(float f) {   // round to nearest even
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)(u >> 16);   // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
  return cdf + x * pdf;
}

struct View {
  bf16* p;
  int ctot, coff, c;   // channel stride, offset, channel count (multiple of 16)
};

// ---------------------------------------------------------------------------
// 3x3 convolution (forward or input-gradient), NT output tiles of 16 channels
// ---------------------------------------------------------------------------
// PX: 16-pixel groups per wave (independent accumulation chains, so a wave
// has PX loads and MFMAs in flight per chunk; each pixel's sum is in the same
// order for any PX)
template <int NT, int PX = 1>
__global__ __launch_bounds__(256) void k_conv3x3(const bf16* __restrict__ in, int in_ctot, int in_coff, int cin,
                                                 const bf16* __restrict__ wg, int flip, const float* __restrict__ bias,
                                                 int act, bf16* out, int out_ctot, int out_coff, int cout_valid,
                                                 bf16* pre, int pre_ctot, float* out_f32, int nimg, int h, int w) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long npix = (long)nimg * h * w;
  // waves past the last pixel stay (the block's barriers need them): their
  // loads read offset 0 and are zeroed, and they store nothing
  const long p0 = ((long)blockIdx.x * 4 + wave) * 16 * PX;
  const int kg = lane >> 4;
  bool va[PX];
  int xa[PX], ya[PX];
  long ba[PX];
#pragma unroll
  for (int g = 0; g < PX; ++g) {
    const long pa = p0 + g * 16 + (lane & 15);
    va[g] = pa < npix;
    xa[g] = (int)(pa % w);
    const long t1 = pa / w;
    ya[g] = (int)(t1 % h);
    ba[g] = t1 / h;
  }
  f32x4 acc[PX][NT];
#pragma unroll
  for (int g = 0; g < PX; ++g)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[g][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  // the block's output tiles: t0 .. t0+NT-1 of gridDim.y * NT (a grid split
  // over the output channels gives small levels more waves)
  const int t0 = blockIdx.y * NT;
  const int nout = gridDim.y * NT * 16;
  // A tap's weight tile of the block (16 NT output rows x cin) is staged in
  // LDS once, with 16-byte loads, instead of every wave gathering it from L2;
  // rows padded by 4 so the 16 rows a B operand reads fall in distinct banks
  __shared__ bf16 wsh[kConvWsh];
  const int wst = cin + 4;
  const int nq = NT * 16 * (cin / 8);   // 8-element pieces of the tile
  for (int tap = 0; tap < 9; ++tap) {
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    bool ok[PX];
    long src[PX];
#pragma unroll
    for (int g = 0; g < PX; ++g) {
      const int ys = ya[g] + dy, xs = xa[g] + dx;
      ok[g] = va[g] && ys >= 0 && ys < h && xs >= 0 && xs < w;
      src[g] = ok[g] ? (((ba[g] * h + ys) * w + xs) * in_ctot + in_coff + 4 * kg) : 0;
    }
    const int wt = flip ? 8 - tap : tap;
    __syncthreads();   // the previous tap's reads of wsh are done
    for (int q = threadIdx.x; q < nq; q += 256) {
      const int row = q / (cin / 8), c8 = q - row * (cin / 8);
      const uint4 v = *(const uint4*)(wg + ((long)wt * nout + t0 * 16 + row) * cin + c8 * 8);
      *(uint2*)(wsh + row * wst + c8 * 8) = uint2{v.x, v.y};   // LDS rows are 8-byte aligned
      *(uint2*)(wsh + row * wst + c8 * 8 + 4) = uint2{v.z, v.w};
    }
    __syncthreads();
    const bf16* wrow = wsh + (lane & 15) * wst + 4 * kg;
    // Input-channel chunks of 16 in groups of four: a group's loads are all
    // issued before its MFMAs (one memory latency per group, not per chunk).
    // The input loads are unconditional (src is 0, a valid offset, outside
    // the image; the value is zeroed after).  The MFMA order is that of a
    // plain chunk-by-chunk loop, so the sums are bitwise the same.
    const int nci = cin >> 4;
    int c = 0;
    for (; c + 4 <= nci; c += 4) {
      bf16x4 a[PX][4], b[4][NT];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
#pragma unroll
        for (int g = 0; g < PX; ++g) a[g][u] = *(const bf16x4*)(in + src[g] + (c + u) * 16);
#pragma unroll
        for (int t = 0; t < NT; ++t) b[u][t] = *(const bf16x4*)(wrow + t * 16 * wst + (c + u) * 16);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int g = 0; g < PX; ++g) {
          if (!ok[g]) a[g][u] = bf16x4{0, 0, 0, 0};
#pragma unroll
          for (int t = 0; t < NT; ++t)
            acc[g][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[g][u], b[u][t], acc[g][t], 0, 0, 0);
        }
    }
    for (; c < nci; ++c) {
      bf16x4 a[PX];
#pragma unroll
      for (int g = 0; g < PX; ++g) {
        a[g] = *(const bf16x4*)(in + src[g] + c * 16);
        if (!ok[g]) a[g] = bf16x4{0, 0, 0, 0};
      }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const bf16x4 b = *(const bf16x4*)(wrow + t * 16 * wst + c * 16);
#pragma unroll
        for (int g = 0; g < PX; ++g)
          acc[g][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[g], b, acc[g][t], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int g = 0; g < PX; ++g)
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int co = (t0 + t) * 16 + (lane & 15);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const long p = p0 + g * 16 + 4 * kg + e;
      if (p < npix && co < cout_valid) {
        float v = acc[g][t][e] + (bias ? bias[co] : 0.f);
        const long o = p * out_ctot + out_coff + co;
        if (pre) pre[p * pre_ctot + co] = f2bf(v);   // pre-activations: own buffer, cout channels
        if (act) v = gelu(v);
        if (out) out[o] = f2bf(v);
        if (out_f32 && co == 0) out_f32[p] = v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// 3x3 convolution on small grids (forward, one 16-channel output tile per
// block).  A small launch costs its chain of dependent memory latencies, so
// all nine taps' weights of the tile are staged in LDS at once (one barrier
// instead of two per tap), and all nine taps' input chunks are loaded before
// the first MFMA.  Per output the MFMAs are issued in k_conv3x3's order (tap
// by tap, input chunks in order), so the sums are bitwise those of k_conv3x3.
// ---------------------------------------------------------------------------
// dn (optional): the next level's input, [nimg][(h + 1) / 2][(w + 1) / 2][dn_ctot]
// channels 0 .. cout_valid - 1: the output at every pixel of even row and
// column (the 2x subsampling k_down makes of the output channels, written
// here instead of by a launch of its own).
template <int CIN>
__global__ __launch_bounds__(256) void k_conv3x3_small(const bf16* __restrict__ in, int in_ctot, int in_coff,
                                                       const bf16* __restrict__ wg, const float* __restrict__ bias,
                                                       int act, bf16* out, int out_ctot, int out_coff, int cout_valid,
                                                       bf16* pre, int pre_ctot, float* out_f32, int nimg, int h,
                                                       int w, bf16* dn, int dn_ctot) {
  constexpr int NCI = CIN / 16;
  constexpr int WST = CIN + 4;   // rows padded by 4 bf16 against bank conflicts
  __shared__ bf16 wsh[9 * 16 * WST];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long npix = (long)nimg * h * w;
  const long p0 = ((long)blockIdx.x * 4 + wave) * 16;
  const int kg = lane >> 4;
  const long pa = p0 + (lane & 15);
  const bool va = pa < npix;
  const int xa = (int)(pa % w);
  const long t1 = pa / w;
  const int ya = (int)(t1 % h);
  const long ba = t1 / h;
  const int t0 = blockIdx.y;
  const int nout = gridDim.y * 16;
  // Every load of the block is issued before the first wait: the tile's
  // weights of all nine taps (16-byte loads into registers), then all nine
  // taps' input chunks; the weights go to LDS once they arrive (the input
  // loads still in flight), and tap t's MFMAs wait only for tap t's inputs.
  // On these small grids a block is one chain of memory latencies, so one
  // round trip instead of one per tap.
  constexpr int NQ = 9 * 16 * (CIN / 8), WQ = (NQ + 255) / 256;
  uint4 wv[WQ];
#pragma unroll
  for (int e = 0; e < WQ; ++e) {
    const int q = min((int)threadIdx.x + e * 256, NQ - 1);   // clamped, unconditional; stored only if in range
    const int tap = q / (16 * (CIN / 8)), rem = q - tap * 16 * (CIN / 8);
    const int row = rem / (CIN / 8), c8 = rem - row * (CIN / 8);
    wv[e] = *(const uint4*)(wg + ((long)tap * nout + t0 * 16 + row) * CIN + c8 * 8);
  }
  // a tap's input chunks (unconditional loads; zeroed outside the image)
  auto fetch = [&](int tap, bf16x4 (&a)[NCI]) -> bool {
    const int ys = ya + tap / 3 - 1, xs = xa + tap % 3 - 1;
    const bool ok = va && ys >= 0 && ys < h && xs >= 0 && xs < w;
    const long src = ok ? (((ba * h + ys) * w + xs) * in_ctot + in_coff + 4 * kg) : 0;
#pragma unroll
    for (int c = 0; c < NCI; ++c) a[c] = *(const bf16x4*)(in + src + c * 16);
    return ok;
  };
  bf16x4 a[9][NCI];
  bool ok[9];
#pragma unroll
  for (int tap = 0; tap < 9; ++tap) ok[tap] = fetch(tap, a[tap]);
  const int co = t0 * 16 + (lane & 15);
  const float bco = (bias && co < cout_valid) ? bias[co] : 0.f;   // with the other loads, not after the MFMAs
#pragma unroll
  for (int e = 0; e < WQ; ++e) {
    const int q = (int)threadIdx.x + e * 256;
    if (q < NQ) {
      const int tap = q / (16 * (CIN / 8)), rem = q - tap * 16 * (CIN / 8);
      const int row = rem / (CIN / 8), c8 = rem - row * (CIN / 8);
      bf16* d = wsh + (tap * 16 + row) * WST + c8 * 8;
      *(uint2*)d = uint2{wv[e].x, wv[e].y};
      *(uint2*)(d + 4) = uint2{wv[e].z, wv[e].w};
    }
  }
  __syncthreads();
  f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int tap = 0; tap < 9; ++tap) {
    const bf16* wrow = wsh + (tap * 16 + (lane & 15)) * WST + 4 * kg;
#pragma unroll
    for (int c = 0; c < NCI; ++c) {
      const bf16x4 av = ok[tap] ? a[tap][c] : bf16x4{0, 0, 0, 0};
      const bf16x4 b = *(const bf16x4*)(wrow + c * 16);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, b, acc, 0, 0, 0);
    }
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const long p = p0 + 4 * kg + e;
    if (p < npix && co < cout_valid) {
      float v = acc[e] + bco;
      const long o = p * out_ctot + out_coff + co;
      if (pre) pre[p * pre_ctot + co] = f2bf(v);
      if (act) v = gelu(v);
      if (out) out[o] = f2bf(v);
      if (out_f32 && co == 0) out_f32[p] = v;
      if (dn) {   // small levels: the pixel index fits 32 bits
        const int pi = (int)p, t1 = pi / w, xw = pi - t1 * w, b = t1 / h, yh = t1 - b * h;
        if (((xw | yh) & 1) == 0)
          dn[(((long)b * ((h + 1) / 2) + (yh >> 1)) * ((w + 1) / 2) + (xw >> 1)) * dn_ctot + co] = f2bf(v);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// weight gradient: dW[tap][ci][co] = sum_p in[p + off(tap)][ci] dz[p][co].
// Block (co tile, ci tile) x pixel chunk; wave-strided 16-pixel steps; MFMA
// with A[co][pixel] = dz, B[pixel][ci] = shifted input.  Partials per chunk.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_wgrad(const bf16* __restrict__ in, int in_ctot, int in_coff, int cin,
                                               const bf16* __restrict__ dz, int dz_ctot, int dz_coff, int cout,
                                               int nimg, int h, int w, long chunk, float* partial) {
  __shared__ float red[4][9][16][17];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nct = cout / 16;
  const int cot = blockIdx.y % nct, cit = blockIdx.y / nct;
  const long npix = (long)nimg * h * w;
  const long pbeg = (long)blockIdx.x * chunk;
  const long pend = pbeg + chunk < npix ? pbeg + chunk : npix;
  const int kg = lane >> 4, l16 = lane & 15;
  f32x4 acc[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (long p0 = pbeg + wave * 16; p0 < pend; p0 += 64) {
    bf16x4 a;
    long pp[4];
    int xx[4], yy[4];
    long bb[4];
    bool vv[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const long p = p0 + 4 * kg + e;
      vv[e] = p < pend;
      pp[e] = p;
      xx[e] = (int)(p % w);
      const long t1 = p / w;
      yy[e] = (int)(t1 % h);
      bb[e] = t1 / h;
      a[e] = vv[e] ? (short)dz[p * dz_ctot + dz_coff + cot * 16 + l16] : (short)0;
    }
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      const int dy = tap / 3 - 1, dx = tap % 3 - 1;
      bf16x4 b;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int ys = yy[e] + dy, xs = xx[e] + dx;
        const bool ok = vv[e] && ys >= 0 && ys < h && xs >= 0 && xs < w;
        b[e] = ok ? (short)in[((bb[e] * h + ys) * w + xs) * in_ctot + in_coff + cit * 16 + l16] : (short)0;
      }
      acc[tap] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[tap], 0, 0, 0);
    }
    (void)pp;
  }
  // acc[tap][e] = D[co = 4 kg + e][ci = l16]
#pragma unroll
  for (int tap = 0; tap < 9; ++tap)
#pragma unroll
    for (int e = 0; e < 4; ++e) red[wave][tap][4 * kg + e][l16] = acc[tap][e];
  __syncthreads();
  for (int k = threadIdx.x; k < 9 * 256; k += 256) {
    const int tap = k / 256, rem = k % 256, co = rem / 16, ci = rem % 16;
    const float s = (red[0][tap][co][ci] + red[1][tap][co][ci]) + (red[2][tap][co][ci] + red[3][tap][co][ci]);
    // partial layout [chunk][tap][ci][co] over the padded cin x cout
    partial[(((long)blockIdx.x * 9 + tap) * cin + cit * 16 + ci) * cout + cot * 16 + co] = s;
  }
}

// sum the chunk partials into the master gradient [9][cin_v][cout_v] (true dims)
__global__ void k_wgrad_reduce(const float* partial, int nchunk, int cin, int cout, int cin_v, int cout_v, float* g) {
  const long n = 9L * cin_v * cout_v;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    const int co = (int)(k % cout_v);
    const long t1 = k / cout_v;
    const int ci = (int)(t1 % cin_v);
    const int tap = (int)(t1 / cin_v);
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c) s += partial[(((long)c * 9 + tap) * cin + ci) * cout + co];
    g[k] = s;
  }
}

// bias gradient: per output channel sum over pixels (chunk partials, then reduce)
__global__ void k_bgrad(const bf16* dz, int ctot, int coff, int cout_v, long npix, long chunk, float* partial) {
  const int co = threadIdx.x & 63;
  const long pbeg = (long)blockIdx.x * chunk;
  const long pend = pbeg + chunk < npix ? pbeg + chunk : npix;
  float s = 0.f;
  if (co < cout_v)
    for (long p = pbeg + (threadIdx.x >> 6); p < pend; p += blockDim.x >> 6) s += bf2f(dz[p * ctot + coff + co]);
  __shared__ float sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < 64 && co < cout_v) {
    float t = 0.f;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += sh[k * 64 + threadIdx.x];
    partial[(long)blockIdx.x * 64 + co] = t;
  }
}
__global__ void k_bgrad_reduce(const float* partial, int nchunk, int cout_v, float* g) {
  const int co = threadIdx.x;
  if (co >= cout_v) return;
  float s = 0.f;
  for (int c = 0; c < nchunk; ++c) s += partial[(long)c * 64 + co];
  g[co] = s;
}

// dz *= gelu'(pre) on a view
__global__ void k_act_bwd(bf16* d, int ctot, int coff, int c, const bf16* pre, int pctot, int pcoff, long npix) {
  const long n = npix * c;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    const long p = k / c;
    const int ch = (int)(k % c);
    const long o = p * ctot + coff + ch;
    d[o] = f2bf(bf2f(d[o]) * gelu_grad(bf2f(pre[p * pctot + pcoff + ch])));
  }
}

// down-sampling by injection and its adjoint (accumulating)
__global__ void k_down(const bf16* f, int fct, int fco, bf16* c, int cct, int cco, int ch, int nimg, int hc, int wc,
                       int hf, int wf) {
  // 32-bit indices (the host keeps every activation tensor below 2^31 values)
  const int n = nimg * hc * wc * ch;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    int t = k / ch;
    const int cc = k - t * ch;
    const int x = t % wc;
    t /= wc;
    const int y = t % hc;
    const int b = t / hc;
    c[((b * hc + y) * wc + x) * cct + cco + cc] = f[((b * hf + 2 * y) * wf + 2 * x) * fct + fco + cc];
  }
}
__global__ void k_down_bwd(const bf16* dc, int cct, int cco, bf16* df, int fct, int fco, int ch, int nimg, int hc,
                           int wc, int hf, int wf) {
  const long n = (long)nimg * hc * wc * ch;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    const int cc = (int)(k % ch);
    long t = k / ch;
    const int x = (int)(t % wc);
    t /= wc;
    const int y = (int)(t % hc);
    const long b = t / hc;
    const long o = ((b * hf + 2 * y) * wf + 2 * x) * fct + fco + cc;
    df[o] = f2bf(bf2f(df[o]) + bf2f(dc[((b * hc + y) * wc + x) * cct + cco + cc]));
  }
}

// linear up-sampling (fine = 2 coarse - 1 per axis) and its adjoint (gather)
// Up-sampling of ch channels (a multiple of 8): one thread per fine pixel and
// 8 channels, 16-byte loads of the 4 coarse neighbours and a 16-byte store
// (channel strides and offsets are multiples of 16 elements).
__global__ void k_up(const bf16* c, int cct, int cco, bf16* f, int fct, int fco, int ch, int nimg, int hc, int wc,
                     int hf, int wf) {
  const int cv = ch / 8;
  // 32-bit indices (the host keeps every activation tensor below 2^31 values)
  const int n = nimg * hf * wf * cv;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const int pix = k / cv;
    const int c8 = (k - pix * cv) * 8;
    const int x = pix % wf;
    const int t = pix / wf;
    const int y = t % hf;
    const int b = t / hf;
    const int y0 = y >> 1, x0 = x >> 1;
    const int y1 = (y & 1) ? y0 + 1 : y0, x1 = (x & 1) ? x0 + 1 : x0;
    auto at = [&](int yy, int xx) {
      return *(const uint4*)(c + ((b * hc + yy) * wc + xx) * cct + cco + c8);
    };
    const uint4 q00 = at(y0, x0), q01 = at(y0, x1), q10 = at(y1, x0), q11 = at(y1, x1);
    const bf16* a00 = (const bf16*)&q00;
    const bf16* a01 = (const bf16*)&q01;
    const bf16* a10 = (const bf16*)&q10;
    const bf16* a11 = (const bf16*)&q11;
    uint4 o;
    bf16* ov = (bf16*)&o;
#pragma unroll
    for (int e = 0; e < 8; ++e)
      ov[e] = f2bf(0.25f * ((bf2f(a00[e]) + bf2f(a01[e])) + (bf2f(a10[e]) + bf2f(a11[e]))));
    *(uint4*)(f + ((b * hf + y) * wf + x) * fct + fco + c8) = o;
  }
}
__global__ void k_up_bwd(const bf16* df, int fct, int fco, bf16* dc, int cct, int cco, int ch, int nimg, int hc,
                         int wc, int hf, int wf) {
  const long n = (long)nimg * hc * wc * ch;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    const int cc = (int)(k % ch);
    long t = k / ch;
    const int x = (int)(t % wc);
    t /= wc;
    const int y = (int)(t % hc);
    const long b = t / hc;
    float s = 0.f;
    for (int ay = -1; ay <= 1; ++ay) {
      const int fy = 2 * y + ay;
      if (fy < 0 || fy >= hf) continue;
      const float wy = ay == 0 ? 1.f : 0.5f;
      for (int ax = -1; ax <= 1; ++ax) {
        const int fx = 2 * x + ax;
        if (fx < 0 || fx >= wf) continue;
        const float wx = ax == 0 ? 1.f : 0.5f;
        s += wy * wx * bf2f(df[((b * hf + fy) * wf + fx) * fct + fco + cc]);
      }
    }
    dc[((b * hc + y) * wc + x) * cct + cco + cc] = f2bf(s);
  }
}

// network input: ch0 = r / scale, ch1, ch2 = coefficient features, rest 0
template <typename Tr>
__global__ void k_input(const Tr* r, const float* scale, const float* f1, const float* f2, int feat_per_img,
                        long npix_img, int nimg, bf16* x0) {
  const long n = npix_img * nimg;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    const long b = k / npix_img, p = k % npix_img;
    const long fo = feat_per_img ? k : p;
    bf16* o = x0 + k * 16;
    o[0] = f2bf((float)((double)r[k] / (double)scale[b]));
    o[1] = f2bf(f1[fo]);
    o[2] = f2bf(f2[fo]);
#pragma unroll
    for (int c = 3; c < 16; ++c) o[c] = 0;
  }
}

// per-image sums: sum r^2 (-> rms scale), or sum z, sum c z for the null-space fit
template <typename Tr>
__global__ void k_img_sums(const Tr* v, long npix_img, int w, int mode, double* part, int nchunk) {
  const int b = blockIdx.y;
  const long chunk = (npix_img + nchunk - 1) / nchunk;
  const long beg = blockIdx.x * chunk, end = beg + chunk < npix_img ? beg + chunk : npix_img;
  double s0 = 0.0, s1 = 0.0;
  for (long p = beg + threadIdx.x; p < end; p += blockDim.x) {
    const double x = (double)v[b * npix_img + p];
    if (mode == 0) {
      s0 += x * x;
    } else {
      const long y = p / w, xx = p % w;
      s0 += x;
      s1 += ((y + xx) & 1) ? -x : x;
    }
  }
  __shared__ double sh[2][256];
  sh[0][threadIdx.x] = s0;
  sh[1][threadIdx.x] = s1;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + off];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[((long)b * nchunk + blockIdx.x) * 2] = sh[0][0];
    part[((long)b * nchunk + blockIdx.x) * 2 + 1] = sh[1][0];
  }
}

// scale[b] = sqrt(mean r^2) (floored)
__global__ void k_scale_from(const double* part, int nchunk, long npix_img, float* scale) {
  const int b = threadIdx.x;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += part[((long)b * nchunk + c) * 2];
  const double rms = sqrt(s / (double)npix_img);
  scale[b] = (float)(rms > 1e-300 ? rms : 1.0);
}

// null-space coefficients (mu, kappa) of each image: least squares on {1, c}
// (mode 0) or on {1} alone (mode 1, kappa = 0)
__global__ void k_null_coef(const double* part, int nchunk, long npix_img, int w, int h, int mode, double* coef) {
  const int b = threadIdx.x;
  double sz = 0.0, scz = 0.0;
  for (int c = 0; c < nchunk; ++c) {
    sz += part[((long)b * nchunk + c) * 2];
    scz += part[((long)b * nchunk + c) * 2 + 1];
  }
  if (mode == 1) {
    coef[2 * b] = sz / (double)npix_img;
    coef[2 * b + 1] = 0.0;
    return;
  }
  // sum of the checkerboard over the grid
  const long ne = ((long)h * w + 1) / 2, no = (long)h * w / 2;
  const double sc = (double)(ne - no), n = (double)npix_img;
  const double det = n * n - sc * sc;
  coef[2 * b] = (n * sz - sc * scz) / det;
  coef[2 * b + 1] = (n * scz - sc * sz) / det;
}

// loss: z projected off the null space vs target; gradient into channel 0 of dout
__global__ void k_loss(const float* z, const float* target, const double* coef, const float* wimg, long npix_img,
                       int w, int nimg, float inv_n, bf16* dout, double* lpart) {
  const long n = npix_img * nimg;
  double l = 0.0;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    const long b = k / npix_img, p = k % npix_img;
    const long y = p / w, x = p % w;
    const double c = ((y + x) & 1) ? -1.0 : 1.0;
    const double zp = (double)z[k] - coef[2 * b] - coef[2 * b + 1] * c;
    const double d = zp - (double)target[k];
    const double wb = wimg ? (double)wimg[b] : 1.0;
    l += wb * d * d;
    bf16* o = dout + k * 16;
    o[0] = f2bf((float)(2.0 * wb * d) * inv_n);
#pragma unroll
    for (int ch = 1; ch < 16; ++ch) o[ch] = 0;
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = l;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) lpart[blockIdx.x] = sh[0];
}

// z_out = scale * z (optionally projected off the null space)
template <typename To>
__global__ void k_output(const float* z, const float* scale, long npix_img, int nimg, To* out) {
  const long n = npix_img * nimg;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x)
    out[k] = (To)((double)z[k] * (double)scale[k / npix_img]);
}

__global__ void k_adam(long n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                       float bc1, float bc2) {
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    w[k] -= lr * (mk / bc1) / (sqrtf(vk / bc2) + eps);
  }
}

// master [9][cin_v][cout_v] fp32 -> forward [9][cout_p][cin_p] and backward
// [9][cin_p][cout_p] bf16, zero padded
__global__ void k_pack_w(const float* wm, int cin_v, int cout_v, int cin_p, int cout_p, bf16* wf, bf16* wb) {
  const long n = 9L * cin_p * cout_p;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    const int co = (int)(k % cout_p);
    const long t1 = k / cout_p;
    const int ci = (int)(t1 % cin_p);
    const int tap = (int)(t1 / cin_p);
    const float v = (ci < cin_v && co < cout_v) ? wm[((long)tap * cin_v + ci) * cout_v + co] : 0.f;
    const bf16 b = f2bf(v);
    wb[k] = b;
    wf[((long)tap * cout_p + co) * cin_p + ci] = b;
  }
}

__global__ void k_split3(const float* in, long n, bf16* x0) {
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    bf16* o = x0 + k * 16;
    o[0] = f2bf(in[3 * k]);
    o[1] = f2bf(in[3 * k + 1]);
    o[2] = f2bf(in[3 * k + 2]);
#pragma unroll
    for (int c = 3; c < 16; ++c) o[c] = 0;
  }
}

__global__ void k_fill_bf16(long n, bf16* a) {
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) a[k] = 0;
}

}  // namespace

// ===========================================================================
// host side: fixed architecture
// ===========================================================================
namespace {

struct LayerSpec {
  int cin_v, cout_v;   // true channel counts
  int act;
};

// 15 convolutions in forward order (see the table in the file header)
constexpr LayerSpec kLayers[15] = {
    {3, 16, 1},  {16, 16, 1},  {16, 32, 1}, {32, 32, 1}, {32, 64, 1},  {64, 64, 1}, {64, 64, 1}, {64, 64, 1},
    {128, 64, 1}, {64, 64, 1}, {96, 32, 1}, {32, 32, 1}, {48, 16, 1}, {16, 16, 1}, {16, 1, 0}};

constexpr int pad16(int c) { return (c + 15) / 16 * 16; }

// a convolution block stages at most 16 NT x cin weights, NT x 16 <= the
// layer's (padded) output channels: forward 16 NT cin <= cin_p cout_p, and
// the input gradient the same with the roles swapped
constexpr bool weight_tiles_fit() {
  for (const LayerSpec& L : kLayers)
    if (pad16(L.cin_v) * pad16(L.cout_v) > 8192) return false;
  return true;
}
static_assert(weight_tiles_fit(), "a layer's weight tile exceeds the convolution's LDS stage");

struct Level {
  int h, w;
};

struct Acts {   // activation / gradient buffers of one (nimg, h, w) shape
  int nimg = 0, h = 0, w = 0;
  Level lv[4];
  std::vector<void*> allocs;
  // activations
  bf16 *X0, *E0a, *CAT0, *D0a, *D0b, *OUT16;
  bf16 *P1, *E1a, *CAT1, *D1a, *D1b;
  bf16 *P2, *E2a, *CAT2, *D2a, *D2b;
  bf16 *P3, *B3a, *B3b;
  // pre-activations (training)
  bf16 *pE0a, *pE0b, *pD0a, *pD0b, *pE1a, *pE1b, *pD1a, *pD1b, *pE2a, *pE2b, *pD2a, *pD2b, *pB3a, *pB3b;
  // gradients
  bf16 *dOUT, *dD0b, *dD0a, *dCAT0, *dD1b, *dD1a, *dCAT1, *dD2b, *dD2a, *dCAT2, *dB3b, *dB3a, *dP3, *dP2, *dP1,
      *dE2a, *dE1a, *dE0a;
  float* zout;       // fp32 network output (channel 0)
  float* scale;      // per image
  double* part;      // per image partial sums
  double* coef;      // per image null-space coefficients
  double* lpart;     // loss partials
  float* feat1;      // per-image features when training (nimg * npix)
  float* feat2;
  bool train = false;
};

}  // namespace

struct pb_nn {
  pb_ctx* ctx = nullptr;
  long nparam = 0;
  long woff[15], boff[15];     // offsets into the master parameter vector
  float* params = nullptr;     // master fp32
  float* grads = nullptr;
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  long adam_t = 0;
  bf16* wf[15];                // forward layout
  bf16* wb[15];                // backward layout
  float* wpart = nullptr;      // wgrad partials
  long wpart_cap = 0;
  float* bpart = nullptr;
  std::vector<void*> allocs;
  std::vector<Acts*> shapes;
  // features of the last operator used for inference
  const pb_op2d* feat_op = nullptr;
  float* f1 = nullptr;
  float* f2 = nullptr;
  long feat_n = 0, feat_cap = 0, feat_version = -1;
  std::vector<pb_op2d*> attached;   // operators this network is attached to
  int null_mode = 0;                // training loss: 0 removes {1, checkerboard}, 1 removes {1}
};

namespace {

template <typename T>
int nn_malloc(pb_nn* nn, std::vector<void*>& list, T** p, long count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, sizeof(T) * (size_t)std::max<long>(count, 1));
  if (e != hipSuccess)
    return pb_fail(nn->ctx, PB_ERR_HIP, "hipMalloc(%ld bytes) failed: %s", (long)(sizeof(T) * count),
                   hipGetErrorString(e));
  list.push_back(q);
  *p = (T*)q;
  return PB_OK;
}

inline int gs(long n, int per = 256) { return (int)std::min<long>((n + per - 1) / per, 65535); }

int pack_weights(pb_nn* nn, hipStream_t s) {
  for (int l = 0; l < 15; ++l) {
    const LayerSpec& L = kLayers[l];
    const int cin_p = pad16(L.cin_v), cout_p = pad16(L.cout_v);
    hipLaunchKernelGGL(k_pack_w, dim3(gs(9L * cin_p * cout_p)), dim3(256), 0, s, nn->params + nn->woff[l], L.cin_v,
                       L.cout_v, cin_p, cout_p, nn->wf[l], nn->wb[l]);
  }
  PB_LAUNCH_CHECK(nn->ctx);
  return PB_OK;
}

Acts* get_shape(pb_nn* nn, int nimg, int h, int w, bool train, int* rc) {
  for (Acts* a : nn->shapes)
    if (a->nimg == nimg && a->h == h && a->w == w && (a->train || !train)) return a;
  // the widest activation (128 channels at full resolution) must index in
  // 32 bits: the pooling and upsampling kernels use int indices
  if (nimg < 1 || h < 1 || w < 1 || (long)nimg * h * w * 128 > 0x7fffffffL) {
    *rc = pb_fail(nn->ctx, PB_ERR_ARG, "network: %d images of %d x %d pixels exceed 2^31 activation values", nimg,
                  h, w);
    return nullptr;
  }
  Acts* a = new Acts();
  a->nimg = nimg;
  a->h = h;
  a->w = w;
  a->train = train;
  a->lv[0] = {h, w};
  for (int l = 1; l < 4; ++l) a->lv[l] = {(a->lv[l - 1].h + 1) / 2, (a->lv[l - 1].w + 1) / 2};
  auto np = [&](int l) { return (long)nimg * a->lv[l].h * a->lv[l].w; };
  int r = 0;
  auto T = [&](bf16** p, int l, int c) { r |= nn_malloc(nn, a->allocs, p, np(l) * c); };
  T(&a->X0, 0, 16);
  T(&a->E0a, 0, 16);
  T(&a->CAT0, 0, 48);
  T(&a->D0a, 0, 16);
  T(&a->D0b, 0, 16);
  T(&a->OUT16, 0, 16);
  T(&a->P1, 1, 16);
  T(&a->E1a, 1, 32);
  T(&a->CAT1, 1, 96);
  T(&a->D1a, 1, 32);
  T(&a->D1b, 1, 32);
  T(&a->P2, 2, 32);
  T(&a->E2a, 2, 64);
  T(&a->CAT2, 2, 128);
  T(&a->D2a, 2, 64);
  T(&a->D2b, 2, 64);
  T(&a->P3, 3, 64);
  T(&a->B3a, 3, 64);
  T(&a->B3b, 3, 64);
  if (train) {
    T(&a->pE0a, 0, 16);
    T(&a->pE0b, 0, 16);
    T(&a->pD0a, 0, 16);
    T(&a->pD0b, 0, 16);
    T(&a->pE1a, 1, 32);
    T(&a->pE1b, 1, 32);
    T(&a->pD1a, 1, 32);
    T(&a->pD1b, 1, 32);
    T(&a->pE2a, 2, 64);
    T(&a->pE2b, 2, 64);
    T(&a->pD2a, 2, 64);
    T(&a->pD2b, 2, 64);
    T(&a->pB3a, 3, 64);
    T(&a->pB3b, 3, 64);
    T(&a->dOUT, 0, 16);
    T(&a->dD0b, 0, 16);
    T(&a->dD0a, 0, 16);
    T(&a->dCAT0, 0, 48);
    T(&a->dD1b, 1, 32);
    T(&a->dD1a, 1, 32);
    T(&a->dCAT1, 1, 96);
    T(&a->dD2b, 2, 64);
    T(&a->dD2a, 2, 64);
    T(&a->dCAT2, 2, 128);
    T(&a->dB3b, 3, 64);
    T(&a->dB3a, 3, 64);
    T(&a->dP3, 3, 64);
    T(&a->dP2, 2, 32);
    T(&a->dP1, 1, 16);
    T(&a->dE2a, 2, 64);
    T(&a->dE1a, 1, 32);
    T(&a->dE0a, 0, 16);
    r |= nn_malloc(nn, a->allocs, &a->feat1, np(0));
    r |= nn_malloc(nn, a->allocs, &a->feat2, np(0));
  }
  r |= nn_malloc(nn, a->allocs, &a->zout, np(0));
  r |= nn_malloc(nn, a->allocs, &a->scale, nimg);
  r |= nn_malloc(nn, a->allocs, &a->part, 2L * nimg * 64);
  r |= nn_malloc(nn, a->allocs, &a->coef, 2L * nimg);
  r |= nn_malloc(nn, a->allocs, &a->lpart, 1024);
  if (r) {
    for (void* p : a->allocs) (void)hipFree(p);
    delete a;
    *rc = PB_ERR_HIP;
    return nullptr;
  }
  nn->shapes.push_back(a);
  *rc = PB_OK;
  return a;
}

// resident 256-thread blocks of k_conv3x3<NT> on the whole device (CUs x
// blocks per CU at the kernel's occupancy), queried once per NT
long conv_slots(int nt) {
  static long slots[5] = {0, 0, 0, 0, 0};
  if (slots[nt] == 0) {
    int dev = 0, cus = 0, per_cu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    hipError_t e = hipErrorInvalidValue;
    if (nt == 1) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_conv3x3<1>, 256, 0);
    else if (nt == 2) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_conv3x3<2>, 256, 0);
    else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_conv3x3<4>, 256, 0);
    slots[nt] = (e == hipSuccess && cus > 0 && per_cu > 0) ? (long)cus * per_cu : 1;
  }
  return slots[nt];
}

void down(Acts* a, int lf, const bf16* f, int fct, int fco, bf16* c, int cct, int cco, int ch, hipStream_t s);

// dn (optional): also the 2x subsampling of the output channels into the
// next level's input dn (channels 0.., dn_ct per pixel), as down() makes it
void conv(pb_nn* nn, Acts* a, int l, int level, const bf16* in, int in_ct, int in_co, bf16* out, int out_ct,
          int out_co, bf16* pre, float* out_f32, hipStream_t s, bf16* dn = nullptr, int dn_ct = 0) {
  const LayerSpec& L = kLayers[l];
  const int cin = pad16(L.cin_v), nt = pad16(L.cout_v) / 16;
  const long npix = (long)a->nimg * a->lv[level].h * a->lv[level].w;
  const float* bias = nn->params + nn->boff[l];
  const int h = a->lv[level].h, w = a->lv[level].w;
  // A wave walks 9 taps x cin/16 chunks of dependent loads.  On small levels
  // the output tiles are split over the grid's y (fewer tiles per block, more
  // waves to hide that latency; the input loads repeat, from cache), with the
  // fewest tiles per block whose grid fits in one round of resident blocks:
  // a second round of a few blocks would double the launch.
  // Large levels take two 16-pixel groups per wave instead (two chains of
  // loads and MFMAs in flight per wave, half the blocks).
  const bool large = npix >= 65536;
  const long bx = (npix + 63) / 64;
  if (npix <= kSmallConvPix) {   // one output tile per block, all taps' weights staged at once
    const dim3 gs_((unsigned)bx, nt), bs_(256);
#define PB_CONVS(CI)                                                                                         \
  hipLaunchKernelGGL((k_conv3x3_small<CI>), gs_, bs_, 0, s, in, in_ct, in_co, nn->wf[l], bias, L.act, out, out_ct, \
                     out_co, L.cout_v, pre, pad16(L.cout_v), out_f32, a->nimg, h, w, dn, dn_ct)
    switch (cin) {
      case 16: PB_CONVS(16); break;
      case 32: PB_CONVS(32); break;
      case 48: PB_CONVS(48); break;
      case 64: PB_CONVS(64); break;
      case 96: PB_CONVS(96); break;
      default: PB_CONVS(128); break;
    }
#undef PB_CONVS
    return;
  }
  int ntb = nt;
  if (!large && nt > 1) {
    ntb = 1;
    for (int c : {1, 2, 4})
      if (c <= nt && nt % c == 0 && bx * (nt / c) <= conv_slots(c)) {
        ntb = c;
        break;
      }
  }
  const dim3 grd((unsigned)(large ? (npix + 127) / 128 : bx), nt / ntb), blk(256);
#define PB_CONV(NT, PX)                                                                                           \
  hipLaunchKernelGGL((k_conv3x3<NT, PX>), grd, blk, 0, s, in, in_ct, in_co, cin, nn->wf[l], 0, bias, L.act, out,   \
                     out_ct, out_co, L.cout_v, pre, pad16(L.cout_v), out_f32, a->nimg, h, w)
  if (large) {
    switch (ntb) {
      case 1: PB_CONV(1, 2); break;
      case 2: PB_CONV(2, 2); break;
      default: PB_CONV(4, 2); break;
    }
  } else {
    switch (ntb) {
      case 1: PB_CONV(1, 1); break;
      case 2: PB_CONV(2, 1); break;
      default: PB_CONV(4, 1); break;
    }
  }
#undef PB_CONV
  if (dn) down(a, level, out, out_ct, out_co, dn, dn_ct, 0, L.cout_v, s);
}

// input gradient of layer l: din (cin_p channels) from dz (cout_p channels)
void dgrad(pb_nn* nn, Acts* a, int l, int level, const bf16* dz, int dz_ct, int dz_co, bf16* din, int din_ct,
           int din_co, hipStream_t s) {
  const LayerSpec& L = kLayers[l];
  const int cout_p = pad16(L.cout_v), cin_p = pad16(L.cin_v);
  const long npix = (long)a->nimg * a->lv[level].h * a->lv[level].w;
  const dim3 grd((unsigned)((npix + 63) / 64)), blk(256);
  const int h = a->lv[level].h, w = a->lv[level].w;
#define PB_DG(NT)                                                                                                 \
  hipLaunchKernelGGL(k_conv3x3<NT>, grd, blk, 0, s, dz, dz_ct, dz_co, cout_p, nn->wb[l], 1, nullptr, 0, din, din_ct, \
                     din_co, cin_p, nullptr, 0, nullptr, a->nimg, h, w)
  switch (cin_p / 16) {
    case 1: PB_DG(1); break;
    case 2: PB_DG(2); break;
    case 3: PB_DG(3); break;
    case 4: PB_DG(4); break;
    case 6: PB_DG(6); break;
    default: PB_DG(8); break;
  }
#undef PB_DG
}

int wgrad(pb_nn* nn, Acts* a, int l, int level, const bf16* in, int in_ct, int in_co, const bf16* dz, int dz_ct,
          int dz_co, hipStream_t s) {
  const LayerSpec& L = kLayers[l];
  const int cin_p = pad16(L.cin_v), cout_p = pad16(L.cout_v);
  const long npix = (long)a->nimg * a->lv[level].h * a->lv[level].w;
  const int h = a->lv[level].h, w = a->lv[level].w;
  const int ntiles = (cin_p / 16) * (cout_p / 16);
  int nchunk = (int)std::max<long>(1, std::min<long>(2048 / ntiles, (npix + 1023) / 1024));
  const long chunk = (npix + nchunk - 1) / nchunk;
  nchunk = (int)((npix + chunk - 1) / chunk);
  const long need = (long)nchunk * 9 * cin_p * cout_p;
  if (need > nn->wpart_cap) return pb_fail(nn->ctx, PB_ERR_STATE, "wgrad partial buffer too small");
  hipLaunchKernelGGL(k_wgrad, dim3(nchunk, ntiles), dim3(256), 0, s, in, in_ct, in_co, cin_p, dz, dz_ct, dz_co, cout_p,
                     a->nimg, h, w, chunk, nn->wpart);
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(gs(9L * L.cin_v * L.cout_v)), dim3(256), 0, s, (const float*)nn->wpart,
                     nchunk, cin_p, cout_p, L.cin_v, L.cout_v, nn->grads + nn->woff[l]);
  const int nbc = (int)std::min<long>(1024, (npix + 255) / 256);
  const long bchunk = (npix + nbc - 1) / nbc;
  hipLaunchKernelGGL(k_bgrad, dim3(nbc), dim3(256), 0, s, dz, dz_ct, dz_co, L.cout_v, npix, bchunk, nn->bpart);
  hipLaunchKernelGGL(k_bgrad_reduce, dim3(1), dim3(64), 0, s, (const float*)nn->bpart, nbc, L.cout_v,
                     nn->grads + nn->boff[l]);
  return PB_OK;
}

void act_bwd(Acts* a, int level, bf16* d, int ct, int co, int c, const bf16* pre, int pct, int pco, hipStream_t s) {
  const long npix = (long)a->nimg * a->lv[level].h * a->lv[level].w;
  hipLaunchKernelGGL(k_act_bwd, dim3(gs(npix * c)), dim3(256), 0, s, d, ct, co, c, pre, pct, pco, npix);
}

void down(Acts* a, int lf, const bf16* f, int fct, int fco, bf16* c, int cct, int cco, int ch, hipStream_t s) {
  const Level F = a->lv[lf], Cc = a->lv[lf + 1];
  const long n = (long)a->nimg * Cc.h * Cc.w * ch;
  hipLaunchKernelGGL(k_down, dim3(gs(n)), dim3(256), 0, s, f, fct, fco, c, cct, cco, ch, a->nimg, Cc.h, Cc.w, F.h, F.w);
}
void down_bwd(Acts* a, int lf, const bf16* dc, int cct, int cco, bf16* df, int fct, int fco, int ch, hipStream_t s) {
  const Level F = a->lv[lf], Cc = a->lv[lf + 1];
  const long n = (long)a->nimg * Cc.h * Cc.w * ch;
  hipLaunchKernelGGL(k_down_bwd, dim3(gs(n)), dim3(256), 0, s, dc, cct, cco, df, fct, fco, ch, a->nimg, Cc.h, Cc.w, F.h,
                     F.w);
}
void up(Acts* a, int lf, const bf16* c, int cct, int cco, bf16* f, int fct, int fco, int ch, hipStream_t s) {
  const Level F = a->lv[lf], Cc = a->lv[lf + 1];
  const long n = (long)a->nimg * F.h * F.w * (ch / 8);   // threads: pixel x 8 channels
  hipLaunchKernelGGL(k_up, dim3(gs(n)), dim3(256), 0, s, c, cct, cco, f, fct, fco, ch, a->nimg, Cc.h, Cc.w, F.h, F.w);
}
void up_bwd(Acts* a, int lf, const bf16* df, int fct, int fco, bf16* dc, int cct, int cco, int ch, hipStream_t s) {
  const Level F = a->lv[lf], Cc = a->lv[lf + 1];
  const long n = (long)a->nimg * Cc.h * Cc.w * ch;
  hipLaunchKernelGGL(k_up_bwd, dim3(gs(n)), dim3(256), 0, s, df, fct, fco, dc, cct, cco, ch, a->nimg, Cc.h, Cc.w, F.h,
                     F.w);
}

// forward from a->X0 to a->zout (fp32 channel 0); train: keep pre-activations
void forward(pb_nn* nn, Acts* a, hipStream_t s) {
  const bool tr = a->train;
  auto P = [&](bf16* p) { return tr ? p : nullptr; };
  conv(nn, a, 0, 0, a->X0, 16, 0, a->E0a, 16, 0, P(a->pE0a), nullptr, s);
  // each encoder level's second convolution also writes the next level's
  // input, its output's 2x subsampling (P1, P2, P3)
  conv(nn, a, 1, 0, a->E0a, 16, 0, a->CAT0, 48, 32, P(a->pE0b), nullptr, s, a->P1, 16);
  conv(nn, a, 2, 1, a->P1, 16, 0, a->E1a, 32, 0, P(a->pE1a), nullptr, s);
  conv(nn, a, 3, 1, a->E1a, 32, 0, a->CAT1, 96, 64, P(a->pE1b), nullptr, s, a->P2, 32);
  conv(nn, a, 4, 2, a->P2, 32, 0, a->E2a, 64, 0, P(a->pE2a), nullptr, s);
  conv(nn, a, 5, 2, a->E2a, 64, 0, a->CAT2, 128, 64, P(a->pE2b), nullptr, s, a->P3, 64);
  conv(nn, a, 6, 3, a->P3, 64, 0, a->B3a, 64, 0, P(a->pB3a), nullptr, s);
  conv(nn, a, 7, 3, a->B3a, 64, 0, a->B3b, 64, 0, P(a->pB3b), nullptr, s);
  up(a, 2, a->B3b, 64, 0, a->CAT2, 128, 0, 64, s);
  conv(nn, a, 8, 2, a->CAT2, 128, 0, a->D2a, 64, 0, P(a->pD2a), nullptr, s);
  conv(nn, a, 9, 2, a->D2a, 64, 0, a->D2b, 64, 0, P(a->pD2b), nullptr, s);
  up(a, 1, a->D2b, 64, 0, a->CAT1, 96, 0, 64, s);
  conv(nn, a, 10, 1, a->CAT1, 96, 0, a->D1a, 32, 0, P(a->pD1a), nullptr, s);
  conv(nn, a, 11, 1, a->D1a, 32, 0, a->D1b, 32, 0, P(a->pD1b), nullptr, s);
  up(a, 0, a->D1b, 32, 0, a->CAT0, 48, 0, 32, s);
  conv(nn, a, 12, 0, a->CAT0, 48, 0, a->D0a, 16, 0, P(a->pD0a), nullptr, s);
  conv(nn, a, 13, 0, a->D0a, 16, 0, a->D0b, 16, 0, P(a->pD0b), nullptr, s);
  conv(nn, a, 14, 0, a->D0b, 16, 0, nullptr, 16, 0, nullptr, a->zout, s);
}

// backward from a->dOUT (channel 0 = dL/dz); fills nn->grads
int backward(pb_nn* nn, Acts* a, hipStream_t s) {
  int rc = 0;
  rc |= wgrad(nn, a, 14, 0, a->D0b, 16, 0, a->dOUT, 16, 0, s);
  dgrad(nn, a, 14, 0, a->dOUT, 16, 0, a->dD0b, 16, 0, s);
  act_bwd(a, 0, a->dD0b, 16, 0, 16, a->pD0b, 16, 0, s);
  rc |= wgrad(nn, a, 13, 0, a->D0a, 16, 0, a->dD0b, 16, 0, s);
  dgrad(nn, a, 13, 0, a->dD0b, 16, 0, a->dD0a, 16, 0, s);
  act_bwd(a, 0, a->dD0a, 16, 0, 16, a->pD0a, 16, 0, s);
  rc |= wgrad(nn, a, 12, 0, a->CAT0, 48, 0, a->dD0a, 16, 0, s);
  dgrad(nn, a, 12, 0, a->dD0a, 16, 0, a->dCAT0, 48, 0, s);
  up_bwd(a, 0, a->dCAT0, 48, 0, a->dD1b, 32, 0, 32, s);
  act_bwd(a, 1, a->dD1b, 32, 0, 32, a->pD1b, 32, 0, s);
  rc |= wgrad(nn, a, 11, 1, a->D1a, 32, 0, a->dD1b, 32, 0, s);
  dgrad(nn, a, 11, 1, a->dD1b, 32, 0, a->dD1a, 32, 0, s);
  act_bwd(a, 1, a->dD1a, 32, 0, 32, a->pD1a, 32, 0, s);
  rc |= wgrad(nn, a, 10, 1, a->CAT1, 96, 0, a->dD1a, 32, 0, s);
  dgrad(nn, a, 10, 1, a->dD1a, 32, 0, a->dCAT1, 96, 0, s);
  up_bwd(a, 1, a->dCAT1, 96, 0, a->dD2b, 64, 0, 64, s);
  act_bwd(a, 2, a->dD2b, 64, 0, 64, a->pD2b, 64, 0, s);
  rc |= wgrad(nn, a, 9, 2, a->D2a, 64, 0, a->dD2b, 64, 0, s);
  dgrad(nn, a, 9, 2, a->dD2b, 64, 0, a->dD2a, 64, 0, s);
  act_bwd(a, 2, a->dD2a, 64, 0, 64, a->pD2a, 64, 0, s);
  rc |= wgrad(nn, a, 8, 2, a->CAT2, 128, 0, a->dD2a, 64, 0, s);
  dgrad(nn, a, 8, 2, a->dD2a, 64, 0, a->dCAT2, 128, 0, s);
  up_bwd(a, 2, a->dCAT2, 128, 0, a->dB3b, 64, 0, 64, s);
  act_bwd(a, 3, a->dB3b, 64, 0, 64, a->pB3b, 64, 0, s);
  rc |= wgrad(nn, a, 7, 3, a->B3a, 64, 0, a->dB3b, 64, 0, s);
  dgrad(nn, a, 7, 3, a->dB3b, 64, 0, a->dB3a, 64, 0, s);
  act_bwd(a, 3, a->dB3a, 64, 0, 64, a->pB3a, 64, 0, s);
  rc |= wgrad(nn, a, 6, 3, a->P3, 64, 0, a->dB3a, 64, 0, s);
  dgrad(nn, a, 6, 3, a->dB3a, 64, 0, a->dP3, 64, 0, s);
  down_bwd(a, 2, a->dP3, 64, 0, a->dCAT2, 128, 64, 64, s);
  act_bwd(a, 2, a->dCAT2, 128, 64, 64, a->pE2b, 64, 0, s);
  rc |= wgrad(nn, a, 5, 2, a->E2a, 64, 0, a->dCAT2, 128, 64, s);
  dgrad(nn, a, 5, 2, a->dCAT2, 128, 64, a->dE2a, 64, 0, s);
  act_bwd(a, 2, a->dE2a, 64, 0, 64, a->pE2a, 64, 0, s);
  rc |= wgrad(nn, a, 4, 2, a->P2, 32, 0, a->dE2a, 64, 0, s);
  dgrad(nn, a, 4, 2, a->dE2a, 64, 0, a->dP2, 32, 0, s);
  down_bwd(a, 1, a->dP2, 32, 0, a->dCAT1, 96, 64, 32, s);
  act_bwd(a, 1, a->dCAT1, 96, 64, 32, a->pE1b, 32, 0, s);
  rc |= wgrad(nn, a, 3, 1, a->E1a, 32, 0, a->dCAT1, 96, 64, s);
  dgrad(nn, a, 3, 1, a->dCAT1, 96, 64, a->dE1a, 32, 0, s);
  act_bwd(a, 1, a->dE1a, 32, 0, 32, a->pE1a, 32, 0, s);
  rc |= wgrad(nn, a, 2, 1, a->P1, 16, 0, a->dE1a, 32, 0, s);
  dgrad(nn, a, 2, 1, a->dE1a, 32, 0, a->dP1, 16, 0, s);
  down_bwd(a, 0, a->dP1, 16, 0, a->dCAT0, 48, 32, 16, s);
  act_bwd(a, 0, a->dCAT0, 48, 32, 16, a->pE0b, 16, 0, s);
  rc |= wgrad(nn, a, 1, 0, a->E0a, 16, 0, a->dCAT0, 48, 32, s);
  dgrad(nn, a, 1, 0, a->dCAT0, 48, 32, a->dE0a, 16, 0, s);
  act_bwd(a, 0, a->dE0a, 16, 0, 16, a->pE0a, 16, 0, s);
  rc |= wgrad(nn, a, 0, 0, a->X0, 16, 0, a->dE0a, 16, 0, s);
  PB_LAUNCH_CHECK(nn->ctx);
  return rc;
}

// per-image rms of v -> a->scale
template <typename Tv>
void rms_scale(Acts* a, const Tv* v, hipStream_t s) {
  const long np = (long)a->h * a->w;
  hipLaunchKernelGGL(k_img_sums<Tv>, dim3(64, a->nimg), dim3(256), 0, s, v, np, a->w, 0, a->part, 64);
  hipLaunchKernelGGL(k_scale_from, dim3(1), dim3(a->nimg), 0, s, (const double*)a->part, 64, np, a->scale);
}

// coefficient features of an operator at its nodes
__global__ void k_op_features(int nx, int ny, const double* wxx, const double* wyy, const double* dinv, float* f1,
                              float* f2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= nx || j >= ny) return;
  const long c00 = (long)j * (nx + 1) + i, c01 = c00 + 1, c10 = c00 + nx + 1, c11 = c10 + 1;
  const double d = fabs(dinv[(long)j * nx + i]);
  f1[(long)j * nx + i] = (float)(d * ((wxx[c00] + wxx[c01]) + (wxx[c10] + wxx[c11])));
  f2[(long)j * nx + i] = (float)(d * ((wyy[c00] + wyy[c01]) + (wyy[c10] + wyy[c11])));
}

int op_features(pb_nn* nn, const pb_op2d* op, hipStream_t s) {
  if (nn->feat_op == op && nn->feat_version == op->version && nn->feat_n == op->n) return PB_OK;
  if (nn->feat_cap < op->n) {
    int rc = nn_malloc(nn, nn->allocs, &nn->f1, op->n);
    rc |= nn_malloc(nn, nn->allocs, &nn->f2, op->n);
    if (rc) return rc;
    nn->feat_cap = op->n;
  }
  hipLaunchKernelGGL(k_op_features, dim3(ceil_div(op->g.nx, 256), op->g.ny), dim3(256), 0, s, op->g.nx, op->g.ny,
                     op->wxx, op->wyy, op->dinv, nn->f1, nn->f2);
  PB_LAUNCH_CHECK(nn->ctx);
  nn->feat_op = op;
  nn->feat_version = op->version;
  nn->feat_n = op->n;
  return PB_OK;
}

// The network's input formed from right-hand sides on a multigrid level's
// unique grid (uny x unx; a periodic axis has no duplicated end node): the
// image's (nh, nw) pixels with the end nodes duplicated, its rms (floored),
// and the input channels r / scale, f1, f2 -- the values k_img_sums,
// k_scale_from and k_input give for the duplicated image, with the same
// sums bit for bit: 64 chunks of the image, each the 256-lane tree of
// k_img_sums (with at most 256 pixels per chunk its first two levels are the
// pair sums below, then a 64-lane butterfly adds the same pairs as the
// tree), then the chunk sums in order.  Block (x, b) writes pixels
// [256 x, 256 x + 256) of image b; every block of an image forms the same
// sum, block 0 stores the scale.
constexpr int kNetInThreads = 256;
constexpr int kNetChunks = 64;   // k_img_sums' chunks per image
__global__ __launch_bounds__(kNetInThreads) void k_net_in(const float* v, int uny, int unx, int nh, int nw,
                                                        const float* f1, const float* f2, float* scale, bf16* x0) {
  __shared__ double csum[kNetChunks];
  __shared__ float sc;
  const int b = blockIdx.y, t = threadIdx.x, lane = t & (kWave - 1), wid = t / kWave;
  const int np = nh * nw;
  const float* vb = v + (long)b * uny * unx;
  // row of pixel p: a float quotient corrected by one either way (exact for
  // the at most 2^14 pixels of a level), not an integer division per value
  const float rnw = 1.f / (float)nw;
  auto src = [&](int p) {
    int j = (int)((float)p * rnw);
    j += (j + 1) * nw <= p ? 1 : 0;
    j -= j * nw > p ? 1 : 0;
    const int i = p - j * nw;
    return (j >= uny ? j - uny : j) * unx + (i >= unx ? i - unx : i);
  };
  const int p = blockIdx.x * kNetInThreads + t;
  const int chunk = (np + kNetChunks - 1) / kNetChunks;   // <= 256 (checked on the host)
  // Every load of the thread first (its pixel, its features, and lane l's
  // values of the wave's 16 chunks, 0 past a chunk's end), then the sums: one
  // round trip instead of one per chunk.
  constexpr int CPW = kNetChunks / (kNetInThreads / kWave);   // chunks per wave
  static_assert(kNetChunks % (kNetInThreads / kWave) == 0, "the chunks divide evenly over the waves");
  const float mine = p < np ? vb[src(p)] : 0.f;
  const float g1 = p < np ? f1[p] : 0.f, g2 = p < np ? f2[p] : 0.f;
  // lane offsets 0, 128, 64, 192 (the tree's pair order); all 64 loads are
  // issued unconditionally at clamped pixels, with no branch between them,
  // and the values past a chunk's end are zeroed after the last load
  float xv[CPW][4];
#pragma unroll
  for (int m = 0; m < CPW; ++m)
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int off = u == 1 ? 128 : u == 2 ? 64 : u == 3 ? 192 : 0;
      const int q = (wid + m * (kNetInThreads / kWave)) * chunk + lane + off;
      xv[m][u] = vb[src(min(q, np - 1))];
    }
#pragma unroll
  for (int m = 0; m < CPW; ++m) {
    const int beg = (wid + m * (kNetInThreads / kWave)) * chunk, end = min(beg + chunk, np);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int off = u == 1 ? 128 : u == 2 ? 64 : u == 3 ? 192 : 0;
      if (off >= chunk || beg + lane + off >= end) xv[m][u] = 0.f;
    }
  }
  // lane l's k_img_sums values: its pixels squared (exact: a float's square
  // has at most 48 significant bits), or 0; tree levels 128 and 64:
  // L[l] = (L[l] + L[l + 128]) + (L[l + 64] + L[l + 192]); then each chunk's
  // 64-lane butterfly, level by level over all chunks of the wave (the
  // chunks' shuffles of a level in flight together)
  double sm[CPW];
#pragma unroll
  for (int m = 0; m < CPW; ++m) {
    auto sq = [&](int u) {
      const double x = (double)xv[m][u];
      return x * x;
    };
    sm[m] = (sq(0) + sq(1)) + (sq(2) + sq(3));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int m = 0; m < CPW; ++m) sm[m] += __shfl_xor(sm[m], off, kWave);
  if (lane == 0)
#pragma unroll
    for (int m = 0; m < CPW; ++m) csum[wid + m * (kNetInThreads / kWave)] = sm[m];
  __syncthreads();
  if (t == 0) {
    double tot = 0.0;
    for (int c = 0; c < kNetChunks; ++c) tot += csum[c];
    const double rms = sqrt(tot / (double)np);
    sc = (float)(rms > 1e-300 ? rms : 1.0);
    if (blockIdx.x == 0) scale[b] = sc;
  }
  __syncthreads();
  if (p >= np) return;
  union {
    bf16 h[16];
    uint4 q[2];
  } o;
  o.h[0] = f2bf((float)((double)mine / (double)sc));
  o.h[1] = f2bf(g1);
  o.h[2] = f2bf(g2);
#pragma unroll
  for (int c = 3; c < 16; ++c) o.h[c] = 0;
  uint4* dst = reinterpret_cast<uint4*>(x0 + ((long)b * np + p) * 16);
  dst[0] = o.q[0];
  dst[1] = o.q[1];
}

}  // namespace

// The network applied to a batch of ``nimg`` right-hand sides v on a
// multigrid level's unique grid (uny x unx nodes, float [nimg][uny][unx];
// the network's nh x nw grid duplicates the end node of a periodic axis)
// that share one pair of feature fields f1, f2 float [nh][nw] (device): one
// launch forms the input (the duplicated image, its rms scale and the input
// channels v / scale, f1, f2), then the forward.  On return *zout holds the
// network's output z, float [nimg][nh][nw], and *scale the per-image rms
// (device buffers of the network, valid until its next call); the result of
// the batch is scale * z, which the multigrid's coarse-level kernel forms.
int pb_nn_forward_level(pb_nn* nn, int nimg, int uny, int unx, int nh, int nw, const float* v, const float* f1,
                        const float* f2, const float** zout, const float** scale, hipStream_t s) {
  if ((nh & 1) == 0 || (nw & 1) == 0 || ((nh - 1) & 7) || ((nw - 1) & 7) || nh < 9 || nw < 9)
    return pb_fail(nn->ctx, PB_ERR_ARG, "network: a %d x %d grid is not 8k + 1 nodes per axis", nh, nw);
  if ((long)nh * nw > 256L * kNetChunks)
    return pb_fail(nn->ctx, PB_ERR_ARG, "network: a %d x %d level exceeds %d pixels", nh, nw, 256 * kNetChunks);
  if (nh < uny || nh > uny + 1 || nw < unx || nw > unx + 1)
    return pb_fail(nn->ctx, PB_ERR_ARG, "network: a %d x %d grid does not duplicate the end nodes of %d x %d", nh, nw,
                   uny, unx);
  int rc = 0;
  Acts* a = get_shape(nn, nimg, nh, nw, false, &rc);
  if (!a) return rc;
  hipLaunchKernelGGL(k_net_in, dim3(ceil_div(nh * nw, kNetInThreads), nimg), dim3(kNetInThreads), 0, s, v, uny, unx,
                     nh, nw, f1, f2, a->scale, a->X0);
  forward(nn, a, s);
  PB_LAUNCH_CHECK(nn->ctx);
  *zout = a->zout;
  *scale = a->scale;
  return PB_OK;
}

// attach bookkeeping: an operator records its network, the network the
// operators it is attached to, so that destroying either side detaches
void pb_nn_track(pb_nn* nn, pb_op2d* op, bool add) {
  if (!nn) return;
  auto& v = nn->attached;
  for (size_t i = 0; i < v.size(); ++i)
    if (v[i] == op) {
      if (add) return;
      v.erase(v.begin() + (long)i);
      return;
    }
  if (add) v.push_back(op);
}

extern "C" int pb_nn_create(pb_ctx* ctx, unsigned long long seed, pb_nn** out) {
  if (!ctx || !out) return pb_fail(ctx, PB_ERR_ARG, "pb_nn_create: null argument");
  pb_nn* nn = new pb_nn();
  nn->ctx = ctx;
  long off = 0;
  for (int l = 0; l < 15; ++l) {
    nn->woff[l] = off;
    off += 9L * kLayers[l].cin_v * kLayers[l].cout_v;
    nn->boff[l] = off;
    off += kLayers[l].cout_v;
  }
  nn->nparam = off;
  int rc = nn_malloc(nn, nn->allocs, &nn->params, off);
  rc |= nn_malloc(nn, nn->allocs, &nn->grads, off);
  rc |= nn_malloc(nn, nn->allocs, &nn->adam_m, off);
  rc |= nn_malloc(nn, nn->allocs, &nn->adam_v, off);
  for (int l = 0; l < 15; ++l) {
    const long sz = 9L * pad16(kLayers[l].cin_v) * pad16(kLayers[l].cout_v);
    rc |= nn_malloc(nn, nn->allocs, &nn->wf[l], sz);
    rc |= nn_malloc(nn, nn->allocs, &nn->wb[l], sz);
  }
  nn->wpart_cap = 2048L * 9 * 16 * 16 + 9L * 128 * 64;
  rc |= nn_malloc(nn, nn->allocs, &nn->wpart, nn->wpart_cap);
  rc |= nn_malloc(nn, nn->allocs, &nn->bpart, 1024L * 64);
  if (rc) {
    pb_nn_destroy(nn);
    return PB_ERR_HIP;
  }
  // He-style initialisation (fixed-seed LCG on the host), zero biases
  std::vector<float> h(off, 0.f);
  unsigned long long st = seed * 6364136223846793005ULL + 1442695040888963407ULL;
  auto uni = [&]() {
    st = st * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(st >> 11) / 9007199254740992.0;
  };
  for (int l = 0; l < 15; ++l) {
    const double fan_in = 9.0 * kLayers[l].cin_v;
    const double bound = sqrt(6.0 / fan_in) * (l == 14 ? 0.1 : 1.0);
    for (long k = 0; k < 9L * kLayers[l].cin_v * kLayers[l].cout_v; ++k)
      h[nn->woff[l] + k] = (float)((2.0 * uni() - 1.0) * bound);
  }
  hipStream_t s = ctx->own_stream;
  if (hipMemcpyAsync(nn->params, h.data(), sizeof(float) * off, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemsetAsync(nn->adam_m, 0, sizeof(float) * off, s) != hipSuccess ||
      hipMemsetAsync(nn->adam_v, 0, sizeof(float) * off, s) != hipSuccess) {
    pb_nn_destroy(nn);
    return pb_fail(ctx, PB_ERR_HIP, "pb_nn_create: initial copy failed");
  }
  rc = pack_weights(nn, s);
  if (rc || hipStreamSynchronize(s) != hipSuccess) {
    pb_nn_destroy(nn);
    return PB_ERR_HIP;
  }
  *out = nn;
  return PB_OK;
}

extern "C" int pb_nn_destroy(pb_nn* nn) {
  if (!nn) return PB_OK;
  if (nn->ctx) {
    (void)hipStreamSynchronize(nn->ctx->stream);
    (void)hipStreamSynchronize(nn->ctx->own_stream);
  }
  for (pb_op2d* op : nn->attached)   // no operator keeps a dangling network
    if (op->nn == nn) op->nn = nullptr;
  nn->attached.clear();
  for (Acts* a : nn->shapes) {
    for (void* p : a->allocs) (void)hipFree(p);
    delete a;
  }
  for (void* p : nn->allocs) (void)hipFree(p);
  delete nn;
  return PB_OK;
}

extern "C" long pb_nn_num_params(const pb_nn* nn) { return nn ? nn->nparam : -1; }

extern "C" int pb_nn_get_params(pb_nn* nn, float* host, float* grads_host) {
  if (!nn || (!host && !grads_host)) return PB_ERR_ARG;
  hipStream_t s = nn->ctx->stream;
  if (host) PB_HIP(nn->ctx, hipMemcpyAsync(host, nn->params, sizeof(float) * nn->nparam, hipMemcpyDeviceToHost, s));
  if (grads_host)
    PB_HIP(nn->ctx, hipMemcpyAsync(grads_host, nn->grads, sizeof(float) * nn->nparam, hipMemcpyDeviceToHost, s));
  PB_HIP(nn->ctx, hipStreamSynchronize(s));
  return PB_OK;
}

extern "C" int pb_nn_set_params(pb_nn* nn, const float* host) {
  if (!nn || !host) return PB_ERR_ARG;
  hipStream_t s = nn->ctx->stream;
  PB_HIP(nn->ctx, hipMemcpyAsync(nn->params, host, sizeof(float) * nn->nparam, hipMemcpyHostToDevice, s));
  int rc = pack_weights(nn, s);
  if (rc) return rc;
  PB_HIP(nn->ctx, hipStreamSynchronize(s));
  nn->adam_t = 0;
  return PB_OK;
}

// Forward (and optionally backward + Adam step) on a batch of nimg images of
// h x w nodes.  inputs: float [nimg][h*w][3] (r, feature 1, feature 2);
// target: float [nimg][h*w] (solution / rms(r), null-space free) or NULL.
// img_weight: float [nimg] per-image loss weights or NULL (all 1).
// out_z: float [nimg][h*w] network output (unscaled) or NULL.  lr > 0 applies
// an Adam step; lr == 0 only computes gradients.  loss: weighted mean squared
// error sum_b w_b sum_p d^2 / (nimg h w).
extern "C" int pb_nn_train_step(pb_nn* nn, int nimg, int h, int w, const float* inputs, const float* target,
                                const float* img_weight, double lr, float* out_z, double* loss) {
  if (!nn || !inputs || nimg < 1 || nimg > 256 || h < 9 || w < 9)
    return pb_fail(nn ? nn->ctx : nullptr, PB_ERR_ARG, "pb_nn_train_step: bad argument");
  if ((h & 1) == 0 || (w & 1) == 0) return pb_fail(nn->ctx, PB_ERR_ARG, "pb_nn_train_step: h and w must be odd");
  hipStream_t s = nn->ctx->stream;
  const bool train = target != nullptr;
  int rc = 0;
  Acts* a = get_shape(nn, nimg, h, w, train, &rc);
  if (!a) return rc;
  const long np = (long)h * w, n = np * nimg;
  // inputs (already normalised by the caller) into the 16-channel X0
  hipLaunchKernelGGL(k_split3, dim3(gs(n)), dim3(256), 0, s, inputs, n, a->X0);
  forward(nn, a, s);
  if (out_z) PB_HIP(nn->ctx, hipMemcpyAsync(out_z, a->zout, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
  if (train) {
    hipLaunchKernelGGL(k_img_sums<float>, dim3(64, nimg), dim3(256), 0, s, (const float*)a->zout, np, w, 1, a->part, 64);
    hipLaunchKernelGGL(k_null_coef, dim3(1), dim3(nimg), 0, s, (const double*)a->part, 64, np, w, h, nn->null_mode,
                       a->coef);
    const int nbl = 1024;
    hipLaunchKernelGGL(k_loss, dim3(nbl), dim3(256), 0, s, (const float*)a->zout, target, (const double*)a->coef,
                       img_weight, np, w, nimg, (float)(1.0 / (double)n), a->dOUT, a->lpart);
    rc = backward(nn, a, s);
    if (rc) return rc;
    if (lr > 0) {
      nn->adam_t += 1;
      const float b1 = 0.9f, b2 = 0.999f;
      const float bc1 = 1.f - powf(b1, (float)nn->adam_t), bc2 = 1.f - powf(b2, (float)nn->adam_t);
      hipLaunchKernelGGL(k_adam, dim3(gs(nn->nparam)), dim3(256), 0, s, nn->nparam, nn->params, (const float*)nn->grads,
                         nn->adam_m, nn->adam_v, (float)lr, b1, b2, 1e-8f, bc1, bc2);
      rc = pack_weights(nn, s);
      if (rc) return rc;
    }
    if (loss) {
      std::vector<double> lp(nbl);
      PB_HIP(nn->ctx, hipMemcpyAsync(lp.data(), a->lpart, sizeof(double) * nbl, hipMemcpyDeviceToHost, s));
      PB_HIP(nn->ctx, hipStreamSynchronize(s));
      double t = 0.0;
      for (double v : lp) t += v;
      *loss = t / (double)n;
    }
  }
  PB_LAUNCH_CHECK(nn->ctx);
  PB_HIP(nn->ctx, hipStreamSynchronize(s));
  return PB_OK;
}

// Network preconditioner on an operator's grid: z = rms(r) * net(r / rms(r),
// features of op).  r, z: device FP64 vectors of op->n values.
extern "C" int pb_nn_apply(pb_nn* nn, const pb_op2d* op, const double* r, double* z) {
  if (!nn || !op || !r || !z) return pb_fail(nn ? nn->ctx : nullptr, PB_ERR_ARG, "pb_nn_apply: null argument");
  if ((op->g.nx & 1) == 0 || (op->g.ny & 1) == 0)
    return pb_fail(nn->ctx, PB_ERR_ARG, "pb_nn_apply: grid dimensions must be odd (2^k + 1 nodes)");
  hipStream_t s = nn->ctx->stream;
  int rc = op_features(nn, op, s);
  if (rc) return rc;
  Acts* a = get_shape(nn, 1, op->g.ny, op->g.nx, false, &rc);
  if (!a) return rc;
  const long np = op->n;
  rms_scale<double>(a, r, s);
  hipLaunchKernelGGL(k_input<double>, dim3(gs(np)), dim3(256), 0, s, r, (const float*)a->scale, (const float*)nn->f1,
                     (const float*)nn->f2, 0, np, 1, a->X0);
  forward(nn, a, s);
  hipLaunchKernelGGL(k_output<double>, dim3(gs(np)), dim3(256), 0, s, (const float*)a->zout, (const float*)a->scale,
                     np, 1, z);
  PB_LAUNCH_CHECK(nn->ctx);
  PB_HIP(nn->ctx, hipStreamSynchronize(s));
  return PB_OK;
}

// Null space the training loss removes from the output before comparing it
// with the target: 0 = {constant, checkerboard} (the 9-point operator's
// pseudo-incompressible null space), 1 = {constant} (a coarse level's).
extern "C" int pb_nn_set_loss_nullspace(pb_nn* nn, int mode) {
  if (!nn || (mode != 0 && mode != 1)) return pb_fail(nn ? nn->ctx : nullptr, PB_ERR_ARG, "bad null-space mode");
  nn->null_mode = mode;
  return PB_OK;
}

// Coefficient features (two float arrays of op->n values) of an operator.
extern "C" int pb_nn_features(pb_nn* nn, const pb_op2d* op, float* f1, float* f2) {
  if (!nn || !op || !f1 || !f2) return pb_fail(nn ? nn->ctx : nullptr, PB_ERR_ARG, "pb_nn_features: null argument");
  hipStream_t s = nn->ctx->stream;
  int rc = op_features(nn, op, s);
  if (rc) return rc;
  PB_HIP(nn->ctx, hipMemcpyAsync(f1, nn->f1, sizeof(float) * op->n, hipMemcpyDeviceToDevice, s));
  PB_HIP(nn->ctx, hipMemcpyAsync(f2, nn->f2, sizeof(float) * op->n, hipMemcpyDeviceToDevice, s));
  PB_HIP(nn->ctx, hipStreamSynchronize(s));
  return PB_OK;
}
