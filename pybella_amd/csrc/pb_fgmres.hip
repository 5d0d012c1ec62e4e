// Flexible GMRES preconditioned by the two-branch multigrid cycle with the
// learned coarse solver (gfx950).  No reference counterpart: it replaces the
// reference's bicgstab call (second_projection.py:230) in the learned mode.
//
// FGMRES(m) (Saad 1993, "A flexible inner-outer preconditioned GMRES") in
// FP64 on the scaled system (D^-1 A) x = b: the basis V and the Hessenberg
// matrix are FP64; the preconditioned directions Z_j = M(V_j) come from the
// FP32 multigrid cycle (pb_mg2d.hpp) whose coarse level is solved by the BF16
// network.  Because the method is flexible, x = x0 + Z y is exact for the
// Z actually produced, so the FP32 / BF16 preconditioner limits only the
// convergence rate, not the attainable FP64 residual.
//
// Each Arnoldi step is the cycle, w = A Z_j (FP64), classical Gram-Schmidt
// applied twice (Giraud et al. 2005) and the normalisation; the Givens
// rotations and the least-squares residual estimate |g_{j+1}| are formed on
// the device by the normalisation kernel's first thread.  Reductions are
// per-block partial records (wave-shuffle block sums) summed in a fixed order
// by a one-block-per-column kernel, or, for the norm, by every block of the
// normalisation (deterministic, no grid-wide fence).  The host reads the
// estimate one iteration behind the device (pinned copy + event), so it never
// waits between iterations; it waits at once only when the convergence rate
// predicts that the step just issued reaches the tolerance.  After a cycle,
// x is updated on the device, the FP64 true residual is recomputed, and a
// cycle that did not reduce it is undone.
#include <algorithm>
#include <cmath>

#include "pb_mg2d.hpp"

using namespace pb2d;

void pb_nn_track(pb_nn* nn, pb_op2d* op, bool add);

namespace {

constexpr int kFgBlocks = 1024;  // blocks of the vector passes (each writes one partial record)
constexpr int kFgThreads = 256;
constexpr int kMaxM = 60;        // restart length cap (K = m + 1 <= 64 column sums per record)

}  // namespace

struct FgWork {
  int m = 0;
  long n = 0;
  double* V = nullptr;       // (m + 1) x n basis
  float* Z = nullptr;        // m x n preconditioned directions
  double* w = nullptr;       // n
  double* w2 = nullptr;      // n
  double* part = nullptr;    // [m + 1][kFgBlocks] partial records, one column per block
  double* part3 = nullptr;   // [kFgBlocks] partial ||w||^2
  double* H = nullptr;       // (m + 1) x m rotated Hessenberg, column-major
  double* cs = nullptr;      // m
  double* sn = nullptr;      // m
  double* g = nullptr;       // m + 1
  double* h1 = nullptr;      // m + 1 (first Gram-Schmidt pass)
  double* h2 = nullptr;      // m + 1 (second pass)
  double* est = nullptr;     // m residual estimates |g_{j+1}|
  double* y = nullptr;       // m least-squares solution
  double* hest = nullptr;    // pinned host copy of est
  hipEvent_t ev[kMaxM] = {};
  std::vector<void*> allocs;
};

namespace {

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// w = (D^-1 A) z in FP64 for an FP32 direction z
struct ZIn {
  using Raw = float;
  const float* z;
  __device__ Raw load(long g) const { return z[g]; }
  __device__ double combine(const Raw& q) const { return (double)q; }
};
struct WOut {
  struct Pre {};
  double* w;
  __device__ Pre prefetch(long) const { return Pre{}; }
  template <class R>
  __device__ void operator()(int, long g, double y, double, const Pre&, const R&) {
    w[g] = y;
  }
};
template <bool CROSS, bool CENTER, int RY>
__global__ __launch_bounds__(NTHR) void kf_matvec(OpDev<double> op, const float* z, double* w) {
  __shared__ TileLds<double, RY> L;
  ZIn in{z};
  WOut out{w};
  stencil_tile2<double, CROSS, CENTER, RY>(op, blockIdx.x, blockIdx.y, in, out, L, [] { return true; });
}

__device__ __forceinline__ long gstride() { return (long)gridDim.x * blockDim.x; }
__device__ __forceinline__ long gfirst() { return blockIdx.x * (long)blockDim.x + threadIdx.x; }

// this block's partial dots of V_0..V_{K-1} with x, stored as entry
// blockIdx.x of rows 0..K-1 of ``part`` ([K][kFgBlocks]); eight vectors per
// pass over x, wave-shuffle block sums
__device__ __forceinline__ void block_dots(const double* V, long n, int K, const double* x, double* part,
                                           double* sh) {
  for (int k0 = 0; k0 < K; k0 += 8) {
    double acc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = 0.0;
    for (long g = gfirst(); g < n; g += gstride()) {
      const double xg = x[g];
      // unconditional loads (vector index clamped to K - 1), the sums guarded:
      // a load behind the guard would wait for its result before the next
      double vq[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) vq[q] = V[(long)min(k0 + q, K - 1) * n + g];
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (k0 + q < K) acc[q] += vq[q] * xg;
    }
    block_sum_w<8, kFgThreads / kWave>(acc, sh);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (k0 + q < K) part[(long)(k0 + q) * kFgBlocks + blockIdx.x] = acc[q];
    }
    __syncthreads();   // sh is reused by the next pass
  }
}

// first Gram-Schmidt pass: partial dots V^T w
__global__ __launch_bounds__(kFgThreads) void kf_dots(const double* V, long n, int K, const double* w, double* part) {
  __shared__ double sh[8 * (kFgThreads / kWave)];
  block_dots(V, n, K, w, part, sh);
}

// h[k] = sum of row k of part (one block per row, fixed order: thread t adds
// entries t, t + 256, ... then a wave-shuffle block sum)
__global__ __launch_bounds__(kFgThreads) void kf_colsum(const double* part, double* h) {
  __shared__ double sh[kFgThreads / kWave];
  const double* row = part + (long)blockIdx.x * kFgBlocks;
  double v[1] = {0.0};
  for (int b = threadIdx.x; b < kFgBlocks; b += kFgThreads) v[0] += row[b];
  block_sum_w<1, kFgThreads / kWave>(v, sh);
  if (threadIdx.x == 0) h[blockIdx.x] = v[0];
}

// w2 = w - V h1, partial dots V^T w2 (the second Gram-Schmidt pass)
__global__ __launch_bounds__(kFgThreads) void kf_orth1(const double* V, long n, int K, const double* h1,
                                                       const double* w, double* w2, double* part) {
  __shared__ double h[kMaxM + 4];
  __shared__ double sh[8 * (kFgThreads / kWave)];
  if ((int)threadIdx.x < K) h[threadIdx.x] = h1[threadIdx.x];
  __syncthreads();
  for (long g = gfirst(); g < n; g += gstride()) {
    double a = w[g];
    for (int k = 0; k < K; ++k) a -= h[k] * V[(long)k * n + g];
    w2[g] = a;
  }
  block_dots(V, n, K, w2, part, sh);
}

// kf_orth1 in one pass over the basis for K <= KM: each element's K basis
// values stay in registers between forming w2 = w - V h1 and the partial
// dots V^T w2.  Per thread the products and sums are kf_orth1's, in the
// same order (the dot of basis vector k over the thread's elements in order,
// then block sums of eight values at a time), so the results are bitwise
// the same; the basis is read once instead of twice.
template <int KM>
__global__ __launch_bounds__(kFgThreads) void kf_orth1r(const double* V, long n, int K, const double* h1,
                                                        const double* w, double* w2, double* part) {
  __shared__ double h[kMaxM + 4];
  __shared__ double sh[8 * (kFgThreads / kWave)];
  if ((int)threadIdx.x < K) h[threadIdx.x] = h1[threadIdx.x];
  __syncthreads();
  double acc[KM];
#pragma unroll
  for (int k = 0; k < KM; ++k) acc[k] = 0.0;
  for (long g = gfirst(); g < n; g += gstride()) {
    double vk[KM];
#pragma unroll
    for (int k = 0; k < KM; ++k) vk[k] = V[(long)min(k, K - 1) * n + g];   // unconditional, used for k < K
    double a = w[g];
#pragma unroll
    for (int k = 0; k < KM; ++k)
      if (k < K) a -= h[k] * vk[k];
    w2[g] = a;
#pragma unroll
    for (int k = 0; k < KM; ++k)
      if (k < K) acc[k] += vk[k] * a;
  }
#pragma unroll
  for (int k0 = 0; k0 < KM; k0 += 8) {
    if (k0 >= K) break;
    double s8[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) s8[q] = k0 + q < KM ? acc[k0 + q] : 0.0;   // KM may be 4
    block_sum_w<8, kFgThreads / kWave>(s8, sh);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (k0 + q < K) part[(long)(k0 + q) * kFgBlocks + blockIdx.x] = s8[q];
    }
    __syncthreads();   // sh is reused by the next pass
  }
}

// kf_orth2 for K <= KM with the element's K basis values loaded together
// (unconditional loads at clamped vector indices) before the sum, in
// kf_orth2's order: bitwise the same
template <int KM>
__global__ __launch_bounds__(kFgThreads) void kf_orth2r(const double* V, long n, int K, const double* h2,
                                                        const double* w2, double* w, double* part3) {
  __shared__ double h[kMaxM + 4];
  __shared__ double sh[kFgThreads / kWave];
  if ((int)threadIdx.x < K) h[threadIdx.x] = h2[threadIdx.x];
  __syncthreads();
  double ss[1] = {0.0};
  for (long g = gfirst(); g < n; g += gstride()) {
    double vk[KM];
#pragma unroll
    for (int k = 0; k < KM; ++k) vk[k] = V[(long)min(k, K - 1) * n + g];
    double a = w2[g];
#pragma unroll
    for (int k = 0; k < KM; ++k)
      if (k < K) a -= h[k] * vk[k];
    w[g] = a;
    ss[0] += a * a;
  }
  block_sum_w<1, kFgThreads / kWave>(ss, sh);
  if (threadIdx.x == 0) part3[blockIdx.x] = ss[0];
}

// w = w2 - V h2, partial ||w||^2
__global__ __launch_bounds__(kFgThreads) void kf_orth2(const double* V, long n, int K, const double* h2,
                                                       const double* w2, double* w, double* part3) {
  __shared__ double h[kMaxM + 4];
  __shared__ double sh[kFgThreads / kWave];
  if ((int)threadIdx.x < K) h[threadIdx.x] = h2[threadIdx.x];
  __syncthreads();
  double ss[1] = {0.0};
  for (long g = gfirst(); g < n; g += gstride()) {
    double a = w2[g];
    for (int k = 0; k < K; ++k) a -= h[k] * V[(long)k * n + g];
    w[g] = a;
    ss[0] += a * a;
  }
  block_sum_w<1, kFgThreads / kWave>(ss, sh);
  if (threadIdx.x == 0) part3[blockIdx.x] = ss[0];
}

// V_{j+1} = w / ||w||; the first thread of block 0 completes Hessenberg column
// j (h1 + h2, ||w||), applies the previous Givens rotations, forms rotation j
// and updates g; est[j] = |g_{j+1}|
__global__ __launch_bounds__(kFgThreads) void kf_norm(int j, int m, long n, const double* part3, const double* w,
                                                      double* vnext, const double* h1, const double* h2, double* H,
                                                      double* cs, double* sn, double* g, double* est) {
  __shared__ double sh[kFgThreads / kWave];
  __shared__ double hn_s;
  double ss[1] = {0.0};
  for (int b = threadIdx.x; b < kFgBlocks; b += kFgThreads) ss[0] += part3[b];
  block_sum_w<1, kFgThreads / kWave>(ss, sh);
  if (threadIdx.x == 0) hn_s = sqrt(ss[0]);
  __syncthreads();
  const double hn = hn_s;
  const double inv = hn > 0.0 ? 1.0 / hn : 0.0;
  for (long q = gfirst(); q < n; q += gstride()) vnext[q] = w[q] * inv;
  if (blockIdx.x == 0) {
    // the column and the previous rotations staged in LDS by the whole block
    // (parallel loads), rotated by one thread there, written back by all
    __shared__ double cols[kMaxM + 2], css[kMaxM], sns[kMaxM];
    for (int i = threadIdx.x; i <= j; i += kFgThreads) cols[i] = h1[i] + h2[i];
    for (int i = threadIdx.x; i < j; i += kFgThreads) {
      css[i] = cs[i];
      sns[i] = sn[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      cols[j + 1] = hn;
      for (int i = 0; i < j; ++i) {
        const double a = cols[i], b = cols[i + 1];
        cols[i] = css[i] * a + sns[i] * b;
        cols[i + 1] = -sns[i] * a + css[i] * b;
      }
      const double a = cols[j], b = cols[j + 1];
      const double r = hypot(a, b);
      const double c = r > 0.0 ? a / r : 1.0, sj = r > 0.0 ? b / r : 0.0;
      cs[j] = c;
      sn[j] = sj;
      cols[j] = r;
      cols[j + 1] = 0.0;
      const double gj = g[j];
      g[j + 1] = -sj * gj;
      g[j] = c * gj;
      est[j] = fabs(-sj * gj);
    }
    __syncthreads();
    double* col = H + (long)j * (m + 1);
    for (int i = threadIdx.x; i <= j + 1; i += kFgThreads) col[i] = cols[i];
  }
}

__global__ void k_fill_zero_d(long n, double* x) {
  for (long q = gfirst(); q < n; q += gstride()) x[q] = 0.0;
}

// V_0 = r / beta, g = (beta, 0, ...)
__global__ void kf_start(long n, int m, const double* r, double beta, double* v0, double* g) {
  const double inv = 1.0 / beta;
  for (long q = gfirst(); q < n; q += gstride()) v0[q] = r[q] * inv;
  if (blockIdx.x == 0) {
    for (int i = threadIdx.x; i <= m; i += blockDim.x) g[i] = i == 0 ? beta : 0.0;
  }
}

// back substitution of the k x k upper-triangular rotated Hessenberg: H y = g
__global__ void kf_backsub(int k, int m, const double* H, const double* g, double* y) {
  if (threadIdx.x != 0) return;
  for (int i = k - 1; i >= 0; --i) {
    double acc = g[i];
    for (int c = i + 1; c < k; ++c) acc -= H[(long)c * (m + 1) + i] * y[c];
    const double d = H[(long)i * (m + 1) + i];
    y[i] = d != 0.0 ? acc / d : 0.0;
  }
}

// x += sign * sum_{i < k} y_i Z_i
__global__ void kf_update(long n, int k, const double* y, double sign, const float* Z, double* x) {
  __shared__ double ys[kMaxM];
  for (int i = threadIdx.x; i < k; i += blockDim.x) ys[i] = sign * y[i];
  __syncthreads();
  for (long q = gfirst(); q < n; q += gstride()) {
    double acc = 0.0;
    for (int i = 0; i < k; ++i) acc += ys[i] * (double)Z[(long)i * n + q];
    x[q] += acc;
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
void fg_free(FgWork* f) {
  if (!f) return;
  for (void* p : f->allocs) (void)hipFree(p);
  if (f->hest) (void)hipHostFree(f->hest);
  for (hipEvent_t& e : f->ev)
    if (e) (void)hipEventDestroy(e);
  delete f;
}

int fg_alloc(pb_op2d* op, int m, FgWork** out) {
  if (op->fg && op->fg->m >= m && op->fg->n == op->n) {
    *out = op->fg;
    return PB_OK;
  }
  fg_free(op->fg);
  op->fg = nullptr;
  FgWork* f = new FgWork();
  f->m = m;
  f->n = op->n;
  auto A = [&](auto** p, long count) {
    void* q = nullptr;
    if (hipMalloc(&q, sizeof(**p) * (size_t)std::max<long>(count, 1)) != hipSuccess) return 1;
    f->allocs.push_back(q);
    *p = (std::remove_reference_t<decltype(*p)>)q;
    return 0;
  };
  const long n = op->n;
  int rc = 0;
  rc |= A(&f->V, (long)(m + 1) * n);
  rc |= A(&f->Z, (long)m * n);
  rc |= A(&f->w, n);
  rc |= A(&f->w2, n);
  rc |= A(&f->part, (long)kFgBlocks * (m + 1));
  rc |= A(&f->part3, kFgBlocks);
  rc |= A(&f->H, (long)(m + 1) * m);
  rc |= A(&f->cs, m);
  rc |= A(&f->sn, m);
  rc |= A(&f->g, m + 1);
  rc |= A(&f->h1, m + 1);
  rc |= A(&f->h2, m + 1);
  rc |= A(&f->est, m);
  rc |= A(&f->y, m);
  void* hp = nullptr;
  if (!rc && hipHostMalloc(&hp, sizeof(double) * m, hipHostMallocDefault) != hipSuccess) rc = 1;
  f->hest = (double*)hp;
  for (int i = 0; i < m && !rc; ++i)
    if (hipEventCreateWithFlags(&f->ev[i], hipEventDisableTiming) != hipSuccess) rc = 1;
  if (rc) {
    fg_free(f);
    return pb_fail(op->ctx, PB_ERR_HIP, "fgmres: workspace allocation failed");
  }
  op->fg = f;
  *out = f;
  return PB_OK;
}

void launch_fg_matvec(const pb_op2d* op, const float* z, double* w, hipStream_t s) {
  const OpDev<double> d = opdev<double>(op);
  const dim3 grd = stencil_grid(op, 4), blk(TX, TYT);
  if (op->cross && op->center) hipLaunchKernelGGL((kf_matvec<true, true, 4>), grd, blk, 0, s, d, z, w);
  else if (op->cross) hipLaunchKernelGGL((kf_matvec<true, false, 4>), grd, blk, 0, s, d, z, w);
  else if (op->center) hipLaunchKernelGGL((kf_matvec<false, true, 4>), grd, blk, 0, s, d, z, w);
  else hipLaunchKernelGGL((kf_matvec<false, false, 4>), grd, blk, 0, s, d, z, w);
}

// One Arnoldi step j: Z_j = M(V_j), then V_{j+1}, Hessenberg column j, est[j].
int arnoldi_step(pb_op2d* op, pb_nn* nn, FgWork* f, int j, const pbmg::CycleOpts& co, hipStream_t s) {
  const long n = op->n;
  const int m = f->m, K = j + 1;
  double* vj = f->V + (long)j * n;
  float* zj = f->Z + (long)j * n;
  int rc = pbmg::cycle(op, nn, vj, zj, co, s);
  if (rc) return rc;
  launch_fg_matvec(op, zj, f->w, s);
  const dim3 gb(kFgBlocks), tb(kFgThreads);
  hipLaunchKernelGGL(kf_dots, gb, tb, 0, s, (const double*)f->V, n, K, (const double*)f->w, f->part);
  hipLaunchKernelGGL(kf_colsum, dim3(K), tb, 0, s, (const double*)f->part, f->h1);
  // the register kernels load KM basis values per element unconditionally:
  // the smallest KM that holds K
#define PB_ORTH1R(KM)                                                                                  \
  hipLaunchKernelGGL(kf_orth1r<KM>, gb, tb, 0, s, (const double*)f->V, n, K, (const double*)f->h1,     \
                     (const double*)f->w, f->w2, f->part)
  if (K <= 4) PB_ORTH1R(4);
  else if (K <= 8) PB_ORTH1R(8);
  else if (K <= 16) PB_ORTH1R(16);
  else if (K <= 32) PB_ORTH1R(32);
  else
    hipLaunchKernelGGL(kf_orth1, gb, tb, 0, s, (const double*)f->V, n, K, (const double*)f->h1, (const double*)f->w,
                       f->w2, f->part);
#undef PB_ORTH1R
  hipLaunchKernelGGL(kf_colsum, dim3(K), tb, 0, s, (const double*)f->part, f->h2);
#define PB_ORTH2R(KM)                                                                                  \
  hipLaunchKernelGGL(kf_orth2r<KM>, gb, tb, 0, s, (const double*)f->V, n, K, (const double*)f->h2,     \
                     (const double*)f->w2, f->w, f->part3)
  if (K <= 4) PB_ORTH2R(4);
  else if (K <= 8) PB_ORTH2R(8);
  else if (K <= 16) PB_ORTH2R(16);
  else if (K <= 32) PB_ORTH2R(32);
  else
    hipLaunchKernelGGL(kf_orth2, gb, tb, 0, s, (const double*)f->V, n, K, (const double*)f->h2, (const double*)f->w2,
                       f->w, f->part3);
#undef PB_ORTH2R
  hipLaunchKernelGGL(kf_norm, gb, tb, 0, s, j, m, n, (const double*)f->part3, (const double*)f->w,
                     f->V + (long)(j + 1) * n, (const double*)f->h1, (const double*)f->h2, f->H, f->cs, f->sn, f->g,
                     f->est);
  PB_LAUNCH_CHECK(op->ctx);
  return PB_OK;
}

// One FGMRES(m) cycle from the residual in op->rtrue (norm rn): adds Z y to x.
// Returns the number of Arnoldi steps used in *k.
int fgmres_cycle(pb_op2d* op, pb_nn* nn, FgWork* f, double* x, int m, double rn, double thr,
                 const pbmg::CycleOpts& co, hipStream_t s, int* kout) {
  const long n = op->n;
  pb_ctx* ctx = op->ctx;
  hipLaunchKernelGGL(kf_start, dim3(kFgBlocks), dim3(kFgThreads), 0, s, n, f->m, (const double*)op->rtrue, rn, f->V,
                     f->g);
  int k = m;
  double prev = rn;   // estimate one step before the last one read
  for (int j = 0; j < m; ++j) {
    int rc = arnoldi_step(op, nn, f, j, co, s);
    if (rc) return rc;
    PB_HIP(ctx, hipMemcpyAsync(f->hest + j, f->est + j, sizeof(double), hipMemcpyDeviceToHost, s));
    PB_HIP(ctx, hipEventRecord(f->ev[j], s));
    if (j == 0) continue;
    // read step j - 1 while step j runs
    PB_HIP(ctx, hipEventSynchronize(f->ev[j - 1]));
    const double e1 = f->hest[j - 1];
    if (e1 <= thr) {
      k = j;
      break;
    }
    // the rate so far predicts step j reaches the tolerance: wait for it now
    const double rate = prev > 0.0 ? e1 / prev : 1.0;
    prev = e1;
    if (e1 * rate <= thr || j + 1 == m) {
      PB_HIP(ctx, hipEventSynchronize(f->ev[j]));
      if (f->hest[j] <= thr) {
        k = j + 1;
        break;
      }
    }
  }
  hipLaunchKernelGGL(kf_backsub, dim3(1), dim3(64), 0, s, k, f->m, (const double*)f->H, (const double*)f->g, f->y);
  hipLaunchKernelGGL(kf_update, dim3(elem_blocks(n)), dim3(256), 0, s, n, k, (const double*)f->y, 1.0,
                     (const float*)f->Z, x);
  PB_LAUNCH_CHECK(ctx);
  *kout = k;
  return PB_OK;
}

}  // namespace

// frees the learned mode's workspace of an operator (called on its destroy)
void pb_fgmres_release(pb_op2d* op) {
  fg_free(op->fg);
  op->fg = nullptr;
  pbmg::release(op);
  if (op->nn) pb_nn_track(op->nn, op, false);
  op->nn = nullptr;
}

extern "C" int pb_op2d_attach_nn(pb_op2d* op, pb_nn* nn) {
  if (!op) return PB_ERR_ARG;
  if (op->nn) pb_nn_track(op->nn, op, false);
  op->nn = nn;
  if (nn) pb_nn_track(nn, op, true);
  return PB_OK;
}

// The learned mode: FGMRES preconditioned by the multigrid cycle whose coarse
// level the attached network solves, for a budget of opts->maxiter Arnoldi
// steps.  Called by pb_bicgstab_solve for PB_SOLVE_NN, which falls back to
// plain BiCGStab when the budget runs out.
int pb_nn_solve(pb_op2d* op, const double* b, double* x, const pb_solve_opts* opts, pb_solve_info* info,
                double (*true_res)(pb_op2d*, const double*, const double*, hipStream_t, int*), hipStream_t s) {
  pb_ctx* ctx = op->ctx;
  pbmg::CycleOpts co;
  co.nu_pre = co.nu_post = opts->sweeps > 0 ? std::min(opts->sweeps, 8) : 2;
  co.omega = opts->omega > 0 ? (float)opts->omega : 0.8f;
  if (opts->use_network == 1) {
    if (!op->nn) return pb_fail(ctx, PB_ERR_STATE, "PB_SOLVE_NN: no network attached to the operator");
    co.coarse = pbmg::COARSE_NET;
  } else if (opts->use_network == 2) {
    co.coarse = pbmg::COARSE_MG;
  } else {
    co.coarse = pbmg::COARSE_SMOOTH;
  }
  const int m = opts->krylov_m > 0 ? std::max(2, std::min(kMaxM, opts->krylov_m)) : 30;
  int rc = pbmg::build(op, s);
  if (rc) return rc;
  FgWork* f = nullptr;
  rc = fg_alloc(op, m, &f);
  if (rc) return rc;
  const long n = op->n;
  const int nb = elem_blocks(n);
  int rcode = 0;
  const double bn = true_res(op, b, op->zero, s, &rcode);
  if (rcode) return rcode;
  if (!opts->x0_given) hipLaunchKernelGGL(k_fill_zero_d, dim3(nb), dim3(256), 0, s, n, x);
  double rn = true_res(op, b, x, s, &rcode);   // op->rtrue = b - A x
  if (rcode) return rcode;
  const double target = opts->tol * bn;
  const int budget = opts->maxiter;
  int total = 0, outer = 0;
  while (rn > target && total < budget && outer < 50) {
    int k = 0;
    // the estimate is the FP64 least-squares residual; 0.7 leaves room for
    // its drift from the recomputed true residual
    rc = fgmres_cycle(op, op->nn, f, x, std::min(m, budget - total), rn, 0.7 * target, co, s, &k);
    if (rc) return rc;
    total += k;
    ++outer;
    const double rn_old = rn;
    rn = true_res(op, b, x, s, &rcode);
    if (rcode) return rcode;
    if (!(rn < rn_old)) {   // no progress: undo this cycle's correction and stop
      hipLaunchKernelGGL(kf_update, dim3(nb), dim3(256), 0, s, n, k, (const double*)f->y, -1.0, (const float*)f->Z,
                         x);
      rn = true_res(op, b, x, s, &rcode);
      if (rcode) return rcode;
      break;
    }
  }
  info->outer = outer;
  info->iters = total;
  info->bnrm2 = bn;
  info->half_step = 0;
  info->resid = rn;
  info->info = rn <= target ? 0 : std::max(total, 1);
  info->true_rel_resid = bn > 0 ? rn / bn : rn;
  return PB_OK;
}
