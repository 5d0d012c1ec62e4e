// Flexible GMRES with the learned preconditioner (gfx950).  No reference
// counterpart: it replaces the reference's bicgstab call
// (second_projection.py:230) when a network is attached to the operator.
//
// Outer loop in FP64 (iterative refinement): r = b - A x (FP64 true residual,
// stopping at ||r|| <= tol ||b||).  Inner loop: FGMRES(m) on A d = r with an
// FP32 Krylov basis V, preconditioned directions Z_j = M(V_j) where M is the
// network (Saad 1993, "A flexible inner-outer preconditioned GMRES") followed
// by optional damped-Jacobi sweeps with the exact diagonal of A; classical
// Gram-Schmidt applied twice per step (Giraud et al. 2005); Givens rotations
// and the small least-squares solve on the host in FP64.  The first outer step
// starts from x0 = M(b) (the network's initial guess).  If the iteration
// budget runs out before the tolerance, plain BiCGStab finishes from the
// current iterate.
#include <type_traits>
#include <vector>

#include "pb_op2d.hpp"

using namespace pb2d;

int pb_nn_apply_f32(pb_nn* nn, const pb_op2d* op, const float* r, float* z, hipStream_t s);

namespace {

struct FgWork {
  int m = 0;
  long n = 0;
  float* V = nullptr;     // (m+1) x n
  float* Z = nullptr;     // m x n
  float* w = nullptr;
  float* t = nullptr;
  float* r32 = nullptr;   // FP32 copy of the outer residual
  float* diag = nullptr;  // exact diagonal of A (FP32)
  double* part = nullptr; // dot partials
  double* h = nullptr;    // device dot results (m+2)
  double* hh = nullptr;   // pinned host mirror
  std::vector<void*> allocs;
};

std::map<const pb_op2d*, FgWork*> g_work;
std::map<const pb_op2d*, pb_nn*> g_attached;

constexpr int kDotBlocks = 128;

// diag(A)[n] = dinv * (hc - sum of the 4 cells' (Wxx + Wyy)): the coefficient
// of p_self in the flux form (each cell's gx and gy change by -1 per unit p_self)
__global__ void k_diag(int nx, int ny, const float* wxx, const float* wyy, const float* hc, const float* dinv,
                       int center, float* d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= nx || j >= ny) return;
  const long c00 = (long)j * (nx + 1) + i, c01 = c00 + 1, c10 = c00 + nx + 1, c11 = c10 + 1;
  const long g = (long)j * nx + i;
  const float s = (wxx[c00] + wxx[c01] + wxx[c10] + wxx[c11]) + (wyy[c00] + wyy[c01] + wyy[c10] + wyy[c11]);
  d[g] = dinv[g] * ((center ? hc[g] : 0.f) - s);
}

// partial dots of w with V[0..k-1]: grid (kDotBlocks, k)
__global__ void k_mdot(const float* V, long n, const float* w, double* part) {
  const int i = blockIdx.y;
  const float* v = V + (long)i * n;
  double s = 0.0;
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x)
    s += (double)v[g] * (double)w[g];
  __shared__ double sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(long)i * gridDim.x + blockIdx.x] = sh[0];
}

__global__ void k_reduce_parts(const double* part, int nb, int k, double* out) {
  const int i = blockIdx.x;
  if (i >= k) return;
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[(long)i * nb + b];
  __shared__ double sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[i] = sh[0];
}

// w -= sum_i h[i] V[i]; partials of w.w into part (one row)
__global__ void k_orth_update(const float* V, long n, int k, const double* h, float* w, double* part) {
  double s = 0.0;
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x) {
    double acc = (double)w[g];
    for (int i = 0; i < k; ++i) acc -= h[i] * (double)V[(long)i * n + g];
    const float wf = (float)acc;
    w[g] = wf;
    s += (double)wf * (double)wf;
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_scale_to(long n, const float* w, const double* nrm2, float* v) {
  const double inv = 1.0 / sqrt(*nrm2);
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x)
    v[g] = (float)((double)w[g] * inv);
}

__global__ void k_combine(const float* Z, long n, int k, const double* y, double* x) {
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int i = 0; i < k; ++i) acc += y[i] * (double)Z[(long)i * n + g];
    x[g] += acc;
  }
}

// z += omega (v - t) / diag
__global__ void k_jacobi(long n, const float* v, const float* t, const float* d, float omega, float* z) {
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x)
    z[g] += omega * (v[g] - t[g]) / d[g];
}

// z = omega v / diag (first Jacobi step from z = 0)
__global__ void k_jacobi0(long n, const float* v, const float* d, float omega, float* z) {
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x)
    z[g] = omega * v[g] / d[g];
}

// partial sums of v.t and t.t (one pair per block)
__global__ void k_dot2(long n, const float* v, const float* t, double* part) {
  double a = 0.0, b = 0.0;
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x) {
    const double tv = (double)t[g];
    a += (double)v[g] * tv;
    b += tv * tv;
  }
  __shared__ double sh[2][256];
  sh[0][threadIdx.x] = a;
  sh[1][threadIdx.x] = b;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + off];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sh[0][0];
    part[2 * blockIdx.x + 1] = sh[1][0];
  }
}

// alpha = (v.t) / (t.t) minimises ||v - alpha A z|| (every block re-reduces the
// partials in the same order, so all blocks see the same alpha); then
// z = alpha z + omega (v - alpha t) / diag
__global__ void k_scale_jacobi(long n, const double* part, int nparts, const float* v, const float* t, const float* d,
                               float omega, float* z) {
  __shared__ float alpha_s;
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int k = 0; k < nparts; ++k) {
      a += part[2 * k];
      b += part[2 * k + 1];
    }
    alpha_s = b > 0.0 ? (float)(a / b) : 1.f;
  }
  __syncthreads();
  const float alpha = alpha_s;
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x)
    z[g] = alpha * z[g] + omega * (v[g] - alpha * t[g]) / d[g];
}

__global__ void k_cast_df(long n, const double* a, float* b) {
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x) b[g] = (float)a[g];
}
__global__ void k_cast_fd(long n, const float* a, double* b) {
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x) b[g] = (double)a[g];
}
__global__ void k_axpy_d(long n, double a, const double* dx, double* x) {
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x) x[g] += a * dx[g];
}

int fg_alloc(pb_op2d* op, int m, FgWork** out) {
  auto it = g_work.find(op);
  if (it != g_work.end() && it->second->m >= m && it->second->n == op->n) {
    *out = it->second;
    return PB_OK;
  }
  if (it != g_work.end()) {
    for (void* p : it->second->allocs) (void)hipFree(p);
    if (it->second->hh) (void)hipHostFree(it->second->hh);
    delete it->second;
    g_work.erase(it);
  }
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
  int rc = 0;
  rc |= A(&f->V, (long)(m + 1) * op->n);
  rc |= A(&f->Z, (long)m * op->n);
  rc |= A(&f->w, op->n);
  rc |= A(&f->t, op->n);
  rc |= A(&f->r32, op->n);
  rc |= A(&f->diag, op->n);
  rc |= A(&f->part, (long)(m + 2) * kDotBlocks);
  rc |= A(&f->h, m + 2);
  void* hp = nullptr;
  if (rc || hipHostMalloc(&hp, sizeof(double) * (m + 2), hipHostMallocDefault) != hipSuccess) {
    for (void* p : f->allocs) (void)hipFree(p);
    delete f;
    return pb_fail(op->ctx, PB_ERR_HIP, "fgmres: workspace allocation failed");
  }
  f->hh = (double*)hp;
  g_work[op] = f;
  *out = f;
  return PB_OK;
}

// norm^2 of a float vector into f->h[0] (device), returned on the host
int fnorm2(pb_op2d* op, FgWork* f, const float* v, hipStream_t s, double* out) {
  hipLaunchKernelGGL(k_mdot, dim3(kDotBlocks, 1), dim3(256), 0, s, v, op->n, v, f->part);
  hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(256), 0, s, (const double*)f->part, kDotBlocks, 1, f->h);
  PB_HIP(op->ctx, hipMemcpyAsync(f->hh, f->h, sizeof(double), hipMemcpyDeviceToHost, s));
  PB_HIP(op->ctx, hipStreamSynchronize(s));
  *out = f->hh[0];
  return PB_OK;
}

// z = M(v): the network, then ``sweeps`` damped-Jacobi steps on A z = v; the
// first sweep also rescales the network output by the alpha that minimises
// ||v - alpha A z|| (computed on the device, no host round trip).  v, z and
// f->t must be distinct buffers.  With ``rescale`` the alpha step is done even
// when sweeps == 0 (the initial guess needs the right magnitude; flexible GMRES
// directions do not).
int precond(pb_op2d* op, pb_nn* nn, FgWork* f, const float* v, float* z, int sweeps, float omega, bool rescale,
            hipStream_t s) {
  const int nb = elem_blocks(op->n);
  if (!nn) {   // Jacobi only: z = omega v / diag, then the remaining sweeps
    hipLaunchKernelGGL(k_jacobi0, dim3(nb), dim3(256), 0, s, op->n, v, (const float*)f->diag, omega, z);
    for (int k = 1; k < sweeps; ++k) {
      void* saved = op->ctx->stream;
      op->ctx->stream = s;
      int rc = pb_op2d_matvec(op, z, f->t, PB_F32);
      op->ctx->stream = (hipStream_t)saved;
      if (rc) return rc;
      hipLaunchKernelGGL(k_jacobi, dim3(nb), dim3(256), 0, s, op->n, v, (const float*)f->t, (const float*)f->diag,
                         omega, z);
    }
    PB_LAUNCH_CHECK(op->ctx);
    return PB_OK;
  }
  int rc = pb_nn_apply_f32(nn, op, v, z, s);
  if (rc) return rc;
  const int steps = (sweeps == 0 && rescale) ? 1 : sweeps;
  if (sweeps == 0) omega = 0.f;
  for (int k = 0; k < steps; ++k) {
    void* saved = op->ctx->stream;
    op->ctx->stream = s;
    rc = pb_op2d_matvec(op, z, f->t, PB_F32);
    op->ctx->stream = (hipStream_t)saved;
    if (rc) return rc;
    if (k == 0) {
      hipLaunchKernelGGL(k_dot2, dim3(kDotBlocks), dim3(256), 0, s, op->n, v, (const float*)f->t, f->part);
      hipLaunchKernelGGL(k_scale_jacobi, dim3(nb), dim3(256), 0, s, op->n, (const double*)f->part, kDotBlocks, v,
                         (const float*)f->t, (const float*)f->diag, omega, z);
    } else {
      hipLaunchKernelGGL(k_jacobi, dim3(nb), dim3(256), 0, s, op->n, v, (const float*)f->t, (const float*)f->diag,
                         omega, z);
    }
  }
  PB_LAUNCH_CHECK(op->ctx);
  return PB_OK;
}

// One FGMRES(m) cycle on A d = r32; adds d to x (FP64).  Returns the number
// of iterations; *est = final residual estimate relative to ||r32||.
int fgmres_cycle(pb_op2d* op, pb_nn* nn, FgWork* f, const float* r32, double* x, int m, double tol_rel, int sweeps,
                 float omega, hipStream_t s, int* iters, double* est) {
  const long n = op->n;
  const int nb = elem_blocks(n);
  double beta2 = 0.0;
  int rc = fnorm2(op, f, r32, s, &beta2);
  if (rc) return rc;
  const double beta = sqrt(beta2);
  *iters = 0;
  *est = 0.0;
  if (beta == 0.0) return PB_OK;
  // V0 = r / beta
  PB_HIP(op->ctx, hipMemcpyAsync(f->h, &beta2, sizeof(double), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_scale_to, dim3(nb), dim3(256), 0, s, n, r32, (const double*)f->h, f->V);
  std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), g(m + 1, 0.0);
  g[0] = beta;
  int j = 0;
  for (; j < m; ++j) {
    float* vj = f->V + (long)j * n;
    float* zj = f->Z + (long)j * n;
    rc = precond(op, nn, f, vj, zj, sweeps, omega, false, s);
    if (rc) return rc;
    void* saved = op->ctx->stream;
    op->ctx->stream = s;
    rc = pb_op2d_matvec(op, zj, f->w, PB_F32);
    op->ctx->stream = (hipStream_t)saved;
    if (rc) return rc;
    std::vector<double> hcol(j + 2, 0.0);
    double wn2 = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
      hipLaunchKernelGGL(k_mdot, dim3(kDotBlocks, j + 1), dim3(256), 0, s, (const float*)f->V, n, (const float*)f->w,
                         f->part);
      hipLaunchKernelGGL(k_reduce_parts, dim3(j + 1), dim3(256), 0, s, (const double*)f->part, kDotBlocks, j + 1, f->h);
      hipLaunchKernelGGL(k_orth_update, dim3(kDotBlocks), dim3(256), 0, s, (const float*)f->V, n, j + 1,
                         (const double*)f->h, f->w, f->part);
      hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(256), 0, s, (const double*)f->part, kDotBlocks, 1,
                         f->h + j + 1);
      PB_HIP(op->ctx, hipMemcpyAsync(f->hh, f->h, sizeof(double) * (j + 2), hipMemcpyDeviceToHost, s));
      PB_HIP(op->ctx, hipStreamSynchronize(s));
      for (int i = 0; i <= j; ++i) hcol[i] += f->hh[i];
      wn2 = f->hh[j + 1];
    }
    const double hn = sqrt(wn2);
    hcol[j + 1] = hn;
    if (hn > 0.0) {
      hipLaunchKernelGGL(k_scale_to, dim3(nb), dim3(256), 0, s, n, (const float*)f->w, (const double*)(f->h + j + 1),
                         f->V + (long)(j + 1) * n);
    }
    // Givens
    for (int i = 0; i < j; ++i) {
      const double a = hcol[i], b = hcol[i + 1];
      hcol[i] = cs[i] * a + sn[i] * b;
      hcol[i + 1] = -sn[i] * a + cs[i] * b;
    }
    const double a = hcol[j], b = hcol[j + 1];
    const double r = std::hypot(a, b);
    cs[j] = r > 0 ? a / r : 1.0;
    sn[j] = r > 0 ? b / r : 0.0;
    hcol[j] = r;
    hcol[j + 1] = 0.0;
    g[j + 1] = -sn[j] * g[j];
    g[j] = cs[j] * g[j];
    for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] = hcol[i];
    *est = std::fabs(g[j + 1]) / beta;
    if (*est <= tol_rel || hn == 0.0) {
      ++j;
      break;
    }
  }
  const int k = j;
  // back substitution H(0:k,0:k) y = g
  std::vector<double> y(k, 0.0);
  for (int i = k - 1; i >= 0; --i) {
    double acc = g[i];
    for (int c = i + 1; c < k; ++c) acc -= H[(size_t)i * m + c] * y[c];
    y[i] = H[(size_t)i * m + i] != 0.0 ? acc / H[(size_t)i * m + i] : 0.0;
  }
  PB_HIP(op->ctx, hipMemcpyAsync(f->h, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_combine, dim3(nb), dim3(256), 0, s, (const float*)f->Z, n, k, (const double*)f->h, x);
  PB_LAUNCH_CHECK(op->ctx);
  PB_HIP(op->ctx, hipStreamSynchronize(s));
  *iters = k;
  return PB_OK;
}

}  // namespace

// frees the flexible-GMRES workspace of an operator (called on its destroy)
void pb_fgmres_release(const pb_op2d* op) {
  auto it = g_work.find(op);
  if (it != g_work.end()) {
    for (void* p : it->second->allocs) (void)hipFree(p);
    if (it->second->hh) (void)hipHostFree(it->second->hh);
    delete it->second;
    g_work.erase(it);
  }
  g_attached.erase(op);
}

extern "C" int pb_op2d_attach_nn(pb_op2d* op, pb_nn* nn) {
  if (!op) return PB_ERR_ARG;
  if (nn) g_attached[op] = nn;
  else g_attached.erase(op);
  return PB_OK;
}

// Solve with the attached network: x0 = M(b), then FGMRES(m) refinement
// cycles; falls back to BiCGStab (fast FP64) past the iteration budget.
// Called by pb_bicgstab_solve for PB_SOLVE_NN.
int pb_nn_solve(pb_op2d* op, const double* b, double* x, const pb_solve_opts* opts, pb_solve_info* info,
                double (*true_res)(pb_op2d*, const double*, const double*, hipStream_t, int*), hipStream_t s) {
  pb_ctx* ctx = op->ctx;
  pb_nn* nn = nullptr;
  if (opts->use_network) {
    auto it = g_attached.find(op);
    if (it == g_attached.end()) return pb_fail(ctx, PB_ERR_STATE, "PB_SOLVE_NN: no network attached to the operator");
    nn = it->second;
  }
  const int m = opts->krylov_m > 0 ? std::max(4, std::min(200, opts->krylov_m)) : 40;
  const int sweeps = opts->sweeps >= 0 ? std::min(opts->sweeps, 16) : 1;
  const float omega = opts->omega > 0 ? (float)opts->omega : 0.6f;
  if (!opts->use_network && sweeps == 0)
    return pb_fail(ctx, PB_ERR_ARG, "PB_SOLVE_NN without the network needs at least one Jacobi sweep");
  FgWork* f = nullptr;
  int rc = fg_alloc(op, m, &f);
  if (rc) return rc;
  const long n = op->n;
  const int nb = elem_blocks(n);
  hipLaunchKernelGGL(k_diag, dim3(ceil_div(op->g.nx, 256), op->g.ny), dim3(256), 0, s, op->g.nx, op->g.ny,
                     (const float*)op->wxx32, (const float*)op->wyy32, (const float*)op->hc32,
                     (const float*)op->dinv32, op->center ? 1 : 0, f->diag);
  const double inner = opts->inner_tol > 0 ? opts->inner_tol : 1e-4;
  const int budget = opts->maxiter;   // total FGMRES iterations before the fallback
  int rcode = 0;
  // ||b|| and the initial guess x0 = M(b) (or the caller's x0)
  const double bn = true_res(op, b, op->zero, s, &rcode);
  if (rcode) return rcode;
  if (!opts->x0_given) {   // V[0], Z[0] are free until the first cycle
    hipLaunchKernelGGL(k_cast_df, dim3(nb), dim3(256), 0, s, n, b, f->V);
    rc = precond(op, nn, f, f->V, f->Z, sweeps, omega, true, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_cast_fd, dim3(nb), dim3(256), 0, s, n, (const float*)f->Z, x);
  }
  double rn = true_res(op, b, x, s, &rcode);   // op->rtrue = b - A x
  if (rcode) return rcode;
  const double target = opts->tol * bn;
  int total = 0, outer = 0;
  while (rn > target && total < budget && outer < 50) {
    hipLaunchKernelGGL(k_cast_df, dim3(nb), dim3(256), 0, s, n, (const double*)op->rtrue, f->r32);
    int its = 0;
    double est = 0.0;
    const double tol_rel = std::max(inner, 0.5 * target / rn);
    rc = fgmres_cycle(op, nn, f, f->r32, x, std::min(m, budget - total), tol_rel, sweeps, omega, s, &its, &est);
    if (rc) return rc;
    total += its;
    ++outer;
    const double rn_old = rn;
    rn = true_res(op, b, x, s, &rcode);
    if (rcode) return rcode;
    if (!(rn < rn_old)) break;   // no progress
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
