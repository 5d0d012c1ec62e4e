// 2D operator handle: assembly on the device and the plain matvec (gfx950).
//
// Reference counterparts (pyBELLA tree):
//   k_assemble_*  operator_coefficients_nodes   second_projection.py:308-356
//                 scale_wall_node_values        second_projection.py:433-464
//                 regime scaling of wcenter     second_projection.py:141-153
//                 multiply_inverse_coriolis(get_coeffs=True) second_projection.py:604-652
//                 precon_diag_prepare           laplacian.py:1284-1356
//   k_matvec      lap2D_gather_new              laplacian.py:546-746
// The flux form of the operator and the tiling are described in pb_op2d.hpp.
#include "pb_op2d.hpp"

using namespace pb2d;

namespace {

__global__ void k_scale_rhs(int nx, int ny, const double* rhs_isc, const double* dinv, double* b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nx || j >= ny) return;
  b[(long)j * nx + i] = rhs_isc[(long)(j + 1) * (nx + 2) + (i + 1)] * dinv[(long)j * nx + i];
}

template <typename T>
struct InVec {
  using Raw = T;
  const T* x;
  __device__ Raw load(long g) const { return x[g]; }
  __device__ T combine(const Raw& v) const { return v; }
};
template <typename T>
struct OutStore {
  struct Pre {};
  T* y;
  __device__ Pre prefetch(long) const { return Pre{}; }
  template <class R>
  __device__ void operator()(int, long g, T v, T, const Pre&, const R&) { y[g] = v; }
};

template <typename T, bool CROSS, bool CENTER, int RY>
__global__ __launch_bounds__(NTHR) void k_matvec(OpDev<T> op, const T* x, T* y) {
  __shared__ TileLds<T, RY> L;
  InVec<T> in{x};
  OutStore<T> out{y};
  stencil_tile2<T, CROSS, CENTER, RY>(op, blockIdx.x, blockIdx.y, in, out, L, [] { return true; });
}

template <typename T>
void launch_matvec(const pb_op2d* op, const T* x, T* y, hipStream_t s) {
  const OpDev<T> d = opdev<T>(op);
  const dim3 grd = stencil_grid(op, 4), blk(TX, TYT);
  if (op->cross) {
    if (op->center) hipLaunchKernelGGL((k_matvec<T, true, true, 4>), grd, blk, 0, s, d, x, y);
    else hipLaunchKernelGGL((k_matvec<T, true, false, 4>), grd, blk, 0, s, d, x, y);
  } else {
    if (op->center) hipLaunchKernelGGL((k_matvec<T, false, true, 4>), grd, blk, 0, s, d, x, y);
    else hipLaunchKernelGGL((k_matvec<T, false, false, 4>), grd, blk, 0, s, d, x, y);
  }
}

// ---------------------------------------------------------------------------
// assembly
// ---------------------------------------------------------------------------
struct AsmParams {
  int nx, ny;          // inner nodes
  int cx, cy;          // cells (nx+3, ny+3)
  double dx, dy, dt;
  double Gammainv, cexp, ccenter, g_over_Msq, nonhydro, wh1, wv, wh2;
  int x_wall, y_wall;
  double regime_scale;   // factor applied to wcenter after wall scaling
  double w_diag;         // 0.75 (2D)
};

// all cells: coeff = Gammainv rhoY Y (-> wplus), nu, rhoY^cexp
__global__ void k_assemble_cells(AsmParams a, const double* rho, const double* rhoY,
                                 const double* dSdy, double* coeff, double* rYpow, double* wplus0,
                                 double* wplus1, double* nu_c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= a.cx || j >= a.cy) return;
  const long c = (long)j * a.cx + i;
  const double rY = rhoY[c];
  const double Y = rY / rho[c];
  const double cf = a.Gammainv * rY * Y;
  coeff[c] = cf;
  rYpow[c] = pow(rY, a.cexp);
  if (wplus0) wplus0[c] = cf;
  if (wplus1) wplus1[c] = cf;
  if (nu_c) nu_c[c] = -a.dt * a.dt * a.g_over_Msq * dSdy[c] * Y;
}

// op-cells (ny+1) x (nx+1): folded weights with Coriolis factors from (rho_c, rhoY_c)
__global__ void k_assemble_opcells(AsmParams a, const double* coeff, const double* rho_c,
                                   const double* rhoY_c, const double* dSdy, double* wxx, double* wyy,
                                   double* wyx, double* wxy, float* wxx32, float* wyy32, float* wyx32,
                                   float* wxy32) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > a.nx || j > a.ny) return;
  const long c = (long)(j + 1) * a.cx + (i + 1);   // reference cell [i+1, j+1]
  double hx = coeff[c], hy = coeff[c];
  if ((a.x_wall && (i == 0 || i == a.nx)) || (a.y_wall && (j == 0 || j == a.ny))) {
    hx = 0.0;
    hy = 0.0;
  }
  const double Y = rhoY_c[c] / rho_c[c];
  const double nu = -a.dt * a.dt * a.g_over_Msq * dSdy[c] * Y;
  const double nh = a.nonhydro, wh1 = a.wh1, wv = a.wv, wh2 = a.wh2;
  const double denom = 1.0 / (wh1 * wh1 + wh2 * wh2 + (nu + nh) * (wv * wv + 1.0));
  const double cxx = (wh1 * wh1 + nu + nh) * denom;
  const double cyy = (nh * (1 + wv * wv)) * denom;
  const double cxy = (nh * (wh1 * wv + wh2)) * denom;
  const double cyx = (wh1 * wv - wh2) * denom;
  const double ox2 = 0.25 / (a.dx * a.dx), oy2 = 0.25 / (a.dy * a.dy), oxy = 0.25 / (a.dx * a.dy);
  const long o = (long)j * (a.nx + 1) + i;
  const double vxx = hx * cxx * ox2, vyy = hy * cyy * oy2, vyx = hy * cxy * oxy, vxy = hx * cyx * oxy;
  wxx[o] = vxx;
  wyy[o] = vyy;
  wyx[o] = vyx;
  wxy[o] = vxy;
  wxx32[o] = (float)vxx;
  wyy32[o] = (float)vyy;
  wyx32[o] = (float)vyx;
  wxy32[o] = (float)vxy;
}

// isc nodes (ny+2) x (nx+2): wcenter (box mean, wall and regime scaling), diag_inv
__global__ void k_assemble_isc(AsmParams a, const double* coeff, const double* rYpow, double* wcenter,
                               double* diag_inv, double* hc, double* dinv, float* hc32, float* dinv32) {
  const int ia = blockIdx.x * blockDim.x + threadIdx.x;   // x index on isc
  const int ib = blockIdx.y * blockDim.y + threadIdx.y;   // y index on isc
  const int ix = a.nx + 2, iy = a.ny + 2;
  if (ia >= ix || ib >= iy) return;
  const long c00 = (long)ib * a.cx + ia, c10 = c00 + 1, c01 = c00 + a.cx, c11 = c01 + 1;
  const double boxp = (rYpow[c00] + rYpow[c10]) + (rYpow[c01] + rYpow[c11]);
  double wc = a.ccenter * boxp / 4.0;
  // scale_wall_node_values(wcenter, x 0.5), second_projection.py:433-464
  if (a.x_wall && (ia == 1 || ia == a.nx) && ib >= 2 && ib <= a.ny - 1) wc *= 0.5;
  if (a.y_wall && (ib == 1 || ib == a.ny)) {
    const bool on = a.x_wall ? (ia == 1) : (ia >= 2 && ia <= a.nx - 1);
    if (on) wc *= 0.5;
  }
  wc *= a.regime_scale;
  const double boxc = (coeff[c00] + coeff[c10]) + (coeff[c01] + coeff[c11]);
  const double wxx = a.w_diag / (a.dx * a.dx), wyy = a.w_diag / (a.dy * a.dy);
  const double wxy = a.w_diag / (a.dx * a.dy), wyx = a.w_diag / (a.dy * a.dx);
  double dg = -wxx * boxc;
  dg -= wyy * boxc;
  dg -= wxy * boxc;
  dg -= wyx * boxc;
  dg += wc;
  const double di = 1.0 / dg;
  const long o = (long)ib * ix + ia;
  if (wcenter) wcenter[o] = wc;
  if (diag_inv) diag_inv[o] = di;
  if (ia >= 1 && ia <= a.nx && ib >= 1 && ib <= a.ny) {
    const long g = (long)(ib - 1) * a.nx + (ia - 1);
    hc[g] = wc;
    dinv[g] = di;
    hc32[g] = (float)wc;
    dinv32[g] = (float)di;
  }
}

}  // namespace

extern "C" int pb_op2d_create(pb_ctx* ctx, const pb_grid2d* grid, pb_op2d** out) {
  if (!ctx || !grid || !out) return pb_fail(ctx, PB_ERR_ARG, "pb_op2d_create: null argument");
  if (grid->nx < 3 || grid->ny < 3)
    return pb_fail(ctx, PB_ERR_ARG, "pb_op2d_create: need nx, ny >= 3 (got %d, %d)", grid->nx, grid->ny);
  if (!(grid->dx > 0) || !(grid->dy > 0)) return pb_fail(ctx, PB_ERR_ARG, "pb_op2d_create: bad spacing");
  if ((long)grid->nx * grid->ny > (1L << 30)) return pb_fail(ctx, PB_ERR_ARG, "pb_op2d_create: grid too large");
  pb_op2d* op = new pb_op2d();
  op->ctx = ctx;
  op->g = *grid;
  op->n = (long)grid->nx * grid->ny;
  op->nc = (long)(grid->nx + 1) * (grid->ny + 1);
  const long ncell = (long)(grid->nx + 3) * (grid->ny + 3);
  int rc = 0;
  rc |= dmalloc(op, &op->wxx, op->nc);
  rc |= dmalloc(op, &op->wyy, op->nc);
  rc |= dmalloc(op, &op->wyx, op->nc);
  rc |= dmalloc(op, &op->wxy, op->nc);
  rc |= dmalloc(op, &op->hc, op->n);
  rc |= dmalloc(op, &op->dinv, op->n);
  rc |= dmalloc(op, &op->wxx32, op->nc);
  rc |= dmalloc(op, &op->wyy32, op->nc);
  rc |= dmalloc(op, &op->wyx32, op->nc);
  rc |= dmalloc(op, &op->wxy32, op->nc);
  rc |= dmalloc(op, &op->hc32, op->n);
  rc |= dmalloc(op, &op->dinv32, op->n);
  rc |= dmalloc(op, &op->coeff, ncell);
  rc |= dmalloc(op, &op->rYpow, ncell);
  if (rc) {
    pb_op2d_destroy(op);
    return PB_ERR_HIP;
  }
  *out = op;
  return PB_OK;
}

void pb_fgmres_release(pb_op2d* op);   // pb_fgmres.hip

extern "C" int pb_op2d_destroy(pb_op2d* op) {
  if (!op) return PB_OK;
  if (op->ctx) {
    (void)hipStreamSynchronize(op->ctx->stream);
    if (op->ctx->own_stream) (void)hipStreamSynchronize(op->ctx->own_stream);
  }
  pb_fgmres_release(op);
  for (auto& kv : op->w64.graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto& kv : op->w32.graphs) (void)hipGraphExecDestroy(kv.second);
  for (void* p : op->allocs) (void)hipFree(p);
  if (op->hstate) (void)hipHostFree(op->hstate);
  if (op->hred) (void)hipHostFree(op->hred);
  delete op;
  return PB_OK;
}

extern "C" int pb_op2d_assemble(pb_op2d* op, const pb_phys* phys, double dt, const double* rho_w,
                                const double* rhoY_w, const double* rho_c, const double* rhoY_c,
                                const double* dSdy, double* wplus0, double* wplus1, double* nu_c,
                                double* wcenter, double* diag_inv) {
  if (!op) return PB_ERR_ARG;
  pb_ctx* ctx = op->ctx;
  if (!phys || !rho_w || !rhoY_w || !rho_c || !rhoY_c || !dSdy)
    return pb_fail(ctx, PB_ERR_ARG, "pb_op2d_assemble: null argument");
  if (!(dt > 0)) return pb_fail(ctx, PB_ERR_ARG, "pb_op2d_assemble: dt must be > 0");
  const pb_grid2d& g = op->g;
  const double gm1 = phys->gamm - 1.0;
  AsmParams a{};
  a.nx = g.nx;
  a.ny = g.ny;
  a.cx = g.nx + 3;
  a.cy = g.ny + 3;
  a.dx = g.dx;
  a.dy = g.dy;
  a.dt = dt;
  a.Gammainv = phys->gamm / gm1;
  a.cexp = 2.0 - phys->gamm;
  a.ccenter = -phys->Msq * (1.0 / gm1) / (dt * dt);
  a.g_over_Msq = phys->g / phys->Msq;
  a.nonhydro = phys->nonhydro;
  a.wh1 = dt * phys->coriolis[0];
  a.wv = dt * phys->coriolis[1];
  a.wh2 = dt * phys->coriolis[2];
  a.x_wall = g.bc_x == PB_BC_WALL;
  a.y_wall = g.bc_y == PB_BC_WALL;
  a.w_diag = 1.0 - 0.5 * 0.5;
  // regime (second_projection.py:141-153)
  if (phys->is_compressible == 1) {
    a.regime_scale = 1.0;
  } else if (phys->is_compressible == 0 && phys->is_arakawa_konor) {
    a.regime_scale = 0.0;
  } else {
    a.regime_scale = phys->compressibility;
  }
  op->center = a.regime_scale != 0.0;
  op->cross = (a.wh1 != 0.0 || a.wh2 != 0.0);
  hipStream_t s = ctx->stream;
  dim3 blk(64, 4);
  hipLaunchKernelGGL(k_assemble_cells, dim3(ceil_div(a.cx, 64), ceil_div(a.cy, 4)), blk, 0, s, a, rho_w,
                     rhoY_w, dSdy, op->coeff, op->rYpow, wplus0, wplus1, nu_c);
  hipLaunchKernelGGL(k_assemble_opcells, dim3(ceil_div(g.nx + 1, 64), ceil_div(g.ny + 1, 4)), blk, 0, s, a,
                     op->coeff, rho_c, rhoY_c, dSdy, op->wxx, op->wyy, op->wyx, op->wxy, op->wxx32,
                     op->wyy32, op->wyx32, op->wxy32);
  hipLaunchKernelGGL(k_assemble_isc, dim3(ceil_div(g.nx + 2, 64), ceil_div(g.ny + 2, 4)), blk, 0, s, a,
                     op->coeff, op->rYpow, wcenter, diag_inv, op->hc, op->dinv, op->hc32, op->dinv32);
  PB_LAUNCH_CHECK(ctx);
  op->assembled = true;
  op->version++;
  return PB_OK;
}

extern "C" int pb_op2d_matvec(pb_op2d* op, const void* x, void* y, int precision) {
  if (!op) return PB_ERR_ARG;
  if (!x || !y) return pb_fail(op->ctx, PB_ERR_ARG, "pb_op2d_matvec: null vector");
  if (!op->assembled) return pb_fail(op->ctx, PB_ERR_STATE, "pb_op2d_matvec: operator not assembled");
  if (precision == PB_F64) {
    launch_matvec<double>(op, (const double*)x, (double*)y, op->ctx->stream);
  } else if (precision == PB_F32) {
    launch_matvec<float>(op, (const float*)x, (float*)y, op->ctx->stream);
  } else {
    return pb_fail(op->ctx, PB_ERR_ARG, "pb_op2d_matvec: unknown precision %d", precision);
  }
  PB_LAUNCH_CHECK(op->ctx);
  return PB_OK;
}

extern "C" int pb_op2d_scale_rhs(pb_op2d* op, const double* rhs_isc, double* b_vec) {
  if (!op) return PB_ERR_ARG;
  if (!rhs_isc || !b_vec) return pb_fail(op->ctx, PB_ERR_ARG, "pb_op2d_scale_rhs: null argument");
  if (!op->assembled) return pb_fail(op->ctx, PB_ERR_STATE, "pb_op2d_scale_rhs: operator not assembled");
  hipLaunchKernelGGL(k_scale_rhs, dim3(ceil_div(op->g.nx, 256), op->g.ny), dim3(256), 0, op->ctx->stream,
                     op->g.nx, op->g.ny, rhs_isc, op->dinv, b_vec);
  PB_LAUNCH_CHECK(op->ctx);
  return PB_OK;
}

