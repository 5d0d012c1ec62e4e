// Cell/node field kernels of the 2D second projection (gfx950).
//
// Layout: every array is row-major [y][x] (x fastest), the transpose of the
// reference's arr[x, y].  With NX, NY inner node counts:
//   cells (NY+3) x (NX+3), nodes (NY+4) x (NX+4), isc (NY+2) x (NX+2).
//
// Reference counterparts (pyBELLA tree):
//   k_bc_cells            set_explicit_boundary_data  src/dycore/utils/boundary.py:7-132
//   k_ghost_x / k_ghost_y set_ghostnodes_p2           src/dycore/utils/boundary.py:296-369
//   k_ghostcells_x / _y   set_ghostcells_p2           src/dycore/utils/boundary.py:284-293
//   k_zero_wall_momenta + k_divergence
//                         divergence_nodes            second_projection.py:525-575
//   k_correction          correction_nodes + grad_nodes + multiply_inverse_coriolis
//                         second_projection.py:280-306, 496-521, 604-646
//   k_scatter_p2, k_add   p2_full scatter and p2_nodes += p2_full
//                         second_projection.py:241-245, 272-276
#include "pb_common.hpp"

namespace {

struct Dims2 {
  int nx, ny;     // inner nodes
  int cx, cy;     // cell array extents  (nx+3, ny+3)
  int nxn, nyn;   // node array extents  (nx+4, ny+4)
  int ix, iy;     // isc extents         (nx+2, ny+2)
};

Dims2 dims_of(const pb_grid2d& g) {
  return Dims2{g.nx, g.ny, g.nx + 3, g.ny + 3, g.nx + 4, g.ny + 4, g.nx + 2, g.ny + 2};
}

// ---------------------------------------------------------------------------
// set_explicit_boundary_data.  Ghost columns copy the inner rows of their
// source column (wrap or mirror, boundary.py:135-216), then every column fills
// its two ghost rows at each y wall from the hydrostatic image formula
// (boundary.py:56-132).  The reference pads x first and then fills the y ghost
// rows of every column with column-local formulas; here the ghost-column rows
// are copied one per thread, and each column's thread copies the four rows its
// own images read before it forms them.
// ---------------------------------------------------------------------------
struct BcParams {
  int periodic_x;
  int gravity_y;    // ud.gravity_strength[1] != 0: hydrostatic images in y
  int periodic_y;   // without gravity: wrap (periodic) or mirror (wall) in y
  int compressible;
  double Gamma_g;   // th.Gamma * g
  double half_dy;   // 0.5 * elem.dy
  double gm1, gm1inv;
  double Sinv[4];   // 1 / stratification(y) of image rows 0, 1, cy-2, cy-1
};

__device__ __forceinline__ void y_image(double* rho, double* rhou, double* rhov, double* rhow,
                                        double* rhoY, double* rhoX, const double* rhoY0,
                                        int cx, int col, int img, int last, int src, double dir,
                                        double S, const BcParams& bp) {
  const long il = (long)last * cx + col, is = (long)src * cx + col, ii = (long)img * cx + col;
  const double Y_last = rhoY[il] / rho[il];
  const double rhoYv_image = -rhov[is] * rhoY[is] / rho[is];
  const double dpi = dir * bp.Gamma_g * bp.half_dy * (1.0 / Y_last + S);
  const double rY = bp.compressible ? pow(pow(rhoY[il], bp.gm1) + dpi, bp.gm1inv) : rhoY0[img];
  const double r = rY * S;
  const double v = rhoYv_image / rY;
  const double u = rhou[is] / rho[is];
  const double w = rhow[is] / rho[is];
  const double X = rhoX[is] / rho[is];
  rho[ii] = r;
  rhou[ii] = r * u;
  rhov[ii] = r * v;
  rhow[ii] = r * w;
  rhoY[ii] = rY;
  rhoX[ii] = r * X;
}

// x-axis ghost column ``col`` (0, 1, cx-2 or cx-1), inner row j: copy of the
// wrapped (periodic) or mirrored (wall) inner column
__device__ __forceinline__ void x_ghost_row(double* rho, double* rhou, double* rhov, double* rhow,
                                            double* rhoY, double* rhoX, int cx, int col, int j,
                                            const BcParams& bp) {
  int srcc;
  double vsign = 1.0;
  bool copy_w = true;
  if (bp.periodic_x) {
    srcc = col < 2 ? cx - 4 + col : col - (cx - 4);  // wrap over cx-4 inner columns
  } else {
    srcc = col < 2 ? 3 - col : 2 * cx - 5 - col;     // symmetric mirror
    vsign = -1.0;                                    // rhov negative-symmetric
    copy_w = false;                                  // rhow not padded on walls
  }
  const long a = (long)j * cx + col, b = (long)j * cx + srcc;
  rho[a] = rho[b];
  rhou[a] = rhou[b];
  rhov[a] = vsign * rhov[b];
  if (copy_w) rhow[a] = rhow[b];
  rhoY[a] = rhoY[b];
  rhoX[a] = rhoX[b];
}

// y ghost row ``row`` (0, 1, cy-2 or cy-1) of column ``col`` on an axis
// without gravity: copy of the wrapped (periodic) or mirrored (wall) inner row
// (boundary.py:40-50 with 135-216 along y)
__device__ __forceinline__ void y_ghost_row(double* rho, double* rhou, double* rhov, double* rhow,
                                            double* rhoY, double* rhoX, int cx, int cy, int col, int row,
                                            const BcParams& bp) {
  int srcr;
  double vsign = 1.0;
  bool copy_w = true;
  if (bp.periodic_y) {
    srcr = row < 2 ? cy - 4 + row : row - (cy - 4);
  } else {
    srcr = row < 2 ? 3 - row : 2 * cy - 5 - row;
    vsign = -1.0;
    copy_w = false;
  }
  const long a = (long)row * cx + col, b = (long)srcr * cx + col;
  rho[a] = rho[b];
  rhou[a] = rhou[b];
  rhov[a] = vsign * rhov[b];
  if (copy_w) rhow[a] = rhow[b];
  rhoY[a] = rhoY[b];
  rhoX[a] = rhoX[b];
}

// the y-axis fill of one column: hydrostatic images on the gravity axis
// (bottom rows 1 then 0, top rows cy-2 then cy-1), else wrap / mirror copies
__device__ __forceinline__ void y_fill_column(double* rho, double* rhou, double* rhov, double* rhow,
                                              double* rhoY, double* rhoX, const double* rhoY0, int cx,
                                              int cy, int col, const BcParams& bp) {
  if (!bp.gravity_y) {
    y_ghost_row(rho, rhou, rhov, rhow, rhoY, rhoX, cx, cy, col, 1, bp);
    y_ghost_row(rho, rhou, rhov, rhow, rhoY, rhoX, cx, cy, col, 0, bp);
    y_ghost_row(rho, rhou, rhov, rhow, rhoY, rhoX, cx, cy, col, cy - 2, bp);
    y_ghost_row(rho, rhou, rhov, rhow, rhoY, rhoX, cx, cy, col, cy - 1, bp);
    return;
  }
  // bottom (direction +1): image 1 <- source 2, image 0 <- source 3
  y_image(rho, rhou, rhov, rhow, rhoY, rhoX, rhoY0, cx, col, 1, 2, 2, 1.0, bp.Sinv[1], bp);
  y_image(rho, rhou, rhov, rhow, rhoY, rhoX, rhoY0, cx, col, 0, 1, 3, 1.0, bp.Sinv[0], bp);
  // top (direction -1)
  y_image(rho, rhou, rhov, rhow, rhoY, rhoX, rhoY0, cx, col, cy - 2, cy - 3, cy - 3, -1.0, bp.Sinv[2], bp);
  y_image(rho, rhou, rhov, rhow, rhoY, rhoX, rhoY0, cx, col, cy - 1, cy - 2, cy - 4, -1.0, bp.Sinv[3], bp);
}

// Threads 0 .. cx-1: column ``col``'s y-axis fill (rows 0, 1, cy-2, cy-1),
// after the x-axis fill of the inner rows it reads (2, 3, cy-4, cy-3) when
// col is an x-ghost column.  Threads cx .. cx + 4 (cy-4) - 1: the x-axis fill
// of the other inner rows of the 4 ghost columns, one row each.  No thread
// reads a value another thread writes: the x fill reads inner columns, the y
// fill reads its own column.
__global__ void k_bc_cells(Dims2 d, BcParams bp, double* rho, double* rhou, double* rhov,
                           double* rhow, double* rhoY, double* rhoX, const double* rhoY0) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int cx = d.cx, cy = d.cy;
  auto image_source_row = [&](int j) { return j == 2 || j == 3 || j == cy - 4 || j == cy - 3; };
  if (t >= cx) {
    const int u = t - cx, g = u & 3, j = 2 + (u >> 2);
    if (j >= cy - 2 || image_source_row(j)) return;
    const int col = g < 2 ? g : cx - 4 + g;
    x_ghost_row(rho, rhou, rhov, rhow, rhoY, rhoX, cx, col, j, bp);
    return;
  }
  const int col = t;
  if (col < 2 || col >= cx - 2) {
    const int rows[4] = {2, 3, cy - 4, cy - 3};
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (rows[k] >= 2 && rows[k] < cy - 2) x_ghost_row(rho, rhou, rhov, rhow, rhoY, rhoX, cx, col, rows[k], bp);
  }
  y_fill_column(rho, rhou, rhov, rhow, rhoY, rhoX, rhoY0, cx, cy, col, bp);
}

// The fill along one axis only, as the advection's split steps call it
// (boundary.py:30-39): axis 0 fills the 4 ghost columns over every row
// (thread = row * 4 + ghost column), axis 1 the y ghost rows of every column.
__global__ void k_bc_axis(Dims2 d, BcParams bp, int axis, double* rho, double* rhou, double* rhov,
                          double* rhow, double* rhoY, double* rhoX, const double* rhoY0) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int cx = d.cx, cy = d.cy;
  if (axis == 0) {
    if (t >= 4 * cy) return;
    const int j = t >> 2, g = t & 3;
    x_ghost_row(rho, rhou, rhov, rhow, rhoY, rhoX, cx, g < 2 ? g : cx - 4 + g, j, bp);
    return;
  }
  if (t >= cx) return;
  y_fill_column(rho, rhou, rhov, rhow, rhoY, rhoX, rhoY0, cx, cy, t, bp);
}

// ---------------------------------------------------------------------------
// set_ghostnodes_p2 on a node array.  Axis by axis as the reference: x over
// every row (including y-ghost rows), then y over every column.
// periodic_plus_one (boundary.py:348-369) on a padded line of n inner values
// at offsets 2..n+1: ghost 0,1 <- inner n-3,n-2; inner 0 <- inner n-1;
// inner n-1 <- old inner 0; ghost n+2,n+3 <- inner 1,2.  'reflect' mirrors
// about the end node.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void fill_line(double* base, long stride, int n, int periodic) {
  double* v = base;
  if (periodic) {
    const double a0 = v[(n - 1) * stride], a1 = v[n * stride], a2 = v[(n + 1) * stride];
    const double b0 = v[2 * stride], b1 = v[3 * stride], b2 = v[4 * stride];
    v[0] = a0;
    v[stride] = a1;
    v[2 * stride] = a2;
    v[(n + 1) * stride] = b0;
    v[(n + 2) * stride] = b1;
    v[(n + 3) * stride] = b2;
  } else {
    v[stride] = v[3 * stride];
    v[0] = v[4 * stride];
    v[(n + 2) * stride] = v[n * stride];
    v[(n + 3) * stride] = v[(n - 1) * stride];
  }
}

__global__ void k_ghost_x(Dims2 d, double* p, int periodic) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= d.nyn) return;
  fill_line(p + (long)row * d.nxn, 1, d.nx, periodic);
}

__global__ void k_ghost_y(Dims2 d, double* p, int periodic) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= d.nxn) return;
  fill_line(p + col, d.nxn, d.ny, periodic);
}

// set_ghostcells_p2 on a cell array (boundary.py:284-293): np.pad 'wrap'
// (periodic) or 'symmetric' (wall) of the two ghost layers, x over every row,
// then y over every column.
__global__ void k_ghostcells_x(Dims2 d, double* p, int periodic) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= d.cy) return;
  const int cx = d.cx;
  double* v = p + (long)row * cx;
  if (periodic) {
    v[0] = v[cx - 4];
    v[1] = v[cx - 3];
    v[cx - 2] = v[2];
    v[cx - 1] = v[3];
  } else {
    v[1] = v[2];
    v[0] = v[3];
    v[cx - 2] = v[cx - 3];
    v[cx - 1] = v[cx - 4];
  }
}

__global__ void k_ghostcells_y(Dims2 d, double* p, int periodic) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= d.cx) return;
  const long s = d.cx;
  double* v = p + col;
  const int cy = d.cy;
  if (periodic) {
    v[0] = v[(cy - 4) * s];
    v[s] = v[(cy - 3) * s];
    v[(cy - 2) * s] = v[2 * s];
    v[(cy - 1) * s] = v[3 * s];
  } else {
    v[s] = v[2 * s];
    v[0] = v[3 * s];
    v[(cy - 2) * s] = v[(cy - 3) * s];
    v[(cy - 1) * s] = v[(cy - 4) * s];
  }
}

__global__ void k_scatter_p2(Dims2 d, const double* x, double* pf) {
  const long n = (long)d.nxn * d.nyn;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    const int j = (int)(k / d.nxn), i = (int)(k % d.nxn);
    const int jj = j - 2, ii = i - 2;
    pf[k] = (jj >= 0 && jj < d.ny && ii >= 0 && ii < d.nx) ? x[(long)jj * d.nx + ii] : 0.0;
  }
}

__global__ void k_add(long n, const double* a, double* b) {
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x)
    b[k] += a[k];
}

// ---------------------------------------------------------------------------
// divergence_nodes
// ---------------------------------------------------------------------------
__global__ void k_zero_wall_momenta(Dims2 d, double* rhou, double* rhov, double* rhow) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= d.cx) return;
  const int rows[4] = {0, 1, d.cy - 2, d.cy - 1};
  for (int k = 0; k < 4; ++k) {
    const long a = (long)rows[k] * d.cx + col;
    rhou[a] = 0.0;
    rhov[a] = 0.0;
    rhow[a] = 0.0;
  }
}

// rhs[b][a] (isc) = (Ux + Vy) / dt with
//   Ux = 0.5 ((U[a+1,b]-U[a,b])/dx + (U[a+1,b+1]-U[a,b+1])/dx),
//   Vy = 0.5 ((V[a,b+1]-V[a,b])/dy + (V[a+1,b+1]-V[a+1,b])/dy),
// U = rhou*Y, V = rhov*Y, Y = rhoY/rho (second_projection.py:554-572).
__global__ void k_divergence(Dims2 d, double dx, double dy, double dt,
                             const double* rho, const double* rhou, const double* rhov,
                             const double* rhoY, double* rhs) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y * blockDim.y + threadIdx.y;
  if (a >= d.ix || b >= d.iy) return;
  const int cx = d.cx;
  const long c00 = (long)b * cx + a, c10 = c00 + 1, c01 = c00 + cx, c11 = c01 + 1;
  const double Y00 = rhoY[c00] / rho[c00], Y10 = rhoY[c10] / rho[c10];
  const double Y01 = rhoY[c01] / rho[c01], Y11 = rhoY[c11] / rho[c11];
  const double U00 = rhou[c00] * Y00, U10 = rhou[c10] * Y10, U01 = rhou[c01] * Y01, U11 = rhou[c11] * Y11;
  const double V00 = rhov[c00] * Y00, V10 = rhov[c10] * Y10, V01 = rhov[c01] * Y01, V11 = rhov[c11] * Y11;
  const double Ux = 0.5 * ((U10 - U00) / dx + (U11 - U01) / dx);
  const double Vy = 0.5 * ((V01 - V00) / dy + (V11 - V10) / dy);
  rhs[(long)b * d.ix + a] = (Ux + Vy) / dt;
}

// ---------------------------------------------------------------------------
// correction_nodes (2D).  Per cell (a, b) of the (nx+3) x (ny+3) cell array,
// with p on nodes (corners (a..a+1, b..b+1)).
// ---------------------------------------------------------------------------
struct CorrParams {
  double dx, dy, dt, Gammainv, updt_chi;
  double g_over_Msq, nonhydro, wh1, wv, wh2;
};

__global__ void k_correction(Dims2 d, CorrParams cp, const double* p, const double* rho,
                             double* rhou, double* rhov, const double* rhoY, double* rhoX,
                             const double* dSdy, double* u_out, double* v_out, double* w_out) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y * blockDim.y + threadIdx.y;
  if (a >= d.cx || b >= d.cy) return;
  const int nxn = d.nxn;
  const double p00 = p[(long)b * nxn + a], p10 = p[(long)b * nxn + a + 1];
  const double p01 = p[(long)(b + 1) * nxn + a], p11 = p[(long)(b + 1) * nxn + a + 1];
  // grad_nodes: signs_x (-,-,+,+), signs_y (-,+,-,+) over (00, 01, 10, 11) in [x,y]
  double Dpx = -p00;
  Dpx += -p01;
  Dpx += p10;
  Dpx += p11;
  double Dpy = -p00;
  Dpy += p01;
  Dpy += -p10;
  Dpy += p11;
  Dpx *= 0.5 / cp.dx;
  Dpy *= 0.5 / cp.dy;
  const double Dpz = 0.0;
  const long c = (long)b * d.cx + a;
  const double r = rho[c], rY = rhoY[c];
  const double thinv = r / rY;
  const double Y = rY / r;
  const double coeff = cp.Gammainv * rY * Y;
  const double u = -cp.dt * coeff * Dpx;
  const double v = -cp.dt * coeff * Dpy;
  const double w = -cp.dt * coeff * Dpz;
  // multiply_inverse_coriolis (second_projection.py:604-646)
  const double nh = cp.nonhydro, wh1 = cp.wh1, wv = cp.wv, wh2 = cp.wh2;
  const double nu = -cp.dt * cp.dt * cp.g_over_Msq * dSdy[c] * Y;
  const double denom = 1.0 / (wh1 * wh1 + wh2 * wh2 + (nu + nh) * (wv * wv + 1.0));
  const double uu = wh1 * wh1 + nu + nh, uv = nh * (wh1 * wv + wh2), uw = wh1 * wh2 - (nu + nh) * wv;
  const double vu = wh1 * wv - wh2, vv = nh * (1 + wv * wv), vw = wh2 * wv + wh1;
  const double wu = wh1 * wh2 + (nu + nh) * wv, wvv = nh * (wh2 * wv - wh1), ww = nu + nh + wh2 * wh2;
  const double U = denom * (uu * u + uv * v + uw * w);
  const double V = denom * (vu * u + vv * v + vw * w);
  const double W = denom * (wu * u + wvv * v + ww * w);
  if (u_out) u_out[c] = U;
  if (v_out) v_out[c] = V;
  if (w_out) w_out[c] = W;
  rhou[c] += thinv * U;
  const double rv_new = rhov[c] + thinv * V;
  rhov[c] = rv_new;
  rhoX[c] += -cp.updt_chi * cp.dt * dSdy[c] * rv_new;
}

BcParams bc_params(const pb_grid2d& g, const pb_phys& ph) {
  const double gm1 = ph.gamm - 1.0;
  BcParams bp{g.bc_x == PB_BC_PERIODIC, ph.g != 0.0, g.bc_y == PB_BC_PERIODIC, ph.is_compressible == 1,
              (gm1 / ph.gamm) * ph.g, 0.5 * g.dy, gm1, 1.0 / gm1, {1.0, 1.0, 1.0, 1.0}};
  if (ph.stratified)
    for (int k = 0; k < 4; ++k) bp.Sinv[k] = ph.strat_inv[k];
  return bp;
}

}  // namespace

extern "C" int pb_bc_cells(pb_ctx* ctx, const pb_grid2d* grid, const pb_phys* phys,
                           double* const fields[6], const double* rhoY0_bg) {
  if (!ctx || !grid || !phys || !fields) return pb_fail(ctx, PB_ERR_ARG, "pb_bc_cells: null argument");
  if (grid->nx < 3 || grid->ny < 5) return pb_fail(ctx, PB_ERR_ARG, "pb_bc_cells: grid too small");
  for (int k = 0; k < 6; ++k)
    if (!fields[k]) return pb_fail(ctx, PB_ERR_ARG, "pb_bc_cells: field %d is null", k);
  if (phys->g != 0.0 && !phys->is_compressible && !rhoY0_bg)
    return pb_fail(ctx, PB_ERR_ARG, "pb_bc_cells: background rhoY0 required when not compressible");
  const Dims2 d = dims_of(*grid);
  const BcParams bp = bc_params(*grid, *phys);
  const int nthr = d.cx + 4 * (d.cy - 4);
  hipLaunchKernelGGL(k_bc_cells, dim3(ceil_div(nthr, 256)), dim3(256), 0, ctx->stream, d, bp,
                     fields[0], fields[1], fields[2], fields[3], fields[4], fields[5], rhoY0_bg);
  PB_LAUNCH_CHECK(ctx);
  return PB_OK;
}

extern "C" int pb_bc_cells_axis(pb_ctx* ctx, const pb_grid2d* grid, const pb_phys* phys,
                                double* const fields[6], const double* rhoY0_bg, int axis) {
  if (!ctx || !grid || !phys || !fields) return pb_fail(ctx, PB_ERR_ARG, "pb_bc_cells_axis: null argument");
  if (axis < 0) return pb_bc_cells(ctx, grid, phys, fields, rhoY0_bg);
  if (axis > 1) return pb_fail(ctx, PB_ERR_ARG, "pb_bc_cells_axis: axis %d (0, 1 or -1)", axis);
  if (grid->nx < 3 || grid->ny < 5) return pb_fail(ctx, PB_ERR_ARG, "pb_bc_cells_axis: grid too small");
  for (int k = 0; k < 6; ++k)
    if (!fields[k]) return pb_fail(ctx, PB_ERR_ARG, "pb_bc_cells_axis: field %d is null", k);
  if (axis == 1 && phys->g != 0.0 && !phys->is_compressible && !rhoY0_bg)
    return pb_fail(ctx, PB_ERR_ARG, "pb_bc_cells_axis: background rhoY0 required when not compressible");
  const Dims2 d = dims_of(*grid);
  const BcParams bp = bc_params(*grid, *phys);
  const int nthr = axis == 0 ? 4 * d.cy : d.cx;
  hipLaunchKernelGGL(k_bc_axis, dim3(ceil_div(nthr, 256)), dim3(256), 0, ctx->stream, d, bp, axis,
                     fields[0], fields[1], fields[2], fields[3], fields[4], fields[5], rhoY0_bg);
  PB_LAUNCH_CHECK(ctx);
  return PB_OK;
}

extern "C" int pb_ghostnodes(pb_ctx* ctx, const pb_grid2d* grid, double* p) {
  if (!ctx || !grid || !p) return pb_fail(ctx, PB_ERR_ARG, "pb_ghostnodes: null argument");
  if (grid->nx < 3 || grid->ny < 3) return pb_fail(ctx, PB_ERR_ARG, "pb_ghostnodes: grid too small");
  const Dims2 d = dims_of(*grid);
  hipLaunchKernelGGL(k_ghost_x, dim3(ceil_div(d.nyn, 64)), dim3(64), 0, ctx->stream, d, p,
                     grid->bc_x == PB_BC_PERIODIC);
  hipLaunchKernelGGL(k_ghost_y, dim3(ceil_div(d.nxn, 64)), dim3(64), 0, ctx->stream, d, p,
                     grid->bc_y == PB_BC_PERIODIC);
  PB_LAUNCH_CHECK(ctx);
  return PB_OK;
}

extern "C" int pb_ghostcells_p2(pb_ctx* ctx, const pb_grid2d* grid, double* p) {
  if (!ctx || !grid || !p) return pb_fail(ctx, PB_ERR_ARG, "pb_ghostcells_p2: null argument");
  if (grid->nx < 3 || grid->ny < 3) return pb_fail(ctx, PB_ERR_ARG, "pb_ghostcells_p2: grid too small");
  const Dims2 d = dims_of(*grid);
  hipLaunchKernelGGL(k_ghostcells_x, dim3(ceil_div(d.cy, 64)), dim3(64), 0, ctx->stream, d, p,
                     grid->bc_x == PB_BC_PERIODIC);
  hipLaunchKernelGGL(k_ghostcells_y, dim3(ceil_div(d.cx, 64)), dim3(64), 0, ctx->stream, d, p,
                     grid->bc_y == PB_BC_PERIODIC);
  PB_LAUNCH_CHECK(ctx);
  return PB_OK;
}

extern "C" int pb_divergence_nodes(pb_ctx* ctx, const pb_grid2d* grid, double dt, double* rho,
                                   double* rhou, double* rhov, double* rhow, const double* rhoY,
                                   double* rhs) {
  if (!ctx || !grid || !rho || !rhou || !rhov || !rhow || !rhoY || !rhs)
    return pb_fail(ctx, PB_ERR_ARG, "pb_divergence_nodes: null argument");
  const Dims2 d = dims_of(*grid);
  if (grid->bc_y == PB_BC_WALL) {
    hipLaunchKernelGGL(k_zero_wall_momenta, dim3(ceil_div(d.cx, 64)), dim3(64), 0, ctx->stream, d,
                       rhou, rhov, rhow);
  }
  dim3 blk(64, 4);
  dim3 grd(ceil_div(d.ix, 64), ceil_div(d.iy, 4));
  hipLaunchKernelGGL(k_divergence, grd, blk, 0, ctx->stream, d, grid->dx, grid->dy, dt, rho, rhou,
                     rhov, rhoY, rhs);
  PB_LAUNCH_CHECK(ctx);
  return PB_OK;
}

extern "C" int pb_correction_nodes(pb_ctx* ctx, const pb_grid2d* grid, const pb_phys* phys, double dt,
                                   double updt_chi, const double* p, const double* rho, double* rhou,
                                   double* rhov, const double* rhoY, double* rhoX, const double* dSdy,
                                   double* u, double* v, double* w) {
  if (!ctx || !grid || !phys || !p || !rho || !rhou || !rhov || !rhoY || !rhoX || !dSdy)
    return pb_fail(ctx, PB_ERR_ARG, "pb_correction_nodes: null argument");
  const Dims2 d = dims_of(*grid);
  const double gm1 = phys->gamm - 1.0;
  CorrParams cp{grid->dx, grid->dy, dt, phys->gamm / gm1, updt_chi, phys->g / phys->Msq,
                phys->nonhydro, dt * phys->coriolis[0], dt * phys->coriolis[1], dt * phys->coriolis[2]};
  dim3 blk(64, 4);
  dim3 grd(ceil_div(d.cx, 64), ceil_div(d.cy, 4));
  hipLaunchKernelGGL(k_correction, grd, blk, 0, ctx->stream, d, cp, p, rho, rhou, rhov, rhoY, rhoX,
                     dSdy, u, v, w);
  PB_LAUNCH_CHECK(ctx);
  return PB_OK;
}

extern "C" int pb_scatter_update_p2(pb_ctx* ctx, const pb_grid2d* grid, const double* x,
                                    double* p2_full, double* p2_nodes) {
  if (!ctx || !grid || !x || !p2_full || !p2_nodes)
    return pb_fail(ctx, PB_ERR_ARG, "pb_scatter_update_p2: null argument");
  const Dims2 d = dims_of(*grid);
  const long n = (long)d.nxn * d.nyn;
  const int nb = (int)std::min<long>(ceil_div((int)n, 256), 2048);
  hipLaunchKernelGGL(k_scatter_p2, dim3(nb), dim3(256), 0, ctx->stream, d, x, p2_full);
  int rc = pb_ghostnodes(ctx, grid, p2_full);
  if (rc) return rc;
  hipLaunchKernelGGL(k_add, dim3(nb), dim3(256), 0, ctx->stream, n, (const double*)p2_full, p2_nodes);
  PB_LAUNCH_CHECK(ctx);
  return pb_ghostnodes(ctx, grid, p2_nodes);
}
