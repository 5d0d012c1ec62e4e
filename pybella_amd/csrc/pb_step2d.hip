// Explicit parts of the 2D time step (gfx950): advective fluxes, the
// recovery + HLL advection split steps, the explicit buoyancy / Coriolis
// parts and the CFL maxima.
//
// Layout (row-major, x fastest; NX, NY inner node counts):
//   cells   (NY+3) x (NX+3)      Sol fields
//   x faces (NY+3) x (NX+4)      flux[0] fields, [y][face]: the reference's
//                                flux[0] array (icy, icx+1) as it lies
//   y faces (NY+4) x (NX+3)      flux[1] fields, [face][x]: the transpose of
//                                the reference's flux[1] (icx, icy+1)
// Face f of a line lies between its cells f-1 and f; faces 0 and n of a line
// of n cells are never written (the reference leaves them at zero).
//
// Reference counterparts (pyBELLA tree):
//   k_adv_flux       recompute_advective_fluxes    numerical_flux.py:5-48
//   k_hll<AXIS>      recovery.do + slopes + limiters(NONE) + get_conservatives
//                    + hll_solver                  recovery.py:6-210,
//                                                  numerical_flux.py:50-99
//   k_update<AXIS>   the conservative update of explicit_step_and_flux
//                                                  explicit.py:150-156
//   k_update_rk      the update loop of advect_rk  explicit.py:219-235
//   k_expl           euler_backward_non_advective_expl_part
//                                                  second_projection.py:72-90
//   k_scale_wall, k_efna_cells, k_efna_nodes
//                    euler_forward_non_advective   second_projection.py:22-69
//   k_cfl_partial, k_cfl_final
//                    the maxima of dynamic_timestep cfl.py:5-35
//
// Every formula keeps the reference's operation order, so that results equal
// the NumPy restatement up to the rounding of pow (ulps) and of the box
// filters the reference evaluates by FFT.
#include "pb_common.hpp"

extern "C" int pb_bc_cells(pb_ctx* ctx, const pb_grid2d* grid, const pb_phys* phys,
                           double* const fields[6], const double* rhoY0_bg);
extern "C" int pb_bc_cells_axis(pb_ctx* ctx, const pb_grid2d* grid, const pb_phys* phys,
                                double* const fields[6], const double* rhoY0_bg, int axis);
extern "C" int pb_ghostnodes(pb_ctx* ctx, const pb_grid2d* grid, double* p);
extern "C" int pb_divergence_nodes(pb_ctx* ctx, const pb_grid2d* grid, double dt, double* rho,
                                   double* rhou, double* rhov, double* rhow, const double* rhoY,
                                   double* rhs);

namespace {

struct Dims {
  int cx, cy;   // cell array extents (nx+3, ny+3)
  int nxn;      // node array width (nx+4)
  int ix, iy;   // isc extents (nx+2, ny+2)
};

Dims dims_of(const pb_grid2d& g) { return Dims{g.nx + 3, g.ny + 3, g.nx + 4, g.nx + 2, g.ny + 2}; }

struct Fields {   // rho, rhou, rhov, rhow, rhoY, rhoX
  double* f[6];
};

// ---------------------------------------------------------------------------
// recompute_advective_fluxes: rhoY u on x faces, rhoY v on y faces, each the
// 2 x 3 cell average with weights (1/2, 1, 1/2) across the face, over 4.
// Thread (X, Y) covers x face (row Y, face X) and y face (face Y, column X).
// ---------------------------------------------------------------------------
__global__ void k_adv_flux(Dims d, const double* __restrict__ rho, const double* __restrict__ rhou,
                           const double* __restrict__ rhov, const double* __restrict__ rhoY,
                           double* __restrict__ fx, double* __restrict__ fy) {
  const int X = blockIdx.x * blockDim.x + threadIdx.x;
  const int Y = blockIdx.y * blockDim.y + threadIdx.y;
  const int cx = d.cx, cy = d.cy;
  if (Y >= 1 && Y <= cy - 2 && X >= 1 && X <= cx - 1) {
    auto a = [&](int i, int j) {
      const long c = (long)j * cx + i;
      return rhoY[c] * rhou[c] / rho[c];
    };
    const int f = X, j = Y;
    double s = 0.5 * a(f - 1, j - 1);
    s += a(f - 1, j);
    s += 0.5 * a(f - 1, j + 1);
    s += 0.5 * a(f, j - 1);
    s += a(f, j);
    s += 0.5 * a(f, j + 1);
    fx[(long)j * (cx + 1) + f] = s / 4.0;
  }
  if (Y >= 1 && Y <= cy - 1 && X >= 1 && X <= cx - 2) {
    auto b = [&](int i, int j) {
      const long c = (long)j * cx + i;
      return rhoY[c] * rhov[c] / rho[c];
    };
    const int f = Y, i = X;
    double s = 0.5 * b(i - 1, f - 1);
    s += b(i, f - 1);
    s += 0.5 * b(i + 1, f - 1);
    s += 0.5 * b(i - 1, f);
    s += b(i, f);
    s += 0.5 * b(i + 1, f);
    fy[(long)f * cx + i] = s / 4.0;
  }
}

// ---------------------------------------------------------------------------
// One advection split step along AXIS (0: x, lines are rows; 1: y, lines are
// columns).
// ---------------------------------------------------------------------------
struct Prim {   // primitives of one cell (variable.py:44-89) and its rhoY
  double u, v, w, X, Yi, rhoY;
};

__device__ __forceinline__ Prim load_prim(const Fields& s, long c) {
  const double r = s.f[0][c];
  Prim p;
  p.rhoY = s.f[4][c];
  double Y = 0.0;
  if (r != 0.0) {
    p.u = s.f[1][c] / r;
    p.v = s.f[2][c] / r;
    p.w = s.f[3][c] / r;
    Y = p.rhoY / r;
    p.X = s.f[5][c] / r;
  } else {
    p.u = p.v = p.w = p.X = 0.0;
  }
  p.Yi = 1.0 / Y;
  return p;
}

// central slope of cell c from its neighbours (limiter NONE):
// ((q[c]-q[c-1]) + (q[c+1]-q[c])) / 2
__device__ __forceinline__ double slope(double qm, double q0, double qp) {
  return 0.5 * ((q0 - qm) + (qp - q0));
}

struct FaceState {   // Lefts / Rights after get_conservatives and primitives
  double u, v, w, X, Y;
};

// get_conservatives (recovery.py:192-210) then primitives (variable.py:44-89)
__device__ __forceinline__ FaceState conserve_prim(double rhoY_face, double u, double v, double w, double X,
                                                   double Y) {
  const double rho = rhoY_face / Y;
  const double ru = u * rho, rv = v * rho, rw = w * rho, rY = Y * rho, rX = X * rho;
  FaceState o;
  if (rho != 0.0) {
    o.u = ru / rho;
    o.v = rv / rho;
    o.w = rw / rho;
    o.Y = rY / rho;
    o.X = rX / rho;
  } else {
    o.u = o.v = o.w = o.Y = o.X = 0.0;
  }
  return o;
}

__device__ __forceinline__ double sign_of(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }

template <int AXIS>
__global__ void k_hll(Dims d, double lam, Fields s, const double* __restrict__ Fa, double* __restrict__ Frho,
                      double* __restrict__ Frhou, double* __restrict__ Frhov, double* __restrict__ Frhow,
                      double* __restrict__ FrhoX) {
  const int X = blockIdx.x * blockDim.x + threadIdx.x;
  const int Yt = blockIdx.y * blockDim.y + threadIdx.y;
  const int f = AXIS == 0 ? X : Yt;    // face along the line
  const int l = AXIS == 0 ? Yt : X;    // line
  const int n = AXIS == 0 ? d.cx : d.cy;
  const int m = AXIS == 0 ? d.cy : d.cx;
  if (l >= m || f < 1 || f > n - 1) return;
  auto cidx = [&](int c) -> long { return AXIS == 0 ? (long)l * d.cx + c : (long)c * d.cx + l; };
  auto fidx = [&](int g) -> long { return AXIS == 0 ? (long)l * (d.cx + 1) + g : (long)g * d.cx + l; };
  // cells f-2, f-1 (left), f (right), f+1
  Prim c0, c1, c2, c3;
  c1 = load_prim(s, cidx(f - 1));
  c2 = load_prim(s, cidx(f));
  const bool sl = f - 1 >= 1 && f - 1 <= n - 2;   // left cell has a slope
  const bool sr = f >= 1 && f <= n - 2;           // right cell has a slope
  if (sl) c0 = load_prim(s, cidx(f - 2));
  if (sr) c3 = load_prim(s, cidx(f + 1));
  const bool inner_line = l >= 1 && l <= m - 2;
  const double Ff = Fa[fidx(f)];
  // face-averaged advection velocity of the left and right cells
  const double ufl = inner_line && sl ? 0.5 * (Fa[fidx(f - 1)] + Ff) / c1.rhoY : 0.0;
  const double ufr = inner_line && sr ? 0.5 * (Ff + Fa[fidx(f + 1)]) / c2.rhoY : 0.0;
  const double al = 1.0 - lam * ufl, ar = 1.0 + lam * ufr;
  // left state: cell f-1 extrapolated to its right face
  const double su_l = sl ? slope(c0.u, c1.u, c2.u) : 0.0;
  const double sv_l = sl ? slope(c0.v, c1.v, c2.v) : 0.0;
  const double sw_l = sl ? slope(c0.w, c1.w, c2.w) : 0.0;
  const double sX_l = sl ? slope(c0.X, c1.X, c2.X) : 0.0;
  const double sY_l = sl ? slope(c0.Yi, c1.Yi, c2.Yi) : 0.0;
  const double Lu = c1.u + 0.5 * su_l * al;
  const double Lv = c1.v + 0.5 * sv_l * al;
  const double Lw = c1.w + 0.5 * sw_l * al;
  const double LX = c1.X + 0.5 * sX_l * al;
  const double LY = 1.0 / (c1.Yi + 0.5 * sY_l * al);
  // right state: cell f extrapolated to its left face
  const double su_r = sr ? slope(c1.u, c2.u, c3.u) : 0.0;
  const double sv_r = sr ? slope(c1.v, c2.v, c3.v) : 0.0;
  const double sw_r = sr ? slope(c1.w, c2.w, c3.w) : 0.0;
  const double sX_r = sr ? slope(c1.X, c2.X, c3.X) : 0.0;
  const double sY_r = sr ? slope(c1.Yi, c2.Yi, c3.Yi) : 0.0;
  const double Ru = c2.u + -0.5 * su_r * ar;
  const double Rv = c2.v + -0.5 * sv_r * ar;
  const double Rw = c2.w + -0.5 * sw_r * ar;
  const double RX = c2.X + -0.5 * sX_r * ar;
  const double RY = 1.0 / (c2.Yi + -0.5 * sY_r * ar);
  // the face's rhoY (recovery.py:94-95), shared by both sides
  const double vl = AXIS == 0 ? c1.u : c1.v, vr = AXIS == 0 ? c2.u : c2.v;
  const double rYf = 0.5 * (c1.rhoY + c2.rhoY) - 0.5 * lam * (vr * c2.rhoY - vl * c1.rhoY);
  const FaceState L = conserve_prim(rYf, Lu, Lv, Lw, LX, LY);
  const FaceState R = conserve_prim(rYf, Ru, Rv, Rw, RX, RY);
  const double up = 0.5 * (1.0 + sign_of(Ff));
  const double a = up / L.Y, b = (1.0 - up) / R.Y;
  const long o = fidx(f);
  Frhou[o] = Ff * (a * L.u + b * R.u);
  Frho[o] = Ff * (a * 1.0 + b * 1.0);
  Frhov[o] = Ff * (a * L.v + b * R.v);
  Frhow[o] = Ff * (a * L.w + b * R.w);
  FrhoX[o] = Ff * (a * L.X + b * R.X);
}

struct FluxPtrs {   // face fluxes rho, rhou, rhov, rhow, rhoY (advective), rhoX
  const double* f[6];
};

template <int AXIS>
__global__ void k_update(Dims d, double lam, Fields s, FluxPtrs F) {
  const int X = blockIdx.x * blockDim.x + threadIdx.x;
  const int Y = blockIdx.y * blockDim.y + threadIdx.y;
  if (X >= d.cx || Y >= d.cy) return;
  const long c = (long)Y * d.cx + X;
  const long fl = AXIS == 0 ? (long)Y * (d.cx + 1) + X : (long)Y * d.cx + X;
  const long fr = AXIS == 0 ? fl + 1 : fl + d.cx;
#pragma unroll
  for (int k = 0; k < 6; ++k) s.f[k][c] = s.f[k][c] + lam * (F.f[k][fl] - F.f[k][fr]);
}

__global__ void k_update_rk(Dims d, double lx, double ly, Fields s, FluxPtrs Fx, FluxPtrs Fy) {
  const int X = blockIdx.x * blockDim.x + threadIdx.x;
  const int Y = blockIdx.y * blockDim.y + threadIdx.y;
  if (X >= d.cx || Y >= d.cy) return;
  const long c = (long)Y * d.cx + X;
  const long xl = (long)Y * (d.cx + 1) + X, yl = (long)Y * d.cx + X;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const double q = s.f[k][c] + lx * (Fx.f[k][xl] - Fx.f[k][xl + 1]);
    s.f[k][c] = q + ly * (Fy.f[k][yl] - Fy.f[k][yl + d.cx]);
  }
}

// ---------------------------------------------------------------------------
// euler_backward_non_advective_expl_part: buoyancy, then the inverse of the
// implicit Coriolis / buoyancy matrix (second_projection.py:604-646) on the
// momenta relative to the background wind.
// ---------------------------------------------------------------------------
struct ExplParams {
  double nonhydro, dt_gM;   // nonhydrostasy, dt * (g / Msq)
  double c_nu;              // -dt^2 * (g / Msq)
  double wh1, wv, wh2;      // dt * coriolis_strength
  double u0, v0, w0;        // background wind
};

__global__ void k_expl(Dims d, ExplParams p, Fields s, const double* __restrict__ dSdy) {
  const int X = blockIdx.x * blockDim.x + threadIdx.x;
  const int Y = blockIdx.y * blockDim.y + threadIdx.y;
  if (X >= d.cx || Y >= d.cy) return;
  const long c = (long)Y * d.cx + X;
  const double r = s.f[0][c], rY = s.f[4][c], rX = s.f[5][c];
  double ru = s.f[1][c], rv = s.f[2][c], rw = s.f[3][c];
  const double dbuoy = rY * (rX / r);
  rv = (p.nonhydro * rv) - p.dt_gM * dbuoy;
  ru = ru + -1.0 * p.u0 * r;
  rv = rv + -1.0 * p.v0 * r;
  rw = rw + -1.0 * p.w0 * r;
  const double Y_ = rY / r;
  const double nu = p.c_nu * dSdy[c] * Y_;
  const double nh = p.nonhydro, wh1 = p.wh1, wv = p.wv, wh2 = p.wh2;
  const double denom = 1.0 / (wh1 * wh1 + wh2 * wh2 + (nu + nh) * (wv * wv + 1.0));
  const double uu = wh1 * wh1 + nu + nh, uv = nh * (wh1 * wv + wh2), uw = wh1 * wh2 - (nu + nh) * wv;
  const double vu = wh1 * wv - wh2, vv = nh * (1 + wv * wv), vw = wh2 * wv + wh1;
  const double wu = wh1 * wh2 + (nu + nh) * wv, wvv = nh * (wh2 * wv - wh1), ww = nu + nh + wh2 * wh2;
  const double U = denom * (uu * ru + uv * rv + uw * rw);
  const double V = denom * (vu * ru + vv * rv + vw * rw);
  const double W = denom * (wu * ru + wvv * rv + ww * rw);
  s.f[1][c] = U + 1.0 * p.u0 * r;
  s.f[2][c] = V + 1.0 * p.v0 * r;
  s.f[3][c] = W + 1.0 * p.w0 * r;
}

// ---------------------------------------------------------------------------
// euler_forward_non_advective
// ---------------------------------------------------------------------------
// scale_wall_node_values (second_projection.py:433-464) on an isc node array:
// x walls scale columns 1 and nx over rows 2..ny-1; y walls scale rows 1 and
// ny over columns 2..nx-1, or over column 1 alone when x has walls too (the
// reference's index list keeps the x wall's last index).
__global__ void k_scale_wall(Dims d, int x_wall, int y_wall, double factor, double* rhs) {
  const int ia = blockIdx.x * blockDim.x + threadIdx.x;
  const int ib = blockIdx.y * blockDim.y + threadIdx.y;
  if (ia >= d.ix || ib >= d.iy) return;
  const int nx = d.ix - 2, ny = d.iy - 2;
  double f = 1.0;
  bool hit = false;
  if (x_wall && (ia == 1 || ia == nx) && ib >= 2 && ib <= ny - 1) hit = true;
  if (hit) f = factor;
  double v = rhs[(long)ib * d.ix + ia];
  if (hit) v *= f;
  if (y_wall && (ib == 1 || ib == ny)) {
    const bool on = x_wall ? (ia == 1) : (ia >= 2 && ia <= nx - 1);
    if (on) v *= factor;
  }
  rhs[(long)ib * d.ix + ia] = v;
}

struct EfnaParams {
  double dt, dx, dy;
  double Ginv, gM, nonhydro;
  double ch1, cv, ch2;       // coriolis_strength
  double u0, v0, w0;
  double ak;                 // 1 - is_ArakawaKonor
};

// momenta and rhoX of every cell (second_projection.py:52-63); p is the node
// pressure before this step's Exner increment
__global__ void k_efna_cells(Dims d, EfnaParams q, Fields s, const double* __restrict__ p,
                             const double* __restrict__ S0c, const double* __restrict__ dSdy) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y * blockDim.y + threadIdx.y;
  if (a >= d.cx || b >= d.cy) return;
  const int nxn = d.nxn;
  const double p00 = p[(long)b * nxn + a], p10 = p[(long)b * nxn + a + 1];
  const double p01 = p[(long)(b + 1) * nxn + a], p11 = p[(long)(b + 1) * nxn + a + 1];
  double Dpx = -p00;
  Dpx += -p01;
  Dpx += p10;
  Dpx += p11;
  double Dpy = -p00;
  Dpy += p01;
  Dpy += -p10;
  Dpy += p11;
  const double dpdx = Dpx * (0.5 / q.dx), dpdy = Dpy * (0.5 / q.dy);
  const long c = (long)b * d.cx + a;
  const double r = s.f[0][c], ru = s.f[1][c], rv = s.f[2][c], rw = s.f[3][c], rY = s.f[4][c], rX = s.f[5][c];
  const double rYovG = q.Ginv * rY;
  const double dbuoy = rY * (rX / r);
  const double drhou = ru - q.u0 * r, drhov = rv - q.v0 * r, drhow = rw - q.w0 * r;
  const double v = rv / r;
  s.f[1][c] = ru - q.dt * (rYovG * dpdx - q.ch2 * drhov + q.cv * drhow);
  s.f[2][c] = rv - q.dt * (rYovG * dpdy + q.gM * dbuoy * q.nonhydro - q.ch1 * drhow + q.ch2 * drhou) * q.ak;
  // rhow - dt * (...) * (ndim == 3) leaves rhow as it is in 2D
  s.f[5][c] = (r * (r / rY - S0c[b])) - q.dt * (v * dSdy[c]) * r;
}

// the compressible Exner increment at inner nodes: p += w (0 - dt dpidP div),
// dpidP = (gamma-1)/Msq * 2x2 mean of rhoY^(gamma-2) (second_projection.py:46-49, 65-68)
__global__ void k_efna_nodes(Dims d, double dt, double gm1_Msq, double gexp, double weight,
                             const double* __restrict__ rhoY, const double* __restrict__ div, double* p) {
  const int ia = blockIdx.x * blockDim.x + threadIdx.x;
  const int ib = blockIdx.y * blockDim.y + threadIdx.y;
  if (ia >= d.ix || ib >= d.iy) return;
  const int cx = d.cx;
  const long c = (long)ib * cx + ia;
  double m = pow(rhoY[c], gexp);
  m += pow(rhoY[c + cx], gexp);
  m += pow(rhoY[c + 1], gexp);
  m += pow(rhoY[c + cx + 1], gexp);
  const double dpidP = gm1_Msq * (m / 4.0);
  const double dp2 = 0.0 - dt * dpidP * div[(long)ib * d.ix + ia];
  const long pn = (long)(ib + 1) * d.nxn + ia + 1;
  p[pn] = p[pn] + weight * dp2;
}

// ---------------------------------------------------------------------------
// dynamic_timestep maxima: |u|, |v|, |w|, |u|+c, |v|+c, |w|+c over every cell,
// c = sqrt(gamma rhoY^gamma / rho) / sqrt(Msq)
// ---------------------------------------------------------------------------
constexpr int kCflThreads = 256;

__device__ __forceinline__ double max_nan(double a, double b) { return (b > a || b != b) ? b : a; }

__device__ __forceinline__ void block_max6(double (&v)[6], double* sh) {
#pragma unroll
  for (int k = 0; k < 6; ++k)
    for (int off = 32; off > 0; off >>= 1) v[k] = max_nan(v[k], __shfl_xor(v[k], off));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
    for (int k = 0; k < 6; ++k) sh[wave * 6 + k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kCflThreads / 64; ++w)
      for (int k = 0; k < 6; ++k) v[k] = max_nan(v[k], sh[w * 6 + k]);
  }
}

__global__ __launch_bounds__(kCflThreads) void k_cfl_partial(long n, double gamm, double Minv,
                                                             const double* __restrict__ rho,
                                                             const double* __restrict__ rhou,
                                                             const double* __restrict__ rhov,
                                                             const double* __restrict__ rhow,
                                                             const double* __restrict__ rhoY,
                                                             double* part) {
  __shared__ double sh[(kCflThreads / 64) * 6];
  double v[6];
  for (int k = 0; k < 6; ++k) v[k] = -1.0;
  for (long c = blockIdx.x * (long)kCflThreads + threadIdx.x; c < n; c += (long)gridDim.x * kCflThreads) {
    const double r = rho[c];
    const double pr = pow(rhoY[c], gamm);
    const double cs = sqrt(gamm * pr / r) * Minv;
    const double u = fabs(rhou[c] / r), vv = fabs(rhov[c] / r), w = fabs(rhow[c] / r);
    v[0] = max_nan(v[0], u);
    v[1] = max_nan(v[1], vv);
    v[2] = max_nan(v[2], w);
    v[3] = max_nan(v[3], u + cs);
    v[4] = max_nan(v[4], vv + cs);
    v[5] = max_nan(v[5], w + cs);
  }
  block_max6(v, sh);
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; ++k) part[(long)blockIdx.x * 6 + k] = v[k];
}

__global__ __launch_bounds__(kCflThreads) void k_cfl_final(int nparts, double* part, double* out) {
  __shared__ double sh[(kCflThreads / 64) * 6];
  double v[6];
  for (int k = 0; k < 6; ++k) v[k] = -1.0;
  for (int b = threadIdx.x; b < nparts; b += kCflThreads)
    for (int k = 0; k < 6; ++k) v[k] = max_nan(v[k], part[(long)b * 6 + k]);
  block_max6(v, sh);
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; ++k) out[k] = v[k];
}

int cfl_blocks(const pb_grid2d& g) {
  const long n = (long)(g.nx + 3) * (g.ny + 3);
  return (int)std::min<long>((n + kCflThreads - 1) / kCflThreads, 1024);
}

int check_grid(pb_ctx* ctx, const pb_grid2d* g, const char* what) {
  if (!g) return pb_fail(ctx, PB_ERR_ARG, "%s: null grid", what);
  if (g->nx < 3 || g->ny < 5) return pb_fail(ctx, PB_ERR_ARG, "%s: grid %d x %d too small", what, g->nx, g->ny);
  if (g->nx > (1 << 20) || g->ny > (1 << 20)) return pb_fail(ctx, PB_ERR_ARG, "%s: grid too large", what);
  return PB_OK;
}

dim3 grid2(int w, int h) { return dim3(ceil_div(w, 64), ceil_div(h, 4)); }
const dim3 kBlk2(64, 4);

}  // namespace

extern "C" int pb_advective_fluxes(pb_ctx* ctx, const pb_grid2d* grid, const double* rho, const double* rhou,
                                   const double* rhov, const double* rhoY, double* fx_rhoY, double* fy_rhoY) {
  if (!ctx) return PB_ERR_ARG;
  if (int rc = check_grid(ctx, grid, "pb_advective_fluxes")) return rc;
  if (!rho || !rhou || !rhov || !rhoY || !fx_rhoY || !fy_rhoY)
    return pb_fail(ctx, PB_ERR_ARG, "pb_advective_fluxes: null argument");
  const Dims d = dims_of(*grid);
  hipLaunchKernelGGL(k_adv_flux, grid2(d.cx + 1, d.cy + 1), kBlk2, 0, ctx->stream, d, rho, rhou, rhov, rhoY,
                     fx_rhoY, fy_rhoY);
  PB_LAUNCH_CHECK(ctx);
  return PB_OK;
}

extern "C" int pb_advect_split(pb_ctx* ctx, const pb_grid2d* grid, const pb_phys* phys, int axis, double lmbda,
                               int rk, double* const cells[6], double* const flux[6], const double* rhoY0_bg) {
  if (!ctx) return PB_ERR_ARG;
  if (int rc = check_grid(ctx, grid, "pb_advect_split")) return rc;
  if (!phys || !cells || !flux) return pb_fail(ctx, PB_ERR_ARG, "pb_advect_split: null argument");
  if (axis != 0 && axis != 1) return pb_fail(ctx, PB_ERR_ARG, "pb_advect_split: axis %d (0 or 1)", axis);
  for (int k = 0; k < 6; ++k)
    if (!cells[k] || !flux[k]) return pb_fail(ctx, PB_ERR_ARG, "pb_advect_split: field %d is null", k);
  const Dims d = dims_of(*grid);
  if (int rc = pb_bc_cells_axis(ctx, grid, phys, cells, rhoY0_bg, axis)) return rc;
  Fields s{{cells[0], cells[1], cells[2], cells[3], cells[4], cells[5]}};
  const double lam_rec = rk ? 0.0 : lmbda;   // recovery.py:43-44
  if (axis == 0)
    hipLaunchKernelGGL(k_hll<0>, grid2(d.cx + 1, d.cy), kBlk2, 0, ctx->stream, d, lam_rec, s,
                       (const double*)flux[4], flux[0], flux[1], flux[2], flux[3], flux[5]);
  else
    hipLaunchKernelGGL(k_hll<1>, grid2(d.cx, d.cy + 1), kBlk2, 0, ctx->stream, d, lam_rec, s,
                       (const double*)flux[4], flux[0], flux[1], flux[2], flux[3], flux[5]);
  PB_LAUNCH_CHECK(ctx);
  if (!rk) {
    FluxPtrs F{{flux[0], flux[1], flux[2], flux[3], flux[4], flux[5]}};
    if (axis == 0)
      hipLaunchKernelGGL(k_update<0>, grid2(d.cx, d.cy), kBlk2, 0, ctx->stream, d, lmbda, s, F);
    else
      hipLaunchKernelGGL(k_update<1>, grid2(d.cx, d.cy), kBlk2, 0, ctx->stream, d, lmbda, s, F);
    PB_LAUNCH_CHECK(ctx);
  }
  return pb_bc_cells_axis(ctx, grid, phys, cells, rhoY0_bg, axis);
}

extern "C" int pb_advect_rk_update(pb_ctx* ctx, const pb_grid2d* grid, const pb_phys* phys, double lx, double ly,
                                   double* const cells[6], double* const fx[6], double* const fy[6],
                                   const double* rhoY0_bg) {
  if (!ctx) return PB_ERR_ARG;
  if (int rc = check_grid(ctx, grid, "pb_advect_rk_update")) return rc;
  if (!phys || !cells || !fx || !fy) return pb_fail(ctx, PB_ERR_ARG, "pb_advect_rk_update: null argument");
  for (int k = 0; k < 6; ++k)
    if (!cells[k] || !fx[k] || !fy[k]) return pb_fail(ctx, PB_ERR_ARG, "pb_advect_rk_update: field %d is null", k);
  const Dims d = dims_of(*grid);
  Fields s{{cells[0], cells[1], cells[2], cells[3], cells[4], cells[5]}};
  FluxPtrs X{{fx[0], fx[1], fx[2], fx[3], fx[4], fx[5]}};
  FluxPtrs Y{{fy[0], fy[1], fy[2], fy[3], fy[4], fy[5]}};
  hipLaunchKernelGGL(k_update_rk, grid2(d.cx, d.cy), kBlk2, 0, ctx->stream, d, lx, ly, s, X, Y);
  PB_LAUNCH_CHECK(ctx);
  return pb_bc_cells(ctx, grid, phys, cells, rhoY0_bg);
}

extern "C" int pb_expl_part(pb_ctx* ctx, const pb_grid2d* grid, const pb_phys* phys, double dt,
                            double* const cells[6], const double* dSdy, const double* rhoY0_bg) {
  if (!ctx) return PB_ERR_ARG;
  if (int rc = check_grid(ctx, grid, "pb_expl_part")) return rc;
  if (!phys || !cells || !dSdy) return pb_fail(ctx, PB_ERR_ARG, "pb_expl_part: null argument");
  for (int k = 0; k < 6; ++k)
    if (!cells[k]) return pb_fail(ctx, PB_ERR_ARG, "pb_expl_part: field %d is null", k);
  const Dims d = dims_of(*grid);
  const double gM = phys->g / phys->Msq;
  ExplParams p{phys->nonhydro, dt * gM, -(dt * dt) * gM, dt * phys->coriolis[0], dt * phys->coriolis[1],
               dt * phys->coriolis[2], phys->wind[0], phys->wind[1], phys->wind[2]};
  Fields s{{cells[0], cells[1], cells[2], cells[3], cells[4], cells[5]}};
  hipLaunchKernelGGL(k_expl, grid2(d.cx, d.cy), kBlk2, 0, ctx->stream, d, p, s, dSdy);
  PB_LAUNCH_CHECK(ctx);
  return pb_bc_cells(ctx, grid, phys, cells, rhoY0_bg);
}

extern "C" int pb_euler_forward(pb_ctx* ctx, const pb_grid2d* grid, const pb_phys* phys, double dt,
                                double* const cells[6], double* p2_nodes, double* rhs, const double* S0c,
                                const double* dSdy, const double* rhoY0_bg) {
  if (!ctx) return PB_ERR_ARG;
  if (int rc = check_grid(ctx, grid, "pb_euler_forward")) return rc;
  if (!phys || !cells || !p2_nodes || !rhs || !S0c || !dSdy)
    return pb_fail(ctx, PB_ERR_ARG, "pb_euler_forward: null argument");
  for (int k = 0; k < 6; ++k)
    if (!cells[k]) return pb_fail(ctx, PB_ERR_ARG, "pb_euler_forward: field %d is null", k);
  const Dims d = dims_of(*grid);
  // divergence (momenta zeroed next to y walls first), not divided by dt
  if (int rc = pb_divergence_nodes(ctx, grid, 1.0, cells[0], cells[1], cells[2], cells[3], cells[4], rhs)) return rc;
  hipLaunchKernelGGL(k_scale_wall, grid2(d.ix, d.iy), kBlk2, 0, ctx->stream, d, grid->bc_x == PB_BC_WALL,
                     grid->bc_y == PB_BC_WALL, 2.0, rhs);
  const double gm1 = phys->gamm - 1.0;
  EfnaParams q{dt, grid->dx, grid->dy, phys->gamm / gm1, phys->g / phys->Msq, phys->nonhydro,
               phys->coriolis[0], phys->coriolis[1], phys->coriolis[2], phys->wind[0], phys->wind[1],
               phys->wind[2], (double)(1 - phys->is_arakawa_konor)};
  Fields s{{cells[0], cells[1], cells[2], cells[3], cells[4], cells[5]}};
  hipLaunchKernelGGL(k_efna_cells, grid2(d.cx, d.cy), kBlk2, 0, ctx->stream, d, q, s, (const double*)p2_nodes,
                     S0c, dSdy);
  hipLaunchKernelGGL(k_efna_nodes, grid2(d.ix, d.iy), kBlk2, 0, ctx->stream, d, dt, gm1 / phys->Msq,
                     phys->gamm - 2.0, phys->compressibility, (const double*)cells[4], (const double*)rhs,
                     p2_nodes);
  PB_LAUNCH_CHECK(ctx);
  if (int rc = pb_ghostnodes(ctx, grid, p2_nodes)) return rc;
  return pb_bc_cells(ctx, grid, phys, cells, rhoY0_bg);
}

extern "C" long pb_cfl_work_words(const pb_grid2d* grid) {
  if (!grid || grid->nx < 1 || grid->ny < 1) return -1;
  return 6L * cfl_blocks(*grid) + 6;
}

extern "C" int pb_cfl_maxima(pb_ctx* ctx, const pb_grid2d* grid, double gamm, double Minv, const double* rho,
                             const double* rhou, const double* rhov, const double* rhow, const double* rhoY,
                             double* work, double* maxima) {
  if (!ctx) return PB_ERR_ARG;
  if (int rc = check_grid(ctx, grid, "pb_cfl_maxima")) return rc;
  if (!rho || !rhou || !rhov || !rhow || !rhoY || !work || !maxima)
    return pb_fail(ctx, PB_ERR_ARG, "pb_cfl_maxima: null argument");
  const long n = (long)(grid->nx + 3) * (grid->ny + 3);
  const int nb = cfl_blocks(*grid);
  hipLaunchKernelGGL(k_cfl_partial, dim3(nb), dim3(kCflThreads), 0, ctx->stream, n, gamm, Minv, rho, rhou, rhov,
                     rhow, rhoY, work);
  hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(kCflThreads), 0, ctx->stream, nb, work, work + 6L * nb);
  PB_LAUNCH_CHECK(ctx);
  PB_HIP(ctx, hipMemcpyAsync(maxima, work + 6L * nb, 6 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  PB_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return PB_OK;
}
