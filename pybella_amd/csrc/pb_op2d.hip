// 2D nodal operator and BiCGStab on gfx950.
//
// Reference counterparts (pyBELLA tree):
//   k_assemble_*  operator_coefficients_nodes   second_projection.py:308-356
//                 scale_wall_node_values        second_projection.py:433-464
//                 regime scaling of wcenter     second_projection.py:141-153
//                 multiply_inverse_coriolis(get_coeffs=True) second_projection.py:604-652
//                 precon_diag_prepare           laplacian.py:1284-1356
//   k_stencil     lap2D_gather_new              laplacian.py:546-746
//   bicgstab_run  scipy.sparse.linalg.bicgstab  (call at second_projection.py:230)
//
// The 9-point operator in flux form.  For op-cell (j, i) (cell rows/cols
// 0..ny, 0..nx; its corners are nodes (j-1..j, i-1..i) with the reference's
// wrap onto the duplicated periodic end node, i-1 -> nx-2 at i=0 and
// i+1 -> 1 at i=nx-1, likewise in y):
//   gx = (p_tr - p_tl) + (p_br - p_bl),  gy = (p_bl - p_tl) + (p_br - p_tr)
//   Fx = Wxx gx + Wyx gy,  Fy = Wyy gy + Wxy gx
// with the folded weights
//   Wxx = hx cxx / (4 dx^2), Wyy = hy cyy / (4 dy^2),
//   Wyx = hy cxy / (4 dx dy), Wxy = hx cyx / (4 dx dy)
// (hx = wplus[0], hy = wplus[1] zeroed on the cells outside walls; the 1/4 is
// the product of the two 0.5 factors of laplacian.py:721-735).  A node sums
// its four cells:  (Fx_TR + Fx_BR - Fx_TL - Fx_BL) + (Fy_BL + Fy_BR - Fy_TL -
// Fy_TR) + hcenter p, times diag_inv — term by term the reference's
// Dxx + Dyy + Dyx + Dxy + hcenter p (laplacian.py:731-739).
#include <algorithm>
#include <cmath>
#include <vector>

#include "pb_common.hpp"

namespace {

constexpr int TX = 64;   // tile width = one wavefront
constexpr int TYT = 4;   // thread rows per block
constexpr int RY = 2;    // node rows per thread
constexpr int TY = TYT * RY;
constexpr int NTHR = TX * TYT;

template <typename T>
struct OpDev {
  const T* wxx;
  const T* wyy;
  const T* wyx;
  const T* wxy;
  const T* hc;
  const T* dinv;
  int nx, ny;
};

// Device state of one Krylov run.  Written only by the last block of a
// reduction kernel; read by every block at kernel start.
struct SolveState {
  double rho, rho1, alpha, omega, beta;
  double atol, bnrm2, resid, tol, rnorm0;
  double rr;          // generic reduction result (true residual runs)
  int iter, maxiter, status, half, legacy;
  unsigned counter;
};

enum { ST_RUN = 0, ST_CONV = 1, ST_HALF = 2, ST_MAXIT = 3, ST_BRK_RHO = -10, ST_BRK_OMEGA = -11 };

__device__ __forceinline__ double eps2() { return 4.930380657631324e-32; }  // DBL_EPSILON^2

// ---------------------------------------------------------------------------
// deterministic grid reduction: per-block partials [NV][nblocks], the last
// block to arrive sums them in block order and runs ``fin`` on thread 0.
// ---------------------------------------------------------------------------
template <int NV, class Fin>
__device__ __forceinline__ void grid_tail(double (&part)[NV], double* sred, double* partials,
                                          SolveState* st, Fin fin) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int nthr = blockDim.x * blockDim.y;
  const int nb = gridDim.x * gridDim.y;
  const int bid = blockIdx.x + blockIdx.y * gridDim.x;
  __shared__ int s_last;
  block_sum<NV>(part, sred);
  if (tid == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[(long)k * nb + bid] = part[k];
    __threadfence();
    const unsigned prev = atomicAdd(&st->counter, 1u);
    s_last = (prev == (unsigned)(nb - 1));
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  double tot[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
    for (int b = tid; b < nb; b += nthr)
      s += __hip_atomic_load(&partials[(long)k * nb + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    tot[k] = s;
  }
  __syncthreads();
  block_sum<NV>(tot, sred);
  if (tid == 0) {
    fin(tot);
    st->counter = 0u;
  }
}

// ---------------------------------------------------------------------------
// the stencil kernel: halo tile of the input expression in LDS, cell fluxes
// in LDS, node gather, then a per-node epilogue that may reduce.
// ---------------------------------------------------------------------------
template <typename T, bool CROSS, bool CENTER, class In, class Epi>
__global__ __launch_bounds__(NTHR) void k_stencil(OpDev<T> op, In in, Epi epi) {
  __shared__ T sp[TY + 2][TX + 2];
  __shared__ T sfx[TY + 1][TX + 1];
  __shared__ T sfy[TY + 1][TX + 1];
  __shared__ double sred[4 * (NTHR / kWave)];
  if (!epi.active()) return;
  in.prepare();
  const int tid = threadIdx.x + TX * threadIdx.y;
  const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY;
  const int nx = op.nx, ny = op.ny;
  for (int k = tid; k < (TY + 2) * (TX + 2); k += NTHR) {
    const int ly = k / (TX + 2), lx = k - ly * (TX + 2);
    int j = j0 - 1 + ly, i = i0 - 1 + lx;
    j = (j < 0) ? ny - 2 : (j == ny ? 1 : j);
    i = (i < 0) ? nx - 2 : (i == nx ? 1 : i);
    T val = T(0);
    if (j < ny && i < nx) val = in((long)j * nx + i);
    sp[ly][lx] = val;
  }
  __syncthreads();
  for (int k = tid; k < (TY + 1) * (TX + 1); k += NTHR) {
    const int cy = k / (TX + 1), cx = k - cy * (TX + 1);
    const int j = j0 + cy, i = i0 + cx;
    T fx = T(0), fy = T(0);
    if (j <= ny && i <= nx) {
      const long c = (long)j * (nx + 1) + i;
      const T tl = sp[cy][cx], tr = sp[cy][cx + 1], bl = sp[cy + 1][cx], br = sp[cy + 1][cx + 1];
      const T gx = (tr - tl) + (br - bl);
      const T gy = (bl - tl) + (br - tr);
      fx = op.wxx[c] * gx;
      fy = op.wyy[c] * gy;
      if (CROSS) {
        fx += op.wyx[c] * gy;
        fy += op.wxy[c] * gx;
      }
    }
    sfx[cy][cx] = fx;
    sfy[cy][cx] = fy;
  }
  __syncthreads();
  double part[Epi::NV > 0 ? Epi::NV : 1];
#pragma unroll
  for (int k = 0; k < (Epi::NV > 0 ? Epi::NV : 1); ++k) part[k] = 0.0;
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int ly = threadIdx.y + TYT * r, lx = threadIdx.x;
    const int j = j0 + ly, i = i0 + lx;
    if (j < ny && i < nx) {
      const long g = (long)j * nx + i;
      T acc = ((sfx[ly][lx + 1] + sfx[ly + 1][lx + 1]) - (sfx[ly][lx] + sfx[ly + 1][lx])) +
              ((sfy[ly + 1][lx] + sfy[ly + 1][lx + 1]) - (sfy[ly][lx] + sfy[ly][lx + 1]));
      const T pc = sp[ly + 1][lx + 1];
      if (CENTER) acc += op.hc[g] * pc;
      const T y = op.dinv[g] * acc;
      epi(g, y, pc, part);
    }
  }
  if constexpr (Epi::NV > 0) {
    grid_tail<Epi::NV>(part, sred, epi.partials, epi.st, epi);
  }
}

// ---- inputs ---------------------------------------------------------------
template <typename T>
struct InVec {
  const T* x;
  __device__ void prepare() {}
  __device__ T operator()(long g) const { return x[g]; }
};

template <typename T>
struct InZeroOr {   // x0 or zero
  const T* x;
  int given;
  __device__ void prepare() {}
  __device__ T operator()(long g) const { return given ? x[g] : T(0); }
};

// p_new = beta (p_old - omega v_old) + r
template <typename T>
struct InPUpdate {
  const T* r;
  const T* p;
  const T* v;
  const SolveState* st;
  T beta, omega;
  __device__ void prepare() {
    beta = (T)st->beta;
    omega = (T)st->omega;
  }
  __device__ T operator()(long g) const { return beta * (p[g] - omega * v[g]) + r[g]; }
};

// s = r - alpha v
template <typename T>
struct InS {
  const T* r;
  const T* v;
  const SolveState* st;
  T alpha;
  __device__ void prepare() { alpha = (T)st->alpha; }
  __device__ T operator()(long g) const { return r[g] - alpha * v[g]; }
};

// ---- epilogues ------------------------------------------------------------
template <typename T>
struct EpiStore {
  static constexpr int NV = 0;
  T* y;
  double* partials;
  SolveState* st;
  __device__ bool active() const { return true; }
  __device__ void operator()(long g, T v, T, double*) const { y[g] = v; }
  __device__ void operator()(double*) const {}
};

// initial residual r = b - A x0; rt = r; p = v = 0; (x = 0 unless given)
template <typename T>
struct EpiInit {
  static constexpr int NV = 2;
  const T* b;
  T* x;
  T* r;
  T* rt;
  T* p0;
  T* v0;
  int x_given;
  double* partials;
  SolveState* st;
  __device__ bool active() const { return true; }
  __device__ void operator()(long g, T y, T, double* part) const {
    const T bg = b[g];
    const T rg = bg - y;
    r[g] = rg;
    rt[g] = rg;
    p0[g] = T(0);
    v0[g] = T(0);
    if (!x_given) x[g] = T(0);
    part[0] += (double)rg * (double)rg;
    part[1] += (double)bg * (double)bg;
  }
  __device__ void operator()(double* tot) const {
    const double rr = tot[0], bb = tot[1];
    st->bnrm2 = sqrt(bb);
    st->rnorm0 = sqrt(rr);
    st->resid = st->rnorm0;
    st->iter = 0;
    st->half = 0;
    st->alpha = 0.0;
    st->omega = 0.0;
    st->beta = 0.0;
    st->rho1 = 1.0;
    st->rho = rr;
    st->atol = st->tol * st->bnrm2;
    if (st->legacy ? (st->rnorm0 <= st->tol) : (st->rnorm0 <= st->atol)) {
      st->status = ST_CONV;
    } else if (fabs(rr) < eps2()) {
      st->status = ST_BRK_RHO;
    } else {
      st->status = ST_RUN;
    }
  }
};

// v = A p_new (p_new written), partial rt . v -> alpha
template <typename T>
struct EpiV {
  static constexpr int NV = 1;
  T* p_new;
  T* v_new;
  const T* rt;
  double* partials;
  SolveState* st;
  __device__ bool active() const { return st->status == ST_RUN; }
  __device__ void operator()(long g, T y, T pc, double* part) const {
    p_new[g] = pc;
    v_new[g] = y;
    part[0] += (double)rt[g] * (double)y;
  }
  __device__ void operator()(double* tot) const {
    const double rv = tot[0];
    if (rv == 0.0) {
      st->status = ST_BRK_OMEGA;
    } else {
      st->alpha = st->rho / rv;
    }
  }
};

// t = A s, partials s.s, t.s, t.t -> early test, omega
template <typename T>
struct EpiT {
  static constexpr int NV = 3;
  T* t;
  double* partials;
  SolveState* st;
  __device__ bool active() const { return st->status == ST_RUN; }
  __device__ void operator()(long g, T y, T s, double* part) const {
    t[g] = y;
    part[0] += (double)s * (double)s;
    part[1] += (double)y * (double)s;
    part[2] += (double)y * (double)y;
  }
  __device__ void operator()(double* tot) const {
    const double snrm = sqrt(tot[0]);
    if (snrm <= st->atol) {
      st->status = ST_HALF;
      st->resid = snrm;
    } else {
      st->omega = tot[1] / tot[2];
    }
  }
};

// true residual r = b - A x (FP64), partial r.r -> st->rr
struct EpiTrueRes {
  static constexpr int NV = 1;
  const double* b;
  double* r;
  double* partials;
  SolveState* st;
  __device__ bool active() const { return true; }
  __device__ void operator()(long g, double y, double, double* part) const {
    const double rg = b[g] - y;
    r[g] = rg;
    part[0] += rg * rg;
  }
  __device__ void operator()(double* tot) const { st->rr = tot[0]; }
};

// ---------------------------------------------------------------------------
// elementwise end of an iteration: x += alpha p + omega s, r = s - omega t,
// partials rt.r, r.r -> convergence, rho, beta
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_bicg_update(long n, T* x, T* r, const T* rt, const T* p,
                                                     const T* v, const T* t, double* partials,
                                                     SolveState* st) {
  __shared__ double sred[2 * (256 / kWave)];
  const int status = st->status;
  if (status != ST_RUN && status != ST_HALF) return;
  const bool half = status == ST_HALF;
  const T alpha = (T)st->alpha, omega = (T)st->omega;
  double part[2] = {0.0, 0.0};
  const long stride = (long)gridDim.x * blockDim.x;
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += stride) {
    if (half) {
      x[g] = x[g] + alpha * p[g];
    } else {
      const T s = r[g] - alpha * v[g];
      T xg = x[g] + alpha * p[g];
      x[g] = xg + omega * s;
      const T rn = s - omega * t[g];
      r[g] = rn;
      part[0] += (double)rt[g] * (double)rn;
      part[1] += (double)rn * (double)rn;
    }
  }
  struct Fin {
    SolveState* st;
    bool half;
    __device__ void operator()(double* tot) const {
      if (half) {
        st->status = ST_CONV;
        st->half = 1;
        return;
      }
      const double rnrm = sqrt(tot[1]);
      st->iter += 1;
      st->resid = rnrm;
      if (rnrm <= st->atol) {
        st->status = ST_CONV;
      } else if (fabs(st->omega) < eps2()) {
        st->status = ST_BRK_OMEGA;
      } else if (st->iter >= st->maxiter) {
        st->status = ST_MAXIT;
      } else {
        st->rho1 = st->rho;
        st->rho = tot[0];
        if (fabs(st->rho) < eps2()) {
          st->status = ST_BRK_RHO;
        } else {
          st->beta = (st->rho / st->rho1) * (st->alpha / st->omega);
        }
      }
    }
  };
  grid_tail<2>(part, sred, partials, st, Fin{st, half});
}

template <typename Tin, typename Tout>
__global__ void k_cast(long n, const Tin* a, Tout* b) {
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x)
    b[g] = (Tout)a[g];
}

template <typename T>
__global__ void k_fill(long n, T* a, T v) {
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x)
    a[g] = v;
}

__global__ void k_accum(long n, const float* dx, double* x) {
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x)
    x[g] += (double)dx[g];
}

__global__ void k_scale_rhs(int nx, int ny, const double* rhs_isc, const double* dinv, double* b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nx || j >= ny) return;
  b[(long)j * nx + i] = rhs_isc[(long)(j + 1) * (nx + 2) + (i + 1)] * dinv[(long)j * nx + i];
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

template <typename T>
struct Work {
  T* x = nullptr;
  T* r = nullptr;
  T* rt = nullptr;
  T* p[2] = {nullptr, nullptr};
  T* v[2] = {nullptr, nullptr};
  T* t = nullptr;
  T* b = nullptr;   // FP32 inner right-hand side
};

}  // namespace

struct pb_op2d {
  pb_ctx* ctx = nullptr;
  pb_grid2d g{};
  long n = 0;      // nodes
  long nc = 0;     // op-cells
  bool cross = false, center = true, assembled = false;
  double* wxx = nullptr;
  double* wyy = nullptr;
  double* wyx = nullptr;
  double* wxy = nullptr;
  double* hc = nullptr;
  double* dinv = nullptr;
  float* wxx32 = nullptr;
  float* wyy32 = nullptr;
  float* wyx32 = nullptr;
  float* wxy32 = nullptr;
  float* hc32 = nullptr;
  float* dinv32 = nullptr;
  double* coeff = nullptr;   // cells scratch
  double* rYpow = nullptr;   // cells scratch
  // solver workspace
  bool have_work = false;
  Work<double> w64;
  Work<float> w32;
  double* rtrue = nullptr;
  double* partials = nullptr;
  SolveState* st = nullptr;      // device, [3]
  SolveState* st_host = nullptr; // pinned, [3]
  std::vector<void*> allocs;
};

namespace {

template <typename T>
int dmalloc(pb_op2d* op, T** p, long count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, sizeof(T) * (size_t)std::max<long>(count, 1));
  if (e != hipSuccess)
    return pb_fail(op->ctx, PB_ERR_HIP, "hipMalloc(%ld bytes) failed: %s", (long)(sizeof(T) * count),
                   hipGetErrorString(e));
  op->allocs.push_back(q);
  *p = (T*)q;
  return PB_OK;
}

template <typename T>
OpDev<T> opdev(const pb_op2d* op);
template <>
OpDev<double> opdev<double>(const pb_op2d* op) {
  return OpDev<double>{op->wxx, op->wyy, op->wyx, op->wxy, op->hc, op->dinv, op->g.nx, op->g.ny};
}
template <>
OpDev<float> opdev<float>(const pb_op2d* op) {
  return OpDev<float>{op->wxx32, op->wyy32, op->wyx32, op->wxy32, op->hc32, op->dinv32, op->g.nx, op->g.ny};
}

inline dim3 stencil_grid(const pb_op2d* op) { return dim3(ceil_div(op->g.nx, TX), ceil_div(op->g.ny, TY)); }
inline int elem_blocks(long n) { return (int)std::min<long>((n + 1023) / 1024, 2048); }

template <typename T, class In, class Epi>
void launch_stencil(const pb_op2d* op, In in, Epi epi) {
  const dim3 grd = stencil_grid(op), blk(TX, TYT);
  hipStream_t s = op->ctx->stream;
  const OpDev<T> d = opdev<T>(op);
  if (op->cross) {
    if (op->center)
      hipLaunchKernelGGL((k_stencil<T, true, true, In, Epi>), grd, blk, 0, s, d, in, epi);
    else
      hipLaunchKernelGGL((k_stencil<T, true, false, In, Epi>), grd, blk, 0, s, d, in, epi);
  } else {
    if (op->center)
      hipLaunchKernelGGL((k_stencil<T, false, true, In, Epi>), grd, blk, 0, s, d, in, epi);
    else
      hipLaunchKernelGGL((k_stencil<T, false, false, In, Epi>), grd, blk, 0, s, d, in, epi);
  }
}

int alloc_work(pb_op2d* op) {
  if (op->have_work) return PB_OK;
  const long n = op->n;
  int rc = 0;
  Work<double>& a = op->w64;
  Work<float>& f = op->w32;
  rc |= dmalloc(op, &a.r, n);
  rc |= dmalloc(op, &a.rt, n);
  rc |= dmalloc(op, &a.p[0], n);
  rc |= dmalloc(op, &a.p[1], n);
  rc |= dmalloc(op, &a.v[0], n);
  rc |= dmalloc(op, &a.v[1], n);
  rc |= dmalloc(op, &a.t, n);
  rc |= dmalloc(op, &f.x, n);
  rc |= dmalloc(op, &f.r, n);
  rc |= dmalloc(op, &f.rt, n);
  rc |= dmalloc(op, &f.p[0], n);
  rc |= dmalloc(op, &f.p[1], n);
  rc |= dmalloc(op, &f.v[0], n);
  rc |= dmalloc(op, &f.v[1], n);
  rc |= dmalloc(op, &f.t, n);
  rc |= dmalloc(op, &f.b, n);
  rc |= dmalloc(op, &op->rtrue, n);
  const dim3 sg = stencil_grid(op);
  const long nbmax = std::max<long>((long)sg.x * sg.y, elem_blocks(n));
  rc |= dmalloc(op, &op->partials, 4 * nbmax);
  rc |= dmalloc(op, &op->st, 3);
  if (rc) return PB_ERR_HIP;
  void* h = nullptr;
  if (hipHostMalloc(&h, sizeof(SolveState) * 3, hipHostMallocDefault) != hipSuccess)
    return pb_fail(op->ctx, PB_ERR_HIP, "hipHostMalloc failed");
  op->st_host = (SolveState*)h;
  PB_HIP(op->ctx, hipMemsetAsync(op->st, 0, sizeof(SolveState) * 3, op->ctx->stream));
  op->have_work = true;
  return PB_OK;
}

// One Krylov run on device vectors (b, x) in precision T with state slot ``slot``.
template <typename T>
int bicgstab_run(pb_op2d* op, Work<T>& w, const T* b, T* x, int x_given, double tol, int maxiter,
                 int legacy, int check_every, int slot, SolveState* out) {
  pb_ctx* ctx = op->ctx;
  hipStream_t s = ctx->stream;
  SolveState* st = op->st + slot;
  SolveState init{};
  init.tol = tol;
  init.maxiter = maxiter;
  init.legacy = legacy;
  init.status = ST_RUN;
  PB_HIP(ctx, hipMemcpyAsync(st, &init, sizeof(SolveState), hipMemcpyHostToDevice, s));
  // initial residual
  launch_stencil<T>(op, InZeroOr<T>{x, x_given},
                    EpiInit<T>{b, x, w.r, w.rt, w.p[0], w.v[0], x_given, op->partials, st});
  PB_LAUNCH_CHECK(ctx);
  const int nb = elem_blocks(op->n);
  int k = 1;
  const int chunk = check_every > 0 ? check_every : 32;
  for (;;) {
    PB_HIP(ctx, hipMemcpyAsync(op->st_host + slot, st, sizeof(SolveState), hipMemcpyDeviceToHost, s));
    PB_HIP(ctx, hipStreamSynchronize(s));
    const SolveState h = op->st_host[slot];
    if (h.status != ST_RUN) {
      *out = h;
      break;
    }
    if (k > maxiter)
      return pb_fail(ctx, PB_ERR_STATE, "bicgstab: %d iterations issued but the run did not finish", maxiter);
    for (int c = 0; c < chunk && k <= maxiter; ++c, ++k) {
      const int po = (k - 1) & 1, pn = k & 1;
      launch_stencil<T>(op, InPUpdate<T>{w.r, w.p[po], w.v[po], st, T(0), T(0)},
                        EpiV<T>{w.p[pn], w.v[pn], w.rt, op->partials, st});
      launch_stencil<T>(op, InS<T>{w.r, w.v[pn], st, T(0)}, EpiT<T>{w.t, op->partials, st});
      hipLaunchKernelGGL(k_bicg_update<T>, dim3(nb), dim3(256), 0, s, op->n, x, w.r, w.rt, w.p[pn],
                         w.v[pn], w.t, op->partials, st);
    }
    PB_LAUNCH_CHECK(ctx);
  }
  return PB_OK;
}

// FP64 true residual r = b - A x; returns ||r||
int true_residual(pb_op2d* op, const double* b, const double* x, double* nrm) {
  SolveState* st = op->st + 2;
  launch_stencil<double>(op, InVec<double>{x}, EpiTrueRes{b, op->rtrue, op->partials, st});
  PB_LAUNCH_CHECK(op->ctx);
  PB_HIP(op->ctx, hipMemcpyAsync(op->st_host + 2, st, sizeof(SolveState), hipMemcpyDeviceToHost,
                                 op->ctx->stream));
  PB_HIP(op->ctx, hipStreamSynchronize(op->ctx->stream));
  *nrm = sqrt(op->st_host[2].rr);
  return PB_OK;
}

int info_of(const SolveState& h) {
  switch (h.status) {
    case ST_CONV: return 0;
    case ST_MAXIT: return h.iter;
    case ST_BRK_RHO: return -10;
    case ST_BRK_OMEGA: return -11;
    default: return h.iter > 0 ? h.iter : 1;
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

extern "C" int pb_op2d_destroy(pb_op2d* op) {
  if (!op) return PB_OK;
  if (op->ctx && op->ctx->stream) (void)hipStreamSynchronize(op->ctx->stream);
  for (void* p : op->allocs) (void)hipFree(p);
  if (op->st_host) (void)hipHostFree(op->st_host);
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
  return PB_OK;
}

extern "C" int pb_op2d_matvec(pb_op2d* op, const void* x, void* y, int precision) {
  if (!op) return PB_ERR_ARG;
  if (!x || !y) return pb_fail(op->ctx, PB_ERR_ARG, "pb_op2d_matvec: null vector");
  if (!op->assembled) return pb_fail(op->ctx, PB_ERR_STATE, "pb_op2d_matvec: operator not assembled");
  if (precision == PB_F64) {
    launch_stencil<double>(op, InVec<double>{(const double*)x}, EpiStore<double>{(double*)y, nullptr, nullptr});
  } else if (precision == PB_F32) {
    launch_stencil<float>(op, InVec<float>{(const float*)x}, EpiStore<float>{(float*)y, nullptr, nullptr});
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

extern "C" int pb_bicgstab_solve(pb_op2d* op, const double* b, double* x, const pb_solve_opts* opts,
                                 pb_solve_info* info) {
  if (!op) return PB_ERR_ARG;
  pb_ctx* ctx = op->ctx;
  if (!b || !x || !opts || !info) return pb_fail(ctx, PB_ERR_ARG, "pb_bicgstab_solve: null argument");
  if (!op->assembled) return pb_fail(ctx, PB_ERR_STATE, "pb_bicgstab_solve: operator not assembled");
  if (!(opts->tol >= 0) || opts->maxiter < 1)
    return pb_fail(ctx, PB_ERR_ARG, "pb_bicgstab_solve: bad tol/maxiter");
  int rc = alloc_work(op);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  pb_solve_info out{};
  if (opts->mode == PB_SOLVE_F64) {
    Work<double>& w = op->w64;
    SolveState h{};
    rc = bicgstab_run<double>(op, w, b, x, opts->x0_given, opts->tol, opts->maxiter,
                              opts->x0_given ? 0 : 1, opts->check_every, 0, &h);
    if (rc) return rc;
    out.info = info_of(h);
    out.iters = h.iter;
    out.outer = 1;
    out.half_step = h.half;
    out.resid = h.resid;
    out.bnrm2 = h.bnrm2;
  } else if (opts->mode == PB_SOLVE_MIXED) {
    const long n = op->n;
    const int nb = elem_blocks(n);
    const double inner = opts->inner_tol > 0 ? opts->inner_tol : 1e-5;
    const int max_outer = opts->max_outer > 0 ? opts->max_outer : 10;
    double bn = 0.0, rn = 0.0;
    if (!opts->x0_given) hipLaunchKernelGGL(k_fill<double>, dim3(nb), dim3(256), 0, s, n, x, 0.0);
    hipLaunchKernelGGL(k_fill<double>, dim3(nb), dim3(256), 0, s, n, op->w64.t, 0.0);
    rc = true_residual(op, b, op->w64.t, &bn);   // ||b||
    if (rc) return rc;
    rc = true_residual(op, b, x, &rn);            // rtrue = b - A x
    if (rc) return rc;
    out.bnrm2 = bn;
    int outer = 0, iters = 0, last_info = 0;
    SolveState h{};
    const double target = opts->tol * bn;
    for (outer = 0; outer < max_outer && rn > target; ++outer) {
      hipLaunchKernelGGL((k_cast<double, float>), dim3(nb), dim3(256), 0, s, n, (const double*)op->rtrue,
                         op->w32.b);
      rc = bicgstab_run<float>(op, op->w32, op->w32.b, op->w32.x, 0, inner, opts->maxiter, 0,
                               opts->check_every, 1, &h);
      if (rc) return rc;
      iters += h.iter;
      last_info = info_of(h);
      hipLaunchKernelGGL(k_accum, dim3(nb), dim3(256), 0, s, n, (const float*)op->w32.x, x);
      rc = true_residual(op, b, x, &rn);
      if (rc) return rc;
      if (h.status < 0 && h.iter == 0) break;   // breakdown before any progress
    }
    out.info = rn <= target ? 0 : (last_info < 0 ? last_info : std::max(iters, 1));
    out.iters = iters;
    out.outer = outer;
    out.half_step = h.half;
    out.resid = rn;
  } else {
    return pb_fail(ctx, PB_ERR_ARG, "pb_bicgstab_solve: mode %d not available on this handle", opts->mode);
  }
  double tr = 0.0;
  rc = true_residual(op, b, x, &tr);
  if (rc) return rc;
  out.true_rel_resid = out.bnrm2 > 0 ? tr / out.bnrm2 : tr;
  *info = out;
  return out.info;
}
