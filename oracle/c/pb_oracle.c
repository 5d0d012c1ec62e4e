/*
 * C restatement of the reference's serial solve path — ORACLE / CPU BASELINE
 * ONLY (test infrastructure; the product never links this).
 *
 *   orc_lap2d      lap2D_gather_new, src/dycore/physics/low_mach/laplacian.py:546-746
 *                  (same per-node loop, index wrap, wall zeroing and operation
 *                  order; compiled without FP contraction)
 *   orc_bicgstab   BiCGStab of Barrett et al. (Templates 1994, sec. 2.3.8) with
 *                  the call semantics of second_projection.py:230 (x0 = 0,
 *                  legacy absolute early exit ||b|| <= tol, stop on the
 *                  recursive residual <= tol*||b||, eps^2 breakdown tests) —
 *                  the same algorithm as oracle/bicgstab.py.
 * Serial, like the reference's Numba loop and single-threaded BLAS-1.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct orc_op2d {
  int nx, ny;
  double dx, dy;
  int x_wall, y_wall;
  const double *hx, *hy, *hc, *cxx, *cyy, *cxy, *cyx, *dinv;  /* cells (ny+1)(nx+1), nodes ny*nx */
} orc_op2d;

void orc_lap2d(const orc_op2d* op, const double* p, double* lap) {
  const int nx = op->nx, ny = op->ny;
  const double oodx = 1.0 / op->dx, oody = 1.0 / op->dy;
  int cnt_x = 0, cnt_y = 0;
  for (int idx = 0; idx < nx * ny; ++idx) {
    const int row = idx / nx;
    const int col = idx - row * nx;
    const int ne = row * (nx + 1) + col;
    const int ne_tl = ne, ne_tr = ne + 1, ne_bl = ne + nx + 1, ne_br = ne + nx + 2;
    int tl = idx - nx - 1, ml = idx - 1, bl = idx + nx - 1;
    int tm = idx - nx, mm = idx, bm = idx + nx;
    int tr = idx - nx + 1, mr = idx + 1, br = idx + nx + 1;
    if (cnt_x == 0) { tl += nx - 1; ml += nx - 1; bl += nx - 1; }
    if (cnt_x == nx - 1) { tr -= nx - 1; mr -= nx - 1; br -= nx - 1; }
    if (cnt_y == 0) { tl += nx * (ny - 1); tm += nx * (ny - 1); tr += nx * (ny - 1); }
    if (cnt_y == ny - 1) { bl -= nx * (ny - 1); bm -= nx * (ny - 1); br -= nx * (ny - 1); }
    const double topleft = p[tl], midleft = p[ml], botleft = p[bl];
    const double topmid = p[tm], midmid = p[mm], botmid = p[bm];
    const double topright = p[tr], midright = p[mr], botright = p[br];
    double hx_tl = op->hx[ne_tl], hx_bl = op->hx[ne_bl], hy_tl = op->hy[ne_tl], hy_bl = op->hy[ne_bl];
    double hx_tr = op->hx[ne_tr], hx_br = op->hx[ne_br], hy_tr = op->hy[ne_tr], hy_br = op->hy[ne_br];
    if (op->x_wall && cnt_x == 0) { hx_tl = hy_tl = hx_bl = hy_bl = 0.0; }
    if (op->x_wall && cnt_x == nx - 1) { hx_tr = hy_tr = hx_br = hy_br = 0.0; }
    if (op->y_wall && cnt_y == 0) { hx_tl = hy_tl = hx_tr = hy_tr = 0.0; }
    if (op->y_wall && cnt_y == ny - 1) { hx_bl = hy_bl = hx_br = hy_br = 0.0; }
    const double Dx_tl = 0.5 * (topmid - topleft + midmid - midleft) * hx_tl;
    const double Dx_tr = 0.5 * (topright - topmid + midright - midmid) * hx_tr;
    const double Dx_bl = 0.5 * (botmid - botleft + midmid - midleft) * hx_bl;
    const double Dx_br = 0.5 * (botright - botmid + midright - midmid) * hx_br;
    const double Dy_tl = 0.5 * (midmid - topmid + midleft - topleft) * hy_tl;
    const double Dy_tr = 0.5 * (midright - topright + midmid - topmid) * hy_tr;
    const double Dy_bl = 0.5 * (botmid - midmid + botleft - midleft) * hy_bl;
    const double Dy_br = 0.5 * (botright - midright + botmid - midmid) * hy_br;
    const double *cxx = op->cxx, *cyy = op->cyy, *cxy = op->cxy, *cyx = op->cyx;
    const double Dxx = 0.5 * (cxx[ne_tr] * Dx_tr - cxx[ne_tl] * Dx_tl + cxx[ne_br] * Dx_br - cxx[ne_bl] * Dx_bl) * oodx * oodx;
    const double Dyy = 0.5 * (cyy[ne_br] * Dy_br - cyy[ne_tr] * Dy_tr + cyy[ne_bl] * Dy_bl - cyy[ne_tl] * Dy_tl) * oody * oody;
    const double Dyx = 0.5 * (cxy[ne_br] * Dy_br - cxy[ne_bl] * Dy_bl + cxy[ne_tr] * Dy_tr - cxy[ne_tl] * Dy_tl) * oody * oodx;
    const double Dxy = 0.5 * (cyx[ne_br] * Dx_br - cyx[ne_tr] * Dx_tr + cyx[ne_bl] * Dx_bl - cyx[ne_tl] * Dx_tl) * oodx * oody;
    double v = Dxx + Dyy + Dyx + Dxy + op->hc[idx] * p[idx];
    lap[idx] = v * op->dinv[idx];
    cnt_x += 1;
    if (cnt_x % nx == 0) { cnt_y += 1; cnt_x = 0; }
  }
}

static double dot(const double* a, const double* b, long n) {
  double s = 0.0;
  for (long i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* Returns info (0 converged, maxiter, -10/-11 breakdown); *niter = full
 * iterations; *half = 1 on the early ||s|| exit; x must hold n doubles. */
int orc_bicgstab(const orc_op2d* op, const double* b, double* x, double tol, int maxiter,
                 int* niter, int* half, double* resid) {
  const long n = (long)op->nx * op->ny;
  double* w = (double*)malloc(sizeof(double) * n * 6);
  if (!w) return -100;
  double *r = w, *rt = w + n, *p = w + 2 * n, *v = w + 3 * n, *s = w + 4 * n, *t = w + 5 * n;
  const double eps2 = 4.930380657631324e-32;
  memset(x, 0, sizeof(double) * n);
  memcpy(r, b, sizeof(double) * n);
  const double bnrm2 = sqrt(dot(b, b, n));
  *niter = 0;
  *half = 0;
  *resid = bnrm2;
  if (bnrm2 <= tol) { free(w); return 0; }
  const double atol = tol * bnrm2;
  memcpy(rt, r, sizeof(double) * n);
  double rho1 = 1.0, alpha = 0.0, omega = 0.0;
  int info = maxiter;
  for (int it = 1; it <= maxiter; ++it) {
    const double rho = dot(rt, r, n);
    if (fabs(rho) < eps2) { info = -10; *niter = it - 1; break; }
    if (it > 1) {
      const double beta = (rho / rho1) * (alpha / omega);
      for (long i = 0; i < n; ++i) {
        double q = p[i] - omega * v[i];
        q = beta * q;
        p[i] = q + r[i];
      }
    } else {
      memcpy(p, r, sizeof(double) * n);
    }
    orc_lap2d(op, p, v);
    const double rv = dot(rt, v, n);
    if (rv == 0.0) { info = -11; *niter = it - 1; break; }
    alpha = rho / rv;
    for (long i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
    const double snrm = sqrt(dot(s, s, n));
    if (snrm <= atol) {
      for (long i = 0; i < n; ++i) x[i] = x[i] + alpha * p[i];
      *half = 1; *niter = it - 1; *resid = snrm; info = 0;
      break;
    }
    orc_lap2d(op, s, t);
    omega = dot(t, s, n) / dot(t, t, n);
    for (long i = 0; i < n; ++i) {
      double xi = x[i] + alpha * p[i];
      x[i] = xi + omega * s[i];
      r[i] = s[i] - omega * t[i];
    }
    const double rnrm = sqrt(dot(r, r, n));
    *niter = it;
    *resid = rnrm;
    if (rnrm <= atol) { info = 0; break; }
    if (fabs(omega) < eps2) { info = -11; break; }
    rho1 = rho;
  }
  free(w);
  return info;
}
