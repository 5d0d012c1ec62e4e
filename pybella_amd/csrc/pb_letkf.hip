// LETKF local analysis with grid-point localisation (gfx950).
//
// For every grid point of one grid type (cells or nodes), the analysis of
// Hunt et al. (2007) over the obs_x x obs_y window around it, as the
// reference's r-localised LETKF computes it (src/data_assimilation/letkf.py:
// analysis.analyse :120-166 on the windows of prepare_rloc :288-426):
//   Y   the members' values at the window's observed locations minus their
//       ensemble mean (the window wraps periodically over the inner grid);
//   w   localisation weight / observation variance per observation (the
//       weight is zero outside a wall);
//   Pa  = (k-1)/rho I + Y^T diag(w) Y      (k x k, in the ensemble space)
//   Pa  = V diag(lambda) V^T               (cyclic Jacobi, in LDS)
//   Wa  = sqrt(k-1) V diag(lambda^-1/2) V^T + 1 wa^T,
//   wa  = Pa^-1 Y^T diag(w) (y_obs - y_mean)
//   x_a^(i) = x_mean + sum_j Wa[i][j] (x^(j) - x_mean)  at the grid point.
//
// One 64-lane wave per grid point.  The window is streamed in chunks of 64
// observation slots (one per lane) through LDS; masked slots carry weight 0.
// Layout: state and analysis (members, attrs, ny, nx), obs and mask
// (attrs, ny, nx), x fastest; loc (obs_x, obs_y) row-major.
#include "pb_common.hpp"

namespace {

constexpr int kLw = 64;   // lanes per grid point

struct LetkfArgs {
  int k, attrs, nx, ny, ox, oy, px, py, wall_x, wall_y;
  double rho;
};

template <int KM>
__global__ __launch_bounds__(kLw) void k_letkf(LetkfArgs a, const double* __restrict__ state,
                                               const double* __restrict__ obs, const double* __restrict__ mask,
                                               const double* __restrict__ covar, const double* __restrict__ loc,
                                               double* __restrict__ out) {
  __shared__ double Ys[kLw][KM + 1];   // a chunk of observation anomalies [slot][member]
  __shared__ double ws[kLw], inn[kLw];
  __shared__ double A[KM][KM + 1], V[KM][KM + 1];
  __shared__ double dv[KM], zv[KM], wav[KM], ls[KM];
  const int lane = threadIdx.x;
  const int pt = blockIdx.x;
  const int ix = pt % a.nx, iy = pt / a.nx;
  const int k = a.k;
  const long plane = (long)a.nx * a.ny;
  constexpr int EPL = (KM * KM + kLw - 1) / kLw;   // Pa entries per lane
  double acc[EPL];
#pragma unroll
  for (int e = 0; e < EPL; ++e) acc[e] = 0.0;
  double dacc = 0.0;
  const int nslot = a.attrs * a.ox * a.oy;
  for (int base = 0; base < nslot; base += kLw) {
    const int s = base + lane;
    double w = 0.0, iv = 0.0;
    if (s < nslot) {
      const int at = s / (a.ox * a.oy), r = s % (a.ox * a.oy), wx = r / a.oy, wy = r % a.oy;
      const int ux = ix + wx - a.px, uy = iy + wy - a.py;
      const int gx = ((ux % a.nx) + a.nx) % a.nx, gy = ((uy % a.ny) + a.ny) % a.ny;
      const long g = (long)gy * a.nx + gx;
      if (mask[at * plane + g] == 0.0) {
        double sum = 0.0;
        for (int m = 0; m < k; ++m) sum += state[((long)m * a.attrs + at) * plane + g];
        const double ym = sum / k;
        for (int m = 0; m < k; ++m) Ys[lane][m] = state[((long)m * a.attrs + at) * plane + g] - ym;
        const bool out_x = a.wall_x && (ux < 0 || ux >= a.nx), out_y = a.wall_y && (uy < 0 || uy >= a.ny);
        const double L = (out_x || out_y) ? 0.0 : loc[wx * a.oy + wy];
        w = L / (covar ? covar[at] : 1.0);
        iv = obs[at * plane + g] - ym;
      } else {
        for (int m = 0; m < k; ++m) Ys[lane][m] = 0.0;
      }
    } else {
      for (int m = 0; m < k; ++m) Ys[lane][m] = 0.0;
    }
    ws[lane] = w;
    inn[lane] = iv;
    __syncthreads();
    const int cnt = min(kLw, nslot - base);
#pragma unroll
    for (int e = 0; e < EPL; ++e) {
      const int idx = lane + e * kLw;
      if (idx < k * k) {
        // (lo, hi) order: the (p, q) and (q, p) sums are the same
        // operations, so Pa is exactly symmetric
        const int p = min(idx / k, idx % k), q = max(idx / k, idx % k);
        double sacc = acc[e];
        for (int j = 0; j < cnt; ++j) sacc += ws[j] * Ys[j][p] * Ys[j][q];
        acc[e] = sacc;
      }
    }
    if (lane < k)
      for (int j = 0; j < cnt; ++j) dacc += ws[j] * Ys[j][lane] * inn[j];
    __syncthreads();
  }
  // Pa and V = I
  const double diag = (k - 1.0) / a.rho;
#pragma unroll
  for (int e = 0; e < EPL; ++e) {
    const int idx = lane + e * kLw;
    if (idx < k * k) {
      const int p = idx / k, q = idx % k;
      A[p][q] = (p == q ? diag : 0.0) + acc[e];
      V[p][q] = p == q ? 1.0 : 0.0;
    }
  }
  if (lane < k) dv[lane] = dacc;
  __syncthreads();
  // cyclic Jacobi sweeps (Golub & Van Loan, Alg. 8.5.2 with the symmetric
  // Schur rotation) until the off-diagonal part is negligible: V collects
  // the rotations, the diagonal of A becomes the eigenvalues
  for (int sweep = 0; sweep < 40; ++sweep) {
    double off = 0.0, dg = 0.0;
    if (lane < k)
      for (int q = 0; q < k; ++q) {
        const double v = A[lane][q];
        if (q == lane) dg += v * v; else off += v * v;
      }
    for (int o = 32; o > 0; o >>= 1) {
      off += __shfl_xor(off, o);
      dg += __shfl_xor(dg, o);
    }
    if (!(off > 1e-30 * dg)) break;
    for (int p = 0; p < k - 1; ++p)
      for (int q = p + 1; q < k; ++q) {
        const double apq = A[p][q];
        if (apq == 0.0) continue;
        const double app = A[p][p], aqq = A[q][q];
        const double tau = (aqq - app) / (2.0 * apq);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double c = 1.0 / sqrt(1.0 + t * t), sn = t * c;
        double arp = 0.0, arq = 0.0, vrp = 0.0, vrq = 0.0;
        if (lane < k) {
          arp = A[lane][p];
          arq = A[lane][q];
          vrp = V[lane][p];
          vrq = V[lane][q];
        }
        __syncthreads();
        if (lane < k) {
          if (lane == p) {
            A[p][p] = app - t * apq;
            A[p][q] = 0.0;
            A[q][p] = 0.0;
          } else if (lane == q) {
            A[q][q] = aqq + t * apq;
          } else {
            const double np_ = c * arp - sn * arq, nq_ = sn * arp + c * arq;
            A[lane][p] = np_;
            A[p][lane] = np_;
            A[lane][q] = nq_;
            A[q][lane] = nq_;
          }
          V[lane][p] = c * vrp - sn * vrq;
          V[lane][q] = sn * vrp + c * vrq;
        }
        __syncthreads();
      }
  }
  // z_e = (V^T d)_e / lambda_e, wa = V z; ls = sqrt(1 / lambda)
  if (lane < k) {
    double z = 0.0;
    for (int i = 0; i < k; ++i) z += V[i][lane] * dv[i];
    zv[lane] = z * (1.0 / A[lane][lane]);
    ls[lane] = sqrt(1.0 / A[lane][lane]);
  }
  __syncthreads();
  if (lane < k) {
    double w = 0.0;
    for (int e = 0; e < k; ++e) w += V[lane][e] * zv[e];
    wav[lane] = w;
  }
  __syncthreads();
  // Wa = sqrt(k-1) V diag(sqrt(1/lambda)) V^T + wa, into the A storage
  const double sk = sqrt(k - 1.0);
#pragma unroll
  for (int e = 0; e < EPL; ++e) {
    const int idx = lane + e * kLw;
    if (idx < k * k) {
      const int i = idx / k, j = idx % k;
      double s = 0.0;
      for (int f = 0; f < k; ++f) s += V[i][f] * ls[f] * V[j][f];
      acc[e] = sk * s + wav[j];
    }
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < EPL; ++e) {
    const int idx = lane + e * kLw;
    if (idx < k * k) A[idx / k][idx % k] = acc[e];
  }
  __syncthreads();
  // analysis at this grid point: member i, attribute at
  const long g = (long)iy * a.nx + ix;
  for (int idx = lane; idx < k * a.attrs; idx += kLw) {
    const int i = idx / a.attrs, at = idx % a.attrs;
    double sum = 0.0;
    for (int m = 0; m < k; ++m) sum += state[((long)m * a.attrs + at) * plane + g];
    const double xm = sum / k;
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += A[i][j] * (state[((long)j * a.attrs + at) * plane + g] - xm);
    out[((long)i * a.attrs + at) * plane + g] = s + xm;
  }
}

}  // namespace

extern "C" int pb_letkf_analyse(pb_ctx* ctx, const pb_letkf_cfg* cfg, const double* state, const double* obs,
                                const double* mask, const double* covar, const double* loc, double* analysis) {
  if (!ctx || !cfg || !state || !obs || !mask || !loc || !analysis)
    return pb_fail(ctx, PB_ERR_ARG, "pb_letkf_analyse: null argument");
  const int k = cfg->members;
  if (k < 2 || k > 32) return pb_fail(ctx, PB_ERR_ARG, "pb_letkf_analyse: members %d (2..32)", k);
  if (cfg->attrs < 1 || cfg->nx < 1 || cfg->ny < 1)
    return pb_fail(ctx, PB_ERR_ARG, "pb_letkf_analyse: empty grid or no attribute");
  if (cfg->obs_x < 1 || cfg->obs_y < 1 || cfg->obs_x % 2 == 0 || cfg->obs_y % 2 == 0 || cfg->obs_x >= cfg->nx ||
      cfg->obs_y >= cfg->ny)
    return pb_fail(ctx, PB_ERR_ARG, "pb_letkf_analyse: window %d x %d must be odd and smaller than the grid",
                   cfg->obs_x, cfg->obs_y);
  if (!(cfg->inflation > 0.0)) return pb_fail(ctx, PB_ERR_ARG, "pb_letkf_analyse: inflation must be > 0");
  if ((long)cfg->nx * cfg->ny > (1L << 31) - 1) return pb_fail(ctx, PB_ERR_ARG, "pb_letkf_analyse: grid too large");
  LetkfArgs a{k, cfg->attrs, cfg->nx, cfg->ny, cfg->obs_x, cfg->obs_y, (cfg->obs_x - 1) / 2, (cfg->obs_y - 1) / 2,
              !cfg->periodic_x, !cfg->periodic_y, cfg->inflation};
  const dim3 grid((unsigned)(cfg->nx * cfg->ny));
  if (k <= 8)
    hipLaunchKernelGGL(k_letkf<8>, grid, dim3(kLw), 0, ctx->stream, a, state, obs, mask, covar, loc, analysis);
  else if (k <= 16)
    hipLaunchKernelGGL(k_letkf<16>, grid, dim3(kLw), 0, ctx->stream, a, state, obs, mask, covar, loc, analysis);
  else
    hipLaunchKernelGGL(k_letkf<32>, grid, dim3(kLw), 0, ctx->stream, a, state, obs, mask, covar, loc, analysis);
  PB_LAUNCH_CHECK(ctx);
  return PB_OK;
}
