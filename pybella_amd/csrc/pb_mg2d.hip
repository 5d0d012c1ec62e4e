// Two-branch multigrid cycle of the 2D nodal operator (gfx950); the design is
// described in pb_mg2d.hpp.
//
// Level 0 runs on the Krylov layout with the operator's own FP32 flux-form
// weights (the LDS-tiled stencil of pb_op2d.hpp); levels >= 1 are FP32
// 9-point stencils on unique grids, one thread per node carrying both
// branches.  The Galerkin stencils are formed in FP64 on the device whenever
// the operator is re-assembled.
#include <algorithm>

#include "pb_mg2d.hpp"

using namespace pb2d;

int pb_nn_forward_level(pb_nn* nn, int nimg, int uny, int unx, int nh, int nw, const float* v, const float* f1,
                        const float* f2, const float** zout, const float** scale, hipStream_t s);

namespace pbmg {
namespace {

// index of node i + d along an axis of n nodes: wrapped (periodic) or
// clamped into range with ok = false (wall; the stencil entry is zero there)
__device__ __forceinline__ int axis_idx(int i, int n, int periodic, bool& ok) {
  if (periodic) {
    ok = true;
    return i < 0 ? i + n : (i >= n ? i - n : i);
  }
  ok = i >= 0 && i < n;
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

// the 3x3 neighbourhood of (j, i), entry k = 3 (dj + 1) + (di + 1)
__device__ __forceinline__ void nbr9(const Dims& d, int j, int i, int (&idx)[9]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    bool oy;
    const int jj = axis_idx(j + a - 1, d.ny, d.py, oy);
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      bool ox;
      const int ii = axis_idx(i + b - 1, d.nx, d.px, ox);
      idx[3 * a + b] = jj * d.nx + ii;
    }
  }
}

// ---------------------------------------------------------------------------
// set-up
// ---------------------------------------------------------------------------

// Level-0 stencil of D^-1 A on the unique grid from the folded flux weights
// (pb_op2d.hpp): per op-cell gx = (tr - tl) + (br - bl), gy = (bl - tl) +
// (br - tr), Fx = Wxx gx + Wyx gy, Fy = Wyy gy + Wxy gx; a node sums its four
// cells' sx Fx + sy Fy (sx = -1 left cells, sy = -1 upper cells), adds hc p
// and is scaled by dinv: lap2D_gather_new (laplacian.py:546-746) term by term.
__global__ void k_expand9(int nxd, Dims u, const double* wxx, const double* wyy, const double* wyx,
                          const double* wxy, const double* hc, const double* dinv, int cross, int center,
                          double* st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= u.nx || j >= u.ny) return;
  double c[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) c[k] = 0.0;
  const int ncx = nxd + 1;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int dcj = q >> 1, dci = q & 1;
    const double sx = dci ? 1.0 : -1.0, sy = dcj ? 1.0 : -1.0;
    const long cc = (long)(j + dcj) * ncx + (i + dci);
    const double a = wxx[cc], b = wyy[cc];
    const double cyx = cross ? wyx[cc] : 0.0, cxy = cross ? wxy[cc] : 0.0;
#pragma unroll
    for (int ca = 0; ca < 2; ++ca) {
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) {
        const double gxc = cb ? 1.0 : -1.0, gyc = ca ? 1.0 : -1.0;
        c[3 * (dcj + ca) + (dci + cb)] += sx * (a * gxc + cyx * gyc) + sy * (b * gyc + cxy * gxc);
      }
    }
  }
  const long g = (long)j * nxd + i;
  if (center) c[4] += hc[g];
  const double di = dinv[g];
  const long plane = (long)u.ny * u.nx, o = (long)j * u.nx + i;
#pragma unroll
  for (int k = 0; k < 9; ++k) st[k * plane + o] = c[k] * di;
}

// 1 / diagonal of level 0 in the Krylov layout (duplicated nodes repeat it)
__global__ void k_id0(int nxd, int nyd, Dims u, const double* st, float* id0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= nxd || j >= nyd) return;
  const int ju = j % u.ny, iu = i % u.nx;
  id0[(long)j * nxd + i] = (float)(1.0 / st[4L * u.ny * u.nx + (long)ju * u.nx + iu]);
}

// coarse parents of the fine node 2 J + q (q in [-2, 2]) as offsets from J
// with their interpolation weights; returns the count (1 or 2)
__device__ __forceinline__ int parents(int q, int (&o)[2], double (&w)[2]) {
  if ((q & 1) == 0) {
    o[0] = q / 2;
    w[0] = 1.0;
    return 1;
  }
  o[0] = (q - 1) / 2;
  o[1] = (q + 1) / 2;
  w[0] = w[1] = 0.5;
  return 2;
}

// Galerkin product R A P at coarse node (J, I): R = P^T / 4 gathers the fine
// nodes 2J + a, 2I + b (a, b in -1..1, weights 1 and 1/2 per axis), each
// fine stencil entry reaches 2J + a + dj, 2I + b + di, and P sends that node
// to its coarse parents.  Also writes the FP32 copy and 1 / centre.  The 81
// fine stencil entries are loaded first (at clamped indices past a wall,
// which the sum skips), so the sum waits for one round trip, not one per
// entry behind its zero test.
__global__ __launch_bounds__(64) void k_rap(Dims f, Dims c, const double* stf, double* stc, float* stc32,
                                            float* idg) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y;
  if (I >= c.nx || J >= c.ny) return;
  double acc[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) acc[k] = 0.0;
  const long pf = (long)f.ny * f.nx;
  double svs[3][3][9];
  bool okv[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    bool oky;
    const int jf = axis_idx(2 * J + a - 1, f.ny, f.py, oky);
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      bool okx;
      const int iF = axis_idx(2 * I + b - 1, f.nx, f.px, okx);
      okv[a][b] = oky && okx;
      const long gf = (long)jf * f.nx + iF;
#pragma unroll
      for (int k = 0; k < 9; ++k) svs[a][b][k] = stf[k * pf + gf];
    }
  }
#pragma unroll
  for (int a = -1; a <= 1; ++a) {
#pragma unroll
    for (int b = -1; b <= 1; ++b) {
      if (!okv[a + 1][b + 1]) continue;
      const double wr = 0.25 * (a == 0 ? 1.0 : 0.5) * (b == 0 ? 1.0 : 0.5);
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        const double sv = svs[a + 1][b + 1][k];
        if (sv == 0.0) continue;
        const int dj = k / 3 - 1, di = k % 3 - 1;
        int oy[2], ox[2];
        double wy[2], wx[2];
        const int ny_ = parents(a + dj, oy, wy), nx_ = parents(b + di, ox, wx);
#pragma unroll
        for (int p = 0; p < ny_; ++p)
#pragma unroll
          for (int q = 0; q < nx_; ++q) acc[3 * (oy[p] + 1) + (ox[q] + 1)] += wr * sv * wy[p] * wx[q];
      }
    }
  }
  const long pc = (long)c.ny * c.nx, o = (long)J * c.nx + I;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    stc[k * pc + o] = acc[k];
    stc32[k * pc + o] = (float)acc[k];
  }
  idg[o] = (float)(1.0 / acc[4]);
}

// network features on the duplicated grid: the stencil's x and y coupling
// relative to its centre, -(W + E) / C and -(N + S) / C
__global__ void k_net_features(Dims u, int nh, int nw, const double* st, float* f1, float* f2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= nw || j >= nh) return;
  const long p = (long)u.ny * u.nx, o = (long)(j % u.ny) * u.nx + (i % u.nx);
  const double cc = st[4 * p + o];
  f1[(long)j * nw + i] = (float)(-(st[3 * p + o] + st[5 * p + o]) / cc);
  f2[(long)j * nw + i] = (float)(-(st[1 * p + o] + st[7 * p + o]) / cc);
}

// ---------------------------------------------------------------------------
// level 0: LDS-tiled stencil launches on the Krylov layout
// ---------------------------------------------------------------------------
struct Nothing {};

// first two sweeps from z = 0: z1 = omega v / d formed for the halo, z2 =
// z1 + omega (v - A z1) / d; also stores the FP32 copy of v
struct FirstIn {
  struct Raw {
    double v;
    float id;
  };
  const double* v;
  const float* id;
  float omega;
  __device__ Raw load(long g) const { return Raw{v[g], id[g]}; }
  __device__ float combine(const Raw& q) const { return omega * (float)q.v * q.id; }
};
struct FirstOut {
  using Pre = Nothing;
  float *v0, *z;
  float omega;
  __device__ Pre prefetch(long) const { return Pre{}; }
  __device__ void operator()(int, long g, float y, float pc, const Pre&, const FirstIn::Raw& q) {
    const float vf = (float)q.v;
    v0[g] = vf;
    z[g] = pc + omega * (vf - y) * q.id;
  }
};

struct VecIn {
  using Raw = float;
  const float* x;
  __device__ Raw load(long g) const { return x[g]; }
  __device__ float combine(const Raw& q) const { return q; }
};
// one sweep: out = z + omega (v - A z) / d
struct SweepOut {
  struct Pre {
    float v, id;
  };
  const float *v0, *id0;
  float* out;
  float omega;
  __device__ Pre prefetch(long g) const { return Pre{v0[g], id0[g]}; }
  template <class R>
  __device__ void operator()(int, long g, float y, float pc, const Pre& e, const R&) {
    out[g] = pc + omega * (e.v - y) * e.id;
  }
};
struct ResidOut {
  using Pre = float;
  const float* v0;
  float* r;
  __device__ Pre prefetch(long g) const { return v0[g]; }
  template <class R>
  __device__ void operator()(int, long g, float y, float, const Pre& vg, const R&) {
    r[g] = vg - y;
  }
};
// z + P e_s + c P e_d formed for every node of the halo (the coarse level-1
// corrections of the two branches gathered at the node's parents)
struct ProlongIn {
  using Raw = float;
  const float* z;
  const float* e;   // [2][c.ny][c.nx]
  Dims u, c;        // level 0 unique grid, level 1
  int nxd;
  float dgain;      // 1: add the checkerboard branch, 0: drop it
  __device__ Raw load(long g) const {
    // 32-bit division (the grid holds < 2^30 nodes); the operator grid has
    // at most one duplicated end node past the unique grid on each axis
    const int j = (int)g / nxd, i = (int)g - j * nxd;
    const int ju = j >= u.ny ? j - u.ny : j, iu = i >= u.nx ? i - u.nx : i;
    const int J0 = ju >> 1, I0 = iu >> 1;
    const int J1 = (ju & 1) ? (J0 + 1 == c.ny && c.py ? 0 : J0 + 1) : J0;
    const int I1 = (iu & 1) ? (I0 + 1 == c.nx && c.px ? 0 : I0 + 1) : I0;
    const long pc = (long)c.ny * c.nx;
    const long a00 = (long)J0 * c.nx + I0, a01 = (long)J0 * c.nx + I1, a10 = (long)J1 * c.nx + I0,
               a11 = (long)J1 * c.nx + I1;
    const float es = 0.25f * ((e[a00] + e[a01]) + (e[a10] + e[a11]));
    const float ed = 0.25f * ((e[pc + a00] + e[pc + a01]) + (e[pc + a10] + e[pc + a11]));
    const float sgn = ((ju + iu) & 1) ? -dgain : dgain;
    return z[g] + es + sgn * ed;
  }
  __device__ float combine(const Raw& q) const { return q; }
};

template <bool CROSS, bool CENTER, int RY>
__global__ __launch_bounds__(NTHR) void k0_first(OpDev<float> op, FirstIn in, FirstOut out) {
  __shared__ TileLds<float, RY> L;
  stencil_tile2<float, CROSS, CENTER, RY>(op, blockIdx.x, blockIdx.y, in, out, L, [] { return true; });
}
template <bool CROSS, bool CENTER, int RY>
__global__ __launch_bounds__(NTHR) void k0_sweep(OpDev<float> op, VecIn in, SweepOut out) {
  __shared__ TileLds<float, RY> L;
  stencil_tile2<float, CROSS, CENTER, RY>(op, blockIdx.x, blockIdx.y, in, out, L, [] { return true; });
}
template <bool CROSS, bool CENTER, int RY>
__global__ __launch_bounds__(NTHR) void k0_resid(OpDev<float> op, VecIn in, ResidOut out) {
  __shared__ TileLds<float, RY> L;
  stencil_tile2<float, CROSS, CENTER, RY>(op, blockIdx.x, blockIdx.y, in, out, L, [] { return true; });
}
template <bool CROSS, bool CENTER, int RY>
__global__ __launch_bounds__(NTHR) void k0_prolong_sweep(OpDev<float> op, ProlongIn in, SweepOut out) {
  __shared__ TileLds<float, RY> L;
  stencil_tile2<float, CROSS, CENTER, RY>(op, blockIdx.x, blockIdx.y, in, out, L, [] { return true; });
}

// z = omega v / d, v0 = v (a single pre-sweep)
__global__ void k0_init(long n, const double* v, const float* id, float omega, float* v0, float* z) {
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x) {
    const float vf = (float)v[g];
    v0[g] = vf;
    z[g] = omega * vf * id[g];
  }
}

// full weighting of r (level 0, Krylov layout) and of c * r onto level 1
// A periodic y axis: the Krylov layout's duplicated end row is node row 0;
// the level-0 sweeps leave it holding their own value of that node, which
// the cycle's result replaces by row 0's (a periodic x axis's duplicated
// column already agrees with column 0 to the last bits)
__global__ void k0_dup_row(int nxd, int nyd, float* z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nxd) z[(long)(nyd - 1) * nxd + i] = z[i];
}

__global__ void k0_restrict2(Dims u, int nxd, Dims c, const float* r, float dgain, float* vc) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y;
  if (I >= c.nx || J >= c.ny) return;
  float s = 0.f, d = 0.f;
#pragma unroll
  for (int a = -1; a <= 1; ++a) {
    bool oky;
    const int jf = axis_idx(2 * J + a, u.ny, u.py, oky);
#pragma unroll
    for (int b = -1; b <= 1; ++b) {
      bool okx;
      const int iF = axis_idx(2 * I + b, u.nx, u.px, okx);
      const float w = (oky && okx) ? 0.25f * (a == 0 ? 1.f : 0.5f) * (b == 0 ? 1.f : 0.5f) : 0.f;
      const float rv = w * r[(long)jf * nxd + iF];
      s += rv;
      d += ((jf + iF) & 1) ? -rv : rv;
    }
  }
  const long o = (long)J * c.nx + I, pc = (long)c.ny * c.nx;
  vc[o] = s;
  vc[pc + o] = dgain * d;
}

// ---------------------------------------------------------------------------
// levels >= 1: one thread per node, both branches
// ---------------------------------------------------------------------------
struct Nb {
  int idx[9];
  float s[9];
};

__device__ __forceinline__ void load_row(const Dims& d, const float* st, long g, Nb& nb) {
  const int j = (int)g / d.nx, i = (int)g - j * d.nx;   // levels hold < 2^31 nodes: 32-bit division
  nbr9(d, j, i, nb.idx);
  const long p = (long)d.ny * d.nx;
#pragma unroll
  for (int k = 0; k < 9; ++k) nb.s[k] = st[k * p + g];
}

// z = omega v / d (a single sweep from zero)
__global__ void kc_init(long n, const float* idg, const float* v, float omega, float* z) {
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x) {
    z[g] = omega * v[g] * idg[g];
    z[n + g] = omega * v[n + g] * idg[g];
  }
}

// All ``sweeps`` Jacobi sweeps from z = 0 on a small level in one block, the
// iterate double-buffered in LDS (the coarsest level's solve).
constexpr int kBlockSweepMax = 2048;   // nodes
constexpr int kBlockSweepThreads = 1024;
__global__ __launch_bounds__(kBlockSweepThreads) void kc_sweeps_block(Dims d, const float* st, const float* idg,
                                                                      const float* v, int sweeps, float omega,
                                                                      float* zout) {
  __shared__ float zb[2][2 * kBlockSweepMax];
  const int n = d.ny * d.nx;
  for (int g = threadIdx.x; g < n; g += kBlockSweepThreads) {
    zb[0][g] = omega * v[g] * idg[g];
    zb[0][n + g] = omega * v[n + g] * idg[g];
  }
  __syncthreads();
  int cur = 0;
  for (int k = 1; k < sweeps; ++k) {
    for (int g = threadIdx.x; g < n; g += kBlockSweepThreads) {
      Nb nb;
      load_row(d, st, g, nb);
      const float ig = idg[g];
#pragma unroll
      for (int ch = 0; ch < 2; ++ch) {
        const float* zc = zb[cur] + ch * n;
        float acc = 0.f;
#pragma unroll
        for (int q = 0; q < 9; ++q) acc += nb.s[q] * zc[nb.idx[q]];
        zb[cur ^ 1][ch * n + g] = zc[g] + omega * (v[ch * n + g] - acc) * ig;
      }
    }
    __syncthreads();
    cur ^= 1;
  }
  for (int g = threadIdx.x; g < n; g += kBlockSweepThreads) {
    zout[g] = zb[cur][g];
    zout[n + g] = zb[cur][n + g];
  }
}

// bilinear value of the coarse field e at fine node (j, i)
__device__ __forceinline__ float interp(const Dims& c, const float* e, int j, int i) {
  const int J0 = j >> 1, I0 = i >> 1;
  const int J1 = (j & 1) ? (J0 + 1 == c.ny && c.py ? 0 : J0 + 1) : J0;
  const int I1 = (i & 1) ? (I0 + 1 == c.nx && c.px ? 0 : I0 + 1) : I0;
  return 0.25f * ((e[(long)J0 * c.nx + I0] + e[(long)J0 * c.nx + I1]) +
                  (e[(long)J1 * c.nx + I0] + e[(long)J1 * c.nx + I1]));
}

// Stencil launches of levels 1 and below on tiles of 64 x 4 nodes (block
// (64, 4)).  tile9 forms the values the stencil reads, both branches, at the
// tile's nodes and its one-node halo into LDS once -- at the halo's wrapped
// or clamped neighbour coordinates, as the stencil's neighbour indices take
// them -- and hands each node of the grid its nine neighbours' values from
// there (entry k = 3 (dj + 1) + (di + 1)).  Per node the arithmetic is that
// of gathering each neighbour's value on its own; the gathers (nine per
// node and branch, 36 more coarse values in the prolongation sweep) were
// what bounded these launches.
constexpr int kPsX = 64, kPsY = 4;
inline dim3 tile_grid(const Dims& d) { return dim3(ceil_div(d.nx, kPsX), ceil_div(d.ny, kPsY)); }
const dim3 kTileBlk(kPsX, kPsY);

// fill(ch, jq, iq, gq): the value at node (jq, iq), flat index gq, of branch
// ch; node(g, zn) with zn[ch][k] for each node g of the grid in the tile
template <class Fill, class Node>
__device__ __forceinline__ void tile9(const Dims& d, Fill fill, Node node) {
  __shared__ float zp[2][kPsY + 2][kPsX + 2];
  const int x0 = blockIdx.x * kPsX, y0 = blockIdx.y * kPsY;
  const int t = threadIdx.y * kPsX + threadIdx.x;
  for (int q = t; q < (kPsY + 2) * (kPsX + 2); q += kPsX * kPsY) {
    const int r = q / (kPsX + 2), cc = q - r * (kPsX + 2);
    const int j = y0 - 1 + r, i = x0 - 1 + cc;
    if (j > d.ny || i > d.nx) continue;   // no node of the grid has it as a neighbour
    bool ok;
    const int jq = axis_idx(j, d.ny, d.py, ok), iq = axis_idx(i, d.nx, d.px, ok);
    const long gq = (long)jq * d.nx + iq;
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) zp[ch][r][cc] = fill(ch, jq, iq, gq);
  }
  __syncthreads();
  const int j = y0 + threadIdx.y, i = x0 + threadIdx.x;
  if (j >= d.ny || i >= d.nx) return;
  float zn[2][9];
#pragma unroll
  for (int ch = 0; ch < 2; ++ch)
#pragma unroll
    for (int k = 0; k < 9; ++k) zn[ch][k] = zp[ch][threadIdx.y + k / 3][threadIdx.x + k % 3];
  node((long)j * d.nx + i, zn);
}

__device__ __forceinline__ void load_st(const Dims& d, const float* st, long g, float (&sk)[9]) {
  const long n = (long)d.ny * d.nx;
#pragma unroll
  for (int k = 0; k < 9; ++k) sk[k] = st[k * n + g];
}

// z = omega v / d, then one sweep (the first two pre-sweeps)
__global__ __launch_bounds__(kPsX * kPsY) void kc_first(Dims d, const float* st, const float* idg, const float* v,
                                                       float omega, float* z) {
  const long n = (long)d.ny * d.nx;
  tile9(
      d, [&](int ch, int, int, long gq) { return omega * v[ch * n + gq] * idg[gq]; },
      [&](long g, const float (&zn)[2][9]) {
        float sk[9];
        load_st(d, st, g, sk);
        const float ig = idg[g];
#pragma unroll
        for (int ch = 0; ch < 2; ++ch) {
          float acc = 0.f;
#pragma unroll
          for (int k = 0; k < 9; ++k) acc += sk[k] * zn[ch][k];
          const float vg = v[ch * n + g];
          z[ch * n + g] = omega * vg * ig + omega * (vg - acc) * ig;
        }
      });
}

// out = z + omega (v - A z) / d
__global__ __launch_bounds__(kPsX * kPsY) void kc_sweep(Dims d, const float* st, const float* idg, const float* v,
                                                       const float* z, float omega, float* out) {
  const long n = (long)d.ny * d.nx;
  tile9(
      d, [&](int ch, int, int, long gq) { return z[ch * n + gq]; },
      [&](long g, const float (&zn)[2][9]) {
        float sk[9];
        load_st(d, st, g, sk);
        const float ig = idg[g];
#pragma unroll
        for (int ch = 0; ch < 2; ++ch) {
          float acc = 0.f;
#pragma unroll
          for (int k = 0; k < 9; ++k) acc += sk[k] * zn[ch][k];
          out[ch * n + g] = zn[ch][4] + omega * (v[ch * n + g] - acc) * ig;
        }
      });
}

// v_c = R (v - A z): the residual of both branches and its full-weighting
// restriction in one launch (no fine residual in memory).  Block (bx, by)
// holds z of its 64 x 4 fine tile with a two-node halo and forms the
// residual on the tile with a one-node halo in LDS, at the halo's wrapped
// or clamped neighbour coordinates as the stencil's neighbour indices take
// them, with kc_resid's arithmetic; then the 2 x 32 coarse nodes whose fine
// centres lie in the tile sum their nine weighted residuals in the order of
// a gather of each fine value on its own.  A fine node past a wall, whose
// weight is 0, adds +0 where the gather added 0 * r of the clamped node:
// the same sum, since a round-to-nearest sum from +0 is never -0.
__global__ __launch_bounds__(kPsX * kPsY) void kc_resid_restrict(Dims d, Dims c, const float* st, const float* v,
                                                                const float* z, float* vc) {
  __shared__ float zp[2][kPsY + 4][kPsX + 4];   // z at rows y0 - 2 .., cols x0 - 2 ..
  __shared__ float rp[2][kPsY + 2][kPsX + 2];   // r at rows y0 - 1 .., cols x0 - 1 ..
  const long n = (long)d.ny * d.nx, pc = (long)c.ny * c.nx;
  const int x0 = blockIdx.x * kPsX, y0 = blockIdx.y * kPsY;
  const int t = threadIdx.y * kPsX + threadIdx.x;
  for (int q = t; q < (kPsY + 4) * (kPsX + 4); q += kPsX * kPsY) {
    const int r = q / (kPsX + 4), cc = q - r * (kPsX + 4);
    const int j = y0 - 2 + r, i = x0 - 2 + cc;
    if (j > d.ny || i > d.nx) continue;   // no residual needed here has it as a neighbour
    bool ok;
    const long gq = (long)axis_idx(j, d.ny, d.py, ok) * d.nx + axis_idx(i, d.nx, d.px, ok);
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) zp[ch][r][cc] = z[ch * n + gq];
  }
  __syncthreads();
  for (int q = t; q < (kPsY + 2) * (kPsX + 2); q += kPsX * kPsY) {
    const int rr = q / (kPsX + 2), rc = q - rr * (kPsX + 2);
    const int j = y0 - 1 + rr, i = x0 - 1 + rc;
    // the restriction reads a residual at a wrapped node, or at an in-range
    // node; past a wall (weight 0) and past the grid's end it reads none
    const bool need = j < d.ny && i < d.nx && (j >= 0 || d.py) && (i >= 0 || d.px);
    if (!need) {
      rp[0][rr][rc] = rp[1][rr][rc] = 0.f;
      continue;
    }
    bool ok;
    const long g = (long)axis_idx(j, d.ny, d.py, ok) * d.nx + axis_idx(i, d.nx, d.px, ok);
    float sk[9];
    load_st(d, st, g, sk);
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) {
      float acc = 0.f;
#pragma unroll
      for (int k = 0; k < 9; ++k) acc += sk[k] * zp[ch][rr + k / 3][rc + k % 3];
      rp[ch][rr][rc] = v[ch * n + g] - acc;
    }
  }
  __syncthreads();
  if (t >= 2 * (kPsX / 2)) return;
  const int J = y0 / 2 + t / (kPsX / 2), I = x0 / 2 + t % (kPsX / 2);
  if (J >= c.ny || I >= c.nx) return;
  float s0 = 0.f, s1 = 0.f;
#pragma unroll
  for (int a = -1; a <= 1; ++a) {
    bool oky;
    axis_idx(2 * J + a, d.ny, d.py, oky);
#pragma unroll
    for (int b = -1; b <= 1; ++b) {
      bool okx;
      axis_idx(2 * I + b, d.nx, d.px, okx);
      const float w = (oky && okx) ? 0.25f * (a == 0 ? 1.f : 0.5f) * (b == 0 ? 1.f : 0.5f) : 0.f;
      const int rr = 2 * J + a - (y0 - 1), rc = 2 * I + b - (x0 - 1);
      s0 += w * rp[0][rr][rc];
      s1 += w * rp[1][rr][rc];
    }
  }
  const long o = (long)J * c.nx + I;
  vc[o] = s0;
  vc[pc + o] = s1;
}

// kc_first followed by kc_resid_restrict, in one launch: t = omega v / d on
// the tile with a three-node halo, kc_first's z (the first two sweeps from
// zero, a stencil over t) on the tile with a two-node halo (at a periodic
// edge at the wrapped node; at a wall the halo position takes its clamped
// node's value), the residual v - A z on the tile with a one-node halo and
// its restriction as in kc_resid_restrict; z of the tile's nodes is stored.
// Per node the arithmetic of the two launches.
__global__ __launch_bounds__(kPsX * kPsY) void kc_first_restrict(Dims d, Dims c, const float* st, const float* idg,
                                                                const float* v, float omega, float* z, float* vc) {
  __shared__ float tp[2][kPsY + 6][kPsX + 6];   // omega v / d at rows y0 - 3 .., cols x0 - 3 ..
  __shared__ float zp[2][kPsY + 4][kPsX + 4];   // z at rows y0 - 2 .., cols x0 - 2 ..
  __shared__ float rp[2][kPsY + 2][kPsX + 2];   // r at rows y0 - 1 .., cols x0 - 1 ..
  const long n = (long)d.ny * d.nx, pc = (long)c.ny * c.nx;
  const int x0 = blockIdx.x * kPsX, y0 = blockIdx.y * kPsY;
  const int t = threadIdx.y * kPsX + threadIdx.x;
  for (int q = t; q < (kPsY + 6) * (kPsX + 6); q += kPsX * kPsY) {
    const int r = q / (kPsX + 6), cc = q - r * (kPsX + 6);
    const int j = y0 - 3 + r, i = x0 - 3 + cc;
    if (j > d.ny + 2 || i > d.nx + 2) continue;   // no z needed here has it as a neighbour
    bool ok;
    const long gq = (long)axis_idx(j, d.ny, d.py, ok) * d.nx + axis_idx(i, d.nx, d.px, ok);
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) tp[ch][r][cc] = omega * v[ch * n + gq] * idg[gq];
  }
  __syncthreads();
  // z where its position is a node of the grid or, on a periodic axis,
  // wraps to one (the residual's neighbours reach two nodes past the tile)
  auto inside = [](int k, int nk, int per) { return per ? (k >= -2 && k <= nk + 1) : (k >= 0 && k < nk); };
  for (int q = t; q < (kPsY + 4) * (kPsX + 4); q += kPsX * kPsY) {
    const int r = q / (kPsX + 4), cc = q - r * (kPsX + 4);
    const int j = y0 - 2 + r, i = x0 - 2 + cc;
    if (!inside(j, d.ny, d.py) || !inside(i, d.nx, d.px)) continue;
    bool ok;
    const long g = (long)axis_idx(j, d.ny, d.py, ok) * d.nx + axis_idx(i, d.nx, d.px, ok);
    float sk[9];
    load_st(d, st, g, sk);
    const float ig = idg[g];
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) {
      float acc = 0.f;
#pragma unroll
      for (int k = 0; k < 9; ++k) acc += sk[k] * tp[ch][r + k / 3][cc + k % 3];
      const float vg = v[ch * n + g];
      zp[ch][r][cc] = omega * vg * ig + omega * (vg - acc) * ig;
    }
  }
  __syncthreads();
  // wall halo one node past the tile's nodes: the clamped node's value (an
  // axis on which the position is inside keeps it)
  for (int q = t; q < (kPsY + 4) * (kPsX + 4); q += kPsX * kPsY) {
    const int r = q / (kPsX + 4), cc = q - r * (kPsX + 4);
    const int j = y0 - 2 + r, i = x0 - 2 + cc;
    if (j < -1 || j > d.ny || i < -1 || i > d.nx) continue;
    if (inside(j, d.ny, d.py) && inside(i, d.nx, d.px)) continue;
    bool ok;
    const int mr = inside(j, d.ny, d.py) ? r : axis_idx(j, d.ny, d.py, ok) - (y0 - 2);
    const int mc = inside(i, d.nx, d.px) ? cc : axis_idx(i, d.nx, d.px, ok) - (x0 - 2);
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) zp[ch][r][cc] = zp[ch][mr][mc];
  }
  __syncthreads();
  {
    const int j = y0 + threadIdx.y, i = x0 + threadIdx.x;
    if (j < d.ny && i < d.nx) {
      const long g = (long)j * d.nx + i;
#pragma unroll
      for (int ch = 0; ch < 2; ++ch) z[ch * n + g] = zp[ch][threadIdx.y + 2][threadIdx.x + 2];
    }
  }
  for (int q = t; q < (kPsY + 2) * (kPsX + 2); q += kPsX * kPsY) {
    const int rr = q / (kPsX + 2), rc = q - rr * (kPsX + 2);
    const int j = y0 - 1 + rr, i = x0 - 1 + rc;
    const bool need = j < d.ny && i < d.nx && (j >= 0 || d.py) && (i >= 0 || d.px);
    if (!need) {
      rp[0][rr][rc] = rp[1][rr][rc] = 0.f;
      continue;
    }
    bool ok;
    const long g = (long)axis_idx(j, d.ny, d.py, ok) * d.nx + axis_idx(i, d.nx, d.px, ok);
    float sk[9];
    load_st(d, st, g, sk);
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) {
      float acc = 0.f;
#pragma unroll
      for (int k = 0; k < 9; ++k) acc += sk[k] * zp[ch][rr + k / 3][rc + k % 3];
      rp[ch][rr][rc] = v[ch * n + g] - acc;
    }
  }
  __syncthreads();
  if (t >= 2 * (kPsX / 2)) return;
  const int J = y0 / 2 + t / (kPsX / 2), I = x0 / 2 + t % (kPsX / 2);
  if (J >= c.ny || I >= c.nx) return;
  float s0 = 0.f, s1 = 0.f;
#pragma unroll
  for (int a = -1; a <= 1; ++a) {
    bool oky;
    axis_idx(2 * J + a, d.ny, d.py, oky);
#pragma unroll
    for (int b = -1; b <= 1; ++b) {
      bool okx;
      axis_idx(2 * I + b, d.nx, d.px, okx);
      const float w = (oky && okx) ? 0.25f * (a == 0 ? 1.f : 0.5f) * (b == 0 ? 1.f : 0.5f) : 0.f;
      const int rr = 2 * J + a - (y0 - 1), rc = 2 * I + b - (x0 - 1);
      s0 += w * rp[0][rr][rc];
      s1 += w * rp[1][rr][rc];
    }
  }
  const long o = (long)J * c.nx + I;
  vc[o] = s0;
  vc[pc + o] = s1;
}

// z' = z + P e, then out = z' + omega (v - A z') / d
__global__ __launch_bounds__(kPsX * kPsY) void kc_prolong_sweep(Dims d, Dims c, const float* st, const float* idg,
                                                               const float* v, const float* z, const float* e,
                                                               float omega, float* out) {
  const long n = (long)d.ny * d.nx, pc = (long)c.ny * c.nx;
  tile9(
      d, [&](int ch, int jq, int iq, long gq) { return z[ch * n + gq] + interp(c, e + ch * pc, jq, iq); },
      [&](long g, const float (&zn)[2][9]) {
        float sk[9];
        load_st(d, st, g, sk);
        const float ig = idg[g];
#pragma unroll
        for (int ch = 0; ch < 2; ++ch) {
          float acc = 0.f;
#pragma unroll
          for (int k = 0; k < 9; ++k) acc += sk[k] * zn[ch][k];
          out[ch * n + g] = zn[ch][4] + omega * (v[ch * n + g] - acc) * ig;
        }
      });
}

// kc_prolong_sweep followed by one kc_sweep, in one launch: z' = z + P e on
// the tile with a two-node halo, the first sweep on the tile with a
// one-node halo (at a periodic edge at the wrapped node; at a wall the halo
// position takes its clamped node's value, which is what the second sweep's
// gather reads there), the second sweep at the tile's nodes.  Per node the
// arithmetic of the two launches.
__global__ __launch_bounds__(kPsX * kPsY) void kc_prolong_sweep2(Dims d, Dims c, const float* st, const float* idg,
                                                                const float* v, const float* z, const float* e,
                                                                float omega, float* out) {
  __shared__ float zp[2][kPsY + 4][kPsX + 4];   // z' at rows y0 - 2 .., cols x0 - 2 ..
  __shared__ float s1[2][kPsY + 2][kPsX + 2];   // first sweep at rows y0 - 1 .., cols x0 - 1 ..
  const long n = (long)d.ny * d.nx, pc = (long)c.ny * c.nx;
  const int x0 = blockIdx.x * kPsX, y0 = blockIdx.y * kPsY;
  const int t = threadIdx.y * kPsX + threadIdx.x;
  for (int q = t; q < (kPsY + 4) * (kPsX + 4); q += kPsX * kPsY) {
    const int r = q / (kPsX + 4), cc = q - r * (kPsX + 4);
    const int j = y0 - 2 + r, i = x0 - 2 + cc;
    if (j > d.ny + 1 || i > d.nx + 1) continue;   // no first-sweep value needed here has it as a neighbour
    bool ok;
    const int jq = axis_idx(j, d.ny, d.py, ok), iq = axis_idx(i, d.nx, d.px, ok);
    const long gq = (long)jq * d.nx + iq;
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) zp[ch][r][cc] = z[ch * n + gq] + interp(c, e + ch * pc, jq, iq);
  }
  __syncthreads();
  // first sweep where its position is a node of the grid or, on a periodic
  // axis, wraps to one
  auto inside = [](int k, int nk, int per) { return per ? (k >= -1 && k <= nk) : (k >= 0 && k < nk); };
  for (int q = t; q < (kPsY + 2) * (kPsX + 2); q += kPsX * kPsY) {
    const int rr = q / (kPsX + 2), rc = q - rr * (kPsX + 2);
    const int j = y0 - 1 + rr, i = x0 - 1 + rc;
    if (!inside(j, d.ny, d.py) || !inside(i, d.nx, d.px)) continue;
    bool ok;
    const long g = (long)axis_idx(j, d.ny, d.py, ok) * d.nx + axis_idx(i, d.nx, d.px, ok);
    float sk[9];
    load_st(d, st, g, sk);
    const float ig = idg[g];
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) {
      float acc = 0.f;
#pragma unroll
      for (int k = 0; k < 9; ++k) acc += sk[k] * zp[ch][rr + k / 3][rc + k % 3];
      s1[ch][rr][rc] = zp[ch][rr + 1][rc + 1] + omega * (v[ch * n + g] - acc) * ig;
    }
  }
  __syncthreads();
  // wall halo: the clamped node's value (its position lies in the same tile)
  for (int q = t; q < (kPsY + 2) * (kPsX + 2); q += kPsX * kPsY) {
    const int rr = q / (kPsX + 2), rc = q - rr * (kPsX + 2);
    const int j = y0 - 1 + rr, i = x0 - 1 + rc;
    if (j < -1 || j > d.ny || i < -1 || i > d.nx) continue;
    if (inside(j, d.ny, d.py) && inside(i, d.nx, d.px)) continue;
    // an axis on which the position is inside keeps it (a periodic one holds
    // the wrapped node's value already); a wall axis clamps
    bool ok;
    const int mr = inside(j, d.ny, d.py) ? rr : axis_idx(j, d.ny, d.py, ok) - (y0 - 1);
    const int mc = inside(i, d.nx, d.px) ? rc : axis_idx(i, d.nx, d.px, ok) - (x0 - 1);
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) s1[ch][rr][rc] = s1[ch][mr][mc];
  }
  __syncthreads();
  const int j = y0 + threadIdx.y, i = x0 + threadIdx.x;
  if (j >= d.ny || i >= d.nx) return;
  const long g = (long)j * d.nx + i;
  float sk[9];
  load_st(d, st, g, sk);
  const float ig = idg[g];
#pragma unroll
  for (int ch = 0; ch < 2; ++ch) {
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < 9; ++k) acc += sk[k] * s1[ch][threadIdx.y + k / 3][threadIdx.x + k % 3];
    out[ch * n + g] = s1[ch][threadIdx.y + 1][threadIdx.x + 1] + omega * (v[ch * n + g] - acc) * ig;
  }
}

// The network level's correction, one block of 1024 threads per branch
// (blockIdx.x), with the level in LDS:
//  1. z = scale * (network output) at every node, a duplicated end node's
//     copies averaged (the network's nh x nw grid duplicates the end node of
//     each periodic axis of the d.ny x d.nx level); each copy's product is
//     rounded to float before the sum (FP contraction off there);
//  2. z *= alpha, alpha = (z . v) / (z . A z), the energy-norm (Galerkin)
//     scale of the direction z: the network predicts the shape of the coarse
//     correction, its scale is fitted here.  A residual fit
//     (A z . v) / (A z . A z) would weight the high-frequency part of A z and
//     shrink the smooth modes the coarse level is there for.  The sums run in
//     double, thread t over nodes t, t + 1024, ..., then the block sum;
//  3. ``sweeps`` damped Jacobi sweeps z += omega (v - A z) / diag.
// The result goes to zfin.  Each step's arithmetic per node is that of the
// separate launches it replaces (output scaling, the fit, kc_sweep).  Thread
// t owns nodes t, t + 1024, t + 2048 and loads everything they need (the
// network's outputs, stencil rows, diag_inv, v) before the first step, so
// the steps run from registers and LDS with no further round trip to memory.
constexpr int kNetFinThreads = 1024;
constexpr int kNetFinMax = 65 * 33;   // nodes of the network level, at most
constexpr int kNetFinNpt = (kNetFinMax + kNetFinThreads - 1) / kNetFinThreads;   // nodes per thread
__global__ __launch_bounds__(kNetFinThreads) void kc_net_finish(Dims d, int nh, int nw, const float* zo,
                                                                const float* scale, const float* st, const float* idg,
                                                                const float* v, int sweeps, float omega,
                                                                float* zfin) {
  __shared__ float zb[2][kNetFinMax];
  __shared__ double sh[2 * (kNetFinThreads / kWave)];
  __shared__ float alpha;
  const int n = d.ny * d.nx, t = threadIdx.x, br = blockIdx.x;
  const float* qb = zo + br * nh * nw;
  const float* vb = v + br * n;
  const float sf = scale[br];
  // loads: per owned node its stencil row, diag_inv, v and the network's
  // values at every copy of the node (q[0..3]: itself, the duplicated
  // column, the duplicated row, the duplicated corner)
  Nb nb[kNetFinNpt];
  float ig[kNetFinNpt], vv[kNetFinNpt], q[kNetFinNpt][4];
  bool dxs[kNetFinNpt], dys[kNetFinNpt];
#pragma unroll
  for (int m = 0; m < kNetFinNpt; ++m) {
    const int g = t + m * kNetFinThreads;
    const int gg = g < n ? g : 0;
    load_row(d, st, gg, nb[m]);
    ig[m] = idg[gg];
    vv[m] = vb[gg];
    const int j = gg / d.nx, i = gg - j * d.nx;
    dxs[m] = d.px && i == 0;
    dys[m] = d.py && j == 0;
    q[m][0] = qb[j * nw + i];
    q[m][1] = dxs[m] ? qb[j * nw + (nw - 1)] : 0.f;
    q[m][2] = dys[m] ? qb[(nh - 1) * nw + i] : 0.f;
    q[m][3] = dxs[m] && dys[m] ? qb[(nh - 1) * nw + (nw - 1)] : 0.f;
  }
#pragma unroll
  for (int m = 0; m < kNetFinNpt; ++m) {
#pragma clang fp contract(off)
    const int g = t + m * kNetFinThreads;
    if (g < n) {
      float a = q[m][0] * sf;
      float cnt = 1.f;
      if (dxs[m]) {
        a += q[m][1] * sf;
        cnt += 1.f;
      }
      if (dys[m]) {
        a += q[m][2] * sf;
        cnt += 1.f;
      }
      if (dxs[m] && dys[m]) {
        a += q[m][3] * sf;
        cnt += 1.f;
      }
      zb[0][g] = a / cnt;
    }
  }
  __syncthreads();
  double acc[2] = {0.0, 0.0};   // z.v, z.Az
#pragma unroll
  for (int m = 0; m < kNetFinNpt; ++m) {
    const int g = t + m * kNetFinThreads;
    if (g < n) {
      float az = 0.f;
#pragma unroll
      for (int k = 0; k < 9; ++k) az += nb[m].s[k] * zb[0][nb[m].idx[k]];
      const double zg = (double)zb[0][g];
      acc[0] += zg * (double)vv[m];
      acc[1] += zg * (double)az;
    }
  }
  block_sum_w<2, kNetFinThreads / kWave>(acc, sh);
  if (t == 0) alpha = acc[1] != 0.0 ? (float)(acc[0] / acc[1]) : 0.f;
  __syncthreads();
  const float al = alpha;
#pragma unroll
  for (int m = 0; m < kNetFinNpt; ++m) {
    const int g = t + m * kNetFinThreads;
    if (g < n) zb[0][g] *= al;
  }
  __syncthreads();
  int cur = 0;
  for (int k = 0; k < sweeps; ++k) {
#pragma unroll
    for (int m = 0; m < kNetFinNpt; ++m) {
      const int g = t + m * kNetFinThreads;
      if (g < n) {
        const float* zc = zb[cur];
        float acc1 = 0.f;
#pragma unroll
        for (int qq = 0; qq < 9; ++qq) acc1 += nb[m].s[qq] * zc[nb[m].idx[qq]];
        zb[cur ^ 1][g] = zc[g] + omega * (vv[m] - acc1) * ig[m];
      }
    }
    __syncthreads();
    cur ^= 1;
  }
#pragma unroll
  for (int m = 0; m < kNetFinNpt; ++m) {
    const int g = t + m * kNetFinThreads;
    if (g < n) zfin[br * n + g] = zb[cur][g];
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
inline int blocks1(long n) { return (int)std::max<long>(1, (n + 255) / 256); }

template <typename T>
int hmalloc(pb_op2d* op, Hier* H, T** p, long count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, sizeof(T) * (size_t)std::max<long>(count, 1));
  if (e != hipSuccess)
    return pb_fail(op->ctx, PB_ERR_HIP, "multigrid: hipMalloc(%ld bytes) failed: %s", (long)(sizeof(T) * count),
                   hipGetErrorString(e));
  H->allocs.push_back(q);
  *p = (T*)q;
  return PB_OK;
}

bool can_coarsen(int n, int periodic) { return periodic ? (n % 2 == 0 && n / 2 >= 3) : (n % 2 == 1 && (n + 1) / 2 >= 3); }
int coarse_n(int n, int periodic) { return periodic ? n / 2 : (n + 1) / 2; }

int layout(pb_op2d* op, Hier* H) {
  const pb_grid2d& g = op->g;
  Dims d0{g.ny, g.nx, g.bc_y == PB_BC_PERIODIC, g.bc_x == PB_BC_PERIODIC};
  if (d0.py) d0.ny -= 1;
  if (d0.px) d0.nx -= 1;
  H->L[0].d = d0;
  H->nlev = 1;
  while (H->nlev < kMaxLevels) {
    const Dims f = H->L[H->nlev - 1].d;
    if (!can_coarsen(f.ny, f.py) || !can_coarsen(f.nx, f.px)) break;
    H->L[H->nlev].d = Dims{coarse_n(f.ny, f.py), coarse_n(f.nx, f.px), f.py, f.px};
    H->nlev += 1;
  }
  if (H->nlev < 2)
    return pb_fail(op->ctx, PB_ERR_ARG, "multigrid: a %d x %d grid has no coarse level", g.nx, g.ny);
  // the network level: the first coarse level whose duplicated grid is 8k+1
  // nodes per axis (>= 9) and at most 65 x 33 nodes, where the network's
  // receptive field (about +-60 nodes) spans the whole grid
  H->nn_level = -1;
  for (int l = 1; l < H->nlev; ++l) {
    const Dims d = H->L[l].d;
    const int h = d.ny + (d.py ? 1 : 0), w = d.nx + (d.px ? 1 : 0);
    if (h >= 9 && w >= 9 && ((h - 1) & 7) == 0 && ((w - 1) & 7) == 0 && (long)h * w <= 65L * 33) {
      H->nn_level = l;
      H->nh = h;
      H->nw = w;
      break;
    }
  }
  int rc = 0;
  for (int l = 0; l < H->nlev; ++l) {
    Level& L = H->L[l];
    const long n = (long)L.d.ny * L.d.nx;
    rc |= hmalloc(op, H, &L.st64, 9 * n);
    if (l >= 1) {
      rc |= hmalloc(op, H, &L.st, 9 * n);
      rc |= hmalloc(op, H, &L.idg, n);
      rc |= hmalloc(op, H, &L.v, 2 * n);
      rc |= hmalloc(op, H, &L.z, 2 * n);
      rc |= hmalloc(op, H, &L.t, 2 * n);
    }
  }
  rc |= hmalloc(op, H, &H->id0, op->n);
  rc |= hmalloc(op, H, &H->v0, op->n);
  rc |= hmalloc(op, H, &H->z0, op->n);
  rc |= hmalloc(op, H, &H->t0, op->n);
  rc |= hmalloc(op, H, &H->r0, op->n);
  if (H->nn_level >= 0) {
    const long nd = (long)H->nh * H->nw;
    rc |= hmalloc(op, H, &H->f1, nd);
    rc |= hmalloc(op, H, &H->f2, nd);
  }
  return rc ? PB_ERR_HIP : PB_OK;
}

#define PB_MG_LAUNCH0(KERNEL, RYV, grd, s, ...)                                                        \
  do {                                                                                                 \
    const dim3 _blk(TX, TYT);                                                                          \
    if (op->cross && op->center) hipLaunchKernelGGL((KERNEL<true, true, RYV>), grd, _blk, 0, s, __VA_ARGS__);   \
    else if (op->cross) hipLaunchKernelGGL((KERNEL<true, false, RYV>), grd, _blk, 0, s, __VA_ARGS__);         \
    else if (op->center) hipLaunchKernelGGL((KERNEL<false, true, RYV>), grd, _blk, 0, s, __VA_ARGS__);        \
    else hipLaunchKernelGGL((KERNEL<false, false, RYV>), grd, _blk, 0, s, __VA_ARGS__);                       \
  } while (0)

constexpr int kRY0 = 4;   // level-0 tile height (rows per thread)

// returns the buffer holding level l's correction (2 channels)
const float* vcycle(pb_op2d* op, Hier* H, pb_nn* nn, int l, int last, const CycleOpts& o, hipStream_t s, int* rc) {
  Level& L = H->L[l];
  const Dims d = L.d;
  const long n = (long)d.ny * d.nx;
  const int nb = blocks1(n);
  float* cur = L.z;
  float* alt = L.t;
  auto swap = [&] { std::swap(cur, alt); };
  if (l == last) {
    if (o.coarse == COARSE_NET && l == H->nn_level && nn) {
      if (n > kNetFinMax || (long)H->nh * H->nw > kNetFinMax) {
        *rc = pb_fail(op->ctx, PB_ERR_ARG, "multigrid: network level of %ld nodes exceeds %d", n, kNetFinMax);
        return nullptr;
      }
      const float *zo = nullptr, *sc = nullptr;
      *rc = pb_nn_forward_level(nn, 2, d.ny, d.nx, H->nh, H->nw, (const float*)L.v, H->f1, H->f2, &zo, &sc, s);
      if (*rc) return nullptr;
      hipLaunchKernelGGL(kc_net_finish, dim3(2), dim3(kNetFinThreads), 0, s, d, H->nh, H->nw, zo, sc,
                         (const float*)L.st, (const float*)L.idg, (const float*)L.v, std::max(0, o.coarse_sweeps),
                         o.omega, cur);
      return cur;
    }
    const int sweeps = std::max(1, o.coarse == COARSE_MG ? o.bottom_sweeps : o.coarse_sweeps);
    if (n <= kBlockSweepMax) {
      hipLaunchKernelGGL(kc_sweeps_block, dim3(1), dim3(kBlockSweepThreads), 0, s, d, (const float*)L.st,
                         (const float*)L.idg, (const float*)L.v, sweeps, o.omega, cur);
      return cur;
    }
    hipLaunchKernelGGL(kc_init, dim3(std::min(nb, 1024)), dim3(256), 0, s, n, (const float*)L.idg,
                       (const float*)L.v, o.omega, cur);
    for (int k = 1; k < sweeps; ++k) {
      hipLaunchKernelGGL(kc_sweep, tile_grid(d), kTileBlk, 0, s, d, (const float*)L.st, (const float*)L.idg,
                         (const float*)L.v, (const float*)cur, o.omega, alt);
      swap();
    }
    return cur;
  }
  // pre-smoothing, residual and restriction
  Level& C = H->L[l + 1];
  if (o.nu_pre == 2) {   // the first two sweeps, the residual and its restriction in one launch
    hipLaunchKernelGGL(kc_first_restrict, tile_grid(d), kTileBlk, 0, s, d, C.d, (const float*)L.st,
                       (const float*)L.idg, (const float*)L.v, o.omega, cur, C.v);
  } else {
    if (o.nu_pre > 2) {
      hipLaunchKernelGGL(kc_first, tile_grid(d), kTileBlk, 0, s, d, (const float*)L.st, (const float*)L.idg,
                         (const float*)L.v, o.omega, cur);
      for (int k = 2; k < o.nu_pre; ++k) {
        hipLaunchKernelGGL(kc_sweep, tile_grid(d), kTileBlk, 0, s, d, (const float*)L.st, (const float*)L.idg,
                           (const float*)L.v, (const float*)cur, o.omega, alt);
        swap();
      }
    } else {
      hipLaunchKernelGGL(kc_init, dim3(std::min(nb, 1024)), dim3(256), 0, s, n, (const float*)L.idg,
                         (const float*)L.v, o.omega, cur);
    }
    hipLaunchKernelGGL(kc_resid_restrict, tile_grid(d), kTileBlk, 0, s, d, C.d, (const float*)L.st,
                       (const float*)L.v, (const float*)cur, C.v);
  }
  const float* e = vcycle(op, H, nn, l + 1, last, o, s, rc);
  if (!e) return nullptr;
  // coarse correction and post-smoothing (at least one sweep)
  int done = 1;
  if (o.nu_post >= 2) {   // the prolongation sweep and the next sweep in one launch
    hipLaunchKernelGGL(kc_prolong_sweep2, tile_grid(d), kTileBlk, 0, s, d, C.d, (const float*)L.st,
                       (const float*)L.idg, (const float*)L.v, (const float*)cur, e, o.omega, alt);
    done = 2;
  } else {
    hipLaunchKernelGGL(kc_prolong_sweep, tile_grid(d), kTileBlk, 0, s, d, C.d, (const float*)L.st,
                       (const float*)L.idg, (const float*)L.v, (const float*)cur, e, o.omega, alt);
  }
  swap();
  for (int k = done; k < o.nu_post; ++k) {
    hipLaunchKernelGGL(kc_sweep, tile_grid(d), kTileBlk, 0, s, d, (const float*)L.st, (const float*)L.idg,
                       (const float*)L.v, (const float*)cur, o.omega, alt);
    swap();
  }
  return cur;
}

}  // namespace

int build(pb_op2d* op, hipStream_t s) {
  if (!op->mg) {
    op->mg = new Hier();
    int rc = layout(op, op->mg);
    if (rc) {
      release(op);
      return rc;
    }
  }
  Hier* H = op->mg;
  if (H->version == op->version) return PB_OK;
  const Dims u = H->L[0].d;
  hipLaunchKernelGGL(k_expand9, dim3(ceil_div(u.nx, 64), u.ny), dim3(64), 0, s, op->g.nx, u, (const double*)op->wxx,
                     (const double*)op->wyy, (const double*)op->wyx, (const double*)op->wxy, (const double*)op->hc,
                     (const double*)op->dinv, op->cross ? 1 : 0, op->center ? 1 : 0, H->L[0].st64);
  hipLaunchKernelGGL(k_id0, dim3(ceil_div(op->g.nx, 64), op->g.ny), dim3(64), 0, s, op->g.nx, op->g.ny, u,
                     (const double*)H->L[0].st64, H->id0);
  for (int l = 0; l + 1 < H->nlev; ++l) {
    const Dims c = H->L[l + 1].d;
    hipLaunchKernelGGL(k_rap, dim3(ceil_div(c.nx, 64), c.ny), dim3(64), 0, s, H->L[l].d, c,
                       (const double*)H->L[l].st64, H->L[l + 1].st64, H->L[l + 1].st, H->L[l + 1].idg);
  }
  if (H->nn_level >= 0) {
    const Level& N = H->L[H->nn_level];
    hipLaunchKernelGGL(k_net_features, dim3(ceil_div(H->nw, 64), H->nh), dim3(64), 0, s, N.d, H->nh, H->nw,
                       (const double*)N.st64, H->f1, H->f2);
  }
  PB_LAUNCH_CHECK(op->ctx);
  H->version = op->version;
  return PB_OK;
}

int cycle(pb_op2d* op, pb_nn* nn, const double* v, float* z, const CycleOpts& o, hipStream_t s) {
  Hier* H = op->mg;
  if (!H || H->version != op->version) return pb_fail(op->ctx, PB_ERR_STATE, "multigrid: hierarchy not built");
  if (o.coarse == COARSE_NET && (!nn || H->nn_level < 0))
    return pb_fail(op->ctx, PB_ERR_STATE, nn ? "multigrid: no coarse level of this grid fits the network"
                                             : "multigrid: no network attached");
  if (o.nu_pre < 1 || o.nu_post < 1) return pb_fail(op->ctx, PB_ERR_ARG, "multigrid: at least one sweep each way");
  const int last = (o.coarse == COARSE_MG || H->nn_level < 0) ? H->nlev - 1 : H->nn_level;
  const OpDev<float> d = opdev<float>(op);
  const dim3 grd = stencil_grid(op, kRY0);
  const Dims u = H->L[0].d, c1 = H->L[1].d;
  const int nxd = op->g.nx;
  const float dgain = o.branches == 2 ? 1.f : 0.f;
  float* cur = H->z0;
  float* alt = H->t0;
  if (o.nu_pre >= 2) {
    FirstIn in{v, H->id0, o.omega};
    FirstOut out{H->v0, cur, o.omega};
    PB_MG_LAUNCH0(k0_first, kRY0, grd, s, d, in, out);
    for (int k = 2; k < o.nu_pre; ++k) {
      PB_MG_LAUNCH0(k0_sweep, kRY0, grd, s, d, VecIn{cur}, SweepOut{H->v0, H->id0, alt, o.omega});
      std::swap(cur, alt);
    }
  } else {
    hipLaunchKernelGGL(k0_init, dim3(elem_blocks(op->n)), dim3(256), 0, s, op->n, v, (const float*)H->id0, o.omega,
                       H->v0, cur);
  }
  PB_MG_LAUNCH0(k0_resid, kRY0, grd, s, d, VecIn{cur}, ResidOut{H->v0, H->r0});
  hipLaunchKernelGGL(k0_restrict2, dim3(ceil_div(c1.nx, 64), c1.ny), dim3(64), 0, s, u, nxd, c1,
                     (const float*)H->r0, dgain, H->L[1].v);
  int rc = PB_OK;
  const float* e = vcycle(op, H, nn, 1, last, o, s, &rc);
  if (!e) return rc ? rc : PB_ERR_STATE;
  ProlongIn pin{cur, e, u, c1, nxd, dgain};
  float* dst = o.nu_post == 1 ? z : alt;
  PB_MG_LAUNCH0(k0_prolong_sweep, kRY0, grd, s, d, pin, SweepOut{H->v0, H->id0, dst, o.omega});
  cur = dst;
  alt = H->z0 == cur ? H->t0 : H->z0;
  for (int k = 1; k < o.nu_post; ++k) {
    float* out = k + 1 == o.nu_post ? z : alt;
    PB_MG_LAUNCH0(k0_sweep, kRY0, grd, s, d, VecIn{cur}, SweepOut{H->v0, H->id0, out, o.omega});
    alt = cur;
    cur = out;
  }
  if (u.py) hipLaunchKernelGGL(k0_dup_row, dim3(ceil_div(nxd, 256)), dim3(256), 0, s, nxd, op->g.ny, z);
  PB_LAUNCH_CHECK(op->ctx);
  return PB_OK;
}

void release(pb_op2d* op) {
  if (!op->mg) return;
  for (void* p : op->mg->allocs) (void)hipFree(p);
  delete op->mg;
  op->mg = nullptr;
}

}  // namespace pbmg

// ---------------------------------------------------------------------------
// C ABI: the cycle and its hierarchy, for tests and network training
// ---------------------------------------------------------------------------
extern "C" int pb_op2d_precond(pb_op2d* op, const double* v, float* z, int coarse, int nu, double omega,
                               int branches) {
  if (!op) return PB_ERR_ARG;
  if (!v || !z) return pb_fail(op->ctx, PB_ERR_ARG, "pb_op2d_precond: null vector");
  if (!op->assembled) return pb_fail(op->ctx, PB_ERR_STATE, "pb_op2d_precond: operator not assembled");
  if (coarse < 0 || coarse > 2 || nu < 1 || nu > 8 || !(omega > 0) || (branches != 1 && branches != 2))
    return pb_fail(op->ctx, PB_ERR_ARG, "pb_op2d_precond: bad option");
  hipStream_t s = op->ctx->stream;
  int rc = pbmg::build(op, s);
  if (rc) return rc;
  pbmg::CycleOpts o;
  o.nu_pre = o.nu_post = nu;
  o.omega = (float)omega;
  o.coarse = coarse;
  o.branches = branches;
  return pbmg::cycle(op, op->nn, v, z, o, s);
}

extern "C" int pb_op2d_mg_info(pb_op2d* op, int* dims, int max_levels, int* nn_level) {
  if (!op || !dims || max_levels < 1) return PB_ERR_ARG;
  if (!op->assembled) return pb_fail(op->ctx, PB_ERR_STATE, "pb_op2d_mg_info: operator not assembled");
  int rc = pbmg::build(op, op->ctx->stream);
  if (rc) return rc;
  const pbmg::Hier* H = op->mg;
  for (int l = 0; l < H->nlev && l < max_levels; ++l) {
    const pbmg::Dims& d = H->L[l].d;
    dims[4 * l] = d.ny;
    dims[4 * l + 1] = d.nx;
    dims[4 * l + 2] = d.py;
    dims[4 * l + 3] = d.px;
  }
  if (nn_level) *nn_level = H->nn_level;
  return H->nlev;
}

extern "C" int pb_op2d_mg_stencil(pb_op2d* op, int level, double* st) {
  if (!op || !st) return PB_ERR_ARG;
  int rc = pbmg::build(op, op->ctx->stream);
  if (rc) return rc;
  const pbmg::Hier* H = op->mg;
  if (level < 0 || level >= H->nlev) return pb_fail(op->ctx, PB_ERR_ARG, "pb_op2d_mg_stencil: no level %d", level);
  const pbmg::Dims& d = H->L[level].d;
  PB_HIP(op->ctx, hipMemcpyAsync(st, H->L[level].st64, sizeof(double) * 9 * (size_t)d.ny * d.nx,
                                 hipMemcpyDeviceToDevice, op->ctx->stream));
  return PB_OK;
}

extern "C" int pb_op2d_mg_features(pb_op2d* op, float* f1, float* f2) {
  if (!op || !f1 || !f2) return PB_ERR_ARG;
  int rc = pbmg::build(op, op->ctx->stream);
  if (rc) return rc;
  const pbmg::Hier* H = op->mg;
  if (H->nn_level < 0) return pb_fail(op->ctx, PB_ERR_STATE, "pb_op2d_mg_features: no network level");
  const size_t nd = (size_t)H->nh * H->nw;
  PB_HIP(op->ctx, hipMemcpyAsync(f1, H->f1, sizeof(float) * nd, hipMemcpyDeviceToDevice, op->ctx->stream));
  PB_HIP(op->ctx, hipMemcpyAsync(f2, H->f2, sizeof(float) * nd, hipMemcpyDeviceToDevice, op->ctx->stream));
  return PB_OK;
}
