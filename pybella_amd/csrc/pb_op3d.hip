// 3D nodal 27-point operator and its BiCGStab solve (gfx950).
//
// Operator (lap3D, laplacian.py:872-1026, restated on a clean layout): nodes
// (nz, ny, nx) with x fastest; cell weights h on the (nz+1, ny+1, nx+1)
// operator cells (cell (k, j, i) has corner nodes k-1..k, j-1..j, i-1..i);
// per cell the differences along each axis summed over the cube's 4 edges,
// F_x = h g_x / (16 dx^2) (and y, z); node value = sum over the 8 cells of
// (+F on the cell's high side of the node, -F on its low side) + hc p, times
// diag_inv.  Periodic axes wrap onto the duplicated end node (i-1 -> n-2 at
// i = 0, i+1 -> 1 at i = n-1, as lap3D's ghost copy, :927-958); wall axes
// zero the weights of the outside cell layers (:933-965).  The x-z Coriolis
// cross term of lap3D (:1017-1022) is zero for every state here
// (coriolis_strength[0] = 0) and is not built.
//
// Kernel shape: a 2.5D march.  A 512-thread block owns a 64 x 8 node tile
// and a chunk of 16 z planes; two node planes (with a one-node halo) and one
// plane of cell fluxes live in LDS.  Each cell plane is computed once and
// used by the node planes below and above it: for node plane k,
//   y = D(k-1/2) + D(k+1/2) + Z(k+1/2) - Z(k-1/2)
// with D the in-plane flux divergence and Z the z-flux sum of the 4 cells
// of a cell plane around the node.
//
// BiCGStab: the two-launch iteration of the 2D solver (x and r updates of
// the previous iteration, p and v = A p in launch 1; s, t = A s and the five
// dots in launch 2).  Each launch first applies the scalar step of the launch
// before it: every block reduces that launch's block partials itself, in the
// same fixed order (deterministic), and block 0 writes the advanced state to
// the other of two state slots.  A one-block kernel applies the last pending
// step at the end of each chunk of iterations.  Mixed precision: FP32 inner
// solves inside FP64 iterative refinement with the FP64 true residual.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <vector>

#include "pb_common.hpp"

namespace {

constexpr int X3 = 64, Y3 = 8, Z3 = 16;
constexpr int NT3 = X3 * Y3;   // threads of a marching block
constexpr int kNV = 5;   // max partials per block

enum : int { S3_RUN = 0, S3_CONV = 1, S3_MAXIT = 3, S3_BRK_RHO = -10, S3_BRK_OMEGA = -11,
             S3_P_CONV = 10, S3_P_HALF = 11, S3_P_MAXIT = 12, S3_P_BRK_RHO = 13, S3_P_BRK_OMEGA = 14 };

// solver state; pend: launch 2 ran and its five dots are not yet folded in
struct KS3 {
  double rho, alpha, omega, beta, atol, bnrm, rnorm, tol;
  int iter, maxiter, status, legacy, pend, pad;
};

template <typename T>
struct Op3 {
  const T* h;
  const T* hc;
  const T* dinv;
  int nx, ny, nz, px, py, pz, center;
  T sx, sy, sz;
};

__device__ __forceinline__ int wrap3(int i, int n, int per) {
  if (i < 0) return per ? n - 2 : 0;
  if (i >= n) return (i == n && per) ? 1 : n - 1;
  return i;
}

// Tile of hardware block h in a march launch of gx * gy * gz blocks, in the
// XCD-aware order: block h runs on XCD h mod 8, and each XCD takes a
// contiguous band of (x tile, y tile, z chunk) tiles so that halo planes and
// rows it shares stay in its L2.  Bijective on [0, gx gy gz).
struct MarchTile {
  int bx, by, bz;
};
__host__ __device__ inline MarchTile march_tile(int h, int gx, int gy, int gz) {
  const int nbk = gx * gy * gz, xx = h % 8, kk8 = h / 8, base = nbk / 8, rem = nbk % 8;
  const int lin = xx < rem ? xx * (base + 1) + kk8 : rem * (base + 1) + (xx - rem) * base + kk8;
  const int byz = lin / gx;
  return MarchTile{lin % gx, byz % gy, byz / gy};
}

__device__ __forceinline__ int tid3() { return threadIdx.x + threadIdx.y * X3; }
__device__ __forceinline__ int bid3() { return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z); }
__device__ __forceinline__ int nblk3() { return gridDim.x * gridDim.y * gridDim.z; }
__device__ __forceinline__ double eps2() { return 4.930380657631324e-32; }

// The march: ``in(g)`` gives the operand at node g (may combine vectors),
// ``pre(g)`` loads the values ``out`` needs at node g, and
// ``out(g, y, pc, e, acc)`` consumes y = (A operand)(g), the operand's own
// value pc and e = pre(g) for the owned nodes; acc collects per-thread partial
// sums.  The next node plane (operand values) is always loaded one step
// ahead.  PF sets where the other loads of a step are issued:
//   2 (FP32 default): one step ahead too -- the next cell plane's weights and
//     the next node's diag_inv, hcenter and pre(g);
//   1 (FP64 default): at the start of their own step, in flight during its
//     barrier (one step ahead would cost the registers of a third block per
//     CU), pre(g) where it is used;
//   0: where they are used (kernels whose accumulators leave no registers).
//   UNR (with PF 2): the march unrolled by two over two register sets, so the
//     loads issued one step ahead are never copied between sets (a copy
//     waits for them at the end of the step that issued them); it costs
//     registers, and pays in the five-dot launch only.
//   GUARD: plane and weight loads only by the threads that hold a value.
//     Otherwise (unless PF 2 without UNR) every thread loads, with clamped
//     indices, so no load sits behind a branch: faster for the matvec and
//     the five-dot launch, slower for launch 1 (measured).
//   ZC: z planes per block (the launch's grid z is ceil(nz / ZC)).
//   RAW: ``in`` is a struct with a Raw type, load(g) -> Raw and
//     combine(Raw) -> operand.  The plane fetch loads the raw values, the
//     combined operand goes to the node plane, and the raw values of the
//     tile's own nodes stay in LDS (one slot per node-plane slot), where
//     ``out(g, y, pc, e, raw, acc)`` finds them: a launch that needs a node's
//     inputs both combined (for the stencil) and raw (for its updates) reads
//     them from memory once.
template <class In, bool RAW, typename T>
struct FetchOf {
  using type = T;
};
template <class In, typename T>
struct FetchOf<In, true, T> {
  using type = typename In::Raw;
};

template <typename T, int NV, int PF = (sizeof(T) == 4 ? 2 : 1), bool UNR = false, bool GUARD = false,
          int ZC = Z3, bool RAW = false, class In, class PreFn, class Out>
__device__ __forceinline__ void march3(const Op3<T>& o, In in, PreFn pre, Out out,
                                       double (&acc)[NV > 0 ? NV : 1]) {
  using PreT = decltype(pre(0L));
  using FT = typename FetchOf<In, RAW, T>::type;
  constexpr bool AHEAD = PF == 2;   // node operands one step ahead
  constexpr bool UNCOND = !GUARD && (!AHEAD || UNR);   // branch-free plane and weight loads
  // three node-plane slots: plane k+1 is stored while threads that are still
  // in step k-1's output stage read plane k-1, so it needs a slot of its own
  __shared__ T P[3][Y3 + 2][X3 + 2];
  __shared__ T F[3][Y3 + 1][X3 + 1];
  __shared__ FT RS[RAW ? 3 : 1][RAW ? Y3 : 1][RAW ? X3 : 1];   // own nodes' raw inputs (RAW)
  const int tx = threadIdx.x, ty = threadIdx.y, tid = tid3();
  const MarchTile mt = march_tile(bid3(), gridDim.x, gridDim.y, gridDim.z);
  const int bzz = mt.bz;
  const int x0 = mt.bx * X3, y0 = mt.by * Y3;
  const int kb = bzz * ZC, ke = min(kb + ZC, o.nz);
  // A block whose chunk starts past the last plane (a grid with more z
  // blocks than ceil(nz / ZC)) loads nothing; its caller's sums stay zero.
  // Only a march whose chunk differs from the default grid's can meet one
  // (the check costs the five-dot march registers, so it is left out there).
  if constexpr (ZC != Z3) {
    if (kb >= o.nz) return;
  }
  const int nx = o.nx, ny = o.ny;
  const int i = x0 + tx, j = y0 + ty;
  const bool own = i < nx && j < ny;
  // a node plane with its halo is NQ values, at most QPT per thread; a cell
  // plane NC weights, at most CPT per thread
  constexpr int NQ = (Y3 + 2) * (X3 + 2), QPT = (NQ + NT3 - 1) / NT3;
  constexpr int NC = (Y3 + 1) * (X3 + 1), CPT = (NC + NT3 - 1) / NT3;
  // Global loads of plane k into registers.  With UNCOND every thread loads
  // QPT values (indices past the plane clamped to its last value, not
  // stored), so no load sits behind a branch and the wait counts stay
  // static; otherwise only the threads with a value load (fewer loads, which
  // the one-step-ahead FP32 marches prefer).
  auto fetch = [&](int k, FT (&r)[QPT]) {
    const long kk = wrap3(k, o.nz, o.pz);
#pragma unroll
    for (int e = 0; e < QPT; ++e) {
      const int q0 = tid + e * NT3;
      if (UNCOND || q0 < NQ) {
        const int q = min(q0, NQ - 1);
        const int jj = q / (X3 + 2), ii = q - jj * (X3 + 2);
        const int gj = wrap3(y0 - 1 + jj, ny, o.py), gi = wrap3(x0 - 1 + ii, nx, o.px);
        if constexpr (RAW) r[e] = in.load((kk * ny + gj) * nx + gi);
        else r[e] = in((kk * ny + gj) * nx + gi);
      }
    }
  };
  auto store = [&](int slot, const FT (&r)[QPT]) {
#pragma unroll
    for (int e = 0; e < QPT; ++e) {
      const int q = tid + e * NT3;
      if (q < NQ) {
        const int jj = q / (X3 + 2), ii = q % (X3 + 2);
        if constexpr (RAW) {
          P[slot][jj][ii] = in.combine(r[e]);
          if (jj >= 1 && jj <= Y3 && ii >= 1 && ii <= X3) RS[slot][jj - 1][ii - 1] = r[e];
        } else {
          P[slot][jj][ii] = r[e];
        }
      }
    }
  };
  auto load_plane = [&](int k, int slot) {
    FT r[QPT];
    fetch(k, r);
    store(slot, r);
  };
  auto fetch_cells = [&](int k, T (&hw)[CPT]) {   // weights of the cell plane between node planes k, k+1
    const long ck = min(k + 1, o.nz);   // cell planes 0..nz; clamped like the node planes
#pragma unroll
    for (int e = 0; e < CPT; ++e) {
      const int q = tid + e * NT3;
      const int cj = q / (X3 + 1), ci = q - cj * (X3 + 1);
      const int cy = y0 + cj, cx = x0 + ci;
      const bool ok = q < NC && cy <= ny && cx <= nx;
      if (UNCOND) {
        const T v = o.h[ok ? (ck * (ny + 1) + cy) * (nx + 1) + cx : 0];   // unconditional load
        hw[e] = ok ? v : T(0);
      } else {
        hw[e] = ok ? o.h[(ck * (ny + 1) + cy) * (nx + 1) + cx] : T(0);
      }
    }
  };
  // cell fluxes between the node planes in slots s0, s1 of cell plane k+1;
  // weights from hw, or (PF == 0) loaded here, each just before its use
  auto cells = [&](int k, const T (&hw)[CPT], int s0, int s1) {
#pragma unroll
    for (int e = 0; e < CPT; ++e) {
      const int q = tid + e * NT3;
      if (q < NC) {
        const int cj = q / (X3 + 1), ci = q - cj * (X3 + 1);
        T hh = hw[e];
        if (PF == 0) {
          const int cy = y0 + cj, cx = x0 + ci;
          hh = (cy <= ny && cx <= nx) ? o.h[((long)min(k + 1, o.nz) * (ny + 1) + cy) * (nx + 1) + cx] : T(0);
        }
        const T a00 = P[s0][cj][ci], a01 = P[s0][cj][ci + 1], a10 = P[s0][cj + 1][ci], a11 = P[s0][cj + 1][ci + 1];
        const T b00 = P[s1][cj][ci], b01 = P[s1][cj][ci + 1], b10 = P[s1][cj + 1][ci], b11 = P[s1][cj + 1][ci + 1];
        const T gx = ((a01 - a00) + (a11 - a10)) + ((b01 - b00) + (b11 - b10));
        const T gy = ((a10 - a00) + (a11 - a01)) + ((b10 - b00) + (b11 - b01));
        const T gz = ((b00 - a00) + (b01 - a01)) + ((b10 - a10) + (b11 - a11));
        F[0][cj][ci] = hh * gx * o.sx;
        F[1][cj][ci] = hh * gy * o.sy;
        F[2][cj][ci] = hh * gz * o.sz;
      }
    }
  };
  struct NodeOps {
    T dinv, hc;
    PreT e;
  };
  auto fetch_node = [&](int k, NodeOps& nd) {   // this thread's node of plane k (node 0 when not owned)
    const long g = own ? ((long)k * ny + j) * nx + i : 0;
    nd.dinv = o.dinv[g];
    nd.hc = o.center ? o.hc[g] : T(0);
    if (AHEAD) nd.e = pre(g);
  };
  auto dz = [&](T& D, T& Z) {
    D = ((F[0][ty][tx + 1] + F[0][ty + 1][tx + 1]) - (F[0][ty][tx] + F[0][ty + 1][tx])) +
        ((F[1][ty + 1][tx] + F[1][ty + 1][tx + 1]) - (F[1][ty][tx] + F[1][ty][tx + 1]));
    Z = (F[2][ty][tx] + F[2][ty][tx + 1]) + (F[2][ty + 1][tx] + F[2][ty + 1][tx + 1]);
  };
  load_plane(kb - 1, 0);
  load_plane(kb, 1);
  FT nxt[QPT];
  fetch(kb + 1, nxt);       // plane kb+1 in flight during the first cell plane
  // two register sets for a step's cell weights and node operands, used in
  // turn (the march is unrolled by two), so a set loaded one step ahead is
  // never copied: a copy would wait for the loads at the end of the step
  // that issued them
  T h0[CPT], hwA[CPT], hwB[CPT];
#pragma unroll
  for (int e = 0; e < CPT; ++e) hwA[e] = hwB[e] = T(0);
  if (PF != 0) fetch_cells(kb - 1, h0);
  NodeOps ndA, ndB;
  if (AHEAD) {
    fetch_cells(kb, hwA);   // cell plane of step kb
    fetch_node(kb, ndA);    // node operands of step kb
  }
  __syncthreads();
  cells(kb - 1, h0, 0, 1);
  __syncthreads();
  T Dp, Zp;
  dz(Dp, Zp);
  // slots: plane kb-1 in 0, kb in 1, kb+1 in 2, kb+2 in 0, ...  No barrier
  // before the store: its slot held plane k-2, whose last reads (step k-2)
  // precede step k-1's barriers, and F's last reads (step k-1's dz) precede
  // this step's first barrier, which orders them before cells(k) writes F
  int sk = 1, sn = 2;
  auto step = [&](int k, T (&hw)[CPT], NodeOps& nd, T (&hwn)[CPT], NodeOps& ndn) {
    store(sn, nxt);           // plane k+1
    if (k + 1 < ke) {         // step k+1's loads in flight during this step's compute
      fetch(k + 2, nxt);
      if (AHEAD) {
        fetch_cells(k + 1, hwn);
        fetch_node(k + 1, ndn);
      }
    }
    if (PF == 1) {            // this step's own loads, in flight during the barrier
      fetch_cells(k, hw);
      fetch_node(k, nd);
    }
    __syncthreads();
    cells(k, hw, sk, sn);
    __syncthreads();
    T Dn, Zn;
    dz(Dn, Zn);
    if (own) {
      const long g = ((long)k * ny + j) * nx + i;
      const T pc = P[sk][ty + 1][tx + 1];
      T y = (Dp + Dn) + (Zn - Zp);
      if constexpr (RAW) {
        static_assert(!RAW || PF == 2, "RAW marches load their node operands one step ahead");
        if (o.center) y += nd.hc * pc;
        y *= nd.dinv;
        out(g, y, pc, nd.e, RS[sk][ty][tx], acc);
      } else if (PF == 0) {
        if (o.center) y += o.hc[g] * pc;
        y *= o.dinv[g];
        out(g, y, pc, pre(g), acc);
      } else {
        if (o.center) y += nd.hc * pc;
        y *= nd.dinv;
        if (AHEAD) out(g, y, pc, nd.e, acc);
        else out(g, y, pc, pre(g), acc);
      }
    }
    Dp = Dn;
    Zp = Zn;
    sk = sn;
    sn = sn == 2 ? 0 : sn + 1;
  };
  if constexpr (AHEAD && UNR) {
    for (int k = kb; k < ke; k += 2) {
      step(k, hwA, ndA, hwB, ndB);
      if (k + 1 < ke) step(k + 1, hwB, ndB, hwA, ndA);
    }
  } else if constexpr (AHEAD) {   // set B (next step's) is copied into set A
    for (int k = kb; k < ke; ++k) {
      step(k, hwA, ndA, hwB, ndB);
#pragma unroll
      for (int e = 0; e < CPT; ++e) hwA[e] = hwB[e];
      ndA = ndB;
    }
  } else {                    // a step's loads are its own: one set suffices
    for (int k = kb; k < ke; ++k) step(k, hwA, ndA, hwB, ndB);
  }
}

// NW: waves per block (8 for the marching blocks of NT3 threads)
template <int NV, int NW = NT3 / kWave>
__device__ __forceinline__ void write_partials3(double (&acc)[NV], double* part) {
  __shared__ double sh[NV * NW];
  block_sum_w<NV, NW>(acc, sh);
  if (tid3() == 0)
#pragma unroll
    for (int v = 0; v < NV; ++v) part[(long)v * nblk3() + bid3()] = acc[v];
}

// threads of the one-block scalar kernels: one record per thread up to 1024 blocks
constexpr int NRT = 1024;

// One-block reduction of the partial records (value v of block b at
// part[v * nb + b], so a wave's loads are contiguous) (NRT threads): thread t
// adds records t, t + NRT, ... in order, then the block sum takes the wave
// sums and the per-wave values in wave order.  The first record is loaded
// unconditionally (clamped to the last one) and zeroed after, so its load is
// not held behind a branch; first_record issues it, sum_records finishes.
// The result is valid in thread 0.
template <int NV>
__device__ __forceinline__ void first_record(const double* part, int nb, double (&out)[NV]) {
  const int b0 = min((int)threadIdx.x, nb - 1);
#pragma unroll
  for (int v = 0; v < NV; ++v) out[v] = part[(long)v * nb + b0];
}

template <int NV>
__device__ __forceinline__ void sum_records(const double* part, int nb, double (&out)[NV]) {
  __shared__ double sh[NV * (NRT / kWave)];
  const int t = threadIdx.x;
  if (t >= nb)
#pragma unroll
    for (int v = 0; v < NV; ++v) out[v] = 0.0;
  for (int b = t + NRT; b < nb; b += NRT)
#pragma unroll
    for (int v = 0; v < NV; ++v) out[v] += part[(long)v * nb + b];
  block_sum_w<NV, NRT / kWave>(out, sh);
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(NT3) void k3_matvec(Op3<T> o, const T* p, T* y) {
  double acc[1] = {0.0};
  march3<T, 0>(
      o, [&](long g) { return p[g]; }, [&](long) { return 0; },
      [&](long g, T v, T, int, double(&)[1]) { y[g] = v; }, acc);
}

// r = b - A x (FP64), r32 = (float) r, partials of r.r
__global__ __launch_bounds__(NT3) void k3_resid(Op3<double> o, const double* x, const double* b, double* r,
                                                float* r32, double* part) {
  double acc[1] = {0.0};
  march3<double, 1>(o, [&](long g) { return x[g]; }, [&](long g) { return b[g]; },
                    [&](long g, double y, double, double bg, double(&a)[1]) {
                      const double rr = bg - y;
                      r[g] = rr;
                      if (r32) r32[g] = (float)rr;
                      a[0] += rr * rr;
                    },
                    acc);
  write_partials3<1>(acc, part);
}

// init: r = b, rt = b, p = v = t = 0, x = 0; partials of b.b
template <typename T>
__global__ void k3_init(long n, const T* b, T* r, T* rt, T* p, T* v, T* t, T* x, double* part) {
  double acc[1] = {0.0};
  for (long g = (long)bid3() * 256 + tid3(); g < n; g += (long)nblk3() * 256) {
    const T bb = b[g];
    r[g] = bb;
    rt[g] = bb;
    p[g] = 0;
    v[g] = 0;
    t[g] = 0;
    x[g] = 0;
    acc[0] += (double)bb * (double)bb;
  }
  write_partials3<1, 4>(acc, part);   // blocks of 256 threads
}

__global__ __launch_bounds__(NRT) void k3_s0(KS3* st, const double* part, int nb, double tol, int maxiter,
                                             int legacy) {
  double s[1];
  first_record<1>(part, nb, s);
  sum_records<1>(part, nb, s);
  if (threadIdx.x == 0) {
    KS3 k{};
    k.rho = s[0];
    k.bnrm = sqrt(s[0]);
    k.tol = tol;
    k.atol = tol * k.bnrm;
    k.maxiter = maxiter;
    k.legacy = legacy;
    k.rnorm = k.bnrm;
    k.alpha = k.omega = k.beta = 0.0;
    k.iter = 0;
    // legacy: return x0 = 0 at once if ||b - A x0|| <= tol (absolute); else b = 0
    k.status = ((legacy && k.bnrm <= tol) || k.bnrm == 0.0) ? S3_CONV : S3_RUN;
    *st = k;
  }
}

// field-wise copies of the state (a whole-struct copy goes through scratch)
__device__ __forceinline__ KS3 load_ks3(const KS3* S) {
  KS3 k;
  k.rho = S->rho;
  k.alpha = S->alpha;
  k.omega = S->omega;
  k.beta = S->beta;
  k.atol = S->atol;
  k.bnrm = S->bnrm;
  k.rnorm = S->rnorm;
  k.tol = S->tol;
  k.iter = S->iter;
  k.maxiter = S->maxiter;
  k.status = S->status;
  k.legacy = S->legacy;
  k.pend = S->pend;
  k.pad = 0;
  return k;
}

__device__ __forceinline__ void store_ks3(KS3* S, const KS3& k) {
  S->rho = k.rho;
  S->alpha = k.alpha;
  S->omega = k.omega;
  S->beta = k.beta;
  S->atol = k.atol;
  S->bnrm = k.bnrm;
  S->rnorm = k.rnorm;
  S->tol = k.tol;
  S->iter = k.iter;
  S->maxiter = k.maxiter;
  S->status = k.status;
  S->legacy = k.legacy;
  S->pend = k.pend;
  S->pad = 0;
}

// The scalar steps between the stencil launches, applied to a state held in
// registers: fold1 after launch 1 (alpha from r~.v), fold2 after launch 2
// (omega, the residual norm and the stopping tests, beta and rho from the
// five dots s.s, t.s, t.t, r~.s, r~.t).
__device__ __forceinline__ void fold1(KS3& k, double rtv) {
  if (rtv == 0.0 || !isfinite(rtv)) k.status = S3_BRK_OMEGA;
  else k.alpha = k.rho / rtv;
}

__device__ __forceinline__ void fold2(KS3& k, const double (&s)[5]) {
  const double ss = s[0], ts = s[1], tt = s[2], rts = s[3], rtt = s[4];
  if (sqrt(ss) <= k.atol) {
    k.status = S3_P_HALF;     // x += alpha p
    k.rnorm = sqrt(ss);
  } else if (tt == 0.0 || !isfinite(tt)) {
    k.status = S3_BRK_OMEGA;
  } else {
    const double om = ts / tt;
    const double rho_new = rts - om * rtt;
    const double rr = fmax(ss - 2.0 * om * ts + om * om * tt, 0.0);
    k.omega = om;
    k.iter += 1;
    k.rnorm = sqrt(rr);
    if (k.rnorm <= k.atol) k.status = S3_P_CONV;
    else if (fabs(om) < eps2()) k.status = S3_P_BRK_OMEGA;
    else if (k.iter >= k.maxiter) k.status = S3_P_MAXIT;
    else if (fabs(rho_new) < eps2()) k.status = S3_P_BRK_RHO;
    else {
      k.beta = (rho_new / k.rho) * (k.alpha / om);
      k.rho = rho_new;
    }
  }
}

// Every block of a marching launch reduces the previous launch's partial
// records (value q of block b at P[q * nb + b]) itself, all in the same
// fixed order (thread t adds records t, t + NT3, ...; wave sums; the NT3/64
// wave values in order), so every block holds bitwise the same sums.  A
// thread's first two records are loaded at kernel entry (indices clamped,
// loads unconditional).
template <int NV>
struct Rec3 {
  double v[2][NV];
  __device__ __forceinline__ void load(const double* P, int nb) {
    const int t = tid3();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int b = min(t + h * NT3, nb - 1);
#pragma unroll
      for (int q = 0; q < NV; ++q) v[h][q] = P[(long)q * nb + b];
    }
  }
};

template <int NV>
__device__ __forceinline__ void sum_rec3(const Rec3<NV>& e, const double* P, int nb, double (&tot)[NV]) {
  constexpr int NW = NT3 / kWave;
  __shared__ double sh[NV * NW];
  const int t = tid3();
  double s[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) s[q] = 0.0;
#pragma unroll
  for (int h = 0; h < 2; ++h)
    if (t + h * NT3 < nb)
#pragma unroll
      for (int q = 0; q < NV; ++q) s[q] += e.v[h][q];
  for (int b = t + 2 * NT3; b < nb; b += NT3)
#pragma unroll
    for (int q = 0; q < NV; ++q) s[q] += P[(long)q * nb + b];
  wave_sum_n<NV>(s);
  if ((t & (kWave - 1)) == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) sh[q * NW + t / kWave] = s[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double w = 0.0;
#pragma unroll
    for (int i = 0; i < NW; ++i) w += sh[q * NW + i];
    tot[q] = w;
  }
}

// z planes per block of launch 1: 24 in both precisions (FP64: the grid
// fits one round of blocks; FP32: 48.7 against 51.3 us with 16-plane chunks
// at 257x129x129, the extra plane per chunk being the smaller share)
template <typename T>
constexpr int kZ1 = 24;

// launch 1's node operand: r, p, v, t raw, p_k = (r - alpha v - omega t) +
// beta (p - omega v) combined
template <typename T>
struct Iter1In {
  struct Raw {
    T v, r, p, t;
  };
  const T *va, *ra, *pa, *t;
  T al, om, be;
  __device__ __forceinline__ Raw load(long g) const { return Raw{va[g], ra[g], pa[g], t[g]}; }
  __device__ __forceinline__ T combine(const Raw& q) const {
    const T rn = (q.r - al * q.v) - om * q.t;
    return rn + be * (q.p - om * q.v);
  }
};

// launch 1 of iteration k: first the scalar step of iteration k-1's launch 2
// (its five dots, when pending), then r_{k-1} = r - alpha v - omega t,
// x += alpha p + omega (r - alpha v), p_k = r_{k-1} + beta (p - omega v),
// v_k = A p_k; partials of rt.v_k into part1.  Reads vector set a, writes set
// b; reads state Sin, block 0 writes the advanced state to Sout.
template <typename T>
__global__ __launch_bounds__(NT3) void k3_iter1(Op3<T> o, const KS3* Sin, KS3* Sout, const double* part5, int nb,
                                                const T* ra, const T* pa, const T* va, const T* t, const T* rt,
                                                T* rb, T* pb, T* vb, T* x, double* part1) {
  Rec3<5> rec;
  rec.load(part5, nb);
  KS3 k = load_ks3(Sin);
  if (k.pend) {
    double s[5];
    sum_rec3<5>(rec, part5, nb, s);
    if (k.status == S3_RUN) fold2(k, s);
    k.pend = 0;
  }
  if (bid3() == 0 && tid3() == 0) store_ks3(Sout, k);
  if (k.status != S3_RUN) return;
  const T al = (T)k.alpha, om = (T)k.omega, be = (T)k.beta;
  double acc[1] = {0.0};
  struct Ops {   // this node's r~ and x, loaded a plane step ahead
    T rt, x;
  };
  // The node plane's r, p, v, t are read once: the operand p_k is combined
  // from them for the stencil, and the tile's own nodes keep them in LDS for
  // the updates.  Node operands one step ahead, over 24-plane chunks.
  march3<T, 1, 2, false, true, kZ1<T>, true>(
      o, Iter1In<T>{va, ra, pa, t, al, om, be}, [&](long g) { return Ops{rt[g], x[g]}; },
      [&](long g, T y, T pn, const Ops& e, const typename Iter1In<T>::Raw& w, double(&a)[1]) {
        const T s = w.r - al * w.v;
        x[g] = e.x + (al * w.p + om * s);
        rb[g] = s - om * w.t;
        pb[g] = pn;
        vb[g] = y;
        a[0] += (double)e.rt * (double)y;
      },
      acc);
  write_partials3<1>(acc, part1);
}

// launch 2: first the scalar step of launch 1 (alpha from the r~.v partials),
// then s = r - alpha v (on the fly), t = A s; partials s.s, t.s, t.t, rt.s,
// rt.t into part5.  Reads state Sin, block 0 writes the advanced state to
// Sout.  Six waves per SIMD keep three marching blocks on a CU (all 765
// blocks of a 257x129x129 launch resident at once): at most 80 VGPRs.
template <typename T>
__global__ __launch_bounds__(NT3) __attribute__((amdgpu_waves_per_eu(6))) void k3_iter2(
    Op3<T> o, const KS3* Sin, KS3* Sout, const double* part1, int nb, const T* r, const T* v, const T* rt, T* t,
    double* part5) {
  Rec3<1> rec;
  rec.load(part1, nb);
  KS3 k = load_ks3(Sin);
  if (k.status == S3_RUN) {
    double s[1];
    sum_rec3<1>(rec, part1, nb, s);
    fold1(k, s[0]);
  }
  k.pend = k.status == S3_RUN ? 1 : 0;
  if (bid3() == 0 && tid3() == 0) store_ks3(Sout, k);
  if (k.status != S3_RUN) return;
  const T al = (T)k.alpha;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  // FP64: five FP64 accumulators; its step loads stay where they are used
  march3<T, 5, sizeof(T) == 4 ? 2 : 0, true>(
      o, [&](long g) { return r[g] - al * v[g]; }, [&](long g) { return rt[g]; },
      [&](long g, T y, T s, T rtg, double(&a)[5]) {
        t[g] = y;
        const double sd = s, td = y, rd = rtg;
        a[0] += sd * sd;
        a[1] += td * sd;
        a[2] += td * td;
        a[3] += rd * sd;
        a[4] += rd * td;
      },
      acc);
  write_partials3<5>(acc, part5);
}

// the pending launch-2 step at the end of a chunk of iterations, so the
// state the host reads is final (one block, in place)
__global__ __launch_bounds__(NRT) void k3_fold(KS3* st, const double* part5, int nb) {
  double s[5];
  first_record<5>(part5, nb, s);
  KS3 k = load_ks3(st);
  if (!k.pend) return;
  sum_records<5>(part5, nb, s);
  if (threadIdx.x == 0) {
    if (k.status == S3_RUN) fold2(k, s);
    k.pend = 0;
    store_ks3(st, k);
  }
}

// pending exit: x += alpha p + omega (r - alpha v)  (omega term dropped for a half step)
template <typename T>
__global__ void k3_finish(long n, const KS3* st, const T* r, const T* p, const T* v, T* x) {
  const T al = (T)st->alpha, om = st->status == S3_P_HALF ? (T)0 : (T)st->omega;
  for (long g = (long)bid3() * 256 + tid3(); g < n; g += (long)nblk3() * 256) x[g] += al * p[g] + om * (r[g] - al * v[g]);
}

template <typename T>
__global__ void k3_cast(long n, const T* a, double* b, int add) {
  for (long g = (long)bid3() * 256 + tid3(); g < n; g += (long)nblk3() * 256) b[g] = add ? b[g] + (double)a[g] : (double)a[g];
}

// assembly -------------------------------------------------------------------
// cell weights h = Gammainv rhoY (rhoY/rho), zeroed on the outside layers of wall axes
__global__ void k3_asm_cells(int nx, int ny, int nz, int px, int py, int pz, double Gammainv, const double* rho,
                             const double* rhoY, double* h64, float* h32) {
  const long nc = (long)(nx + 1) * (ny + 1) * (nz + 1);
  for (long c = (long)bid3() * 256 + tid3(); c < nc; c += (long)nblk3() * 256) {
    const int cx = (int)(c % (nx + 1));
    const long t = c / (nx + 1);
    const int cy = (int)(t % (ny + 1)), cz = (int)(t / (ny + 1));
    const double Y = rhoY[c] / rho[c];
    double h = Gammainv * rhoY[c] * Y;
    if ((!px && (cx == 0 || cx == nx)) || (!py && (cy == 0 || cy == ny)) || (!pz && (cz == 0 || cz == nz))) h = 0.0;
    h64[c] = h;
    h32[c] = (float)h;
  }
}

// node terms: wcenter (2x2x2 mean of rhoY^(2-gamma), wall planes halved on
// inner nodes of the other axes, regime scale), diag_inv from the unzeroed
// box sum of the cell weights (laplacian.py:1284-1356, 3D coefficient 1/16)
__global__ void k3_asm_nodes(int nx, int ny, int nz, int px, int py, int pz, double Gammainv, double gamm,
                             double ccenter, double regime, double dx, double dy, double dz, const double* rho,
                             const double* rhoY, double* hc64, float* hc32, double* dinv64, float* dinv32,
                             double* wc_out, double* di_out) {
  const long n = (long)nx * ny * nz;
  for (long g = (long)bid3() * 256 + tid3(); g < n; g += (long)nblk3() * 256) {
    const int i = (int)(g % nx);
    const long t = g / nx;
    const int j = (int)(t % ny), k = (int)(t / ny);
    double br = 0.0, bh = 0.0;
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int c = 0; c < 2; ++c) {
          const long cc = ((long)(k + a) * (ny + 1) + (j + b)) * (nx + 1) + (i + c);
          br += pow(rhoY[cc], 2.0 - gamm);
          bh += Gammainv * rhoY[cc] * (rhoY[cc] / rho[cc]);
        }
    double wc = ccenter * br / 8.0;
    const bool ix = i > 0 && i < nx - 1, iy = j > 0 && j < ny - 1, iz = k > 0 && k < nz - 1;
    if (!px && !ix && iy && iz) wc *= 0.5;
    if (!py && !iy && ix && iz) wc *= 0.5;
    if (!pz && !iz && ix && iy) wc *= 0.5;
    wc *= regime;
    const double diag = wc - 0.0625 * bh * (1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz) + 2.0 / (dx * dy));
    const double di = 1.0 / diag;
    hc64[g] = wc;
    hc32[g] = (float)wc;
    dinv64[g] = di;
    dinv32[g] = (float)di;
    if (wc_out) wc_out[g] = wc;
    if (di_out) di_out[g] = di;
  }
}

__global__ void k3_cast_down(long n, const double* a, float* b) {
  for (long g = (long)bid3() * 256 + tid3(); g < n; g += (long)nblk3() * 256) b[g] = (float)a[g];
}

__global__ void k3_fill(long n, double* a, double v) {
  for (long g = (long)bid3() * 256 + tid3(); g < n; g += (long)nblk3() * 256) a[g] = v;
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct pb_op3d {
  pb_ctx* ctx = nullptr;
  pb_grid3d g{};
  long n = 0, nc = 0;
  bool assembled = false, center = true;
  double *h64 = nullptr, *hc64 = nullptr, *dinv64 = nullptr;
  float *h32 = nullptr, *hc32 = nullptr, *dinv32 = nullptr;
  // Krylov sets [precision][set]: r, p, v; shared t, rt, x, b
  void* r[2][2] = {};
  void* p[2][2] = {};
  void* v[2][2] = {};
  void* t[2] = {};
  void* rt[2] = {};
  void* x[2] = {};
  void* b32 = nullptr;
  double* r64 = nullptr;
  double* part = nullptr;    // block partials: launch 2's five dots, residual and init sums
  double* part1 = nullptr;   // block partials of launch 1 (r~.v)
  KS3* st = nullptr;         // two state slots: launch 1 advances st[0] into st[1], launch 2 back
  KS3* hst = nullptr;
  bool have_work = false;
  std::vector<void*> allocs;
  std::map<int, hipGraphExec_t> graphs;   // captured iteration chunks, by (precision, chunk, centre flag)
};

namespace {

template <typename T>
int dalloc3(pb_op3d* op, T** p, long count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, sizeof(T) * (size_t)std::max<long>(count, 1));
  if (e != hipSuccess)
    return pb_fail(op->ctx, PB_ERR_HIP, "hipMalloc(%ld bytes) failed: %s", (long)(sizeof(T) * count),
                   hipGetErrorString(e));
  op->allocs.push_back(q);
  *p = (T*)q;
  return PB_OK;
}

dim3 march_grid(const pb_op3d* op, int zc = Z3) {
  return dim3(ceil_div(op->g.nx, X3), ceil_div(op->g.ny, Y3), ceil_div(op->g.nz, zc));
}
int march_blocks(const pb_op3d* op, int zc = Z3) {
  const dim3 gr = march_grid(op, zc);
  return (int)(gr.x * gr.y * gr.z);
}
// A march launch must have exactly the tiles of its kernel's chunk of zc
// planes: fewer z blocks would leave planes out, more would start chunks
// past the last plane (whose loads the march does not guard).
int march_check(const pb_op3d* op, const dim3& grd, int zc) {
  const dim3 want = march_grid(op, zc);
  if (grd.x != want.x || grd.y != want.y || grd.z != want.z)
    return pb_fail(op->ctx, PB_ERR_ARG, "3D march: grid (%u, %u, %u) does not tile %d x %d x %d nodes in %d-plane chunks",
                   grd.x, grd.y, grd.z, op->g.nx, op->g.ny, op->g.nz, zc);
  return PB_OK;
}
dim3 elem_grid3(long n) { return dim3((unsigned)std::min<long>((n + 255) / 256, 4096)); }
const dim3 kBlk3(X3, Y3);

template <typename T>
Op3<T> op3(const pb_op3d* op);
template <>
Op3<double> op3<double>(const pb_op3d* op) {
  const pb_grid3d& g = op->g;
  return Op3<double>{op->h64, op->hc64, op->dinv64, g.nx, g.ny, g.nz,
                     g.bc_x == PB_BC_PERIODIC, g.bc_y == PB_BC_PERIODIC, g.bc_z == PB_BC_PERIODIC, op->center ? 1 : 0,
                     0.0625 / (g.dx * g.dx), 0.0625 / (g.dy * g.dy), 0.0625 / (g.dz * g.dz)};
}
template <>
Op3<float> op3<float>(const pb_op3d* op) {
  const pb_grid3d& g = op->g;
  return Op3<float>{op->h32, op->hc32, op->dinv32, g.nx, g.ny, g.nz,
                    g.bc_x == PB_BC_PERIODIC, g.bc_y == PB_BC_PERIODIC, g.bc_z == PB_BC_PERIODIC, op->center ? 1 : 0,
                    (float)(0.0625 / (g.dx * g.dx)), (float)(0.0625 / (g.dy * g.dy)), (float)(0.0625 / (g.dz * g.dz))};
}

int alloc_work3(pb_op3d* op) {
  if (op->have_work) return PB_OK;
  const long n = op->n;
  int rc = 0;
  for (int s = 0; s < 2; ++s) {
    rc |= dalloc3(op, (double**)&op->r[0][s], n);
    rc |= dalloc3(op, (double**)&op->p[0][s], n);
    rc |= dalloc3(op, (double**)&op->v[0][s], n);
    rc |= dalloc3(op, (float**)&op->r[1][s], n);
    rc |= dalloc3(op, (float**)&op->p[1][s], n);
    rc |= dalloc3(op, (float**)&op->v[1][s], n);
  }
  rc |= dalloc3(op, (double**)&op->t[0], n);
  rc |= dalloc3(op, (double**)&op->rt[0], n);
  rc |= dalloc3(op, (double**)&op->x[0], n);
  rc |= dalloc3(op, (float**)&op->t[1], n);
  rc |= dalloc3(op, (float**)&op->rt[1], n);
  rc |= dalloc3(op, (float**)&op->x[1], n);
  rc |= dalloc3(op, (float**)&op->b32, n);
  rc |= dalloc3(op, &op->r64, n);
  rc |= dalloc3(op, &op->part, (long)std::max(march_blocks(op), 4096) * kNV);
  rc |= dalloc3(op, &op->part1, (long)std::max(march_blocks(op), 4096));
  rc |= dalloc3(op, &op->st, 2);
  if (rc) return rc;
  void* hp = nullptr;
  PB_HIP(op->ctx, hipHostMalloc(&hp, sizeof(KS3), hipHostMallocDefault));
  op->hst = (KS3*)hp;
  op->have_work = true;
  return PB_OK;
}

template <typename T>
struct Prec;
template <>
struct Prec<double> {
  static constexpr int k = 0;
};
template <>
struct Prec<float> {
  static constexpr int k = 1;
};

// BiCGStab on (D^-1 A) x = b with x0 = 0; b, x are the workspace's own
// vectors of precision T.  Returns PB_OK and the final solver state.
template <typename T>
int krylov3(pb_op3d* op, const T* b, double tol, int maxiter, int legacy, int check_every, hipStream_t s,
            KS3* out) {
  const int P = Prec<T>::k;
  const long n = op->n;
  const Op3<T> o = op3<T>(op);
  const dim3 mg = march_grid(op);
  const int nbm = march_blocks(op);   // launch 2's blocks (five-dot records)
  const dim3 mg1 = march_grid(op, kZ1<T>);
  const int nb1 = march_blocks(op, kZ1<T>);   // launch 1's blocks (r~.v records)
  if (int rc = march_check(op, mg, Z3)) return rc;       // launch 2 marches Z3-plane chunks
  if (int rc = march_check(op, mg1, kZ1<T>)) return rc;  // launch 1 marches kZ1<T>-plane chunks
  const dim3 eg = elem_grid3(n);
  T* t = (T*)op->t[P];
  T* rt = (T*)op->rt[P];
  T* x = (T*)op->x[P];
  hipLaunchKernelGGL(k3_init<T>, eg, dim3(256), 0, s, n, b, (T*)op->r[P][0], rt, (T*)op->p[P][0], (T*)op->v[P][0], t,
                     x, op->part);
  hipLaunchKernelGGL(k3_s0, dim3(1), dim3(NRT), 0, s, op->st, (const double*)op->part, (int)eg.x, tol, maxiter,
                     legacy);
  PB_LAUNCH_CHECK(op->ctx);
  // an even chunk starts every replay at the same buffer parity, so one
  // captured graph per (precision, chunk) serves every chunk of every solve
  const int chunk = 2 * std::max(1, (check_every > 0 ? check_every : 32) / 2);
  // each launch applies the scalar step of the launch before it (the state
  // moves st[0] -> st[1] -> st[0]); the chunk ends with the pending step of
  // its last launch 2, so st[0] is final when the host reads it
  KS3* S0 = op->st;
  KS3* S1 = op->st + 1;
  auto launch_chunk = [&](long it0) {
    for (int c = 0; c < chunk; ++c) {
      const long it = it0 + c + 1;
      const int a = (int)((it - 1) & 1), bset = (int)(it & 1);
      hipLaunchKernelGGL(k3_iter1<T>, mg1, kBlk3, 0, s, o, (const KS3*)S0, S1, (const double*)op->part, nbm,
                         (const T*)op->r[P][a], (const T*)op->p[P][a], (const T*)op->v[P][a], (const T*)t,
                         (const T*)rt, (T*)op->r[P][bset], (T*)op->p[P][bset], (T*)op->v[P][bset], x, op->part1);
      hipLaunchKernelGGL(k3_iter2<T>, mg, kBlk3, 0, s, o, (const KS3*)S1, S0, (const double*)op->part1, nb1,
                         (const T*)op->r[P][bset], (const T*)op->v[P][bset], (const T*)rt, t, op->part);
    }
    hipLaunchKernelGGL(k3_fold, dim3(1), dim3(NRT), 0, s, S0, (const double*)op->part, nbm);
  };
  const char* no_graph = std::getenv("PYBELLA_AMD_NO_GRAPH");
  const bool graphs = !(no_graph && no_graph[0] == '1') && s != nullptr;
  long it = 0;
  for (;;) {
    if (graphs) {
      // the captured arguments hold the operator's centre flag, which a
      // re-assembly in the other regime changes
      const int key = (chunk * 2 + P) * 2 + (op->center ? 1 : 0);
      auto g = op->graphs.find(key);
      if (g == op->graphs.end()) {
        hipGraph_t gr = nullptr;
        hipGraphExec_t ex = nullptr;
        PB_HIP(op->ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        launch_chunk(0);
        PB_HIP(op->ctx, hipStreamEndCapture(s, &gr));
        PB_HIP(op->ctx, hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
        (void)hipGraphDestroy(gr);
        g = op->graphs.emplace(key, ex).first;
      }
      PB_HIP(op->ctx, hipGraphLaunch(g->second, s));
    } else {
      launch_chunk(it);
    }
    it += chunk;
    PB_LAUNCH_CHECK(op->ctx);
    PB_HIP(op->ctx, hipMemcpyAsync(op->hst, op->st, sizeof(KS3), hipMemcpyDeviceToHost, s));
    PB_HIP(op->ctx, hipStreamSynchronize(s));
    KS3 h = *op->hst;
    if (h.status == S3_RUN) continue;
    if (h.status >= S3_P_CONV) {   // the exit's x update with the vectors of iteration h.iter (or h.iter+1)
      // the vectors of the iteration that stopped: launch 1 of iteration k wrote set k & 1; a half-step
      // exit stops inside iteration iter + 1, the others after iteration iter
      const long k = h.status == S3_P_HALF ? (long)h.iter + 1 : (long)h.iter;
      const int set = (int)(k & 1);
      hipLaunchKernelGGL(k3_finish<T>, eg, dim3(256), 0, s, n, (const KS3*)op->st, (const T*)op->r[P][set],
                         (const T*)op->p[P][set], (const T*)op->v[P][set], x);
      PB_LAUNCH_CHECK(op->ctx);
      switch (h.status) {
        case S3_P_CONV:
        case S3_P_HALF: h.status = S3_CONV; break;
        case S3_P_MAXIT: h.status = S3_MAXIT; break;
        case S3_P_BRK_RHO: h.status = S3_BRK_RHO; break;
        default: h.status = S3_BRK_OMEGA; break;
      }
    }
    *out = h;
    return PB_OK;
  }
}

int info3(const KS3& h) {
  switch (h.status) {
    case S3_CONV: return 0;
    case S3_BRK_RHO: return -10;
    case S3_BRK_OMEGA: return -11;
    default: return std::max(h.iter, 1);
  }
}

// ||b - A x|| in FP64 into op->r64 (and op->b32 as FP32)
int resid3(pb_op3d* op, const double* x, const double* b, hipStream_t s, double* nrm) {
  const int nbm = march_blocks(op);
  if (int rc = march_check(op, march_grid(op), Z3)) return rc;
  hipLaunchKernelGGL(k3_resid, march_grid(op), kBlk3, 0, s, op3<double>(op), x, b, op->r64, (float*)op->b32,
                     op->part);
  PB_LAUNCH_CHECK(op->ctx);
  std::vector<double> h((size_t)nbm);
  PB_HIP(op->ctx, hipMemcpyAsync(h.data(), op->part, sizeof(double) * nbm, hipMemcpyDeviceToHost, s));
  PB_HIP(op->ctx, hipStreamSynchronize(s));
  double t = 0.0;
  for (double v : h) t += v;
  *nrm = std::sqrt(t);
  return PB_OK;
}

}  // namespace

extern "C" int pb_op3d_create(pb_ctx* ctx, const pb_grid3d* grid, pb_op3d** out) {
  if (!ctx || !grid || !out) return pb_fail(ctx, PB_ERR_ARG, "pb_op3d_create: null argument");
  const pb_grid3d& g = *grid;
  if (g.nx < 3 || g.ny < 3 || g.nz < 3 || !(g.dx > 0) || !(g.dy > 0) || !(g.dz > 0))
    return pb_fail(ctx, PB_ERR_ARG, "pb_op3d_create: bad grid %d x %d x %d", g.nx, g.ny, g.nz);
  pb_op3d* op = new pb_op3d();
  op->ctx = ctx;
  op->g = g;
  op->n = (long)g.nx * g.ny * g.nz;
  op->nc = (long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  int rc = dalloc3(op, &op->h64, op->nc);
  rc |= dalloc3(op, &op->h32, op->nc);
  rc |= dalloc3(op, &op->hc64, op->n);
  rc |= dalloc3(op, &op->hc32, op->n);
  rc |= dalloc3(op, &op->dinv64, op->n);
  rc |= dalloc3(op, &op->dinv32, op->n);
  if (rc) {
    pb_op3d_destroy(op);
    return PB_ERR_HIP;
  }
  *out = op;
  return PB_OK;
}

extern "C" int pb_op3d_destroy(pb_op3d* op) {
  if (!op) return PB_OK;
  if (op->ctx) (void)hipStreamSynchronize(op->ctx->stream);
  if (op->ctx) (void)hipStreamSynchronize(op->ctx->own_stream);
  for (auto& kv : op->graphs) (void)hipGraphExecDestroy(kv.second);
  for (void* p : op->allocs) (void)hipFree(p);
  if (op->hst) (void)hipHostFree(op->hst);
  delete op;
  return PB_OK;
}

extern "C" int pb_op3d_assemble(pb_op3d* op, const pb_phys* phys, double dt, const double* rho_c,
                                const double* rhoY_c, double* wcenter, double* diag_inv) {
  if (!op || !phys || !rho_c || !rhoY_c || !(dt > 0))
    return pb_fail(op ? op->ctx : nullptr, PB_ERR_ARG, "pb_op3d_assemble: bad argument");
  const pb_grid3d& g = op->g;
  hipStream_t s = op->ctx->stream;
  const double gm = phys->gamm, Gammainv = gm / (gm - 1.0), gm1inv = 1.0 / (gm - 1.0);
  const double ccenter = -phys->Msq * gm1inv / (dt * dt);
  // regime (second_projection.py:139-158): compressible -> 1, pseudo-incompressible -> compressibility
  const double regime = phys->is_compressible ? 1.0 : (phys->is_arakawa_konor ? 0.0 : phys->compressibility);
  op->center = regime != 0.0;
  const int px = g.bc_x == PB_BC_PERIODIC, py = g.bc_y == PB_BC_PERIODIC, pz = g.bc_z == PB_BC_PERIODIC;
  hipLaunchKernelGGL(k3_asm_cells, elem_grid3(op->nc), dim3(256), 0, s, g.nx, g.ny, g.nz, px, py, pz, Gammainv, rho_c,
                     rhoY_c, op->h64, op->h32);
  hipLaunchKernelGGL(k3_asm_nodes, elem_grid3(op->n), dim3(256), 0, s, g.nx, g.ny, g.nz, px, py, pz, Gammainv, gm,
                     ccenter, regime, g.dx, g.dy, g.dz, rho_c, rhoY_c, op->hc64, op->hc32, op->dinv64, op->dinv32,
                     wcenter, diag_inv);
  PB_LAUNCH_CHECK(op->ctx);
  op->assembled = true;
  return PB_OK;
}

// The tile (x tile, y tile, z chunk) every hardware block of a march launch
// of gx * gy * gz blocks processes: out[3 h .. 3 h + 2].  Host only (no GPU
// call); it runs the kernels' own mapping, for checking it.
extern "C" int pb_op3d_march_map(int gx, int gy, int gz, int* out) {
  if (gx < 1 || gy < 1 || gz < 1 || !out) return PB_ERR_ARG;
  const int nb = gx * gy * gz;
  for (int h = 0; h < nb; ++h) {
    const MarchTile t = march_tile(h, gx, gy, gz);
    out[3 * h] = t.bx;
    out[3 * h + 1] = t.by;
    out[3 * h + 2] = t.bz;
  }
  return PB_OK;
}

extern "C" int pb_op3d_matvec(pb_op3d* op, const void* x, void* y, int precision) {
  if (!op || !x || !y) return pb_fail(op ? op->ctx : nullptr, PB_ERR_ARG, "pb_op3d_matvec: null argument");
  if (!op->assembled) return pb_fail(op->ctx, PB_ERR_STATE, "pb_op3d_matvec: operator not assembled");
  hipStream_t s = op->ctx->stream;
  if (int rc = march_check(op, march_grid(op), Z3)) return rc;
  if (precision == PB_F64)
    hipLaunchKernelGGL(k3_matvec<double>, march_grid(op), kBlk3, 0, s, op3<double>(op), (const double*)x, (double*)y);
  else
    hipLaunchKernelGGL(k3_matvec<float>, march_grid(op), kBlk3, 0, s, op3<float>(op), (const float*)x, (float*)y);
  PB_LAUNCH_CHECK(op->ctx);
  return PB_OK;
}

// Solve (D^-1 A) x = b.  PB_SOLVE_F64 / PB_SOLVE_F64_FAST: FP64 BiCGStab with
// the reference's legacy tolerance; PB_SOLVE_MIXED: FP32 BiCGStab on the
// residual inside FP64 iterative refinement until the FP64 true residual
// ||b - A x|| <= tol ||b|| (a refinement that does not reduce it switches to
// an FP64 inner solve).  Synchronous.
extern "C" int pb_op3d_solve(pb_op3d* op, const double* b, double* x, const pb_solve_opts* opts,
                             pb_solve_info* info) {
  if (!op || !b || !x || !opts || !info) return pb_fail(op ? op->ctx : nullptr, PB_ERR_ARG, "pb_op3d_solve: null");
  if (!op->assembled) return pb_fail(op->ctx, PB_ERR_STATE, "pb_op3d_solve: operator not assembled");
  if (!(opts->tol >= 0) || opts->maxiter < 1) return pb_fail(op->ctx, PB_ERR_ARG, "pb_op3d_solve: bad tol/maxiter");
  if (opts->x0_given) return pb_fail(op->ctx, PB_ERR_ARG, "pb_op3d_solve: x0 must be zero");
  int rc = alloc_work3(op);
  if (rc) return rc;
  hipStream_t s = op->ctx->stream;
  if (s == nullptr) {   // the null stream cannot be captured: run on the context's own stream
    (void)hipStreamSynchronize(nullptr);
    s = op->ctx->own_stream;
  }
  const long n = op->n;
  const dim3 eg = elem_grid3(n);
  pb_solve_info out{};
  if (opts->mode == PB_SOLVE_F64 || opts->mode == PB_SOLVE_F64_FAST) {
    KS3 h{};
    rc = krylov3<double>(op, b, opts->tol, opts->maxiter, 1, opts->check_every, s, &h);
    if (rc) return rc;
    PB_HIP(op->ctx, hipMemcpyAsync(x, op->x[0], sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    out.info = info3(h);
    out.iters = h.iter;
    out.outer = 1;
    out.half_step = 0;
    out.resid = h.rnorm;
    out.bnrm2 = h.bnrm;
  } else if (opts->mode == PB_SOLVE_MIXED) {
    // 1e-5: two refinements at 257x129x129 where 1e-4 took three (FP32
    // BiCGStab reaches it without stagnating on the 27-point operator)
    const double inner = opts->inner_tol > 0 ? opts->inner_tol : 1e-5;
    const int max_outer = opts->max_outer > 0 ? opts->max_outer : 20;
    hipLaunchKernelGGL(k3_fill, eg, dim3(256), 0, s, n, x, 0.0);
    double bn = 0.0;
    rc = resid3(op, x, b, s, &bn);   // r = b, b32 = (float) b
    if (rc) return rc;
    const double target = opts->tol * bn;
    double rn = bn;
    int iters = 0, outer = 0, last = 0;
    bool fp64 = false;
    while (rn > target && outer < max_outer && iters < opts->maxiter) {
      KS3 h{};
      const double tol_rel = std::max(inner, 0.5 * target / rn);
      if (!fp64) {
        rc = krylov3<float>(op, (const float*)op->b32, tol_rel, opts->maxiter - iters, 0, opts->check_every, s, &h);
        if (rc) return rc;
        hipLaunchKernelGGL(k3_cast<float>, eg, dim3(256), 0, s, n, (const float*)op->x[1], x, 1);
      } else {
        rc = krylov3<double>(op, (const double*)op->r64, tol_rel, opts->maxiter - iters, 0, opts->check_every, s,
                             &h);
        if (rc) return rc;
        hipLaunchKernelGGL(k3_cast<double>, eg, dim3(256), 0, s, n, (const double*)op->x[0], x, 1);
      }
      PB_LAUNCH_CHECK(op->ctx);
      iters += h.iter;
      last = info3(h);
      ++outer;
      const double rn_old = rn;
      rc = resid3(op, x, b, s, &rn);
      if (rc) return rc;
      if (!(rn < 0.5 * rn_old)) fp64 = true;   // FP32 refinement stalled
    }
    out.info = rn <= target ? 0 : (last < 0 ? last : std::max(iters, 1));
    out.iters = iters;
    out.outer = outer;
    out.resid = rn;
    out.bnrm2 = bn;
  } else {
    return pb_fail(op->ctx, PB_ERR_ARG, "pb_op3d_solve: mode %d not available in 3D", opts->mode);
  }
  double tr = 0.0;
  rc = resid3(op, x, b, s, &tr);
  if (rc) return rc;
  out.true_rel_resid = out.bnrm2 > 0 ? tr / out.bnrm2 : tr;
  *info = out;
  return out.info;
}

// Average device time (microseconds) of one launch of a 3D solver kernel on
// pattern-filled vectors: 0 matvec, 1 launch 1 of the iteration, 2 launch 2.
extern "C" int pb_op3d_time_kernel(pb_op3d* op, int kernel, int precision, int reps, double* avg_us) {
  if (!op || !avg_us || reps < 1 || kernel < 0 || kernel > 2)
    return pb_fail(op ? op->ctx : nullptr, PB_ERR_ARG, "pb_op3d_time_kernel: bad argument");
  if (!op->assembled) return pb_fail(op->ctx, PB_ERR_STATE, "pb_op3d_time_kernel: operator not assembled");
  int rc = alloc_work3(op);
  if (rc) return rc;
  hipStream_t s = op->ctx->stream;
  const long n = op->n;
  const dim3 eg = elem_grid3(n), mg = march_grid(op);
  const int P = precision == PB_F64 ? 0 : 1;
  if ((rc = march_check(op, mg, Z3))) return rc;
  if ((rc = march_check(op, march_grid(op, P == 0 ? kZ1<double> : kZ1<float>), P == 0 ? kZ1<double> : kZ1<float>)))
    return rc;
  // vectors r, p, v, t, rt, x of both sets hold a constant pattern; the state
  // carries small fixed scalars so repeated launches stay finite
  if (P == 0) {
    for (void* vec : {op->r[0][0], op->p[0][0], op->v[0][0], op->t[0], op->rt[0], op->x[0], op->r[0][1]})
      hipLaunchKernelGGL(k3_fill, eg, dim3(256), 0, s, n, (double*)vec, 1e-3);
  } else {
    hipLaunchKernelGGL(k3_fill, eg, dim3(256), 0, s, n, op->r64, 1e-3);
    for (void* vec : {op->r[1][0], op->p[1][0], op->v[1][0], op->t[1], op->rt[1], op->x[1], op->r[1][1]})
      hipLaunchKernelGGL(k3_cast_down, eg, dim3(256), 0, s, n, (const double*)op->r64, (float*)vec);
  }
  // Each timed launch reads state st[0] and writes st[1], so repeated
  // launches see the same state.  Launch 1 folds pending five-dot records
  // (1, 0.5, 1, 1, 0.5) per block: omega 0.5, beta 1 with rho = 0.75 nb;
  // launch 2 folds r~.v records of 1: alpha = rho / nb.  The solve keeps
  // running (atol 0) and every value stays finite.
  const int nbm = march_blocks(op);
  KS3 k{};
  k.rho = 0.75 * nbm;
  k.alpha = 0.5;
  k.omega = 0.25;
  k.beta = 0.125;
  k.maxiter = 1 << 30;
  k.status = S3_RUN;
  k.pend = kernel == 1 ? 1 : 0;
  *op->hst = k;
  PB_HIP(op->ctx, hipMemcpyAsync(op->st, op->hst, sizeof(KS3), hipMemcpyHostToDevice, s));
  std::vector<double> rec5((size_t)nbm * kNV), rec1((size_t)nbm, 1.0);
  for (int b = 0; b < nbm; ++b) {   // value-major records, as write_partials3 stores them
    const double pat[kNV] = {1.0, 0.5, 1.0, 1.0, 0.5};
    for (int q = 0; q < kNV; ++q) rec5[(size_t)q * nbm + b] = pat[q];
  }
  PB_HIP(op->ctx, hipMemcpyAsync(op->part, rec5.data(), sizeof(double) * rec5.size(), hipMemcpyHostToDevice, s));
  PB_HIP(op->ctx, hipMemcpyAsync(op->part1, rec1.data(), sizeof(double) * rec1.size(), hipMemcpyHostToDevice, s));
  PB_HIP(op->ctx, hipStreamSynchronize(s));
  auto launch = [&]() {
    if (kernel == 0) {
      if (P == 0)
        hipLaunchKernelGGL(k3_matvec<double>, mg, kBlk3, 0, s, op3<double>(op), (const double*)op->p[0][0],
                           (double*)op->v[0][1]);
      else
        hipLaunchKernelGGL(k3_matvec<float>, mg, kBlk3, 0, s, op3<float>(op), (const float*)op->p[1][0],
                           (float*)op->v[1][1]);
    } else if (kernel == 1) {
      if (P == 0)
        hipLaunchKernelGGL(k3_iter1<double>, march_grid(op, kZ1<double>), kBlk3, 0, s, op3<double>(op),
                           (const KS3*)op->st, op->st + 1,
                           (const double*)op->part, nbm, (const double*)op->r[0][0], (const double*)op->p[0][0],
                           (const double*)op->v[0][0], (const double*)op->t[0], (const double*)op->rt[0],
                           (double*)op->r[0][1], (double*)op->p[0][1], (double*)op->v[0][1], (double*)op->x[0],
                           op->part1);
      else
        hipLaunchKernelGGL(k3_iter1<float>, march_grid(op, kZ1<float>), kBlk3, 0, s, op3<float>(op),
                           (const KS3*)op->st, op->st + 1,
                           (const double*)op->part, nbm, (const float*)op->r[1][0], (const float*)op->p[1][0],
                           (const float*)op->v[1][0], (const float*)op->t[1], (const float*)op->rt[1],
                           (float*)op->r[1][1], (float*)op->p[1][1], (float*)op->v[1][1], (float*)op->x[1],
                           op->part1);
    } else {   // reads part1's records and writes its own five-dot records to part
      if (P == 0)
        hipLaunchKernelGGL(k3_iter2<double>, mg, kBlk3, 0, s, op3<double>(op), (const KS3*)op->st, op->st + 1,
                           (const double*)op->part1, march_blocks(op, kZ1<double>), (const double*)op->r[0][0],
                           (const double*)op->v[0][0],
                           (const double*)op->rt[0], (double*)op->t[0], op->part);
      else
        hipLaunchKernelGGL(k3_iter2<float>, mg, kBlk3, 0, s, op3<float>(op), (const KS3*)op->st, op->st + 1,
                           (const double*)op->part1, march_blocks(op, kZ1<float>), (const float*)op->r[1][0],
                           (const float*)op->v[1][0],
                           (const float*)op->rt[1], (float*)op->t[1], op->part);
    }
  };
  launch();
  PB_LAUNCH_CHECK(op->ctx);
  hipEvent_t e0, e1;
  PB_HIP(op->ctx, hipEventCreate(&e0));
  PB_HIP(op->ctx, hipEventCreate(&e1));
  PB_HIP(op->ctx, hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) launch();
  PB_HIP(op->ctx, hipEventRecord(e1, s));
  PB_HIP(op->ctx, hipEventSynchronize(e1));
  float ms = 0.f;
  PB_HIP(op->ctx, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  PB_LAUNCH_CHECK(op->ctx);
  *avg_us = 1000.0 * ms / reps;
  return PB_OK;
}
