// 2D nodal operator handle and the LDS-tiled stencil machinery (gfx950).
//
// The 9-point operator of lap2D_gather_new (laplacian.py:546-746) in flux
// form.  For op-cell (j, i) (rows 0..ny, cols 0..nx; corners are nodes
// (j-1..j, i-1..i), with the reference's wrap onto the duplicated periodic end
// node: i-1 -> nx-2 at i=0, i+1 -> 1 at i=nx-1, likewise in y):
//   gx = (p_tr - p_tl) + (p_br - p_bl),  gy = (p_bl - p_tl) + (p_br - p_tr)
//   Fx = Wxx gx + Wyx gy,                Fy = Wyy gy + Wxy gx
// with folded weights Wxx = hx cxx/(4 dx^2), Wyy = hy cyy/(4 dy^2),
// Wyx = hy cxy/(4 dx dy), Wxy = hx cyx/(4 dx dy) (hx = wplus[0], hy = wplus[1]
// zeroed on cells outside walls; the 1/4 is the product of the two 0.5
// factors of laplacian.py:721-735).  A node sums its four cells:
//   (Fx_TR + Fx_BR - Fx_TL - Fx_BL) + (Fy_BL + Fy_BR - Fy_TL - Fy_TR) + hc p,
// times diag_inv — term by term Dxx + Dyy + Dyx + Dxy + hcenter p
// (laplacian.py:731-739).
//
// Tiles are 64 nodes wide (one wavefront per row) and 4*RY rows tall; a block
// of 64x4 threads loads the (rows+2) x 66 halo tile of its input expression
// into LDS once (coalesced: one row of 66 values per wavefront pass), forms
// every cell flux of the tile once into LDS, then each thread gathers RY
// nodes.  Halo re-reads between neighbouring tiles are the only redundancy.
#pragma once
#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

#include "pb_common.hpp"

namespace pb2d {

constexpr int TX = 64;    // tile width = one wavefront
constexpr int TYT = 4;    // thread rows per block
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

// LDS of one tile: halo of the input, cell fluxes.
template <typename T, int RY>
struct TileLds {
  static constexpr int TY = TYT * RY;
  T sp[(TY + 2) * (TX + 2)];
  T sfx[(TY + 1) * (TX + 1)];
  T sfy[(TY + 1) * (TX + 1)];
};

// Apply the operator on tile (bx, by): y = diag_inv * (A in)(g) for each node g
// of the tile, consumed by ``out``.  Every global load is issued up front: the
// raw halo values (In::load), this thread's cell weights and its nodes'
// epilogue operands (Out::prefetch, diag_inv, hcenter) are requested before the
// first barrier, then ``prologue()`` runs (it may reduce the previous launch's
// partials and set the scalars that In::combine and Out use; returning false
// ends the block), so the latencies of all of them overlap.  Loads are
// unconditional (out-of-grid threads read node 0 and discard it), so no branch
// separates them and the waits before the prologue cover only what it needs.
//
// Row assignment: in pass q < RY a thread loads the raw values of its own
// output row q (halo row 1 + ty + TYT q), so the epilogue receives them as
// ``out(r, g, y, pc, pre, raw)`` (r: the thread's row 0..RY-1) without loading
// them again; one more pass loads the top (ty == 0) and bottom (ty == 1) halo
// rows.  The 2 x HR edge-column values are loaded by the first 2 HR threads,
// one each, so the other lanes spend neither load instructions nor registers
// on them.
template <typename T, bool CROSS, bool CENTER, int RY, class In, class Out, class Pro>
__device__ __forceinline__ void stencil_tile2(const OpDev<T>& op, int bx, int by, In& in, Out& out,
                                              TileLds<T, RY>& L, Pro&& prologue) {
  constexpr int TY = TYT * RY;
  constexpr int HR = TY + 2;
  constexpr int NPASS = RY + 1;
  constexpr int NCELLS = (TY + 1) * (TX + 1);
  constexpr int NCP = (NCELLS + NTHR - 1) / NTHR;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int tid = tx + TX * ty;
  const int i0 = bx * TX, j0 = by * TY;
  const int nx = op.nx, ny = op.ny;
  int ic = i0 + tx;
  ic = (ic == nx) ? 1 : ic;
  const bool col_ok = ic < nx;
  typename In::Raw raw[NPASS];
  bool rok[NPASS];
  int hrow[NPASS];
#pragma unroll
  for (int q = 0; q < NPASS; ++q) {
    const int hl = q < RY ? 1 + ty + TYT * q : (ty == 0 ? 0 : (ty == 1 ? HR - 1 : -1));
    hrow[q] = hl;
    int j = j0 - 1 + hl;
    j = (j < 0) ? ny - 2 : (j == ny ? 1 : j);
    rok[q] = hl >= 0 && j < ny;
    raw[q] = in.load((rok[q] && col_ok) ? (long)j * nx + ic : 0);   // node 0 when unused
  }
  // The two edge columns of all HR halo rows: thread tid < 2 HR loads the
  // value of halo row tid / 2, left (even tid) or right (odd tid) edge.
  static_assert(2 * HR <= NTHR, "edge loads need 2 HR threads");
  const bool edge_thr = tid < 2 * HR;
  const int ehl = tid >> 1, eside = tid & 1;
  bool eok = false;
  typename In::Raw eraw;
  {
    int ie = eside ? i0 + TX : i0 - 1;
    ie = (ie < 0) ? nx - 2 : (ie == nx ? 1 : ie);
    int j = j0 - 1 + ehl;
    j = (j < 0) ? ny - 2 : (j == ny ? 1 : j);
    eok = edge_thr && ie < nx && j < ny;
    eraw = in.load(eok ? (long)j * nx + ie : 0);
  }
  T wa[NCP], wb[NCP], wc[NCP], wd[NCP];
#pragma unroll
  for (int c = 0; c < NCP; ++c) {
    const int k = tid + c * NTHR;
    const int cy = k / (TX + 1), cx = k - cy * (TX + 1);
    const int j = j0 + cy, i = i0 + cx;
    const bool cok = k < NCELLS && j <= ny && i <= nx;
    const long cc = cok ? (long)j * (nx + 1) + i : 0;
    wa[c] = cok ? op.wxx[cc] : T(0);
    wb[c] = cok ? op.wyy[cc] : T(0);
    wc[c] = (CROSS && cok) ? op.wyx[cc] : T(0);
    wd[c] = (CROSS && cok) ? op.wxy[cc] : T(0);
  }
  typename Out::Pre pre[RY];
  T dv[RY], hv[RY];
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int j = j0 + ty + TYT * r, i = i0 + tx;
    const bool nok = j < ny && i < nx;
    const long g = nok ? (long)j * nx + i : 0;
    pre[r] = out.prefetch(g);
    dv[r] = nok ? op.dinv[g] : T(0);
    hv[r] = (CENTER && nok) ? op.hc[g] : T(0);
  }
  if (!prologue()) return;
#pragma unroll
  for (int q = 0; q < NPASS; ++q) {
    const int hl = hrow[q];
    if (hl >= 0) {
      L.sp[hl * (TX + 2) + tx + 1] = (rok[q] && col_ok) ? in.combine(raw[q]) : T(0);
    }
  }
  if (edge_thr) L.sp[ehl * (TX + 2) + (eside ? TX + 1 : 0)] = eok ? in.combine(eraw) : T(0);
  __syncthreads();
#pragma unroll
  for (int c = 0; c < NCP; ++c) {
    const int k = tid + c * NTHR;
    if (k < NCELLS) {
      const int cy = k / (TX + 1), cx = k - cy * (TX + 1);
      const T* sp = L.sp + cy * (TX + 2) + cx;
      const T tl = sp[0], tr = sp[1], bl = sp[TX + 2], br = sp[TX + 3];
      const T gx = (tr - tl) + (br - bl);
      const T gy = (bl - tl) + (br - tr);
      T fx = wa[c] * gx, fy = wb[c] * gy;
      if (CROSS) {
        fx += wc[c] * gy;
        fy += wd[c] * gx;
      }
      L.sfx[k] = fx;
      L.sfy[k] = fy;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int ly = ty + TYT * r, lx = tx;
    const int j = j0 + ly, i = i0 + lx;
    if (j < ny && i < nx) {
      const long g = (long)j * nx + i;
      const int c0 = ly * (TX + 1) + lx, c1 = c0 + (TX + 1);
      T acc = ((L.sfx[c0 + 1] + L.sfx[c1 + 1]) - (L.sfx[c0] + L.sfx[c1])) +
              ((L.sfy[c1] + L.sfy[c1 + 1]) - (L.sfy[c0] + L.sfy[c0 + 1]));
      const T pc = L.sp[(ly + 1) * (TX + 2) + lx + 1];
      if (CENTER) acc += hv[r] * pc;
      out(r, g, dv[r] * acc, pc, pre[r], raw[r]);
    }
  }
}

// The first two partial records of this thread (b = tid, tid + NTHR), loaded
// before the tile's own loads so that their latency overlaps them: vector
// loads complete in issue order, so records requested after the tile would
// wait for the whole tile to arrive first.
template <int NV>
struct EarlyPartials {
  double v[2][NV];
  // Unconditional loads of a clamped record (nb >= 1): a guarded load would
  // share its branch with the guarded add of sum_partials, and the add would
  // then wait for the load right here.
  __device__ __forceinline__ void load(const double* P, int nb) {
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int b = min(tid + h * NTHR, nb - 1);
#pragma unroll
      for (int k = 0; k < NV; ++k) v[h][k] = P[(long)k * nb + b];
    }
  }
};

// Sum ``nb`` block partials of NV values in a fixed order; the result is
// valid in every thread.  sred: NV * (NTHR/64).  Records are stored value
// by value, [k * nb + b] (written by write_partials with the same nb), so
// the 64 lanes of a load read 64 consecutive words: with [b * NV + k] a
// five- or nine-value record set spreads one load over 20-36 cache lines.
// ``early`` holds records tid and tid + NTHR (EarlyPartials::load on the same
// P and nb); thread tid adds its records b = tid, tid + NTHR, ... in order.
template <int NV>
__device__ __forceinline__ void sum_partials(const EarlyPartials<NV>& early, const double* P, int nb,
                                             double (&tot)[NV], double* sred) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
#pragma unroll
  for (int k = 0; k < NV; ++k) tot[k] = 0.0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (tid + h * NTHR < nb) {
#pragma unroll
      for (int k = 0; k < NV; ++k) tot[k] += early.v[h][k];
    }
  }
  for (int b = tid + 2 * NTHR; b < nb; b += NTHR) {
#pragma unroll
    for (int k = 0; k < NV; ++k) tot[k] += P[(long)k * nb + b];
  }
  // wave sums into sred, then every thread adds them in wave order (the same
  // order in every thread and block); sred is next written after the tile's
  // own barriers.
  constexpr int NW = NTHR / kWave;
  const int lane = tid & (kWave - 1), wid = tid / kWave;
  wave_sum_n<NV>(tot);
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    if (lane == 0) sred[k * NW + wid] = tot[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) s += sred[k * NW + w];
    tot[k] = s;
  }
}

template <int NV>
__device__ __forceinline__ void sum_partials(const double* P, int nb, double (&tot)[NV], double* sred) {
  EarlyPartials<NV> early;
  early.load(P, nb);
  sum_partials<NV>(early, P, nb, tot, sred);
}

// Block sum of NV thread values, written as this block's record (plain
// stores, value k at P[k * nb + bid]).
template <int NV>
__device__ __forceinline__ void write_partials(double (&v)[NV], double* sred, double* P, int nb, int bid) {
  block_sum_w<NV, NTHR / kWave>(v, sred);
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  if (tid == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) P[(long)k * nb + bid] = v[k];
  }
}

// Device state of a Krylov run, double-buffered: a kernel reads S[in] and its
// block 0 writes S[out], so no block ever reads a value written in its own
// launch.  Scalars are derived redundantly (identically) in every block from
// the previous launch's block partials, so no grid-wide fence is needed.
struct KState {
  double rho, alpha, omega, beta;
  double atol, bnrm2, resid, tol, rnorm0;
  int iter, maxiter, status, half, legacy, phase, x_upd;
};

// A state record as 8-byte words.  Loaded at a kernel's entry from a uniform
// address that nothing in the kernel writes, the words are scalar loads held
// in SGPRs; the thread that writes the next record stores them back and then
// overwrites the fields it changes, with no load at the kernel's end.
struct StateWords {
  static_assert(sizeof(KState) % 8 == 0, "KState is copied in 8-byte words");
  long long w[sizeof(KState) / 8];
};

__device__ __forceinline__ StateWords load_state_words(const KState* src) {
  StateWords r;
  const long long* s = reinterpret_cast<const long long*>(src);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(KState) / 8); ++i) r.w[i] = s[i];
  return r;
}

__device__ __forceinline__ void store_state_words(KState* dst, const StateWords& r) {
  long long* d = reinterpret_cast<long long*>(dst);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(KState) / 8); ++i) d[i] = r.w[i];
}

enum {
  ST_RUN = 0,
  ST_CONV = 1,
  ST_MAXIT = 3,
  ST_BRK_RHO = -10,
  ST_BRK_OMEGA = -11,
  // pending: the final x update is still to be applied by the next launch
  ST_P_CONV = 10,
  ST_P_HALF = 11,
  ST_P_MAXIT = 12,
  ST_P_BRK_RHO = 13,
  ST_P_BRK_OMEGA = 14,
  // final, one-launch iteration only: the two copies of a duplicated periodic
  // end node differ in the initial residual or the operator, so the solve
  // runs the two-launch iteration instead
  ST_SEAM = -20,
};
enum { PH_INIT = 0, PH_A = 1, PH_B = 2, PH_L3 = 3, PH_F = 4 };

__host__ __device__ inline bool st_final(int s) {
  return s == ST_CONV || s == ST_MAXIT || s == ST_BRK_RHO || s == ST_BRK_OMEGA || s == ST_SEAM;
}

template <typename T>
struct Work {
  T* r[2] = {nullptr, nullptr};
  T* p[2] = {nullptr, nullptr};
  T* v[2] = {nullptr, nullptr};
  T* t = nullptr;
  T* rt = nullptr;
  T* b = nullptr;          // right-hand side copy in T
  double* x = nullptr;     // iterate (always FP64)
  double* PA = nullptr;    // partials of the A launch (1 value)
  double* PB = nullptr;    // partials of the init / B launch (<= 5 values)
  double* PC = nullptr;    // partials of the exact-mode update launch (2 values)
  // one-launch iteration (FP64 only): the second t buffer (s, p, v use r, p,
  // v), u = A^T r~, and its partials (10 values), double-buffered
  T* t2 = nullptr;
  T* u = nullptr;
  double* PF[2] = {nullptr, nullptr};
  KState* S = nullptr;     // [3]
  std::map<int, hipGraphExec_t> graphs;   // (variant, chunk) key -> instantiated graph
};

}  // namespace pb2d

namespace pbmg {
struct Hier;
}
struct FgWork;

struct pb_op2d {
  pb_ctx* ctx = nullptr;
  pb_grid2d g{};
  long n = 0;      // nodes
  long nc = 0;     // op-cells
  bool cross = false, center = true, assembled = false;
  long version = 0;   // bumped by every assembly (keys caches derived from the weights)
  int ry = 4;      // tile rows per thread (4 or 8)
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
  // solver workspace (allocated at the first solve)
  bool have_work = false;
  pb2d::Work<double> w64;
  pb2d::Work<float> w32;
  double* rtrue = nullptr;   // FP64 true residual
  double* xcorr = nullptr;   // FP64 correction (mixed mode)
  double* zero = nullptr;    // FP64 zeros
  double* PT = nullptr;      // partials of true-residual launches
  double* red = nullptr;     // [2] true-residual result slots
  pb2d::KState* hstate = nullptr;   // pinned host mirror
  double* hred = nullptr;           // pinned host mirror of red
  std::vector<void*> allocs;
  // learned-preconditioner mode: multigrid hierarchy, flexible-GMRES
  // workspace and the attached network (owned / referenced by this operator)
  pbmg::Hier* mg = nullptr;
  FgWork* fg = nullptr;
  pb_nn* nn = nullptr;
};

namespace pb2d {

template <typename T>
inline OpDev<T> opdev(const pb_op2d* op);
template <>
inline OpDev<double> opdev<double>(const pb_op2d* op) {
  return OpDev<double>{op->wxx, op->wyy, op->wyx, op->wxy, op->hc, op->dinv, op->g.nx, op->g.ny};
}
template <>
inline OpDev<float> opdev<float>(const pb_op2d* op) {
  return OpDev<float>{op->wxx32, op->wyy32, op->wyx32, op->wxy32, op->hc32, op->dinv32, op->g.nx, op->g.ny};
}

inline dim3 stencil_grid(const pb_op2d* op, int ry) {
  return dim3(ceil_div(op->g.nx, TX), ceil_div(op->g.ny, TYT * ry));
}

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

inline int elem_blocks(long n) { return (int)std::min<long>((n + 1023) / 1024, 2048); }

}  // namespace pb2d
