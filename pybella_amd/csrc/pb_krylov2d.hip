// BiCGStab on the 2D nodal operator, two fused stencil launches per iteration
// (gfx950).  Replaces the reference's scipy.sparse.linalg.bicgstab call
// (second_projection.py:230; SciPy 1.7.3 BiCGSTAB of Barrett et al.,
// Templates 1994, sec. 2.3.8, M = I).
//
// Iteration k of the Templates algorithm
//   p_k = r_{k-1} + beta (p_{k-1} - omega v_{k-1}),  v_k = A p_k,
//   alpha = rho / (rt.v_k),  s = r_{k-1} - alpha v_k,  t = A s,
//   omega = (t.s)/(t.t),  x += alpha p_k + omega s,  r_k = s - omega t,
// is split into two launches:
//   A(k): x update of iteration k-1, r_{k-1} = s_{k-1} - omega t formed on the
//         fly, p_k, v_k = A p_k; block partials of rt.v_k;
//   B(k): alpha from A's partials; s = r - alpha v on the fly, t = A s; block
//         partials of s.s, t.s, t.t, rt.s, rt.t.
// The scalars of the next iteration follow from B's sums without another pass:
//   ||r_k||^2 = s.s - 2 omega t.s + omega^2 t.t,  rho_k = rt.s - omega rt.t.
// Stopping follows the reference's legacy semantics: x0 = 0, return at once if
// ||b - A x0|| <= tol (absolute), stop at ||s|| <= tol ||b|| (x += alpha p) or
// ||r|| <= tol ||b||, eps^2 breakdown tests on rho and omega.
//
// Each launch derives its scalars itself: every block sums the previous
// launch's block partials in block order (deterministic, no grid fence), and
// block 0 stores the new state in the other half of a double buffer.  x is
// kept in FP64 also when the Krylov vectors are FP32.
#include <cstddef>
#include <cstdlib>
#include <type_traits>

#include "pb_op2d.hpp"

using namespace pb2d;

namespace {

__device__ __forceinline__ double eps2() { return 4.930380657631324e-32; }   // DBL_EPSILON^2

__device__ __forceinline__ int block_id() { return blockIdx.x + blockIdx.y * gridDim.x; }
__device__ __forceinline__ int nblocks() { return gridDim.x * gridDim.y; }

// XCD-aware tile order.  Hardware block h is dispatched to XCD h mod 8 (8
// XCDs, each with its own L2).  Giving XCD x the contiguous tile range
// [start(x), start(x) + count(x)) keeps vertically adjacent tiles, which share
// halo rows, in one L2.  Bijective on [0, n); partial records stay indexed by
// the tile number, so sums and iterates do not depend on the mapping.
__device__ __forceinline__ int xcd_tile(int h, int n) {
  constexpr int NX = 8;
  const int x = h % NX, k = h / NX, base = n / NX, rem = n % NX;
  return x < rem ? x * (base + 1) + k : rem * (base + 1) + (x - rem) * base + k;
}
__device__ __forceinline__ int thread_id() { return threadIdx.x + threadIdx.y * blockDim.x; }

// ---------------------------------------------------------------------------
// init: r0 = b - A x0, rt = r0, p = v = t = 0, x = x0 (or 0); records
// (r.r, b.b, 0...) of NREC values
// ---------------------------------------------------------------------------
struct Empty {};

template <typename T>
struct InitIn {
  using Raw = double;
  const double* x0;
  int given;
  __device__ Raw load(long g) const { return given ? x0[g] : 0.0; }
  __device__ T combine(const Raw& r) const { return (T)r; }
};
template <typename T>
struct InitOut {
  using Pre = T;
  const T* b;
  T *r, *rt, *p, *v, *t;
  double* x;
  const double* x0;
  int given;
  double rr, bb;
  __device__ Pre prefetch(long g) const { return b[g]; }
  template <class R>
  __device__ void operator()(int, long g, T y, T, const Pre& bg, const R&) {
    const T rg = bg - y;
    r[g] = rg;
    rt[g] = rg;
    p[g] = T(0);
    v[g] = T(0);
    t[g] = T(0);
    x[g] = given ? x0[g] : 0.0;
    rr += (double)rg * (double)rg;
    bb += (double)bg * (double)bg;
  }
};

// tile (bx, by) of an operator whose launch has nb blocks, this one numbered bid
template <int NREC, typename T, bool CROSS, bool CENTER, int RY>
__device__ __forceinline__ void init_body(OpDev<T>& op, InitIn<T>& in, InitOut<T>& out, double* P, int bx, int by,
                                          int bid, int nb) {
  __shared__ TileLds<T, RY> L;
  __shared__ double sred[8 * (NTHR / kWave)];
  out.rr = 0.0;
  out.bb = 0.0;
  stencil_tile2<T, CROSS, CENTER, RY>(op, bx, by, in, out, L, [] { return true; });
  double v[NREC];
#pragma unroll
  for (int k = 0; k < NREC; ++k) v[k] = 0.0;
  v[0] = out.rr;
  v[1] = out.bb;
  write_partials<NREC>(v, sred, P, nb, bid);
}

template <typename T, bool CROSS, bool CENTER, int RY>
__global__ __launch_bounds__(NTHR) void k_init5(OpDev<T> op, InitIn<T> in, InitOut<T> out, double* P) {
  const int tile = xcd_tile(block_id(), nblocks());
  init_body<5, T, CROSS, CENTER, RY>(op, in, out, P, tile % gridDim.x, tile / gridDim.x, tile, nblocks());
}
template <typename T, bool CROSS, bool CENTER, int RY>
__global__ __launch_bounds__(NTHR) void k_init2(OpDev<T> op, InitIn<T> in, InitOut<T> out, double* P) {
  const int tile = xcd_tile(block_id(), nblocks());
  init_body<2, T, CROSS, CENTER, RY>(op, in, out, P, tile % gridDim.x, tile / gridDim.x, tile, nblocks());
}

// ---------------------------------------------------------------------------
// scalar steps shared by every block
// ---------------------------------------------------------------------------
__device__ __forceinline__ KState scalars_after_init(KState st, double rr, double bb) {
  st.bnrm2 = sqrt(bb);
  st.rnorm0 = sqrt(rr);
  st.resid = st.rnorm0;
  st.iter = 0;
  st.half = 0;
  st.alpha = 0.0;
  st.omega = 0.0;
  st.beta = 0.0;
  st.x_upd = 0;
  st.rho = rr;
  st.atol = st.tol * st.bnrm2;
  if (st.legacy ? (st.rnorm0 <= st.tol) : (st.rnorm0 <= st.atol)) {
    st.status = ST_CONV;
  } else if (fabs(rr) < eps2()) {
    st.status = ST_BRK_RHO;
  } else {
    st.status = ST_RUN;
  }
  return st;
}

__device__ __forceinline__ KState scalars_after_b(KState st, const double (&tot)[5]) {
  const double ss = tot[0], ts = tot[1], tt = tot[2], rts = tot[3], rtt = tot[4];
  const double snrm = sqrt(ss);
  st.x_upd = 1;
  if (snrm <= st.atol) {
    st.status = ST_P_HALF;
    st.resid = snrm;
    return st;
  }
  const double om = ts / tt;
  if (!isfinite(om) || !isfinite(snrm)) {
    st.status = ST_BRK_OMEGA;
    return st;
  }
  st.omega = om;
  const double rr = ss - 2.0 * om * ts + om * om * tt;
  const double rho_new = rts - om * rtt;
  st.iter += 1;
  st.resid = sqrt(fmax(rr, 0.0));
  if (st.resid <= st.atol) {
    st.status = ST_P_CONV;
  } else if (fabs(om) < eps2()) {
    st.status = ST_P_BRK_OMEGA;
  } else if (st.iter >= st.maxiter) {
    st.status = ST_P_MAXIT;
  } else {
    const double beta = (rho_new / st.rho) * (st.alpha / om);
    st.rho = rho_new;
    if (fabs(rho_new) < eps2() || !isfinite(beta)) {
      st.status = ST_P_BRK_RHO;
    } else {
      st.beta = beta;
      st.status = ST_RUN;
    }
  }
  return st;
}

__device__ __forceinline__ KState scalars_after_l3(KState st, double rtr, double rr) {
  const double rnrm = sqrt(rr);
  st.iter += 1;
  st.resid = rnrm;
  if (!isfinite(rnrm) || !isfinite(rtr)) {
    st.status = ST_BRK_RHO;
  } else if (rnrm <= st.atol) {
    st.status = ST_CONV;
  } else if (fabs(st.omega) < eps2()) {
    st.status = ST_BRK_OMEGA;
  } else if (st.iter >= st.maxiter) {
    st.status = ST_MAXIT;
  } else if (fabs(rtr) < eps2()) {
    st.status = ST_BRK_RHO;
  } else {
    st.beta = (rtr / st.rho) * (st.alpha / st.omega);
    st.rho = rtr;
    st.status = ST_RUN;
  }
  return st;
}

__device__ __forceinline__ int finalize(int pending) {
  switch (pending) {
    case ST_P_CONV: return ST_CONV;
    case ST_P_HALF: return ST_CONV;
    case ST_P_MAXIT: return ST_MAXIT;
    case ST_P_BRK_RHO: return ST_BRK_RHO;
    case ST_P_BRK_OMEGA: return ST_BRK_OMEGA;
    default: return pending;
  }
}

// ---------------------------------------------------------------------------
// fast variant, launch A(k)
// ---------------------------------------------------------------------------
template <typename T>
struct IterPtrs {
  const T* r_prev;
  const T* p_prev;
  const T* v_prev;
  const T* t;
  const T* rt;
  T* r_new;
  T* p_new;
  T* v_new;
  double* x;
};

template <typename T>
struct AIn {
  struct Raw {
    T r, p, v, t;
  };
  const T *r, *p, *v, *t;
  T alpha, omega, beta;
  __device__ Raw load(long g) const { return Raw{r[g], p[g], v[g], t[g]}; }
  __device__ T combine(const Raw& q) const {
    const T s = q.r - alpha * q.v;
    const T rn = s - omega * q.t;
    return rn + beta * (q.p - omega * q.v);
  }
};

template <typename T>
struct AOut {
  struct Pre {
    T rt;
    double x;
  };
  IterPtrs<T> q;
  T alpha, omega;
  double alpha_d, omega_d;
  int xupd;
  double rv;
  __device__ Pre prefetch(long g) const { return Pre{q.rt[g], q.x[g]}; }
  // e: this node's prefetched operands; w: its raw r, p, v, t (loaded for the tile)
  __device__ void operator()(int, long g, T y, T pn, const Pre& e, const typename AIn<T>::Raw& w) {
    const T s = w.r - alpha * w.v;
    const T rn = s - omega * w.t;
    if (xupd) {
      const double xg = e.x + alpha_d * (double)w.p;
      __builtin_nontemporal_store(xg + omega_d * (double)s, q.x + g);
    }
    __builtin_nontemporal_store(rn, q.r_new + g);
    __builtin_nontemporal_store(pn, q.p_new + g);
    __builtin_nontemporal_store(y, q.v_new + g);
    rv += (double)e.rt * (double)y;
  }
};

template <typename T, bool CROSS, bool CENTER, int RY>
__device__ __forceinline__ void iter_a_body(const OpDev<T>& op, const IterPtrs<T>& q, const KState* Sin, KState* Sout,
                                            const double* PB, double* PA, int bx, int by, int bid, int nb) {
  __shared__ TileLds<T, RY> L;
  __shared__ double sred[8 * (NTHR / kWave)];
  const int tid = thread_id();
  AIn<T> in{q.r_prev, q.p_prev, q.v_prev, q.t, T(0), T(0), T(0)};
  AOut<T> out{q, T(0), T(0), 0.0, 0.0, 0, 0.0};
  KState st = *Sin;
  EarlyPartials<5> early;
  early.load(PB, nb);
  auto pro = [&]() -> bool {
    double tot[5];
    sum_partials<5>(early, PB, nb, tot, sred);
    if (st_final(st.status)) {
      if (bid == 0 && tid == 0) *Sout = st;
      return false;
    }
    st = (st.phase == PH_INIT) ? scalars_after_init(st, tot[0], tot[1]) : scalars_after_b(st, tot);
    st.phase = PH_A;
    if (st.status != ST_RUN) {
      if (st.status >= ST_P_CONV) {   // final x update of a finished run
        const bool half = st.status == ST_P_HALF;
        const T alpha = (T)st.alpha;
        constexpr int TY = TYT * RY;
#pragma unroll
        for (int r = 0; r < RY; ++r) {
          const int j = by * TY + threadIdx.y + TYT * r, i = bx * TX + threadIdx.x;
          if (j < op.ny && i < op.nx) {
            const long g = (long)j * op.nx + i;
            const double xg = q.x[g] + st.alpha * (double)q.p_prev[g];
            if (half) {
              q.x[g] = xg;
            } else {
              const T s = q.r_prev[g] - alpha * q.v_prev[g];
              q.x[g] = xg + st.omega * (double)s;
            }
          }
        }
        if (half) st.half = 1;
        st.status = finalize(st.status);
      }
      if (bid == 0 && tid == 0) *Sout = st;
      return false;
    }
    in.alpha = out.alpha = (T)st.alpha;
    in.omega = out.omega = (T)st.omega;
    in.beta = (T)st.beta;
    out.alpha_d = st.alpha;
    out.omega_d = st.omega;
    out.xupd = st.x_upd;
    return true;
  };
  stencil_tile2<T, CROSS, CENTER, RY>(op, bx, by, in, out, L, pro);
  if (st.status != ST_RUN || st_final(st.status)) return;
  double v[1] = {out.rv};
  write_partials<1>(v, sred, PA, nb, bid);
  if (bid == 0 && tid == 0) *Sout = st;
}

template <typename T, bool CROSS, bool CENTER, int RY>
__global__ __launch_bounds__(NTHR) void k_iter_a(OpDev<T> op, IterPtrs<T> q, const KState* Sin, KState* Sout,
                                                 const double* PB, double* PA) {
  const int tile = xcd_tile(block_id(), nblocks());
  iter_a_body<T, CROSS, CENTER, RY>(op, q, Sin, Sout, PB, PA, tile % gridDim.x, tile / gridDim.x, tile, nblocks());
}

// ---------------------------------------------------------------------------
// launch B(k) (shared by both variants)
// ---------------------------------------------------------------------------
template <typename T>
struct BIn {
  struct Raw {
    T r, v;
  };
  const T *r, *v;
  T alpha;
  __device__ Raw load(long g) const { return Raw{r[g], v[g]}; }
  __device__ T combine(const Raw& q) const { return q.r - alpha * q.v; }
};
// t = y is held per row and stored by flush() after the block's reduction:
// stores issued before the reduction would make its cross-lane steps wait
// for them (their registers are reused by the reduction).
template <typename T, int RY>
struct BOut {
  using Pre = T;
  T* t;
  const T* rt;
  double ss, ts, tt, rts, rtt;
  T tv[RY];
  unsigned tg[RY];   // node of row r, or ~0u for a row outside the grid
  __device__ Pre prefetch(long g) const { return rt[g]; }
  template <class R>
  __device__ void operator()(int r, long g, T y, T s, const Pre& rtg, const R&) {
    tv[r] = y;
    tg[r] = (unsigned)g;
    const double sd = (double)s, yd = (double)y, rd = (double)rtg;
    ss += sd * sd;
    ts += yd * sd;
    tt += yd * yd;
    rts += rd * sd;
    rtt += rd * yd;
  }
  __device__ void flush() const {
#pragma unroll
    for (int r = 0; r < RY; ++r)
      if (tg[r] != ~0u) __builtin_nontemporal_store(tv[r], t + tg[r]);
  }
};

template <typename T, bool CROSS, bool CENTER, int RY>
__device__ __forceinline__ void iter_b_body(const OpDev<T>& op, const T* r, const T* v, const T* rt, T* t,
                                            const KState* Sin, KState* Sout, const double* PA, double* PB, int bx,
                                            int by, int bid, int nb) {
  __shared__ TileLds<T, RY> L;
  __shared__ double sred[8 * (NTHR / kWave)];
  const int tid = thread_id();
  BIn<T> in{r, v, T(0)};
  BOut<T, RY> out{t, rt, 0.0, 0.0, 0.0, 0.0, 0.0, {}, {}};
#pragma unroll
  for (int r = 0; r < RY; ++r) out.tg[r] = ~0u;
  // The state record as scalar words (see StateWords); rho and the status are
  // read from it, the writer thread stores it back with alpha and the phase.
  const StateWords sw = load_state_words(Sin);
  const double rho = Sin->rho;
  const int status = Sin->status;
  EarlyPartials<1> early;
  early.load(PA, nb);
  const bool writer = bid == 0 && tid == 0;
  double alpha = 0.0;
  bool run = false;
  auto pro = [&]() -> bool {
    double tot[1];
    sum_partials<1>(early, PA, nb, tot, sred);
    if (status != ST_RUN) {
      if (writer) store_state_words(Sout, sw);
      return false;
    }
    alpha = rho / tot[0];
    if (tot[0] == 0.0 || !isfinite(alpha)) {
      if (writer) {
        store_state_words(Sout, sw);
        Sout->phase = PH_B;
        Sout->status = ST_BRK_OMEGA;
      }
      return false;
    }
    in.alpha = (T)alpha;
    run = true;
    return true;
  };
  stencil_tile2<T, CROSS, CENTER, RY>(op, bx, by, in, out, L, pro);
  if (!run) return;
  double vals[5] = {out.ss, out.ts, out.tt, out.rts, out.rtt};
  write_partials<5>(vals, sred, PB, nb, bid);
  out.flush();
  if (writer) {
    store_state_words(Sout, sw);
    Sout->phase = PH_B;
    Sout->alpha = alpha;
  }
}

template <typename T, bool CROSS, bool CENTER, int RY>
__global__ __launch_bounds__(NTHR) void k_iter_b(OpDev<T> op, const T* r, const T* v, const T* rt, T* t,
                                                 const KState* Sin, KState* Sout, const double* PA, double* PB) {
  const int tile = xcd_tile(block_id(), nblocks());
  iter_b_body<T, CROSS, CENTER, RY>(op, r, v, rt, t, Sin, Sout, PA, PB, tile % gridDim.x, tile / gridDim.x, tile,
                                    nblocks());
}

// ---------------------------------------------------------------------------
// exact variant: three launches per iteration with rho = rt.r and ||r||
// formed from the stored r as the reference does.
//   L1(k): beta from L3's partials (or init); p_k = beta (p - omega v) + r,
//          v_k = A p_k; partials rt.v
//   L2(k): the B launch (alpha; s = r - alpha v, t = A s; dots)
//   L3(k): ||s|| test, omega; x += alpha p + omega s; r = s - omega t;
//          partials rt.r, r.r
// ---------------------------------------------------------------------------
template <typename T>
struct L1In {
  struct Raw {
    T r, p, v;
  };
  const T *r, *p, *v;
  T omega, beta;
  __device__ Raw load(long g) const { return Raw{r[g], p[g], v[g]}; }
  __device__ T combine(const Raw& q) const { return beta * (q.p - omega * q.v) + q.r; }
};
template <typename T>
struct L1Out {
  using Pre = T;
  T *p_new, *v_new;
  const T* rt;
  double rv;
  __device__ Pre prefetch(long g) const { return rt[g]; }
  template <class R>
  __device__ void operator()(int, long g, T y, T pn, const Pre& rtg, const R&) {
    p_new[g] = pn;
    v_new[g] = y;
    rv += (double)rtg * (double)y;
  }
};

template <typename T, bool CROSS, bool CENTER, int RY>
__global__ __launch_bounds__(NTHR) void k_ex_p(OpDev<T> op, const T* r, const T* p_prev, const T* v_prev,
                                               T* p_new, T* v_new, const T* rt, const KState* Sin, KState* Sout,
                                               const double* PC, double* PA) {
  __shared__ TileLds<T, RY> L;
  __shared__ double sred[8 * (NTHR / kWave)];
  const int tid = thread_id(), nb = nblocks(), bid = xcd_tile(block_id(), nb);
  L1In<T> in{r, p_prev, v_prev, T(0), T(0)};
  L1Out<T> out{p_new, v_new, rt, 0.0};
  KState st = *Sin;
  EarlyPartials<2> early;
  early.load(PC, nb);
  bool run = false;
  auto pro = [&]() -> bool {
    double tot[2];
    sum_partials<2>(early, PC, nb, tot, sred);
    if (st_final(st.status)) {
      if (bid == 0 && tid == 0) *Sout = st;
      return false;
    }
    st = (st.phase == PH_INIT) ? scalars_after_init(st, tot[0], tot[1]) : scalars_after_l3(st, tot[0], tot[1]);
    st.phase = PH_A;
    if (st.status != ST_RUN) {
      if (bid == 0 && tid == 0) *Sout = st;
      return false;
    }
    in.omega = (T)st.omega;
    in.beta = (T)st.beta;
    run = true;
    return true;
  };
  stencil_tile2<T, CROSS, CENTER, RY>(op, bid % gridDim.x, bid / gridDim.x, in, out, L, pro);
  if (!run) return;
  double v[1] = {out.rv};
  write_partials<1>(v, sred, PA, nb, bid);
  if (bid == 0 && tid == 0) *Sout = st;
}

template <typename T, int RY>
__global__ __launch_bounds__(NTHR) void k_ex_update(int nx, int ny, T* r, const T* p, const T* v, const T* t,
                                                    const T* rt, double* x, const KState* Sin, KState* Sout,
                                                    const double* PB, double* PC) {
  __shared__ double sred[8 * (NTHR / kWave)];
  const int tid = thread_id(), nb = nblocks(), bid = xcd_tile(block_id(), nb);
  const int bx = bid % gridDim.x, by = bid / gridDim.x;
  constexpr int TY = TYT * RY;
  // The state record as scalar words (see StateWords); the writer thread
  // stores it back with the fields this launch changes.
  const StateWords sw = load_state_words(Sin);
  const int status = Sin->status;
  const double atol = Sin->atol, alpha_d = Sin->alpha;
  const bool writer = bid == 0 && tid == 0;
  EarlyPartials<5> early;
  early.load(PB, nb);
  T er[RY], ep[RY], ev[RY], et[RY], ert[RY];
  double ex[RY];
  bool ok[RY];
  long gg[RY];
  // unconditional loads (threads past the grid read node 0, unused): one
  // straight run of loads with no branch between the rows
#pragma unroll
  for (int q = 0; q < RY; ++q) {
    const int j = by * TY + threadIdx.y + TYT * q, i = bx * TX + threadIdx.x;
    ok[q] = j < ny && i < nx;
    const long g = ok[q] ? (long)j * nx + i : 0;
    gg[q] = g;
    er[q] = r[g];
    ep[q] = p[g];
    ev[q] = v[g];
    et[q] = t[g];
    ert[q] = rt[g];
    ex[q] = x[g];
  }
  double tot[5];
  sum_partials<5>(early, PB, nb, tot, sred);
  if (status != ST_RUN) {
    if (writer) store_state_words(Sout, sw);
    return;
  }
  const double snrm = sqrt(tot[0]);
  const bool half = snrm <= atol;
  double omega_d = 0.0;
  if (!half) {
    omega_d = tot[1] / tot[2];
    if (!isfinite(omega_d) || !isfinite(snrm)) {
      if (writer) {
        store_state_words(Sout, sw);
        Sout->status = ST_BRK_OMEGA;
      }
      return;
    }
  }
  const T alpha = (T)alpha_d, omega = (T)omega_d;
  double rtr = 0.0, rr = 0.0;
#pragma unroll
  for (int q = 0; q < RY; ++q) {
    if (ok[q]) {
      const long g = gg[q];
      const double xg = ex[q] + alpha_d * (double)ep[q];
      if (half) {
        x[g] = xg;
      } else {
        const T s = er[q] - alpha * ev[q];
        __builtin_nontemporal_store(xg + omega_d * (double)s, x + g);
        const T rn = s - omega * et[q];
        __builtin_nontemporal_store(rn, r + g);
        rtr += (double)ert[q] * (double)rn;
        rr += (double)rn * (double)rn;
      }
    }
  }
  double vals[2] = {rtr, rr};
  __syncthreads();   // every wave has read the prologue's sums out of sred
  write_partials<2>(vals, sred, PC, nb, bid);
  if (writer) {
    store_state_words(Sout, sw);
    Sout->phase = PH_L3;
    if (half) {
      Sout->resid = snrm;
      Sout->half = 1;
      Sout->status = ST_CONV;
    } else {
      Sout->omega = omega_d;
    }
  }
}

// ---------------------------------------------------------------------------
// one-launch variant (PB_SOLVE_F64_FUSED)
//
// alpha = rho / (rt.v_k) needs a grid-wide sum between v_k = A p_k and
// s_k = r - alpha v_k.  With u = A^T rt, formed once per solve, rt.v_k =
// u.p_k, and u.p_k follows from the sums of the launch before:
//   u.p_k = (u.s - omega u.t) + beta (u.p_{k-1} - omega u.v_{k-1}).
// So iteration k is one launch F(k):
//   omega, ||r_{k-1}||, rho and beta from F(k-1)'s sums of s.s, t.s, t.t,
//   rt.s, rt.t (as the B launch's five dots give them), alpha_k = rho / u.p_k;
//   r_{k-1} = s - omega t and p_k = r_{k-1} + beta (p_{k-1} - omega v_{k-1})
//   on the tile with a two-node halo; x += alpha_{k-1} p_{k-1} + omega s;
//   v_k = A p_k on the tile and the one-node ring around it; s_k = r -
//   alpha_k v_k there; t_k = A s_k on the tile; sums s.s, t.s, t.t, rt.s,
//   rt.t, u.s, u.t, u.p, u.v of iteration k.
// Positions outside the grid take the values of the nodes the wrap maps them
// to (period n - 1 on both axes, as the one-node halo of the operator: -1 ->
// n-2, n -> 1).  On a wall axis those positions only meet cells of zero
// weight.  On a periodic axis the ring node at -1 stands for node n-2 and
// takes node 0 as its neighbour where node n-2 has node n-1: exact while the
// two copies of the duplicated end node are equal in every vector, which
// holds when they are equal in the initial residual and the operator's
// diag_inv, hcenter and seam cells (every later vector is formed by the same
// operations on equal values).  The launch that forms u counts mismatches of
// those, and F(1) ends a solve that has any with ST_SEAM.
// ---------------------------------------------------------------------------
// records: s.s, t.s, t.t, rt.s, rt.t, u.s, u.t, u.p, u.v; the adjoint
// launch's records carry the seam mismatches in the t.s slot
constexpr int NFREC = 9;

// Positions reach at most two nodes before the grid and one tile past it.
// The wrap is exact for -2 <= i <= 2n - 2, which covers every position whose
// value reaches a node of the grid (at most two past the last node); farther
// positions (tiles wider than the grid) are clamped to the last node: any
// finite value serves there.
__device__ __forceinline__ int wrap_node(int i, int n) {
  const int m = n - 1;
  return i < 0 ? i + m : (i >= n ? min(i - m, m) : i);
}
__device__ __forceinline__ int wrap_cell(int c, int n) {   // op-cells 0..n, same rule
  const int m = n - 1;
  return c < 0 ? c + m : (c > n ? min(c - m, n) : c);
}

// (A^T rt)_j for node (j, i).  Node i's row of A gathers op-cells (J, I),
// J in {j, j+1}, I in {i, i+1}, which have node i as a corner in their own
// (unwrapped) coordinates; a cell's flux reads its corners through the wrap.
// So node j receives from every cell one of whose corners wraps onto j: per
// axis, cell i (j as its right / lower corner), cell i + 1 (left / upper),
// and at the seams cell 0 (left corner -1 -> n-2) and cell n (right corner
// n -> 1).  Each such cell passes on the weighted values z = diag_inv rt of
// its unwrapped corners that are nodes.
template <bool CROSS, bool CENTER>
__device__ __forceinline__ double adjoint_at(const OpDev<double>& op, const double* rt, int j, int i) {
  const int nx = op.nx, ny = op.ny;
  const long g = (long)j * nx + i;
  double u = CENTER ? op.hc[g] * (op.dinv[g] * rt[g]) : 0.0;
  int cxs[4], kxs[4], ncx = 0, cys[4], kys[4], ncy = 0;
  cxs[ncx] = i, kxs[ncx++] = 1;
  cxs[ncx] = i + 1, kxs[ncx++] = 0;
  if (i == nx - 2) cxs[ncx] = 0, kxs[ncx++] = 0;
  if (i == 1) cxs[ncx] = nx, kxs[ncx++] = 1;
  cys[ncy] = j, kys[ncy++] = 1;
  cys[ncy] = j + 1, kys[ncy++] = 0;
  if (j == ny - 2) cys[ncy] = 0, kys[ncy++] = 0;
  if (j == 1) cys[ncy] = ny, kys[ncy++] = 1;
  for (int a = 0; a < ncy; ++a) {
    for (int b = 0; b < ncx; ++b) {
      const int cy = cys[a], cx = cxs[b];
      const long cc = (long)cy * (nx + 1) + cx;
      const double gxk = kxs[b] ? 1.0 : -1.0, gyk = kys[a] ? 1.0 : -1.0;
      double cfx = op.wxx[cc] * gxk, cfy = op.wyy[cc] * gyk;
      if (CROSS) {
        cfx += op.wyx[cc] * gyk;
        cfy += op.wxy[cc] * gxk;
      }
#pragma unroll
      for (int iy = 0; iy < 2; ++iy) {
        const int nj = cy - 1 + iy;
        if (nj < 0 || nj >= ny) continue;
#pragma unroll
        for (int ix = 0; ix < 2; ++ix) {
          const int ni = cx - 1 + ix;
          if (ni < 0 || ni >= nx) continue;
          const long gn = (long)nj * nx + ni;
          const double coef = (ix == 0 ? cfx : -cfx) + (iy == 0 ? cfy : -cfy);
          u += coef * (op.dinv[gn] * rt[gn]);
        }
      }
    }
  }
  return u;
}

// mismatches between the two copies of the duplicated end nodes that node
// (j, i) is responsible for: its own pair when it lies on the first column
// (row) of a periodic axis, and the pairs of seam cells of its cell row
// (column)
template <bool CROSS, bool CENTER>
__device__ __forceinline__ double seam_mismatch(const OpDev<double>& op, const double* rt, int j, int i, int per_x,
                                                int per_y) {
  const int nx = op.nx, ny = op.ny;
  double m = 0.0;
  auto ne = [&](double a, double b) { m += (a == b) ? 0.0 : 1.0; };
  auto nodes = [&](long g0, long g1) {
    ne(rt[g0], rt[g1]);
    ne(op.dinv[g0], op.dinv[g1]);
    if (CENTER) ne(op.hc[g0], op.hc[g1]);
  };
  auto cells = [&](long c0, long c1) {
    ne(op.wxx[c0], op.wxx[c1]);
    ne(op.wyy[c0], op.wyy[c1]);
    if (CROSS) {
      ne(op.wyx[c0], op.wyx[c1]);
      ne(op.wxy[c0], op.wxy[c1]);
    }
  };
  const long W = nx + 1;
  if (per_x && i == 0) {
    nodes((long)j * nx, (long)j * nx + nx - 1);
    for (int J = j; J <= (j == ny - 1 ? ny : j); ++J) {
      cells(J * W + 0, J * W + nx - 1);
      cells(J * W + 1, J * W + nx);
    }
  }
  if (per_y && j == 0) {
    nodes(i, (long)(ny - 1) * nx + i);
    for (int I = i; I <= (i == nx - 1 ? nx : i); ++I) {
      cells(I, (long)(ny - 1) * W + I);
      cells(W + I, (long)ny * W + I);
    }
  }
  return m;
}

// u = A^T rt on the tiles of the one-launch iteration, after the init launch:
// the scalars after init from its records, records (rt.rt, seam mismatches,
// 0, rt.rt, 0, u.rt, 0, 0, 0) for F(1)
template <bool CROSS, bool CENTER, int RY>
__global__ __launch_bounds__(NTHR) void k_adj_f(OpDev<double> op, const double* rt, double* u, int per_x, int per_y,
                                                const KState* Sin, KState* Sout, const double* PB, double* Pout) {
  __shared__ double sred[NFREC * (NTHR / kWave)];
  const int tid = thread_id(), nb = nblocks(), bid = xcd_tile(block_id(), nb);
  const int bx = bid % gridDim.x, by = bid / gridDim.x;
  constexpr int TY = TYT * RY;
  double tot[5];
  sum_partials<5>(PB, nb, tot, sred);
  const KState st = scalars_after_init(*Sin, tot[0], tot[1]);
  double acc[NFREC];
#pragma unroll
  for (int k = 0; k < NFREC; ++k) acc[k] = 0.0;
#pragma unroll
  for (int q = 0; q < RY; ++q) {
    const int j = by * TY + threadIdx.y + TYT * q, i = bx * TX + threadIdx.x;
    if (j < op.ny && i < op.nx) {
      const long g = (long)j * op.nx + i;
      const double uj = adjoint_at<CROSS, CENTER>(op, rt, j, i), r = rt[g];
      u[g] = uj;
      acc[0] += r * r;
      acc[3] += r * r;
      acc[5] += uj * r;
      acc[1] += seam_mismatch<CROSS, CENTER>(op, rt, j, i, per_x, per_y);
    }
  }
  __syncthreads();   // every wave has read the init sums out of sred
  write_partials<NFREC>(acc, sred, Pout, nb, bid);
  if (bid == 0 && tid == 0) *Sout = st;
}

template <int RY>
struct FusedLds {
  static constexpr int TY = TYT * RY;
  static constexpr int H2R = TY + 4, H2C = TX + 4;   // nodes with a two-node halo
  static constexpr int H1R = TY + 2, H1C = TX + 2;   // nodes with a one-node ring
  static constexpr int C1R = TY + 3, C1C = TX + 3;   // cells between the H2 nodes
  double P[H2R * H2C];   // p_k
  double R[H1R * H1C];   // r_{k-1}
  double S[H1R * H1C];   // s_k
  double FX[C1R * C1C], FY[C1R * C1C];
};

struct FPtrs {
  const double *s0, *t0, *p0, *v0;   // iteration k-1
  double *s1, *t1, *p1, *v1;         // iteration k
  const double *rt, *u;
  double* x;
};

template <bool CROSS, bool CENTER, int RY>
__global__ __launch_bounds__(NTHR) void k_iter_f(OpDev<double> op, FPtrs q, const KState* Sin, KState* Sout,
                                                 const double* Pin, double* Pout) {
  using Lt = FusedLds<RY>;
  static_assert(TYT == 4, "the halo-row pass gives each of the four thread rows one of the four halo rows");
  constexpr int TY = Lt::TY, H2C = Lt::H2C, H1C = Lt::H1C, C1C = Lt::C1C;
  constexpr int NC1 = Lt::C1R * C1C, NCP = (NC1 + NTHR - 1) / NTHR;
  constexpr int NRING = 2 * H1C + 2 * TY, NEDGE = 4 * Lt::H2R;
  static_assert(NRING <= NTHR && NEDGE <= NTHR, "one ring node and one edge value per thread");
  __shared__ Lt L;
  __shared__ double sred[NFREC * (NTHR / kWave)];
  const int tid = thread_id(), nb = nblocks(), bid = xcd_tile(block_id(), nb);
  const int bx = bid % gridDim.x, by = bid / gridDim.x;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int nx = op.nx, ny = op.ny, i0 = bx * TX, j0 = by * TY;
  const bool writer = bid == 0 && tid == 0;
  KState st = *Sin;
  EarlyPartials<NFREC> early;
  early.load(Pin, nb);
  auto gw = [&](int j, int i) { return (unsigned)(wrap_node(j, ny) * nx + wrap_node(i, nx)); };
  // iteration k-1's s, t, p, v on the two-node halo tile: the own rows
  // (q < RY), one of the four halo rows (q = RY: rows 0, 1, TY+2, TY+3 by
  // thread row) and one value of the four edge columns (q = RY + 1)
  constexpr int NQ = RY + 2;
  double rs[NQ], rt_[NQ], rp[NQ], rv[NQ];
  int hr[NQ], hc[NQ];
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    if (k < RY) {
      hr[k] = 2 + ty + TYT * k;
      hc[k] = 2 + tx;
    } else if (k == RY) {
      hr[k] = ty < 2 ? ty : TY + ty;
      hc[k] = 2 + tx;
    } else {
      const int e = tid < NEDGE ? tid : 0, side = e & 3;
      hr[k] = e >> 2;
      hc[k] = side < 2 ? side : TX + side;
    }
    const unsigned g = gw(j0 - 2 + hr[k], i0 - 2 + hc[k]);
    rs[k] = q.s0[g];
    rt_[k] = q.t0[g];
    rp[k] = q.p0[g];
    rv[k] = q.v0[g];
  }
  // own positions: diag_inv, hcenter (every position: positions past the
  // grid feed the ring); x, rt, u are loaded after the scalar step
  double dv[RY], hv[RY];
  bool ok[RY];
  unsigned og[RY];
#pragma unroll
  for (int k = 0; k < RY; ++k) {
    const int j = j0 + ty + TYT * k, i = i0 + tx;
    ok[k] = j < ny && i < nx;
    og[k] = ok[k] ? (unsigned)(j * nx + i) : 0u;
    const unsigned g = gw(j, i);
    dv[k] = op.dinv[g];
    hv[k] = CENTER ? op.hc[g] : 0.0;
  }
  // this thread's ring node (H1 coordinates)
  int ra = 0, rb = 0;
  if (tid < H1C) {
    ra = 0, rb = tid;
  } else if (tid < 2 * H1C) {
    ra = TY + 1, rb = tid - H1C;
  } else if (tid < 2 * H1C + TY) {
    ra = 1 + tid - 2 * H1C, rb = 0;
  } else {
    ra = 1 + min(tid - 2 * H1C - TY, TY - 1), rb = H1C - 1;
  }
  const bool ring = tid < NRING;
  const unsigned rg = gw(j0 - 1 + ra, i0 - 1 + rb);
  const double rdv = op.dinv[rg], rhc = CENTER ? op.hc[rg] : 0.0;
  // cell weights of the C1 cells this thread forms
  double wa[NCP], wb[NCP], wc[NCP], wd[NCP];
#pragma unroll
  for (int c = 0; c < NCP; ++c) {
    const int k = min(tid + c * NTHR, NC1 - 1);
    const int cy = k / C1C, cx = k - cy * C1C;
    const unsigned cc = (unsigned)(wrap_cell(j0 - 1 + cy, ny) * (nx + 1) + wrap_cell(i0 - 1 + cx, nx));
    wa[c] = op.wxx[cc];
    wb[c] = op.wyy[cc];
    wc[c] = CROSS ? op.wyx[cc] : 0.0;
    wd[c] = CROSS ? op.wxy[cc] : 0.0;
  }
  // scalars (every thread the same)
  double tot[NFREC];
  sum_partials<NFREC>(early, Pin, nb, tot, sred);
  const double alpha_old = st.alpha;
  bool run = false, xup = false, half = false;
  double al = 0.0, om = 0.0, be = 0.0;
  if (!st_final(st.status)) {
    if (st.phase == PH_INIT) {
      if (tot[1] != 0.0) {
        st.status = ST_SEAM;
      } else {
        const double a = st.rho / tot[5];
        if (tot[5] == 0.0 || !isfinite(a)) st.status = ST_BRK_OMEGA;
        else al = a, run = true;
      }
    } else {
      const double t5[5] = {tot[0], tot[1], tot[2], tot[3], tot[4]};
      st = scalars_after_b(st, t5);
      if (st.status == ST_RUN) {
        om = st.omega;
        be = st.beta;
        xup = true;
        const double up = (tot[5] - om * tot[6]) + be * (tot[7] - om * tot[8]);
        const double a = st.rho / up;
        if (up == 0.0 || !isfinite(a)) st.status = ST_BRK_OMEGA;   // after iteration k-1's x update
        else al = a, run = true;
      } else if (st.status >= ST_P_CONV) {   // final x update of a finished run
        xup = true;
        half = st.status == ST_P_HALF;
        om = st.omega;
        if (half) st.half = 1;
        st.status = finalize(st.status);
      }
    }
  }
  st.phase = PH_F;
  if (!run) {
    if (xup) {
#pragma unroll
      for (int k = 0; k < RY; ++k) {
        if (ok[k]) {
          const double xg = q.x[og[k]] + alpha_old * rp[k];
          q.x[og[k]] = half ? xg : xg + om * rs[k];
        }
      }
    }
    if (writer) *Sout = st;
    return;
  }
  st.alpha = al;
  // r_{k-1} and p_k on the halo tile
  double pr[RY], rr[RY];
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    if (k <= RY || tid < NEDGE) {
      const double r = rs[k] - om * rt_[k];
      const double p = r + be * (rp[k] - om * rv[k]);
      L.P[hr[k] * H2C + hc[k]] = p;
      if (hr[k] >= 1 && hr[k] <= TY + 2 && hc[k] >= 1 && hc[k] <= TX + 2) L.R[(hr[k] - 1) * H1C + hc[k] - 1] = r;
      if (k < RY) {
        pr[k] = p;
        rr[k] = r;
      }
    }
  }
  // iteration k-1's x update is applied at the end of the launch; x, rt and
  // u arrive during the two products
  double ox[RY], ort[RY], ou[RY];
#pragma unroll
  for (int k = 0; k < RY; ++k) {
    ox[k] = q.x[og[k]];
    ort[k] = q.rt[og[k]];
    ou[k] = q.u[og[k]];
  }
  __syncthreads();
  // v_k = A p_k: cell fluxes, then the own nodes and the ring
#pragma unroll
  for (int c = 0; c < NCP; ++c) {
    const int k = tid + c * NTHR;
    if (k < NC1) {
      const int cy = k / C1C, cx = k - cy * C1C;
      const double* sp = L.P + cy * H2C + cx;
      const double tl = sp[0], tr = sp[1], bl = sp[H2C], br = sp[H2C + 1];
      const double gx = (tr - tl) + (br - bl), gy = (bl - tl) + (br - tr);
      double fx = wa[c] * gx, fy = wb[c] * gy;
      if (CROSS) {
        fx += wc[c] * gy;
        fy += wd[c] * gx;
      }
      L.FX[k] = fx;
      L.FY[k] = fy;
    }
  }
  __syncthreads();
  auto gather = [&](int a, int b) {   // flux sum of H1 node (a, b)
    const int c0 = a * C1C + b, c1 = c0 + C1C;
    return ((L.FX[c0 + 1] + L.FX[c1 + 1]) - (L.FX[c0] + L.FX[c1])) +
           ((L.FY[c1] + L.FY[c1 + 1]) - (L.FY[c0] + L.FY[c0 + 1]));
  };
  double vv[RY], ss[RY];
#pragma unroll
  for (int k = 0; k < RY; ++k) {
    const int a = 1 + ty + TYT * k, b = 1 + tx;
    double acc = gather(a, b);
    if (CENTER) acc += hv[k] * pr[k];
    vv[k] = dv[k] * acc;
    ss[k] = rr[k] - al * vv[k];
    L.S[a * H1C + b] = ss[k];
  }
  if (ring) {
    double acc = gather(ra, rb);
    if (CENTER) acc += rhc * L.P[(ra + 1) * H2C + rb + 1];
    L.S[ra * H1C + rb] = L.R[ra * H1C + rb] - al * (rdv * acc);
  }
  __syncthreads();
  // t_k = A s_k: fluxes of the cells inside the ring, then the own nodes
#pragma unroll
  for (int c = 0; c < NCP; ++c) {
    const int k = tid + c * NTHR;
    if (k < NC1) {
      const int cy = k / C1C, cx = k - cy * C1C;
      if (cy >= 1 && cy <= TY + 1 && cx >= 1 && cx <= TX + 1) {
        const double* sp = L.S + (cy - 1) * H1C + (cx - 1);
        const double tl = sp[0], tr = sp[1], bl = sp[H1C], br = sp[H1C + 1];
        const double gx = (tr - tl) + (br - bl), gy = (bl - tl) + (br - tr);
        double fx = wa[c] * gx, fy = wb[c] * gy;
        if (CROSS) {
          fx += wc[c] * gy;
          fy += wd[c] * gx;
        }
        L.FX[k] = fx;
        L.FY[k] = fy;
      }
    }
  }
  __syncthreads();
  double acc[NFREC];
#pragma unroll
  for (int k = 0; k < NFREC; ++k) acc[k] = 0.0;
#pragma unroll
  for (int k = 0; k < RY; ++k) {
    const int a = 1 + ty + TYT * k, b = 1 + tx;
    double y = gather(a, b);
    if (CENTER) y += hv[k] * ss[k];
    const double t = dv[k] * y;
    if (ok[k]) {
      const unsigned g = og[k];
      if (xup) __builtin_nontemporal_store((ox[k] + alpha_old * rp[k]) + om * rs[k], q.x + g);
      __builtin_nontemporal_store(pr[k], q.p1 + g);
      __builtin_nontemporal_store(vv[k], q.v1 + g);
      __builtin_nontemporal_store(ss[k], q.s1 + g);
      __builtin_nontemporal_store(t, q.t1 + g);
      const double s = ss[k];
      acc[0] += s * s;
      acc[1] += t * s;
      acc[2] += t * t;
      acc[3] += ort[k] * s;
      acc[4] += ort[k] * t;
      acc[5] += ou[k] * s;
      acc[6] += ou[k] * t;
      acc[7] += ou[k] * pr[k];
      acc[8] += ou[k] * vv[k];
    }
  }
  write_partials<NFREC>(acc, sred, Pout, nb, bid);
  if (writer) *Sout = st;
}

// ---------------------------------------------------------------------------
// FP64 true residual r = b - A x with partials of r.r, and a one-block sum
// ---------------------------------------------------------------------------
struct TrIn {
  using Raw = double;
  const double* x;
  __device__ Raw load(long g) const { return x[g]; }
  __device__ double combine(const Raw& v) const { return v; }
};
struct TrOut {
  using Pre = double;
  const double* b;
  double* r;
  double rr;
  __device__ Pre prefetch(long g) const { return b[g]; }
  template <class R>
  __device__ void operator()(int, long g, double y, double, const Pre& bg, const R&) {
    const double rg = bg - y;
    r[g] = rg;
    rr += rg * rg;
  }
};

template <bool CROSS, bool CENTER, int RY>
__global__ __launch_bounds__(NTHR) void k_true_res(OpDev<double> op, const double* x, const double* b, double* r,
                                                   double* PT) {
  __shared__ TileLds<double, RY> L;
  __shared__ double sred[8 * (NTHR / kWave)];
  TrIn in{x};
  TrOut out{b, r, 0.0};
  const int tile = xcd_tile(block_id(), nblocks());
  stencil_tile2<double, CROSS, CENTER, RY>(op, tile % gridDim.x, tile / gridDim.x, in, out, L, [] { return true; });
  double v[1] = {out.rr};
  write_partials<1>(v, sred, PT, nblocks(), tile);
}

__global__ __launch_bounds__(NTHR) void k_sum1(const double* P, int nb, double* out) {
  __shared__ double sred[8 * (NTHR / kWave)];
  double tot[1];
  sum_partials<1>(P, nb, tot, sred);
  if (thread_id() == 0) *out = tot[0];
}

template <typename Tin, typename Tout>
__global__ void k_cast(long n, const Tin* a, Tout* b) {
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x)
    b[g] = (Tout)a[g];
}

__global__ void k_axpy(long n, double a, const double* dx, double* x) {
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x)
    x[g] += a * dx[g];
}

template <typename T>
__global__ void k_fill_pattern(long n, T* a, double phase) {
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x)
    a[g] = (T)(0.5 + 0.25 * sin(0.37 * (double)g + phase));
}

__global__ void k_fill_const(long n, double* a, double v) {
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x) a[g] = v;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <typename T>
Work<T>& work_of(pb_op2d* op);
template <>
Work<double>& work_of<double>(pb_op2d* op) { return op->w64; }
template <>
Work<float>& work_of<float>(pb_op2d* op) { return op->w32; }

template <typename T>
int alloc_vecs(pb_op2d* op, Work<T>& w, int nb) {
  const long n = op->n;
  int rc = 0;
  for (int k = 0; k < 2; ++k) {
    rc |= dmalloc(op, &w.r[k], n);
    rc |= dmalloc(op, &w.p[k], n);
    rc |= dmalloc(op, &w.v[k], n);
  }
  rc |= dmalloc(op, &w.t, n);
  rc |= dmalloc(op, &w.rt, n);
  rc |= dmalloc(op, &w.b, n);
  rc |= dmalloc(op, &w.x, n);
  rc |= dmalloc(op, &w.PA, nb);
  rc |= dmalloc(op, &w.PB, 5L * nb);
  rc |= dmalloc(op, &w.PC, 2L * nb);
  if (std::is_same<T, double>::value) {   // one-launch iteration (FP64 only)
    rc |= dmalloc(op, &w.t2, n);
    rc |= dmalloc(op, &w.u, n);
    rc |= dmalloc(op, &w.PF[0], (long)NFREC * nb);
    rc |= dmalloc(op, &w.PF[1], (long)NFREC * nb);
  }
  rc |= dmalloc(op, &w.S, 3);
  return rc;
}

int alloc_work(pb_op2d* op) {
  if (op->have_work) return PB_OK;
  const dim3 sg = stencil_grid(op, 2);   // the finest tiling bounds the block count
  const int nb = (int)(sg.x * sg.y);
  int rc = alloc_vecs<double>(op, op->w64, nb);
  rc |= alloc_vecs<float>(op, op->w32, nb);
  rc |= dmalloc(op, &op->rtrue, op->n);
  rc |= dmalloc(op, &op->xcorr, op->n);
  rc |= dmalloc(op, &op->zero, op->n);
  rc |= dmalloc(op, &op->PT, nb);
  rc |= dmalloc(op, &op->red, 2);
  if (rc) return PB_ERR_HIP;
  void* h = nullptr;
  if (hipHostMalloc(&h, sizeof(KState) * 2, hipHostMallocDefault) != hipSuccess)
    return pb_fail(op->ctx, PB_ERR_HIP, "hipHostMalloc failed");
  op->hstate = (KState*)h;
  if (hipHostMalloc(&h, sizeof(double) * 2, hipHostMallocDefault) != hipSuccess)
    return pb_fail(op->ctx, PB_ERR_HIP, "hipHostMalloc failed");
  op->hred = (double*)h;
  hipLaunchKernelGGL(k_fill_const, dim3(elem_blocks(op->n)), dim3(256), 0, op->ctx->own_stream, op->n, op->zero,
                     0.0);
  PB_HIP(op->ctx, hipStreamSynchronize(op->ctx->own_stream));
  op->have_work = true;
  return PB_OK;
}

// kernel dispatch over (CROSS, CENTER) and the tile height RY = op->ry
// (2, 3, 4, 5, 6, 8 rows per thread; see pb_op2d_set_tile)
#define PB_DISPATCH_RY(KERNEL, T, RYV, grd, s, ...)                                                    \
  do {                                                                                                \
    const dim3 _blk(TX, TYT);                                                                         \
    if (_c && _h) hipLaunchKernelGGL((KERNEL<T, true, true, RYV>), grd, _blk, 0, s, __VA_ARGS__);     \
    else if (_c) hipLaunchKernelGGL((KERNEL<T, true, false, RYV>), grd, _blk, 0, s, __VA_ARGS__);     \
    else if (_h) hipLaunchKernelGGL((KERNEL<T, false, true, RYV>), grd, _blk, 0, s, __VA_ARGS__);     \
    else hipLaunchKernelGGL((KERNEL<T, false, false, RYV>), grd, _blk, 0, s, __VA_ARGS__);            \
  } while (0)

#define PB_RY_SWITCH(RYVAL, M, ...)       \
  do {                                   \
    switch (RYVAL) {                     \
      case 2: M(2, __VA_ARGS__); break;  \
      case 3: M(3, __VA_ARGS__); break;  \
      case 5: M(5, __VA_ARGS__); break;  \
      case 6: M(6, __VA_ARGS__); break;  \
      case 8: M(8, __VA_ARGS__); break;  \
      default: M(4, __VA_ARGS__); break; \
    }                                    \
  } while (0)
#define PB_DISPATCH_AT(RYV, KERNEL, T, grd, s, ...) PB_DISPATCH_RY(KERNEL, T, RYV, grd, s, __VA_ARGS__)
#define PB_DISPATCH(KERNEL, T, op, grd, s, ...)                                  \
  do {                                                                          \
    const bool _c = (op)->cross, _h = (op)->center;                            \
    PB_RY_SWITCH((op)->ry, PB_DISPATCH_AT, KERNEL, T, grd, s, __VA_ARGS__);     \
  } while (0)

template <typename T>
IterPtrs<T> iter_ptrs(Work<T>& w, int k) {
  const int po = (k - 1) & 1, pn = k & 1;
  return IterPtrs<T>{w.r[po], w.p[po], w.v[po], w.t, w.rt, w.r[pn], w.p[pn], w.v[pn], w.x};
}

template <typename T>
void launch_a(pb_op2d* op, Work<T>& w, int k, hipStream_t s) {
  const dim3 grd = stencil_grid(op, op->ry);
  const OpDev<T> d = opdev<T>(op);
  PB_DISPATCH(k_iter_a, T, op, grd, s, d, iter_ptrs(w, k), (const KState*)w.S, w.S + 1, (const double*)w.PB, w.PA);
}

template <typename T>
void launch_b(pb_op2d* op, Work<T>& w, int k, hipStream_t s, int sin = 1, int sout = 0, bool exact = false) {
  const dim3 grd = stencil_grid(op, op->ry);
  const OpDev<T> d = opdev<T>(op);
  const int pn = k & 1;
  const T* r = exact ? w.r[0] : w.r[pn];   // the exact variant keeps r in place
  PB_DISPATCH(k_iter_b, T, op, grd, s, d, r, (const T*)w.v[pn], (const T*)w.rt, w.t,
              (const KState*)(w.S + sin), w.S + sout, (const double*)w.PA, w.PB);
}

// exact variant, iteration k: L1 reads S2 writes S0, L2 reads S0 writes S1,
// L3 reads S1 writes S2.  r lives in w.r[0]; p and v alternate with k.
template <typename T>
void launch_l1(pb_op2d* op, Work<T>& w, int k, hipStream_t s) {
  const dim3 grd = stencil_grid(op, op->ry);
  const OpDev<T> d = opdev<T>(op);
  const int po = (k - 1) & 1, pn = k & 1;
  PB_DISPATCH(k_ex_p, T, op, grd, s, d, (const T*)w.r[0], (const T*)w.p[po], (const T*)w.v[po], w.p[pn],
              w.v[pn], (const T*)w.rt, (const KState*)(w.S + 2), w.S, (const double*)w.PC, w.PA);
}

template <typename T>
void launch_l3(pb_op2d* op, Work<T>& w, int k, hipStream_t s) {
  const dim3 grd = stencil_grid(op, op->ry), blk(TX, TYT);
  const int pn = k & 1;
#define PB_L3_AT(RYV, unused)                                                                                     \
  hipLaunchKernelGGL((k_ex_update<T, RYV>), grd, blk, 0, s, op->g.nx, op->g.ny, w.r[0], (const T*)w.p[pn],           \
                     (const T*)w.v[pn], (const T*)w.t, (const T*)w.rt, w.x, (const KState*)(w.S + 1), w.S + 2,       \
                     (const double*)w.PB, w.PC)
  PB_RY_SWITCH(op->ry, PB_L3_AT, 0);
#undef PB_L3_AT
}

// iteration variants of a Krylov run
enum { VAR_FAST = 0, VAR_EXACT = 1, VAR_FUSED = 2 };

// tile height of the one-launch iteration: 5 rows per thread (20-row tiles;
// 209-249 VGPRs, two waves per SIMD, no spills)
constexpr int kFusedRY = 5;
inline int fused_ry(const pb_op2d*) { return kFusedRY; }

#define PB_FUSED_RY(KERNEL, RYV, grd, s, ...)                                                     \
  do {                                                                                          \
    const dim3 _blk(TX, TYT);                                                                   \
    if (_c && _h) hipLaunchKernelGGL((KERNEL<true, true, RYV>), grd, _blk, 0, s, __VA_ARGS__);  \
    else if (_c) hipLaunchKernelGGL((KERNEL<true, false, RYV>), grd, _blk, 0, s, __VA_ARGS__);  \
    else if (_h) hipLaunchKernelGGL((KERNEL<false, true, RYV>), grd, _blk, 0, s, __VA_ARGS__);  \
    else hipLaunchKernelGGL((KERNEL<false, false, RYV>), grd, _blk, 0, s, __VA_ARGS__);         \
  } while (0)
#define PB_FUSED_DISPATCH(KERNEL, op, grd, s, ...)                         \
  do {                                                                    \
    const bool _c = (op)->cross, _h = (op)->center;                       \
    PB_FUSED_RY(KERNEL, kFusedRY, grd, s, __VA_ARGS__);                   \
  } while (0)

// u = A^T r~ after the init launch (reads S0 and the init records in PB,
// writes S1 and the records PF[0])
void launch_adj(pb_op2d* op, Work<double>& w, hipStream_t s) {
  const dim3 grd = stencil_grid(op, fused_ry(op));
  const OpDev<double> d = opdev<double>(op);
  const int px = op->g.bc_x == PB_BC_PERIODIC, py = op->g.bc_y == PB_BC_PERIODIC;
  PB_FUSED_DISPATCH(k_adj_f, op, grd, s, d, (const double*)w.rt, w.u, px, py, (const KState*)w.S, w.S + 1,
                    (const double*)w.PB, w.PF[0]);
}

// F(k): reads the vectors of set (k-1)&1, state S[k&1] and records PF[(k-1)&1];
// writes set k&1, S[(k+1)&1] and PF[k&1] (s lives in r, t in t / t2)
void launch_f(pb_op2d* op, Work<double>& w, int k, hipStream_t s) {
  const dim3 grd = stencil_grid(op, fused_ry(op));
  const OpDev<double> d = opdev<double>(op);
  const int a = (k - 1) & 1, b = k & 1;
  FPtrs q{w.r[a], a ? w.t2 : w.t, w.p[a], w.v[a], w.r[b], b ? w.t2 : w.t, w.p[b], w.v[b], w.rt, w.u, w.x};
  PB_FUSED_DISPATCH(k_iter_f, op, grd, s, d, q, (const KState*)(w.S + (k & 1)), w.S + ((k + 1) & 1),
                    (const double*)w.PF[a], w.PF[b]);
}

// chunk of c iterations starting at an odd k (the buffer parity pattern
// repeats every two iterations).  fast: B(k), A(k+1); exact: L2(k), L3(k),
// L1(k+1); fused: F(k+1)
template <typename T>
void launch_chunk(pb_op2d* op, Work<T>& w, int c, int variant, hipStream_t s) {
  for (int i = 0; i < c; ++i) {
    const int k = 1 + i;
    if (variant == VAR_EXACT) {
      launch_b<T>(op, w, k, s, 0, 1, true);
      launch_l3<T>(op, w, k, s);
      launch_l1<T>(op, w, k + 1, s);
    } else if (variant == VAR_FUSED) {
      if constexpr (std::is_same<T, double>::value) launch_f(op, w, k + 1, s);
    } else {
      launch_b<T>(op, w, k, s);
      launch_a<T>(op, w, k + 1, s);
    }
  }
}

bool graphs_enabled() {
  const char* e = std::getenv("PYBELLA_AMD_NO_GRAPH");
  return !(e && e[0] == '1');
}

template <typename T>
int run_chunk(pb_op2d* op, Work<T>& w, int c, int variant, hipStream_t s) {
  if (graphs_enabled() && s != nullptr) {
    // captured launches are specialised on the tile height and on the
    // operator's (cross, center) flags, which a re-assembly may change
    const int key = (((c * 3 + variant) * 16 + op->ry) * 2 + (op->cross ? 1 : 0)) * 2 + (op->center ? 1 : 0);
    auto it = w.graphs.find(key);
    if (it == w.graphs.end()) {
      hipGraph_t gr = nullptr;
      hipGraphExec_t ex = nullptr;
      PB_HIP(op->ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      launch_chunk<T>(op, w, c, variant, s);
      PB_HIP(op->ctx, hipStreamEndCapture(s, &gr));
      PB_HIP(op->ctx, hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
      (void)hipGraphDestroy(gr);
      it = w.graphs.emplace(key, ex).first;
    }
    PB_HIP(op->ctx, hipGraphLaunch(it->second, s));
  } else {
    launch_chunk<T>(op, w, c, variant, s);
  }
  PB_LAUNCH_CHECK(op->ctx);
  return PB_OK;
}

__global__ void k_put_state(KState* dst, KState v) { *dst = v; }

// One Krylov run: solves A x = b (b, x internal T / FP64 buffers of ``w``).
// x0: optional FP64 initial guess (device); result in w.x.  ``host``: where
// the state is read back to (two records); null = the operator's pinned
// mirror, with the chunks replayed as cached hipGraphs.  With a caller's
// buffer the chunks are launched directly, so the run allocates nothing.
template <typename T>
int krylov_run(pb_op2d* op, Work<T>& w, const T* b, const double* x0, double tol, int maxiter, int legacy,
               int check_every, hipStream_t s, KState* out, int variant = VAR_FAST, KState* host = nullptr) {
  pb_ctx* ctx = op->ctx;
  if (variant == VAR_FUSED && !std::is_same<T, double>::value)
    return pb_fail(ctx, PB_ERR_ARG, "bicgstab: the one-launch iteration is FP64 only");
  const bool exact = variant == VAR_EXACT, fused = variant == VAR_FUSED;
  KState* hs = host ? host : op->hstate;
  KState init{};
  init.tol = tol;
  init.maxiter = maxiter;
  init.legacy = legacy;
  init.status = ST_RUN;
  init.phase = PH_INIT;
  // the first launch of an iteration reads S0 (fast) or S2 (exact); the
  // record travels as a kernel argument
  hipLaunchKernelGGL(k_put_state, dim3(1), dim3(1), 0, s, w.S + (exact ? 2 : 0), init);
  {
    // the one-launch iteration's adjoint launch sums the init records over
    // its own tiling
    const int iry = fused ? fused_ry(op) : op->ry;
    const dim3 grd = stencil_grid(op, iry);
    const OpDev<T> d = opdev<T>(op);
    InitIn<T> in{x0, x0 != nullptr};
    InitOut<T> o{b, w.r[0], w.rt, w.p[0], w.v[0], w.t, w.x, x0, x0 != nullptr, 0.0, 0.0};
    const bool _c = op->cross, _h = op->center;
    if (exact) PB_RY_SWITCH(iry, PB_DISPATCH_AT, k_init2, T, grd, s, d, in, o, w.PC);
    else PB_RY_SWITCH(iry, PB_DISPATCH_AT, k_init5, T, grd, s, d, in, o, w.PB);
  }
  if (exact) {
    launch_l1<T>(op, w, 1, s);
  } else if (fused) {
    if constexpr (std::is_same<T, double>::value) {
      launch_adj(op, w, s);
      launch_f(op, w, 1, s);
    }
  } else {
    launch_a<T>(op, w, 1, s);
  }
  PB_LAUNCH_CHECK(ctx);
  const int chk = exact ? 0 : (fused ? 0 : 1);   // slot written by the iteration's first launch
  const int chunk = std::max(2, (check_every > 0 ? check_every : 32) & ~1);
  long issued = 1;
  for (;;) {
    PB_HIP(ctx, hipMemcpyAsync(hs + 1, w.S + chk, sizeof(KState), hipMemcpyDeviceToHost, s));
    PB_HIP(ctx, hipStreamSynchronize(s));
    const KState h = hs[1];
    if (st_final(h.status)) {
      *out = h;
      return PB_OK;
    }
    if (issued > (long)maxiter + 2L * chunk + 2)
      return pb_fail(ctx, PB_ERR_STATE, "bicgstab: %ld iterations issued without a final state (status %d)",
                     issued, h.status);
    if (host) {
      launch_chunk<T>(op, w, chunk, variant, s);
      PB_LAUNCH_CHECK(ctx);
    } else {
      int rc = run_chunk<T>(op, w, chunk, variant, s);
      if (rc) return rc;
    }
    issued += chunk;
  }
}

// Buffers of a true-residual evaluation: r = b - A x (n), block partials
// (one per tile of the 4-row tiling), the device sum and its host copy.
struct TrBufs {
  double* r;
  double* PT;
  double* red;
  double* hred;
};

int true_residual_in(pb_op2d* op, const TrBufs& tb, const double* b, const double* x, hipStream_t s, double* nrm) {
  const dim3 grd = stencil_grid(op, 4), blk(TX, TYT);
  const OpDev<double> d = opdev<double>(op);
  const bool c = op->cross, h = op->center;
  if (c && h) hipLaunchKernelGGL((k_true_res<true, true, 4>), grd, blk, 0, s, d, x, b, tb.r, tb.PT);
  else if (c) hipLaunchKernelGGL((k_true_res<true, false, 4>), grd, blk, 0, s, d, x, b, tb.r, tb.PT);
  else if (h) hipLaunchKernelGGL((k_true_res<false, true, 4>), grd, blk, 0, s, d, x, b, tb.r, tb.PT);
  else hipLaunchKernelGGL((k_true_res<false, false, 4>), grd, blk, 0, s, d, x, b, tb.r, tb.PT);
  hipLaunchKernelGGL(k_sum1, dim3(1), dim3(TX, TYT), 0, s, (const double*)tb.PT, (int)(grd.x * grd.y), tb.red);
  PB_LAUNCH_CHECK(op->ctx);
  PB_HIP(op->ctx, hipMemcpyAsync(tb.hred, tb.red, sizeof(double), hipMemcpyDeviceToHost, s));
  PB_HIP(op->ctx, hipStreamSynchronize(s));
  *nrm = sqrt(tb.hred[0]);
  return PB_OK;
}

int true_residual(pb_op2d* op, const double* b, const double* x, hipStream_t s, double* nrm) {
  return true_residual_in(op, TrBufs{op->rtrue, op->PT, op->red, op->hred}, b, x, s, nrm);
}

// Iteration cap of one FP32 inner run of the mixed mode: FP32 BiCGStab can
// stagnate for thousands of iterations depending on the rounding order, and
// restarting it from the FP64 true residual (the next refinement step) breaks
// the stagnation.  2 sqrt(n) is ~2.5x the iterations an inner run to 1e-3
// takes on the stratified 1025x513 solve.
int inner_cap(const pb_op2d* op, int maxiter) {
  const int cap = std::max(100, (int)(2.0 * std::sqrt((double)op->n)));
  return std::min(maxiter, cap);
}

int info_of(const KState& h) {
  switch (h.status) {
    case ST_CONV: return 0;
    case ST_MAXIT: return std::max(h.iter, 1);
    case ST_BRK_RHO: return -10;
    case ST_BRK_OMEGA: return -11;
    default: return std::max(h.iter, 1);
  }
}

}  // namespace

// FP64 true residual norm for other translation units (op->rtrue = b - A x)
double pb2d_true_residual(pb_op2d* op, const double* b, const double* x, hipStream_t s, int* rc) {
  double nrm = 0.0;
  *rc = true_residual(op, b, x, s, &nrm);
  return nrm;
}

int pb_nn_solve(pb_op2d* op, const double* b, double* x, const pb_solve_opts* opts, pb_solve_info* info,
                double (*true_res)(pb_op2d*, const double*, const double*, hipStream_t, int*), hipStream_t s);

namespace {

// the stream solves run on: the caller's when it can be captured, else our own
hipStream_t solve_stream(pb_ctx* ctx) {
  if (ctx->stream != nullptr) return ctx->stream;
  (void)hipStreamSynchronize(nullptr);
  return ctx->own_stream;
}

}  // namespace

namespace {

// FP64 finish of the mixed mode from the current iterate x (op->rtrue holds
// b - A x, rn its norm): BiCGStab on A d = r_true to ||r|| <= target, x += d.
// A run that ends short of the target (a breakdown near convergence, where
// rho = r~.r falls below the absolute eps^2 test, or a recursive residual
// that drifted from the true one) is restarted from the new true residual,
// at most 8 times; a correction that does not reduce ||b - A x|| is undone.
int fp64_finish(pb_op2d* op, const double* b, double* x, double target, int maxiter, int check_every,
                hipStream_t s, double* rn, int* iters, int* last_info) {
  pb_ctx* ctx = op->ctx;
  const long n = op->n;
  const int nb = elem_blocks(n);
  int left = maxiter;
  for (int restart = 0; restart < 8 && *rn > target && left > 0; ++restart) {
    KState h64{};
    PB_HIP(ctx, hipMemcpyAsync(op->xcorr, op->rtrue, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    int rc = krylov_run<double>(op, op->w64, op->xcorr, nullptr, target / std::max(*rn, 1e-300), left, 0,
                                check_every, s, &h64);
    if (rc) return rc;
    *iters += h64.iter;
    left -= std::max(h64.iter, 1);
    *last_info = info_of(h64);
    const double old = *rn;
    hipLaunchKernelGGL(k_axpy, dim3(nb), dim3(256), 0, s, n, 1.0, (const double*)op->w64.x, x);
    rc = true_residual(op, b, x, s, rn);
    if (rc) return rc;
    if (!(*rn < old)) {
      hipLaunchKernelGGL(k_axpy, dim3(nb), dim3(256), 0, s, n, -1.0, (const double*)op->w64.x, x);
      rc = true_residual(op, b, x, s, rn);
      if (rc) return rc;
      break;
    }
  }
  return PB_OK;
}

}  // namespace

extern "C" int pb_bicgstab_solve(pb_op2d* op, const double* b, double* x, const pb_solve_opts* opts,
                                 pb_solve_info* info) {
  if (!op) return PB_ERR_ARG;
  pb_ctx* ctx = op->ctx;
  if (!b || !x || !opts || !info) return pb_fail(ctx, PB_ERR_ARG, "pb_bicgstab_solve: null argument");
  if (!op->assembled) return pb_fail(ctx, PB_ERR_STATE, "pb_bicgstab_solve: operator not assembled");
  // PB_SOLVE_NN accepts a zero flexible-GMRES budget: initial guess M(b), then BiCGStab
  if (!(opts->tol >= 0) || opts->maxiter < (opts->mode == PB_SOLVE_NN ? 0 : 1))
    return pb_fail(ctx, PB_ERR_ARG, "pb_bicgstab_solve: bad tol/maxiter");
  int rc = alloc_work(op);
  if (rc) return rc;
  hipStream_t s = solve_stream(ctx);
  const long n = op->n;
  const int nb = elem_blocks(n);
  pb_solve_info out{};
  if (opts->mode == PB_SOLVE_F64 || opts->mode == PB_SOLVE_F64_FAST || opts->mode == PB_SOLVE_F64_FUSED) {
    KState h{};
    const int var = opts->mode == PB_SOLVE_F64 ? VAR_EXACT : (opts->mode == PB_SOLVE_F64_FUSED ? VAR_FUSED : VAR_FAST);
    rc = krylov_run<double>(op, op->w64, b, opts->x0_given ? x : nullptr, opts->tol, opts->maxiter,
                            opts->x0_given ? 0 : 1, opts->check_every, s, &h, var);
    if (rc) return rc;
    if (h.status == ST_SEAM) {   // duplicated end nodes differ: the two-launch iteration
      out.fallback = 1;
      rc = krylov_run<double>(op, op->w64, b, opts->x0_given ? x : nullptr, opts->tol, opts->maxiter,
                              opts->x0_given ? 0 : 1, opts->check_every, s, &h, VAR_FAST);
      if (rc) return rc;
    }
    PB_HIP(ctx, hipMemcpyAsync(x, op->w64.x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    out.info = info_of(h);
    out.iters = h.iter;
    out.outer = 1;
    out.half_step = h.half;
    out.resid = h.resid;
    out.bnrm2 = h.bnrm2;
  } else if (opts->mode == PB_SOLVE_MIXED) {
    const double inner = opts->inner_tol > 0 ? opts->inner_tol : 1e-5;
    const int max_outer = opts->max_outer > 0 ? opts->max_outer : 10;
    if (!opts->x0_given) hipLaunchKernelGGL(k_fill_const, dim3(nb), dim3(256), 0, s, n, x, 0.0);
    double bn = 0.0, rn = 0.0;
    rc = true_residual(op, b, op->zero, s, &bn);
    if (rc) return rc;
    rc = true_residual(op, b, x, s, &rn);   // op->rtrue = b - A x
    if (rc) return rc;
    out.bnrm2 = bn;
    const double target = opts->tol * bn;
    int outer = 0, iters = 0, last_info = 0;
    KState h{};
    bool fallback = false;
    Work<float>& w = op->w32;
    for (outer = 0; outer < max_outer && rn > target; ++outer) {
      hipLaunchKernelGGL((k_cast<double, float>), dim3(nb), dim3(256), 0, s, n, (const double*)op->rtrue, w.b);
      rc = krylov_run<float>(op, w, w.b, nullptr, inner, inner_cap(op, opts->maxiter), 0, opts->check_every, s, &h);
      if (rc) return rc;
      iters += h.iter;
      last_info = info_of(h);
      const double rn_old = rn;
      hipLaunchKernelGGL(k_axpy, dim3(nb), dim3(256), 0, s, n, 1.0, (const double*)w.x, x);
      rc = true_residual(op, b, x, s, &rn);
      if (rc) return rc;
      if (!(rn < rn_old)) {   // no progress: undo the correction
        hipLaunchKernelGGL(k_axpy, dim3(nb), dim3(256), 0, s, n, -1.0, (const double*)w.x, x);
        rc = true_residual(op, b, x, s, &rn);
        if (rc) return rc;
      }
      if (!(rn < 0.5 * rn_old)) {
        fallback = true;
        break;
      }
    }
    if (fallback && rn > target) {   // finish in FP64 from the current iterate
      rc = fp64_finish(op, b, x, target, opts->maxiter, opts->check_every, s, &rn, &iters, &last_info);
      if (rc) return rc;
      ++outer;
    }
    out.info = rn <= target ? 0 : (last_info < 0 ? last_info : std::max(iters, 1));
    out.iters = iters;
    out.outer = outer;
    out.half_step = h.half;
    out.resid = rn;
  } else if (opts->mode == PB_SOLVE_NN) {
    rc = pb_nn_solve(op, b, x, opts, &out, pb2d_true_residual, s);
    if (rc) return rc;
    if (out.info != 0) {
      // budget exhausted: plain BiCGStab (two-launch iteration) from the
      // current iterate, restarted from the new iterate after a breakdown
      // (near convergence rho = r~.r can fall below the absolute eps^2 test)
      const int fmax = opts->fallback_maxiter > 0 ? opts->fallback_maxiter : 6000;
      out.fallback = 1;
      int left = fmax;
      for (int restart = 0; restart < 8 && left > 0; ++restart) {
        KState h{};
        rc = krylov_run<double>(op, op->w64, b, x, opts->tol, left, 0, opts->check_every, s, &h, VAR_FAST);
        if (rc) return rc;
        PB_HIP(ctx, hipMemcpyAsync(x, op->w64.x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        out.iters += h.iter;
        left -= std::max(h.iter, 1);
        out.info = info_of(h);
        out.resid = h.resid;
        if (out.info >= 0) break;
      }
    }
  } else {
    return pb_fail(ctx, PB_ERR_ARG, "pb_bicgstab_solve: mode %d not available on this handle", opts->mode);
  }
  double tr = 0.0;
  rc = true_residual(op, b, x, s, &tr);
  if (rc) return rc;
  out.true_rel_resid = out.bnrm2 > 0 ? tr / out.bnrm2 : tr;
  *info = out;
  return out.info;
}

namespace {
constexpr long kAlignWords = 32;   // 256-byte alignment of every scratch region
inline long align_words(long w) { return (w + kAlignWords - 1) / kAlignWords * kAlignWords; }
constexpr long kStateWords = (long)((3 * sizeof(KState) + sizeof(double) - 1) / sizeof(double));

// Words of static scratch an FP64 solve of ``op`` needs: eight Krylov vectors
// (r, p, v double-buffered, t, r~), the partial records of the finest tiling
// (A 1, B 5, exact-update 2, true residual 1 per tile), the true-residual sum
// and three state records.
long noalloc_words(const pb_op2d* op) {
  const dim3 sg = stencil_grid(op, 2);
  const long nb = (long)sg.x * sg.y;
  return 8 * align_words(op->n) + align_words(nb) + align_words(5 * nb) + align_words(2 * nb) + align_words(nb) +
         align_words(2) + align_words(kStateWords);
}
}  // namespace

extern "C" int pb_bicgstab_solve_noalloc(pb_op2d* op, const double* b, double* x, const pb_solve_opts* opts,
                                         pb_solve_info* info) {
  if (!op) return PB_ERR_ARG;
  pb_ctx* ctx = op->ctx;
  if (!b || !x || !opts || !info) return pb_fail(ctx, PB_ERR_ARG, "pb_bicgstab_solve_noalloc: null argument");
  if (!op->assembled) return pb_fail(ctx, PB_ERR_STATE, "pb_bicgstab_solve_noalloc: operator not assembled");
  if (opts->mode != PB_SOLVE_F64 && opts->mode != PB_SOLVE_F64_FAST)
    return pb_fail(ctx, PB_ERR_ARG, "pb_bicgstab_solve_noalloc: mode %d; only PB_SOLVE_F64 and PB_SOLVE_F64_FAST",
                   opts->mode);
  if (!(opts->tol >= 0) || opts->maxiter < 1)
    return pb_fail(ctx, PB_ERR_ARG, "pb_bicgstab_solve_noalloc: bad tol/maxiter");
  if (!ctx->scratch)
    return pb_fail(ctx, PB_ERR_STATE, "pb_bicgstab_solve_noalloc: static scratch unavailable on device %d (%s)",
                   ctx->device, hipGetErrorString((hipError_t)ctx->scratch_err));
  const long need = noalloc_words(op);
  const long have = (long)(ctx->scratch_bytes / sizeof(double));
  if (need > have)
    return pb_fail(ctx, PB_ERR_ARG,
                   "pb_bicgstab_solve_noalloc: a %d x %d grid needs %ld bytes of scratch, the static scratch holds "
                   "%ld (about %ld nodes at most)",
                   op->g.nx, op->g.ny, need * 8, have * 8, (have - 4096) / 8);
  // carve the scratch: every region is written before it is read within one
  // call (vectors by the init launch, partial records by the launch before
  // their readers, the state record by k_put_state), so nothing is cleared
  double* base = ctx->scratch;
  long off = 0;
  auto take = [&](long words) {
    double* p = base + off;
    off += align_words(words);
    return p;
  };
  const dim3 sg = stencil_grid(op, 2);
  const long nb = (long)sg.x * sg.y;
  const long n = op->n;
  Work<double> w;
  for (int k = 0; k < 2; ++k) {
    w.r[k] = take(n);
    w.p[k] = take(n);
    w.v[k] = take(n);
  }
  w.t = take(n);
  w.rt = take(n);
  w.x = x;   // the iterate lives in the caller's output
  w.PA = take(nb);
  w.PB = take(5 * nb);
  w.PC = take(2 * nb);
  double* PT = take(nb);
  double* red = take(2);
  w.S = reinterpret_cast<KState*>(take(kStateWords));
  hipStream_t s = solve_stream(ctx);
  KState hbuf[2];
  KState h{};
  const bool exact = opts->mode == PB_SOLVE_F64;
  int rc = krylov_run<double>(op, w, b, opts->x0_given ? x : nullptr, opts->tol, opts->maxiter,
                              opts->x0_given ? 0 : 1, opts->check_every, s, &h, exact ? VAR_EXACT : VAR_FAST, hbuf);
  if (rc) return rc;
  pb_solve_info out{};
  out.info = info_of(h);
  out.iters = h.iter;
  out.outer = 1;
  out.half_step = h.half;
  out.resid = h.resid;
  out.bnrm2 = h.bnrm2;
  double hred = 0.0, tr = 0.0;
  rc = true_residual_in(op, TrBufs{w.t, PT, red, &hred}, b, x, s, &tr);   // t is free after the run
  if (rc) return rc;
  out.true_rel_resid = out.bnrm2 > 0 ? tr / out.bnrm2 : tr;
  *info = out;
  return out.info;
}

namespace {
template <typename T>
int time_kernel_t(pb_op2d* op, int kernel, int reps, hipStream_t s, double* avg_us) {
  pb_ctx* ctx = op->ctx;
  Work<T>& w = work_of<T>(op);
  const long n = op->n;
  const int nb = elem_blocks(n);
  const dim3 sg = stencil_grid(op, op->ry);
  const int nbs = (int)(sg.x * sg.y);
  T* vecs[] = {w.r[0], w.r[1], w.p[0], w.p[1], w.v[0], w.v[1], w.t, w.rt, w.b};
  for (int k = 0; k < 9; ++k) hipLaunchKernelGGL(k_fill_pattern<T>, dim3(nb), dim3(256), 0, s, n, vecs[k], 0.7 * k);
  hipLaunchKernelGGL(k_fill_pattern<double>, dim3(nb), dim3(256), 0, s, n, w.x, 0.1);
  // partials that keep the state running: B totals (1, .5, 1, .3, .1), A total 1
  std::vector<double> pb(5L * nbs), pa(nbs, 1.0 / nbs);   // record layout [value][block]
  const double tb[5] = {1.0, 0.5, 1.0, 0.3, 0.1};
  for (int k = 0; k < 5; ++k)
    for (int b = 0; b < nbs; ++b) pb[(long)k * nbs + b] = tb[k] / nbs;
  PB_HIP(ctx, hipMemcpyAsync(w.PB, pb.data(), sizeof(double) * pb.size(), hipMemcpyHostToDevice, s));
  PB_HIP(ctx, hipMemcpyAsync(w.PA, pa.data(), sizeof(double) * pa.size(), hipMemcpyHostToDevice, s));
  KState st{};
  st.rho = 1.0;
  st.alpha = 0.5;
  st.omega = 0.5;
  st.beta = 0.5;
  st.atol = -1.0;
  st.tol = 0.0;
  st.maxiter = 0x7fffffff;
  st.status = ST_RUN;
  st.x_upd = 1;
  KState sts[2] = {st, st};
  sts[0].phase = PH_B;   // read by A
  sts[1].phase = PH_A;   // read by B
  PB_HIP(ctx, hipMemcpyAsync(w.S, sts, sizeof(sts), hipMemcpyHostToDevice, s));
  PB_HIP(ctx, hipStreamSynchronize(s));
  auto launch = [&]() {
    switch (kernel) {
      case 0:
        pb_op2d_matvec(op, w.p[0], w.v[1], sizeof(T) == 8 ? PB_F64 : PB_F32);
        break;
      case 1: launch_a<T>(op, w, 2, s); break;
      case 2: launch_b<T>(op, w, 1, s); break;
      default: {
        const dim3 grd = stencil_grid(op, 4), blk(TX, TYT);
        const OpDev<double> d = opdev<double>(op);
        hipLaunchKernelGGL((k_true_res<false, false, 4>), grd, blk, 0, s, d, w.x, op->zero, op->rtrue, op->PT);
      }
    }
  };
  for (int i = 0; i < 3; ++i) launch();
  hipEvent_t e0, e1;
  PB_HIP(ctx, hipEventCreate(&e0));
  PB_HIP(ctx, hipEventCreate(&e1));
  PB_HIP(ctx, hipEventRecord(e0, s));
  for (int i = 0; i < reps; ++i) launch();
  PB_HIP(ctx, hipEventRecord(e1, s));
  PB_HIP(ctx, hipEventSynchronize(e1));
  float ms = 0.f;
  PB_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  PB_LAUNCH_CHECK(ctx);
  KState chk[2];
  PB_HIP(ctx, hipMemcpy(chk, w.S, sizeof(chk), hipMemcpyDeviceToHost));
  if ((kernel == 1 && chk[1].status != ST_RUN) || (kernel == 2 && chk[0].status != ST_RUN))
    return pb_fail(ctx, PB_ERR_STATE, "pb_op2d_time_kernel: timing state left the running status");
  *avg_us = 1000.0 * ms / reps;
  return PB_OK;
}
}  // namespace

extern "C" int pb_op2d_time_kernel(pb_op2d* op, int kernel, int precision, int reps, double* avg_us) {
  if (!op) return PB_ERR_ARG;
  if (!avg_us || reps < 1 || kernel < 0 || kernel > 3)
    return pb_fail(op->ctx, PB_ERR_ARG, "pb_op2d_time_kernel: bad argument");
  if (!op->assembled) return pb_fail(op->ctx, PB_ERR_STATE, "pb_op2d_time_kernel: operator not assembled");
  int rc = alloc_work(op);
  if (rc) return rc;
  hipStream_t s = solve_stream(op->ctx);
  if (kernel == 0) {
    // the plain matvec runs on the context stream
    s = op->ctx->stream;
  }
  if (precision == PB_F64) return time_kernel_t<double>(op, kernel, reps, s, avg_us);
  if (precision == PB_F32) return time_kernel_t<float>(op, kernel, reps, s, avg_us);
  return pb_fail(op->ctx, PB_ERR_ARG, "pb_op2d_time_kernel: unknown precision %d", precision);
}

// Average device time of the A (kernel 1) or B (kernel 2) launch inside a real
// FP64 two-launch solve of b (x0 = 0, tolerance tol): the solve's iterations
// are launched directly, and at the start of each chunk of iterations the
// chosen launch is repeated back to back on the live state between two HIP
// events on the solve stream; *launches = how many were timed.  The solve's
// result is the fast mode's (same kernels and order); x receives it.
extern "C" int pb_op2d_time_in_solve(pb_op2d* op, const double* b, double* x, double tol, int kernel,
                                     double* avg_us, int* launches) {
  if (!op) return PB_ERR_ARG;
  pb_ctx* ctx = op->ctx;
  if (!b || !x || !avg_us || !launches || (kernel != 1 && kernel != 2) || !(tol > 0))
    return pb_fail(ctx, PB_ERR_ARG, "pb_op2d_time_in_solve: bad argument");
  if (!op->assembled) return pb_fail(ctx, PB_ERR_STATE, "pb_op2d_time_in_solve: operator not assembled");
  int rc = alloc_work(op);
  if (rc) return rc;
  hipStream_t s = solve_stream(ctx);
  Work<double>& w = op->w64;
  KState init{};
  init.tol = tol;
  init.maxiter = 20000;
  init.legacy = 1;
  init.status = ST_RUN;
  init.phase = PH_INIT;
  hipLaunchKernelGGL(k_put_state, dim3(1), dim3(1), 0, s, w.S, init);
  {
    const dim3 grd = stencil_grid(op, op->ry);
    const OpDev<double> d = opdev<double>(op);
    InitIn<double> in{nullptr, 0};
    InitOut<double> o{b, w.r[0], w.rt, w.p[0], w.v[0], w.t, w.x, nullptr, 0, 0.0, 0.0};
    PB_DISPATCH(k_init5, double, op, grd, s, d, in, o, w.PB);
  }
  // At the start of every chunk whose residual is still far above the
  // target, the chosen launch is repeated kBurst times back to back between
  // two events, in its place in the iteration: A and B write the same
  // outputs from the same inputs each time, except A's in-place x update,
  // which is saved before the burst and restored after it.  So the events
  // time the kernel itself on the live state, without a gap per launch.
  constexpr int kBurst = 8;
  std::vector<hipEvent_t> ev;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> bursts;
  auto burst = [&](bool is_a, int k) -> int {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    PB_HIP(ctx, hipEventCreate(&e0));
    ev.push_back(e0);
    PB_HIP(ctx, hipEventCreate(&e1));
    ev.push_back(e1);
    if (is_a) PB_HIP(ctx, hipMemcpyAsync(op->xcorr, w.x, sizeof(double) * op->n, hipMemcpyDeviceToDevice, s));
    PB_HIP(ctx, hipEventRecord(e0, s));
    for (int r = 0; r < kBurst; ++r) {
      if (is_a) launch_a<double>(op, w, k, s);
      else launch_b<double>(op, w, k, s);
    }
    PB_HIP(ctx, hipEventRecord(e1, s));
    if (is_a) PB_HIP(ctx, hipMemcpyAsync(w.x, op->xcorr, sizeof(double) * op->n, hipMemcpyDeviceToDevice, s));
    bursts.emplace_back(e0, e1);
    return PB_OK;
  };
  launch_a<double>(op, w, 1, s);
  KState h{};
  const int chunk = 32;
  for (long issued = 1;; issued += chunk) {
    PB_HIP(ctx, hipMemcpyAsync(op->hstate + 1, w.S + 1, sizeof(KState), hipMemcpyDeviceToHost, s));
    PB_HIP(ctx, hipStreamSynchronize(s));
    h = op->hstate[1];
    if (st_final(h.status)) break;
    if (issued > 20000 + 2L * chunk)
      return pb_fail(ctx, PB_ERR_STATE, "pb_op2d_time_in_solve: no final state after %ld iterations", issued);
    // one iteration cannot take the residual from 100x the target to the exit
    const bool live = h.status == ST_RUN && h.resid > 100.0 * h.atol;
    for (int i = 0; i < chunk; ++i) {   // B(k), A(k + 1) from an odd k: the parity pattern of launch_chunk
      const int k = 1 + i;
      if (i == 0 && live && kernel == 2 && (rc = burst(false, k))) return rc;
      launch_b<double>(op, w, k, s);
      if (i == 0 && live && kernel == 1 && (rc = burst(true, k + 1))) return rc;
      launch_a<double>(op, w, k + 1, s);
    }
    PB_LAUNCH_CHECK(ctx);
  }
  int ntimed = 0;
  double total_ms = 0.0;
  for (const auto& pr : bursts) {
    float ms = 0.f;
    PB_HIP(ctx, hipEventElapsedTime(&ms, pr.first, pr.second));
    total_ms += ms;
    ntimed += kBurst;
  }
  for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  PB_HIP(ctx, hipMemcpyAsync(x, w.x, sizeof(double) * op->n, hipMemcpyDeviceToDevice, s));
  PB_HIP(ctx, hipStreamSynchronize(s));
  if (h.status != ST_CONV) return pb_fail(ctx, PB_ERR_STATE, "pb_op2d_time_in_solve: solve ended with status %d", h.status);
  *launches = ntimed;
  *avg_us = ntimed ? 1e3 * total_ms / ntimed : 0.0;
  return PB_OK;
}

extern "C" int pb_op2d_set_tile(pb_op2d* op, int rows_per_thread) {
  if (!op) return PB_ERR_ARG;
  if (rows_per_thread < 2 || rows_per_thread > 8 || rows_per_thread == 7)
    return pb_fail(op->ctx, PB_ERR_ARG, "pb_op2d_set_tile: rows per thread must be 2, 3, 4, 5, 6 or 8");
  op->ry = rows_per_thread;
  return PB_OK;
}

// ---------------------------------------------------------------------------
// Batched uneven solves (ensemble members, possibly of different grid sizes).
// Every member keeps its own operator, workspace, tile height and kernels, so
// its iterates are those of a single solve.  A chunk of iterations of ALL
// members is captured as one hipGraph: a fork from the caller's stream into
// one branch per member stream, joined back at the end.  Each replay advances
// every member by the chunk; members that have finished run no-op launches
// (their state is final).  The host reads all members' states once per chunk.
// No reference counterpart: the reference solves ensemble members one after
// the other (src/__main__.py:223-254).
// ---------------------------------------------------------------------------
struct pb_batch {
  pb_ctx* ctx = nullptr;
  std::vector<pb_op2d*> ops;
  void* mtab[2] = {nullptr, nullptr};      // device member tables (FP64, FP32) of the flattened launches
  KState* dstates = nullptr;               // device gather target, one per member
  std::vector<char> htab[2];               // host images of the member tables
  std::map<long, hipGraphExec_t> flat_graphs;
  int flat_sig = -1;                       // (ry, cross, center) the flat graphs were captured with
  std::vector<hipStream_t> streams;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> joins;
  std::map<long, hipGraphExec_t> graphs;   // key: precision, chunk
  std::vector<int> ry_captured;            // member launch signatures the graphs were captured with
  KState* hst = nullptr;                   // pinned, one per member
  // a user-facing batch is split into groups (each a batch of its own) whose
  // working sets fit the Infinity Cache together; groups are solved in turn
  std::vector<pb_batch*> groups;
  std::vector<int> first;                  // index of each group's first member
};

namespace {

// ---- flattened launches: one grid over the tiles of every member ----------
// Block b (after the XCD-aware remap of the whole grid) belongs to the member
// whose tile range [begin, begin + nbx * nby) contains it; within the member it is tile (local % nbx, local / nbx) with
// block number ``local``, exactly the numbering of the member's own launch, so
// partial sums and iterates are those of a single solve.  All members share
// the template parameters (tile height, cross, center) of the launch.
template <typename T>
struct BMember {
  OpDev<T> op;
  const T* b;
  T* r[2];
  T* p[2];
  T* v[2];
  T* t;
  T* rt;
  double* x;
  KState* S;
  double* PA;
  double* PB;
  int begin, nbx, nby, pad;
};

// Device view of a BMember with every pointer typed in the global address
// space (same layout).  Pointers read from the member table are otherwise
// generic, and generic (flat) loads also count against the scalar-load
// counter, so each wait for a scalar load would wait for the whole tile.
#define PB_GLOBAL __attribute__((address_space(1)))
template <typename T>
struct BMemberG {
  PB_GLOBAL const T *wxx, *wyy, *wyx, *wxy, *hc, *dinv;
  int nx, ny;
  PB_GLOBAL const T* b;
  PB_GLOBAL T* r[2];
  PB_GLOBAL T* p[2];
  PB_GLOBAL T* v[2];
  PB_GLOBAL T* t;
  PB_GLOBAL T* rt;
  PB_GLOBAL double* x;
  PB_GLOBAL KState* S;
  PB_GLOBAL double* PA;
  PB_GLOBAL double* PB;
  int begin, nbx, nby, pad;
};
static_assert(sizeof(BMemberG<double>) == sizeof(BMember<double>), "member view layout");
static_assert(sizeof(BMemberG<float>) == sizeof(BMember<float>), "member view layout");
static_assert(offsetof(BMemberG<double>, begin) == offsetof(BMember<double>, begin), "member view layout");

template <typename P>
__device__ __forceinline__ P* gen(PB_GLOBAL P* p) {
  return (P*)p;
}

template <typename T>
__device__ __forceinline__ OpDev<T> member_op(const BMemberG<T>& G) {
  return OpDev<T>{gen(G.wxx), gen(G.wyy), gen(G.wyx), gen(G.wxy), gen(G.hc), gen(G.dinv), G.nx, G.ny};
}

template <typename T>
__device__ __forceinline__ int find_member(const BMember<T>* mem, int nmem, int blk) {
  int lo = 0, hi = nmem - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (mem[mid].begin <= blk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <typename T, bool CROSS, bool CENTER, int RY>
__global__ __launch_bounds__(NTHR) void k_init5_batch(const BMember<T>* __restrict__ mem, int nmem) {
  const int g = xcd_tile(blockIdx.x, gridDim.x);
  const BMemberG<T>& M = reinterpret_cast<const BMemberG<T>&>(mem[find_member(mem, nmem, g)]);
  if (st_final(M.S[0].status)) return;   // parked member (uniform over the block)
  const int local = g - M.begin;
  OpDev<T> op = member_op(M);
  InitIn<T> in{nullptr, 0};
  InitOut<T> o{gen(M.b), gen(M.r[0]), gen(M.rt), gen(M.p[0]), gen(M.v[0]), gen(M.t), gen(M.x), nullptr, 0, 0.0, 0.0};
  init_body<5, T, CROSS, CENTER, RY>(op, in, o, gen(M.PB), local % M.nbx, local / M.nbx, local, M.nbx * M.nby);
}

template <typename T, bool CROSS, bool CENTER, int RY>
__global__ __launch_bounds__(NTHR) void k_iter_a_batch(const BMember<T>* __restrict__ mem, int nmem, int k) {
  const int g = xcd_tile(blockIdx.x, gridDim.x);
  const BMemberG<T>& M = reinterpret_cast<const BMemberG<T>&>(mem[find_member(mem, nmem, g)]);
  const int local = g - M.begin, po = (k - 1) & 1, pn = k & 1;
  const IterPtrs<T> q{gen(M.r[po]), gen(M.p[po]), gen(M.v[po]), gen(M.t), gen(M.rt),
                      gen(M.r[pn]), gen(M.p[pn]), gen(M.v[pn]), gen(M.x)};
  iter_a_body<T, CROSS, CENTER, RY>(member_op(M), q, gen(M.S), gen(M.S) + 1, gen(M.PB), gen(M.PA), local % M.nbx,
                                    local / M.nbx, local, M.nbx * M.nby);
}

template <typename T, bool CROSS, bool CENTER, int RY>
__global__ __launch_bounds__(NTHR) void k_iter_b_batch(const BMember<T>* __restrict__ mem, int nmem, int k) {
  const int g = xcd_tile(blockIdx.x, gridDim.x);
  const BMemberG<T>& M = reinterpret_cast<const BMemberG<T>&>(mem[find_member(mem, nmem, g)]);
  const int local = g - M.begin, pn = k & 1;
  iter_b_body<T, CROSS, CENTER, RY>(member_op(M), gen(M.r[pn]), gen(M.v[pn]), gen(M.rt), gen(M.t), gen(M.S) + 1,
                                    gen(M.S), gen(M.PA), gen(M.PB), local % M.nbx, local / M.nbx, local,
                                    M.nbx * M.nby);
}

template <typename T>
__global__ void k_gather_states(const BMember<T>* mem, int nmem, int slot, KState* out) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < nmem) out[m] = mem[m].S[slot];
}

// launch helper over the batch's uniform (cross, center, ry)
#define PB_FLAT_AT(RYV, KERNEL, T, cross, center, grd, s, ...)                                          \
  do {                                                                                                 \
    const dim3 _blk(TX, TYT);                                                                          \
    if (cross && center) hipLaunchKernelGGL((KERNEL<T, true, true, RYV>), grd, _blk, 0, s, __VA_ARGS__); \
    else if (cross) hipLaunchKernelGGL((KERNEL<T, true, false, RYV>), grd, _blk, 0, s, __VA_ARGS__);     \
    else if (center) hipLaunchKernelGGL((KERNEL<T, false, true, RYV>), grd, _blk, 0, s, __VA_ARGS__);    \
    else hipLaunchKernelGGL((KERNEL<T, false, false, RYV>), grd, _blk, 0, s, __VA_ARGS__);               \
  } while (0)
#define PB_FLAT(KERNEL, T, cross, center, ry, grd, s, ...) \
  PB_RY_SWITCH(ry, PB_FLAT_AT, KERNEL, T, cross, center, grd, s, __VA_ARGS__)

// the batch can use the flattened launches when every member has the same
// tile height and operator kind; returns the (ry, cross, center) signature or -1
int flat_signature(const pb_batch* bt) {
  const pb_op2d* o0 = bt->ops[0];
  for (const pb_op2d* op : bt->ops)
    if (op->ry != o0->ry || op->cross != o0->cross || op->center != o0->center) return -1;
  return o0->ry * 4 + (o0->cross ? 2 : 0) + (o0->center ? 1 : 0);
}

template <typename T>
int krylov_batch_flat(pb_batch* bt, const std::vector<char>& active, const std::vector<const T*>& b,
                      const std::vector<double>& tol, int maxiter, int legacy, int check_every, hipStream_t s,
                      std::vector<KState>& out) {
  const int P = std::is_same<T, float>::value ? 1 : 0;
  const int M = (int)bt->ops.size();
  const pb_op2d* o0 = bt->ops[0];
  const int ry = o0->ry;
  const bool cross = o0->cross, center = o0->center;
  const int sig = flat_signature(bt);
  if (sig != bt->flat_sig) {   // captured launches are specialised on the signature
    for (auto& kv : bt->flat_graphs) (void)hipGraphExecDestroy(kv.second);
    bt->flat_graphs.clear();
    bt->flat_sig = sig;
  }
  // member table (device pointers are fixed for the life of the workspaces)
  std::vector<BMember<T>> tab((size_t)M);
  int total = 0;
  for (int m = 0; m < M; ++m) {
    pb_op2d* op = bt->ops[m];
    Work<T>& w = work_of<T>(op);
    const dim3 g = stencil_grid(op, ry);
    BMember<T>& e = tab[(size_t)m];
    e.op = opdev<T>(op);
    e.b = b[m];
    for (int q = 0; q < 2; ++q) {
      e.r[q] = w.r[q];
      e.p[q] = w.p[q];
      e.v[q] = w.v[q];
    }
    e.t = w.t;
    e.rt = w.rt;
    e.x = w.x;
    e.S = w.S;
    e.PA = w.PA;
    e.PB = w.PB;
    e.begin = total;
    e.nbx = (int)g.x;
    e.nby = (int)g.y;
    e.pad = 0;
    total += (int)(g.x * g.y);
  }
  std::vector<char>& ht = bt->htab[P];
  ht.assign((const char*)tab.data(), (const char*)tab.data() + sizeof(BMember<T>) * tab.size());
  PB_HIP(bt->ctx, hipMemcpyAsync(bt->mtab[P], ht.data(), ht.size(), hipMemcpyHostToDevice, s));
  for (int m = 0; m < M; ++m) {
    KState init{};
    init.tol = tol[(size_t)m];
    init.maxiter = maxiter;
    init.legacy = legacy;
    init.phase = PH_INIT;
    init.status = active[(size_t)m] ? ST_RUN : ST_CONV;
    bt->hst[m] = init;
    Work<T>& w = work_of<T>(bt->ops[m]);
    PB_HIP(bt->ctx, hipMemcpyAsync(w.S, bt->hst + m, sizeof(KState), hipMemcpyHostToDevice, s));
    if (!active[(size_t)m])
      PB_HIP(bt->ctx, hipMemcpyAsync(w.S + 1, bt->hst + m, sizeof(KState), hipMemcpyHostToDevice, s));
  }
  const BMember<T>* dt = (const BMember<T>*)bt->mtab[P];
  const dim3 grd((unsigned)total);
  PB_FLAT(k_init5_batch, T, cross, center, ry, grd, s, dt, M);
  PB_FLAT(k_iter_a_batch, T, cross, center, ry, grd, s, dt, M, 1);
  PB_LAUNCH_CHECK(bt->ctx);
  const int chunk = std::max(2, (check_every > 0 ? check_every : 32) & ~1);
  const bool graphs = graphs_enabled() && s != nullptr;
  auto issue_chunk = [&]() -> int {
    for (int i = 0; i < chunk; ++i) {
      PB_FLAT(k_iter_b_batch, T, cross, center, ry, grd, s, dt, M, 1 + i);
      PB_FLAT(k_iter_a_batch, T, cross, center, ry, grd, s, dt, M, 2 + i);
    }
    return PB_OK;
  };
  hipGraphExec_t ex = nullptr;
  if (graphs) {
    const long key = ((long)chunk * 2 + P) * 65536 + total;
    auto it = bt->flat_graphs.find(key);
    if (it == bt->flat_graphs.end()) {
      hipGraph_t gr = nullptr;
      PB_HIP(bt->ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      issue_chunk();
      PB_HIP(bt->ctx, hipStreamEndCapture(s, &gr));
      PB_HIP(bt->ctx, hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
      (void)hipGraphDestroy(gr);
      it = bt->flat_graphs.emplace(key, ex).first;
    }
    ex = it->second;
  }
  long issued = 1;
  for (;;) {
    hipLaunchKernelGGL(k_gather_states<T>, dim3((M + 63) / 64), dim3(64), 0, s, dt, M, 1, bt->dstates);
    PB_HIP(bt->ctx, hipMemcpyAsync(bt->hst, bt->dstates, sizeof(KState) * M, hipMemcpyDeviceToHost, s));
    PB_HIP(bt->ctx, hipStreamSynchronize(s));
    bool done = true;
    for (int m = 0; m < M; ++m) done = done && st_final(bt->hst[m].status);
    if (done) break;
    if (issued > (long)maxiter + 2L * chunk + 2)
      return pb_fail(bt->ctx, PB_ERR_STATE, "batch bicgstab: %ld iterations issued without final states", issued);
    if (graphs) PB_HIP(bt->ctx, hipGraphLaunch(ex, s));
    else issue_chunk();
    PB_LAUNCH_CHECK(bt->ctx);
    issued += chunk;
  }
  out.assign(bt->hst, bt->hst + M);
  return PB_OK;
}

template <typename T>
int batch_graph(pb_batch* bt, int chunk, hipStream_t s, hipGraphExec_t* out) {
  std::vector<int> ry;   // launch signature of every member: tile height and (cross, center)
  for (pb_op2d* op : bt->ops) ry.push_back(op->ry * 4 + (op->cross ? 2 : 0) + (op->center ? 1 : 0));
  if (ry != bt->ry_captured) {   // a member's tile or operator kind changed: the launches are stale
    for (auto& kv : bt->graphs) (void)hipGraphExecDestroy(kv.second);
    bt->graphs.clear();
    bt->ry_captured = ry;
  }
  const long key = (long)chunk * 2 + (std::is_same<T, float>::value ? 1 : 0);
  auto it = bt->graphs.find(key);
  if (it == bt->graphs.end()) {
    hipGraph_t gr = nullptr;
    hipGraphExec_t ex = nullptr;
    PB_HIP(bt->ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    PB_HIP(bt->ctx, hipEventRecord(bt->fork, s));
    for (size_t m = 0; m < bt->ops.size(); ++m) {
      PB_HIP(bt->ctx, hipStreamWaitEvent(bt->streams[m], bt->fork, 0));
      launch_chunk<T>(bt->ops[m], work_of<T>(bt->ops[m]), chunk, VAR_FAST, bt->streams[m]);
      PB_HIP(bt->ctx, hipEventRecord(bt->joins[m], bt->streams[m]));
      PB_HIP(bt->ctx, hipStreamWaitEvent(s, bt->joins[m], 0));
    }
    PB_HIP(bt->ctx, hipStreamEndCapture(s, &gr));
    PB_HIP(bt->ctx, hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
    (void)hipGraphDestroy(gr);
    it = bt->graphs.emplace(key, ex).first;
  }
  *out = it->second;
  return PB_OK;
}

// One batched Krylov run (fast two-launch iteration, x0 = 0): member m solves
// A_m x = b[m] to tolerance tol[m] (legacy or relative), if active[m].
template <typename T>
int krylov_batch(pb_batch* bt, const std::vector<char>& active, const std::vector<const T*>& b,
                 const std::vector<double>& tol, int maxiter, int legacy, int check_every, hipStream_t s,
                 std::vector<KState>& out) {
  // one launch over every member's tiles when they share the launch signature
  if (flat_signature(bt) >= 0) return krylov_batch_flat<T>(bt, active, b, tol, maxiter, legacy, check_every, s, out);
  const size_t M = bt->ops.size();
  for (size_t m = 0; m < M; ++m) {
    pb_op2d* op = bt->ops[m];
    Work<T>& w = work_of<T>(op);
    KState init{};
    init.tol = tol[m];
    init.maxiter = maxiter;
    init.legacy = legacy;
    init.phase = PH_INIT;
    init.status = active[m] ? ST_RUN : ST_CONV;
    bt->hst[m] = init;
    if (active[m]) {
      PB_HIP(bt->ctx, hipMemcpyAsync(w.S, bt->hst + m, sizeof(KState), hipMemcpyHostToDevice, s));
      const dim3 grd = stencil_grid(op, op->ry);
      const OpDev<T> d = opdev<T>(op);
      InitIn<T> in{nullptr, false};
      InitOut<T> o{b[m], w.r[0], w.rt, w.p[0], w.v[0], w.t, w.x, nullptr, false, 0.0, 0.0};
      PB_DISPATCH(k_init5, T, op, grd, s, d, in, o, w.PB);
      launch_a<T>(op, w, 1, s);
    } else {   // parked: both state slots final, so every launch of the chunk is a no-op
      PB_HIP(bt->ctx, hipMemcpyAsync(w.S, bt->hst + m, sizeof(KState), hipMemcpyHostToDevice, s));
      PB_HIP(bt->ctx, hipMemcpyAsync(w.S + 1, bt->hst + m, sizeof(KState), hipMemcpyHostToDevice, s));
    }
  }
  PB_LAUNCH_CHECK(bt->ctx);
  const int chunk = std::max(2, (check_every > 0 ? check_every : 32) & ~1);
  hipGraphExec_t ex = nullptr;
  const bool graphs = graphs_enabled() && s != nullptr;
  if (graphs) {
    int rc = batch_graph<T>(bt, chunk, s, &ex);
    if (rc) return rc;
  }
  long issued = 1;
  for (;;) {
    for (size_t m = 0; m < M; ++m)
      PB_HIP(bt->ctx, hipMemcpyAsync(bt->hst + m, work_of<T>(bt->ops[m]).S + 1, sizeof(KState),
                                     hipMemcpyDeviceToHost, s));
    PB_HIP(bt->ctx, hipStreamSynchronize(s));
    bool done = true;
    for (size_t m = 0; m < M; ++m) done = done && st_final(bt->hst[m].status);
    if (done) break;
    if (issued > (long)maxiter + 2L * chunk + 2)
      return pb_fail(bt->ctx, PB_ERR_STATE, "batch bicgstab: %ld iterations issued without final states", issued);
    if (graphs) {
      PB_HIP(bt->ctx, hipGraphLaunch(ex, s));
    } else {
      for (size_t m = 0; m < M; ++m) launch_chunk<T>(bt->ops[m], work_of<T>(bt->ops[m]), chunk, VAR_FAST, s);
    }
    PB_LAUNCH_CHECK(bt->ctx);
    issued += chunk;
  }
  out.assign(bt->hst, bt->hst + M);
  return PB_OK;
}

}  // namespace

namespace {
int batch_group_new(pb_ctx* ctx, pb_op2d* const* ops, int count, pb_batch** out);
int batch_group_solve(pb_batch* bt, const double* const* b, double* const* x, const pb_solve_opts* opts,
                      pb_solve_info* infos, hipStream_t s);
void batch_group_free(pb_batch* bt);

// Bytes a member's FP64 solve keeps live: 9 Krylov vectors in T, the FP64
// iterate and ~4 operator words per node.  Groups are filled up to half of
// MI355X's 256 MB Infinity Cache so the lockstep iterations of a group stay
// cache-resident (members bigger than that form a group alone).
constexpr double kBatchCacheBudget = 128.0 * 1024 * 1024;
double member_bytes(const pb_op2d* op) { return (double)op->n * 8.0 * 14.0; }
}  // namespace

extern "C" int pb_batch_create(pb_ctx* ctx, pb_op2d* const* ops, int count, pb_batch** out) {
  if (!ctx || !ops || !out || count < 1 || count > 4096)
    return pb_fail(ctx, PB_ERR_ARG, "pb_batch_create: bad argument");
  for (int m = 0; m < count; ++m) {
    if (!ops[m] || ops[m]->ctx != ctx) return pb_fail(ctx, PB_ERR_ARG, "pb_batch_create: member %d invalid", m);
    for (int q = 0; q < m; ++q)
      if (ops[q] == ops[m]) return pb_fail(ctx, PB_ERR_ARG, "pb_batch_create: operator %d repeated", m);
  }
  pb_batch* top = new pb_batch();
  top->ctx = ctx;
  top->ops.assign(ops, ops + count);
  int beg = 0;
  double bytes = 0.0;
  for (int m = 0; m <= count; ++m) {
    const double mb = m < count ? member_bytes(ops[m]) : 0.0;
    if (m == count || (m > beg && bytes + mb > kBatchCacheBudget)) {
      pb_batch* g = nullptr;
      const int rc = batch_group_new(ctx, ops + beg, m - beg, &g);
      if (rc) {
        pb_batch_destroy(top);
        return rc;
      }
      top->groups.push_back(g);
      top->first.push_back(beg);
      beg = m;
      bytes = 0.0;
    }
    bytes += mb;
  }
  *out = top;
  return PB_OK;
}

namespace {
int batch_group_new(pb_ctx* ctx, pb_op2d* const* ops, int count, pb_batch** out) {
  pb_batch* bt = new pb_batch();
  bt->ctx = ctx;
  bt->ops.assign(ops, ops + count);
  auto fail = [&](const char* what) {
    batch_group_free(bt);
    return pb_fail(ctx, PB_ERR_HIP, "pb_batch_create: %s failed", what);
  };
  for (int m = 0; m < count; ++m) {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return fail("hipStreamCreate");
    bt->streams.push_back(st);
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return fail("hipEventCreate");
    bt->joins.push_back(ev);
    if (alloc_work(ops[m])) return fail("workspace allocation");
  }
  if (hipEventCreateWithFlags(&bt->fork, hipEventDisableTiming) != hipSuccess) return fail("hipEventCreate");
  void* hp = nullptr;
  if (hipHostMalloc(&hp, sizeof(KState) * count, hipHostMallocDefault) != hipSuccess) return fail("hipHostMalloc");
  bt->hst = (KState*)hp;
  if (hipMalloc(&bt->mtab[0], sizeof(BMember<double>) * count) != hipSuccess ||
      hipMalloc(&bt->mtab[1], sizeof(BMember<float>) * count) != hipSuccess ||
      hipMalloc((void**)&bt->dstates, sizeof(KState) * count) != hipSuccess)
    return fail("hipMalloc");
  *out = bt;
  return PB_OK;
}

void batch_group_free(pb_batch* bt) {
  if (!bt) return;
  if (bt->ctx) (void)hipStreamSynchronize(bt->ctx->stream);
  for (hipStream_t st : bt->streams) (void)hipStreamSynchronize(st);
  for (auto& kv : bt->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto& kv : bt->flat_graphs) (void)hipGraphExecDestroy(kv.second);
  for (void* p : bt->mtab)
    if (p) (void)hipFree(p);
  if (bt->dstates) (void)hipFree(bt->dstates);
  for (hipStream_t st : bt->streams) (void)hipStreamDestroy(st);
  for (hipEvent_t ev : bt->joins) (void)hipEventDestroy(ev);
  if (bt->fork) (void)hipEventDestroy(bt->fork);
  if (bt->hst) (void)hipHostFree(bt->hst);
  delete bt;
}
}  // namespace

extern "C" int pb_batch_destroy(pb_batch* bt) {
  if (!bt) return PB_OK;
  for (pb_batch* g : bt->groups) batch_group_free(g);
  delete bt;
  return PB_OK;
}

// Solve every member: (D_m^-1 A_m) x[m] = b[m], x0 = 0.  Modes:
//   PB_SOLVE_F64_FAST  FP64 two-launch BiCGStab, legacy tolerance (as a single
//                      PB_SOLVE_F64_FAST solve of each member);
//   PB_SOLVE_MIXED     FP32 Krylov (FP64 iterate) inside FP64 iterative
//                      refinement to ||b - A x|| <= tol ||b||; the inner runs
//                      of all members still refining share one graph per
//                      chunk; a member whose refinement stalls finishes alone
//                      in FP64.
// Synchronous; infos[m] as pb_bicgstab_solve's.  Returns 0 when every member
// converged, else the number of members that did not.
extern "C" int pb_batch_solve(pb_batch* bt, const double* const* b, double* const* x, const pb_solve_opts* opts,
                              pb_solve_info* infos) {
  if (!bt) return PB_ERR_ARG;
  pb_ctx* ctx = bt->ctx;
  if (!b || !x || !opts || !infos) return pb_fail(ctx, PB_ERR_ARG, "pb_batch_solve: null argument");
  if (!(opts->tol >= 0) || opts->maxiter < 1) return pb_fail(ctx, PB_ERR_ARG, "pb_batch_solve: bad tol/maxiter");
  if (opts->x0_given) return pb_fail(ctx, PB_ERR_ARG, "pb_batch_solve: x0 must be zero");
  const size_t M = bt->ops.size();
  for (size_t m = 0; m < M; ++m) {
    if (!b[m] || !x[m]) return pb_fail(ctx, PB_ERR_ARG, "pb_batch_solve: member %d has a null vector", (int)m);
    if (!bt->ops[m]->assembled) return pb_fail(ctx, PB_ERR_STATE, "pb_batch_solve: member %d not assembled", (int)m);
  }
  if (opts->mode != PB_SOLVE_F64_FAST && opts->mode != PB_SOLVE_MIXED)
    return pb_fail(ctx, PB_ERR_ARG, "pb_batch_solve: mode %d not batched", opts->mode);
  hipStream_t s = solve_stream(ctx);
  int failed = 0;
  for (size_t g = 0; g < bt->groups.size(); ++g) {
    const int f = bt->first[g];
    const int rc = batch_group_solve(bt->groups[g], b + f, x + f, opts, infos + f, s);
    if (rc < 0) return rc;
    failed += rc;
  }
  return failed;
}

namespace {
// one group: members solved together (flattened launches when they share the
// launch signature, else one graph branch per member)
int batch_group_solve(pb_batch* bt, const double* const* b, double* const* x, const pb_solve_opts* opts,
                      pb_solve_info* infos, hipStream_t s) {
  pb_ctx* ctx = bt->ctx;
  const size_t M = bt->ops.size();
  std::vector<KState> st;
  std::vector<char> active(M, 1);
  int failed = 0;
  if (opts->mode == PB_SOLVE_F64_FAST) {
    std::vector<double> tol(M, opts->tol);
    int rc = krylov_batch<double>(bt, active, std::vector<const double*>(b, b + M), tol, opts->maxiter, 1,
                                  opts->check_every, s, st);
    if (rc) return rc;
    for (size_t m = 0; m < M; ++m) {
      pb_op2d* op = bt->ops[m];
      PB_HIP(ctx, hipMemcpyAsync(x[m], op->w64.x, sizeof(double) * op->n, hipMemcpyDeviceToDevice, s));
      pb_solve_info out{};
      out.info = info_of(st[m]);
      out.iters = st[m].iter;
      out.outer = 1;
      out.half_step = st[m].half;
      out.resid = st[m].resid;
      out.bnrm2 = st[m].bnrm2;
      double tr = 0.0;
      rc = true_residual(op, b[m], x[m], s, &tr);
      if (rc) return rc;
      out.true_rel_resid = out.bnrm2 > 0 ? tr / out.bnrm2 : tr;
      infos[m] = out;
      failed += out.info != 0;
    }
    return failed;
  }
  if (opts->mode != PB_SOLVE_MIXED) return pb_fail(ctx, PB_ERR_ARG, "pb_batch_solve: mode %d not batched", opts->mode);
  const double inner = opts->inner_tol > 0 ? opts->inner_tol : 1e-5;
  const int max_outer = opts->max_outer > 0 ? opts->max_outer : 10;
  std::vector<double> bn(M), rn(M), target(M);
  std::vector<int> iters(M, 0), outer(M, 0), last(M, 0);
  std::vector<char> stalled(M, 0);
  int rc = 0;
  for (size_t m = 0; m < M; ++m) {
    pb_op2d* op = bt->ops[m];
    hipLaunchKernelGGL(k_fill_const, dim3(elem_blocks(op->n)), dim3(256), 0, s, op->n, x[m], 0.0);
    rc = true_residual(op, b[m], x[m], s, &bn[m]);   // op->rtrue = b
    if (rc) return rc;
    rn[m] = bn[m];
    target[m] = opts->tol * bn[m];
  }
  for (int round = 0; round < max_outer; ++round) {
    bool any = false;
    std::vector<const float*> bf(M);
    for (size_t m = 0; m < M; ++m) {
      pb_op2d* op = bt->ops[m];
      active[m] = rn[m] > target[m] && !stalled[m];
      any = any || active[m];
      bf[m] = op->w32.b;
      if (active[m])
        hipLaunchKernelGGL((k_cast<double, float>), dim3(elem_blocks(op->n)), dim3(256), 0, s, op->n,
                           (const double*)op->rtrue, op->w32.b);
    }
    if (!any) break;
    int cap = 0;
    for (size_t m = 0; m < M; ++m) cap = std::max(cap, inner_cap(bt->ops[m], opts->maxiter));
    rc = krylov_batch<float>(bt, active, bf, std::vector<double>(M, inner), cap, 0, opts->check_every, s,
                             st);
    if (rc) return rc;
    for (size_t m = 0; m < M; ++m) {
      if (!active[m]) continue;
      pb_op2d* op = bt->ops[m];
      const int nb = elem_blocks(op->n);
      iters[m] += st[m].iter;
      last[m] = info_of(st[m]);
      ++outer[m];
      const double old = rn[m];
      hipLaunchKernelGGL(k_axpy, dim3(nb), dim3(256), 0, s, op->n, 1.0, (const double*)op->w32.x, x[m]);
      rc = true_residual(op, b[m], x[m], s, &rn[m]);
      if (rc) return rc;
      if (!(rn[m] < old)) {   // no progress: undo the correction
        hipLaunchKernelGGL(k_axpy, dim3(nb), dim3(256), 0, s, op->n, -1.0, (const double*)op->w32.x, x[m]);
        rc = true_residual(op, b[m], x[m], s, &rn[m]);
        if (rc) return rc;
      }
      if (!(rn[m] < 0.5 * old)) stalled[m] = 1;
    }
  }
  for (size_t m = 0; m < M; ++m) {
    pb_op2d* op = bt->ops[m];
    if (stalled[m] && rn[m] > target[m]) {   // finish alone in FP64 from the current iterate
      rc = fp64_finish(op, b[m], x[m], target[m], opts->maxiter, opts->check_every, s, &rn[m], &iters[m], &last[m]);
      if (rc) return rc;
      ++outer[m];
    }
    pb_solve_info out{};
    out.info = rn[m] <= target[m] ? 0 : (last[m] < 0 ? last[m] : std::max(iters[m], 1));
    out.iters = iters[m];
    out.outer = outer[m];
    out.resid = rn[m];
    out.bnrm2 = bn[m];
    out.true_rel_resid = bn[m] > 0 ? rn[m] / bn[m] : rn[m];
    out.fallback = stalled[m];
    infos[m] = out;
    failed += out.info != 0;
  }
  return failed;
}
}  // namespace
