// Two-branch multigrid cycle of the 2D nodal operator (gfx950), the
// preconditioner around the learned coarse solver.  No reference
// counterpart: pyBELLA preconditions its BiCGStab with the diagonal only
// (laplacian.py:1284-1356, second_projection.py:205-230).
//
// The scaled 9-point operator D^-1 A couples, with zero Coriolis and equal x
// and y weights, only nodes of equal parity i + j: a smooth field f and the
// checkerboard-modulated c * f are equally low-energy.  Full weighting onto
// the even-even nodes sees only the sum of the two sublattices' envelopes, so
// the fine level restricts r and c * r (two branches, carried as two channels
// through every coarser level) and adds back P e_s + c * P e_d.  Levels >= 1
// hold Galerkin operators R A P (R = P^T / 4, P bilinear) as 9-point stencils
// on unique node grids: a periodic axis keeps n - 1 nodes and wraps (the
// Krylov layout duplicates the periodic end node), a wall axis keeps all.
//
// The coarse solve is, per mode: the network on one coarse level (both
// branches as a batch of two images, on the level's grid with the periodic
// column duplicated), Jacobi sweeps on that level (the same cycle without the
// network), or the V-cycle continued to the coarsest level.
#pragma once
#include <vector>

#include "pb_op2d.hpp"

namespace pbmg {

constexpr int kMaxLevels = 16;

struct Dims {
  int ny, nx;   // node rows, node columns of the unique grid
  int py, px;   // periodic axes
};

struct Level {
  Dims d{};
  double* st64 = nullptr;   // stencil [9][ny][nx], entry k multiplies z[j + k/3 - 1][i + k%3 - 1]
  float* st = nullptr;      // FP32 copy (levels >= 1)
  float* idg = nullptr;     // 1 / centre entry (levels >= 1)
  float* v = nullptr;       // [2][ny][nx] right-hand sides (the two branches)
  float* z = nullptr;       // [2][ny][nx] iterate
  float* t = nullptr;       // [2][ny][nx] second iterate buffer
};

struct Hier {
  long version = -1;        // operator assembly the stencils were built from
  int nlev = 0;             // levels (0 = the operator's grid)
  int nn_level = -1;        // level solved by the network, -1 if no level fits it
  Level L[kMaxLevels];
  // level 0 in the Krylov layout (ny x nx with the duplicated periodic nodes)
  float* id0 = nullptr;     // 1 / diagonal of D^-1 A
  float* v0 = nullptr;      // FP32 copy of the cycle's input
  float* z0 = nullptr;
  float* t0 = nullptr;
  float* r0 = nullptr;
  // the network level with its periodic columns (rows) duplicated: nh x nw
  int nh = 0, nw = 0;
  float *f1 = nullptr, *f2 = nullptr;   // coefficient features [nh][nw]
  std::vector<void*> allocs;
};

enum { COARSE_NET = 0, COARSE_SMOOTH = 1, COARSE_MG = 2 };

struct CycleOpts {
  int nu_pre = 2, nu_post = 2;   // Jacobi sweeps before / after the coarse correction
  float omega = 0.8f;            // Jacobi damping
  int coarse = COARSE_NET;
  int coarse_sweeps = 2;         // sweeps after the network / of the smoother-only coarse level
  int bottom_sweeps = 40;        // sweeps on the coarsest level (COARSE_MG)
  int branches = 2;              // 1 drops the checkerboard branch (for tests)
};

// (Re)build the hierarchy if the operator was re-assembled since.
int build(pb_op2d* op, hipStream_t s);
// z = M v: v an FP64 Krylov vector, z FP32 (both in the Krylov layout).
int cycle(pb_op2d* op, pb_nn* nn, const double* v, float* z, const CycleOpts& o, hipStream_t s);
void release(pb_op2d* op);

}  // namespace pbmg
