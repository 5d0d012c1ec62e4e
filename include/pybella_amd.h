/*
 * pybella_amd — C ABI of the MI355X (gfx950) nodal elliptic pressure solve.
 *
 * Every array argument is a DEVICE pointer owned by the caller, float64 unless
 * stated, row-major with x fastest, i.e. the transpose of pyBELLA's
 * ``arr[x, y]`` numpy layout.  Shapes for a grid of nx * ny inner nodes
 * (nx = ud.inx, ny = ud.iny):
 *   cells  (ny+3) x (nx+3)    Sol.rho ... Sol.rhoX, mpv.wplus[d], mpv.u/v/w,
 *                             dSdy, nu_c        (elem.sc, two ghost layers)
 *   nodes  (ny+4) x (nx+4)    mpv.p2_nodes      (node.sc)
 *   isc    (ny+2) x (nx+2)    mpv.rhs, mpv.wcenter, diag_inv (node.isc)
 *   vector  ny    x  nx       the Krylov vector p2_nodes[2:-2,2:-2].T.ravel()
 * Calls are asynchronous on the context's stream unless stated; no call
 * allocates device memory except *_create and the first solve on a handle.
 * Return value: 0 ok, < 0 error (pb_last_error gives the message),
 * > 0 only where documented.
 *
 * Reference interfaces replaced (pyBELLA tree, file:line):
 *   pb_bc_cells            set_explicit_boundary_data   src/dycore/utils/boundary.py:7
 *   pb_ghostnodes          set_ghostnodes_p2            src/dycore/utils/boundary.py:296
 *   pb_ghostcells_p2       set_ghostcells_p2            src/dycore/utils/boundary.py:284
 *   pb_op2d_assemble       operator_coefficients_nodes  second_projection.py:308
 *                          + multiply_inverse_coriolis(get_coeffs=True) :604
 *                          + precon_diag_prepare        laplacian.py:1284
 *                          + stencil_9pt_numba_test     laplacian.py:496
 *   pb_op2d_matvec         LinearOperator.matvec -> lap2D_gather_new laplacian.py:546
 *   pb_divergence_nodes    divergence_nodes             second_projection.py:525
 *   pb_correction_nodes    correction_nodes             second_projection.py:280
 *   pb_scatter_update_p2   p2_full scatter + p2_nodes += second_projection.py:241-277
 *   pb_bicgstab_solve      scipy.sparse.linalg.bicgstab call second_projection.py:230
 *   pb_op3d_*              stencil_27pt / lap3D         laplacian.py:872-1026
 */
#ifndef PYBELLA_AMD_H
#define PYBELLA_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define PB_OK 0
#define PB_ERR_ARG (-1)
#define PB_ERR_HIP (-2)
#define PB_ERR_STATE (-3)

#define PB_BC_PERIODIC 0
#define PB_BC_WALL 1

/* precision selector for matvec / solver vectors */
#define PB_F64 0
#define PB_F32 1

/* solver modes */
#define PB_SOLVE_F64 0      /* FP64 BiCGStab, reference semantics: rho = rt.r and ||r||
                               from the stored r (3 launches per iteration)      */
#define PB_SOLVE_MIXED 1    /* FP32 BiCGStab inside FP64 iterative refinement  */
#define PB_SOLVE_NN 2       /* FGMRES preconditioned by the multigrid cycle with
                               the network as its coarse solver                 */
#define PB_SOLVE_F64_FAST 3 /* FP64 BiCGStab, rho and ||r|| from s/t dots
                               (2 launches per iteration)                      */
#define PB_SOLVE_F64_FUSED 4 /* FP64 BiCGStab in one launch per iteration:
                               rho and ||r|| as PB_SOLVE_F64_FAST, r~.v_k as
                               u.p_k with u = A^T r~ formed once per solve.
                               Needs the two copies of each duplicated periodic
                               end node equal (initial residual, diag_inv,
                               hcenter, seam cell weights); otherwise the
                               solve runs PB_SOLVE_F64_FAST (info.fallback 1) */

typedef struct pb_ctx pb_ctx;
typedef struct pb_op2d pb_op2d;
typedef struct pb_op3d pb_op3d;
typedef struct pb_batch pb_batch;
typedef struct pb_op3d pb_op3d;
typedef struct pb_batch pb_batch;
typedef struct pb_nn pb_nn;

typedef struct pb_grid2d {
  int nx, ny;          /* inner node counts (ud.inx, ud.iny)            */
  double dx, dy;       /* node spacings (node.dx, node.dy)              */
  int bc_x, bc_y;      /* PB_BC_PERIODIC or PB_BC_WALL (RAYLEIGH->WALL) */
} pb_grid2d;

typedef struct pb_phys {
  double gamm;             /* ud.gamm                                    */
  double Msq;              /* ud.Msq                                     */
  double g;                /* ud.gravity_strength[1]                     */
  double coriolis[3];      /* ud.coriolis_strength                       */
  double nonhydro;         /* ud.nonhydrostasy                           */
  int is_compressible;     /* ud.is_compressible                         */
  double compressibility;  /* ud.compressibility                         */
  int is_arakawa_konor;    /* ud.is_ArakawaKonor                         */
  double wind[3];          /* ud.u/v/w_wind_speed (time-step kernels)     */
  int stratified;          /* 1: strat_inv below holds 1/stratification(y)
                              at the gravity-axis ghost rows; 0: 1 there   */
  double strat_inv[4];     /* cell ghost rows 0, 1, ny+1, ny+2 (bottom to
                              top) of the y images (boundary.py:83)        */
} pb_phys;

/* 3D node grid: nx * ny * nz inner nodes (y the gravity axis), spacings,
 * boundary type per axis. */
typedef struct pb_grid3d {
  int nx, ny, nz;
  double dx, dy, dz;
  int bc_x, bc_y, bc_z;
} pb_grid3d;

typedef struct pb_solve_opts {
  int mode;            /* PB_SOLVE_*                                          */
  double tol;          /* ud.tol (legacy semantics in PB_SOLVE_F64)           */
  int maxiter;         /* ud.max_iterations (per Krylov run); PB_SOLVE_NN: the
                          flexible-GMRES iteration budget before the fallback */
  double inner_tol;    /* PB_SOLVE_MIXED: inner relative tol                  */
  int max_outer;       /* PB_SOLVE_MIXED: refinement cap                      */
  int check_every;     /* iterations between host convergence checks (0=auto) */
  int x0_given;        /* 0: x0 = 0 as the reference; 1: use x as the guess   */
  int fallback_maxiter; /* PB_SOLVE_NN: BiCGStab iterations after the budget  */
  int krylov_m;        /* PB_SOLVE_NN: FGMRES restart length (<=0: 30, max 60) */
  int sweeps;          /* PB_SOLVE_NN: Jacobi sweeps before and after each
                          coarse correction of the cycle (<=0: 2, max 8)       */
  double omega;        /* PB_SOLVE_NN: Jacobi damping (<=0: 0.8)               */
  int use_network;     /* PB_SOLVE_NN coarse solve: 1 the network on its level;
                          0 Jacobi sweeps on that level (the same cycle without
                          the network, for A/B); 2 the cycle continued to the
                          coarsest level (no network)                          */
} pb_solve_opts;

typedef struct pb_solve_info {
  int info;            /* 0 converged, >0 not converged (iterations), <0 breakdown */
  int iters;           /* total Krylov iterations (callback count)                 */
  int outer;           /* refinement steps taken (mixed / NN)                      */
  int half_step;       /* 1 if the last run stopped at the early ||s|| test        */
  double resid;        /* last recursive residual norm                             */
  double bnrm2;        /* ||b||                                                    */
  double true_rel_resid; /* ||b - A x|| / ||b|| recomputed in FP64                 */
  int fallback;        /* PB_SOLVE_NN: 1 if plain BiCGStab had to finish the solve;
                          PB_SOLVE_F64_FUSED: 1 if the seams sent it to the
                          two-launch iteration                               */
} pb_solve_info;

int pb_version(void);
int pb_ctx_create(int device, pb_ctx** out);
int pb_ctx_destroy(pb_ctx* ctx);
/* Issue all later work on hip_stream (NULL = the null stream).  A new context
 * uses a non-blocking stream of its own until this is called. */
int pb_ctx_set_stream(pb_ctx* ctx, void* hip_stream);
int pb_ctx_synchronize(pb_ctx* ctx);
const char* pb_last_error(const pb_ctx* ctx);

/* cell ghost fill: wrap / mirror in x, hydrostatic images in y (gravity axis).
 * fields[6] = rho, rhou, rhov, rhow, rhoY, rhoX (cells); rhoY0_bg = the cell
 * background column mpv.HydroState.rhoY0 (ny+3), used when !is_compressible. */
int pb_bc_cells(pb_ctx* ctx, const pb_grid2d* grid, const pb_phys* phys,
                double* const fields[6], const double* rhoY0_bg);

/* the same fill along one axis only (axis 0: x, 1: y; -1: both, as
 * pb_bc_cells), as the advection's split steps call it (boundary.py:30-39).
 * Without gravity (phys->g == 0) y is filled like x: wrap or mirror. */
int pb_bc_cells_axis(pb_ctx* ctx, const pb_grid2d* grid, const pb_phys* phys,
                     double* const fields[6], const double* rhoY0_bg, int axis);

/* cell ghost fill of a scalar (ny+3)x(nx+3) cell array such as mpv.p2_cells:
 * wrap (periodic) or mirror (wall), x then y.  Replaces set_ghostcells_p2,
 * src/dycore/utils/boundary.py:284-293. */
int pb_ghostcells_p2(pb_ctx* ctx, const pb_grid2d* grid, double* p);

/* node ghost fill of a (ny+4)x(nx+4) node array (periodic_plus_one / reflect) */
int pb_ghostnodes(pb_ctx* ctx, const pb_grid2d* grid, double* p);

int pb_op2d_create(pb_ctx* ctx, const pb_grid2d* grid, pb_op2d** out);
int pb_op2d_destroy(pb_op2d* op);

/* Assemble the operator on the device.  rho_w/rhoY_w: state for wplus and
 * wcenter (Sol0 when given, else Sol); rho_c/rhoY_c: state for the Coriolis
 * factors (Sol).  dSdy: cells.  Outputs (any may be NULL): wplus0, wplus1,
 * nu_c (cells), wcenter, diag_inv (isc).  wcenter is scaled by the regime
 * (x compressibility for pseudo-incompressible) before diag_inv is formed, as
 * second_projection.py:141-153 does before precon_diag_prepare. */
int pb_op2d_assemble(pb_op2d* op, const pb_phys* phys, double dt,
                     const double* rho_w, const double* rhoY_w,
                     const double* rho_c, const double* rhoY_c, const double* dSdy,
                     double* wplus0, double* wplus1, double* nu_c,
                     double* wcenter, double* diag_inv);

/* y = diag_inv * (A x); precision PB_F64 (double*) or PB_F32 (float*) */
int pb_op2d_matvec(pb_op2d* op, const void* x, void* y, int precision);

/* b_vec = diag_inv[1:-1,1:-1] * rhs[1:-1,1:-1] as a Krylov vector (rhs: isc) */
int pb_op2d_scale_rhs(pb_op2d* op, const double* rhs_isc, double* b_vec);

/* Solve (D^-1 A) x = b.  Synchronous (returns after the solve).  Returns the
 * info code (0 converged, >0 iterations when not converged, <0 breakdown) or
 * a PB_ERR_* code. */
int pb_bicgstab_solve(pb_op2d* op, const double* b, double* x,
                      const pb_solve_opts* opts, pb_solve_info* info);

/* The FP64 solve of pb_bicgstab_solve (modes PB_SOLVE_F64 and
 * PB_SOLVE_F64_FAST only) without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device or pinned memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its Krylov vectors, partial records and state records live in
 * a static device array of 40 MiB per device (looked up at pb_ctx_create),
 * which holds grids up to about 650,000 nodes (1025 x 513 fits); a larger
 * grid returns PB_ERR_ARG.  The iterate is kept in x itself; iterations are
 * launched directly (no hipGraph) and the state is read back into host stack
 * memory.  Iterates, x and the info equal pb_bicgstab_solve's in the same
 * mode and tile height.  Every call on one device shares that array and it
 * keeps the last call's contents: calls that overlap on one device (two
 * streams or host threads) race and must be serialised by the caller.
 * Synchronous; returns the info code as pb_bicgstab_solve. */
int pb_bicgstab_solve_noalloc(pb_op2d* op, const double* b, double* x,
                              const pb_solve_opts* opts, pb_solve_info* info);

/* Tile height of the solver's stencil launches: 64 x (4 * rows_per_thread)
 * nodes per 256-thread block; rows_per_thread in {2, 3, 4, 5, 6, 8} (default
 * 4).  The height sets the block count, and so how the grid fills the CUs'
 * block slots; used by the tile autotuner.  No reference counterpart. */
int pb_op2d_set_tile(pb_op2d* op, int rows_per_thread);

/* Average device time (microseconds) of one launch of a solver kernel, timed
 * with HIP events over ``reps`` back-to-back launches on the context stream on
 * pattern-filled workspace vectors.  kernel: 0 plain matvec, 1 BiCGStab
 * p-update + matvec + rt.v reduction, 2 s + matvec + (s.s, t.s, t.t), 3 x/r
 * update + (rt.r, r.r).  Synchronous.  Measurement aid; no reference
 * counterpart.  Kernels 1 and 2 are the two launches of the PB_SOLVE_F64_FAST
 * iteration; kernel 3 is the FP64 true-residual stencil. */
int pb_op2d_time_kernel(pb_op2d* op, int kernel, int precision, int reps, double* avg_us);

/* Average device time (microseconds) of the A (kernel 1) or B (kernel 2)
 * launch of the PB_SOLVE_F64_FAST iteration inside a real solve of b to
 * tolerance tol (x0 = 0; iterations launched directly, no graph): at the
 * start of each chunk of 32 iterations whose residual is above 100x the
 * target, that launch is repeated 8 times back to back on the live state,
 * in its place in the iteration, between HIP events on the solve stream
 * (the repeats write identical outputs; A's in-place x update is saved and
 * restored); *launches = the number timed; x receives the solution.
 * Synchronous.  Measurement aid; no reference counterpart. */
int pb_op2d_time_in_solve(pb_op2d* op, const double* b, double* x, double tol, int kernel,
                          double* avg_us, int* launches);

/* divergence of (rhou, rhov)*Y at nodes -> rhs (isc), divided by dt.
 * Zeroes rhou/rhov/rhow in the two cell rows next to y walls first. */
int pb_divergence_nodes(pb_ctx* ctx, const pb_grid2d* grid, double dt,
                        double* rho, double* rhou, double* rhov, double* rhow,
                        const double* rhoY, double* rhs);

/* momentum correction from node pressure p (nodes): rhou, rhov += ...,
 * rhoX -= updt_chi dt dSdy rhov; mpv u, v, w (cells) receive the corrections. */
int pb_correction_nodes(pb_ctx* ctx, const pb_grid2d* grid, const pb_phys* phys,
                        double dt, double updt_chi, const double* p,
                        const double* rho, double* rhou, double* rhov,
                        const double* rhoY, double* rhoX, const double* dSdy,
                        double* u, double* v, double* w);

/* p2_full = scatter(x) with node ghosts; p2_nodes += p2_full; ghost fill.
 * p2_full (nodes) receives the increment. */
int pb_scatter_update_p2(pb_ctx* ctx, const pb_grid2d* grid, const double* x,
                         double* p2_full, double* p2_nodes);

/* ---- explicit parts of the time step ----
 * Face-flux arrays: x faces (ny+3) x (nx+4) in [y][face] (the reference's
 * flux[0] as it lies), y faces (ny+4) x (nx+3) in [face][x] (the transpose of
 * the reference's flux[1]).  flux[6] = rho, rhou, rhov, rhow, rhoY, rhoX face
 * fluxes, rhoY being the advective flux.  Faces 0 and n of a line are never
 * written: the caller zeroes them once.  cells[6] = rho, rhou, rhov, rhow,
 * rhoY, rhoX as for pb_bc_cells.  All asynchronous on the context stream
 * except pb_cfl_maxima. */

/* rhoY-weighted advective face fluxes into the rhoY entries of both flux
 * sets (inner faces and lines only).  Replaces recompute_advective_fluxes,
 * numerical_flux.py:5-48. */
int pb_advective_fluxes(pb_ctx* ctx, const pb_grid2d* grid, const double* rho, const double* rhou,
                        const double* rhov, const double* rhoY, double* fx_rhoY, double* fy_rhoY);

/* One advection split step along axis (0 x, 1 y): ghost cells along the
 * axis, the second-order recovery of face states (limiter NONE) and the
 * upwind (HLL) fluxes of rho, rhou, rhov, rhow, rhoX, then, unless rk, the
 * conservative update with lmbda = dt/dx, then ghost cells along the axis
 * again.  rk: the one-stage Runge-Kutta variant (recovery without the
 * time-centring term, no update).  Replaces explicit_step_and_flux,
 * explicit.py:75-165, with recovery.do (recovery.py:6-210) and hll_solver
 * (numerical_flux.py:50-99). */
int pb_advect_split(pb_ctx* ctx, const pb_grid2d* grid, const pb_phys* phys, int axis, double lmbda,
                    int rk, double* const cells[6], double* const flux[6], const double* rhoY0_bg);

/* The update loop of advect_rk: q += lx (fx_l - fx_r), then q += ly (fy_l -
 * fy_r), then every ghost cell (explicit.py:219-237). */
int pb_advect_rk_update(pb_ctx* ctx, const pb_grid2d* grid, const pb_phys* phys, double lx, double ly,
                        double* const cells[6], double* const fx[6], double* const fy[6],
                        const double* rhoY0_bg);

/* Buoyancy, then the inverse Coriolis / buoyancy matrix on the momenta
 * relative to the wind, then every ghost cell.  Replaces
 * euler_backward_non_advective_expl_part, second_projection.py:72-90. */
int pb_expl_part(pb_ctx* ctx, const pb_grid2d* grid, const pb_phys* phys, double dt,
                 double* const cells[6], const double* dSdy, const double* rhoY0_bg);

/* The explicit forward step: rhs (isc) = nodal divergence (momenta zeroed
 * next to y walls first), doubled on wall node rows; momenta and rhoX
 * updated from the old p2_nodes, buoyancy and Coriolis; p2_nodes +=
 * compressibility * (-dt dpidP rhs); node and cell ghosts.  S0c: the cell
 * background column (ny+3).  Replaces euler_forward_non_advective,
 * second_projection.py:22-69. */
int pb_euler_forward(pb_ctx* ctx, const pb_grid2d* grid, const pb_phys* phys, double dt,
                     double* const cells[6], double* p2_nodes, double* rhs, const double* S0c,
                     const double* dSdy, const double* rhoY0_bg);

/* Device work size (doubles) of pb_cfl_maxima for a grid. */
long pb_cfl_work_words(const pb_grid2d* grid);

/* maxima[6] (host) = max over every cell of |u|, |v|, |w|, |u|+c, |v|+c,
 * |w|+c with c = sqrt(gamm rhoY^gamm / rho) * Minv, Minv = 1/sqrt(Msq): the
 * reductions of dynamic_timestep (cfl.py:5-35).  work: pb_cfl_work_words
 * doubles of device memory.  Synchronous. */
int pb_cfl_maxima(pb_ctx* ctx, const pb_grid2d* grid, double gamm, double Minv, const double* rho,
                  const double* rhou, const double* rhov, const double* rhow, const double* rhoY,
                  double* work, double* maxima);

/* ---- LETKF local analysis (data assimilation of an ensemble) ----
 * One grid type (cells or nodes) of nx * ny inner points; every point is
 * analysed over the obs_x * obs_y window around it (odd sizes, periodic
 * wrap over the inner grid, zero localisation weight outside a wall).
 * state, analysis: (members, attrs, ny, nx); obs, mask: (attrs, ny, nx),
 * mask != 0 where a location is not observed; covar: (attrs) observation
 * variances or NULL (1); loc: (obs_x, obs_y) localisation weights.  Members
 * 2..32.  Replaces prepare_rloc.analyse_by_grid_type with analysis.analyse,
 * src/data_assimilation/letkf.py:288-426, 120-166. */
typedef struct pb_letkf_cfg {
  int members, attrs, nx, ny;
  int obs_x, obs_y;
  int periodic_x, periodic_y;
  double inflation;    /* dap.inflation_factor (rho)                      */
} pb_letkf_cfg;

int pb_letkf_analyse(pb_ctx* ctx, const pb_letkf_cfg* cfg, const double* state, const double* obs,
                     const double* mask, const double* covar, const double* loc, double* analysis);

/* ---- learned preconditioner (no reference counterpart: pyBELLA has none) ----
 * Fixed architecture: 4-level convolutional encoder-decoder, 3x3 convolutions,
 * 16-32-64-64 channels, GELU, 15 layers, BF16 MFMA kernels with FP32
 * accumulation and FP32 master parameters (layout per layer: W[9][cin][cout]
 * then bias[cout]).  Grids must have odd node counts (2^k + 1).  Every
 * network call below runs on the context's stream (pb_ctx_set_stream). */
int pb_nn_create(pb_ctx* ctx, unsigned long long seed, pb_nn** out);
int pb_nn_destroy(pb_nn* nn);
long pb_nn_num_params(const pb_nn* nn);
/* host copies of the parameters and / or the last computed gradients */
int pb_nn_get_params(pb_nn* nn, float* host_params, float* host_grads);
int pb_nn_set_params(pb_nn* nn, const float* host_params);
/* One batch: inputs float[nimg][h*w][3] (normalised r, feature 1, feature 2),
 * target float[nimg][h*w] or NULL (forward only), img_weight float[nimg] or
 * NULL (all 1), out_z float[nimg][h*w] or NULL (device pointers).  With a
 * target: gradients of the per-image weighted mean squared error of the output
 * (constant and checkerboard components removed) and, if lr > 0, one Adam
 * step.  Synchronous. */
int pb_nn_train_step(pb_nn* nn, int nimg, int h, int w, const float* inputs, const float* target,
                     const float* img_weight, double lr, float* out_z, double* loss);
/* z = rms(r) * net(r / rms(r), features(op)); r, z: FP64 device vectors. */
int pb_nn_apply(pb_nn* nn, const pb_op2d* op, const double* r, double* z);
/* the two coefficient features (float device arrays of nx*ny) of an operator */
int pb_nn_features(pb_nn* nn, const pb_op2d* op, float* f1, float* f2);
/* attach a network to an operator: PB_SOLVE_NN then uses it as the coarse
 * solver of the multigrid cycle that preconditions flexible GMRES (NULL
 * detaches).  pb_nn_destroy detaches the network from every operator. */
int pb_op2d_attach_nn(pb_op2d* op, pb_nn* nn);
/* Null space the training loss removes from the output: 0 {constant,
 * checkerboard} (default), 1 {constant} (a coarse level's operator). */
int pb_nn_set_loss_nullspace(pb_nn* nn, int mode);

/* ---- the multigrid cycle of PB_SOLVE_NN (no reference counterpart) ----
 * The scaled 9-point operator couples, without Coriolis, only nodes of equal
 * parity i + j; the cycle restricts r and checkerboard * r from the fine grid
 * (two branches, carried as two channels), continues with Galerkin operators
 * R A P (9-point stencils on unique grids: a periodic axis keeps n - 1 nodes)
 * and solves one coarse level (the first whose grid, periodic nodes
 * duplicated, is 8k+1 nodes per axis and at most 65 x 33) with the network.
 * One application z = M v (v FP64, z FP32 Krylov vectors): coarse 0 the
 * network, 1 Jacobi sweeps on that level, 2 the cycle continued to the
 * coarsest level; nu Jacobi sweeps each way (damping omega); branches 1 or 2.
 * Test and training aid. */
int pb_op2d_precond(pb_op2d* op, const double* v, float* z, int coarse, int nu, double omega, int branches);
/* Levels of the hierarchy: dims[4 l + 0..3] = ny, nx, periodic y, periodic x
 * (unique grid) for l < max_levels; *nn_level = the network's level or -1.
 * Returns the level count. */
int pb_op2d_mg_info(pb_op2d* op, int* dims, int max_levels, int* nn_level);
/* FP64 stencil [9][ny][nx] of a level (device pointer); entry k multiplies
 * the node at row offset k / 3 - 1, column offset k % 3 - 1. */
int pb_op2d_mg_stencil(pb_op2d* op, int level, double* st);
/* The network level's input features on its duplicated grid (device float
 * arrays of nh * nw values): -(W + E) / C and -(N + S) / C of its stencil. */
int pb_op2d_mg_features(pb_op2d* op, float* f1, float* f2);

/* ---- Batched uneven 2D solves (ensemble members of any grid sizes) ----
 * No reference counterpart: the reference solves members one after the other
 * (src/__main__.py:223-254, each through second_projection.py:230).  A batch
 * holds distinct assembled operators (one per member, each with its own tile
 * height); a solve captures one chunk of iterations of every member as a
 * single hipGraph (one branch per member) and replays it until all members
 * are final.  Per member the iterates equal a single solve's. */
int pb_batch_create(pb_ctx* ctx, pb_op2d* const* ops, int count, pb_batch** out);
int pb_batch_destroy(pb_batch* batch);
/* b[m], x[m]: device FP64 Krylov vectors of member m.  Modes
 * PB_SOLVE_F64_FAST and PB_SOLVE_MIXED.  Synchronous.  Returns 0 when every
 * member converged, > 0 the number that did not, < 0 an error. */
int pb_batch_solve(pb_batch* batch, const double* const* b, double* const* x, const pb_solve_opts* opts,
                   pb_solve_info* infos);

/* ---- 3D 27-point operator (stencil_27pt / lap3D, laplacian.py:872-1026) ----
 * Layout: nodes (nz, ny, nx) x fastest; operator cells (nz+1, ny+1, nx+1)
 * (cell (k, j, i) between nodes k-1..k, j-1..j, i-1..i; the outer layers are
 * ghost cells).  Periodic axes wrap onto the duplicated end node; wall axes
 * zero the outside cell layers.  The x-z Coriolis cross term is not built
 * (coriolis_strength[0] = 0 in every 3D configuration). */
int pb_op3d_create(pb_ctx* ctx, const pb_grid3d* grid, pb_op3d** out);
int pb_op3d_destroy(pb_op3d* op);
/* Coefficients from the cell state (operator_coefficients_nodes,
 * second_projection.py:308-356, every wplus[d] = Gammainv rhoY Y; the regime
 * scaling :139-158) and the 3D diagonal (precon_diag_prepare,
 * laplacian.py:1284-1356).  wcenter / diag_inv: optional node outputs. */
int pb_op3d_assemble(pb_op3d* op, const pb_phys* phys, double dt, const double* rho_cells,
                     const double* rhoY_cells, double* wcenter, double* diag_inv);
/* y = diag_inv * (A x); precision PB_F64 or PB_F32 */
int pb_op3d_matvec(pb_op3d* op, const void* x, void* y, int precision);
/* Solve (D^-1 A) x = b with x0 = 0 (the reference's bicgstab call,
 * second_projection.py:230, on the 27-point operator).  Modes PB_SOLVE_F64
 * (legacy tolerance) and PB_SOLVE_MIXED (FP32 inner solves in FP64
 * refinement).  Synchronous; returns the info code as pb_bicgstab_solve. */
int pb_op3d_solve(pb_op3d* op, const double* b, double* x, const pb_solve_opts* opts, pb_solve_info* info);
/* The tile of every block of a 3D stencil march launched with gx * gy * gz
 * blocks of 64 x 8 nodes and one z chunk each (the launch's gz is
 * ceil(nz / chunk)): out[3 h + 0..2] = x tile, y tile, z chunk of hardware
 * block h, in the XCD-aware order the kernels use.  Host only, no GPU call;
 * a check of the block mapping.  No reference counterpart. */
int pb_op3d_march_map(int gx, int gy, int gz, int* out);
/* Average device time (microseconds) of a 3D kernel: 0 matvec, 1 iteration
 * launch 1 (x, r, p updates + v = A p + rt.v), 2 launch 2 (t = A s + 5 dots).
 * Measurement aid; no reference counterpart. */
int pb_op3d_time_kernel(pb_op3d* op, int kernel, int precision, int reps, double* avg_us);

#ifdef __cplusplus
}
#endif
#endif /* PYBELLA_AMD_H */
