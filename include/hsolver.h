/*
 * hsolver.h — C ABI of the MI355X (gfx950) SIMPLE conjugate-heat-transfer step.
 *
 * One hs_ctx owns the device copy of one structured, staggered, non-uniform grid:
 * the fluid region (velocities on faces, pressure / pressure correction / fluid
 * temperature on cells) and the solid ("mobs") temperature, on one GPU and one
 * HIP stream.  Every entry point is stream-ordered and synchronises before it
 * returns; host buffers are owned by the caller.  One ctx per host thread.
 *
 * The entry points mirror the member functions of the reference's header-only
 * solver structs, which are the reference's only interface for this path:
 *   FluidSolver  fluid_solver.h:52-182   (init :882, solve :1401, setForNext :1634)
 *   MobSolver    mobs_solver.h:31-77     (init :116, solve :303, setForNext :290)
 *   Solver       solver.h:24-45          (solve :44, exchange :62-148)
 * Each declaration below names the reference function it replaces.
 *
 * Layouts (all x-fastest, "padded" = one ghost layer on every side, exactly the
 * reference's padded work arrays, fluid_solver.h:93-109):
 *   P  cells  (nx+2)(ny+2)(nz+2)   p_temp, p_prime, t_temp, solid t_temp
 *   U  x-faces (nx+3)(ny+2)(nz+2)  u_star, up
 *   V  y-faces (nx+2)(ny+3)(nz+2)  v_star, vp
 *   W  z-faces (nx+2)(ny+2)(nz+3)  w_star, wp
 *   C  cells  nx*ny*nz (unpadded)   LocalCell.data of p (distributed_mesh.h:17-20)
 * Padded index of (i,j,k) in layout X with extents (ex,ey,ez) is i + ex*(j + ey*k).
 */
#ifndef HSOLVER_H
#define HSOLVER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes (the reference returns void everywhere; these replace its
 *      silent failure modes, e.g. ignored PCG non-convergence :1792-1793) ---- */
#define HS_OK          0
#define HS_E_ARG      -1   /* bad argument / unsupported grid */
#define HS_E_HIP      -2   /* HIP runtime error */
#define HS_E_NOCONV   -3   /* linear solve hit maxit (only if params.fail_on_noconv) */
#define HS_E_NAN      -4   /* NaN/Inf detected in a residual */
#define HS_E_IO       -5   /* file could not be read */
#define HS_E_STATE    -6   /* call out of order (e.g. solve before init) */
#define HS_E_NOMEM    -7   /* device allocation failed */

/* ---- physics mode ----
 * HS_MODE_REF    reference-faithful: lid-driven cavity, every quirk of the
 *                reference kept (frozen p' matrix, never-reset GS residual,
 *                swapped upwind labels in y/z momentum, leaky min-side walls,
 *                uniform-mean solid conduction, split fluid/solid temperature
 *                with a one-cell ghost exchange).
 * HS_MODE_CONFIG parameterised channel / cavity physics: inlet, outflow,
 *                no-penetration at every wall and solid face, symmetric
 *                area-consistent p' operator rebuilt every iteration,
 *                monolithic conjugate energy equation with harmonic-mean
 *                face conductivity and a per-cell heat source. */
#define HS_MODE_REF     0
#define HS_MODE_CONFIG  1

/* sweep ordering of the under-relaxed momentum / energy sweeps */
#define HS_ORDER_LEX    0   /* lexicographic Gauss-Seidel, identical to the reference's k,j,i loops */
#define HS_ORDER_RB     1   /* red-black Gauss-Seidel (colour = (i+j+k)&1 of the padded index) */

/* pressure-correction solver */
#define HS_PRECOND_NONE 0   /* CG with identity preconditioner: HYPRE ParCSR PCG as coded */
#define HS_PRECOND_GMG  1   /* flexible CG preconditioned by geometric multigrid */
#define HS_PRECOND_AMG  2   /* flexible CG preconditioned by aggregation (algebraic) multigrid:
                               the slot the reference reserves for BoomerAMG
                               (fluid_solver.h:1775-1787, 1876-1888) */

#define HS_PREC_FP64    0
#define HS_PREC_MIXED   1   /* FP32 multigrid inside FP64 flexible CG */

typedef struct {
    int32_t nx, ny, nz;               /* cells */
    const double *xv, *yv, *zv;       /* vertex coordinates, nx+1 / ny+1 / nz+1 (.in xgrid/ygrid/zgrid) */
    const uint8_t *cell_type;         /* nx*ny*nz, x-fastest: 0 fluid, 1 solid */
} hs_grid;

typedef struct {
    int32_t mode;                     /* HS_MODE_* */
    int32_t order;                    /* HS_ORDER_* */
    double  rho;                      /* RHO  (fluid_solver.h:23) */
    double  nu;                       /* NU   (fluid_solver.h:25) */
    double  alpha_f;                  /* ALPHA fluid diffusivity (fluid_solver.h:30); Mode C: k_fluid / (rho c_p) */
    double  k_ratio;                  /* Mode C: k_solid / k_fluid (per-cell k overrides, hs_set_conductivity) */
    double  u_lid;                    /* U0 lid speed (fluid_solver.h:22) */
    double  u_in;                     /* Mode C: inlet u on the x=0 faces (0 = wall) */
    int32_t outflow;                  /* Mode C: 1 = zero-gradient outflow on the x=max faces */
    int32_t lid;                      /* Mode C: 1 = moving lid on the y=max wall */
    double  relax_u, relax_p, relax_t, relax_ts;   /* UNDER_RELAX (fluid_solver.h:35, mobs_solver.h:27) */
    double  mom_tol, t_tol, ts_tol;   /* GS stop: CONVERGENCE 2e-3 (fluid), CONVERGENCE2 1e-3 (solid) */
    int32_t mom_sweeps_max, t_sweeps_max, ts_sweeps_max;   /* GS_STEP 100 */
    double  p_tol;                    /* HYPRE_PCGSetTol (fluid_solver.h:1874) */
    int32_t p_maxit;                  /* HYPRE_PCGSetMaxIter (fluid_solver.h:1873) */
    int32_t freeze_p_matrix;          /* 1 = keep the iteration-1 p' matrix (reference values1d behaviour) */
    int32_t precond;                  /* HS_PRECOND_* */
    int32_t precision;                /* HS_PREC_* */
    int32_t mg_levels;                /* multigrid levels including the finest; 0 = down to <= 4 cells a side */
    int32_t mg_pre, mg_post;          /* RB-GS sweeps before / after coarse correction; -1 (default)
                                         = by grid: V(3,3) and V(1,2) on the finest level when no
                                         cell's aspect ratio exceeds 8, else V(6,6) everywhere */
    int32_t mg_coarse_sweeps;         /* sweeps on the coarsest level when no dense solve */
    int32_t mg_dense_coarse;          /* N > 0: coarsest level (<= 4096 cells) solved with a dense
                                         inverse built on the matrix cores (FP32 under the mixed
                                         V-cycle, FP64 under the FP64 one), rebuilt at most every N
                                         hierarchy builds, sooner when a p' solve takes 2 or more FCG
                                         iterations more than the first one after the last rebuild
                                         (0: red-black sweeps on the coarsest level) */
    double  t_fluid_init;             /* fluid t_temp fill, 293 (fluid_solver.h:108) */
    double  t_solid_init;             /* solid t_temp fill, T1 = 333 (mobs_solver.h:55) */
    double  t_wall;                   /* Mode C: Dirichlet temperature of inlet / walls */
    int32_t fail_on_noconv;           /* 1 = return HS_E_NOCONV when the p' solve hits maxit */
    int32_t mg_dense_refine;          /* FP64 iterative-refinement steps of the dense coarse solve
                                         against the current coarse operator (default 1: it recovers
                                         the digits the FP32 inverse loses, and the change of the
                                         operator since either inverse was built) */
    int32_t mg_sweep_rows;            /* tile of the fused red-black sweep: 0 = by level width
                                         (256x4, 128x8 or 64x8 columns); 4, 8, 16 = 64-wide tiles
                                         of that many rows; -1 = two in-place colour passes */
    int32_t mg_pre0, mg_post0;        /* sweeps on the finest level (-1: mg_pre / mg_post, or
                                         the by-grid choice above) */
    int32_t mg_cycle;                 /* visits of each level below the finest per cycle: 1 V, 2 W */
    float   mg_omega;                 /* over-relaxation of the pressure multigrid's red-black
                                         sweeps on the finest level, x + omega (x_GS - x);
                                         1 = Gauss-Seidel; <= 0 (default) = by grid: 1.2, or
                                         1.7 when a cell's aspect ratio exceeds 8 */
    float   mg_omega_coarse;          /* the same on the levels below the finest (down to the
                                         single-workgroup tail, which stays at 1); <= 0
                                         (default) = by grid: 1.6, or 1.7 as above */
    float   e_omega;                  /* the same for the energy multigrid (hs_energy_solve,
                                         hs_energy_solve_rhs, hs_adjoint_energy); <= 0 (default)
                                         = 1, Gauss-Seidel: the energy operator is not
                                         symmetric and values above 1 can diverge */
    int32_t mg_wide_offsets;          /* 1 = the fused multigrid sweeps address every level with
                                         64-bit byte offsets, the form used for level arrays of
                                         2 GiB and more (0, default: 32-bit below 2 GiB) */
    int32_t e_pre, e_post;            /* sweeps of the energy multigrid before / after its coarse
                                         correction; -1 (default) = mg_pre / mg_post */
    int32_t p_warm;                   /* initial iterate of the multigrid p' solve of a SIMPLE
                                         iteration: 0 (default) = 0, as the reference
                                         (fluid_solver.h:1766); 1 = the previous iteration's p';
                                         2, 3 = the line / parabola through the last 2 / 3 p'
                                         solutions (2 x1 - x2, 3 x1 - 3 x2 + x3; fewer while fewer
                                         are held).  The stop rule stays ||r|| / ||b|| < p_tol, and a
                                         start no better than 0 is replaced by 0 */
    int32_t e_precond;                /* preconditioner of the energy solves (hs_energy_solve,
                                         hs_energy_solve_rhs, hs_adjoint_energy): HS_PRECOND_GMG
                                         (default; 0 means the same) or HS_PRECOND_AMG */
    /* aggregation multigrid (precond or e_precond = HS_PRECOND_AMG): the structured finest
       level, then levels of aggregates formed by two passes of pairwise matching along
       strong couplings, Galerkin coarse operators, a dense inverse on the matrix cores once
       a level has at most 4096 rows (FP64 refinement), V-cycle */
    int32_t amg_sweeps0;              /* red-black Gauss-Seidel sweeps on the finest level before and
                                         after the coarse correction (<= 0: 2) */
    int32_t amg_sweeps;               /* Jacobi sweeps (weight 2/3) before and after on every
                                         aggregated level (<= 0: 3) */
    float   amg_theta;                /* strong coupling: the pair's symmetrised coupling over the
                                         geometric mean of the two diagonals is at least this
                                         fraction of the row's largest (<= 0: 0.1; every coupling
                                         below 64K rows) */
    float   amg_alpha;                /* scaling of every prolongated coarse correction (<= 0: 1.4
                                         for the p' solve, 1.6 for the energy solves) */
    float   amg_omega;                /* over-relaxation of the p' cycle's finest-level sweeps (<= 0:
                                         1; the energy cycle's use e_omega) */
    int32_t amg_pad;
} hs_params;

typedef struct {
    int32_t mom_sweeps[3];            /* GS sweeps run by solveX/Y/Zmomentum this iteration */
    int32_t t_sweeps;                 /* fluid (Mode R) / conjugate (Mode C) energy sweeps */
    int32_t ts_sweeps;                /* solid sweeps (Mode R) */
    int32_t p_iters;                  /* p' Krylov iterations */
    int32_t p_converged;              /* 1 if the stop rule was met before maxit */
    int32_t pad0;
    double  p_relres;                 /* sqrt(<r,r>/<b,b>) of the recurrence at exit */
    double  p_true_relres;            /* ||b - A x||_2 / ||b||_2 recomputed in FP64 after the solve */
    double  mass_imbalance;           /* ||b||_2 of the p' right-hand side (continuity residual) */
    double  energy_residual;          /* |sum of energy cell residuals| / max(sum q V, tiny) (Mode C) */
    double  heat_in;                  /* sum q V over solid cells (Mode C) */
    double  mom_residual[3];          /* last GS residual value of each momentum component */
    double  t_residual, ts_residual;
} hs_stats;

/* array identifiers for hs_get_array / hs_put_array (layouts above) */
#define HS_A_U        0   /* u_star  U */
#define HS_A_V        1   /* v_star  V */
#define HS_A_W        2   /* w_star  W */
#define HS_A_UP       3   /* up      U  (momentum diagonal, fluid_solver.h:625) */
#define HS_A_VP       4   /* vp      V */
#define HS_A_WP       5   /* wp      W */
#define HS_A_P        6   /* p       C  (cell pressure, LocalCell.data) */
#define HS_A_PTEMP    7   /* p_temp  P */
#define HS_A_PPRIME   8   /* p_prime P */
#define HS_A_TF       9   /* fluid t_temp P (Mode C: the conjugate temperature) */
#define HS_A_TS      10   /* solid t_temp P (Mode R; same padded layout as the fluid) */
#define HS_A_Q       11   /* heat source per cell, C */
#define HS_A_K       12   /* conductivity per cell, C */
#define HS_A_B       13   /* p' right-hand side, P */
#define HS_A_COUNT   14

typedef struct hs_ctx hs_ctx;

/* ---- lifecycle ---- */
void        hs_params_default(int32_t mode, hs_params* p);
/* FluidSolver::FluidSolver fluid_solver.h:93-109 + MobSolver::MobSolver mobs_solver.h:50-56:
 * copies the grid to the device, allocates the padded work arrays, fills t_temp. */
int         hs_create(const hs_grid* g, const hs_params* p, hs_ctx** out);
void        hs_destroy(hs_ctx* c);
const char* hs_strerror(int code);
/* sizeof(hs_grid), sizeof(hs_params), sizeof(hs_stats): lets bindings check their layouts */
int         hs_abi_sizes(int32_t* sizes3);
int64_t     hs_array_size(const hs_ctx* c, int32_t which);
/* The context's parameters as it runs them: the multigrid sweep counts and over-relaxation
 * left to hs_create (-1 / <= 0) come back with the values it chose from the grid.  No
 * counterpart in the reference, whose solver settings are compile-time macros. */
int         hs_get_params(const hs_ctx* c, hs_params* p);

/* ---- the reference's solver methods ---- */
/* FluidSolver::init fluid_solver.h:882-1237 + MobSolver::init mobs_solver.h:116-288:
 * boundary conditions, interface lists, zeroed interface face velocities. */
int hs_init(hs_ctx* c);
/* FluidSolver::solve fluid_solver.h:1401-1633: x/y/z momentum, p' assembly + solve,
 * velocity / pressure correction, fluid energy. */
int hs_fluid_solve(hs_ctx* c, hs_stats* st);
/* MobSolver::solve mobs_solver.h:303-371 (solid conduction sweeps). */
int hs_solid_solve(hs_ctx* c, hs_stats* st);
/* Solver::solve exchange block solver.h:62-148 (fluid <-> solid interface temperatures). */
int hs_exchange(hs_ctx* c);
/* FluidSolver::setForNext fluid_solver.h:1634-1682 + MobSolver::setForNext mobs_solver.h:290-301. */
int hs_set_for_next(hs_ctx* c);
/* One outer iteration of Solver::solve solver.h:56-154: the four calls above in order. */
int hs_step(hs_ctx* c, hs_stats* st);
/* n outer iterations; st may be NULL or hold n entries. */
int hs_run(hs_ctx* c, int32_t n, hs_stats* st);

/* ---- finer phases (the member functions FluidSolver::solve calls) ---- */
/* solveXmomentum :577 / solveYmomentum :491 / solveZmomentum :407 (comp 0/1/2). */
int hs_momentum(hs_ctx* c, int32_t comp, int32_t* sweeps, double* residual);
/* HYPREIJMatrix_SolveCorrectionPressure :1713-1838 (assembly :1949-2096 / :2107-2194 + PCG :1789-1790). */
int hs_pressure_correction(hs_ctx* c, hs_stats* st);
/* velocity correction :1459-1571 + pressure update :1574-1622. */
int hs_correct(hs_ctx* c);
/* solveTemp :794-881 (Mode R) / conjugate energy sweeps (Mode C). */
int hs_fluid_energy(hs_ctx* c, int32_t* sweeps, double* residual);

/* ---- linear algebra on the current p' operator (tests, host drivers) ---- */
/* y = A x on the P layout (non-fluid entries of x are ignored, of y are set to 0). */
int hs_pressure_apply(hs_ctx* c, const double* x_host, double* y_host);
/* solve A x = b on the P layout with the configured Krylov method; x on entry is
 * ignored (x0 = 0 as in the reference, fluid_solver.h:1766). */
int hs_pressure_solve(hs_ctx* c, const double* b_host, double* x_host, hs_stats* st);

/* hs_pressure_solve for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart
 * in the reference.  b and x are device pointers of the P layout on the context's GPU;
 * the solve's control block and reduction partials live in a fixed device array of
 * 256.5 KiB per GPU, held until the process ends, so the call allocates nothing.
 * The result and statistics are those of hs_pressure_solve on the same input.
 * Returns HS_E_ARG when the grid's reductions need more partials than the array holds
 * (more than 16384 z-marching column blocks, e.g. nx*ny beyond about 4.2M cells per
 * plane), HS_E_STATE before the p' operator exists or, for HS_PRECOND_GMG, before the
 * first multigrid solve on the context has built its hierarchy.  Calls that overlap on
 * one GPU (other threads, other contexts) share the array and race. */
int hs_pressure_solve_static(hs_ctx* c, const double* b_dev, double* x_dev, hs_stats* st);

/* applyCellBoundaryCondition fluid_solver.h:324-405 on its own (Mode R only):
 * boundary-cell pressure copies, ghost p_temp fill, p(0,0,0) = 0. */
int hs_cell_bc(hs_ctx* c);

/* ---- data ---- */
int hs_get_array(hs_ctx* c, int32_t which, double* host, int64_t n);
int hs_put_array(hs_ctx* c, int32_t which, const double* host, int64_t n);
/* velocities on unpadded faces (u: (nx+1)*ny*nz, v: nx*(ny+1)*nz, w: nx*ny*(nz+1)),
 * p and t on cells (t: fluid temperature in fluid cells, solid temperature in solid
 * cells).  Any pointer may be NULL. */
int hs_get_fields(hs_ctx* c, double* u, double* v, double* w, double* p, double* t);
int hs_put_fields(hs_ctx* c, const double* u, const double* v, const double* w,
                  const double* p, const double* t);
/* Output packing of Solver::solve solver.h:160-213: one OutputData record
 * (fluid_solver.h:38-45) per cell, in the global cell order main.cpp:95-101 writes.
 *   loc  3 per cell: the cell's minimum vertex (solver.h:178-180)
 *   vel  3 per cell: fluid cells their west / south / bottom face velocity (solver.h:172-174),
 *        solid cells 0
 *   p    fluid cells the cell pressure, solid cells 0
 *   t    fluid cells the fluid output temperature, with 293 on cells at the domain
 *        boundary in HS_MODE_REF (fluid_solver.h:394-404); solid cells the solid
 *        temperature (mobs_solver.h:290-301); HS_MODE_CONFIG: the conjugate temperature
 *   id   fluid cells their p' row number, fluid cells counted in the same order
 *        (fluid_solver.h:1966; 0 before the first p' assembly), solid cells -1
 * Any pointer may be NULL. */
int hs_get_output(hs_ctx* c, double* loc, double* vel, double* p, double* t, int32_t* id);
/* Mode C per-cell heat source q (C layout) and conductivity k (C layout). */
int hs_set_source(hs_ctx* c, const double* q);
int hs_set_conductivity(hs_ctx* c, const double* k);

/* ---- energy equation on a frozen flow and its discrete adjoint (Mode C) ----
 * No counterpart in the reference: its temperatures only see under-relaxed sweeps
 * (solveTemp fluid_solver.h:794-881, MobSolver::solveTemperature mobs_solver.h:467-500),
 * whose fixed point is the linear system solved here.  Unknowns: every interior cell
 * (fluid and solid).  Operator: the conjugate energy stencil of the Mode C sweeps
 * (harmonic-mean face conductance, upwind convection with the current face velocities)
 * with the zero-gradient outflow ghost folded into the diagonal; right-hand side:
 * q V plus the Dirichlet ghost temperatures of inlet and walls.  Solver: BiCGStab in
 * FP64 preconditioned by a geometric multigrid V-cycle (FP32 when
 * params.precision == HS_PREC_MIXED) built from the same face data; the transpose
 * solve reuses the hierarchy with the convective fluxes reversed.
 * Vectors below are cell arrays in the C layout (nx*ny*nz, x fastest); `mem` says
 * whether pointers are host memory or device memory of the context's GPU. */
#define HS_MEM_HOST    0
#define HS_MEM_DEVICE  1
typedef struct {
    int32_t iters;                    /* BiCGStab iterations */
    int32_t converged;                /* 1 if ||r|| <= tol ||b|| was met */
    int32_t breakdown;                /* 1 if a BiCGStab scalar vanished and restarts ran out */
    int32_t restarts;                 /* restarts with the current residual as shadow vector */
    double  relres;                   /* recurrence ||r|| / ||b|| at exit */
    double  true_relres;              /* ||b - A x|| / ||b|| recomputed in FP64 */
} hs_solve_info;
/* Solve the energy equation on the current flow, q and k; the result replaces the
 * temperature field (HS_A_TF), outflow ghost included. */
int hs_energy_solve(hs_ctx* c, double tol, int32_t maxit, hs_solve_info* info);
/* Discrete adjoint at the current (solved) temperature: solves A^T lambda = dJdT and
 * returns dJ/dq = lambda V and dJ/dk (all cells; FP64).  Either output may be NULL. */
int hs_adjoint_energy(hs_ctx* c, const double* dJdT, double* dJdq, double* dJdk, double tol,
                      int32_t maxit, int32_t mem, hs_solve_info* info);
/* A x = b (transpose = 0) or A^T x = b (transpose = 1) for a caller right-hand side b
 * with the current operator (tangent / adjoint sensitivities); x0 = 0. */
int hs_energy_solve_rhs(hs_ctx* c, int32_t transpose, const double* b, double* x, double tol,
                        int32_t maxit, int32_t mem, hs_solve_info* info);
/* y = A x (transpose = 0) or A^T x (transpose = 1) of the energy operator. */
int hs_energy_apply(hs_ctx* c, int32_t transpose, const double* x, double* y, int32_t mem);
/* Global energy balance of the current temperature field (Mode C; no counterpart in the
 * reference, which computes no balance): *heat_in = sum q V,
 * *residual = |sum over cells of the energy equation's residual| / |heat_in|, the figure
 * hs_stats.energy_residual reports after each hs_step. */
int hs_energy_balance(hs_ctx* c, double* heat_in, double* residual);
/* Cell arrays in the C layout: HS_A_TF (temperature, interior), HS_A_Q, HS_A_K. */
int hs_get_cells(hs_ctx* c, int32_t which, double* dst, int32_t mem);
int hs_put_cells(hs_ctx* c, int32_t which, const double* src, int32_t mem);

/* ---- live kernel timing (no counterpart in the reference; its only timer is clock()
 *      around outer iterations, solver.h:53-54,156-158) ---- */
#define HS_K_MG_SMOOTH0  1   /* one red-black Gauss-Seidel sweep (fused colours) on the finest multigrid level,
                                 from a nonzero iterate (the zero-guess first sweep is not counted) */
#define HS_K_SPMV        2   /* FP64 p' operator apply + <p,q> of the Krylov iteration */
#define HS_K_MOM_RB      3   /* one red-black colour pass of x-momentum */
#define HS_K_E_SMOOTH0   4   /* one red-black colour pass on the finest energy multigrid level */
/* record HIP events on the context's stream around every launch of one kernel (0: off) */
int hs_prof_set(hs_ctx* c, int32_t kernel);
/* launches timed and their summed duration in ms since hs_prof_set */
int hs_prof_get(hs_ctx* c, int64_t* count, double* total_ms);
/* plain device copy of nbytes (reps times): the known-size reference for hardware byte
 * counters; *avg_ms = mean copy time */
int hs_copy_bench(int64_t nbytes, int32_t reps, double* avg_ms);
/* multigrid level dimensions (3 ints per level); returns the number of levels */
int hs_mg_levels(hs_ctx* c, int32_t* dims, int32_t max_levels);
/* aggregation multigrid levels (no counterpart in the reference: HYPRE keeps BoomerAMG's
 * hierarchy internal) below the finest one (which = 0: the p' hierarchy, 1: the
 * energy hierarchy): rows and stored off-diagonal entries, 2 per level, coarsest last;
 * returns the number of levels (0 before the first build) */
int hs_amg_levels(hs_ctx* c, int32_t which, int64_t* rows_nnz, int32_t max_levels);

#ifdef __cplusplus
}
#endif
#endif /* HSOLVER_H */
