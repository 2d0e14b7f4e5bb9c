/*
 * oracle.h — CPU FP64 restatement of the reference's SIMPLE conjugate-heat-transfer
 * step.  TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the CPU-baseline
 * leg of the benchmark link it as the checker; the product never does.
 *
 * The API mirrors include/hsolver.h one-for-one with an `or_` prefix and takes the
 * same hs_grid / hs_params / hs_stats structs, so a test drives the oracle and the
 * GPU library with the same call sequence and compares the same arrays.
 */
#ifndef HS_ORACLE_H
#define HS_ORACLE_H
#include "../include/hsolver.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx or_ctx;

int     or_create(const hs_grid* g, const hs_params* p, or_ctx** out);
void    or_destroy(or_ctx* c);
int64_t or_array_size(const or_ctx* c, int32_t which);
int     or_init(or_ctx* c);
int     or_fluid_solve(or_ctx* c, hs_stats* st);
int     or_solid_solve(or_ctx* c, hs_stats* st);
int     or_exchange(or_ctx* c);
int     or_set_for_next(or_ctx* c);
int     or_step(or_ctx* c, hs_stats* st);
int     or_run(or_ctx* c, int32_t n, hs_stats* st);
int     or_momentum(or_ctx* c, int32_t comp, int32_t* sweeps, double* residual);
int     or_pressure_correction(or_ctx* c, hs_stats* st);
int     or_correct(or_ctx* c);
int     or_fluid_energy(or_ctx* c, int32_t* sweeps, double* residual);
int     or_pressure_apply(or_ctx* c, const double* x, double* y);
int     or_pressure_solve(or_ctx* c, const double* b, double* x, hs_stats* st);
int     or_get_array(or_ctx* c, int32_t which, double* host, int64_t n);
int     or_put_array(or_ctx* c, int32_t which, const double* host, int64_t n);
int     or_get_fields(or_ctx* c, double* u, double* v, double* w, double* p, double* t);
int     or_set_source(or_ctx* c, const double* q);
int     or_set_conductivity(or_ctx* c, const double* k);
/* sequential k,j,i boundary-cell pressure copy (fluid_solver.h:324-405) on its own */
int     or_cell_bc(or_ctx* c);
/* export the current p' operator as CSR over fluid rows in k,j,i order
 * (row numbering of initMap fluid_solver.h:1891-1913; column order of
 * initColsAndSetValues :2003-2083).  Pass NULL arrays to query nrows / nnz. */
int     or_export_csr(or_ctx* c, int64_t* nrows, int64_t* nnz, int64_t* rowptr,
                      int64_t* cols, double* vals, double* rhs, int64_t* row_cell);
/* energy-equation operator of Mode C on the current (frozen) flow:
 * A T = s over all interior cells, C layout, CSR, the outflow ghost folded into
 * the diagonal (the sweep's fixed point); used by the adjoint checks. */
int     or_export_energy_csr(or_ctx* c, int64_t* nnz, int64_t* rowptr, int64_t* cols,
                             double* vals, double* rhs);

/* energy equation on the frozen flow and its discrete adjoint (see hsolver.h);
 * FP64 BiCGStab with a diagonal preconditioner on the assembled CSR and its
 * explicit transpose; host memory only */
int     or_energy_solve(or_ctx* c, double tol, int32_t maxit, hs_solve_info* info);
int     or_adjoint_energy(or_ctx* c, const double* dJdT, double* dJdq, double* dJdk, double tol,
                          int32_t maxit, int32_t mem, hs_solve_info* info);
int     or_energy_solve_rhs(or_ctx* c, int32_t transpose, const double* b, double* x, double tol,
                            int32_t maxit, int32_t mem, hs_solve_info* info);
int     or_energy_apply(or_ctx* c, int32_t transpose, const double* x, double* y, int32_t mem);
/* Mode C global energy balance of the current field, as hs_energy_balance reports it:
 * *heat_in = sum q V, *residual = |sum over cells of the energy row residual| / |heat_in|
 * (the direct per-cell sum of the step statistics). */
int     or_energy_balance(or_ctx* c, double* heat_in, double* residual);
int     or_get_cells(or_ctx* c, int32_t which, double* dst, int32_t mem);
int     or_put_cells(or_ctx* c, int32_t which, const double* src, int32_t mem);

/* host threads of the p' CG (default 1: the serial order; n > 1 splits it into z-slabs,
 * for the CPU baseline timing) */
void    or_set_threads(int32_t n);
/* test infrastructure of the oracle's own pins: commit planted mistake n (0 = none) */
void    or_test_plant(int32_t n);

#ifdef __cplusplus
}
#endif
#endif
