// hs_internal.h — device context and helpers shared by the gfx950 kernels and the
// C-ABI host code of libhsolver_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <algorithm>
#include <cstdio>
#include <vector>

#include "../../include/hsolver.h"

#define HS_CHECK(expr)                                                              \
    do {                                                                            \
        hipError_t _e = (expr);                                                     \
        if (_e != hipSuccess) {                                                     \
            fprintf(stderr, "hsolver: HIP error %s at %s:%d\n", hipGetErrorString(_e), \
                    __FILE__, __LINE__);                                            \
            return HS_E_HIP;                                                        \
        }                                                                           \
    } while (0)

#define HS_TRY(expr)                  \
    do {                              \
        int _r = (expr);              \
        if (_r != HS_OK) return _r;   \
    } while (0)

// Grid extents and strides for the padded layouts (x fastest).
struct Dims {
    int nx, ny, nz;
    // P layout (cells): (nx+2)(ny+2)(nz+2); W layout shares the i,j strides with P.
    long sPy, sPz;
    long sUy, sUz;   // U layout (nx+3)(ny+2)(nz+2)
    long sVy, sVz;   // V layout (nx+2)(ny+3)(nz+2)
    long nP, nU, nV, nW, nC;
};

inline Dims make_dims(int nx, int ny, int nz) {
    Dims d;
    d.nx = nx; d.ny = ny; d.nz = nz;
    d.sPy = nx + 2; d.sPz = (long)(nx + 2) * (ny + 2);
    d.sUy = nx + 3; d.sUz = (long)(nx + 3) * (ny + 2);
    d.sVy = nx + 2; d.sVz = (long)(nx + 2) * (ny + 3);
    d.nP = d.sPz * (nz + 2);
    d.nU = d.sUz * (nz + 2);
    d.nV = d.sVz * (nz + 2);
    d.nW = d.sPz * (nz + 3);
    d.nC = (long)nx * ny * nz;
    return d;
}

__host__ __device__ inline long idxP(const Dims& d, long i, long j, long k) { return i + d.sPy * j + d.sPz * k; }
__host__ __device__ inline long idxU(const Dims& d, long i, long j, long k) { return i + d.sUy * j + d.sUz * k; }
__host__ __device__ inline long idxV(const Dims& d, long i, long j, long k) { return i + d.sVy * j + d.sVz * k; }
__host__ __device__ inline long idxW(const Dims& d, long i, long j, long k) { return i + d.sPy * j + d.sPz * k; }
__host__ __device__ inline long idxC(const Dims& d, long i, long j, long k) { return i + (long)d.nx * (j + (long)d.ny * k); }

// Cell classes stored per padded cell (P layout).
enum : uint8_t { CELL_GHOST = 0, CELL_FLUID = 1, CELL_SOLID = 2 };

// 1D geometry per padded index: width of the cell, of the minus neighbour, of the plus
// neighbour (the reference's xp/xw/xe, yp/ys/yn, zp/zf/zb with its end fallbacks).
struct Geo {
    const double *xp, *xw, *xe, *yp, *ys, *yn, *zp, *zf, *zb;
};

// Scalar state of one Krylov solve or one sweep loop, resident in device memory so the
// control decisions are taken on the GPU without a host round trip per iteration.
struct SolveCtl {
    double bb;         // <b,b>
    double gamma;      // <r,z>
    double gamma_old;
    double alpha;
    double beta;
    double rr;         // <r,r> at the last update
    double sdotp;
    double eps;        // tol^2
    int    iters;
    int    done;       // 1: stop launching (converged / breakdown / maxit / non-finite)
    int    converged;
    int    maxit;
    int    nan;        // 1: a Krylov reduction came out NaN or Inf (the solve stopped there)
    int    cold;       // 1: a warm start's residual was not below ||b||: start from zero instead
    double red[8];     // generic reduction results
    double cg_g[2];    // unpreconditioned CG: <r_k, r_k> in slot k & 1 (a kernel that forms
                       // the next one writes the other slot while its blocks read this one)
};

struct SweepCtl {
    double residual;   // reference accumulator (Mode R) or last sweep value
    double global;     // sqrt(...) compared with the tolerance
    double ncount;     // normaliser N
    double tol;
    int    step;
    int    maxsteps;
    int    active;     // 1 while another sweep is due
    int    quirk;      // 1: accumulator never reset (Mode R, fluid); 2: solid variant
};

constexpr int RED_THREADS = 256;
constexpr int MAX_RED_BLOCKS = 4096;

// Multigrid level (Mode C pressure operator on coarser grids); arrays are float or
// double (f32) in the padded layout (nx+2)(ny+2)(nz+2) of the level.
struct MGLevel {
    int level;
    int nx, ny, nz;
    long sy, sz, n;
    bool f32;
    void *cw, *cs, *cb, *g, *x, *b;
    void* x2;          // second iterate buffer of the out-of-place fused red-black sweep
    uint8_t* act;      // 1 where the level operator has a row (D > 0), set at every build
};

// Control block of the energy-equation BiCGStab (device resident, see hs_energy.hip).
struct BiCtl {
    double bb, rr, rho, alpha, omega, beta, eps, ts, tt, ss;
    double rhn;        // <r^, r^> of the current shadow residual
    double red[4];
    int iters, maxit, done, converged, breakdown, half;
    int skip;          // 1 once the rest of the iteration must not run (done, half step, restart)
    int restart;       // 1: the next iteration starts with r^ = r (after a (near) breakdown)
    int restarts;      // restarts so far
    int wbrk;          // omega could not be formed this iteration
};

// One level of the energy-equation hierarchy: face conductances d* and mass fluxes f*
// on the min face of every cell (padded layout of the level), iterate x, right-hand side b.
struct ELevel {
    int nx, ny, nz;
    long sy, sz, n;
    void *dx, *dy, *dz, *fx, *fy, *fz, *x, *b;
    void* x2;          // second iterate buffer (the fused sweep runs out of place, then they swap)
    bool own_faces;    // false when the faces alias the FP64 fine operator
};

struct AmgHier;

struct EnergyState {
    bool f32 = false;
    std::vector<ELevel*> lv;
    double *dx = nullptr, *dy = nullptr, *dz = nullptr, *fx = nullptr, *fy = nullptr, *fz = nullptr;  // FP64 fine operator
    double* rhs = nullptr;                     // assembled right-hand side (P layout)
    double *x = nullptr, *r = nullptr, *rh = nullptr, *p = nullptr, *v = nullptr, *t = nullptr,
           *ph = nullptr, *sh = nullptr;       // BiCGStab vectors (P layout)
    double* stage = nullptr;                   // nx*ny*nz staging for host buffers
    double* stage2 = nullptr;
    double* stage3 = nullptr;
    BiCtl* ctl = nullptr;
    BiCtl* h_ctl = nullptr;                    // pinned mirror
    int last_iters = 0;
    long build_id = 0;                         // operator builds so far
    AmgHier* amg = nullptr;                    // aggregation multigrid (e_precond = HS_PRECOND_AMG)
};

// Dense inverse of the coarsest multigrid level (hs_dense.hip), in the scalar of the
// hierarchy: FP32 under the mixed-precision V-cycle, FP64 under the FP64 one.
struct DenseCoarse {
    int np = 0;                    // order, padded to a multiple of 64
    bool f64 = false;              // the matrix scalar is double (else float)
    void* M = nullptr;             // np x np, row-major: the level matrix, then its inverse
    void* C = nullptr;             // np x 64 panel of the Gauss-Jordan step
    void* P = nullptr;             // 64 x 64 pivot-block inverse
    double *r = nullptr, *z = nullptr;   // right-hand side, solution (FP64)
    void* xf = nullptr;                  // input of the inverse apply, in the matrix scalar
    long builds = 0, factorizations = 0;
    long since = 0;                // hierarchy builds since the last factorization
    int iters_ref = -1;            // FCG iterations of the first solve after it (-1: none yet)
};

// Aggregation multigrid (hs_amg.hip): the finest level is a structured 7-point operator
// (the pressure multigrid's level 0 or the energy hierarchy's level 0); every coarser
// level is a CSR matrix over aggregates of the level above it.
struct AmgLevel {
    long n = 0, nnz = 0;            // rows, stored off-diagonal entries
    int *rp = nullptr, *ci = nullptr;
    double *vd = nullptr, *dd = nullptr;   // entries and diagonal in FP64 (setup, refinement)
    void* dg = nullptr;                    // diagonal in the cycle's scalar (alias when FP64)
    // the cycle's copy of the entries: sliced ELL, 64 rows (one wavefront) per slice, each
    // slice padded to its longest row and stored slot-major, so that the 64 lanes of a
    // wavefront read 64 consecutive entries
    int *so = nullptr, *sw = nullptr, *sc = nullptr;   // slice offsets, widths, columns
    void* sv = nullptr;                    // entries in the cycle's scalar
    long nsell = 0;
    int wb = 0, wbi = 0;                   // bounds of the row lengths (level, intermediate)
    void *x = nullptr, *x2 = nullptr, *b = nullptr, *r = nullptr;
    int* agg = nullptr;             // rows of the level above -> rows of this level (-1: none)
    int* memb = nullptr;            // 4 per row: rows of the level above in this aggregate (-1: none)
    // the first pairwise pass's intermediate matrix (kept for the numeric rebuild)
    long ni = 0, nnzi = 0;
    int *rpi = nullptr, *cii = nullptr, *membi = nullptr, *aggi = nullptr;
    double *vdi = nullptr, *ddi = nullptr;
    int* membf = nullptr;           // 2 per intermediate row: rows of the level above
    int* aggb = nullptr;            // intermediate rows -> rows of this level
};

struct AmgHier {
    bool f32 = false;
    int kind = -1;                  // 0: p' operator, 1: energy A, 2: energy A^T
    long n0 = 0;                    // rows of the structured finest level (padded cells)
    std::vector<AmgLevel*> lv;      // aggregated levels, coarsest last
    DenseCoarse* dense = nullptr;   // inverse of the coarsest level when it is small enough
    int* scratch = nullptr;         // matching / numbering work (ints)
    long scratch_n = 0;
    int* dflag = nullptr;           // device: [0] row-width overflow, [2..3] scan total (64-bit)
    int* h_flag = nullptr;          // pinned mirror
    long builds = 0, aggregations = 0, since = 0;
    int iters_ref = -1;             // p' FCG iterations of the first solve after an aggregation
    long built_id = -1;             // energy: the operator build the hierarchy was made from
};

struct hs_ctx {
    hs_params prm;
    Dims d;
    hipStream_t stream = nullptr;
    int device = 0;
    bool inited = false;
    bool have_matrix = false;
    // warm starts of the SIMPLE p' solve: the consecutive solutions held (kx = the last,
    // khist[0] the one before, khist[1] the one before that)
    int kx_hist = 0;
    double* khist[2] = {nullptr, nullptr};
    bool defer_balance = false;  // hs_step reads the energy balance at its final synchronisation
    // fixed-count sweep loops inside hs_step (Mode C: no tolerance) queue their control
    // block into h_wdefer[id] and hs_step reads the counts and residuals at its final
    // synchronisation; wdefer[id] marks a slot to read
    SweepCtl* h_wdefer = nullptr;       // pinned, 5 slots (momentum x/y/z, energy, solid)
    bool wdefer[5] = {false, false, false, false, false};

    // host copies
    std::vector<double> X, Y, Z;
    std::vector<uint8_t> type;          // C layout
    int solid_regions = 0;
    int sb[6] = {0, -1, 0, -1, 0, -1};  // solid bbox (padded, inclusive)
    double n_solved[3] = {0, 0, 0};     // Mode C solved-face counts
    double n_tcells = 0;                // Mode C energy cells (fluid + solid)
    int gs_last[5] = {0, 0, 0, 0, 0};   // sweeps the last Gauss-Seidel loop of each kind ran

    // device geometry and classes
    double* geo = nullptr;              // 9 arrays of length max(n)+3 packed
    Geo g{};
    uint8_t* cls = nullptr;             // P layout CELL_*

    // fields
    double *u = nullptr, *v = nullptr, *w = nullptr, *up = nullptr, *vp = nullptr, *wp = nullptr;
    double *p = nullptr, *psnap = nullptr, *ptemp = nullptr, *pprime = nullptr, *tf = nullptr, *ts = nullptr;
    double *q = nullptr, *k = nullptr, *b = nullptr;
    // the heat input sum q V of the energy balance, kept on the device and recomputed only
    // after q changed (q_version counts the writes of q)
    double* qsum = nullptr;
    long q_version = 0, qsum_version = -1;
    // p' between the solve and the correction: the solve's output kx stands for the
    // pprime array (no copy), and the pprime buffer itself stays zero unless written
    // through hs_put_array (pprime_dirty)
    bool pprime_in_kx = false, pprime_dirty = false;
    // p' operator: Mode R per-cell row coefficients (P), Mode C per-face (U/V/W)
    double *ax = nullptr, *ay = nullptr, *az = nullptr;
    // Krylov work (P layout)
    double *kr = nullptr, *kp = nullptr, *ks = nullptr, *kx = nullptr, *kz = nullptr, *kq = nullptr;
    double* kp2 = nullptr;   // second direction buffer of flexible CG (fused update + apply)
    double* kb = nullptr;               // right-hand side staged from the host
    // per-point squared change for lexicographic sweeps (max layout size)
    double* dres = nullptr;
    // reductions
    double* partials = nullptr;
    SolveCtl* sctl = nullptr;
    SweepCtl* wctl = nullptr;
    SolveCtl* h_sctl = nullptr;         // pinned mirrors
    SweepCtl* h_wctl = nullptr;
    // interface pairs (Mode R)
    long* if_fluid = nullptr;
    long* if_solid = nullptr;
    long n_if = 0;

    // multigrid hierarchy (created on demand)
    std::vector<MGLevel*> mg;
    // energy equation solve / adjoint state (created on demand)
    EnergyState* en = nullptr;
    // dense coarsest-level inverse (created on demand)
    DenseCoarse* dense = nullptr;
    // aggregation multigrid of the p' operator (HS_PRECOND_AMG, created on demand)
    AmgHier* amgp = nullptr;
    int last_iters = 0;

    // live per-kernel timing: HIP events around every launch of one chosen kernel
    int prof_kernel = 0;
    std::vector<hipEvent_t> prof_ev;    // pairs (start, stop)
    int prof_used = 0;                  // events recorded since the last fold
    long prof_count = 0;
    double prof_ms = 0.0;
};

int prof_begin(hs_ctx* c, int kernel);   // returns slot or -1
int prof_end(hs_ctx* c, int slot);

// kinds of the one-block reduction finish (k_finish)
enum FinishKind : int {
    FIN_SWEEP = 0,      // Gauss-Seidel loop control (wctl)
    FIN_CG_SETUP,       // <b,b>
    FIN_CG_ALPHA,       // <p,Ap> -> alpha, breakdown checks (flexible CG's first apply)
    FIN_CG_UPDATE,      // <r,r> of CG update `it` -> HYPRE stop rule, beta (end of a batch)
    FIN_DOT,            // plain sum -> *out
    FIN_FCG_SETUP,      // <b,b>
    FIN_FCG_START,      // <r,z> -> gamma
    FIN_FCG_UPDATE,     // <r,r> -> stop rule
    FIN_FCG_BETA,       // <z,r>, <z,q> -> beta, gamma
    FIN_DOT2,           // two plain sums -> out[0], out[1]
    FIN_FCG_SETUP_WARM, // <b,b>, <r,r> of a nonzero initial iterate
    FIN_DIFF,           // plain sum minus out[1] -> out[0]
};

// the array that holds p' now (the solve's output, or the pprime buffer)
inline double* pprime_now(const hs_ctx* c) { return c->pprime_in_kx ? c->kx : c->pprime; }

// ---- kernel launchers (defined in the .hip files) ----
int launch_finish(hs_ctx* c, int kind, int n, double* out);
// FIN_CG_UPDATE over the n partials at part, for CG update number it (0-based)
int launch_finish_cg(hs_ctx* c, const double* part, int n, int it);
int launch_fill(hs_ctx* c, double* a, long n, double v);
int launch_copy(hs_ctx* c, double* dst, const double* src, long n);
int bc_x_ref(hs_ctx* c);
int bc_y_ref(hs_ctx* c);
int bc_z_ref(hs_ctx* c);
int cell_bc_ref(hs_ctx* c);
int ptemp_from_p_all(hs_ctx* c);
int bc_cfg_init(hs_ctx* c);
int bc_cfg_outflow(hs_ctx* c, int comp);
int zero_interface_faces(hs_ctx* c);
int momentum_sweep(hs_ctx* c, int comp);            // one sweep incl. residual -> wctl
int sweep_begin(hs_ctx* c, double n, double tol, int maxsteps, int quirk);
int assemble_matrix(hs_ctx* c);
int assemble_rhs(hs_ctx* c);
int spmv(hs_ctx* c, const double* x, double* y);
int dot(hs_ctx* c, const double* a, const double* b, double* out_host);
// with_resid: the true residual ||b - Ax|| / ||b|| and ||b|| (p_true_relres, mass_imbalance)
// are formed on the device before each readback of the solve's control block and read back
// with it, so a converged solve returns them without another synchronisation
int cg_solve(hs_ctx* c, const double* b, double* x, hs_stats* st, bool with_resid = false);
int cg_spmv_dot(hs_ctx* c, const void* p, bool f32, double* q);
// flexible CG's warm start: r = b - A x, level-0 right-hand side (f32: float) = r, the
// level-0 iterate zeroed when zero_x0, and FIN_FCG_SETUP_WARM's <b,b>, <r,r>
int warm_residual(hs_ctx* c, const double* b, const double* x, double* r, void* b0, void* x0, bool zero_x0,
                  bool f32);
int fcg_spmv_dot(hs_ctx* c, const void* z, bool f32, const void* p, void* pn, double* q);
int mask_fluid(hs_ctx* c, double* x);
int true_relres(hs_ctx* c, const double* b, const double* x, double* relres, double* bnorm);
// queues ||b - Ax||^2 and ||b||^2 into sctl->red[0], red[1] (no synchronisation)
int launch_resid_norms(hs_ctx* c, const double* b, const double* x);
void resid_from_red(const SolveCtl* h, hs_stats* st);
int correct_and_update(hs_ctx* c);   // velocity correction, p += relax_p p', p' = 0
int energy_sweep(hs_ctx* c);
int solid_sweep(hs_ctx* c);
int exchange_kernel(hs_ctx* c);
int energy_balance(hs_ctx* c, double* heat, double* resid);
// the same sums queued with an asynchronous copy into h_sctl->red[1] (residual) and red[2]
// (heat input), read after the caller's next synchronisation
int energy_balance_async(hs_ctx* c);
int gmg_setup(hs_ctx* c);
int fcg_solve(hs_ctx* c, const double* b, double* x, hs_stats* st, bool with_resid = false);
// hs_pressure_solve's device part on the right-hand side staged in c->kb -> c->kx
int pressure_solve_staged(hs_ctx* c, hs_stats* st);
void gmg_free(hs_ctx* c);
// dense coarsest-level solve (hs_dense.hip)
int dense_max_cells();
int dense_setup(hs_ctx* c, const MGLevel* m, bool f32);
int dense_solve(hs_ctx* c, const MGLevel* m, bool f32);
void dense_free(hs_ctx* c);
// dense inverse of a CSR level (order n <= dense_max_cells()): factor, and the per-cycle
// apply x = A^-1 b (b, x in the scalar of the inverse) with `refine` FP64 refinement steps
// against the level's FP64 entries; kernels return at once while *skip != 0
int dense_factor_csr(hs_ctx* c, DenseCoarse** D, int n, const int* rp, const int* ci, const double* vd,
                     const double* dd, bool f32);
int dense_solve_csr(hs_ctx* c, DenseCoarse* D, int n, const int* rp, const int* ci, const double* vd,
                    const double* dd, const void* b, void* x, bool f32, int refine, const int* skip);
void dense_release(DenseCoarse** D);
// aggregation multigrid (hs_amg.hip).  The structured finest level is the pressure
// multigrid's level 0 (amg_*_p) or the energy hierarchy's level 0 with A (tr = 0) or A^T
// (tr = 1) (amg_*_e).  amg_setup_* builds or refreshes the aggregated levels from it;
// amg_coarse_* runs the part of one cycle below the finest level: restrict b - A x of
// the finest level, cycle the aggregated levels, add the scaled correction to x.
int amg_setup_p(hs_ctx* c);
int amg_coarse_p(hs_ctx* c, const int* skip);
int amg_setup_e(hs_ctx* c, ELevel* l0, int tr);
int amg_coarse_e(hs_ctx* c, ELevel* l0, int tr, const int* skip);
void amg_free(AmgHier** H);
// energy equation on the frozen flow (hs_energy.hip); vectors in the C layout are device pointers
int energy_solve(hs_ctx* c, double tol, int maxit, hs_solve_info* info);
int energy_adjoint(hs_ctx* c, const double* dJdT, double* dJdq, double* dJdk, double tol, int maxit,
                   hs_solve_info* info);
int energy_apply(hs_ctx* c, int transpose, const double* x, double* y);
int energy_solve_rhs(hs_ctx* c, int transpose, const double* b, double* x, double tol, int maxit, hs_solve_info* info);
void energy_free(hs_ctx* c);
int energy_stage(hs_ctx* c, double** s1, double** s2, double** s3);   // nx*ny*nz device staging
int cells_from_padded(hs_ctx* c, const double* P, double* C);        // interior of a P array -> C layout
int cells_to_padded(hs_ctx* c, const double* C, double* P);          // C layout -> interior of a P array
