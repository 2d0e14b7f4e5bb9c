// hs_api.cpp — C ABI of libhsolver_amd (declared in include/hsolver.h).
//
// Host-side sequencing of the reference's solver methods on one HIP stream.  Each
// entry point corresponds to a reference member function (cited at its definition);
// all arithmetic runs in the gfx950 kernels of hs_kernels.hip / hs_mg.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include "hs_internal.h"

namespace {

/* fluid_solver.h:1248-1261 cell widths with the reference's end fallbacks. */
void axis_geom(const std::vector<double>& V, int n, double* pw, double* mw, double* plw) {
    const long S = n + 1;
    for (long i = 0; i < n + 3; ++i) { pw[i] = mw[i] = plw[i] = 0.0; }
    for (long i = 1; i <= n + 1; ++i) {
        const double c = (i < S) ? (V[i] - V[i - 1]) : (V[i - 1] - V[i - 2]);
        const double m = (i - 2 >= 0) ? (V[i - 1] - V[i - 2]) : c;
        const double p = (i + 1 < S) ? (V[i + 1] - V[i]) : c;
        pw[i] = c; mw[i] = m; plw[i] = p;
    }
}

template <typename T>
int dalloc(T** p, size_t n) {
    if (hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return HS_E_NOMEM;
    if (hipMemset(*p, 0, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return HS_E_HIP;
    return HS_OK;
}

double* array_ptr(hs_ctx* c, int which) {
    switch (which) {
        case HS_A_U: return c->u; case HS_A_V: return c->v; case HS_A_W: return c->w;
        case HS_A_UP: return c->up; case HS_A_VP: return c->vp; case HS_A_WP: return c->wp;
        case HS_A_P: return c->p; case HS_A_PTEMP: return c->ptemp; case HS_A_PPRIME: return pprime_now(c);
        case HS_A_TF: return c->tf; case HS_A_TS: return c->ts; case HS_A_Q: return c->q;
        case HS_A_K: return c->k; case HS_A_B: return c->b;
    }
    return nullptr;
}

int64_t array_len(const hs_ctx* c, int which) {
    const Dims& d = c->d;
    switch (which) {
        case HS_A_U: case HS_A_UP: return d.nU;
        case HS_A_V: case HS_A_VP: return d.nV;
        case HS_A_W: case HS_A_WP: return d.nW;
        case HS_A_P: case HS_A_Q: case HS_A_K: return d.nC;
        case HS_A_PTEMP: case HS_A_PPRIME: case HS_A_TF: case HS_A_TS: case HS_A_B: return d.nP;
    }
    return -1;
}

// connected solid regions (6-connectivity, mesh.cpp:298-454); bbox of the first
int solid_regions(hs_ctx* c) {
    const Dims& d = c->d;
    std::vector<int> lab(d.nC, -1);
    std::vector<long> stack;
    int nreg = 0;
    for (long c0 = 0; c0 < d.nC; ++c0) {
        if (c->type[c0] != 1 || lab[c0] >= 0) continue;
        int bb[6] = {1 << 30, -1, 1 << 30, -1, 1 << 30, -1};
        stack.assign(1, c0);
        lab[c0] = nreg;
        while (!stack.empty()) {
            const long q = stack.back();
            stack.pop_back();
            const long i = q % d.nx, j = (q / d.nx) % d.ny, k = q / ((long)d.nx * d.ny);
            bb[0] = std::min<long>(bb[0], i + 1); bb[1] = std::max<long>(bb[1], i + 1);
            bb[2] = std::min<long>(bb[2], j + 1); bb[3] = std::max<long>(bb[3], j + 1);
            bb[4] = std::min<long>(bb[4], k + 1); bb[5] = std::max<long>(bb[5], k + 1);
            const long nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
            for (auto& t : nb) {
                if (t[0] < 0 || t[0] >= d.nx || t[1] < 0 || t[1] >= d.ny || t[2] < 0 || t[2] >= d.nz) continue;
                const long cn = idxC(d, t[0], t[1], t[2]);
                if (c->type[cn] == 1 && lab[cn] < 0) { lab[cn] = nreg; stack.push_back(cn); }
            }
        }
        if (nreg == 0) std::memcpy(c->sb, bb, sizeof bb);
        ++nreg;
    }
    return nreg;
}

// interface face pairs in the reference's list order: x faces (k,j,i), y faces, z faces
// (fluid_solver.h:900-1229 / mobs_solver.h:121-268).
void interface_pairs(const hs_ctx* c, std::vector<long>& F, std::vector<long>& S) {
    const Dims& d = c->d;
    auto fl = [&](long i, long j, long k) { return c->type[idxC(d, i - 1, j - 1, k - 1)] == 0; };
    for (long k = 1; k <= d.nz; ++k)
        for (long j = 1; j <= d.ny; ++j)
            for (long i = 2; i <= d.nx; ++i) {
                const bool fa = fl(i - 1, j, k), fb = fl(i, j, k);
                if (fa && !fb) { F.push_back(idxP(d, i - 1, j, k)); S.push_back(idxP(d, i, j, k)); }
                else if (!fa && fb) { F.push_back(idxP(d, i, j, k)); S.push_back(idxP(d, i - 1, j, k)); }
            }
    for (long k = 1; k <= d.nz; ++k)
        for (long j = 2; j <= d.ny; ++j)
            for (long i = 1; i <= d.nx; ++i) {
                const bool fa = fl(i, j - 1, k), fb = fl(i, j, k);
                if (fa && !fb) { F.push_back(idxP(d, i, j - 1, k)); S.push_back(idxP(d, i, j, k)); }
                else if (!fa && fb) { F.push_back(idxP(d, i, j, k)); S.push_back(idxP(d, i, j - 1, k)); }
            }
    for (long k = 2; k <= d.nz; ++k)
        for (long j = 1; j <= d.ny; ++j)
            for (long i = 1; i <= d.nx; ++i) {
                const bool fa = fl(i, j, k - 1), fb = fl(i, j, k);
                if (fa && !fb) { F.push_back(idxP(d, i, j, k - 1)); S.push_back(idxP(d, i, j, k)); }
                else if (!fa && fb) { F.push_back(idxP(d, i, j, k)); S.push_back(idxP(d, i, j, k - 1)); }
            }
}

double count_solved(const hs_ctx* c, int comp) {
    const Dims& d = c->d;
    auto fl = [&](long i, long j, long k) {
        return i >= 1 && i <= d.nx && j >= 1 && j <= d.ny && k >= 1 && k <= d.nz && c->type[idxC(d, i - 1, j - 1, k - 1)] == 0;
    };
    double n = 0;
    for (long k = 1; k <= d.nz + (comp == 2); ++k)
        for (long j = 1; j <= d.ny + (comp == 1); ++j)
            for (long i = 1; i <= d.nx + (comp == 0); ++i) {
                const bool a = comp == 0 ? fl(i - 1, j, k) : (comp == 1 ? fl(i, j - 1, k) : fl(i, j, k - 1));
                n += (a && fl(i, j, k)) ? 1 : 0;
            }
    return n;
}

bool modeR(const hs_ctx* c) { return c->prm.mode == HS_MODE_REF; }

// read the sweep control block back (one synchronisation)
int read_wctl(hs_ctx* c) {
    HS_CHECK(hipMemcpyAsync(c->h_wctl, c->wctl, sizeof(SweepCtl), hipMemcpyDeviceToHost, c->stream));
    HS_CHECK(hipStreamSynchronize(c->stream));
    return HS_OK;
}

int apply_bc_comp(hs_ctx* c, int comp) {
    if (modeR(c)) return comp == 0 ? bc_x_ref(c) : (comp == 1 ? bc_y_ref(c) : bc_z_ref(c));
    return bc_cfg_outflow(c, comp);
}

// Runs a Gauss-Seidel loop whose stop rule lives in wctl.  The device gates every sweep
// (and its control step) on wctl->active, and the boundary kernels after a sweep are
// idempotent, so a sweep queued past the stop changes nothing.  The host therefore queues
// all sweeps when the count is fixed (tolerance 0), and otherwise as many as this loop
// (`id`) ran last time before its first read of the control block, then one at a time.
template <typename SweepFn, typename AfterFn>
int gs_loop(hs_ctx* c, int id, double n, double tol, int maxsteps, int quirk, SweepFn sweep, AfterFn after,
            int* sweeps, double* resid) {
    HS_TRY(sweep_begin(c, n, tol, maxsteps, quirk));
    const bool fixed = !(tol > 0.0) && quirk == 0;
    const int first = fixed ? maxsteps : std::min(maxsteps, std::max(1, c->gs_last[id]));
    int s = 0;
    for (; s < first; ++s) {
        HS_TRY(sweep());
        HS_TRY(after());
    }
    if (fixed && c->defer_balance && c->h_wdefer) {
        // the sweep count is fixed: nothing to decide on the host, so no synchronisation
        // here; hs_step reads the control block with its final one
        HS_CHECK(hipMemcpyAsync(&c->h_wdefer[id], c->wctl, sizeof(SweepCtl), hipMemcpyDeviceToHost, c->stream));
        c->wdefer[id] = true;
        return HS_OK;
    }
    HS_TRY(read_wctl(c));
    while (c->h_wctl->active && s < maxsteps) {
        HS_TRY(sweep());
        HS_TRY(after());
        ++s;
        HS_TRY(read_wctl(c));
    }
    c->gs_last[id] = c->h_wctl->step;
    if (sweeps) *sweeps = c->h_wctl->step;
    if (resid) *resid = c->h_wctl->global;
    if (std::isnan(c->h_wctl->global)) return HS_E_NAN;
    return HS_OK;
}

int do_momentum(hs_ctx* c, int comp, int* sweeps, double* resid) {
    const Dims& d = c->d;
    double n;
    int quirk;
    if (modeR(c)) {
        const double cx = d.nx, cy = d.ny, cz = d.nz;
        n = comp == 0 ? cz * cy * (cx + 1) : (comp == 1 ? cz * (cy + 1) * cx : (cz + 1) * cy * cx);
        quirk = 1;
    } else {
        n = std::max(1.0, c->n_solved[comp]);
        quirk = 0;
    }
    return gs_loop(
        c, comp, n, c->prm.mom_tol, c->prm.mom_sweeps_max, quirk, [&] { return momentum_sweep(c, comp); },
        [&] { return apply_bc_comp(c, comp); }, sweeps, resid);
}

int do_energy(hs_ctx* c, int* sweeps, double* resid) {
    const Dims& d = c->d;
    if (modeR(c))
        return gs_loop(c, 3, (double)d.nx * d.ny * d.nz, c->prm.t_tol, c->prm.t_sweeps_max, 1,
                       [&] { return energy_sweep(c); }, [&] { return HS_OK; }, sweeps, resid);
    return gs_loop(c, 3, std::max(1.0, c->n_tcells), c->prm.t_tol, c->prm.t_sweeps_max, 0,
                   [&] { return energy_sweep(c); }, [&] { return bc_cfg_outflow(c, 3); }, sweeps, resid);
}

int do_pressure(hs_ctx* c, hs_stats* st) {
    if (!c->have_matrix || !(modeR(c) && c->prm.freeze_p_matrix)) HS_TRY(assemble_matrix(c));
    HS_TRY(assemble_rhs(c));
    hs_stats loc{};
    int e;
    // the true residual of the solution is formed before the solver's last readback
    if (c->prm.precond != HS_PRECOND_NONE) e = fcg_solve(c, c->b, c->kx, &loc, true);
    else e = cg_solve(c, c->b, c->kx, &loc, true);
    c->kx_hist = e == HS_OK ? std::min(c->kx_hist + 1, 3) : 0;
    if (e != HS_OK) return e;
    // p' is kx until the correction consumes it (the pprime buffer is not copied into)
    c->pprime_in_kx = true;
    if (st) {
        st->p_iters = loc.p_iters; st->p_converged = loc.p_converged; st->p_relres = loc.p_relres;
        st->p_true_relres = loc.p_true_relres; st->mass_imbalance = loc.mass_imbalance;
    }
    if (c->prm.fail_on_noconv && !loc.p_converged) return HS_E_NOCONV;
    return HS_OK;
}

int do_correct(hs_ctx* c) {
    HS_TRY(correct_and_update(c));
    if (!modeR(c)) HS_TRY(bc_cfg_outflow(c, 0));
    return HS_OK;
}

}  // namespace

// The solve of hs_pressure_solve on the right-hand side staged in kb: entries outside
// fluid cells zeroed, the configured Krylov method into kx, the true residual and ||b||
// recomputed.  The reductions use c->partials and c->sctl as they are at the call.
int pressure_solve_staged(hs_ctx* c, hs_stats* st) {
    // kx is about to hold another system's solution: a p' still held there moves to the
    // pprime buffer first
    if (c->pprime_in_kx) {
        HS_TRY(launch_copy(c, c->pprime, c->kx, c->d.nP));
        c->pprime_in_kx = false;
        c->pprime_dirty = true;
    }
    c->kx_hist = 0;
    HS_TRY(mask_fluid(c, c->kb));
    hs_stats loc{};
    const int e = (c->prm.precond != HS_PRECOND_NONE) ? fcg_solve(c, c->kb, c->kx, &loc)
                                                     : cg_solve(c, c->kb, c->kx, &loc);
    if (e != HS_OK) return e;
    double rel, bn;
    HS_TRY(true_relres(c, c->kb, c->kx, &rel, &bn));
    loc.p_true_relres = rel;
    loc.mass_imbalance = bn;
    if (st) *st = loc;
    return HS_OK;
}

// Records a start event if kernel is the one being timed; the pool is folded (summed
// after a sync on its last event) when full, so any number of launches can be timed.
int prof_begin(hs_ctx* c, int kernel) {
    if (c->prof_kernel != kernel || c->prof_ev.empty()) return -1;
    if (c->prof_used + 2 > (int)c->prof_ev.size()) {
        if (hipEventSynchronize(c->prof_ev[c->prof_used - 1]) != hipSuccess) return -1;
        for (int e = 0; e < c->prof_used; e += 2) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, c->prof_ev[e], c->prof_ev[e + 1]) == hipSuccess) {
                c->prof_ms += ms;
                c->prof_count += 1;
            }
        }
        c->prof_used = 0;
    }
    const int slot = c->prof_used;
    if (hipEventRecord(c->prof_ev[slot], c->stream) != hipSuccess) return -1;
    c->prof_used += 2;
    return slot;
}

int prof_end(hs_ctx* c, int slot) {
    if (slot < 0) return HS_OK;
    HS_CHECK(hipEventRecord(c->prof_ev[slot + 1], c->stream));
    return HS_OK;
}

extern "C" {

void hs_params_default(int32_t mode, hs_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->mode = mode;
    p->order = HS_ORDER_LEX;
    p->rho = 1.0;
    p->nu = 0.001;
    p->alpha_f = 0.001;
    p->k_ratio = 1.0;
    p->u_lid = 0.1;
    p->relax_u = p->relax_p = p->relax_t = p->relax_ts = 0.3;
    p->mom_tol = 2e-3;
    p->t_tol = 2e-3;
    p->ts_tol = 1e-3;
    p->mom_sweeps_max = p->t_sweeps_max = p->ts_sweeps_max = 100;
    p->p_tol = 1e-5;
    p->p_maxit = 1000;
    p->freeze_p_matrix = 1;
    p->precond = HS_PRECOND_NONE;
    p->precision = HS_PREC_FP64;
    p->mg_levels = 0;   // 0: coarsen until the smallest dimension is <= 4
    p->mg_pre = -1;      /* -1: chosen at hs_create from the grid's cell aspect ratio */
    p->mg_post = -1;
    p->mg_coarse_sweeps = 40;
    p->mg_dense_coarse = 0;
    p->mg_dense_refine = 1;
    p->mg_sweep_rows = 0;
    p->mg_pre0 = -1;
    p->mg_post0 = -1;
    p->mg_cycle = 1;
    p->mg_omega = 0.0f;  /* <= 0: chosen at hs_create from the grid's cell aspect ratio */
    p->mg_omega_coarse = 0.0f;
    p->e_pre = p->e_post = -1;   /* by grid at hs_create */
    p->e_omega = 0.0f;   /* <= 0: 1 (Gauss-Seidel) at hs_create */
    p->t_fluid_init = 293.0;
    p->t_solid_init = 333.0;
    p->t_wall = 293.0;
    if (mode == HS_MODE_CONFIG) {
        p->freeze_p_matrix = 0;
        p->relax_u = 0.7;
        p->relax_p = 0.3;
        p->relax_t = 1.0;
        p->relax_ts = 1.0;
        p->mom_sweeps_max = 1;
        p->t_sweeps_max = 1;
        p->ts_sweeps_max = 0;
        p->mom_tol = 0.0;
        p->t_tol = 0.0;
        p->p_tol = 1e-8;
        p->outflow = 1;
        p->u_lid = 0.0;
        p->u_in = 1.0;
    }
}

int hs_abi_sizes(int32_t* sizes3) {
    if (!sizes3) return HS_E_ARG;
    sizes3[0] = (int32_t)sizeof(hs_grid);
    sizes3[1] = (int32_t)sizeof(hs_params);
    sizes3[2] = (int32_t)sizeof(hs_stats);
    return HS_OK;
}

const char* hs_strerror(int code) {
    switch (code) {
        case HS_OK: return "ok";
        case HS_E_ARG: return "invalid argument or unsupported grid";
        case HS_E_HIP: return "HIP runtime error";
        case HS_E_NOCONV: return "pressure-correction solve did not converge";
        case HS_E_NAN: return "NaN or Inf in a residual";
        case HS_E_IO: return "file could not be read";
        case HS_E_STATE: return "call out of order";
        case HS_E_NOMEM: return "device allocation failed";
    }
    return "unknown error";
}

// Largest ratio of a cell's longest to shortest side: a cell takes one width per axis, so
// the extreme is the widest width of one axis over the narrowest of another.
static double max_cell_aspect(const hs_ctx* c) {
    const std::vector<double>* ax[3] = {&c->X, &c->Y, &c->Z};
    double wmax[3], wmin[3];
    for (int a = 0; a < 3; ++a) {
        wmax[a] = 0.0; wmin[a] = INFINITY;
        for (size_t i = 0; i + 1 < ax[a]->size(); ++i) {
            const double w = (*ax[a])[i + 1] - (*ax[a])[i];
            wmax[a] = std::max(wmax[a], w);
            wmin[a] = std::min(wmin[a], w);
        }
    }
    double r = 1.0;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            if (a != b) r = std::max(r, wmax[a] / wmin[b]);
    return r;
}

// Multigrid smoothing left unset (sweep counts -1, over-relaxation <= 0) is chosen from
// the grid, as measured:
//   near-isotropic cells: one pre- and two post-sweeps on the finest level, V(3,3) below,
//     omega 1.2 on the finest level and 1.6 below (the 256^3 box: 9 FCG iterations in both
//     precisions on every SIMPLE step; 1.3 on the finest level takes 9-10, V(2,1) on the
//     finest level with V(2,2) below takes 11, the same at omega 1 takes 13, and 1.5 or
//     1.7 below the finest level take 10);
//   cells of aspect ratio above 8 couple far more strongly across their thin side, which
//     the point smoother has to carry: V(6,6) and omega 1.7 on every level (the 33.5M-cell
//     wall-stretched channel at aspect 50: 10 FCG iterations, against 22 at omega 1 and
//     66 with the isotropic choice).
static void resolve_sweeps(hs_ctx* c) {
    hs_params& p = c->prm;
    const bool aniso = max_cell_aspect(c) > 8.0;
    if (p.mg_pre < 0) {
        p.mg_pre = aniso ? 6 : 3;
        if (!aniso && p.mg_pre0 < 0) p.mg_pre0 = 1;
    }
    if (p.mg_post < 0) {
        p.mg_post = aniso ? 6 : 3;
        if (!aniso && p.mg_post0 < 0) p.mg_post0 = 2;
    }
    if (!(p.mg_omega > 0.0f)) p.mg_omega = aniso ? 1.7f : 1.2f;
    if (!(p.mg_omega_coarse > 0.0f)) p.mg_omega_coarse = aniso ? 1.7f : 1.6f;
    // energy multigrid sweeps: the pressure cycle's counts.  On the aspect-50 channel V(8,8)
    // took the energy and adjoint solves of one frozen flow from 6.0 to 5.1 s (176 + 187 to
    // 125 + 119 BiCGStab iterations) and those of a later flow from 5.27 to 5.33 s (165 + 157
    // to 136 + 122), so the counts stay shared unless set
    if (p.e_pre < 0) p.e_pre = p.mg_pre;
    if (p.e_post < 0) p.e_post = p.mg_post;
    // energy multigrid (BiCGStab preconditioner): Gauss-Seidel.  The upwind convection
    // operator is not symmetric, so over-relaxation has no convergence guarantee: on the
    // aspect-50 channel omega 1.4 cut the energy and adjoint solves from 142 + 148 to
    // 94 + 105 BiCGStab iterations on one frozen flow and left the adjoint diverging on a
    // later one
    if (!(p.e_omega > 0.0f)) p.e_omega = 1.0f;
    // aggregation multigrid, measured on the 33.5M-cell finned channel (aspect 50, k ratio
    // 1000): two red-black sweeps before and after on the finest level, three Jacobi sweeps
    // on every aggregated level, theta 0.1, the prolongated correction scaled by 1.4 (p') or
    // 1.6 (energy; left <= 0 here and chosen by use).  Energy solve: (1, 2) sweeps at scaling
    // 1 took 242 BiCGStab iterations, at 1.6 109; (2, 3) at 1.6 68, with theta 0.1 59
    // (1.16 s against the geometric hierarchy's 163 in 2.63 s).  p' solve (FCG per SIMPLE
    // iteration): (1, 2) at 1: 48 iterations, (2, 3) at 1: 42, at 1.4: 26, at 1.6: 29, at
    // 1.8: 38
    if (p.amg_sweeps0 <= 0) p.amg_sweeps0 = 2;
    if (p.amg_sweeps <= 0) p.amg_sweeps = 3;
    if (!(p.amg_theta > 0.0f)) p.amg_theta = 0.1f;
    if (!(p.amg_omega > 0.0f)) p.amg_omega = 1.0f;
    if (p.e_precond == 0) p.e_precond = HS_PRECOND_GMG;
}

int hs_create(const hs_grid* g, const hs_params* p, hs_ctx** out) {
    if (!g || !p || !out || g->nx < 3 || g->ny < 3 || g->nz < 3 || !g->xv || !g->yv || !g->zv || !g->cell_type)
        return HS_E_ARG;
    if (p->mode != HS_MODE_REF && p->mode != HS_MODE_CONFIG) return HS_E_ARG;
    if (p->precond < HS_PRECOND_NONE || p->precond > HS_PRECOND_AMG) return HS_E_ARG;
    if (p->e_precond < 0 || p->e_precond > HS_PRECOND_AMG) return HS_E_ARG;   // 0 and GMG: geometric
    hs_ctx* c = new (std::nothrow) hs_ctx();
    if (!c) return HS_E_NOMEM;
    c->prm = *p;
    c->d = make_dims(g->nx, g->ny, g->nz);
    const Dims& d = c->d;
    c->X.assign(g->xv, g->xv + d.nx + 1);
    c->Y.assign(g->yv, g->yv + d.ny + 1);
    c->Z.assign(g->zv, g->zv + d.nz + 1);
    resolve_sweeps(c);
    c->type.assign(g->cell_type, g->cell_type + d.nC);
    for (auto t : c->type) if (t > 1) { delete c; return HS_E_ARG; }
    c->solid_regions = solid_regions(c);
    if (modeR(c)) {
        // one fluid region spanning the domain, at most one solid region (Q16)
        if (c->solid_regions > 1) { delete c; return HS_E_ARG; }
        bool any[6] = {false, false, false, false, false, false};
        for (long k = 0; k < d.nz; ++k)
            for (long j = 0; j < d.ny; ++j)
                for (long i = 0; i < d.nx; ++i)
                    if (c->type[idxC(d, i, j, k)] == 0) {
                        any[0] |= i == 0; any[1] |= i == d.nx - 1; any[2] |= j == 0;
                        any[3] |= j == d.ny - 1; any[4] |= k == 0; any[5] |= k == d.nz - 1;
                    }
        for (bool a : any) if (!a) { delete c; return HS_E_ARG; }
    } else {
        for (int comp = 0; comp < 3; ++comp) c->n_solved[comp] = count_solved(c, comp);
        c->n_tcells = (double)d.nC;
    }
    if (hipGetDevice(&c->device) != hipSuccess) { delete c; return HS_E_HIP; }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return HS_E_HIP; }

    int e = HS_OK;
    auto A = [&](double** ptr, long n) { if (e == HS_OK) e = dalloc(ptr, (size_t)n); };
    const long gl = std::max(d.nx, std::max(d.ny, d.nz)) + 3;
    A(&c->geo, 9 * gl);
    A(&c->u, d.nU); A(&c->up, d.nU); A(&c->v, d.nV); A(&c->vp, d.nV); A(&c->w, d.nW); A(&c->wp, d.nW);
    A(&c->p, d.nC); A(&c->psnap, d.nC); A(&c->q, d.nC); A(&c->k, d.nC);
    A(&c->ptemp, d.nP); A(&c->pprime, d.nP); A(&c->tf, d.nP); A(&c->ts, d.nP); A(&c->b, d.nP);
    A(&c->qsum, 1);
    if (modeR(c)) { A(&c->ax, d.nP); A(&c->ay, d.nP); A(&c->az, d.nP); }
    else { A(&c->ax, d.nU); A(&c->ay, d.nV); A(&c->az, d.nW); }
    A(&c->kb, d.nP); A(&c->kr, d.nP); A(&c->kp, d.nP); A(&c->kp2, d.nP); A(&c->ks, d.nP); A(&c->kx, d.nP); A(&c->kz, d.nP); A(&c->kq, d.nP);
    A(&c->dres, std::max(std::max(d.nU, d.nV), std::max(d.nW, d.nP)));
    {
        // block partials: two per block of the largest z-marching grid (64 x 4 column tiles,
        // z chunks as the launchers choose them), and at least 64K
        const long tiles = (long)((d.nx + 1 + 63) / 64) * ((d.ny + 1 + 3) / 4);
        long want = 4096 / tiles;
        if (want < 1) want = 1;
        const long kc = (d.nz + 1 + want - 1) / want, gz = (d.nz + 1 + kc - 1) / kc;
        A(&c->partials, std::max(65536L, 2 * tiles * gz + 1024));
    }
    if (e == HS_OK) e = dalloc(&c->cls, (size_t)d.nP);
    if (e == HS_OK) e = dalloc(&c->sctl, 1);
    if (e == HS_OK) e = dalloc(&c->wctl, 1);
    if (e == HS_OK && hipHostMalloc((void**)&c->h_sctl, sizeof(SolveCtl), 0) != hipSuccess) e = HS_E_NOMEM;
    if (e == HS_OK && hipHostMalloc((void**)&c->h_wctl, sizeof(SweepCtl), 0) != hipSuccess) e = HS_E_NOMEM;
    if (e == HS_OK && hipHostMalloc((void**)&c->h_wdefer, 5 * sizeof(SweepCtl), 0) != hipSuccess) e = HS_E_NOMEM;
    if (e != HS_OK) { hs_destroy(c); return e; }

    // geometry
    std::vector<double> geo(9 * gl, 0.0);
    axis_geom(c->X, d.nx, &geo[0 * gl], &geo[1 * gl], &geo[2 * gl]);
    axis_geom(c->Y, d.ny, &geo[3 * gl], &geo[4 * gl], &geo[5 * gl]);
    axis_geom(c->Z, d.nz, &geo[6 * gl], &geo[7 * gl], &geo[8 * gl]);
    const int gl_n[9] = {d.nx, d.nx, d.nx, d.ny, d.ny, d.ny, d.nz, d.nz, d.nz};
    (void)gl_n;
    if (hipMemcpy(c->geo, geo.data(), geo.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) { hs_destroy(c); return HS_E_HIP; }
    c->g.xp = c->geo + 0 * gl; c->g.xw = c->geo + 1 * gl; c->g.xe = c->geo + 2 * gl;
    c->g.yp = c->geo + 3 * gl; c->g.ys = c->geo + 4 * gl; c->g.yn = c->geo + 5 * gl;
    c->g.zp = c->geo + 6 * gl; c->g.zf = c->geo + 7 * gl; c->g.zb = c->geo + 8 * gl;

    // cell classes
    std::vector<uint8_t> cls(d.nP, CELL_GHOST);
    for (long k = 1; k <= d.nz; ++k)
        for (long j = 1; j <= d.ny; ++j)
            for (long i = 1; i <= d.nx; ++i)
                cls[idxP(d, i, j, k)] = c->type[idxC(d, i - 1, j - 1, k - 1)] == 0 ? CELL_FLUID : CELL_SOLID;
    if (hipMemcpy(c->cls, cls.data(), d.nP, hipMemcpyHostToDevice) != hipSuccess) { hs_destroy(c); return HS_E_HIP; }

    // interface pairs
    std::vector<long> F, S;
    interface_pairs(c, F, S);
    c->n_if = (long)F.size();
    if (c->n_if > 0) {
        if (dalloc(&c->if_fluid, F.size()) != HS_OK || dalloc(&c->if_solid, S.size()) != HS_OK) { hs_destroy(c); return HS_E_NOMEM; }
        hipMemcpy(c->if_fluid, F.data(), F.size() * sizeof(long), hipMemcpyHostToDevice);
        hipMemcpy(c->if_solid, S.data(), S.size() * sizeof(long), hipMemcpyHostToDevice);
    }

    // initial fills (fluid_solver.h:98-108, mobs_solver.h:55)
    std::vector<double> kk(d.nC);
    for (long i = 0; i < d.nC; ++i) kk[i] = c->type[i] == 1 ? p->alpha_f * p->k_ratio : p->alpha_f;
    if (hipMemcpy(c->k, kk.data(), d.nC * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) { hs_destroy(c); return HS_E_HIP; }
    if (modeR(c)) {
        e = launch_fill(c, c->tf, d.nP, p->t_fluid_init);
    } else {
        std::vector<double> t(d.nP, p->t_fluid_init);
        for (long k = 0; k < d.nz + 2; ++k)
            for (long j = 0; j < d.ny + 2; ++j)
                for (long i = 0; i < d.nx + 2; ++i)
                    if (i == 0 || j == 0 || k == 0 || i == d.nx + 1 || j == d.ny + 1 || k == d.nz + 1)
                        t[idxP(d, i, j, k)] = p->t_wall;
        if (hipMemcpy(c->tf, t.data(), d.nP * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) { hs_destroy(c); return HS_E_HIP; }
    }
    if (e == HS_OK) e = launch_fill(c, c->ts, d.nP, p->t_solid_init);
    if (e == HS_OK && hipStreamSynchronize(c->stream) != hipSuccess) e = HS_E_HIP;
    if (e != HS_OK) { hs_destroy(c); return e; }
    *out = c;
    return HS_OK;
}

void hs_destroy(hs_ctx* c) {
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    gmg_free(c);
    energy_free(c);
    double* bufs[] = {c->geo, c->u, c->up, c->v, c->vp, c->w, c->wp, c->p, c->psnap, c->q, c->qsum, c->k, c->ptemp,
                      c->pprime, c->tf, c->ts, c->b, c->ax, c->ay, c->az, c->kr, c->kp, c->kp2, c->ks, c->kx,
                      c->kz, c->kq, c->kb, c->dres, c->partials, c->khist[0], c->khist[1]};
    for (double* b : bufs) if (b) hipFree(b);
    if (c->cls) hipFree(c->cls);
    if (c->sctl) hipFree(c->sctl);
    if (c->wctl) hipFree(c->wctl);
    for (hipEvent_t e : c->prof_ev) hipEventDestroy(e);
    if (c->if_fluid) hipFree(c->if_fluid);
    if (c->if_solid) hipFree(c->if_solid);
    if (c->h_sctl) hipHostFree(c->h_sctl);
    if (c->h_wctl) hipHostFree(c->h_wctl);
    if (c->h_wdefer) hipHostFree(c->h_wdefer);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int64_t hs_array_size(const hs_ctx* c, int32_t which) { return c ? array_len(c, which) : -1; }

int hs_get_params(const hs_ctx* c, hs_params* p) {
    if (!c || !p) return HS_E_ARG;
    *p = c->prm;
    return HS_OK;
}

/* FluidSolver::init fluid_solver.h:882-1237 + MobSolver::init mobs_solver.h:116-288. */
int hs_init(hs_ctx* c) {
    if (!c) return HS_E_ARG;
    if (modeR(c)) {
        HS_TRY(bc_x_ref(c)); HS_TRY(bc_y_ref(c)); HS_TRY(bc_z_ref(c));
        HS_TRY(cell_bc_ref(c));
        HS_TRY(zero_interface_faces(c));
    } else {
        HS_TRY(bc_cfg_init(c));
    }
    HS_CHECK(hipStreamSynchronize(c->stream));
    c->inited = true;
    c->kx_hist = 0;
    return HS_OK;
}

/* solveXmomentum :577 / solveYmomentum :491 / solveZmomentum :407. */
int hs_momentum(hs_ctx* c, int32_t comp, int32_t* sweeps, double* residual) {
    if (!c || comp < 0 || comp > 2) return HS_E_ARG;
    int s = 0; double r = 0;
    HS_TRY(do_momentum(c, comp, &s, &r));
    if (sweeps) *sweeps = s;
    if (residual) *residual = r;
    return HS_OK;
}

int hs_pressure_correction(hs_ctx* c, hs_stats* st) {
    if (!c) return HS_E_ARG;
    return do_pressure(c, st);
}

int hs_correct(hs_ctx* c) {
    if (!c) return HS_E_ARG;
    HS_TRY(do_correct(c));
    HS_CHECK(hipStreamSynchronize(c->stream));
    return HS_OK;
}

int hs_fluid_energy(hs_ctx* c, int32_t* sweeps, double* residual) {
    if (!c) return HS_E_ARG;
    int s = 0; double r = 0;
    HS_TRY(do_energy(c, &s, &r));
    if (sweeps) *sweeps = s;
    if (residual) *residual = r;
    return HS_OK;
}

/* FluidSolver::solve fluid_solver.h:1401-1633. */
int hs_fluid_solve(hs_ctx* c, hs_stats* st) {
    if (!c) return HS_E_ARG;
    if (!c->inited) return HS_E_STATE;
    hs_stats loc{};
    if (st) { loc.ts_sweeps = st->ts_sweeps; loc.ts_residual = st->ts_residual; }
    for (int comp = 0; comp < 3; ++comp) HS_TRY(do_momentum(c, comp, &loc.mom_sweeps[comp], &loc.mom_residual[comp]));
    HS_TRY(do_pressure(c, &loc));
    HS_TRY(do_correct(c));
    HS_TRY(do_energy(c, &loc.t_sweeps, &loc.t_residual));
    if (!modeR(c)) {
        if (c->defer_balance) {
            HS_TRY(energy_balance_async(c));   // hs_step reads it after its last synchronisation
        } else {
            double heat = 0, res = 0;
            HS_TRY(energy_balance(c, &heat, &res));
            loc.heat_in = heat;
            loc.energy_residual = std::fabs(res) / std::max(std::fabs(heat), 1e-300);
        }
    }
    if (st) *st = loc;
    return HS_OK;
}

/* MobSolver::solve mobs_solver.h:303-371. */
int hs_solid_solve(hs_ctx* c, hs_stats* st) {
    if (!c) return HS_E_ARG;
    if (!c->inited) return HS_E_STATE;
    int s = 0; double r = 0;
    if (modeR(c) && c->solid_regions > 0) {
        const double n = (double)(c->sb[1] - c->sb[0] + 1) * (c->sb[3] - c->sb[2] + 1) * (c->sb[5] - c->sb[4] + 1);
        HS_TRY(gs_loop(c, 4, n, c->prm.ts_tol, c->prm.ts_sweeps_max, 2, [&] { return solid_sweep(c); },
                       [&] { return HS_OK; }, &s, &r));
    }
    if (st) { st->ts_sweeps = s; st->ts_residual = r; }
    return HS_OK;
}

/* Solver::solve exchange block solver.h:62-148. */
int hs_exchange(hs_ctx* c) {
    if (!c) return HS_E_ARG;
    if (modeR(c)) HS_TRY(exchange_kernel(c));
    HS_CHECK(hipStreamSynchronize(c->stream));
    return HS_OK;
}

/* FluidSolver::setForNext fluid_solver.h:1634-1682 + MobSolver::setForNext. */
int hs_set_for_next(hs_ctx* c) {
    if (!c) return HS_E_ARG;
    if (modeR(c)) {
        HS_TRY(bc_x_ref(c)); HS_TRY(bc_y_ref(c)); HS_TRY(bc_z_ref(c));
        HS_TRY(cell_bc_ref(c));
        HS_TRY(ptemp_from_p_all(c));
    } else {
        for (int comp = 0; comp < 4; ++comp) HS_TRY(bc_cfg_outflow(c, comp));
    }
    HS_CHECK(hipStreamSynchronize(c->stream));
    return HS_OK;
}

int hs_step(hs_ctx* c, hs_stats* st) {
    if (!c) return HS_E_ARG;
    hs_stats loc{};
    c->defer_balance = true;
    const int e = hs_fluid_solve(c, &loc);
    c->defer_balance = false;
    HS_TRY(e);
    HS_TRY(hs_solid_solve(c, &loc));
    HS_TRY(hs_exchange(c));
    HS_TRY(hs_set_for_next(c));   // synchronises: the deferred energy balance is in h_sctl
    if (!modeR(c)) {
        const double res = c->h_sctl->red[1], heat = c->h_sctl->red[2];
        loc.heat_in = heat;
        loc.energy_residual = std::fabs(res) / std::max(std::fabs(heat), 1e-300);
    }
    // the fixed-count sweep loops' deferred control blocks (also read by set_for_next's sync)
    bool nan = false;
    for (int id = 0; id < 5; ++id) {
        if (!c->wdefer[id]) continue;
        c->wdefer[id] = false;
        const SweepCtl& w = c->h_wdefer[id];
        c->gs_last[id] = w.step;
        if (id < 3) { loc.mom_sweeps[id] = w.step; loc.mom_residual[id] = w.global; }
        else if (id == 3) { loc.t_sweeps = w.step; loc.t_residual = w.global; }
        else { loc.ts_sweeps = w.step; loc.ts_residual = w.global; }
        nan = nan || std::isnan(w.global);
    }
    if (nan) return HS_E_NAN;
    if (st) *st = loc;
    return HS_OK;
}

int hs_run(hs_ctx* c, int32_t n, hs_stats* st) {
    if (!c || n < 0) return HS_E_ARG;
    for (int32_t it = 0; it < n; ++it) HS_TRY(hs_step(c, st ? st + it : nullptr));
    return HS_OK;
}

int hs_pressure_apply(hs_ctx* c, const double* x_host, double* y_host) {
    if (!c || !x_host || !y_host) return HS_E_ARG;
    if (!c->have_matrix) return HS_E_STATE;
    const long n = c->d.nP;
    HS_CHECK(hipMemcpyAsync(c->kb, x_host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HS_TRY(mask_fluid(c, c->kb));
    HS_TRY(spmv(c, c->kb, c->kq));
    HS_CHECK(hipMemcpyAsync(y_host, c->kq, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HS_CHECK(hipStreamSynchronize(c->stream));
    return HS_OK;
}

int hs_pressure_solve(hs_ctx* c, const double* b_host, double* x_host, hs_stats* st) {
    if (!c || !b_host || !x_host) return HS_E_ARG;
    if (!c->have_matrix) return HS_E_STATE;
    const long n = c->d.nP;
    HS_CHECK(hipMemcpyAsync(c->kb, b_host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    hs_stats loc{};
    HS_TRY(pressure_solve_staged(c, &loc));
    HS_CHECK(hipMemcpyAsync(x_host, c->kx, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HS_CHECK(hipStreamSynchronize(c->stream));
    if (st) *st = loc;
    return HS_OK;
}

int hs_get_array(hs_ctx* c, int32_t which, double* host, int64_t n) {
    if (!c || !host) return HS_E_ARG;
    const int64_t s = array_len(c, which);
    if (s < 0 || n != s) return HS_E_ARG;
    HS_CHECK(hipMemcpyAsync(host, array_ptr(c, which), s * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HS_CHECK(hipStreamSynchronize(c->stream));
    return HS_OK;
}

int hs_put_array(hs_ctx* c, int32_t which, const double* host, int64_t n) {
    if (!c || !host) return HS_E_ARG;
    const int64_t s = array_len(c, which);
    if (s < 0 || n != s) return HS_E_ARG;
    if (which == HS_A_PPRIME) {   // the written values are p' from now on
        c->pprime_in_kx = false;
        c->pprime_dirty = true;
    }
    HS_CHECK(hipMemcpyAsync(array_ptr(c, which), host, s * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (which == HS_A_Q) ++c->q_version;
    HS_CHECK(hipStreamSynchronize(c->stream));
    return HS_OK;
}

int hs_get_fields(hs_ctx* c, double* u, double* v, double* w, double* p, double* t) {
    if (!c) return HS_E_ARG;
    const Dims& d = c->d;
    std::vector<double> U(d.nU), V(d.nV), W(d.nW), P(d.nC), TF(d.nP), TS(d.nP);
    HS_TRY(hs_get_array(c, HS_A_U, U.data(), d.nU));
    HS_TRY(hs_get_array(c, HS_A_V, V.data(), d.nV));
    HS_TRY(hs_get_array(c, HS_A_W, W.data(), d.nW));
    HS_TRY(hs_get_array(c, HS_A_P, P.data(), d.nC));
    HS_TRY(hs_get_array(c, HS_A_TF, TF.data(), d.nP));
    HS_TRY(hs_get_array(c, HS_A_TS, TS.data(), d.nP));
    const long nx = d.nx, ny = d.ny, nz = d.nz;
    for (long k = 0; k < nz; ++k)
        for (long j = 0; j < ny; ++j) {
            for (long i = 0; i <= nx; ++i) if (u) u[i + (nx + 1) * (j + ny * k)] = U[idxU(d, i + 1, j + 1, k + 1)];
            for (long i = 0; i < nx; ++i) {
                const long cc = idxC(d, i, j, k);
                if (p) p[cc] = P[cc];
                if (t) {
                    const long c2 = idxP(d, i + 1, j + 1, k + 1);
                    t[cc] = (modeR(c) && c->type[cc] == 1) ? TS[c2] : TF[c2];
                }
            }
        }
    for (long k = 0; k < nz; ++k)
        for (long j = 0; j <= ny; ++j)
            for (long i = 0; i < nx; ++i) if (v) v[i + nx * (j + (ny + 1) * k)] = V[idxV(d, i + 1, j + 1, k + 1)];
    for (long k = 0; k <= nz; ++k)
        for (long j = 0; j < ny; ++j)
            for (long i = 0; i < nx; ++i) if (w) w[i + nx * (j + ny * k)] = W[idxW(d, i + 1, j + 1, k + 1)];
    return HS_OK;
}

/* Output packing of Solver::solve solver.h:160-213 into OutputData records
 * (fluid_solver.h:38-45), one per cell in the order of main.cpp's resultMap (global cell
 * index, x fastest; main.cpp:95-101).  Fluid cells: velocity on the cell's west / south /
 * bottom faces (solver.h:172-174), p, the fluid output temperature with the domain-boundary
 * cells at 293 (applyCellBoundaryCondition fluid_solver.h:394-404, Mode R), and the p'
 * row number as id (fluid_solver.h:1966, :2120; 0 until a p' matrix exists, the value of
 * the zero-filled id vector, fluid_solver.h:97).  Solid cells: the solid temperature and
 * the OutputData defaults elsewhere (velocity and p 0, id -1; solver.h:186-197). */
int hs_get_output(hs_ctx* c, double* loc, double* vel, double* p, double* t, int32_t* id) {
    if (!c) return HS_E_ARG;
    const Dims& d = c->d;
    const long nx = d.nx, ny = d.ny, nz = d.nz;
    std::vector<double> u((nx + 1) * ny * nz), v(nx * (ny + 1) * nz), w(nx * ny * (nz + 1)), pc(d.nC), tc(d.nC);
    HS_TRY(hs_get_fields(c, u.data(), v.data(), w.data(), pc.data(), tc.data()));
    const double kBoundaryT = 293.0;   // fluid_solver.h:402
    int32_t row = 0;
    for (long k = 0; k < nz; ++k)
        for (long j = 0; j < ny; ++j)
            for (long i = 0; i < nx; ++i) {
                const long q = idxC(d, i, j, k);
                const bool fluid = c->type[q] == 0;
                if (loc) { loc[3 * q] = c->X[i]; loc[3 * q + 1] = c->Y[j]; loc[3 * q + 2] = c->Z[k]; }
                if (vel) {
                    vel[3 * q] = fluid ? u[i + (nx + 1) * (j + ny * k)] : 0.0;
                    vel[3 * q + 1] = fluid ? v[i + nx * (j + (ny + 1) * k)] : 0.0;
                    vel[3 * q + 2] = fluid ? w[i + nx * (j + ny * k)] : 0.0;
                }
                if (p) p[q] = fluid ? pc[q] : 0.0;
                if (t) {
                    const bool edge = i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1;
                    t[q] = (fluid && edge && modeR(c)) ? kBoundaryT : tc[q];
                }
                if (id) id[q] = fluid ? (c->have_matrix ? row : 0) : -1;
                if (fluid) ++row;
            }
    return HS_OK;
}

int hs_put_fields(hs_ctx* c, const double* u, const double* v, const double* w, const double* p, const double* t) {
    if (!c) return HS_E_ARG;
    const Dims& d = c->d;
    std::vector<double> U(d.nU), V(d.nV), W(d.nW), P(d.nC), TF(d.nP), TS(d.nP), PT(d.nP);
    HS_TRY(hs_get_array(c, HS_A_U, U.data(), d.nU));
    HS_TRY(hs_get_array(c, HS_A_V, V.data(), d.nV));
    HS_TRY(hs_get_array(c, HS_A_W, W.data(), d.nW));
    HS_TRY(hs_get_array(c, HS_A_P, P.data(), d.nC));
    HS_TRY(hs_get_array(c, HS_A_TF, TF.data(), d.nP));
    HS_TRY(hs_get_array(c, HS_A_TS, TS.data(), d.nP));
    HS_TRY(hs_get_array(c, HS_A_PTEMP, PT.data(), d.nP));
    const long nx = d.nx, ny = d.ny, nz = d.nz;
    for (long k = 0; k < nz; ++k)
        for (long j = 0; j < ny; ++j) {
            for (long i = 0; i <= nx; ++i) if (u) U[idxU(d, i + 1, j + 1, k + 1)] = u[i + (nx + 1) * (j + ny * k)];
            for (long i = 0; i < nx; ++i) {
                const long cc = idxC(d, i, j, k), c2 = idxP(d, i + 1, j + 1, k + 1);
                if (p) { P[cc] = p[cc]; PT[c2] = p[cc]; }
                if (t) { if (modeR(c) && c->type[cc] == 1) TS[c2] = t[cc]; else TF[c2] = t[cc]; }
            }
        }
    for (long k = 0; k < nz; ++k)
        for (long j = 0; j <= ny; ++j)
            for (long i = 0; i < nx; ++i) if (v) V[idxV(d, i + 1, j + 1, k + 1)] = v[i + nx * (j + (ny + 1) * k)];
    for (long k = 0; k <= nz; ++k)
        for (long j = 0; j < ny; ++j)
            for (long i = 0; i < nx; ++i) if (w) W[idxW(d, i + 1, j + 1, k + 1)] = w[i + nx * (j + ny * k)];
    HS_TRY(hs_put_array(c, HS_A_U, U.data(), d.nU));
    HS_TRY(hs_put_array(c, HS_A_V, V.data(), d.nV));
    HS_TRY(hs_put_array(c, HS_A_W, W.data(), d.nW));
    HS_TRY(hs_put_array(c, HS_A_P, P.data(), d.nC));
    HS_TRY(hs_put_array(c, HS_A_PTEMP, PT.data(), d.nP));
    HS_TRY(hs_put_array(c, HS_A_TF, TF.data(), d.nP));
    HS_TRY(hs_put_array(c, HS_A_TS, TS.data(), d.nP));
    return HS_OK;
}

int hs_set_source(hs_ctx* c, const double* q) {
    if (!c || !q) return HS_E_ARG;
    return hs_put_array(c, HS_A_Q, q, c->d.nC);
}

int hs_set_conductivity(hs_ctx* c, const double* k) {
    if (!c || !k) return HS_E_ARG;
    return hs_put_array(c, HS_A_K, k, c->d.nC);
}

/* ---- energy equation on a frozen flow and its adjoint (no reference counterpart) ---- */
namespace {
bool mem_ok(int32_t mem) { return mem == HS_MEM_HOST || mem == HS_MEM_DEVICE; }
// device copy of a caller vector of nC doubles: the pointer itself when it is device
// memory, else the staging buffer `stage` after an upload
int stage_in(hs_ctx* c, const double* src, int32_t mem, double* stage, const double** dev) {
    if (mem == HS_MEM_DEVICE) { *dev = src; return HS_OK; }
    HS_CHECK(hipMemcpyAsync(stage, src, c->d.nC * sizeof(double), hipMemcpyHostToDevice, c->stream));
    *dev = stage;
    return HS_OK;
}
int stage_out(hs_ctx* c, const double* dev, double* dst, int32_t mem) {
    if (mem == HS_MEM_DEVICE || !dst) return HS_OK;
    HS_CHECK(hipMemcpyAsync(dst, dev, c->d.nC * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return HS_OK;
}
}  // namespace

int hs_energy_solve(hs_ctx* c, double tol, int32_t maxit, hs_solve_info* info) {
    if (!c) return HS_E_ARG;
    if (!c->inited) return HS_E_STATE;
    return energy_solve(c, tol, maxit, info);
}

int hs_adjoint_energy(hs_ctx* c, const double* dJdT, double* dJdq, double* dJdk, double tol, int32_t maxit,
                      int32_t mem, hs_solve_info* info) {
    if (!c || !dJdT || !mem_ok(mem)) return HS_E_ARG;
    if (!c->inited) return HS_E_STATE;
    double *s1, *s2, *s3;
    HS_TRY(energy_stage(c, &s1, &s2, &s3));
    const double* g;
    HS_TRY(stage_in(c, dJdT, mem, s1, &g));
    double* oq = (mem == HS_MEM_DEVICE) ? dJdq : (dJdq ? s2 : nullptr);
    double* ok = (mem == HS_MEM_DEVICE) ? dJdk : (dJdk ? s3 : nullptr);
    HS_TRY(energy_adjoint(c, g, oq, ok, tol, maxit, info));
    if (dJdq) HS_TRY(stage_out(c, oq, dJdq, mem));
    if (dJdk) HS_TRY(stage_out(c, ok, dJdk, mem));
    HS_CHECK(hipStreamSynchronize(c->stream));
    return HS_OK;
}

int hs_energy_solve_rhs(hs_ctx* c, int32_t transpose, const double* b, double* x, double tol, int32_t maxit,
                        int32_t mem, hs_solve_info* info) {
    if (!c || !b || !x || !mem_ok(mem)) return HS_E_ARG;
    if (!c->inited) return HS_E_STATE;
    double *s1, *s2;
    HS_TRY(energy_stage(c, &s1, &s2, nullptr));
    const double* bd;
    HS_TRY(stage_in(c, b, mem, s1, &bd));
    double* xd = (mem == HS_MEM_DEVICE) ? x : s2;
    HS_TRY(energy_solve_rhs(c, transpose, bd, xd, tol, maxit, info));
    HS_TRY(stage_out(c, xd, x, mem));
    HS_CHECK(hipStreamSynchronize(c->stream));
    return HS_OK;
}

int hs_energy_apply(hs_ctx* c, int32_t transpose, const double* x, double* y, int32_t mem) {
    if (!c || !x || !y || !mem_ok(mem)) return HS_E_ARG;
    if (!c->inited) return HS_E_STATE;
    double *s1, *s2;
    HS_TRY(energy_stage(c, &s1, &s2, nullptr));
    const double* xd;
    HS_TRY(stage_in(c, x, mem, s1, &xd));
    double* yd = (mem == HS_MEM_DEVICE) ? y : s2;
    HS_TRY(energy_apply(c, transpose, xd, yd));
    HS_TRY(stage_out(c, yd, y, mem));
    HS_CHECK(hipStreamSynchronize(c->stream));
    return HS_OK;
}

int hs_get_cells(hs_ctx* c, int32_t which, double* dst, int32_t mem) {
    if (!c || !dst || !mem_ok(mem)) return HS_E_ARG;
    const hipMemcpyKind kind = mem == HS_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    if (which == HS_A_Q || which == HS_A_K) {
        HS_CHECK(hipMemcpyAsync(dst, which == HS_A_Q ? c->q : c->k, c->d.nC * sizeof(double), kind, c->stream));
    } else if (which == HS_A_TF) {
        double* s1;
        HS_TRY(energy_stage(c, &s1, nullptr, nullptr));
        double* out = mem == HS_MEM_DEVICE ? dst : s1;
        HS_TRY(cells_from_padded(c, c->tf, out));
        HS_TRY(stage_out(c, out, dst, mem));
    } else {
        return HS_E_ARG;
    }
    HS_CHECK(hipStreamSynchronize(c->stream));
    return HS_OK;
}

int hs_put_cells(hs_ctx* c, int32_t which, const double* src, int32_t mem) {
    if (!c || !src || !mem_ok(mem)) return HS_E_ARG;
    const hipMemcpyKind kind = mem == HS_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (which == HS_A_Q || which == HS_A_K) {
        HS_CHECK(hipMemcpyAsync(which == HS_A_Q ? c->q : c->k, src, c->d.nC * sizeof(double), kind, c->stream));
        if (which == HS_A_Q) ++c->q_version;
    } else if (which == HS_A_TF) {
        double* s1;
        HS_TRY(energy_stage(c, &s1, nullptr, nullptr));
        const double* in;
        HS_TRY(stage_in(c, src, mem, s1, &in));
        HS_TRY(cells_to_padded(c, in, c->tf));
    } else {
        return HS_E_ARG;
    }
    HS_CHECK(hipStreamSynchronize(c->stream));
    return HS_OK;
}

/* applyCellBoundaryCondition fluid_solver.h:324-405 on its own (tests / host drivers). */
int hs_cell_bc(hs_ctx* c) {
    if (!c) return HS_E_ARG;
    if (!modeR(c)) return HS_E_STATE;
    HS_TRY(cell_bc_ref(c));
    HS_CHECK(hipStreamSynchronize(c->stream));
    return HS_OK;
}

}  // extern "C"

extern "C" {

/* Live kernel timing (no counterpart in the reference, whose only timer is clock()
 * around outer iterations, solver.h:53-54,156-158): kernel = HS_K_* starts recording
 * HIP events on the context's stream around every launch of that kernel; 0 stops. */
int hs_prof_set(hs_ctx* c, int32_t kernel) {
    if (!c) return HS_E_ARG;
    HS_CHECK(hipStreamSynchronize(c->stream));
    if (c->prof_ev.empty() && kernel != 0) {
        c->prof_ev.resize(2048);
        for (auto& e : c->prof_ev) HS_CHECK(hipEventCreate(&e));
    }
    c->prof_kernel = kernel;
    c->prof_used = 0;
    c->prof_count = 0;
    c->prof_ms = 0.0;
    return HS_OK;
}

/* number of timed launches and their summed duration (ms) since hs_prof_set */
int hs_prof_get(hs_ctx* c, int64_t* count, double* total_ms) {
    if (!c) return HS_E_ARG;
    HS_CHECK(hipStreamSynchronize(c->stream));
    for (int e = 0; e < c->prof_used; e += 2) {
        float ms = 0.f;
        HS_CHECK(hipEventElapsedTime(&ms, c->prof_ev[e], c->prof_ev[e + 1]));
        c->prof_ms += ms;
        c->prof_count += 1;
    }
    c->prof_used = 0;
    if (count) *count = c->prof_count;
    if (total_ms) *total_ms = c->prof_ms;
    return HS_OK;
}

/* active cells of each multigrid level (for byte counts); returns the level count */
int hs_mg_levels(hs_ctx* c, int32_t* dims, int32_t max_levels) {
    if (!c || max_levels < 0 || (max_levels > 0 && !dims)) return HS_E_ARG;
    int n = 0;
    for (MGLevel* m : c->mg) {
        if (n >= max_levels) break;
        dims[3 * n] = m->nx; dims[3 * n + 1] = m->ny; dims[3 * n + 2] = m->nz;
        ++n;
    }
    return n;
}

int hs_energy_balance(hs_ctx* c, double* heat_in, double* residual) {
    if (!c || !heat_in || !residual) return HS_E_ARG;
    if (c->prm.mode != HS_MODE_CONFIG) return HS_E_ARG;
    if (!c->inited) return HS_E_STATE;
    double heat = 0.0, res = 0.0;
    HS_TRY(energy_balance(c, &heat, &res));
    *heat_in = heat;
    *residual = std::fabs(res) / std::max(std::fabs(heat), 1e-300);
    return HS_OK;
}

int hs_amg_levels(hs_ctx* c, int32_t which, int64_t* rows_nnz, int32_t max_levels) {
    if (!c || (which != 0 && which != 1) || max_levels < 0 || (max_levels > 0 && !rows_nnz)) return HS_E_ARG;
    const AmgHier* H = which == 0 ? c->amgp : (c->en ? c->en->amg : nullptr);
    if (!H) return 0;
    int n = 0;
    for (const AmgLevel* L : H->lv) {
        if (n >= max_levels) break;
        rows_nnz[2 * n] = L->n;
        rows_nnz[2 * n + 1] = L->nnz;
        ++n;
    }
    return n;
}

}  // extern "C"

extern "C" {

/* Plain device copy of nbytes, reps times, on a private stream: the known-size
 * reference that hardware byte counters (FETCH_SIZE / WRITE_SIZE) are scaled by.
 * Returns the average copy time in ms. */
int hs_copy_bench(int64_t nbytes, int32_t reps, double* avg_ms) {
    if (nbytes <= 0 || reps <= 0) return HS_E_ARG;
    hs_ctx tmp;
    HS_CHECK(hipStreamCreateWithFlags(&tmp.stream, hipStreamNonBlocking));
    double *a = nullptr, *b = nullptr;
    const long n = (long)(nbytes / 8);
    if (hipMalloc((void**)&a, n * 8) != hipSuccess || hipMalloc((void**)&b, n * 8) != hipSuccess) {
        if (a) (void)hipFree(a);
        (void)hipStreamDestroy(tmp.stream);
        return HS_E_NOMEM;
    }
    int e = launch_fill(&tmp, a, n, 1.0);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, tmp.stream);
    for (int r = 0; r < reps && e == HS_OK; ++r) e = launch_copy(&tmp, b, a, n);
    (void)hipEventRecord(e1, tmp.stream);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (avg_ms) *avg_ms = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipStreamDestroy(tmp.stream);
    tmp.stream = nullptr;
    return e;
}

}  // extern "C"
