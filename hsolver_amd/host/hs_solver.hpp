// hs_solver.hpp — C++ host shim with the reference's solver shapes over the C ABI.
//
// The reference's interface is its header-only structs (fluid_solver.h:52-182,
// mobs_solver.h:31-77, solver.h:24-45); a host that drove them keeps the same calls:
//   FluidSolver::init / solve / setForNext,  MobSolver::init / solve / setForNext,
//   Solver::addFluids / addMobs / solve   (call order of solver.h:44-155).
// On one rank the fluid region and the solid region live in one hs_ctx (one GPU), so
// both solver objects forward to the same context; the exchange block of
// Solver::solve (solver.h:62-148) is one hs_exchange call.
#pragma once

#include <chrono>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hsolver.h"
#include "../../include/hsolver_mesh.h"

namespace hsolver {

inline void check(int rc, const char* what) {
    if (rc != HS_OK) throw std::runtime_error(std::string(what) + ": " + hs_strerror(rc));
}

// Owns the device state of one grid (RAII around hs_create / hs_destroy).
class Context {
public:
    Context(const hs_grid& g, const hs_params& p) { check(hs_create(&g, &p, &c_), "hs_create"); }
    ~Context() { hs_destroy(c_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    hs_ctx* get() const { return c_; }
    bool inited = false;

private:
    hs_ctx* c_ = nullptr;
};

// fluid_solver.h:52 FluidSolver
struct FluidSolver {
    explicit FluidSolver(Context& ctx) : ctx(ctx) {}
    void init() {                       // fluid_solver.h:882
        if (!ctx.inited) { check(hs_init(ctx.get()), "hs_init"); ctx.inited = true; }
    }
    void solve() { check(hs_fluid_solve(ctx.get(), &stats), "hs_fluid_solve"); }   // :1401
    void setForNext() { pending_next = true; }                                      // :1634
    Context& ctx;
    hs_stats stats{};
    bool pending_next = false;
};

// mobs_solver.h:31 MobSolver
struct MobSolver {
    explicit MobSolver(Context& ctx) : ctx(ctx) {}
    void init() {                       // mobs_solver.h:116
        if (!ctx.inited) { check(hs_init(ctx.get()), "hs_init"); ctx.inited = true; }
    }
    void solve() { check(hs_solid_solve(ctx.get(), &stats), "hs_solid_solve"); }   // :303
    void setForNext() { pending_next = true; }                                      // :290
    Context& ctx;
    hs_stats stats{};
    bool pending_next = false;
};

// solver.h:24 Solver
struct Solver {
    std::vector<FluidSolver*> fluids;
    std::vector<MobSolver*> mobs;
    std::vector<hs_stats> history;
    double seconds_after_warmup = 0.0;   // wall time of iterations warmup..n-1

    void addFluids(FluidSolver* f) { fluids.push_back(f); }
    void addMobs(MobSolver* m) { mobs.push_back(m); }

    // solver.h:44-155: init all, then per outer iteration fluid solve, solid solve,
    // interface exchange, setForNext.  The reference times iterations 10..99 with
    // clock(); this keeps wall time over iterations [warmup, n).
    void solve(int n = 100, int warmup = 10) {
        for (auto* f : fluids) f->init();
        for (auto* m : mobs) m->init();
        auto t0 = std::chrono::steady_clock::now();
        for (int step = 0; step < n; ++step) {
            if (step == warmup) t0 = std::chrono::steady_clock::now();
            for (auto* f : fluids) f->solve();
            for (auto* m : mobs) m->solve();
            Context* any = !fluids.empty() ? &fluids[0]->ctx : (!mobs.empty() ? &mobs[0]->ctx : nullptr);
            if (any) check(hs_exchange(any->get()), "hs_exchange");
            for (auto* f : fluids) f->setForNext();
            for (auto* m : mobs) m->setForNext();
            if (any) check(hs_set_for_next(any->get()), "hs_set_for_next");
            for (auto* f : fluids) f->pending_next = false;
            for (auto* m : mobs) m->pending_next = false;
            hs_stats st = fluids.empty() ? hs_stats{} : fluids[0]->stats;
            if (!mobs.empty()) { st.ts_sweeps = mobs[0]->stats.ts_sweeps; st.ts_residual = mobs[0]->stats.ts_residual; }
            history.push_back(st);
        }
        const auto t1 = std::chrono::steady_clock::now();
        if (n > warmup) seconds_after_warmup = std::chrono::duration<double>(t1 - t0).count();
    }
};

// Output gather + VTK writer of main.cpp:101-156 on one rank: one point per cell at
// the cell's minimum vertex (solver.h:178-180), velocity sampled on the cell's
// west/south/bottom faces (solver.h:172-174), with DIMENSIONS set from the grid and
// data printed at round-trip precision (%.17g).
inline void write_vtk(const char* path, hs_ctx* c, const hs_grid& g) {
    const long nx = g.nx, ny = g.ny, nz = g.nz, n = nx * ny * nz;
    std::vector<double> u((nx + 1) * ny * nz), v(nx * (ny + 1) * nz), w(nx * ny * (nz + 1)), p(n), t(n);
    check(hs_get_fields(c, u.data(), v.data(), w.data(), p.data(), t.data()), "hs_get_fields");
    FILE* f = std::fopen(path, "w");
    if (!f) throw std::runtime_error(std::string("cannot write ") + path);
    std::fprintf(f, "# vtk DataFile Version 3.0\nSIMPLE conjugate heat transfer\nASCII\nDATASET STRUCTURED_GRID\n");
    std::fprintf(f, "DIMENSIONS %ld %ld %ld\nPOINTS %ld double\n", nx, ny, nz, n);
    for (long k = 0; k < nz; ++k)
        for (long j = 0; j < ny; ++j)
            for (long i = 0; i < nx; ++i) std::fprintf(f, "%.12f %.12f %.12f\n", g.xv[i], g.yv[j], g.zv[k]);
    std::fprintf(f, "POINT_DATA %ld\nSCALARS pressure double 1\nLOOKUP_TABLE default\n", n);
    for (long q = 0; q < n; ++q) std::fprintf(f, "%.17g\n", p[q]);
    std::fprintf(f, "VECTORS velocity double\n");
    for (long k = 0; k < nz; ++k)
        for (long j = 0; j < ny; ++j)
            for (long i = 0; i < nx; ++i)
                std::fprintf(f, "%.17g %.17g %.17g\n", u[i + (nx + 1) * (j + ny * k)], v[i + nx * (j + (ny + 1) * k)],
                             w[i + nx * (j + ny * k)]);
    std::fprintf(f, "SCALARS temperature double 1\nLOOKUP_TABLE default\n");
    for (long q = 0; q < n; ++q) std::fprintf(f, "%.17g\n", t[q]);
    std::fprintf(f, "SCALARS id int 1\nLOOKUP_TABLE default\n");
    for (long q = 0; q < n; ++q) std::fprintf(f, "%ld\n", g.cell_type[q] == 0 ? q : -1L);
    std::fclose(f);
}

}  // namespace hsolver
