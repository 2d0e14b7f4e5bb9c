// hs_solver.hpp — C++ host shim with the reference's solver shapes over the C ABI.
//
// The reference's interface is its header-only structs (fluid_solver.h:52-182,
// mobs_solver.h:31-77, solver.h:24-45); a host that drove them keeps the same calls:
//   FluidSolver::init / solve / setForNext,  MobSolver::init / solve / setForNext,
//   Solver::addFluids / addMobs / solve   (call order of solver.h:44-155),
// and reads the same output members (result / idArray / idToData, solver.h:160-213).
// On one rank the fluid region and the solid region live in one hs_ctx (one GPU), so
// both solver objects forward to the same context; the exchange block of
// Solver::solve (solver.h:62-148) is one hs_exchange call.
#pragma once

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hsolver.h"
#include "../../include/hsolver_mesh.h"

namespace hsolver {

inline void check(int rc, const char* what) {
    if (rc != HS_OK) throw std::runtime_error(std::string(what) + ": " + hs_strerror(rc));
}

// fluid_solver.h:38-45: one output record per cell (defaults: solid cells keep them)
struct OutputData {
    double ux = 0, uy = 0, uz = 0, p = 0;
    double t = 0;
    double locationx = 0.0, locationy = 0.0, locationz = 0.0;
    int globalid = -1;
};

// Owns the device state of one grid (RAII around hs_create / hs_destroy).
class Context {
public:
    Context(const hs_grid& g, const hs_params& p) : nx(g.nx), ny(g.ny), nz(g.nz) {
        check(hs_create(&g, &p, &c_), "hs_create");
    }
    ~Context() { hs_destroy(c_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    hs_ctx* get() const { return c_; }
    bool inited = false;
    const int nx, ny, nz;

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
    // :1634 t <- t_temp, the velocity / cell boundary conditions, p_temp <- p.  One call
    // per call of the reference's method, wherever the host's loop makes it.
    void setForNext() { check(hs_set_for_next(ctx.get()), "hs_set_for_next"); }
    Context& ctx;
    hs_stats stats{};
    // solver.h:160-213 output of this solver's cells (fluid cells), keyed by the global
    // 1-based cell index (distributed_mesh.h:320-330)
    std::vector<OutputData> result;
    std::vector<int64_t> idArray;
    std::map<int, OutputData> idToData;
};

// mobs_solver.h:31 MobSolver
struct MobSolver {
    explicit MobSolver(Context& ctx) : ctx(ctx) {}
    void init() {                       // mobs_solver.h:116
        if (!ctx.inited) { check(hs_init(ctx.get()), "hs_init"); ctx.inited = true; }
    }
    void solve() { check(hs_solid_solve(ctx.get(), &stats), "hs_solid_solve"); }   // :303
    // :290 t <- t_temp on the solid cells.  The solid temperature is one device array
    // here, read directly by hs_get_output / hs_get_fields, so there is no copy to make.
    void setForNext() {}
    Context& ctx;
    hs_stats stats{};
    std::vector<OutputData> result;     // solid cells (solver.h:186-212)
    std::vector<int64_t> idArray;
    std::map<int, OutputData> idToData;
};

// solver.h:24 Solver
struct Solver {
    std::vector<FluidSolver*> fluids;
    std::vector<MobSolver*> mobs;
    int id = 0;
    std::vector<hs_stats> history;
    double seconds_after_warmup = 0.0;   // wall time of iterations warmup..n-1

    explicit Solver(int id_ = 0) : id(id_) {}
    void addFluids(FluidSolver* f) { fluids.push_back(f); }
    void addMobs(MobSolver* m) { mobs.push_back(m); }

    // solver.h:44-155: init all, then per outer iteration fluid solve, solid solve,
    // interface exchange, fluid setForNext, solid setForNext; then the output packing
    // (solver.h:160-213).  The reference runs 100 iterations and times 10..99 with
    // clock(); this keeps wall time over iterations [warmup, n).
    void solve(int n = 100, int warmup = 10) {
        if (n < 1) throw std::invalid_argument("Solver::solve: iteration count must be at least 1");
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
            hs_stats st = fluids.empty() ? hs_stats{} : fluids[0]->stats;
            if (!mobs.empty()) { st.ts_sweeps = mobs[0]->stats.ts_sweeps; st.ts_residual = mobs[0]->stats.ts_residual; }
            history.push_back(st);
        }
        const auto t1 = std::chrono::steady_clock::now();
        if (n > warmup) seconds_after_warmup = std::chrono::duration<double>(t1 - t0).count();
        getResult();
    }

    // solver.h:160-213: every fluid solver's result holds its fluid cells, every solid
    // solver's its solid cells, each ordered by global cell index.
    void getResult() {
        Context* ctx = !fluids.empty() ? &fluids[0]->ctx : (!mobs.empty() ? &mobs[0]->ctx : nullptr);
        if (!ctx) return;
        const long n = (long)ctx->nx * ctx->ny * ctx->nz;
        std::vector<double> loc(3 * n), vel(3 * n), p(n), t(n);
        std::vector<int32_t> ids(n);
        check(hs_get_output(ctx->get(), loc.data(), vel.data(), p.data(), t.data(), ids.data()), "hs_get_output");
        auto fill = [&](std::map<int, OutputData>& m, std::vector<OutputData>& res, std::vector<int64_t>& idx,
                        bool fluid) {
            m.clear();
            for (long q = 0; q < n; ++q) {
                if ((ids[q] >= 0) != fluid) continue;
                OutputData o;
                if (fluid) {
                    o.ux = vel[3 * q]; o.uy = vel[3 * q + 1]; o.uz = vel[3 * q + 2];
                    o.p = p[q];
                    o.globalid = ids[q];
                }
                o.t = t[q];
                o.locationx = loc[3 * q]; o.locationy = loc[3 * q + 1]; o.locationz = loc[3 * q + 2];
                m[(int)(q + 1)] = o;
            }
            res.clear();
            idx.clear();
            for (auto& kv : m) { idx.push_back(kv.first); res.push_back(kv.second); }
        };
        for (auto* f : fluids) fill(f->idToData, f->result, f->idArray, true);
        for (auto* m : mobs) fill(m->idToData, m->result, m->idArray, false);
    }
};

// The VTK writer of main.cpp:54-156 on one rank: the solvers' results merged by global
// cell index (main.cpp:95-101) and written in the reference's layout and number format
// (std::fixed, 12 decimals; id as an integer).  DIMENSIONS comes from the grid (the
// reference prints a fixed 50 50 50, main.cpp:109).
inline void write_vtk(const char* path, const Solver& s, int nx, int ny, int nz) {
    std::map<int, OutputData> all;
    for (auto* f : s.fluids) for (size_t i = 0; i < f->idArray.size(); ++i) all[(int)f->idArray[i]] = f->result[i];
    for (auto* m : s.mobs) for (size_t i = 0; i < m->idArray.size(); ++i) all[(int)m->idArray[i]] = m->result[i];
    FILE* f = std::fopen(path, "w");
    if (!f) throw std::runtime_error(std::string("cannot write ") + path);
    std::fprintf(f, "# vtk DataFile Version 3.0\nMPI gathered data with scalar p\nASCII\nDATASET STRUCTURED_GRID\n");
    std::fprintf(f, "DIMENSIONS %d %d %d\nPOINTS %zu double\n", nx, ny, nz, all.size());
    for (auto& e : all) std::fprintf(f, "%.12f %.12f %.12f\n", e.second.locationx, e.second.locationy, e.second.locationz);
    std::fprintf(f, "POINT_DATA %zu\nSCALARS pressure double 1\nLOOKUP_TABLE default\n", all.size());
    for (auto& e : all) std::fprintf(f, "%.12f\n", e.second.p);
    std::fprintf(f, "VECTORS velocity double\n");
    for (auto& e : all) std::fprintf(f, "%.12f %.12f %.12f\n", e.second.ux, e.second.uy, e.second.uz);
    std::fprintf(f, "SCALARS temperature double 1\nLOOKUP_TABLE default\n");
    for (auto& e : all) std::fprintf(f, "%.12f\n", e.second.t);
    std::fprintf(f, "SCALARS id int 1\nLOOKUP_TABLE default\n");
    for (auto& e : all) std::fprintf(f, "%d\n", e.second.globalid);
    std::fclose(f);
}

// Raw fields at round-trip precision (%.17g): every cell's p, the fluid / solid
// temperature arrays as solved (no boundary override), the west / south / bottom face
// velocities of every cell, and the linear cell index of fluid cells as id.  Not the
// reference's format; for checking fields bit for bit.
inline void write_vtk_raw(const char* path, hs_ctx* c, const hs_grid& g) {
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
