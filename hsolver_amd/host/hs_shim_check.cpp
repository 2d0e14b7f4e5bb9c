// hs_shim_check.cpp — checks the C++ shim (hs_solver.hpp) against the C ABI on one mesh.
//
//   hsolver_shim_check <file.in> [iterations=5] [--order lex|rb]
//
// Three contexts on the same grid and parameters (the reference's physics, HS_MODE_REF):
//   A  a host loop written out in the order of the reference's Solver::solve
//      (solver.h:44-155) over FluidSolver / MobSolver objects: every fluid solve, every
//      solid solve, the exchange, every fluid setForNext, every solid setForNext;
//   B  hs_init + hs_run(n);
//   C  hsolver::Solver::solve(n), then its output members (result / idArray / idToData,
//      solver.h:160-213).
// Every device array of A and C must equal B's bit for bit, and C's output records must
// cover every cell once (fluid records in the fluid solver, solid ones in the solid
// solver) with the values of hs_get_output.  Exit 0 when all hold, 3 on a mismatch,
// 1 on usage / input errors, 2 on solver errors.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "hs_solver.hpp"

namespace {

std::vector<double> array(hs_ctx* c, int which) {
    std::vector<double> a((size_t)hs_array_size(c, which));
    hsolver::check(hs_get_array(c, which, a.data(), (int64_t)a.size()), "hs_get_array");
    return a;
}

int compare(hs_ctx* a, hs_ctx* b, const char* what) {
    int bad = 0;
    for (int w = 0; w < HS_A_COUNT; ++w) {
        const std::vector<double> x = array(a, w), y = array(b, w);
        if (x.size() != y.size() || std::memcmp(x.data(), y.data(), x.size() * sizeof(double)) != 0) {
            std::fprintf(stderr, "mismatch: %s, array %d\n", what, w);
            ++bad;
        }
    }
    return bad;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <file.in> [iterations>=1] [--order lex|rb]\n", argv[0]);
        return 1;
    }
    int iters = 5, order = HS_ORDER_LEX;
    for (int a = 2; a < argc; ++a) {
        if (!std::strcmp(argv[a], "--order") && a + 1 < argc) {
            order = !std::strcmp(argv[++a], "rb") ? HS_ORDER_RB : HS_ORDER_LEX;
        } else {
            char* end = nullptr;
            const long n = std::strtol(argv[a], &end, 10);
            if (end == argv[a] || *end != '\0' || n < 1 || n > 100000) {
                std::fprintf(stderr, "bad iteration count %s\n", argv[a]);
                return 1;
            }
            iters = (int)n;
        }
    }
    hs_mesh* mesh = nullptr;
    if (hs_mesh_read(argv[1], &mesh) != HS_OK) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 1;
    }
    int32_t nx, ny, nz, nmobs, nwalls, nfr, nsr;
    hs_mesh_info(mesh, &nx, &ny, &nz, &nmobs, &nwalls, &nfr, &nsr);
    hs_grid g;
    hs_mesh_grid(mesh, &g);
    hs_params p;
    hs_params_default(HS_MODE_REF, &p);
    p.order = order;
    int bad = 0;
    try {
        hsolver::Context A(g, p), B(g, p), C(g, p);
        {
            hsolver::FluidSolver fluid(A);
            hsolver::MobSolver mob(A);
            std::vector<hsolver::FluidSolver*> fluids{&fluid};
            std::vector<hsolver::MobSolver*> mobs;
            if (nsr > 0) mobs.push_back(&mob);
            for (auto* f : fluids) f->init();
            for (auto* m : mobs) m->init();
            for (int step = 0; step < iters; ++step) {
                for (auto* f : fluids) f->solve();
                for (auto* m : mobs) m->solve();
                hsolver::check(hs_exchange(A.get()), "hs_exchange");
                for (auto* f : fluids) f->setForNext();
                for (auto* m : mobs) m->setForNext();
            }
        }
        hsolver::check(hs_init(B.get()), "hs_init");
        hsolver::check(hs_run(B.get(), iters, nullptr), "hs_run");
        hsolver::FluidSolver fluid(C);
        hsolver::MobSolver mob(C);
        hsolver::Solver solver;
        solver.addFluids(&fluid);
        if (nsr > 0) solver.addMobs(&mob);
        solver.solve(iters, 0);
        bad += compare(A.get(), B.get(), "reference-order struct loop vs hs_run");
        bad += compare(C.get(), B.get(), "Solver::solve vs hs_run");
        // output members against hs_get_output
        const long n = (long)nx * ny * nz;
        std::vector<double> loc(3 * n), vel(3 * n), pp(n), tt(n);
        std::vector<int32_t> ids(n);
        hsolver::check(hs_get_output(B.get(), loc.data(), vel.data(), pp.data(), tt.data(), ids.data()), "hs_get_output");
        long nfluid = 0;
        for (long q = 0; q < n; ++q) nfluid += ids[q] >= 0;
        const size_t nsolid_out = nsr > 0 ? mob.result.size() : 0;
        if (fluid.result.size() != (size_t)nfluid || fluid.idArray.size() != fluid.result.size() ||
            (nsr > 0 && nsolid_out != (size_t)(n - nfluid))) {
            std::fprintf(stderr, "mismatch: output record counts (fluid %zu of %ld, solid %zu)\n", fluid.result.size(),
                         nfluid, nsolid_out);
            ++bad;
        }
        auto check_rec = [&](const hsolver::OutputData& o, long q, bool fl) {
            const bool ok = o.t == tt[q] && o.locationx == loc[3 * q] && o.locationy == loc[3 * q + 1] &&
                            o.locationz == loc[3 * q + 2] &&
                            (fl ? (o.ux == vel[3 * q] && o.uy == vel[3 * q + 1] && o.uz == vel[3 * q + 2] &&
                                   o.p == pp[q] && o.globalid == ids[q])
                                : (o.ux == 0 && o.uy == 0 && o.uz == 0 && o.p == 0 && o.globalid == -1));
            if (!ok) { std::fprintf(stderr, "mismatch: output record of cell %ld\n", q); ++bad; }
        };
        for (size_t r = 0; r < fluid.result.size() && bad < 10; ++r) check_rec(fluid.result[r], fluid.idArray[r] - 1, true);
        for (size_t r = 0; nsr > 0 && r < mob.result.size() && bad < 10; ++r) check_rec(mob.result[r], mob.idArray[r] - 1, false);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        hs_mesh_free(mesh);
        return 2;
    }
    hs_mesh_free(mesh);
    if (bad) return 3;
    std::printf("shim check: %d iterations, %dx%dx%d, struct loop and Solver::solve bit-identical to hs_run, "
                "output records match hs_get_output\n", iters, nx, ny, nz);
    return 0;
}
