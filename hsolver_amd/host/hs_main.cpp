// hs_main.cpp — host driver with the shape of the reference's main.cpp:10-161 on one
// rank: read the `.in` mesh, build the fluid and solid solvers, run Solver::solve,
// write result.vtk.
//
//   hsolver_main <file.in> [iterations=100] [out.vtk=result.vtk] [--order lex|rb] [--raw]
//
// result.vtk holds the reference's output records (Solver::solve solver.h:160-213 packing,
// main.cpp:101-156 layout and number format); --raw writes the solved arrays instead, at
// round-trip precision.
//
// The physics is the reference's (HS_MODE_REF, every quirk kept); decomposition is
// collapsed to one rank and one GPU, so the processor files of the decomposer are not
// needed.  Exit code 0 on success, 1 on usage / input errors, 2 on solver errors.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "hs_solver.hpp"

static int usage(const char* prog) {
    std::fprintf(stderr,
                 "usage: %s <file.in> [iterations>=1] [out.vtk] [--order lex|rb] [--proc DIR] [--raw]\n"
                 "       %s <file.in> --decompose DIR   (write DIR/processor0/ and exit)\n",
                 prog, prog);
    return 1;
}

int main(int argc, char** argv) {
    if (argc < 2) return usage(argv[0]);
    const char* in = argv[1];
    int iters = 100;
    std::string out = "result.vtk";
    int order = HS_ORDER_LEX;
    int pos = 0;
    const char* proc_dir = nullptr;
    const char* decompose_dir = nullptr;
    bool raw = false;
    for (int a = 2; a < argc; ++a) {
        if (!std::strcmp(argv[a], "--order") && a + 1 < argc) {
            const char* o = argv[++a];
            if (!std::strcmp(o, "rb")) order = HS_ORDER_RB;
            else if (!std::strcmp(o, "lex")) order = HS_ORDER_LEX;
            else return usage(argv[0]);
        } else if (!std::strcmp(argv[a], "--proc") && a + 1 < argc) {
            proc_dir = argv[++a];
        } else if (!std::strcmp(argv[a], "--decompose") && a + 1 < argc) {
            decompose_dir = argv[++a];
        } else if (!std::strcmp(argv[a], "--raw")) {
            raw = true;
        } else if (argv[a][0] == '-' && argv[a][1] == '-') {
            return usage(argv[0]);                     // unknown flag, or a flag without its value
        } else if (pos == 0) {
            char* end = nullptr;
            const long n = std::strtol(argv[a], &end, 10);
            if (end == argv[a] || *end != '\0' || n < 1 || n > 100000000) return usage(argv[0]);
            iters = (int)n;
            ++pos;
        } else if (pos == 1) {
            out = argv[a];
            ++pos;
        } else {
            return usage(argv[0]);
        }
    }
    hs_mesh* mesh = nullptr;
    int rc = hs_mesh_read(in, &mesh);
    if (rc != HS_OK) {
        std::fprintf(stderr, "cannot read %s: %s\n", in, hs_strerror(rc));
        return 1;
    }
    if (decompose_dir) {
        // decompose.cpp:5-31 with one part
        rc = hs_mesh_write_proc(mesh, decompose_dir);
        if (rc != HS_OK) std::fprintf(stderr, "cannot write %s/processor0: %s\n", decompose_dir, hs_strerror(rc));
        else std::printf("Info: wrote %s/processor0 (cells, faces, boundary)\n", decompose_dir);
        hs_mesh_free(mesh);
        return rc == HS_OK ? 0 : 1;
    }
    if (proc_dir) {
        // main.cpp:18-21: the rank's region cells, faces and boundaries
        rc = hs_mesh_read_proc(mesh, proc_dir);
        if (rc != HS_OK) {
            std::fprintf(stderr, "cannot read %s/processor0: %s\n", proc_dir, hs_strerror(rc));
            hs_mesh_free(mesh);
            return 1;
        }
        std::printf("Info: cell types from %s/processor0\n", proc_dir);
    }
    int32_t nx, ny, nz, nmobs, nwalls, nfr, nsr;
    hs_mesh_info(mesh, &nx, &ny, &nz, &nmobs, &nwalls, &nfr, &nsr);
    std::printf("Info: mesh %dx%dx%d, %d mobs boxes, %d fluid / %d solid regions\n", nx, ny, nz, nmobs, nfr, nsr);
    hs_grid g;
    hs_mesh_grid(mesh, &g);
    hs_params p;
    hs_params_default(HS_MODE_REF, &p);
    p.order = order;
    try {
        hsolver::Context ctx(g, p);
        hsolver::FluidSolver fluid(ctx);
        hsolver::MobSolver mob(ctx);
        hsolver::Solver solver;
        solver.addFluids(&fluid);
        if (nsr > 0) solver.addMobs(&mob);
        const int warm = iters > 10 ? 10 : 0;
        solver.solve(iters, warm);
        const hs_stats& last = solver.history.back();
        std::printf("Info: %d iterations; last: p' CG %d iters (relres %.3e), sweeps u/v/w %d/%d/%d, T %d, solid %d\n",
                    iters, last.p_iters, last.p_relres, last.mom_sweeps[0], last.mom_sweeps[1], last.mom_sweeps[2],
                    last.t_sweeps, last.ts_sweeps);
        if (iters > warm)
            std::printf("Info: wall time %.3f s over iterations %d-%d, average %.4f s\n", solver.seconds_after_warmup,
                        warm, iters - 1, solver.seconds_after_warmup / (iters - warm));
        if (raw) hsolver::write_vtk_raw(out.c_str(), ctx.get(), g);
        else hsolver::write_vtk(out.c_str(), solver, nx, ny, nz);
        std::printf("VTK file written: %s\n", out.c_str());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        hs_mesh_free(mesh);
        return 2;
    }
    hs_mesh_free(mesh);
    return 0;
}
