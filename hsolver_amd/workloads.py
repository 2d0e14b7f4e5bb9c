"""The benchmark configurations other than the headline one, run on the GPU.

    python -m hsolver_amd.workloads c1 [--iters 200]
    python -m hsolver_amd.workloads c2 [--iters 500]
    python -m hsolver_amd.workloads c5 [--iters 50] [--precond gmg|amg] [--energy gmg,amg]

c1  64x16x16 channel, solid cube, Re 100, k ratio 100, q = 1: the reference's solver
    settings (lexicographic sweeps, unpreconditioned CG to 1e-8) and the fast path
    (red-black sweeps, multigrid FCG), SIMPLE iterations/s and balance residuals.
c2  256x64x64 channel (1.05M cells), heated cube with seeded q, Re 400: the FP64 drop-in
    (red-black sweeps, FP64 unpreconditioned CG = the reference's solver) and FP64
    multigrid FCG; SIMPLE iterations/s and the HBM rate of the p' SpMV and of one
    x-momentum colour pass, timed live with HIP events around every launch:
      SpMV: the first of each solve (k_cg_spmv) 41 B per cell (p, three face-coefficient
      arrays, q written, class); the others form p = r + beta p inside the apply
      (k_fcg_spmv<double>) at 57 B per cell (r and the old p read, the new p and q
      written, three face-coefficient arrays, class)
      momentum colour pass (k_mom_rb<0>): 41 B per x-face (u, v, w, p_temp read,
      u and a_P written on the colour, class)
c5  512x256x256 channel (33.5M cells), y/z tanh-stretched to cell aspect ratio 50, 64
    heated plate fins (20% of the cells), Re 2000, k ratio 1000: mixed-precision SIMPLE
    (multigrid FCG with the dense coarse solve; --precond gmg|amg picks the geometric or
    the aggregation multigrid), then on the frozen flow the energy solve and the adjoint
    of the fins' temperature p-norm (p = 8) with respect to the fin heat sources, checked
    by central differences on seeded fin cells, once per energy preconditioner in
    --energy (default: geometric, then aggregation multigrid).
Each prints one JSON line.  Inputs are synthetic (configs.py), resident in HBM.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time

from . import abi, configs, core

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[workloads] {msg}", file=sys.stderr, flush=True)


def timed_run(case, order, precond, precision, iters, prof_kernel=0):
    lib = core.library()
    lib.hs_prof_set.argtypes = [C.c_void_p, C.c_int32]
    lib.hs_prof_get.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    p = abi.Params.from_buffer_copy(case.params)
    p.order, p.precond, p.precision = order, precond, precision
    ctx = core.create(case.grid, p, case.q)
    ctx.init()
    ctx.step()                                   # warm-up (allocations, hierarchy)
    if prof_kernel:
        lib.hs_prof_set(ctx.h, prof_kernel)
    hist = []
    t0 = time.perf_counter()
    for s in range(iters):
        hist.append(ctx.step())
        if (s + 1) % 100 == 0:
            log(f"{case.name}: {s + 1}/{iters} iterations, mass imbalance {hist[-1]['mass_imbalance']:.3e}")
    dt = time.perf_counter() - t0
    prof = None
    if prof_kernel:
        n, ms = C.c_int64(), C.c_double()
        lib.hs_prof_get(ctx.h, C.byref(n), C.byref(ms))
        prof = (n.value, ms.value)
        lib.hs_prof_set(ctx.h, 0)
    ctx.close()
    return dt, hist, prof


def summary(dt, hist, iters):
    last = hist[-1]
    return {"simple_iters_per_s": round(iters / dt, 3), "ms_per_iter": round(1e3 * dt / iters, 3),
            "p_iters_last": last["p_iters"], "p_iters_mean": round(sum(h["p_iters"] for h in hist) / len(hist), 1),
            "mass_imbalance": last["mass_imbalance"], "energy_residual": last["energy_residual"],
            "p_true_relres_max": max(h["p_true_relres"] for h in hist)}


def rate(prof, bytes_per_launch):
    n, ms = prof
    avg = ms / max(n, 1)
    gbs = bytes_per_launch / (avg * 1e-3) / 1e9
    return {"launches": n, "avg_us": round(avg * 1e3, 2), "GB_s": round(gbs, 1), "frac_of_8TBs": round(gbs / HBM_PEAK_GBS, 4)}


def run_c1(iters):
    case = configs.c1()
    out = {"workload": "c1 channel 64x16x16, solid cube, Re=100, k ratio 100", "iters": iters}
    dt, hist, _ = timed_run(case, abi.HS_ORDER_LEX, abi.HS_PRECOND_NONE, abi.HS_PREC_FP64, iters)
    out["reference_settings"] = summary(dt, hist, iters)
    for name, prec in (("fast_fp64", abi.HS_PREC_FP64), ("fast_mixed", abi.HS_PREC_MIXED)):
        dt, hist, _ = timed_run(case, abi.HS_ORDER_RB, abi.HS_PRECOND_GMG, prec, iters)
        out[name] = summary(dt, hist, iters)
    return out


def run_c2(iters):
    case = configs.c2()
    nx, ny, nz = case.shape
    ncell = nx * ny * nz
    out = {"workload": f"c2 channel {nx}x{ny}x{nz}, heated cube, Re=400", "iters": iters}
    dt, hist, prof = timed_run(case, abi.HS_ORDER_RB, abi.HS_PRECOND_NONE, abi.HS_PREC_FP64, iters,
                               prof_kernel=2)
    out["fp64_dropin_cg"] = summary(dt, hist, iters)
    n_first = iters                              # one solve per timed SIMPLE iteration
    n, ms = prof
    per_launch = (n_first * 41.0 + (n - n_first) * 57.0) * ncell / max(n, 1)
    out["fp64_dropin_cg"]["spmv"] = rate(prof, per_launch)
    dt, hist, prof = timed_run(case, abi.HS_ORDER_RB, abi.HS_PRECOND_GMG, abi.HS_PREC_FP64, iters,
                               prof_kernel=3)
    out["fp64_gmg"] = summary(dt, hist, iters)
    out["fp64_gmg"]["momentum_colour_pass"] = rate(prof, 41.0 * (nx + 1) * ny * nz)
    return out


PRECONDS = {"gmg": abi.HS_PRECOND_GMG, "amg": abi.HS_PRECOND_AMG}


def energy_and_adjoint(ctx, case, fd_cells, tag):
    """Energy solve, adjoint of the fins' temperature p-norm (p = 8) with respect to the fin
    heat sources, and central differences on fd_cells seeded fin cells, on ctx's flow."""
    import numpy as np
    from .adjoint import central_difference, cell_volumes
    nx, ny, nz = case.shape
    shape = (nz, ny, nx)
    solid = case.grid.cell_type.reshape(shape) == 1
    V = cell_volumes(case)
    t0 = time.perf_counter()
    fwd = ctx.energy_solve(1e-11, 2000)
    t_fwd = time.perf_counter() - t0
    T = ctx.cells("tf")
    S = float(np.sum(np.where(solid, T ** 8 * V, 0.0)))
    J = S ** 0.125
    dJdT = np.where(solid, J ** -7 * T ** 7 * V, 0.0)
    t0 = time.perf_counter()
    dq, dk, adj = ctx.adjoint_energy(dJdT, 1e-11, 2000)
    t_adj = time.perf_counter() - t0
    log(f"c5 [{tag}] energy solve {fwd}, {t_fwd:.2f} s; adjoint {adj}, {t_adj:.2f} s; J = {J:.6e}")
    q0, k0 = ctx.cells("q").copy(), ctx.cells("k").copy()
    idx = np.flatnonzero(solid.ravel())
    pick = idx[(configs.splitmix64(55, fd_cells) * idx.size).astype(np.int64)]
    errs = []
    for c in pick:
        cell = np.unravel_index(c, shape)
        fd = central_difference(ctx, "q", cell, 1e-6 * q0[cell], q0, k0, solid, V, tol=1e-11, maxit=2000)
        errs.append(abs(fd - dq[cell]) / abs(dq[cell]))
        log(f"c5 [{tag}] central difference at fin cell {tuple(int(v) for v in cell)}: {fd:.9e} vs adjoint {dq[cell]:.9e}")
    ctx.put_cells("q", q0)
    return {"J": J, "forward": fwd, "adjoint": adj, "forward_s": round(t_fwd, 3), "adjoint_s": round(t_adj, 3),
            "fd_cells": fd_cells, "fd_max_rel_err_q": max(errs) if errs else None}


def run_c5(iters, fd_cells=4, precond="gmg", energy=("gmg", "amg")):
    case = configs.c5()
    nx, ny, nz = case.shape
    out = {"workload": case.workload, "cells": nx * ny * nz, "solid_fraction": round(float(case.grid.cell_type.mean()), 4),
           "max_cell_aspect": round(configs.max_cell_aspect(case.grid.xv, case.grid.yv, case.grid.zv), 2),
           "iters": iters, "p_precond": precond}
    p = abi.Params.from_buffer_copy(case.params)
    p.order, p.precond, p.precision = abi.HS_ORDER_RB, PRECONDS[precond], abi.HS_PREC_MIXED
    p.mg_dense_coarse = 50
    p.e_precond = PRECONDS[energy[0]]
    ctx = core.create(case.grid, p, case.q)
    ctx.init()
    for w in range(2):
        t0 = time.perf_counter()
        st = ctx.step()
        log(f"c5 warm-up {w + 1}: p' iters {st['p_iters']}, true relres {st['p_true_relres']:.2e}, "
            f"{time.perf_counter() - t0:.2f} s")
    hist = []
    t0 = time.perf_counter()
    for s in range(iters):
        hist.append(ctx.step())
        if (s + 1) % 10 == 0:
            log(f"c5: {s + 1}/{iters} SIMPLE iterations, p' iters {hist[-1]['p_iters']}, "
                f"mass imbalance {hist[-1]['mass_imbalance']:.3e}")
    dt = time.perf_counter() - t0
    out["mixed_simple"] = summary(dt, hist, iters)
    if precond == "amg":
        out["mixed_simple"]["amg_levels"] = ctx.amg_levels(0)
    # energy solve and adjoint on the frozen flow, with each energy preconditioner asked for
    for k, name in enumerate(energy):
        if k == 0:
            c2 = ctx
        else:
            p2 = abi.Params.from_buffer_copy(p)
            p2.e_precond = PRECONDS[name]
            c2 = core.create(case.grid, p2, case.q)
            c2.init()
            c2.load_state(ctx.state())
        out[f"adjoint_{name}"] = energy_and_adjoint(c2, case, fd_cells, name)
        if name == "amg":
            out[f"adjoint_{name}"]["amg_levels"] = c2.amg_levels(1)
        if k > 0:
            c2.close()
    ctx.close()
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("config", choices=["c1", "c2", "c5"])
    ap.add_argument("--iters", type=int, default=None)
    ap.add_argument("--precond", choices=sorted(PRECONDS), default="gmg", help="c5: preconditioner of the p' solves")
    ap.add_argument("--energy", default="gmg,amg",
                    help="c5: energy / adjoint preconditioners to run on the frozen flow, comma-separated")
    ap.add_argument("--fd-cells", type=int, default=4, help="c5: central differences per energy preconditioner")
    args = ap.parse_args(argv)
    if args.config == "c1":
        out = run_c1(args.iters or 200)
    elif args.config == "c2":
        out = run_c2(args.iters or 500)
    else:
        out = run_c5(args.iters or 50, args.fd_cells, args.precond, tuple(args.energy.split(",")))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
