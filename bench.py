#!/usr/bin/env python3
"""SIMPLE conjugate-heat-transfer iterations/s on one MI355X.

Workload (the metric's configuration): 256^3 unit box with 27 seeded solid blocks,
conductivity ratio 1000, lid + inlet forcing at Re = 1000, configuration physics,
red-black momentum / energy sweeps, p' solved to a 1e-8 relative residual (||b - A p'|| /
||b||), each solve starting from the parabola through the last three SIMPLE iterations'
p' (3 x1 - 3 x2 + x3, or zero when that is no better; --cold: from zero).
  value      = mixed-precision path (FP32 multigrid V-cycle inside FP64 flexible CG; the
               16^3 coarsest level solved with a dense FP32 inverse built on the matrix
               cores plus FP64 refinement, rebuilt at most every --dense hierarchy builds and
               sooner when the FCG iteration count rises by 2)
  fp64_value = the same step with the V-cycle in FP64 and the dense coarse inverse in FP64
               (v_mfma_f64_16x16x4f64), rebuilt on the same rule
The default 50 timed steps after 5 warmup steps include one rebuild of each inverse
(hierarchy build 51), so the rate carries its amortised cost.
Each timed step is one full SIMPLE outer iteration (x/y/z momentum, p' assembly +
solve, velocity / pressure correction, conjugate energy sweep, set-for-next), all
inputs resident in HBM.

roofline: the dominant kernel (one fused red-black Gauss-Seidel sweep of the finest
multigrid level, both colours, out of place), timed live with HIP events on the
library's stream around every launch in the timed region.  Algorithmic bytes per
launch = 28 B per cell in FP32: the three face-coupling arrays, ground, right-hand side
and the old iterate read (24 B), the new iterate written (4 B).

cpu_baseline: the CPU oracle (reference algorithm: lexicographic Gauss-Seidel sweeps,
unpreconditioned CG with HYPRE's stop rule), the first SIMPLE iteration of the same
workload timed in full, the CG on the host threads of one GPU's share (16, or fewer if
the host has fewer), the sweeps on one thread.  A baseline, not the target.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "SIMPLE iters/s at 256³ CHT; mixed vs FP64 GPU at 1e-8 rel p-residual"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def gpu_run(case, precision, precond, steps, warmup, prof_kernel=None, order=None, dense=0, fields_out=None,
            step_ms=None, warm=0, energy_out=None):
    from hsolver_amd import abi, core
    lib = core.library()
    lib.hs_prof_set.argtypes = [C.c_void_p, C.c_int32]
    lib.hs_prof_get.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    p = abi.Params.from_buffer_copy(case.params)
    p.precision = precision
    p.precond = precond
    p.order = abi.HS_ORDER_RB if order is None else order
    p.mg_dense_coarse = dense
    p.p_warm = warm
    p.e_precond = abi.HS_PRECOND_AMG         # only the energy solve after the timed region uses it
    ctx = core.create(case.grid, p, case.q)
    ctx.init()
    for w in range(warmup):
        st = ctx.step()
        log(f"warmup {w + 1}/{warmup}: p' iters {st['p_iters']} true relres {st['p_true_relres']:.2e}")
    if prof_kernel:
        lib.hs_prof_set(ctx.h, prof_kernel)
    stats = []
    t0 = time.perf_counter()
    tp = t0
    for s in range(steps):
        stats.append(ctx.step())        # synchronises at its end
        if step_ms is not None:
            tn = time.perf_counter()
            step_ms.append(round((tn - tp) * 1e3, 3))
            tp = tn
    t1 = time.perf_counter()
    prof = None
    if prof_kernel:
        n, ms = C.c_int64(), C.c_double()
        lib.hs_prof_get(ctx.h, C.byref(n), C.byref(ms))
        prof = (n.value, ms.value)
        lib.hs_prof_set(ctx.h, 0)
    if fields_out is not None:
        fields_out.update(ctx.fields())      # after the timed region
    if energy_out is not None:
        # after the timed region: the global energy balance the SIMPLE sweeps left, then the
        # conjugate energy equation solved on the final (frozen) flow and its balance
        _, r0 = ctx.energy_balance()
        te = time.perf_counter()
        info = ctx.energy_solve(1e-10, 1000)
        te = time.perf_counter() - te
        _, r1 = ctx.energy_balance()
        energy_out.update({"after_simple_steps": r0, "after_energy_solve": r1,
                           "energy_solve": {"preconditioner": "aggregation multigrid", "iters": info["iters"],
                                            "true_relres": info["true_relres"], "s": round(te, 3)}})
    ctx.close()
    return (t1 - t0), stats, prof


def field_distance(a, b):
    import numpy as np
    out = {}
    for k in ("u", "v", "w", "p", "t"):
        if k in a and k in b:
            den = float(np.linalg.norm(b[k]))
            out[k] = float(np.linalg.norm(a[k] - b[k]) / den) if den > 0 else float(np.linalg.norm(a[k]))
    return out


def pmc_traffic(n):
    """HBM bytes per launch of the roofline kernel, from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by what the counter reads for a plain
    512 MiB device copy).  Counters cannot be read inside a timed run, so the figure is the
    profiled one; None for other grid sizes."""
    f = ROOT / "profiles" / "k_mg_rbsweep_float_c3_256.md"
    if n != 256 or not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def cpu_threads() -> int:
    """Host threads for the CPU leg: the GPU box's CPU share (16 per GPU), at most what the
    host has."""
    return max(1, min(16, os.cpu_count() or 1))


def cpu_sample(case, threads):
    """The oracle (the reference algorithm) on the host: the first SIMPLE iteration of the
    workload in full, lexicographic sweeps on one thread, the p' CG (unpreconditioned, HYPRE's
    stop rule) on `threads` z-slab threads.  Returns (seconds, CG iterations, sweep seconds)."""
    from hsolver_amd import abi
    from hsolver_amd.core import Context
    olib = C.CDLL(str(ROOT / "oracle" / "liboracle.so"))
    abi.declare(olib, "or")
    olib.or_set_threads.argtypes = [C.c_int32]
    olib.or_set_threads(threads)
    p = abi.Params.from_buffer_copy(case.params)
    p.order = abi.HS_ORDER_LEX
    p.precond = abi.HS_PRECOND_NONE
    ctx = Context(olib, "or", case.grid, p)
    ctx.set_source(case.q)
    ctx.init()
    t0 = time.perf_counter()
    for comp in range(3):
        ctx.momentum(comp)
    t_mom = time.perf_counter() - t0
    st = ctx.pressure_correction()
    ctx.correct()
    ctx.fluid_energy()
    ctx.solid_solve()
    ctx.exchange()
    ctx.set_for_next()
    t = time.perf_counter() - t0
    ctx.close()
    olib.or_set_threads(1)
    log(f"cpu oracle first step: {t:.2f} s ({st['p_iters']} CG iterations on {threads} threads, momentum "
        f"sweeps {t_mom:.2f} s)")
    return t, st["p_iters"], t_mom


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256, help="grid cells per side (the workload is 256)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 multigrid comparison run")
    ap.add_argument("--no-cg", action="store_true", help="skip the unpreconditioned FP64 CG drop-in step")
    ap.add_argument("--dense", type=int, default=50, metavar="N",
                    help="coarsest level (16^3) solved with a dense inverse built on the matrix cores (FP32 on "
                         "the mixed path, FP64 on the FP64 one), rebuilt at most every N hierarchy builds and "
                         "sooner when the FCG iteration count rises by 2 (0: red-black sweeps there)")
    ap.add_argument("--cold", action="store_true",
                    help="start every multigrid p' solve from zero instead of extrapolating the last SIMPLE p' solutions")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="also run a plain device copy of BYTES (reference for byte counters)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {max(args.gpus, world)}); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    from hsolver_amd import abi, configs
    if args.copy_calibration:
        from hsolver_amd import core
        lib = core.library()
        lib.hs_copy_bench.argtypes = [C.c_int64, C.c_int32, C.POINTER(C.c_double)]
        ms = C.c_double()
        rc = lib.hs_copy_bench(args.copy_calibration, 3, C.byref(ms))
        log(f"copy calibration: {args.copy_calibration} B in {ms.value:.3f} ms "
            f"({2 * args.copy_calibration / ms.value / 1e6:.0f} GB/s read+write), rc {rc}")
    t_case = time.perf_counter()
    case = configs.c3(args.n)
    nx, ny, nz = case.shape
    ncell = nx * ny * nz
    solid_frac = float(case.grid.cell_type.mean())
    log(f"workload {case.workload}: {ncell} cells, solid fraction {solid_frac:.3f} "
        f"(built in {time.perf_counter() - t_case:.1f} s)")

    f_mixed, f_fp64 = {}, {}
    ms_mixed, ms_fp64 = [], []
    energy = {}
    dt_mixed, st_mixed, prof = gpu_run(case, abi.HS_PREC_MIXED, abi.HS_PRECOND_GMG, args.steps, args.warmup,
                                       prof_kernel=1, dense=args.dense, fields_out=f_mixed, step_ms=ms_mixed,
                                       warm=0 if args.cold else 3, energy_out=energy)
    log(f"energy balance after the SIMPLE steps {energy['after_simple_steps']:.3e}, after the energy solve "
        f"{energy['after_energy_solve']:.3e} ({energy['energy_solve']})")
    value = args.steps / dt_mixed
    log(f"mixed: {value:.2f} SIMPLE it/s, p' iters {[s['p_iters'] for s in st_mixed]}")

    fp64_value = None
    st_fp64 = []
    if not args.no_fp64:
        dt64, st_fp64, _ = gpu_run(case, abi.HS_PREC_FP64, abi.HS_PRECOND_GMG, args.steps, args.warmup,
                                   dense=args.dense, fields_out=f_fp64, step_ms=ms_fp64, warm=0 if args.cold else 3)
        fp64_value = args.steps / dt64
        log(f"fp64 gmg: {fp64_value:.2f} SIMPLE it/s, p' iters {[s['p_iters'] for s in st_fp64]}")

    cg_step = None
    if not args.no_cg:
        # the reference's solver as coded (HYPRE PCG, identity preconditioner), first step
        dtcg, stcg, _ = gpu_run(case, abi.HS_PREC_FP64, abi.HS_PRECOND_NONE, 1, 0)
        cg_step = {"s_per_step": dtcg, "p_iters": stcg[0]["p_iters"],
                   "p_true_relres": stcg[0]["p_true_relres"]}
        log(f"fp64 cg (reference solver) first step: {dtcg:.2f} s, {stcg[0]['p_iters']} iterations")

    n_launch, ms_total = prof
    avg_ms = ms_total / max(n_launch, 1)
    bytes_per_launch = 28.0 * ncell
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(args.n),
                "kernel": "k_mg_rbsweep<float,256,4> (finest level, one full fused red-black sweep)",
                "launches": n_launch, "avg_us": round(avg_ms * 1e3, 2),
                "bytes_per_launch": bytes_per_launch}

    cpu = None
    if not args.no_cpu and int(os.environ.get("RANK", "0")) == 0:
        nt = cpu_threads()
        t_full, n_cg, t_mom = cpu_sample(case, nt)
        cpu = {"value": round(1.0 / t_full, 6), "unit": "SIMPLE iters/s", "cores": nt, "kind": "port",
               "sample": (f"CPU oracle (the reference algorithm: lexicographic Gauss-Seidel sweeps, unpreconditioned "
                          f"CG with HYPRE's stop rule), the first SIMPLE iteration of the same {nx}x{ny}x{nz} "
                          f"workload timed in full: {n_cg} CG iterations on {nt} threads (z-slabs, per-slab partial "
                          f"sums), the sweeps on one thread ({t_mom:.2f} s of momentum sweeps); host "
                          f"os.cpu_count() = {os.cpu_count()}"),
               "s_per_iter": round(t_full, 2), "host_cpus": os.cpu_count()}

    relres = [s["p_true_relres"] for s in st_mixed]
    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "SIMPLE iters/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1000.0 / value, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32",
        "data": "synthetic",
        "config": {"workload": case.workload, "grid": [nx, ny, nz], "solid_fraction": round(solid_frac, 4),
                   "p_tol": case.params.p_tol, "krylov": "flexible CG",
                   "p_start": "zero" if args.cold else "parabola through the last 3 SIMPLE p' solutions (3 x1 - 3 x2 + x3; zero if no better)", "preconditioner": "GMG, V(1,2) on the finest level and V(3,3) below, fused red-black SOR (omega 1.2 on the finest level, 1.6 below)",
                   "coarse_mixed": (f"16^3 dense FP32 inverse on MFMA, rebuilt at most every {args.dense} builds "
                                    "(sooner on an FCG iteration rise of 2), 1 FP64 refinement step"
                                    if args.dense else "single-workgroup RB-GS tail"),
                   "coarse_fp64": (f"16^3 dense FP64 inverse on MFMA (v_mfma_f64_16x16x4f64), rebuilt at most "
                                   f"every {args.dense} builds (sooner on an FCG iteration rise of 2), 1 FP64 refinement step"
                                   if args.dense else "single-workgroup RB-GS tail"),
                   "sweep_order": "red-black", "physics": "configuration (inlet+lid+outflow, conjugate energy)"},
        "fp64_value": round(fp64_value, 4) if fp64_value else None,
        "mixed_over_fp64": round(value / fp64_value, 4) if fp64_value else None,
        # both paths run the same warmup + timed SIMPLE iterations from the same start, each
        # solving p' to 1e-8: relative L2 distance of the mixed path's fields from the FP64 one
        "mixed_vs_fp64_rel_l2": field_distance(f_mixed, f_fp64) if f_fp64 else None,
        # wall time of each timed step (the step holding the inverse's rebuild stands out)
        "step_ms_mixed": {"median": sorted(ms_mixed)[len(ms_mixed) // 2], "max": max(ms_mixed)} if ms_mixed else None,
        "step_ms_fp64": {"median": sorted(ms_fp64)[len(ms_fp64) // 2], "max": max(ms_fp64)} if ms_fp64 else None,
        "p_iters_mixed": [s["p_iters"] for s in st_mixed],
        "p_iters_fp64": [s["p_iters"] for s in st_fp64],
        "p_true_relres_max": max(relres) if relres else None,
        "mass_imbalance_last": st_mixed[-1]["mass_imbalance"] if st_mixed else None,
        "energy_residual_last": st_mixed[-1]["energy_residual"] if st_mixed else None,
        "energy_balance": energy,
        "fp64_cg_first_step": cg_step,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
