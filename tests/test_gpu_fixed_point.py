"""The benchmarked path against the reference's algorithm where ordering and solver
drop out, and the headline geometry as a trajectory.

  * c1 to its steady state twice: the fast path (red-black sweeps, FP32 multigrid inside
    FP64 flexible CG, MFMA dense coarse solve) on the GPU, and the reference's order
    (lexicographic Gauss-Seidel, unpreconditioned CG with HYPRE's stop rule) in the
    oracle.  At the fixed point of SIMPLE the sweep order and the Krylov method no
    longer enter: velocity, pressure (about its mean: only its differences enter the
    momentum equations) and, after the energy equation is solved on the converged flow,
    temperature agree to 1e-6, and both close the global energy balance;
  * the 256^3 configuration's geometry at 64^3 (27 blocks, lid + inlet, Re 1000,
    k ratio 1000): 10 SIMPLE iterations of the FP64 and of the mixed multigrid paths
    against the oracle's red-black sweeps with its CG, all at 1e-10.
"""
import numpy as np
import pytest

from hsolver_amd import abi, configs

from helpers import energy_balance, gpu_ctx, oracle_ctx, rel_l2

pytestmark = pytest.mark.gpu

MASS_TOL = 1e-11


def converge(ctx, max_steps, label):
    for it in range(max_steps):
        st = ctx.step()
        if it % 200 == 199:
            print(f"[fixed point] {label} step {it + 1}: mass imbalance {st['mass_imbalance']:.3e}", flush=True)
        if st["mass_imbalance"] < MASS_TOL:
            return it + 1, st
    return max_steps, st


@pytest.mark.timeout(600)   # about a minute of oracle CPU time (1400 lexicographic SIMPLE iterations)
def test_c1_fixed_point_fast_path_vs_reference_order(hs_lib, oracle_lib):
    case = configs.c1()
    gp = abi.Params.from_buffer_copy(case.params)
    gp.order, gp.precond, gp.precision, gp.mg_dense_coarse = (abi.HS_ORDER_RB, abi.HS_PRECOND_GMG,
                                                              abi.HS_PREC_MIXED, 50)
    op = abi.Params.from_buffer_copy(case.params)
    op.order, op.precond = abi.HS_ORDER_LEX, abi.HS_PRECOND_NONE
    g, o = gpu_ctx(hs_lib, case, gp), oracle_ctx(oracle_lib, case, op)
    g.init(); o.init()
    ng, sg = converge(g, 4000, "gpu rb+mixed gmg")
    no, so = converge(o, 4000, "oracle lex+cg")
    assert sg["mass_imbalance"] < MASS_TOL and so["mass_imbalance"] < MASS_TOL, (sg, so)
    # temperature: the energy equation on the converged (frozen) flow, solved exactly
    ig, io = g.energy_solve(1e-13, 2000), o.energy_solve(1e-13, 50000)
    assert ig["converged"] == 1 and io["converged"] == 1, (ig, io)
    fg, fo = g.fields(), o.fields()
    # the momentum equations see pressure differences only, and nothing fixes the level
    # in this physics (p accumulates each path's own corrections): compare p about its
    # mean over the fluid cells, and report the level difference
    fl = case.grid.cell_type.reshape(fg["p"].shape) == 0
    shift = float(np.mean(fg["p"][fl]) - np.mean(fo["p"][fl]))
    for f in (fg, fo):
        f["p"] = np.where(fl, f["p"] - np.mean(f["p"][fl]), 0.0)
    errs = {key: rel_l2(fg[key], fo[key]) for key in fg}
    print(f"[fixed point] c1: {ng} GPU / {no} oracle SIMPLE iterations to mass imbalance {MASS_TOL:g}; "
          f"rel-L2 GPU vs oracle {errs} (p about its mean; level difference {shift:.3e})", flush=True)
    assert all(e < 1e-6 for e in errs.values()), errs
    for label, ctx in (("gpu", g), ("oracle", o)):
        q_in, q_out, q_mass = energy_balance(case, ctx.cells("tf"), ctx.get("u"), ctx.get("v"), ctx.get("w"),
                                             ctx.cells("k"), ctx.cells("q"))
        print(f"[fixed point] {label} energy balance: in {q_in:.12e} out {q_out:.12e} "
              f"(residual mass imbalance carries {q_mass:.3e})", flush=True)
        assert abs(q_in - q_out + q_mass) < 1e-7 * q_in, (label, q_in, q_out, q_mass)   # 4.5e-9 measured
        assert abs(q_mass) < 1e-4 * q_in
    g.close()
    o.close()


@pytest.fixture(scope="module")
def c3_64_oracle(oracle_lib):
    """10 SIMPLE iterations of the oracle (red-black sweeps, CG at 1e-10) on the 256^3
    configuration's geometry at 64^3 (about 25 s of CPU)."""
    case = configs.c3(64)
    case.params.order = abi.HS_ORDER_RB
    case.params.p_tol = 1e-10
    o = oracle_ctx(oracle_lib, case)
    o.init()
    hist = []
    for it in range(10):
        so = o.step()
        assert so["p_converged"] == 1, (it, so)
        hist.append(so)
        print(f"[c3 64] oracle step {it}: CG {so['p_iters']} iterations", flush=True)
    f = o.fields()
    o.close()
    return case, f, hist


@pytest.mark.timeout(600)
@pytest.mark.parametrize("precision,warm", [(abi.HS_PREC_FP64, 0), (abi.HS_PREC_MIXED, 0), (abi.HS_PREC_MIXED, 3),
                                            (abi.HS_PREC_FP64, 3)], ids=["fp64", "mixed", "mixed_warm", "fp64_warm"])
def test_c3_geometry_trajectory_vs_oracle(hs_lib, c3_64_oracle, precision, warm):
    """warm = 3: each p' solve starts from the parabola through the last three solutions
    (p_warm); the trajectory stays within the same bound of the oracle's zero-start CG."""
    case, fo, ho = c3_64_oracle
    gp = abi.Params.from_buffer_copy(case.params)
    gp.precond, gp.precision = abi.HS_PRECOND_GMG, precision
    gp.p_warm = warm
    gp.mg_dense_coarse = 50 if precision == abi.HS_PREC_MIXED else 0
    g = gpu_ctx(hs_lib, case, gp)
    g.init()
    for it in range(10):
        sg = g.step()
        assert sg["p_true_relres"] < 1e-9, (it, sg)
        print(f"[c3 64] step {it}: FCG {sg['p_iters']} iterations, mass imbalance {sg['mass_imbalance']:.6e} "
              f"(oracle {ho[it]['mass_imbalance']:.6e})", flush=True)
    fg = g.fields()
    errs = {key: rel_l2(fg[key], fo[key]) for key in fg}
    print(f"[c3 64] {precision}: rel-L2 vs oracle after 10 iterations {errs}", flush=True)
    assert all(e < 1e-7 for e in errs.values()), errs
    g.close()
