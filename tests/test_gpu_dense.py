"""The dense coarsest-level solve (MFMA Gauss-Jordan inverse, FP32 under the mixed V-cycle,
FP64 on v_mfma_f64_16x16x4f64 under the FP64 one, each with one FP64 refinement step).

  * a grid of at most 4096 cells is its own coarsest level: the preconditioner is then
    the dense solve alone, so flexible CG must converge in two iterations to the
    oracle's tight CG solution — this checks the inverse built on the matrix cores
    entry by entry (through its action) and the refinement;
  * on a multi-level hierarchy the dense coarse solve gives the oracle's solution and
    the SIMPLE trajectory of the oracle;
  * a stale inverse (rebuilt every 4 hierarchy builds) still meets the stop rule: the
    FP64 refinement corrects against the current coarse operator.
"""
import numpy as np
import pytest

from hsolver_amd import abi, configs

from helpers import gpu_ctx, oracle_ctx, rel_l2

pytestmark = pytest.mark.gpu


def first_pressure_system(hs_lib, oracle_lib, case, gp):
    g, o = gpu_ctx(hs_lib, case, gp), oracle_ctx(oracle_lib, case)
    g.init(); o.init()
    for comp in range(3):
        g.momentum(comp); o.momentum(comp)
    g.pressure_correction(); o.pressure_correction()
    b = o.get("b")
    op = abi.Params.from_buffer_copy(case.params)
    op.p_tol = 1e-13
    op.p_maxit = 20000
    o2 = oracle_ctx(oracle_lib, case, op)
    o2.init()
    for comp in range(3):
        o2.momentum(comp)
    o2.pressure_correction()
    xo, _ = o2.pressure_solve(b)
    return g, b, xo


@pytest.mark.parametrize("precision", [abi.HS_PREC_MIXED, abi.HS_PREC_FP64], ids=["mixed", "fp64"])
def test_dense_only_level_is_an_exact_preconditioner(hs_lib, oracle_lib, precision):
    case = configs.c3(16)                      # 16^3 = 4096 cells: one level, solved densely
    case.params.order = abi.HS_ORDER_RB
    gp = abi.Params.from_buffer_copy(case.params)
    gp.precond = abi.HS_PRECOND_GMG
    gp.precision = precision
    gp.mg_dense_coarse = 1
    gp.p_tol = 1e-10
    print(f"[dense] c3 16^3 single level, precision {precision}", flush=True)
    g, b, xo = first_pressure_system(hs_lib, oracle_lib, case, gp)
    xg, sg = g.pressure_solve(b)
    # the FP64 inverse is exact to round-off: one FCG iteration; the FP32 one with its FP64
    # refinement step: at most two
    assert sg["p_converged"] == 1 and sg["p_iters"] <= (2 if precision == abi.HS_PREC_MIXED else 1), sg
    assert sg["p_true_relres"] < 2e-10
    assert rel_l2(xg, xo) < 1e-8


@pytest.mark.parametrize("precision", [abi.HS_PREC_MIXED, abi.HS_PREC_FP64], ids=["mixed", "fp64"])
def test_dense_coarse_multilevel_solution(hs_lib, oracle_lib, precision):
    case = configs.c3(64)                      # levels 64, 32, 16 (dense)
    case.params.order = abi.HS_ORDER_RB
    gp = abi.Params.from_buffer_copy(case.params)
    gp.precond = abi.HS_PRECOND_GMG
    gp.precision = precision
    gp.mg_dense_coarse = 1
    gp.p_tol = 1e-10
    print(f"[dense] c3 64^3 three levels, precision {precision}", flush=True)
    g, b, xo = first_pressure_system(hs_lib, oracle_lib, case, gp)
    xg, sg = g.pressure_solve(b)
    assert sg["p_converged"] == 1 and sg["p_true_relres"] < 2e-10, sg
    assert sg["p_iters"] < 20, sg
    assert rel_l2(xg, xo) < 1e-7


@pytest.mark.parametrize("every", [1, 4])
def test_dense_coarse_simple_trajectory(hs_lib, oracle_lib, every):
    case = configs.c2(scale=4)                 # 64x16x16: levels 64x16x16, 32x8x8 (dense)
    case.params.order = abi.HS_ORDER_RB
    case.params.p_tol = 1e-10
    gp = abi.Params.from_buffer_copy(case.params)
    gp.precond = abi.HS_PRECOND_GMG
    gp.precision = abi.HS_PREC_MIXED
    gp.mg_dense_coarse = every
    print(f"[dense] c2/4 trajectory, inverse rebuilt every {every} builds", flush=True)
    g, o = gpu_ctx(hs_lib, case, gp), oracle_ctx(oracle_lib, case)
    g.init(); o.init()
    for it in range(10):
        sg = g.step(); o.step()
        assert sg["p_true_relres"] < 1e-9, (it, sg)
    fg, fo = g.fields(), o.fields()
    for k in fg:
        assert rel_l2(fg[k], fo[k]) < 1e-6, k
