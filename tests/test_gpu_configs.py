"""The benchmark configurations end to end on the GPU against the oracle.

  * c1 (64x16x16 channel, solid cube, Re 100, k ratio 100, q = 1): 200 SIMPLE
    iterations with the reference's solver settings (lexicographic sweeps,
    unpreconditioned CG at 1e-8), GPU vs the oracle's CPU path; every field within
    1e-6 relative L2, mass and energy balance residuals reported;
  * the same 200 iterations with the fast path (red-black sweeps, multigrid-preconditioned
    FCG, FP64 or mixed V-cycle) against the oracle's red-black sweeps with its
    unpreconditioned CG: the two Krylov methods meet the same 1e-8 stop rule, so the
    trajectories agree to 1e-6.
"""
import numpy as np
import pytest

from hsolver_amd import abi, configs

from helpers import gpu_ctx, oracle_ctx, rel_l2

pytestmark = pytest.mark.gpu

ITERS = 200


def oracle_run(oracle_lib, order):
    case = configs.c1()
    case.params.order = order
    o = oracle_ctx(oracle_lib, case)
    o.init()
    hist = [o.step() for _ in range(ITERS)]
    f = o.fields()
    o.close()
    return configs.c1(), f, hist


@pytest.fixture(scope="module")
def c1_oracle(oracle_lib):
    return oracle_run(oracle_lib, abi.HS_ORDER_LEX)


@pytest.fixture(scope="module")
def c1_oracle_rb(oracle_lib):
    return oracle_run(oracle_lib, abi.HS_ORDER_RB)


def run_gpu(hs_lib, case, order, precond, precision=abi.HS_PREC_FP64):
    p = abi.Params.from_buffer_copy(case.params)
    p.order, p.precond, p.precision = order, precond, precision
    g = gpu_ctx(hs_lib, case, p)
    g.init()
    hist = [g.step() for _ in range(ITERS)]
    f = g.fields()
    g.close()
    return f, hist


def test_c1_reference_settings_200_iterations(hs_lib, c1_oracle):
    case, fo, ho = c1_oracle
    fg, hg = run_gpu(hs_lib, case, abi.HS_ORDER_LEX, abi.HS_PRECOND_NONE)
    errs = {k: rel_l2(fg[k], fo[k]) for k in fg}
    print(f"[c1] 200 iterations, LEX + CG: rel-L2 {errs}; mass imbalance GPU {hg[-1]['mass_imbalance']:.3e} "
          f"oracle {ho[-1]['mass_imbalance']:.3e}; energy residual {hg[-1]['energy_residual']:.3e}", flush=True)
    assert all(e < 1e-6 for e in errs.values()), errs
    assert hg[-1]["mass_imbalance"] == pytest.approx(ho[-1]["mass_imbalance"], rel=1e-3, abs=1e-12)


@pytest.mark.parametrize("precision", [abi.HS_PREC_FP64, abi.HS_PREC_MIXED], ids=["fp64", "mixed"])
def test_c1_fast_path_200_iterations(hs_lib, c1_oracle_rb, precision):
    case, fo, ho = c1_oracle_rb
    fg, hg = run_gpu(hs_lib, case, abi.HS_ORDER_RB, abi.HS_PRECOND_GMG, precision)
    errs = {k: rel_l2(fg[k], fo[k]) for k in fg}
    print(f"[c1] 200 iterations, RB + GMG ({precision}): rel-L2 vs oracle RB + CG {errs}", flush=True)
    assert all(e < 1e-6 for e in errs.values()), errs
