"""The benchmark configurations end to end on the GPU against the oracle.

  * c1 (64x16x16 channel, solid cube, Re 100, k ratio 100, q = 1): 200 SIMPLE
    iterations with the reference's solver settings (lexicographic sweeps,
    unpreconditioned CG at 1e-8), GPU vs the oracle's CPU path; every field within
    1e-6 relative L2, mass and energy balance residuals reported;
  * the c5 geometry (stretched grid, plate fins) at 64x32x32: SIMPLE, energy solve and
    adjoint against the oracle;
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


def test_c5_geometry_reduced(hs_lib, oracle_lib):
    """The c5 geometry at 64x32x32 (tanh-stretched y/z at cell aspect ratio 50, plate fins):
    10 SIMPLE iterations of the mixed fast path against the oracle's red-black + CG run,
    then the energy solve and the fins' p-norm adjoint against the oracle's FP64 ones."""
    from helpers import cell_volumes, solid_pnorm
    case = configs.c5(64, 32, 32, nfins=8)
    case.params.order = abi.HS_ORDER_RB
    case.params.p_tol = 1e-10
    gp = abi.Params.from_buffer_copy(case.params)
    gp.precond, gp.precision, gp.mg_dense_coarse = abi.HS_PRECOND_GMG, abi.HS_PREC_MIXED, 1
    g, o = gpu_ctx(hs_lib, case, gp), oracle_ctx(oracle_lib, case)
    g.init(); o.init()
    for it in range(10):
        sg = g.step(); o.step()
        assert sg["p_true_relres"] < 1e-9, (it, sg)
    fg, fo = g.fields(), o.fields()
    errs = {k: rel_l2(fg[k], fo[k]) for k in fg}
    print(f"[c5 reduced] 10 iterations, rel-L2 vs oracle {errs}", flush=True)
    assert all(e < 1e-6 for e in errs.values()), errs
    o.load_state(g.state())
    ig = g.energy_solve(1e-12, 3000)
    io = o.energy_solve(1e-12, 20000)
    assert ig["converged"] == 1 and io["converged"] == 1, (ig, io)
    assert rel_l2(g.cells("tf"), o.cells("tf")) < 1e-9
    nx, ny, nz = case.shape
    solid = case.grid.cell_type.reshape(nz, ny, nx) == 1
    _, dJdT = solid_pnorm(o.cells("tf"), solid, cell_volumes(case))
    dq_g, dk_g, _ = g.adjoint_energy(dJdT, 1e-12, 3000)
    dq_o, dk_o, _ = o.adjoint_energy(dJdT, 1e-12, 20000)
    assert rel_l2(dq_g, dq_o) < 1e-8 and rel_l2(dk_g, dk_o) < 1e-7


@pytest.mark.timeout(600)   # the oracle's share of this test is CPU work of 30-60 s
def test_c2_full_size_fp64_dropin_vs_oracle(hs_lib, oracle_lib):
    """c2 at its full 256x64x64 size: 10 SIMPLE iterations of the FP64 drop-in (red-black
    sweeps + the reference's unpreconditioned CG, 1e-8) on the GPU and in the oracle, from
    the same start; CG iteration counts per step within one, fields within 1e-8 (the
    Krylov round-off at this tolerance is the only difference).  The multigrid paths (FP64
    and mixed precision, flexible CG to the same 1e-8) against the same oracle fields: within
    1e-6 (they stop at a different point inside the tolerance)."""
    case = configs.c2()
    case.params.order = abi.HS_ORDER_RB
    case.params.p_tol = 1e-8
    g, o = gpu_ctx(hs_lib, case), oracle_ctx(oracle_lib, case)
    g.init(); o.init()
    for it in range(10):
        sg, so = g.step(), o.step()
        print(f"[c2 full] step {it}: CG iterations GPU {sg['p_iters']} oracle {so['p_iters']}", flush=True)
        assert abs(sg["p_iters"] - so["p_iters"]) <= 1, (it, sg["p_iters"], so["p_iters"])
    fg, fo = g.fields(), o.fields()
    errs = {k: rel_l2(fg[k], fo[k]) for k in fg}
    print(f"[c2 full] rel-L2 vs oracle after 10 iterations {errs}", flush=True)
    assert all(e < 1e-8 for e in errs.values()), errs
    g.close()
    for prec, name in ((abi.HS_PREC_FP64, "fp64"), (abi.HS_PREC_MIXED, "mixed")):
        p = abi.Params.from_buffer_copy(case.params)
        p.precond, p.precision = abi.HS_PRECOND_GMG, prec
        p.mg_dense_coarse = 50
        h = gpu_ctx(hs_lib, case, p)
        h.init()
        hist = [h.step() for _ in range(10)]
        assert all(st["p_converged"] == 1 and st["p_true_relres"] < 1e-8 for st in hist), hist
        fh = h.fields()
        errs = {k: rel_l2(fh[k], fo[k]) for k in fh}
        print(f"[c2 full] multigrid {name}: FCG iterations {[st['p_iters'] for st in hist]}, rel-L2 vs oracle {errs}",
              flush=True)
        assert all(e < 1e-6 for e in errs.values()), (name, errs)
        h.close()
    o.close()
