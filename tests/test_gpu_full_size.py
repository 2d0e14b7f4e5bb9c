"""The headline workload at its full size (256^3 unit box, 27 solid blocks, Re 1000, k ratio
1000, p' to 1e-8) on the GPU, checked through properties that do not need the CPU oracle
(which takes about 80 s per SIMPLE iteration at this size):

  * every p' solve of the mixed and the FP64 multigrid paths converges, and the true
    residual ||b - A p'|| / ||b||, recomputed in FP64 after each solve, meets the 1e-8
    stop rule;
  * after the same SIMPLE iterations the mixed path's fields are within 1e-6 relative L2
    of the FP64 path's, and the mass imbalance agrees to 1e-6;
  * the fused smoother sweep and two in-place colour passes give the same bits over a
    whole SIMPLE iteration;
  * the FP64 drop-in (the reference's unpreconditioned CG) solves the first p' system to
    the same true-residual bar.
"""
import numpy as np
import pytest

from hsolver_amd import abi, configs

from helpers import gpu_ctx, rel_l2

pytestmark = pytest.mark.gpu

STEPS = 4


def run(hs_lib, precision, precond=abi.HS_PRECOND_GMG, rows=0, steps=STEPS):
    case = configs.c3(256)
    p = abi.Params.from_buffer_copy(case.params)
    p.order, p.precond, p.precision = abi.HS_ORDER_RB, precond, precision
    p.mg_dense_coarse = 50 if precision == abi.HS_PREC_MIXED else 0
    p.mg_sweep_rows = rows
    g = gpu_ctx(hs_lib, case, p)
    g.init()
    hist = [g.step() for _ in range(steps)]
    f = g.fields()
    g.close()
    return f, hist


@pytest.fixture(scope="module")
def fp64_run(hs_lib):
    return run(hs_lib, abi.HS_PREC_FP64)


def test_full_size_mixed_against_fp64(hs_lib, fp64_run):
    f64, h64 = fp64_run
    fm, hm = run(hs_lib, abi.HS_PREC_MIXED)
    for h in h64 + hm:
        assert h["p_converged"] == 1 and h["p_true_relres"] <= 1e-8, h
    errs = {k: rel_l2(fm[k], f64[k]) for k in f64}
    print(f"[256^3] {STEPS} SIMPLE iterations, FCG iterations mixed {[h['p_iters'] for h in hm]} "
          f"FP64 {[h['p_iters'] for h in h64]}; mixed vs FP64 rel-L2 {errs}", flush=True)
    assert max(errs.values()) < 1e-6, errs
    assert hm[-1]["mass_imbalance"] == pytest.approx(h64[-1]["mass_imbalance"], rel=1e-6)
    for f in fm.values():
        assert np.all(np.isfinite(f))


def test_full_size_fused_sweep_bit_equal_to_colour_passes(hs_lib):
    fa, ha = run(hs_lib, abi.HS_PREC_MIXED, rows=0, steps=1)
    fb, hb = run(hs_lib, abi.HS_PREC_MIXED, rows=-1, steps=1)
    assert ha[0]["p_iters"] == hb[0]["p_iters"]
    for k in fa:
        assert np.array_equal(fa[k], fb[k]), k


def test_full_size_dropin_cg_first_step(hs_lib):
    _, h = run(hs_lib, abi.HS_PREC_FP64, precond=abi.HS_PRECOND_NONE, steps=1)
    print(f"[256^3] drop-in CG: {h[0]['p_iters']} iterations, true relres {h[0]['p_true_relres']:.3e}", flush=True)
    assert h[0]["p_converged"] == 1 and h[0]["p_true_relres"] <= 1e-8


def test_c5_full_size_adjoint_against_central_differences(hs_lib):
    """The 512x256x256 stretched channel (33.5M cells, aspect 50, 64 heated fins): 8 mixed
    SIMPLE iterations converge every p' solve to the 1e-8 true-residual bar; on the frozen
    flow the energy solve and the adjoint of the fins' temperature p-norm (p = 8) converge,
    and dJ/dq at two seeded fin cells agrees with a central difference to 1e-5."""
    from hsolver_amd import core
    from hsolver_amd.adjoint import central_difference, cell_volumes
    case = configs.c5()
    nx, ny, nz = case.shape
    p = abi.Params.from_buffer_copy(case.params)
    p.order, p.precond, p.precision = abi.HS_ORDER_RB, abi.HS_PRECOND_GMG, abi.HS_PREC_MIXED
    p.mg_dense_coarse = 50
    ctx = core.create(case.grid, p, case.q)
    ctx.init()
    for it in range(8):
        st = ctx.step()
        assert st["p_converged"] == 1 and st["p_true_relres"] <= 1e-8, (it, st)
    shape = (nz, ny, nx)
    solid = case.grid.cell_type.reshape(shape) == 1
    V = cell_volumes(case)
    fwd = ctx.energy_solve(1e-11, 2000)
    assert fwd["converged"] == 1 and fwd["true_relres"] < 1e-9, fwd
    T = ctx.cells("tf")
    S = float(np.sum(np.where(solid, T ** 8 * V, 0.0)))
    J = S ** 0.125
    dJdT = np.where(solid, J ** -7 * T ** 7 * V, 0.0)
    dq, dk, adj = ctx.adjoint_energy(dJdT, 1e-11, 2000)
    assert adj["converged"] == 1 and adj["true_relres"] < 1e-9, adj
    q0, k0 = ctx.cells("q").copy(), ctx.cells("k").copy()
    idx = np.flatnonzero(solid.ravel())
    pick = idx[(configs.splitmix64(77, 2) * idx.size).astype(np.int64)]
    for c in pick:
        cell = np.unravel_index(c, shape)
        fd = central_difference(ctx, "q", cell, 1e-6 * q0[cell], q0, k0, solid, V, tol=1e-11, maxit=2000)
        err = abs(fd - dq[cell]) / abs(dq[cell])
        print(f"[c5 full] fin cell {tuple(int(v) for v in cell)}: central difference {fd:.9e}, adjoint {dq[cell]:.9e}, "
              f"rel err {err:.2e}", flush=True)
        assert err < 1e-5
    ctx.close()
