"""Aggregation multigrid in the preconditioner slot (HS_PRECOND_AMG for the p' solve,
e_precond = HS_PRECOND_AMG for the energy solves) against the oracle.

  * p' solve on the wall-stretched finned channel (the c5 geometry at 64x32x32, cell
    aspect ratio 50) and on the c3 geometry (64^3, 27 blocks): the oracle's tight CG
    solution, FP64 and mixed hierarchies;
  * a 10-iteration SIMPLE trajectory with the AMG p' solve against the oracle, with the
    kept aggregates refilled on most steps and aggregated again every 4 builds;
  * energy solve and discrete adjoint on a frozen flow of the finned channel against the
    oracle's FP64 solution and transpose solve;
  * the hierarchy is deterministic: two contexts give bit-identical solves.
"""
import numpy as np
import pytest

from hsolver_amd import abi, configs

from helpers import cell_volumes, energy_balance, gpu_ctx, oracle_ctx, rel_l2, solid_pnorm

pytestmark = pytest.mark.gpu

PRECS = [abi.HS_PREC_FP64, abi.HS_PREC_MIXED]
PIDS = ["fp64", "mixed"]


def amg_params(case, precision, dense_every=50):
    gp = abi.Params.from_buffer_copy(case.params)
    gp.order = abi.HS_ORDER_RB
    gp.precond = abi.HS_PRECOND_AMG
    gp.precision = precision
    gp.mg_dense_coarse = dense_every
    return gp


def first_pressure_system(hs_lib, oracle_lib, case, gp):
    """GPU context after the first momentum sweeps and p' assembly, the oracle's right-hand
    side of that system and its solution by the oracle's CG to 1e-13."""
    g, o = gpu_ctx(hs_lib, case, gp), oracle_ctx(oracle_lib, case)
    g.init(); o.init()
    for comp in range(3):
        g.momentum(comp); o.momentum(comp)
    g.pressure_correction(); o.pressure_correction()
    b = o.get("b")
    op = abi.Params.from_buffer_copy(case.params)
    op.p_tol, op.p_maxit = 1e-13, 100000
    o2 = oracle_ctx(oracle_lib, case, op)
    o2.init()
    for comp in range(3):
        o2.momentum(comp)
    o2.pressure_correction()
    xo, _ = o2.pressure_solve(b)
    return g, b, xo


@pytest.mark.parametrize("precision", PRECS, ids=PIDS)
@pytest.mark.parametrize("geometry", ["c5_64", "c3_64"])
def test_amg_pressure_solution(hs_lib, oracle_lib, geometry, precision):
    if geometry == "c5_64":
        case = configs.c5(64, 32, 32)
    else:
        case = configs.c3(64)
    case.params.order = abi.HS_ORDER_RB
    gp = amg_params(case, precision)
    gp.p_tol = 1e-10
    print(f"[amg] p' solve {geometry} precision {precision}", flush=True)
    g, b, xo = first_pressure_system(hs_lib, oracle_lib, case, gp)
    xg, sg = g.pressure_solve(b)
    print(f"[amg] {geometry} {precision}: {sg['p_iters']} FCG iterations", flush=True)
    assert sg["p_converged"] == 1 and sg["p_true_relres"] < 2e-10, sg
    assert sg["p_iters"] < 60, sg
    assert rel_l2(xg, xo) < 1e-7


@pytest.mark.parametrize("precision", PRECS, ids=PIDS)
def test_amg_simple_trajectory_matches_oracle(hs_lib, oracle_lib, precision):
    case = configs.c5(64, 32, 32)
    case.params.order = abi.HS_ORDER_RB
    case.params.p_tol = 1e-10
    case.params.p_maxit = 100000
    gp = amg_params(case, precision, dense_every=4)
    print(f"[amg] c5 64x32x32 trajectory precision {precision}", flush=True)
    g, o = gpu_ctx(hs_lib, case, gp), oracle_ctx(oracle_lib, case)
    g.init(); o.init()
    for it in range(10):
        sg = g.step(); o.step()
        assert sg["p_converged"] == 1 and sg["p_true_relres"] < 1e-9, (it, sg)
    fg, fo = g.fields(), o.fields()
    for k in fg:
        assert rel_l2(fg[k], fo[k]) < 1e-6, k


def test_amg_pressure_deterministic(hs_lib, oracle_lib):
    case = configs.c5(64, 32, 32)
    case.params.order = abi.HS_ORDER_RB
    gp = amg_params(case, abi.HS_PREC_MIXED)
    gp.p_tol = 1e-10
    g1, b, _ = first_pressure_system(hs_lib, oracle_lib, case, gp)
    g2 = gpu_ctx(hs_lib, case, gp)
    g2.init()
    for comp in range(3):
        g2.momentum(comp)
    g2.pressure_correction()
    x1, s1 = g1.pressure_solve(b)
    x2, s2 = g2.pressure_solve(b)
    assert s1["p_iters"] == s2["p_iters"]
    assert np.array_equal(x1, x2)


@pytest.fixture(scope="module")
def frozen(hs_lib, oracle_lib):
    """Finned channel 64x32x32 after 20 SIMPLE iterations (GPU, geometric multigrid),
    state copied into the oracle."""
    case = configs.c5(64, 32, 32)
    p = abi.Params.from_buffer_copy(case.params)
    p.order = abi.HS_ORDER_RB
    p.precond = abi.HS_PRECOND_GMG
    p.precision = abi.HS_PREC_FP64
    p.mg_dense_coarse = 50
    print("[amg] frozen flow c5 64x32x32", flush=True)
    g = gpu_ctx(hs_lib, case, p)
    g.init()
    for _ in range(20):
        g.step()
    st = g.state()
    g.close()
    o = oracle_ctx(oracle_lib, case)
    o.init()
    o.load_state(st)
    yield case, st, o
    o.close()


def energy_ctx(hs_lib, case, st, precision):
    p = abi.Params.from_buffer_copy(case.params)
    p.order = abi.HS_ORDER_RB
    p.precond = abi.HS_PRECOND_GMG
    p.precision = precision
    p.e_precond = abi.HS_PRECOND_AMG
    g = gpu_ctx(hs_lib, case, p)
    g.init()
    g.load_state(st)
    return g


@pytest.mark.parametrize("precision", PRECS, ids=PIDS)
def test_amg_energy_solve_matches_oracle(hs_lib, frozen, precision):
    case, st, o = frozen
    g = energy_ctx(hs_lib, case, st, precision)
    ig = g.energy_solve(tol=1e-12, maxit=400)
    io = o.energy_solve(tol=1e-12, maxit=20000)
    print(f"[amg] energy {precision}: {ig['iters']} BiCGStab iterations", flush=True)
    assert ig["converged"] == 1 and ig["true_relres"] < 1e-11, ig
    assert io["converged"] == 1, io
    assert rel_l2(g.cells("tf"), o.cells("tf")) < 1e-9
    g.close()


@pytest.mark.parametrize("precision", PRECS, ids=PIDS)
def test_amg_adjoint_matches_oracle(hs_lib, frozen, precision):
    case, st, o = frozen
    g = energy_ctx(hs_lib, case, st, precision)
    g.energy_solve(tol=1e-13, maxit=400)
    o.energy_solve(tol=1e-13, maxit=20000)
    nx, ny, nz = case.shape
    solid = np.asarray(case.grid.cell_type).reshape(nz, ny, nx) == 1
    T = g.cells("tf")
    _, dJdT = solid_pnorm(T, solid, cell_volumes(case))
    dq, dk, info = g.adjoint_energy(dJdT, tol=1e-13, maxit=400)
    dqo, dko, info_o = o.adjoint_energy(dJdT, tol=1e-13, maxit=20000)
    print(f"[amg] adjoint {precision}: {info['iters']} BiCGStab iterations", flush=True)
    assert info["converged"] == 1 and info_o["converged"] == 1
    assert rel_l2(dq, dqo) < 1e-9
    assert rel_l2(dk, dko) < 1e-8
    g.close()


def test_energy_balance_after_amg_solve(hs_lib, frozen):
    """hs_energy_balance on the frozen flow's temperature: the global balance that the SIMPLE
    sweeps leave open closes after the energy solve, and the heat input and outflow agree
    with the physics sum built in numpy (heat in = sum q V; out = outflow enthalpy and
    Dirichlet-face conduction, with the residual mass imbalance's share)."""
    case, st, o = frozen
    g = energy_ctx(hs_lib, case, st, abi.HS_PREC_MIXED)
    h0, r0 = g.energy_balance()
    info = g.energy_solve(tol=1e-12, maxit=400)
    h1, r1 = g.energy_balance()
    q_in, q_out, q_mass = energy_balance(case, g.cells("tf"), g.get("u"), g.get("v"), g.get("w"), g.cells("k"),
                                         g.cells("q"))
    print(f"[amg] energy balance before {r0:.3e}, after the solve {r1:.3e} ({info['iters']} iterations); "
          f"numpy: in {q_in:.12e} out {q_out:.12e} mass share {q_mass:.3e}", flush=True)
    assert info["converged"] == 1
    assert h0 == h1 and abs(h1 - q_in) < 1e-12 * q_in
    assert r0 > 1e-3 and r1 < 1e-8
    assert abs(q_in - q_out + q_mass) < 1e-7 * q_in
    g.close()
