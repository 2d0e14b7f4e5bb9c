"""GPU parity: libhsolver_amd vs the CPU oracle on identical inputs.

Per-phase checks start both sides from the same state and require bit-identical
results for every Gauss-Seidel / Jacobi point update (both sides evaluate each
update in the reference's operation order, -ffp-contract=off); Krylov solves, whose
dot products are summed in a different order, are compared to 1e-10 relative.
"""
import numpy as np
import pytest

from hsolver_amd import abi, configs

from helpers import gpu_ctx, oracle_ctx, rel_l2, small_ref_case

pytestmark = pytest.mark.gpu

NAMES = ("u", "v", "w", "up", "vp", "wp", "p", "ptemp", "pprime", "tf", "ts")


def assert_same(a_ctx, b_ctx, names=NAMES, what=""):
    for n in names:
        a, b = a_ctx.get(n), b_ctx.get(n)
        if not np.array_equal(a, b):
            diff = np.abs(a - b)
            idx = np.unravel_index(np.argmax(diff), diff.shape)
            raise AssertionError(f"{what}: array {n} differs, max |d|={diff.max():.3e} at {idx} "
                                 f"(gpu {a[idx]!r} oracle {b[idx]!r})")


def sync(dst, src, names=NAMES):
    for n in names:
        dst.put(n, src.get(n))


@pytest.mark.parametrize("order", [abi.HS_ORDER_LEX, abi.HS_ORDER_RB])
def test_mode_r_phases_bitexact(hs_lib, oracle_lib, order):
    case = small_ref_case(order=order)
    g, o = gpu_ctx(hs_lib, case), oracle_ctx(oracle_lib, case)
    g.init(); o.init()
    assert_same(g, o, what="init")
    for it in range(3):
        for comp in range(3):
            sg, so = g.momentum(comp), o.momentum(comp)
            assert sg[0] == so[0], (it, comp, sg, so)
            assert_same(g, o, what=f"iter {it} momentum {comp}")
        stg, sto = g.pressure_correction(), o.pressure_correction()
        assert abs(stg["p_iters"] - sto["p_iters"]) <= 1, (stg, sto)
        assert rel_l2(g.get("pprime"), o.get("pprime")) < 1e-10
        assert rel_l2(g.get("b"), o.get("b")) == 0.0
        sync(g, o, ("pprime",))
        g.correct(); o.correct()
        assert_same(g, o, what=f"iter {it} correct")
        sg, so = g.fluid_energy(), o.fluid_energy()
        assert sg[0] == so[0]
        assert_same(g, o, what=f"iter {it} energy")
        sg, so = g.solid_solve(), o.solid_solve()
        assert sg["ts_sweeps"] == so["ts_sweeps"]
        assert_same(g, o, what=f"iter {it} solid")
        g.exchange(); o.exchange()
        assert_same(g, o, what=f"iter {it} exchange")
        g.set_for_next(); o.set_for_next()
        assert_same(g, o, what=f"iter {it} set_for_next")


@pytest.mark.parametrize("order", [abi.HS_ORDER_LEX, abi.HS_ORDER_RB])
def test_mode_r_steps(hs_lib, oracle_lib, order):
    case = small_ref_case(nx=16, ny=14, nz=12, block=(6, 9, 5, 8, 4, 7), order=order)
    g, o = gpu_ctx(hs_lib, case), oracle_ctx(oracle_lib, case)
    g.init(); o.init()
    for it in range(6):
        sg, so = g.step(), o.step()
        assert sg["mom_sweeps"] == so["mom_sweeps"] and sg["t_sweeps"] == so["t_sweeps"]
    fg, fo = g.fields(), o.fields()
    for k in fg:
        assert rel_l2(fg[k], fo[k]) < 1e-9, k


def test_cell_bc_matches_literal_loop(hs_lib, oracle_lib):
    case = small_ref_case(nx=9, ny=7, nz=6, block=None)
    g, o = gpu_ctx(hs_lib, case), oracle_ctx(oracle_lib, case)
    rng = np.random.default_rng(7)
    for trial in range(3):
        p = rng.standard_normal(9 * 7 * 6)
        pt = rng.standard_normal(g.size(abi.HS_A_PTEMP))
        for c in (g, o):
            c.put("p", p); c.put("ptemp", pt)
            c.cell_bc()
        assert_same(g, o, names=("p", "ptemp"), what=f"cell_bc trial {trial}")


@pytest.mark.parametrize("order", [abi.HS_ORDER_RB, abi.HS_ORDER_LEX])
def test_mode_c_steps(hs_lib, oracle_lib, order):
    case = configs.c2(scale=8)          # 32x8x8 channel with the heated cube
    case.params.order = order
    case.params.p_tol = 1e-12           # Krylov round-off far below the comparison bar
    g, o = gpu_ctx(hs_lib, case), oracle_ctx(oracle_lib, case)
    g.init(); o.init()
    assert_same(g, o, what="init")
    for it in range(5):
        sg, so = g.step(), o.step()
        assert abs(sg["p_iters"] - so["p_iters"]) <= 1
    fg, fo = g.fields(), o.fields()
    for k in fg:
        assert rel_l2(fg[k], fo[k]) < 1e-9, k


@pytest.mark.parametrize("name", ["c2_32x8x8", "c3_32", "c5_64x32x32"])
def test_energy_balance_matches_oracle(hs_lib, oracle_lib, name):
    """hs_energy_balance (a telescoped sum: each cell's -T m_c plus its boundary faces'
    terms) against the oracle's direct sum of every cell row's residual, on one state:
    after 3 SIMPLE steps (the outflow ghost set by the end of the step), and after the
    sweeps, correction and energy sweep of a fourth (the ghost one sweep old and the
    corrected flow not yet divergence-free to round-off).  Channel with inlet, outflow
    and walls; lid-driven box with inlet; wall-stretched finned channel."""
    case = {"c2_32x8x8": lambda: configs.c2(scale=8), "c3_32": lambda: configs.c3(32),
            "c5_64x32x32": lambda: configs.c5(64, 32, 32)}[name]()
    case.params.order = abi.HS_ORDER_RB
    g = gpu_ctx(hs_lib, case)
    g.init()
    for _ in range(3):
        g.step()
    o = oracle_ctx(oracle_lib, case)
    o.init()
    for phase in ("end of step", "after the energy sweep"):
        if phase != "end of step":
            for comp in range(3):
                g.momentum(comp)
            g.pressure_correction()
            g.correct()
            g.fluid_energy()
        o.load_state(g.state())
        hg, rg = g.energy_balance()
        ho, ro = o.energy_balance()
        print(f"[balance] {name} {phase}: heat {hg:.12e} / {ho:.12e}, residual {rg:.12e} / {ro:.12e}", flush=True)
        assert abs(hg - ho) <= 1e-13 * abs(ho)
        assert ro > 1e-6
        assert abs(rg - ro) <= 1e-10 * max(ro, 1.0)
    g.close()
    o.close()


def test_mode_c_phases_bitexact(hs_lib, oracle_lib):
    case = configs.c2(scale=8)
    case.params.order = abi.HS_ORDER_RB
    g, o = gpu_ctx(hs_lib, case), oracle_ctx(oracle_lib, case)
    g.init(); o.init()
    for it in range(2):
        for comp in range(3):
            g.momentum(comp); o.momentum(comp)
            assert_same(g, o, what=f"iter {it} momentum {comp}")
        g.pressure_correction(); o.pressure_correction()
        assert rel_l2(g.get("pprime"), o.get("pprime")) < 1e-10
        sync(g, o, ("pprime",))
        g.correct(); o.correct()
        assert_same(g, o, what=f"iter {it} correct")
        g.fluid_energy(); o.fluid_energy()
        assert_same(g, o, what=f"iter {it} energy")
        g.set_for_next(); o.set_for_next()
        assert_same(g, o, what=f"iter {it} set_for_next")


@pytest.mark.parametrize("precond", [abi.HS_PRECOND_NONE, abi.HS_PRECOND_GMG])
def test_pprime_array_between_solve_and_correction(hs_lib, oracle_lib, precond):
    """The p' array between the pressure solve and the correction reads back as the solve's
    solution (the oracle's to 1e-9, both solved to 1e-12); a pressure solve of another right-hand side in
    between changes neither it nor what the correction does; after the correction it
    reads back as zero.  (A p' written through hs_put_array before the correction is the
    one the correction uses: test_mode_c_phases_bitexact.)"""
    case = configs.c2(scale=8)
    case.params.order = abi.HS_ORDER_RB
    case.params.p_tol = 1e-12           # two solvers' solutions agree far below the bar
    p = abi.Params.from_buffer_copy(case.params)
    p.precond = precond
    g1, g2, o = gpu_ctx(hs_lib, case, p), gpu_ctx(hs_lib, case, p), oracle_ctx(oracle_lib, case)
    for c in (g1, g2, o):
        c.init()
        for comp in range(3):
            c.momentum(comp)
        c.pressure_correction()
    x1 = g1.get("pprime")
    assert rel_l2(x1, o.get("pprime")) < 1e-9
    b = np.random.default_rng(11).standard_normal(x1.shape)
    xs, _ = g2.pressure_solve(b)                 # another system through the solver's buffers
    assert rel_l2(xs, x1) > 1e-3
    assert np.array_equal(g2.get("pprime"), x1)
    g1.correct(); g2.correct()
    assert_same(g1, g2, what="correction after another solve")
    assert not np.any(g1.get("pprime")) and not np.any(g2.get("pprime"))
    for c in (g1, g2, o):
        c.close()


def test_pressure_operator_and_cg(hs_lib, oracle_lib):
    case = configs.c2(scale=4)          # 64x16x16
    case.params.order = abi.HS_ORDER_RB
    g, o = gpu_ctx(hs_lib, case), oracle_ctx(oracle_lib, case)
    g.init(); o.init()
    for comp in range(3):
        g.momentum(comp); o.momentum(comp)
    g.pressure_correction(); o.pressure_correction()
    rng = np.random.default_rng(3)
    x = rng.standard_normal(g.size(abi.HS_A_B))
    yg, yo = g.pressure_apply(x), o.pressure_apply(x)
    assert np.array_equal(yg, yo)
    b = o.get("b")
    xg, sg = g.pressure_solve(b)
    xo, so = o.pressure_solve(b)
    assert abs(sg["p_iters"] - so["p_iters"]) <= 1
    assert sg["p_true_relres"] < 1.1 * case.params.p_tol
    assert rel_l2(xg, xo) < 1e-9


@pytest.mark.parametrize("precision", [abi.HS_PREC_FP64, abi.HS_PREC_MIXED])
def test_gmg_fcg_solution(hs_lib, oracle_lib, precision):
    case = configs.c2(scale=4)          # 64x16x16 channel
    case.params.order = abi.HS_ORDER_RB
    gp = abi.Params.from_buffer_copy(case.params)
    gp.precond = abi.HS_PRECOND_GMG
    gp.precision = precision
    gp.p_tol = 1e-10
    g, o = gpu_ctx(hs_lib, case, gp), oracle_ctx(oracle_lib, case)
    g.init(); o.init()
    for comp in range(3):
        g.momentum(comp); o.momentum(comp)
    g.pressure_correction(); o.pressure_correction()
    b = o.get("b")
    xg, sg = g.pressure_solve(b)
    op = abi.Params.from_buffer_copy(case.params)
    op.p_tol = 1e-13
    op.p_maxit = 20000
    o2 = oracle_ctx(oracle_lib, case, op)
    o2.init()
    for comp in range(3):
        o2.momentum(comp)
    o2.pressure_correction()
    xo, so = o2.pressure_solve(b)
    assert sg["p_converged"] == 1
    assert sg["p_true_relres"] < 2e-10
    assert sg["p_iters"] < 40, sg
    assert rel_l2(xg, xo) < 1e-7


@pytest.mark.parametrize("precision", [abi.HS_PREC_FP64, abi.HS_PREC_MIXED])
def test_gmg_simple_trajectory_matches_oracle(hs_lib, oracle_lib, precision):
    case = configs.c2(scale=8)
    case.params.order = abi.HS_ORDER_RB
    case.params.p_tol = 1e-10
    gp = abi.Params.from_buffer_copy(case.params)
    gp.precond = abi.HS_PRECOND_GMG
    gp.precision = precision
    g, o = gpu_ctx(hs_lib, case, gp), oracle_ctx(oracle_lib, case)
    g.init(); o.init()
    for it in range(10):
        sg = g.step(); o.step()
        assert sg["p_true_relres"] < 1e-9
    fg, fo = g.fields(), o.fields()
    for k in fg:
        assert rel_l2(fg[k], fo[k]) < 1e-6, k


@pytest.mark.parametrize("precision", [abi.HS_PREC_FP64, abi.HS_PREC_MIXED])
def test_gmg_w_cycle_solution(hs_lib, oracle_lib, precision):
    """The W-cycle option (mg_cycle = 2: each level below the finest visited twice, the
    second visit starting from the first one's iterate, not from a zero guess) gives the
    same p' solution as the oracle's tight CG, in no more FCG iterations than the V-cycle."""
    case = configs.c2(scale=4)          # 64x16x16: levels 64x16x16, 32x8x8, 16x4x4
    case.params.order = abi.HS_ORDER_RB
    o = oracle_ctx(oracle_lib, case)
    o.init()
    for comp in range(3):
        o.momentum(comp)
    o.pressure_correction()
    b = o.get("b")
    op = abi.Params.from_buffer_copy(case.params)
    op.p_tol, op.p_maxit = 1e-13, 20000
    o2 = oracle_ctx(oracle_lib, case, op)
    o2.init()
    for comp in range(3):
        o2.momentum(comp)
    o2.pressure_correction()
    xo, _ = o2.pressure_solve(b)
    iters = {}
    for cyc in (1, 2):
        gp = abi.Params.from_buffer_copy(case.params)
        gp.precond, gp.precision, gp.p_tol, gp.mg_cycle = abi.HS_PRECOND_GMG, precision, 1e-10, cyc
        g = gpu_ctx(hs_lib, case, gp)
        g.init()
        for comp in range(3):
            g.momentum(comp)
        g.pressure_correction()
        xg, sg = g.pressure_solve(b)
        g.close()
        assert sg["p_converged"] == 1 and sg["p_true_relres"] < 2e-10, (cyc, sg)
        assert rel_l2(xg, xo) < 1e-7, cyc
        iters[cyc] = sg["p_iters"]
    assert iters[2] <= iters[1], iters
