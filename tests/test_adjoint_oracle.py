"""Pins the oracle's energy equation as a linear system and its discrete adjoint.

The operator is the Mode C conjugate energy stencil on a frozen flow; the checks are
independent of the oracle's own solver:
  * its solve matches scipy's direct solve on the exported matrix, and the solution is
    a fixed point of the oracle's own energy sweep (which applies the outflow ghost
    copy instead of the folded diagonal);
  * the adjoint variable matches scipy's direct solve with the transposed matrix, and
    the matrix-free apply passes the dot-product test <y, A x> = <A^T y, x>;
  * dJ/dq and dJ/dk match central finite differences of J (relative step 1e-6) on
    solid cells, interface cells, wall cells, inlet and outflow cells — a dropped face
    term, a wrong sign or the wrong neighbour fails.  The difference J(+h) - J(-h) is
    formed without cancellation: with A_+ T_+ = b_+ and A_- T_- = b_-,
        A_+ (T_+ - T_-) = (b_+ - b_-) - (A_+ - A_-) T_-
    is solved directly for the temperature difference (the product taken on T shifted by
    the perturbed cell's value, which row-sum preservation allows), and the p-norm
    difference is evaluated with expm1 / log1p.  This is the exact nonlinear difference (no
    linearisation), accurate to ~1e-10 relative even where dJ/dk is tiny.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from hsolver_amd import abi, configs

from helpers import cell_volumes, export_energy, oracle_ctx, rel_l2, solid_pnorm

N = 10


@pytest.fixture(scope="module")
def frozen(oracle_lib):
    """c4 geometry at 10^3 (inlet, outflow, heated cube, k ratio 100), 12 SIMPLE steps;
    temperatures are rises above the inlet / wall value (t_wall = 0)."""
    case = configs.c4(N)
    case.params.order = abi.HS_ORDER_LEX
    ctx = oracle_ctx(oracle_lib, case)
    ctx.init()
    for _ in range(12):
        ctx.step()
    return case, ctx


def system(ctx, q, k):
    ctx.put_cells("q", q)
    ctx.put_cells("k", k)
    return export_energy(ctx, N ** 3)


def central_difference(ctx, case, qp, kp, qm, km, p=8):
    """J(qp, kp) - J(qm, km) through direct solves, free of cancellation."""
    Am, bm = system(ctx, qm, km)
    Ap, bp = system(ctx, qp, kp)
    Tm = spla.spsolve(Am.tocsc(), bm)
    # a k change at an interior cell keeps every row sum, so T may be shifted by that
    # cell's value before the product, which removes the cancellation
    dA = Ap - Am
    changed = np.flatnonzero(np.abs(dA).sum(axis=1).A1 > 0)
    keeps_rows = changed.size and np.abs(dA.sum(axis=1)).max() <= 1e-9 * np.abs(dA).max()
    shift = Tm[changed[0]] if keeps_rows else 0.0
    dT = spla.spsolve(Ap.tocsc(), (bp - bm) - dA @ (Tm - shift))
    solid = case.grid.cell_type.ravel() == 1
    V = cell_volumes(case).ravel()[solid]
    t, dt = Tm[solid], dT[solid]
    S = np.sum(t ** p * V)
    dS = np.sum(t ** p * np.expm1(p * np.log1p(dt / t)) * V)
    return S ** (1.0 / p) * np.expm1(np.log1p(dS / S) / p)


def test_energy_solve_matches_direct_and_is_sweep_fixed_point(frozen):
    case, ctx = frozen
    A, b = export_energy(ctx, N ** 3)
    Td = spla.spsolve(A.tocsc(), b)
    info = ctx.energy_solve(tol=1e-13, maxit=2000)
    assert info["converged"] == 1 and info["true_relres"] < 1e-12
    T = ctx.cells("tf").ravel()
    assert rel_l2(T, Td) < 1e-11
    # one more relaxation-free sweep changes nothing: the folded system is the sweep's fixed point
    sweeps, res = ctx.fluid_energy()
    assert sweeps == 1 and res < 1e-10
    assert rel_l2(ctx.cells("tf").ravel(), Td) < 1e-11


def test_energy_operator_is_not_symmetric_and_apply_matches_csr(frozen):
    case, ctx = frozen
    A, _ = export_energy(ctx, N ** 3)
    assert abs(A - A.T).max() > 1e-6          # convection makes it non-symmetric
    rng = np.random.default_rng(7)
    x, y = rng.standard_normal(N ** 3), rng.standard_normal(N ** 3)
    Ax = ctx.energy_apply(x).ravel()
    Aty = ctx.energy_apply(y, transpose=True).ravel()
    assert rel_l2(Ax, A @ x) < 1e-14
    assert rel_l2(Aty, A.T @ y) < 1e-14
    assert abs(y @ Ax - Aty @ x) <= 1e-12 * abs(y @ Ax)


def test_adjoint_variable_matches_direct_transpose_solve(frozen):
    case, ctx = frozen
    ctx.energy_solve(tol=1e-13, maxit=2000)
    T = ctx.cells("tf")
    solid = case.grid.cell_type.reshape(N, N, N) == 1
    V = cell_volumes(case)
    _, g = solid_pnorm(T, solid, V)
    dq, dk, info = ctx.adjoint_energy(g, tol=1e-13, maxit=2000)
    assert info["converged"] == 1
    A, _ = export_energy(ctx, N ** 3)
    lam = spla.spsolve(A.T.tocsc(), g.ravel())
    assert rel_l2(dq.ravel(), lam * V.ravel()) < 1e-10


def test_adjoint_gradients_match_central_differences(frozen):
    case, ctx = frozen
    q0, k0 = ctx.cells("q").copy(), ctx.cells("k").copy()
    ctx.energy_solve(tol=1e-13, maxit=2000)
    T = ctx.cells("tf")
    solid = case.grid.cell_type.reshape(N, N, N) == 1
    _, g = solid_pnorm(T, solid, cell_volumes(case))
    dq, dk, _ = ctx.adjoint_energy(g, tol=1e-13, maxit=2000)
    lo = int(round(N * 0.375))          # first solid index (0-based) of the centred 2^3 cube
    cells_q = [(lo, lo, lo), (lo + 1, lo, lo + 1), (lo, lo + 1, lo + 1)]
    cells_k = cells_q + [(lo, lo, lo - 1),      # fluid cell touching the solid (interface faces)
                         (0, 4, 5),             # wall cell on z = 0 (boundary face)
                         (5, 0, 3),             # wall cell on y = 0
                         (3, 5, N - 1),         # outflow cell (x = max)
                         (4, 4, 0)]             # inlet cell (x = 0)
    for (kz, jy, ix) in cells_q:
        h = 1e-6 * q0[kz, jy, ix]
        qp, qm = q0.copy(), q0.copy()
        qp[kz, jy, ix] += h
        qm[kz, jy, ix] -= h
        fd = central_difference(ctx, case, qp, k0, qm, k0) / (2 * h)
        assert fd == pytest.approx(dq[kz, jy, ix], rel=1e-7), ("q", kz, jy, ix)
    for (kz, jy, ix) in cells_k:
        h = 1e-6 * k0[kz, jy, ix]
        kp, km = k0.copy(), k0.copy()
        kp[kz, jy, ix] += h
        km[kz, jy, ix] -= h
        fd = central_difference(ctx, case, q0, kp, q0, km) / (2 * h)
        # (lambda_P - lambda_nb) differences amplify the transpose solve's 1e-13 residual tolerance
        assert fd == pytest.approx(dk[kz, jy, ix], rel=1e-6), ("k", kz, jy, ix, fd, dk[kz, jy, ix])
    ctx.put_cells("q", q0)
    ctx.put_cells("k", k0)


def test_solve_based_central_difference_matches_matrix_difference(frozen):
    """hsolver_amd.adjoint.central_difference (operator applies + a right-hand-side solve,
    the form the GPU check uses) against the exported-matrix difference above, on
    interior, interface and boundary cells."""
    from hsolver_amd.adjoint import central_difference as solve_difference
    case, ctx = frozen
    q0, k0 = ctx.cells("q").copy(), ctx.cells("k").copy()
    solid = case.grid.cell_type.reshape(N, N, N) == 1
    V = cell_volumes(case)
    lo = int(round(N * 0.375))
    # interior solid and interface cells (shifted product), then boundary cells, where the
    # k change moves the boundary face's conductance and the product is taken unshifted:
    # a wall cell on z = 0, one on y = 0, an outflow cell (x = max) and an inlet cell
    for which, cell in (("q", (lo, lo + 1, lo)), ("k", (lo + 1, lo + 1, lo)), ("k", (lo, lo, lo - 1)),
                        ("k", (0, 4, 5)), ("k", (5, 0, 3)), ("k", (3, 5, N - 1)), ("k", (4, 4, 0))):
        base = q0 if which == "q" else k0
        h = 1e-6 * base[cell]
        a = solve_difference(ctx, which, cell, h, q0, k0, solid, V, tol=1e-14, maxit=5000)
        plus, minus = base.copy(), base.copy()
        plus[cell] += h
        minus[cell] -= h
        if which == "q":
            b = central_difference(ctx, case, plus, k0, minus, k0) / (2 * h)
        else:
            b = central_difference(ctx, case, q0, plus, q0, minus) / (2 * h)
        assert a == pytest.approx(b, rel=1e-8), (which, cell)
    ctx.put_cells("q", q0)
    ctx.put_cells("k", k0)


def test_shifted_product_is_wrong_at_a_boundary_cell(frozen):
    """The shift is what the boundary rule switches off: forced on at a wall cell, the
    difference misses the boundary face's change and disagrees with the matrix difference."""
    from hsolver_amd.adjoint import central_difference as solve_difference
    case, ctx = frozen
    q0, k0 = ctx.cells("q").copy(), ctx.cells("k").copy()
    solid = case.grid.cell_type.reshape(N, N, N) == 1
    V = cell_volumes(case)
    cell = (0, 4, 5)
    h = 1e-6 * k0[cell]
    plus, minus = k0.copy(), k0.copy()
    plus[cell] += h
    minus[cell] -= h
    want = central_difference(ctx, case, q0, plus, q0, minus) / (2 * h)
    good = solve_difference(ctx, "k", cell, h, q0, k0, solid, V, tol=1e-14, maxit=5000)
    bad = solve_difference(ctx, "k", cell, h, q0, k0, solid, V, tol=1e-14, maxit=5000, shift=True)
    assert good == pytest.approx(want, rel=1e-8)
    assert abs(bad - want) > 1e-3 * abs(want)
    ctx.put_cells("q", q0)
    ctx.put_cells("k", k0)
