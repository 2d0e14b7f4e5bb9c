"""The fused red-black sweep against the two in-place colour passes it replaces.

The multigrid smoother's fused kernel (one pass per sweep, first colour recomputed on a
one-cell tile ring, operator streamed through registers) performs the same operations
in the same order as the per-colour passes, so whole SIMPLE trajectories with the
multigrid-preconditioned pressure solve must agree bit for bit, for every tile shape:
64-wide tiles of 4, 8, 16 rows (the tile-side ring cells are interior cells of the
level) and the automatic 256 / 128 / 64-wide tiles, on level sizes that are not
multiples of the tiles, in both precisions.
"""
import numpy as np
import pytest

from hsolver_amd import abi, configs

from helpers import gpu_ctx

pytestmark = pytest.mark.gpu

CASES = {
    "c3_90": lambda: configs.c3(90),            # levels 90, 45, 23, 12: odd, partial tiles
    "c2_64x16x16": lambda: configs.c2(scale=4),
    "c3_37": lambda: configs.c3(37),
    "c3_160": lambda: configs.c3(160),          # 256-wide tiles with 96 idle columns
}


def trajectory(hs_lib, case, precision, rows, steps=3, omega=(1.0, 1.0), wide=0):
    p = abi.Params.from_buffer_copy(case.params)
    p.mg_omega, p.mg_omega_coarse = omega
    p.mg_wide_offsets = wide
    p.order = abi.HS_ORDER_RB
    p.precond = abi.HS_PRECOND_GMG
    p.precision = precision
    p.mg_sweep_rows = rows
    p.mg_dense_coarse = 10 if precision == abi.HS_PREC_MIXED else 0
    g = gpu_ctx(hs_lib, case, p)
    g.init()
    iters = [g.step()["p_iters"] for _ in range(steps)]
    f = g.fields()
    g.close()
    return f, iters


@pytest.mark.parametrize("precision", [abi.HS_PREC_MIXED, abi.HS_PREC_FP64], ids=["mixed", "fp64"])
@pytest.mark.parametrize("name", list(CASES))
def test_fused_sweep_bit_equal_to_colour_passes(hs_lib, name, precision):
    case = CASES[name]()
    print(f"[sweep] {name} precision {precision}", flush=True)
    ref, it_ref = trajectory(hs_lib, case, precision, -1)
    for rows in (0, 4, 8, 16):
        f, it = trajectory(hs_lib, case, precision, rows)
        assert it == it_ref, (rows, it, it_ref)
        for k in ref:
            assert np.array_equal(f[k], ref[k]), (rows, k)


@pytest.mark.parametrize("precision", [abi.HS_PREC_MIXED, abi.HS_PREC_FP64], ids=["mixed", "fp64"])
@pytest.mark.parametrize("name", ["c3_90", "c2_64x16x16"])
def test_over_relaxed_fused_sweep_bit_equal_to_colour_passes(hs_lib, name, precision):
    """The same with over-relaxed smoothing (1.25 on the finest level, 1.5 below): the
    fused sweep and the colour passes apply x + omega (x_GS - x) alike, level by level."""
    case = CASES[name]()
    ref, it_ref = trajectory(hs_lib, case, precision, -1, omega=(1.25, 1.5))
    for rows in (0, 8):
        f, it = trajectory(hs_lib, case, precision, rows, omega=(1.25, 1.5))
        assert it == it_ref, (rows, it, it_ref)
        for k in ref:
            assert np.array_equal(f[k], ref[k]), (rows, k)


@pytest.mark.parametrize("e_omega", [1.0, 1.3], ids=["gauss_seidel", "over_relaxed"])
@pytest.mark.parametrize("precision", [abi.HS_PREC_MIXED, abi.HS_PREC_FP64], ids=["mixed", "fp64"])
def test_fused_energy_sweep_bit_equal_to_colour_passes(hs_lib, precision, e_omega):
    """The energy multigrid's fused sweep against its two in-place colour passes, on one
    frozen flow: the forward energy solve (A) and the adjoint solve (A^T) give the same
    iteration counts and the same bits.  24^3 (tiles 64 x 8 with side cells in the
    domain) and 150^3 (256-wide tiles on the finest level, partial tiles in y, 128 x 8
    tiles on the next).  Gauss-Seidel sweeps and over-relaxed ones."""
    from helpers import cell_volumes, solid_pnorm
    for case in (configs.c4(24), configs.c3(150)):
        p = abi.Params.from_buffer_copy(case.params)
        p.order, p.precond, p.precision = abi.HS_ORDER_RB, abi.HS_PRECOND_GMG, precision
        p.e_omega = e_omega
        g = gpu_ctx(hs_lib, case, p)
        g.init()
        for _ in range(5):
            g.step()
        state = g.state()
        g.close()
        res = {}
        for rows in (-1, 0):
            p.mg_sweep_rows = rows
            h = gpu_ctx(hs_lib, case, p)
            h.init()
            h.load_state(state)
            fwd = h.energy_solve(tol=1e-12, maxit=400)
            T = h.cells("tf")
            solid = case.grid.cell_type.reshape(T.shape) == 1
            _, dJdT = solid_pnorm(T, solid, cell_volumes(case))
            dq, dk, adj = h.adjoint_energy(dJdT, tol=1e-12, maxit=400)
            res[rows] = (fwd["iters"], adj["iters"], T, dq, dk)
            h.close()
        print(f"[energy sweep] {case.name} precision {precision}: forward {res[0][0]}, adjoint {res[0][1]} iterations",
              flush=True)
        assert res[0][0] == res[-1][0] and res[0][1] == res[-1][1], (res[0][:2], res[-1][:2])
        for a, b in zip(res[0][2:], res[-1][2:]):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", [abi.HS_PREC_MIXED, abi.HS_PREC_FP64], ids=["mixed", "fp64"])
def test_wide_offset_sweeps_bit_equal(hs_lib, precision):
    """The 64-bit-offset form of the fused sweeps, which levels of 2 GiB and more take,
    forced on small grids against the 32-bit form: the same SIMPLE trajectories (pressure
    multigrid, default over-relaxation) and the same energy and adjoint solves, bit for bit."""
    from helpers import cell_volumes, solid_pnorm
    case = CASES["c3_90"]()
    ref, it_ref = trajectory(hs_lib, case, precision, 0, omega=(1.25, 1.5))
    f, it = trajectory(hs_lib, case, precision, 0, omega=(1.25, 1.5), wide=1)
    print(f"[sweep] wide offsets, precision {precision}: FCG iterations {it} against {it_ref}", flush=True)
    assert it == it_ref
    for k in ref:
        assert np.array_equal(f[k], ref[k]), k
    ecase = configs.c4(24)
    p = abi.Params.from_buffer_copy(ecase.params)
    p.order, p.precond, p.precision = abi.HS_ORDER_RB, abi.HS_PRECOND_GMG, precision
    g = gpu_ctx(hs_lib, ecase, p)
    g.init()
    for _ in range(5):
        g.step()
    state = g.state()
    g.close()
    res = {}
    for wide in (0, 1):
        p.mg_wide_offsets = wide
        h = gpu_ctx(hs_lib, ecase, p)
        h.init()
        h.load_state(state)
        fwd = h.energy_solve(tol=1e-12, maxit=400)
        T = h.cells("tf")
        solid = ecase.grid.cell_type.reshape(T.shape) == 1
        _, dJdT = solid_pnorm(T, solid, cell_volumes(ecase))
        dq, dk, adj = h.adjoint_energy(dJdT, tol=1e-12, maxit=400)
        res[wide] = (fwd["iters"], adj["iters"], T, dq, dk)
        h.close()
    assert res[0][:2] == res[1][:2], (res[0][:2], res[1][:2])
    for a, b in zip(res[0][2:], res[1][2:]):
        assert np.array_equal(a, b)


def test_smoothing_chosen_by_grid(hs_lib):
    """hs_create fills the multigrid settings left unset from the grid's cell aspect ratio:
    one pre- and two post-sweeps on the finest level, V(3,3) below and over-relaxation
    1.2 / 1.6 on a near-isotropic grid; V(6,6) and 1.7 everywhere on a stretched one;
    explicit values are kept."""
    from hsolver_amd import core
    iso = configs.c3(32)
    p = abi.Params.from_buffer_copy(iso.params)
    p.precond = abi.HS_PRECOND_GMG
    ctx = core.create(iso.grid, p, iso.q)
    r = ctx.run_params()
    ctx.close()
    assert (r.mg_pre0, r.mg_post0, r.mg_pre, r.mg_post) == (1, 2, 3, 3)
    assert (round(r.mg_omega, 4), round(r.mg_omega_coarse, 4), r.e_omega) == (1.2, 1.6, 1.0)
    assert (r.e_pre, r.e_post) == (3, 3)
    # aggregation multigrid: two finest-level sweeps, three Jacobi sweeps, theta 0.1,
    # Gauss-Seidel on the finest level; the correction scaling stays chosen by use
    assert (r.amg_sweeps0, r.amg_sweeps, round(r.amg_theta, 4), r.amg_omega, r.amg_alpha) == (2, 3, 0.1, 1.0, 0.0)
    assert r.e_precond == abi.HS_PRECOND_GMG
    st = configs.c5(32, 16, 16, nfins=2)
    assert configs.max_cell_aspect(st.grid.xv, st.grid.yv, st.grid.zv) > 8
    p = abi.Params.from_buffer_copy(st.params)
    p.precond = abi.HS_PRECOND_GMG
    ctx = core.create(st.grid, p, st.q)
    r = ctx.run_params()
    ctx.close()
    assert (r.mg_pre, r.mg_post, r.mg_pre0, r.mg_post0) == (6, 6, -1, -1)
    assert (round(r.mg_omega, 4), round(r.mg_omega_coarse, 4)) == (1.7, 1.7)
    p = abi.Params.from_buffer_copy(iso.params)
    p.mg_pre, p.mg_post, p.mg_omega, p.mg_omega_coarse = 2, 2, 1.0, 1.0
    ctx = core.create(iso.grid, p, iso.q)
    r = ctx.run_params()
    ctx.close()
    assert (r.mg_pre, r.mg_post, r.mg_pre0, r.mg_post0, r.mg_omega, r.mg_omega_coarse) == (2, 2, -1, -1, 1.0, 1.0)
