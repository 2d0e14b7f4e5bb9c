"""hs_pressure_solve_static (the p' solve with its scratch in a fixed device array, for
callers that cannot allocate) against hs_pressure_solve on the same context and
right-hand side: the same kernels on the same data, so the solutions and statistics
must be identical.  Two calls in a row per solver, the second on a smaller grid; a
multigrid context whose hierarchy does not exist yet and a grid whose reductions need
more partials than the array holds are refused, and so is the aggregation multigrid."""
import ctypes as C

import numpy as np
import pytest

from hsolver_amd import abi, configs

from helpers import gpu_ctx

pytestmark = pytest.mark.gpu

SOLVERS = [(abi.HS_PRECOND_NONE, abi.HS_PREC_FP64), (abi.HS_PRECOND_GMG, abi.HS_PREC_FP64),
           (abi.HS_PRECOND_GMG, abi.HS_PREC_MIXED)]


def static_solve(lib, ctx, b):
    import torch
    dev = torch.device("cuda:0")
    bd = torch.from_numpy(np.ascontiguousarray(b, dtype=np.float64).ravel()).to(dev)
    xd = torch.full_like(bd, 7.0)
    torch.cuda.synchronize()
    st = abi.Stats()
    fn = lib.hs_pressure_solve_static
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(abi.Stats)]
    rc = fn(ctx.h, C.c_void_p(bd.data_ptr()), C.c_void_p(xd.data_ptr()), C.byref(st))
    return rc, xd.cpu().numpy(), st.as_dict()


def channel(hs_lib, scale, precond, precision, steps=1):
    case = configs.c2(scale=scale)
    p = abi.Params.from_buffer_copy(case.params)
    p.order, p.precond, p.precision = abi.HS_ORDER_RB, precond, precision
    p.mg_dense_coarse = 50 if precision == abi.HS_PREC_MIXED else 0
    g = gpu_ctx(hs_lib, case, p)
    g.init()
    for _ in range(steps):
        g.step()
    return g


@pytest.mark.parametrize("precond,precision", SOLVERS)
def test_static_matches_pressure_solve(hs_lib, precond, precision):
    import torch  # noqa: F401
    rng = np.random.default_rng(5)
    for scale in (4, 8):                               # 64x16x16, then the smaller 32x8x8
        g = channel(hs_lib, scale, precond, precision)
        shape = abi.padded_shape("b", g.nx, g.ny, g.nz)
        b = rng.standard_normal(shape)
        x_ref, st_ref = g.pressure_solve(b)
        rc, x, st = static_solve(hs_lib, g, b)
        assert rc == 0, rc
        assert st_ref["p_converged"] == 1 and st["p_iters"] == st_ref["p_iters"]
        assert np.array_equal(x, x_ref.ravel())
        for key in ("p_relres", "p_true_relres", "mass_imbalance"):
            assert st[key] == st_ref[key], key
        # and once more with the context's own buffers after the static call
        x_again, st_again = g.pressure_solve(b)
        assert np.array_equal(x_again, x_ref) and st_again["p_iters"] == st_ref["p_iters"]
        g.close()


def test_static_refuses_missing_hierarchy(hs_lib):
    g = channel(hs_lib, 8, abi.HS_PRECOND_GMG, abi.HS_PREC_MIXED, steps=0)
    rc, _, _ = static_solve(hs_lib, g, np.zeros(abi.padded_shape("b", g.nx, g.ny, g.nz)))
    assert rc == abi.HS_E_STATE
    g.close()


def test_static_refuses_aggregation_multigrid(hs_lib):
    """The aggregation multigrid allocates when it aggregates again, so the no-allocation
    entry point refuses it (the context's own solve still works)."""
    g = channel(hs_lib, 8, abi.HS_PRECOND_AMG, abi.HS_PREC_MIXED)
    b = np.random.default_rng(6).standard_normal(abi.padded_shape("b", g.nx, g.ny, g.nz))
    rc, _, _ = static_solve(hs_lib, g, b)
    assert rc == abi.HS_E_ARG
    _, st = g.pressure_solve(b)
    assert st["p_converged"] == 1
    g.close()


def test_static_refuses_grid_beyond_scratch(hs_lib):
    # 1024 x 4100 cells per plane: 16 x 1025 column tiles, so the residual pass writes
    # 2 x 16400 partials, more than the array's 32768
    nx, ny, nz = 1024, 4100, 3
    grid = abi.GridArrays(configs.uniform_axis(nx, 1.0), configs.uniform_axis(ny, 4.0), configs.uniform_axis(nz, 0.01),
                          np.zeros(nx * ny * nz, dtype=np.uint8))
    p = abi.params_default(abi.HS_MODE_CONFIG)
    from hsolver_amd.core import Context
    g = Context(hs_lib, "hs", grid, p)
    rc, _, _ = static_solve(hs_lib, g, np.zeros((nz + 2) * (ny + 2) * (nx + 2)))
    assert rc == abi.HS_E_ARG
    g.close()
