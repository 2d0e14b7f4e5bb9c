"""Error paths of the p' solve and the energy solves, forced from the host.

The reference ignores PCG non-convergence (fluid_solver.h:1792-1793) and has no NaN
check; here a Krylov reduction that is not finite stops the solve at once and the call
returns HS_E_NAN, a solve that reaches maxit returns HS_E_NOCONV when fail_on_noconv is
set, and the adjoint wrapper raises when a forward or adjoint solve did not converge.
NaN values only ever enter arithmetic (no index depends on a value), so every kernel
runs to its normal end."""
import numpy as np
import pytest

from hsolver_amd import abi, configs
from hsolver_amd.core import HSError

from helpers import gpu_ctx

pytestmark = pytest.mark.gpu

SOLVERS = [(abi.HS_PRECOND_NONE, abi.HS_PREC_FP64), (abi.HS_PRECOND_GMG, abi.HS_PREC_FP64),
           (abi.HS_PRECOND_GMG, abi.HS_PREC_MIXED)]


def channel(hs_lib, precond, precision, **over):
    case = configs.c2(scale=8)
    p = abi.Params.from_buffer_copy(case.params)
    p.order, p.precond, p.precision = abi.HS_ORDER_RB, precond, precision
    for k, v in over.items():
        setattr(p, k, v)
    g = gpu_ctx(hs_lib, case, p)
    g.init()
    return g


@pytest.mark.parametrize("precond,precision", SOLVERS)
def test_noconv_with_fail_on_noconv(hs_lib, precond, precision):
    g = channel(hs_lib, precond, precision, p_maxit=1, fail_on_noconv=1)
    with pytest.raises(HSError) as e:
        g.step()
    assert e.value.code == abi.HS_E_NOCONV
    g.close()
    # without the flag the same solve returns normally and reports p_converged = 0
    g = channel(hs_lib, precond, precision, p_maxit=1)
    st = g.step()
    assert st["p_converged"] == 0 and st["p_iters"] == 1
    g.close()


@pytest.mark.parametrize("precond,precision", SOLVERS)
def test_nan_in_rhs_stops_the_solve(hs_lib, precond, precision):
    g = channel(hs_lib, precond, precision)
    g.step()                                   # assembles the operator (and the hierarchy)
    shape = abi.padded_shape("b", g.nx, g.ny, g.nz)
    b = np.zeros(shape)
    b[1:-1, 1:-1, 1:-1] = 1.0
    b[3, 3, 3] = np.nan
    with pytest.raises(HSError) as e:
        g.pressure_solve(b)
    assert e.value.code == abi.HS_E_NAN
    # the context stays usable: the next solve with a finite right-hand side converges
    b[3, 3, 3] = 1.0
    x, st = g.pressure_solve(b)
    assert st["p_converged"] == 1 and np.all(np.isfinite(x))
    g.close()


@pytest.mark.parametrize("precond,precision", SOLVERS)
def test_nan_velocity_in_pressure_correction(hs_lib, precond, precision):
    g = channel(hs_lib, precond, precision)
    g.step()
    u = g.get("u")
    u[4, 4, 5] = np.nan                        # an interior face: the p' right-hand side picks it up
    g.put("u", u)
    with pytest.raises(HSError) as e:
        g.pressure_correction()
    assert e.value.code == abi.HS_E_NAN
    g.close()


def test_nan_velocity_in_step(hs_lib):
    g = channel(hs_lib, abi.HS_PRECOND_GMG, abi.HS_PREC_MIXED)
    g.step()
    u = g.get("u")
    u[4, 4, 5] = np.nan
    g.put("u", u)
    with pytest.raises(HSError) as e:
        g.step()
    assert e.value.code == abi.HS_E_NAN
    g.close()


def test_adjoint_wrapper_refuses_unconverged_solves(hs_lib):
    import torch
    from hsolver_amd.adjoint import EnergySolve, EnergySolver, SolveError, cell_volumes, solid_pnorm
    case = configs.c4(12)
    case.params.order = abi.HS_ORDER_RB
    g = gpu_ctx(hs_lib, case)
    g.init()
    for _ in range(3):
        g.step()
    dev = torch.device("cuda:0")
    shape = (12, 12, 12)
    q = torch.from_numpy(case.q.reshape(shape)).to(dev).requires_grad_(True)
    k = torch.from_numpy(g.cells("k")).to(dev).requires_grad_(True)
    with pytest.raises(SolveError):
        EnergySolve.apply(q, k, EnergySolver(g, tol=1e-13, maxit=1))
    # converged forward, adjoint capped at one iteration
    solver = EnergySolver(g, tol=1e-13, maxit=200)
    solid = torch.from_numpy(case.grid.cell_type.reshape(shape) == 1).to(dev)
    V = torch.from_numpy(cell_volumes(case)).to(dev)
    J = solid_pnorm(EnergySolve.apply(q, k, solver), solid, V)
    solver.maxit = 1
    with pytest.raises(SolveError):
        J.backward()
    g.close()
