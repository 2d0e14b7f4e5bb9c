"""The oracle against the committed scipy golden vectors (tests/golden/make_fixtures.py).

The fixtures hold right-hand sides and scipy spsolve solutions; the matrices are rebuilt
here by the oracle, so a change in its assembly breaks A x = b at the stored solution, and
its solvers (HYPRE-semantics CG, the energy solve and its transpose) are checked against
the direct solutions.  scipy is not needed at test time.
"""
from pathlib import Path

import numpy as np
import pytest

from hsolver_amd import abi, configs

from helpers import cell_volumes, export_energy, export_p, oracle_ctx, rel_l2, small_ref_case

GOLD = Path(__file__).resolve().parent / "golden"


def first_p_system(lib, case):
    o = oracle_ctx(lib, case)
    o.init()
    for comp in range(3):
        o.momentum(comp)
    o.pressure_correction()
    return o, export_p(o)


@pytest.mark.parametrize("name, make", [("p_c1", configs.c1), ("p_refcase", small_ref_case)])
def test_pressure_system_golden(oracle_lib, name, make):
    fx = np.load(GOLD / f"{name}.npz")
    case = make()
    o, (A, b, cell) = first_p_system(oracle_lib, case)
    assert np.array_equal(cell, fx["row_cell"])
    assert rel_l2(b, fx["b"]) < 1e-14
    assert np.linalg.norm(A @ fx["x"] - fx["b"]) <= 1e-11 * np.linalg.norm(fx["b"])
    if case.params.mode == abi.HS_MODE_CONFIG:
        # the symmetric Mode C operator: the oracle's CG (HYPRE stop rule) reaches the direct solution
        xo, st = o.pressure_solve(o.get("b"))
        assert st["p_converged"] == 1
        x_rows = xo.reshape(-1)[cell]          # row_cell holds P-layout indices
        assert rel_l2(x_rows, fx["x"]) < 1e-6


def test_energy_and_adjoint_golden(oracle_lib):
    fx = np.load(GOLD / "e_c4.npz")
    flow = np.load(GOLD / "e_c4_flow.npz")
    n = 10
    case = configs.c4(n)
    case.params.order = abi.HS_ORDER_LEX
    o = oracle_ctx(oracle_lib, case)
    o.init()
    o.load_state({k: flow[k] for k in ("u", "v", "w", "tf")})
    A, b = export_energy(o, n ** 3)
    assert rel_l2(b, fx["b"]) < 1e-14
    assert np.linalg.norm(A @ fx["T"] - fx["b"]) <= 1e-11 * np.linalg.norm(fx["b"])
    assert np.linalg.norm(A.T @ fx["lam"] - fx["dJdT"]) <= 1e-11 * np.linalg.norm(fx["dJdT"])
    info = o.energy_solve(1e-13, 5000)
    assert info["converged"] == 1
    assert rel_l2(o.cells("tf").ravel(), fx["T"]) < 1e-11
    dq, dk, ia = o.adjoint_energy(fx["dJdT"], 1e-13, 5000)
    assert ia["converged"] == 1
    assert rel_l2(dq.ravel(), fx["lam"] * cell_volumes(case).ravel()) < 1e-10
