"""Writes the scipy golden vectors under tests/golden/ (run in the build container only).

Each fixture holds a right-hand side and its scipy.sparse.linalg.spsolve solution for a
system the oracle assembles:
  p_c1.npz      c1, first SIMPLE iteration: the Mode C p' system over fluid cells
  p_refcase.npz a 12x10x8 Mode R cavity on a stretched grid with one solid block: the
                reference's (non-symmetric) p' system of iteration 1
  e_c4.npz      c4 at 10^3 after 12 SIMPLE iterations: the frozen-flow energy system and
                its transpose (the adjoint) with dJ/dT of the solid p-norm
The matrices are not stored: tests/test_golden.py rebuilds them with the oracle, so the
fixtures pin the oracle's assembly (A x = b at the stored solution) and its solvers.

    python tests/golden/make_fixtures.py
"""
import ctypes as C
import sys
from pathlib import Path

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from hsolver_amd import abi, configs  # noqa: E402
from helpers import cell_volumes, export_energy, export_p, oracle_ctx, small_ref_case, solid_pnorm  # noqa: E402


def oracle_lib():
    lib = C.CDLL(str(ROOT / "oracle" / "liboracle.so"))
    abi.declare(lib, "or")
    return lib


def p_system(lib, case):
    o = oracle_ctx(lib, case)
    o.init()
    for comp in range(3):
        o.momentum(comp)
    o.pressure_correction()
    A, b, cell = export_p(o)
    return A, b, cell


def main():
    lib = oracle_lib()
    case = configs.c1()
    A, b, cell = p_system(lib, case)
    x = spla.spsolve(A.tocsc(), b)
    np.savez_compressed(HERE / "p_c1.npz", b=b, x=x, row_cell=cell)
    case = small_ref_case()
    A, b, cell = p_system(lib, case)
    x = spla.spsolve(A.tocsc(), b)
    np.savez_compressed(HERE / "p_refcase.npz", b=b, x=x, row_cell=cell)
    n = 10
    case = configs.c4(n)
    case.params.order = abi.HS_ORDER_LEX
    o = oracle_ctx(lib, case)
    o.init()
    for _ in range(12):
        o.step()
    A, b = export_energy(o, n ** 3)
    T = spla.spsolve(A.tocsc(), b)
    solid = case.grid.cell_type.reshape(n, n, n) == 1
    _, g = solid_pnorm(T.reshape(n, n, n), solid, cell_volumes(case))
    lam = spla.spsolve(A.T.tocsc(), g.ravel())
    np.savez_compressed(HERE / "e_c4.npz", b=b, T=T, dJdT=g.ravel(), lam=lam)
    # the frozen flow of the energy fixture, so the test rebuilds the same operator
    st = o.state(("u", "v", "w", "tf"))
    np.savez_compressed(HERE / "e_c4_flow.npz", **st)
    for f in sorted(HERE.glob("*.npz")):
        print(f.name, f.stat().st_size, "bytes")


if __name__ == "__main__":
    main()
