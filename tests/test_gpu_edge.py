"""Degenerate and ragged grids, GPU against the oracle in the benchmark physics (Mode C):

  * an all-solid box: no p' unknowns (an empty system), conduction only;
  * an all-fluid box: no solid cells, no interface;
  * the thinnest domains accepted: three cells in z, and three cells in y and z (a line);
    (hs_create refuses fewer than three cells along any axis, tests/test_abi.py);
  * a ragged 33x17x9 box with a block touching two walls.

For each, a few SIMPLE iterations of the geometric and the aggregation multigrid paths (FP64
and mixed) and of the drop-in CG against the oracle's red-black sweeps with its CG: fields within 1e-6 relative L2 (the
two Krylov solvers meet the same stop rule), the empty system finishing without iterations.
"""
import numpy as np
import pytest

from hsolver_amd import abi, configs

from helpers import gpu_ctx, oracle_ctx, rel_l2

pytestmark = pytest.mark.gpu


def make(name, nx, ny, nz, boxes, full_solid=False):
    ct = configs.box_cells(nx, ny, nz, boxes)
    if full_solid:
        ct[:] = 1
    p = configs.config_params(re=200.0, k_ratio=50.0, lid=True, inlet=True)
    q = np.where(ct.reshape(-1) == 1, 1.0, 0.0)
    return configs.Case(name, configs.uniform_axis(nx, nx / 16.0), configs.uniform_axis(ny, ny / 16.0),
                        configs.uniform_axis(nz, nz / 16.0), ct, p, q, name)


CASES = {
    "all_solid": lambda: make("all_solid", 12, 10, 8, [], full_solid=True),
    "all_fluid": lambda: make("all_fluid", 20, 12, 10, []),
    "slab_z3": lambda: make("slab_z3", 24, 12, 3, [(10, 13, 4, 7, 1, 3)]),
    "line_yz3": lambda: make("line_yz3", 40, 3, 3, []),
    "ragged": lambda: make("ragged", 33, 17, 9, [(20, 26, 1, 6, 4, 9)]),
}

PATHS = {
    "gmg_fp64": (abi.HS_PRECOND_GMG, abi.HS_PREC_FP64),
    "gmg_mixed": (abi.HS_PRECOND_GMG, abi.HS_PREC_MIXED),
    "dropin_cg": (abi.HS_PRECOND_NONE, abi.HS_PREC_FP64),
    "amg_fp64": (abi.HS_PRECOND_AMG, abi.HS_PREC_FP64),
    "amg_mixed": (abi.HS_PRECOND_AMG, abi.HS_PREC_MIXED),
}


@pytest.mark.parametrize("path", list(PATHS))
@pytest.mark.parametrize("name", list(CASES))
def test_degenerate_and_ragged_grids(hs_lib, oracle_lib, name, path):
    case = CASES[name]()
    case.params.order = abi.HS_ORDER_RB
    case.params.p_tol = 1e-10
    gp = abi.Params.from_buffer_copy(case.params)
    gp.precond, gp.precision = PATHS[path]
    gp.mg_dense_coarse = 10 if gp.precision == abi.HS_PREC_MIXED or gp.precond == abi.HS_PRECOND_AMG else 0
    print(f"[edge] {name} {case.shape} {path}", flush=True)
    g, o = gpu_ctx(hs_lib, case, gp), oracle_ctx(oracle_lib, case)
    g.init()
    o.init()
    for it in range(4):
        sg, so = g.step(), o.step()
        assert sg["p_converged"] == 1, (it, sg)
        if name == "all_solid":
            assert sg["p_iters"] == 0 and sg["mass_imbalance"] == 0.0, sg
    fg, fo = g.fields(), o.fields()
    g.close()
    o.close()
    for k in fg:
        assert np.all(np.isfinite(fg[k])), k
        assert rel_l2(fg[k], fo[k]) < 1e-6, (k, rel_l2(fg[k], fo[k]))

