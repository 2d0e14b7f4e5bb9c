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


def trajectory(hs_lib, case, precision, rows, steps=3):
    p = abi.Params.from_buffer_copy(case.params)
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
