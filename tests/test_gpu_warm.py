"""The warm-started p' solve (hs_params.p_warm): every SIMPLE iteration's solve meets the
same stop rule from the previous solution, takes no more FCG iterations than the
zero start, and hs_pressure_solve (another right-hand side) never starts warm."""
import numpy as np
import pytest

from hsolver_amd import abi, configs

from helpers import gpu_ctx, rel_l2

pytestmark = pytest.mark.gpu


def run(hs_lib, warm, steps=12):
    case = configs.c1()
    p = abi.Params.from_buffer_copy(case.params)
    p.order, p.precond, p.precision, p.p_warm = abi.HS_ORDER_RB, abi.HS_PRECOND_GMG, abi.HS_PREC_MIXED, warm
    p.mg_dense_coarse = 0     # no inverse rebuilds that follow each context's iteration counts
    g = gpu_ctx(hs_lib, case, p)
    g.init()
    hist = [g.step() for _ in range(steps)]
    return case, g, hist


def test_warm_start_meets_the_stop_rule_with_fewer_iterations(hs_lib):
    print("[warm] c1 mixed, 12 SIMPLE iterations cold and warm", flush=True)
    case, cold, hc = run(hs_lib, 0)
    _, warm, hw = run(hs_lib, 1)
    for it, (a, b) in enumerate(zip(hc, hw)):
        assert b["p_converged"] == 1 and b["p_true_relres"] < case.params.p_tol, (it, b)
        assert b["p_iters"] <= a["p_iters"], (it, a["p_iters"], b["p_iters"])
    assert hw[0]["p_iters"] == hc[0]["p_iters"]          # the first solve has nothing to start from
    assert sum(s["p_iters"] for s in hw) < sum(s["p_iters"] for s in hc)
    fc, fw = cold.fields(), warm.fields()
    errs = {k: rel_l2(fw[k], fc[k]) for k in fc}
    print(f"[warm] iterations cold {[s['p_iters'] for s in hc]} warm {[s['p_iters'] for s in hw]}; "
          f"fields rel-L2 {errs}", flush=True)
    assert all(e < 1e-6 for e in errs.values()), errs
    cold.close()
    warm.close()


def test_pressure_solve_is_never_warm(hs_lib):
    """hs_pressure_solve on the same right-hand side gives the same bits in a warm and a
    cold context with identical state (it always starts from zero)."""
    case, cold, _ = run(hs_lib, 0, steps=3)
    _, warm, _ = run(hs_lib, 1, steps=3)
    warm.load_state(cold.state())
    rng = np.random.default_rng(11)
    b = rng.standard_normal(abi.padded_shape("b", cold.nx, cold.ny, cold.nz))
    for g in (cold, warm):
        g.step()              # the same matrix in both (warm's solve of this step started warm)
    warm.load_state(cold.state())
    xc, sc = cold.pressure_solve(b)
    xw, sw = warm.pressure_solve(b)
    assert sc["p_iters"] == sw["p_iters"] and np.array_equal(xc, xw)
    cold.close()
    warm.close()
