"""Warm-started p' solves (hs_params.p_warm 1-3: the previous SIMPLE iteration's p', or the
line / parabola through the last two / three): every solve meets the same stop rule, the
warm trajectories stay close to the zero-start one and take fewer FCG iterations, a start
no better than zero falls back to zero (then the solve is the zero start's, bit for bit),
and hs_pressure_solve (another right-hand side) never starts warm."""
import numpy as np
import pytest

from hsolver_amd import abi, configs

from helpers import gpu_ctx, rel_l2

pytestmark = pytest.mark.gpu


def make(hs_lib, warm):
    case = configs.c1()
    p = abi.Params.from_buffer_copy(case.params)
    p.order, p.precond, p.precision, p.p_warm = abi.HS_ORDER_RB, abi.HS_PRECOND_GMG, abi.HS_PREC_MIXED, warm
    p.mg_dense_coarse = 0     # no inverse rebuilds that follow each context's iteration counts
    g = gpu_ctx(hs_lib, case, p)
    g.init()
    return case, g


@pytest.mark.parametrize("warm", [1, 2, 3])
def test_warm_start_meets_the_stop_rule_with_fewer_iterations(hs_lib, warm):
    print(f"[warm] c1 mixed, 12 SIMPLE iterations from zero and with p_warm = {warm}", flush=True)
    case, cold = make(hs_lib, 0)
    _, hot = make(hs_lib, warm)
    hc = [cold.step() for _ in range(12)]
    hw = [hot.step() for _ in range(12)]
    for it, (a, b) in enumerate(zip(hc, hw)):
        assert b["p_converged"] == 1 and b["p_true_relres"] < case.params.p_tol, (it, b)
        assert b["p_iters"] <= a["p_iters"] + (0 if warm == 1 else 1), (it, a["p_iters"], b["p_iters"])
    assert hw[0]["p_iters"] == hc[0]["p_iters"]          # the first solve has nothing to start from
    assert sum(s["p_iters"] for s in hw) < sum(s["p_iters"] for s in hc)
    fc, fw = cold.fields(), hot.fields()
    errs = {k: rel_l2(fw[k], fc[k]) for k in fc}
    print(f"[warm] iterations zero {[s['p_iters'] for s in hc]} warm {[s['p_iters'] for s in hw]}; "
          f"fields rel-L2 {errs}", flush=True)
    assert all(e < 1e-6 for e in errs.values()), errs
    cold.close()
    hot.close()


def test_start_worse_than_zero_falls_back_to_zero(hs_lib):
    """After four steps the velocities are scaled by 1e-2, so the p' right-hand side
    shrinks a hundredfold and the parabola through the held solutions leaves a residual
    far above ||b||: that solve must be the zero start's, bit for bit."""
    _, cold = make(hs_lib, 0)
    _, hot = make(hs_lib, 3)
    for _ in range(4):
        cold.step()
        hot.step()
    st = cold.state()
    for key in ("u", "v", "w"):
        st[key] = st[key] * 1e-2
    cold.load_state(st)
    hot.load_state(st)
    a, b = cold.step(), hot.step()
    assert a["p_iters"] == b["p_iters"] and a["p_converged"] == b["p_converged"] == 1
    sa, sb = cold.state(), hot.state()
    for key in sa:
        assert np.array_equal(sa[key], sb[key]), key
    cold.close()
    hot.close()


def test_pressure_solve_is_never_warm(hs_lib):
    """hs_pressure_solve on the same right-hand side gives the same bits in a warm and a
    zero-start context with identical state and matrix (it always starts from zero)."""
    _, cold = make(hs_lib, 0)
    _, hot = make(hs_lib, 3)
    for _ in range(3):
        cold.step()
        hot.step()
    hot.load_state(cold.state())
    for g in (cold, hot):
        g.step()              # the same matrix in both (hot's solve of this step started warm)
    hot.load_state(cold.state())
    rng = np.random.default_rng(11)
    b = rng.standard_normal(abi.padded_shape("b", cold.nx, cold.ny, cold.nz))
    xc, sc = cold.pressure_solve(b)
    xw, sw = hot.pressure_solve(b)
    assert sc["p_iters"] == sw["p_iters"] and np.array_equal(xc, xw)
    cold.close()
    hot.close()
