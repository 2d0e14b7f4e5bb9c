"""GPU energy solve and discrete adjoint against the oracle on the same frozen flow.

  * forward solve and the matrix-free operator / transpose against the oracle's
    assembled system (FP64 and mixed-precision V-cycles);
  * dJ/dq and dJ/dk against the oracle's FP64 transpose solve (full gradients);
  * central differences (relative step 1e-6) on 32 seeded solid cells and on wall,
    inlet, outflow and interface cells, through the GPU forward solve (J(+h) - J(-h)
    formed by solving for the temperature difference, hsolver_amd.adjoint.central_difference);
  * the torch.autograd path (EnergySolve + solid_pnorm, J.backward()) against the
    C-ABI adjoint, including a backward after the context was re-solved.
"""
import numpy as np
import pytest

from hsolver_amd import abi, configs

from helpers import cell_volumes, gpu_ctx, oracle_ctx, rel_l2, solid_pnorm

pytestmark = pytest.mark.gpu

N = 24
TOL = 1e-13


@pytest.fixture(scope="module", params=[abi.HS_PREC_FP64, abi.HS_PREC_MIXED], ids=["fp64", "mixed"])
def pair(request, hs_lib, oracle_lib):
    case = configs.c4(N)
    p = abi.Params.from_buffer_copy(case.params)
    p.order = abi.HS_ORDER_RB
    p.precond = abi.HS_PRECOND_GMG
    p.precision = request.param
    print(f"[adjoint] c4 {N}^3 precision {request.param}", flush=True)
    g = gpu_ctx(hs_lib, case, p)
    g.init()
    for _ in range(20):
        g.step()
    o = oracle_ctx(oracle_lib, case)
    o.init()
    o.load_state(g.state())
    yield case, g, o
    g.close()
    o.close()


def test_energy_solve_matches_oracle(pair):
    case, g, o = pair
    ig = g.energy_solve(tol=TOL, maxit=300)
    io = o.energy_solve(tol=TOL, maxit=5000)
    assert ig["converged"] == 1 and ig["true_relres"] < 1e-12, ig
    assert io["converged"] == 1, io
    assert rel_l2(g.cells("tf"), o.cells("tf")) < 1e-10
    # the solution is the energy sweep's fixed point on the GPU too
    s, res = g.fluid_energy()
    assert res < 1e-9


def test_energy_operator_matches_oracle(pair):
    case, g, o = pair
    rng = np.random.default_rng(3)
    x, y = rng.standard_normal((N, N, N)), rng.standard_normal((N, N, N))
    Ax = g.energy_apply(x)
    Aty = g.energy_apply(y, transpose=True)
    assert rel_l2(Ax, o.energy_apply(x)) < 1e-13
    assert rel_l2(Aty, o.energy_apply(y, transpose=True)) < 1e-13
    assert abs(np.vdot(y, Ax) - np.vdot(Aty, x)) <= 1e-12 * abs(np.vdot(y, Ax))


def test_adjoint_matches_oracle_transpose_solve(pair):
    case, g, o = pair
    g.energy_solve(tol=TOL, maxit=300)
    o.energy_solve(tol=TOL, maxit=5000)
    solid = case.grid.cell_type.reshape(N, N, N) == 1
    _, dJdT = solid_pnorm(o.cells("tf"), solid, cell_volumes(case))
    dq_g, dk_g, info = g.adjoint_energy(dJdT, tol=TOL, maxit=300)
    dq_o, dk_o, info_o = o.adjoint_energy(dJdT, tol=TOL, maxit=5000)
    assert info["converged"] == 1 and info_o["converged"] == 1, (info, info_o)
    assert rel_l2(dq_g, dq_o) < 1e-9
    assert rel_l2(dk_g, dk_o) < 1e-8


def test_central_differences_on_seeded_solid_cells(pair):
    from hsolver_amd.adjoint import central_difference
    case, g, o = pair
    q0, k0 = g.cells("q").copy(), g.cells("k").copy()
    solid = case.grid.cell_type.reshape(N, N, N) == 1
    V = cell_volumes(case)
    g.energy_solve(tol=TOL, maxit=300)
    _, dJdT = solid_pnorm(g.cells("tf"), solid, V)
    dq, dk, _ = g.adjoint_energy(dJdT, tol=TOL, maxit=300)
    idx = np.flatnonzero(solid.ravel())
    pick = idx[(configs.splitmix64(44, 32) * idx.size).astype(np.int64)]
    worst = {"q": 0.0, "k": 0.0}
    for c in pick:
        cell = np.unravel_index(c, (N, N, N))
        for which, base, grad in (("q", q0, dq), ("k", k0, dk)):
            fd = central_difference(g, which, cell, 1e-6 * base[cell], q0, k0, solid, V, tol=TOL, maxit=300)
            err = abs(fd - grad[cell]) / abs(grad[cell])
            worst[which] = max(worst[which], err)
            assert err < 1e-5, (which, cell, fd, grad[cell])
    g.put_cells("q", q0)
    g.put_cells("k", k0)
    print(f"[adjoint] worst central-difference relative error {worst}", flush=True)


def test_central_differences_on_boundary_cells(pair):
    """dJ/dk on cells whose rows hold a boundary face (the Dirichlet-face and outflow
    branches of the gradient) and on fluid cells at the solid's faces, against central
    differences of the GPU forward solve; on the boundary the difference takes the
    unshifted product (the k change moves the boundary face's conductance)."""
    from hsolver_amd.adjoint import central_difference
    case, g, o = pair
    q0, k0 = g.cells("q").copy(), g.cells("k").copy()
    solid = case.grid.cell_type.reshape(N, N, N) == 1
    V = cell_volumes(case)
    g.energy_solve(tol=TOL, maxit=300)
    _, dJdT = solid_pnorm(g.cells("tf"), solid, V)
    _, dk, _ = g.adjoint_energy(dJdT, tol=TOL, maxit=300)
    lo = int(round(N * 0.375))
    m = N // 2
    cells = [(0, m, m), (N - 1, m, m), (m, 0, m), (m, N - 1, m),    # walls z = 0 / max, y = 0 / max
             (m, m, 0), (m, m + 1, N - 1),                           # inlet, outflow
             (lo, lo, lo - 1), (lo - 1, lo + 1, lo + 1)]             # fluid at the solid's faces
    # (edge cells are left out: their dJ/dk, ~1e-17 to 1e-21 here, is below what the
    # solve tolerance resolves)
    for cell in cells:
        fd = central_difference(g, "k", cell, 1e-6 * k0[cell], q0, k0, solid, V, tol=TOL, maxit=300)
        err = abs(fd - dk[cell]) / abs(dk[cell])
        assert err < 1e-5, (cell, fd, dk[cell])
    g.put_cells("q", q0)
    g.put_cells("k", k0)


def test_autograd_function(pair):
    import torch
    from hsolver_amd.adjoint import EnergySolve, EnergySolver
    from hsolver_amd.adjoint import solid_pnorm as torch_pnorm
    case, g, o = pair
    dev = torch.device("cuda:0")
    q0, k0 = g.cells("q").copy(), g.cells("k").copy()
    solver = EnergySolver(g, tol=TOL, maxit=300)
    solid = torch.from_numpy(case.grid.cell_type.reshape(N, N, N) == 1).to(dev)
    V = torch.from_numpy(cell_volumes(case)).to(dev)
    q = torch.from_numpy(q0).to(dev).requires_grad_(True)
    k = torch.from_numpy(k0).to(dev).requires_grad_(True)
    T = EnergySolve.apply(q, k, solver)
    J = torch_pnorm(T, solid, V)
    # a second forward with another source before the first backward
    q2 = torch.from_numpy(1.5 * q0).to(dev)
    EnergySolve.apply(q2, k.detach(), solver)
    J.backward()
    # reference: the C-ABI adjoint at the first state (host buffers)
    g.put_cells("q", q0)
    g.put_cells("k", k0)
    g.energy_solve(tol=TOL, maxit=300)
    Th = g.cells("tf")
    assert rel_l2(T.detach().cpu().numpy(), Th) < 1e-12
    Jh, dJdT = solid_pnorm(Th, case.grid.cell_type.reshape(N, N, N) == 1, cell_volumes(case))
    assert J.item() == pytest.approx(Jh, rel=1e-12)
    dq, dk, _ = g.adjoint_energy(dJdT, tol=TOL, maxit=300)
    assert rel_l2(q.grad.cpu().numpy(), dq) < 1e-10
    assert rel_l2(k.grad.cpu().numpy(), dk) < 1e-10


@pytest.mark.timeout(600)   # the oracle's share of this test is CPU work of 30-60 s
def test_c4_128_full_gradient_vs_oracle(hs_lib, oracle_lib):
    """The adjoint workload at full size: converged 128^3 flow (mixed-precision SIMPLE),
    frozen; forward solve and full dJ/dq, dJ/dk of both GPU precisions against the
    oracle's FP64 solves on the same state."""
    import os
    from hsolver_amd.adjoint import converge_flow
    os.environ["HS_ORACLE_PROGRESS"] = "1"
    n = 128
    case = configs.c4(n)
    p = abi.Params.from_buffer_copy(case.params)
    p.order = abi.HS_ORDER_RB
    p.precond = abi.HS_PRECOND_GMG
    p.precision = abi.HS_PREC_MIXED
    g = gpu_ctx(hs_lib, case, p)
    g.init()
    hist = converge_flow(g, 3000, 1e-9, log=lambda m: print(f"[adjoint] {m}", flush=True))
    print(f"[adjoint] flow: {len(hist)} iterations, mass imbalance {hist[-1]:.3e}", flush=True)
    state = g.state()
    o = oracle_ctx(oracle_lib, case)
    o.init()
    o.load_state(state)
    io = o.energy_solve(tol=TOL, maxit=20000)
    assert io["converged"] == 1, io
    To = o.cells("tf")
    solid = case.grid.cell_type.reshape(n, n, n) == 1
    _, dJdT = solid_pnorm(To, solid, cell_volumes(case))
    dq_o, dk_o, info_o = o.adjoint_energy(dJdT, tol=TOL, maxit=20000)
    assert info_o["converged"] == 1, info_o
    for prec in (abi.HS_PREC_MIXED, abi.HS_PREC_FP64):
        p.precision = prec
        h = gpu_ctx(hs_lib, case, p)
        h.init()
        h.load_state(state)
        ig = h.energy_solve(tol=TOL, maxit=400)
        assert ig["converged"] == 1, ig
        eT = rel_l2(h.cells("tf"), To)
        dq, dk, ia = h.adjoint_energy(dJdT, tol=TOL, maxit=400)
        assert ia["converged"] == 1, ia
        eq, ek = rel_l2(dq, dq_o), rel_l2(dk, dk_o)
        print(f"[adjoint] 128^3 precision {prec}: forward {ig['iters']} it, adjoint {ia['iters']} it; "
              f"rel-L2 T {eT:.2e}, dJ/dq {eq:.2e}, dJ/dk {ek:.2e}", flush=True)
        assert eT < 1e-10 and eq < 1e-9 and ek < 1e-8
        h.close()
    g.close()
    o.close()
