"""Discrete adjoint of the conjugate energy equation behind torch.autograd.

On a converged, frozen flow the Mode C energy equation is one linear system
A(k, u) T = b(q).  `EnergySolve.apply(q, k, solver)` returns T (every cell) and
carries the adjoint: its backward solves A^T lambda = dJ/dT on the GPU (hand-written
BiCGStab + multigrid kernels in libhsolver_amd) and returns
    dJ/dq = lambda V,     dJ/dk = -lambda^T (dA/dk T - db/dk)
in FP64.  Any torch objective of T therefore gets exact discrete gradients with
respect to the per-cell heat source and conductivity; `solid_pnorm` is the p-norm of
solid temperature used by the adjoint workload.

Tensors are float64 CUDA tensors of shape (nz, ny, nx) on the library's device; the
library reads and writes them in place through device pointers (HS_MEM_DEVICE).
There is no CPU path: without the HIP library the solver cannot be created.

Run as a module for the 128^3 adjoint workload (converged flow on the c4 geometry,
forward solve, adjoint, central differences on seeded solid cells):
    python -m hsolver_amd.adjoint --n 128
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time

import numpy as np
import torch

from . import abi, configs, core


def hip_runtimes() -> list:
    """Paths of the HIP runtime libraries mapped into this process."""
    try:
        with open("/proc/self/maps") as f:
            return sorted({ln.split()[-1] for ln in f if "libamdhip64" in ln})
    except OSError:
        return []


class SolveError(RuntimeError):
    """A forward or adjoint energy solve stopped before its tolerance (maxit reached or a
    BiCGStab breakdown): the temperature or the gradients it would return are not solutions."""

    def __init__(self, what: str, info: dict):
        super().__init__(f"{what} did not converge: {info}")
        self.info = info


def _require_converged(what: str, info: dict, tol: float, maxit: int) -> None:
    if not info["converged"] or info["breakdown"]:
        raise SolveError(f"{what} (tol {tol:g}, maxit {maxit})", info)


class EnergySolver:
    """A GPU context whose flow is frozen; q and k are the differentiable inputs.

    torch and libhsolver_amd must share one HIP runtime: import torch before the library
    is first loaded (torch bundles its own runtime, and the library then binds to it)."""

    def __init__(self, ctx: core.Context, tol: float = 1e-12, maxit: int = 500):
        rts = hip_runtimes()
        if len(rts) > 1:
            raise RuntimeError("two HIP runtimes are loaded (" + ", ".join(rts) + "): import torch before "
                               "libhsolver_amd is loaded so that both use torch's runtime")
        self.ctx = ctx
        self.tol = tol
        self.maxit = maxit
        self.shape = (ctx.nz, ctx.ny, ctx.nx)
        self.version = 0          # bumped by every forward solve
        self.last_info = None

    def _dev(self, t: torch.Tensor) -> int:
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and tuple(t.shape) == self.shape):
            raise ValueError(f"expected a contiguous float64 CUDA tensor of shape {self.shape}")
        return t.data_ptr()

    def solve(self, q: torch.Tensor, k: torch.Tensor) -> torch.Tensor:
        torch.cuda.current_stream().synchronize()
        self.ctx._call("put_cells", abi.HS_A_Q, C.c_void_p(self._dev(q)), abi.HS_MEM_DEVICE)
        self.ctx._call("put_cells", abi.HS_A_K, C.c_void_p(self._dev(k)), abi.HS_MEM_DEVICE)
        info = abi.SolveInfo()
        self.ctx._call("energy_solve", C.c_double(self.tol), self.maxit, C.byref(info))
        self.last_info = info.as_dict()
        _require_converged("energy solve", self.last_info, self.tol, self.maxit)
        T = torch.empty(self.shape, dtype=torch.float64, device=q.device)
        self.ctx._call("get_cells", abi.HS_A_TF, C.c_void_p(T.data_ptr()), abi.HS_MEM_DEVICE)
        self.version += 1
        return T

    def adjoint(self, gT: torch.Tensor):
        torch.cuda.current_stream().synchronize()
        dq = torch.empty(self.shape, dtype=torch.float64, device=gT.device)
        dk = torch.empty_like(dq)
        info = abi.SolveInfo()
        self.ctx._call("adjoint_energy", C.c_void_p(self._dev(gT)), C.c_void_p(dq.data_ptr()),
                       C.c_void_p(dk.data_ptr()), C.c_double(self.tol), self.maxit, abi.HS_MEM_DEVICE,
                       C.byref(info))
        self.last_adjoint_info = info.as_dict()
        _require_converged("adjoint solve", self.last_adjoint_info, self.tol, self.maxit)
        return dq, dk


class EnergySolve(torch.autograd.Function):
    """T = A(k)^-1 b(q) on the frozen flow; backward by the discrete adjoint."""

    @staticmethod
    def forward(fctx, q, k, solver: EnergySolver):
        q = q.detach().contiguous()
        k = k.detach().contiguous()
        T = solver.solve(q, k)
        fctx.solver = solver
        fctx.version = solver.version
        fctx.save_for_backward(q, k)
        return T

    @staticmethod
    def backward(fctx, gT):
        solver = fctx.solver
        if solver.version != fctx.version:
            # the context was re-solved since this forward: restore its q, k, T first
            q, k = fctx.saved_tensors
            solver.solve(q, k)
            fctx.version = solver.version
        dq, dk = solver.adjoint(gT.detach().to(torch.float64).contiguous())
        return dq, dk, None


def solid_pnorm(T: torch.Tensor, solid: torch.Tensor, V: torch.Tensor, p: float = 8.0) -> torch.Tensor:
    """J = (sum over solid cells of T^p V)^(1/p) (a smooth maximum of solid temperature)."""
    return torch.sum(torch.where(solid, T.clamp_min(0.0) ** p * V, torch.zeros_like(T))) ** (1.0 / p)


def cell_volumes(case) -> np.ndarray:
    g = case.grid
    dx, dy, dz = np.diff(g.xv), np.diff(g.yv), np.diff(g.zv)
    return dz[:, None, None] * dy[None, :, None] * dx[None, None, :]


def pnorm_difference(T: np.ndarray, dT: np.ndarray, solid: np.ndarray, V: np.ndarray, p: float = 8.0) -> float:
    """J(T + dT) - J(T) for J = (sum_solid T^p V)^(1/p), evaluated without cancellation
    (expm1 / log1p); solid temperatures must be positive."""
    t, dt, v = T[solid], dT[solid], V[solid]
    S = np.sum(t ** p * v)
    dS = np.sum(t ** p * np.expm1(p * np.log1p(dt / t)) * v)
    return float(S ** (1.0 / p) * np.expm1(np.log1p(dS / S) / p))


def central_difference(ctx: core.Context, which: str, cell: tuple, h: float, q0: np.ndarray, k0: np.ndarray,
                       solid: np.ndarray, V: np.ndarray, tol: float = 1e-13, maxit: int = 400,
                       p: float = 8.0, shift: bool | None = None) -> float:
    """(J(x + h) - J(x - h)) / 2h for x = q or k of one cell.

    The two states are not differenced directly: with A_- T_- = b_- and A_+ T_+ = b_+,
        A_+ (T_+ - T_-) = (b_+ - b_-) - (A_+ - A_-) T_-,
    where b_+ - b_- = (q_+ - q_-) V at the cell for a q change.  b does not depend on k
    away from the boundary, nor at a boundary cell when the inlet / wall temperature
    t_wall is 0 (its ghost term is a_face * 0); other boundary cases raise ValueError.
    A change of k at a cell away from the boundary keeps every row sum of A, so
    (A_+ - A_-) T_- = (A_+ - A_-)(T_- - T_c), applied at k_+ and k_- to the field shifted
    by the cell's own temperature: its rounding is then relative to the local
    temperature differences, not to T.  At a cell on the domain boundary the k change
    also moves its boundary face's conductance and with it the row sum, so the product
    is taken on T_- itself (shift=None picks by the cell's position; True / False
    force it).  This is the exact nonlinear difference, with an error relative to the
    difference itself instead of to J."""
    nz, ny, nx = q0.shape
    on_boundary = any(c == 0 or c == n - 1 for c, n in zip(cell, (nz, ny, nx)))
    if shift is None:
        shift = not on_boundary
    if which == "k" and on_boundary and ctx.params.t_wall != 0.0:
        raise ValueError("a k difference at a boundary cell needs t_wall = 0 (b then does not depend on k)")
    qm, qp, km, kp = q0.copy(), q0.copy(), k0.copy(), k0.copy()
    if which == "q":
        qm[cell] -= h
        qp[cell] += h
    else:
        km[cell] -= h
        kp[cell] += h
    ctx.put_cells("q", qm)
    ctx.put_cells("k", km)
    info = ctx.energy_solve(tol, maxit)
    if not info["converged"]:
        raise RuntimeError(f"energy solve did not converge: {info}")
    Tm = ctx.cells("tf")
    if which == "q":
        r = np.zeros_like(Tm)
        r[cell] = (qp[cell] - qm[cell]) * V[cell]
    else:
        shifted = Tm - Tm[cell] if shift else Tm
        r = ctx.energy_apply(shifted)            # A_- (T_- - T_c), or A_- T_-
    ctx.put_cells("q", qp)
    ctx.put_cells("k", kp)
    if which == "k":
        r = r - ctx.energy_apply(shifted)        # -(A_+ - A_-) T_-
    dT, info = ctx.energy_solve_rhs(r, tol=tol, maxit=maxit)
    if not info["converged"]:
        raise RuntimeError(f"difference solve did not converge: {info}")
    return pnorm_difference(Tm, dT, solid, V, p) / (2.0 * h)


def converge_flow(ctx: core.Context, max_steps: int, mass_tol: float, log=None) -> list:
    """SIMPLE iterations until the continuity residual falls below mass_tol."""
    hist = []
    for s in range(max_steps):
        st = ctx.step()
        hist.append(st["mass_imbalance"])
        if log and (s % 50 == 0):
            log(f"step {s}: mass imbalance {st['mass_imbalance']:.3e}, p' iters {st['p_iters']}")
        if st["mass_imbalance"] < mass_tol:
            break
    return hist


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="adjoint of the solid-temperature p-norm on the c4 geometry")
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--steps", type=int, default=3000, help="max SIMPLE iterations for the frozen flow")
    ap.add_argument("--mass-tol", type=float, default=1e-9)
    ap.add_argument("--cells", type=int, default=32, help="seeded solid cells for central differences")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--energy-precond", choices=["gmg", "amg"], default="gmg",
                    help="preconditioner of the energy and adjoint BiCGStab: geometric or aggregation multigrid")
    args = ap.parse_args(argv)

    def log(m):
        print(f"[adjoint] {m}", file=sys.stderr, flush=True)

    case = configs.c4(args.n)
    p = abi.Params.from_buffer_copy(case.params)
    p.order = abi.HS_ORDER_RB
    p.precond = abi.HS_PRECOND_GMG
    p.precision = abi.HS_PREC_MIXED if args.precision == "mixed" else abi.HS_PREC_FP64
    p.e_precond = abi.HS_PRECOND_AMG if args.energy_precond == "amg" else abi.HS_PRECOND_GMG
    ctx = core.create(case.grid, p, case.q)
    ctx.init()
    t0 = time.perf_counter()
    hist = converge_flow(ctx, args.steps, args.mass_tol, log)
    t_flow = time.perf_counter() - t0
    log(f"flow: {len(hist)} SIMPLE iterations in {t_flow:.1f} s, mass imbalance {hist[-1]:.3e}")

    dev = torch.device("cuda:0")
    shape = (args.n,) * 3
    q = torch.from_numpy(case.q.reshape(shape)).to(dev)
    k = torch.from_numpy(ctx.cells("k")).to(dev)
    solid = torch.from_numpy(case.grid.cell_type.reshape(shape) == 1).to(dev)
    V = torch.from_numpy(cell_volumes(case)).to(dev)
    solver = EnergySolver(ctx, tol=1e-13, maxit=400)

    qv, kv = q.clone().requires_grad_(True), k.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    T = EnergySolve.apply(qv, kv, solver)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    fwd_info = solver.last_info
    J = solid_pnorm(T, solid, V)
    t0 = time.perf_counter()
    J.backward()
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    adj_info = solver.last_adjoint_info
    gq, gk = qv.grad, kv.grad
    log(f"J = {J.item():.12e}; forward {fwd_info}, {t_fwd * 1e3:.1f} ms; adjoint {adj_info}, {t_adj * 1e3:.1f} ms")

    # central differences, relative step 1e-6, on seeded solid cells
    q0, k0 = ctx.cells("q").copy(), ctx.cells("k").copy()
    solid_np = case.grid.cell_type.reshape(shape) == 1
    V_np = cell_volumes(case)
    idx = np.flatnonzero(solid_np.ravel())
    pick = idx[(configs.splitmix64(44, args.cells) * idx.size).astype(np.int64)]
    errs_q, errs_k = [], []
    gq_np, gk_np = gq.cpu().numpy(), gk.cpu().numpy()
    t0 = time.perf_counter()
    for c in pick:
        cell = np.unravel_index(c, shape)
        for which, base, grad, errs in (("q", q0, gq_np, errs_q), ("k", k0, gk_np, errs_k)):
            fd = central_difference(ctx, which, cell, 1e-6 * base[cell], q0, k0, solid_np, V_np)
            errs.append(abs(fd - grad[cell]) / abs(grad[cell]))
    log(f"central differences on {args.cells} cells in {time.perf_counter() - t0:.1f} s: "
        f"max rel err q {max(errs_q):.2e}, k {max(errs_k):.2e}")
    out = {"workload": f"c4 {args.n}^3 adjoint", "flow_iters": len(hist), "mass_imbalance": hist[-1],
           "J": J.item(), "forward": fwd_info, "adjoint": adj_info,
           "forward_ms": round(t_fwd * 1e3, 2), "adjoint_ms": round(t_adj * 1e3, 2),
           "fd_cells": int(args.cells), "fd_max_rel_err_q": max(errs_q), "fd_max_rel_err_k": max(errs_k),
           "precision": args.precision, "energy_precond": args.energy_precond}
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
