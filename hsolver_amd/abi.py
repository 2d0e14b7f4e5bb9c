"""ctypes mirror of include/hsolver.h (structs, constants) and the library loader.

The struct layouts here must match the header field for field; tests check
ctypes.sizeof against the C sizes the library reports (hs_abi_sizes).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HS_OK, HS_E_ARG, HS_E_HIP, HS_E_NOCONV, HS_E_NAN, HS_E_IO, HS_E_STATE, HS_E_NOMEM = 0, -1, -2, -3, -4, -5, -6, -7
HS_MODE_REF, HS_MODE_CONFIG = 0, 1
HS_ORDER_LEX, HS_ORDER_RB = 0, 1
HS_PRECOND_NONE, HS_PRECOND_GMG, HS_PRECOND_AMG = 0, 1, 2
HS_PREC_FP64, HS_PREC_MIXED = 0, 1
HS_MEM_HOST, HS_MEM_DEVICE = 0, 1

(HS_A_U, HS_A_V, HS_A_W, HS_A_UP, HS_A_VP, HS_A_WP, HS_A_P, HS_A_PTEMP, HS_A_PPRIME,
 HS_A_TF, HS_A_TS, HS_A_Q, HS_A_K, HS_A_B) = range(14)
ARRAY_NAMES = ["u", "v", "w", "up", "vp", "wp", "p", "ptemp", "pprime", "tf", "ts", "q", "k", "b"]


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("xv", C.POINTER(C.c_double)), ("yv", C.POINTER(C.c_double)),
                ("zv", C.POINTER(C.c_double)), ("cell_type", C.POINTER(C.c_uint8))]


class Params(C.Structure):
    _fields_ = [("mode", C.c_int32), ("order", C.c_int32),
                ("rho", C.c_double), ("nu", C.c_double), ("alpha_f", C.c_double),
                ("k_ratio", C.c_double), ("u_lid", C.c_double), ("u_in", C.c_double),
                ("outflow", C.c_int32), ("lid", C.c_int32),
                ("relax_u", C.c_double), ("relax_p", C.c_double), ("relax_t", C.c_double),
                ("relax_ts", C.c_double),
                ("mom_tol", C.c_double), ("t_tol", C.c_double), ("ts_tol", C.c_double),
                ("mom_sweeps_max", C.c_int32), ("t_sweeps_max", C.c_int32),
                ("ts_sweeps_max", C.c_int32),
                ("p_tol", C.c_double), ("p_maxit", C.c_int32), ("freeze_p_matrix", C.c_int32),
                ("precond", C.c_int32), ("precision", C.c_int32),
                ("mg_levels", C.c_int32), ("mg_pre", C.c_int32), ("mg_post", C.c_int32),
                ("mg_coarse_sweeps", C.c_int32), ("mg_dense_coarse", C.c_int32),
                ("t_fluid_init", C.c_double), ("t_solid_init", C.c_double), ("t_wall", C.c_double),
                ("fail_on_noconv", C.c_int32), ("mg_dense_refine", C.c_int32),
                ("mg_sweep_rows", C.c_int32), ("mg_pre0", C.c_int32), ("mg_post0", C.c_int32),
                ("mg_cycle", C.c_int32), ("mg_omega", C.c_float), ("mg_omega_coarse", C.c_float),
                ("e_omega", C.c_float), ("mg_wide_offsets", C.c_int32), ("e_pre", C.c_int32), ("e_post", C.c_int32),
                ("p_warm", C.c_int32), ("e_precond", C.c_int32), ("amg_sweeps0", C.c_int32),
                ("amg_sweeps", C.c_int32), ("amg_theta", C.c_float), ("amg_alpha", C.c_float),
                ("amg_omega", C.c_float), ("amg_pad", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("mom_sweeps", C.c_int32 * 3), ("t_sweeps", C.c_int32), ("ts_sweeps", C.c_int32),
                ("p_iters", C.c_int32), ("p_converged", C.c_int32), ("pad0", C.c_int32),
                ("p_relres", C.c_double), ("p_true_relres", C.c_double),
                ("mass_imbalance", C.c_double), ("energy_residual", C.c_double),
                ("heat_in", C.c_double), ("mom_residual", C.c_double * 3),
                ("t_residual", C.c_double), ("ts_residual", C.c_double)]

    def as_dict(self) -> dict:
        return {
            "mom_sweeps": list(self.mom_sweeps), "t_sweeps": self.t_sweeps,
            "ts_sweeps": self.ts_sweeps, "p_iters": self.p_iters, "p_converged": self.p_converged,
            "p_relres": self.p_relres, "p_true_relres": self.p_true_relres,
            "mass_imbalance": self.mass_imbalance, "energy_residual": self.energy_residual,
            "heat_in": self.heat_in, "mom_residual": list(self.mom_residual),
            "t_residual": self.t_residual, "ts_residual": self.ts_residual,
        }


class SolveInfo(C.Structure):
    _fields_ = [("iters", C.c_int32), ("converged", C.c_int32), ("breakdown", C.c_int32),
                ("restarts", C.c_int32), ("relres", C.c_double), ("true_relres", C.c_double)]

    def as_dict(self) -> dict:
        return {"iters": self.iters, "converged": self.converged, "breakdown": self.breakdown,
                "restarts": self.restarts, "relres": self.relres, "true_relres": self.true_relres}


def params_default(mode: int) -> Params:
    """Python restatement of hs_params_default (checked against the C one in tests).

    Mode R constants: fluid_solver.h:21-37, mobs_solver.h:15-29, solver.h:52,
    fluid_solver.h:1871-1875.
    """
    p = Params()
    p.mode = mode
    p.order = HS_ORDER_LEX
    p.rho = 1.0
    p.nu = 0.001
    p.alpha_f = 0.001
    p.k_ratio = 1.0
    p.u_lid = 0.1
    p.u_in = 0.0
    p.outflow = 0
    p.lid = 0
    p.relax_u = p.relax_p = p.relax_t = p.relax_ts = 0.3
    p.mom_tol = 2e-3
    p.t_tol = 2e-3
    p.ts_tol = 1e-3
    p.mom_sweeps_max = p.t_sweeps_max = p.ts_sweeps_max = 100
    p.p_tol = 1e-5
    p.p_maxit = 1000
    p.freeze_p_matrix = 1
    p.precond = HS_PRECOND_NONE
    p.precision = HS_PREC_FP64
    p.mg_levels = 0
    p.mg_pre = -1
    p.mg_post = -1
    p.mg_coarse_sweeps = 40
    p.mg_dense_coarse = 0
    p.mg_dense_refine = 1
    p.mg_sweep_rows = 0
    p.mg_pre0 = -1
    p.mg_post0 = -1
    p.mg_cycle = 1
    p.mg_omega = 0.0
    p.mg_omega_coarse = 0.0
    p.e_pre = p.e_post = -1
    p.e_omega = 0.0
    p.t_fluid_init = 293.0
    p.t_solid_init = 333.0
    p.t_wall = 293.0
    p.fail_on_noconv = 0
    if mode == HS_MODE_CONFIG:
        p.freeze_p_matrix = 0
        p.relax_u = 0.7
        p.relax_p = 0.3
        p.relax_t = 1.0
        p.relax_ts = 1.0
        p.mom_sweeps_max = 1
        p.t_sweeps_max = 1
        p.ts_sweeps_max = 0
        p.mom_tol = 0.0
        p.t_tol = 0.0
        p.p_tol = 1e-8
        p.outflow = 1
        p.u_lid = 0.0
        p.u_in = 1.0
    return p


def padded_sizes(nx: int, ny: int, nz: int) -> dict:
    nu = (nx + 3) * (ny + 2) * (nz + 2)
    nv = (nx + 2) * (ny + 3) * (nz + 2)
    nw = (nx + 2) * (ny + 2) * (nz + 3)
    npc = (nx + 2) * (ny + 2) * (nz + 2)
    nc = nx * ny * nz
    return {"u": nu, "v": nv, "w": nw, "up": nu, "vp": nv, "wp": nw, "p": nc, "ptemp": npc,
            "pprime": npc, "tf": npc, "ts": npc, "q": nc, "k": nc, "b": npc}


def padded_shape(name: str, nx: int, ny: int, nz: int) -> tuple:
    """(z, y, x) numpy shape of an array in its layout."""
    if name in ("u", "up"):
        return (nz + 2, ny + 2, nx + 3)
    if name in ("v", "vp"):
        return (nz + 2, ny + 3, nx + 2)
    if name in ("w", "wp"):
        return (nz + 3, ny + 2, nx + 2)
    if name in ("p", "q", "k"):
        return (nz, ny, nx)
    return (nz + 2, ny + 2, nx + 2)


class GridArrays:
    """Owns numpy buffers behind a Grid struct (keeps them alive)."""

    def __init__(self, xv, yv, zv, cell_type):
        self.xv = np.ascontiguousarray(xv, dtype=np.float64)
        self.yv = np.ascontiguousarray(yv, dtype=np.float64)
        self.zv = np.ascontiguousarray(zv, dtype=np.float64)
        self.nx, self.ny, self.nz = len(self.xv) - 1, len(self.yv) - 1, len(self.zv) - 1
        ct = np.ascontiguousarray(cell_type, dtype=np.uint8).reshape(-1)
        if ct.size != self.nx * self.ny * self.nz:
            raise ValueError("cell_type size does not match the vertex arrays")
        self.cell_type = ct
        self.struct = Grid(self.nx, self.ny, self.nz,
                           self.xv.ctypes.data_as(C.POINTER(C.c_double)),
                           self.yv.ctypes.data_as(C.POINTER(C.c_double)),
                           self.zv.ctypes.data_as(C.POINTER(C.c_double)),
                           self.cell_type.ctypes.data_as(C.POINTER(C.c_uint8)))


def dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libhsolver_amd.so"


def load_library(path: os.PathLike | None = None) -> C.CDLL:
    """Load the HIP library.  Raises if it has not been built: there is no CPU path.

    A process that also uses torch.cuda must import torch first: torch ships its own
    HIP runtime, and the library binds to an already loaded one (loading the library
    first maps /opt/rocm's runtime beside torch's, and torch then finds no device)."""
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: run __graft_entry__.build() (the HIP library is required)")
    lib = C.CDLL(str(p))
    declare(lib, "hs")
    return lib


def declare(lib: C.CDLL, prefix: str) -> None:
    """Attach argtypes for the hs_/or_ entry points that exist in lib."""
    P, I32, I64, D = C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_double)
    PS = C.POINTER(Stats)
    PI = C.POINTER(SolveInfo)
    D_ = C.c_double
    sig = {
        "create": (C.c_int, [C.POINTER(Grid), C.POINTER(Params), C.POINTER(P)]),
        "destroy": (None, [P]),
        "array_size": (I64, [P, I32]),
        "get_params": (C.c_int, [P, C.POINTER(Params)]),
        "init": (C.c_int, [P]),
        "fluid_solve": (C.c_int, [P, PS]),
        "solid_solve": (C.c_int, [P, PS]),
        "exchange": (C.c_int, [P]),
        "set_for_next": (C.c_int, [P]),
        "step": (C.c_int, [P, PS]),
        "run": (C.c_int, [P, I32, PS]),
        "momentum": (C.c_int, [P, I32, C.POINTER(I32), D]),
        "pressure_correction": (C.c_int, [P, PS]),
        "correct": (C.c_int, [P]),
        "fluid_energy": (C.c_int, [P, C.POINTER(I32), D]),
        "pressure_apply": (C.c_int, [P, D, D]),
        "pressure_solve": (C.c_int, [P, D, D, PS]),
        "get_array": (C.c_int, [P, I32, D, I64]),
        "put_array": (C.c_int, [P, I32, D, I64]),
        "get_fields": (C.c_int, [P, D, D, D, D, D]),
        "put_fields": (C.c_int, [P, D, D, D, D, D]),
        "get_output": (C.c_int, [P, D, D, D, D, C.POINTER(I32)]),
        "set_source": (C.c_int, [P, D]),
        "set_conductivity": (C.c_int, [P, D]),
        "cell_bc": (C.c_int, [P]),
        "energy_solve": (C.c_int, [P, D_, I32, PI]),
        "adjoint_energy": (C.c_int, [P, P, P, P, D_, I32, I32, PI]),
        "energy_apply": (C.c_int, [P, I32, P, P, I32]),
        "energy_solve_rhs": (C.c_int, [P, I32, P, P, D_, I32, I32, PI]),
        "get_cells": (C.c_int, [P, I32, P, I32]),
        "put_cells": (C.c_int, [P, I32, P, I32]),
        "strerror": (C.c_char_p, [C.c_int]),
        "params_default": (None, [I32, C.POINTER(Params)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, f"{prefix}_{name}", None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
