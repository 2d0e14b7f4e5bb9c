"""Shared test helpers: oracle / GPU contexts on the same case, small grids."""
import ctypes as C

import numpy as np
import scipy.sparse as sp

from hsolver_amd import abi, configs
from hsolver_amd.core import Context


def oracle_ctx(lib, case, params=None):
    p = params if params is not None else case.params
    ctx = Context(lib, "or", case.grid, p)
    if np.any(case.q):
        ctx.set_source(case.q)
    return ctx


def gpu_ctx(lib, case, params=None):
    p = params if params is not None else case.params
    ctx = Context(lib, "hs", case.grid, p)
    if np.any(case.q):
        ctx.set_source(case.q)
    return ctx


def small_ref_case(nx=12, ny=10, nz=8, block=(5, 7, 4, 6, 3, 5), stretch=True, order=abi.HS_ORDER_LEX):
    """Mode R cavity with one solid block on a mildly stretched grid."""
    def ax(n, s):
        t = np.linspace(0, 1, n + 1)
        return t + (0.08 * np.sin(np.pi * t) * s if stretch else 0.0)
    ct = configs.box_cells(nx, ny, nz, [block]) if block else np.zeros((nz, ny, nx), np.uint8)
    p = configs.mode_r_params()
    p.order = order
    return configs.Case("small_r", ax(nx, 1), ax(ny, 1), ax(nz, 1), ct, p)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def export_energy(ctx, n):
    """The oracle's Mode C energy operator on the current flow as scipy CSR + rhs."""
    lib = ctx.lib
    I64 = C.c_int64
    lib.or_export_energy_csr.argtypes = [C.c_void_p, C.POINTER(I64), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    nz = I64()
    assert lib.or_export_energy_csr(ctx.h, C.byref(nz), None, None, None, None) == 0
    rp = np.zeros(n + 1, np.int64)
    cols = np.zeros(nz.value, np.int64)
    vals = np.zeros(nz.value)
    rhs = np.zeros(n)
    assert lib.or_export_energy_csr(ctx.h, C.byref(nz), rp.ctypes.data, cols.ctypes.data, vals.ctypes.data,
                                    rhs.ctypes.data) == 0
    return sp.csr_matrix((vals, cols, rp), shape=(n, n)), rhs


def solid_pnorm(T, solid, V, p=8):
    """J = (sum over solid cells of T^p V)^(1/p) and dJ/dT (numpy, FP64)."""
    T = np.asarray(T, dtype=np.float64)
    S = np.sum(np.where(solid, T ** p * V, 0.0))
    J = S ** (1.0 / p)
    g = np.where(solid, J ** (1 - p) * T ** (p - 1) * V, 0.0)
    return J, g


def cell_volumes(case):
    g = case.grid
    dx, dy, dz = np.diff(g.xv), np.diff(g.yv), np.diff(g.zv)
    return dz[:, None, None] * dy[None, :, None] * dx[None, None, :]


def export_p(ctx):
    """The oracle's p' operator (fluid rows, initMap order) as scipy CSR + rhs + row cells."""
    lib = ctx.lib
    I64 = C.c_int64
    lib.or_export_csr.argtypes = [C.c_void_p, C.POINTER(I64), C.POINTER(I64), C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_void_p]
    nr, nz = I64(), I64()
    assert lib.or_export_csr(ctx.h, C.byref(nr), C.byref(nz), None, None, None, None, None) == 0
    rp = np.zeros(nr.value + 1, np.int64)
    cols = np.zeros(nz.value, np.int64)
    vals = np.zeros(nz.value)
    rhs = np.zeros(nr.value)
    cell = np.zeros(nr.value, np.int64)
    assert lib.or_export_csr(ctx.h, C.byref(nr), C.byref(nz), rp.ctypes.data, cols.ctypes.data, vals.ctypes.data,
                             rhs.ctypes.data, cell.ctypes.data) == 0
    return sp.csr_matrix((vals, cols, rp), shape=(nr.value, nr.value)), rhs, cell


def pack_reference(grid, st, mode_r=True, assembled=True):
    """The reference's output records (solver.h:160-213, OutputData fluid_solver.h:38-45)
    built in numpy from raw padded arrays `st` (Context.state()): per cell in global cell
    order (x fastest), the minimum vertex, the west / south / bottom face velocity, p, t
    and id.  Fluid cells: t from the fluid temperature with 293 on the domain-boundary
    cells in Mode R (fluid_solver.h:394-404), id = p' row number counted over fluid
    cells in the same order (initMap fluid_solver.h:1891-1905; 0 before the first
    assembly).  Solid cells: the solid temperature (Mode R) or the conjugate one, zero
    velocity and p, id -1."""
    nx, ny, nz = grid.nx, grid.ny, grid.nz
    ct = np.asarray(grid.cell_type).reshape(nz, ny, nx)
    fluid = ct == 0
    u = st["u"][1:nz + 1, 1:ny + 1, 1:nx + 1]
    v = st["v"][1:nz + 1, 1:ny + 1, 1:nx + 1]
    w = st["w"][1:nz + 1, 1:ny + 1, 1:nx + 1]
    tf = st["tf"][1:nz + 1, 1:ny + 1, 1:nx + 1]
    ts = st["ts"][1:nz + 1, 1:ny + 1, 1:nx + 1]
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    edge = (i == 0) | (j == 0) | (k == 0) | (i == nx - 1) | (j == ny - 1) | (k == nz - 1)
    if mode_r:
        t = np.where(fluid, np.where(edge, 293.0, tf), ts)
    else:
        t = tf.copy()
    rows = np.cumsum(fluid.ravel()) - 1
    ids = np.where(fluid.ravel(), rows if assembled else 0, -1).astype(np.int32)
    vel = np.stack([np.where(fluid, a, 0.0).ravel() for a in (u, v, w)], axis=1)
    loc = np.stack([np.asarray(grid.xv)[i.ravel()], np.asarray(grid.yv)[j.ravel()],
                    np.asarray(grid.zv)[k.ravel()]], axis=1)
    return {"loc": loc, "vel": vel, "p": np.where(fluid, st["p"], 0.0).ravel(), "t": t.ravel(), "id": ids}


def energy_balance(case, T, u, v, w, k, q):
    """Global energy balance of the Mode C conjugate energy equation, built from the
    physics: heat in (sum q V) against what leaves (outflow enthalpy rho u A T_cell on the
    x = max faces minus the inflow enthalpy rho u A T_wall on the x = 0 faces, plus
    conduction k A (T_cell - T_wall) / width through the Dirichlet ghosts of the inlet
    plane and the four walls; the zero-gradient outflow ghost conducts nothing).  On an
    exactly divergence-free flow the two are equal for the exact discrete solution.  A
    flow with a residual mass imbalance m_c per cell (net outward rho u A) adds
    sum T_c m_c: the upwind rows carry the inflow flux on their diagonal.  T, k, q:
    (nz, ny, nx) cell arrays; u, v, w: the padded face arrays.
    Returns (heat in, heat out, sum T_c m_c)."""
    p = case.params
    nx, ny, nz = case.shape
    g = case.grid
    dx, dy, dz = np.diff(g.xv), np.diff(g.yv), np.diff(g.zv)
    V = dz[:, None, None] * dy[None, :, None] * dx[None, None, :]
    Ayz = dz[:, None] * dy[None, :]
    Axz = dz[:, None] * dx[None, :]
    Axy = dy[:, None] * dx[None, :]
    Tw = p.t_wall
    heat_in = float(np.sum(q * V))
    out = p.rho * np.sum(u[1:nz + 1, 1:ny + 1, nx + 1] * Ayz * T[:, :, -1]) if p.outflow else 0.0
    out -= p.rho * np.sum(u[1:nz + 1, 1:ny + 1, 1] * Ayz) * Tw
    out += np.sum(k[:, :, 0] * Ayz * (T[:, :, 0] - Tw) / dx[0])
    if not p.outflow:
        out += np.sum(k[:, :, -1] * Ayz * (T[:, :, -1] - Tw) / dx[-1])
    out += np.sum(k[:, 0, :] * Axz * (T[:, 0, :] - Tw) / dy[0]) + np.sum(k[:, -1, :] * Axz * (T[:, -1, :] - Tw) / dy[-1])
    out += np.sum(k[0] * Axy * (T[0] - Tw) / dz[0]) + np.sum(k[-1] * Axy * (T[-1] - Tw) / dz[-1])
    uc = u[1:nz + 1, 1:ny + 1, 1:nx + 2]
    vc = v[1:nz + 1, 1:ny + 2, 1:nx + 1]
    wc = w[1:nz + 2, 1:ny + 1, 1:nx + 1]
    m = p.rho * ((uc[:, :, 1:] - uc[:, :, :-1]) * Ayz[:, :, None] + (vc[:, 1:, :] - vc[:, :-1, :]) * Axz[:, None, :]
                 + (wc[1:] - wc[:-1]) * Axy[None, :, :])
    return heat_in, float(out), float(np.sum(T * m))
