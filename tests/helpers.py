"""Shared test helpers: oracle / GPU contexts on the same case, small grids."""
import numpy as np

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
