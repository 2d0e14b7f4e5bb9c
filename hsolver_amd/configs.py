"""Grid, geometry and parameter generators for the benchmark configurations.

Configurations (synthetic inputs of the shapes the benchmark names; seeds fixed):
  c1  64x16x16 channel on [0,4]x[0,1]^2, centred solid cube cells 31-34 / 7-10 (1-based),
      q = 1, k_s/k_f = 100, Re = 100, Pr = 0.7, alpha_u = 0.7, alpha_p = 0.3, tol 1e-8
  c2  256x64x64 channel, cube cells 121-136 / 25-40, q ~ U[0.5,1.5] per solid cell
      (seed 2), Re = 400
  c3  256^3 unit box, 27 seeded solid blocks, k ratio 1000, lid + inlet, Re = 1000
  c4  128^3 on the c2 geometry scaled to a unit cube (adjoint case)
  c5  512x256x256 channel, 64 seeded plate fins (20% of the cells), tanh wall-stretched
      y/z grid with cell aspect ratio up to 50, Re 2000, k ratio 1000
Reference-faithful inputs come from the .in files through hsolver_amd.mesh.

Random numbers use splitmix64 -> (x >> 11) * 2^-53 so every platform draws the same
values (no std::uniform_real_distribution / numpy default generator dependence).
"""
from __future__ import annotations

import numpy as np

from . import abi


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n uniforms in [0,1) from splitmix64 (Steele, Lea, Flood 2014)."""
    out = np.empty(n, dtype=np.float64)
    x = seed & 0xFFFFFFFFFFFFFFFF
    mask = 0xFFFFFFFFFFFFFFFF
    for i in range(n):
        x = (x + 0x9E3779B97F4A7C15) & mask
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z = z ^ (z >> 31)
        out[i] = (z >> 11) * (1.0 / 9007199254740992.0)
    return out


def uniform_axis(n: int, length: float) -> np.ndarray:
    return np.linspace(0.0, length, n + 1)


def tanh_axis(n: int, length: float, beta: float) -> np.ndarray:
    """Two-sided tanh wall clustering (Vinokur-style), symmetric about the centre."""
    s = np.linspace(-1.0, 1.0, n + 1)
    return 0.5 * length * (1.0 + np.tanh(beta * s) / np.tanh(beta))


def box_cells(nx, ny, nz, boxes) -> np.ndarray:
    """cell_type array (z, y, x) with solid = 1 inside 1-based inclusive boxes
    (the .in `mobs` convention, mesh.cpp:102-110)."""
    ct = np.zeros((nz, ny, nx), dtype=np.uint8)
    for (i0, i1, j0, j1, k0, k1) in boxes:
        ct[k0 - 1:k1, j0 - 1:j1, i0 - 1:i1] = 1
    return ct


class Case:
    """A grid + parameters + per-cell source, ready for hs_create / or_create."""

    def __init__(self, name, xv, yv, zv, cell_type, params, q=None, workload=""):
        self.name = name
        self.grid = abi.GridArrays(xv, yv, zv, cell_type)
        self.params = params
        nx, ny, nz = self.grid.nx, self.grid.ny, self.grid.nz
        self.q = np.zeros(nx * ny * nz) if q is None else np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        self.workload = workload

    @property
    def shape(self):
        return (self.grid.nx, self.grid.ny, self.grid.nz)


def config_params(re: float, pr: float = 0.7, k_ratio: float = 100.0, u_ref: float = 1.0,
                  length: float = 1.0, tol: float = 1e-8, inlet: bool = True, lid: bool = False) -> abi.Params:
    p = abi.params_default(abi.HS_MODE_CONFIG)
    p.nu = u_ref * length / re
    p.alpha_f = p.nu / pr
    p.k_ratio = k_ratio
    p.u_in = u_ref if inlet else 0.0
    p.outflow = 1 if inlet else 0
    p.lid = 1 if lid else 0
    p.u_lid = u_ref if lid else 0.0
    p.p_tol = tol
    p.p_maxit = 20000
    return p


def c1() -> Case:
    nx, ny, nz = 64, 16, 16
    ct = box_cells(nx, ny, nz, [(31, 34, 7, 10, 7, 10)])
    p = config_params(re=100.0, k_ratio=100.0)
    q = np.where(ct.reshape(-1) == 1, 1.0, 0.0)
    return Case("c1", uniform_axis(nx, 4.0), uniform_axis(ny, 1.0), uniform_axis(nz, 1.0), ct, p, q,
                "channel 64x16x16, solid cube, Re=100")


def c2(scale: int = 1) -> Case:
    """256x64x64 channel (scale=1); scale=2/4 give 128x32x32 / 64x16x16 for tests."""
    nx, ny, nz = 256 // scale, 64 // scale, 64 // scale
    lo, hi = (121 - 1) // scale + 1, 136 // scale
    jlo, jhi = (25 - 1) // scale + 1, 40 // scale
    ct = box_cells(nx, ny, nz, [(lo, hi, jlo, jhi, jlo, jhi)])
    p = config_params(re=400.0, k_ratio=100.0)
    solid = ct.reshape(-1) == 1
    r = splitmix64(2, int(solid.sum()))
    q = np.zeros(nx * ny * nz)
    q[solid] = 0.5 + r
    return Case("c2", uniform_axis(nx, 4.0), uniform_axis(ny, 1.0), uniform_axis(nz, 1.0), ct, p, q,
                f"channel {nx}x{ny}x{nz}, solid cube, Re=400")


def c3_blocks(n: int, seed: int = 3, nblocks: int = 27, smin: float = 0.05, smax: float = 0.15,
              target_fraction: float = 0.15):
    """27 seeded axis-aligned blocks in the unit box.

    Blocks of side U[0.05,0.15] cannot reach a 15% solid fraction (27 E[s^3] ~ 3.4%),
    so each block is elongated along x; the elongation factor starts at
    target / (27 E[s^3]) and grows in 5% steps until the snapped, overlapping blocks
    cover at least the target fraction of the cells.  Returns (boxes, fraction)."""
    r = splitmix64(seed, nblocks * 6)
    es3 = (smax ** 4 - smin ** 4) / (4 * (smax - smin))
    ex0 = target_fraction / (nblocks * es3)
    for step in range(60):
        ex = ex0 * (1.05 ** step)
        boxes = []
        for b in range(nblocks):
            sx = min(0.95, (smin + (smax - smin) * r[6 * b + 0]) * ex)
            sy = smin + (smax - smin) * r[6 * b + 1]
            sz = smin + (smax - smin) * r[6 * b + 2]
            cx = sx / 2 + (1 - sx) * r[6 * b + 3]
            cy = 0.1 + sy / 2 + (0.8 - sy) * r[6 * b + 4]   # keep blocks off the lid / floor
            cz = sz / 2 + (1 - sz) * r[6 * b + 5]
            i0 = max(1, int(round((cx - sx / 2) * n)) + 1)
            i1 = min(n, max(i0, int(round((cx + sx / 2) * n))))
            j0 = max(2, int(round((cy - sy / 2) * n)) + 1)
            j1 = min(n - 1, max(j0, int(round((cy + sy / 2) * n))))
            k0 = max(2, int(round((cz - sz / 2) * n)) + 1)
            k1 = min(n - 1, max(k0, int(round((cz + sz / 2) * n))))
            # keep the inlet and outlet columns fluid so the channel stays open
            i0, i1 = max(i0, 2), min(i1, n - 1)
            boxes.append((i0, i1, j0, j1, k0, k1))
        frac = float(box_cells(n, n, n, boxes).mean())
        if frac >= target_fraction:
            break
    return boxes, frac


def c3(n: int = 256) -> Case:
    boxes, _ = c3_blocks(n)
    ct = box_cells(n, n, n, boxes)
    p = config_params(re=1000.0, k_ratio=1000.0, lid=True, inlet=True)
    solid = ct.reshape(-1) == 1
    r = splitmix64(33, int(solid.sum())) if n <= 64 else None
    q = np.zeros(n ** 3)
    q[solid] = 1.0 if r is None else 0.5 + r
    return Case("c3", uniform_axis(n, 1.0), uniform_axis(n, 1.0), uniform_axis(n, 1.0), ct, p, q,
                f"unit box {n}^3, 27 blocks, lid+inlet Re=1000, k ratio 1000")


def c4(n: int = 128) -> Case:
    """c2 geometry (a centred cube of 1/4 the channel height) in a unit cube.

    Temperatures are rises above the inlet / wall value (t_wall = 0): the energy
    equation is linear, so this is the same problem as any other wall temperature,
    and the p-norm objective then measures the heating rather than the offset."""
    lo, hi = int(round(n * 0.375)) + 1, int(round(n * 0.625))
    ct = box_cells(n, n, n, [(lo, hi, lo, hi, lo, hi)])
    p = config_params(re=400.0, k_ratio=100.0)
    p.t_wall = 0.0
    solid = ct.reshape(-1) == 1
    r = splitmix64(4, int(solid.sum()))
    q = np.zeros(n ** 3)
    q[solid] = 0.5 + r
    return Case("c4", uniform_axis(n, 1.0), uniform_axis(n, 1.0), uniform_axis(n, 1.0), ct, p, q,
                f"unit cube {n}^3, centred heated cube, Re=400")


def max_cell_aspect(xv, yv, zv) -> float:
    """Largest ratio of a cell's longest to shortest side: a cell takes one width per
    axis, so the extreme is the widest width of one axis over the narrowest of another."""
    d = [np.diff(v) for v in (xv, yv, zv)]
    return float(max(d[a].max() / d[b].min() for a in range(3) for b in range(3) if a != b))


def c5_grid(nx=512, ny=256, nz=256, aspect=50.0):
    """[0,2] x [0,1]^2 with y and z tanh-clustered to both walls; the clustering factor
    is the one (bisection) whose largest cell aspect ratio equals `aspect`."""
    xv = uniform_axis(nx, 2.0)
    lo, hi = 0.1, 6.0
    for _ in range(60):
        beta = 0.5 * (lo + hi)
        yv, zv = tanh_axis(ny, 1.0, beta), tanh_axis(nz, 1.0, beta)
        if max_cell_aspect(xv, yv, zv) > aspect:
            hi = beta
        else:
            lo = beta
    return xv, tanh_axis(ny, 1.0, lo), tanh_axis(nz, 1.0, lo), lo


def c5_fins(nx, ny, nz, nfins=64, seed=5, target_fraction=0.20):
    """Plate fins standing on the y = 0 wall: height 0.6 of the channel, seeded x / z
    positions, thickness nx/64 cells in x; the z extent grows (5% steps) until the fins
    cover `target_fraction` of the cells.  Inlet and outlet columns stay fluid.
    Returns (boxes, fraction)."""
    r = splitmix64(seed, nfins * 2)
    fin_h = int(round(0.6 * ny))
    tx = max(1, nx // 64)
    tz = max(1, int(target_fraction * nx * nz / (nfins * tx)))
    for _ in range(80):
        boxes = []
        for f in range(nfins):
            i0 = 2 + int(r[2 * f] * (nx - tx - 3))
            k0 = 1 + int(r[2 * f + 1] * (nz - tz))
            boxes.append((i0, i0 + tx - 1, 1, fin_h, k0, min(nz, k0 + tz - 1)))
        # every fin spans y = 1..fin_h, so the solid fraction is that of the x-z footprint
        foot = np.zeros((nz, nx), dtype=bool)
        for (a0, a1, _, _, b0, b1) in boxes:
            foot[b0 - 1:b1, a0 - 1:a1] = True
        frac = float(foot.mean()) * fin_h / ny
        if frac >= target_fraction:
            break
        tz = min(nz, int(tz * 1.05) + 1)
    return boxes, frac


def c5(nx=512, ny=256, nz=256, nfins=64, seed=5, aspect=50.0) -> Case:
    """512x256x256 channel, wall-stretched y/z grid (cell aspect ratio up to 50), 64
    seeded heated plate fins on the floor (20% of the cells solid), Re 2000, k ratio 1000."""
    xv, yv, zv, beta = c5_grid(nx, ny, nz, aspect)
    boxes, _ = c5_fins(nx, ny, nz, nfins, seed)
    ct = box_cells(nx, ny, nz, boxes)
    p = config_params(re=2000.0, k_ratio=1000.0, length=1.0)
    p.t_wall = 0.0                       # temperatures as rises above the inlet / wall value
    q = np.where(ct.reshape(-1) == 1, 1.0, 0.0)
    return Case("c5", xv, yv, zv, ct, p, q, f"channel {nx}x{ny}x{nz}, {nfins} fins, Re=2000, aspect<={aspect:g}")


def mode_r_params() -> abi.Params:
    return abi.params_default(abi.HS_MODE_REF)
