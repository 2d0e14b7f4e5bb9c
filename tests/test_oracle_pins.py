"""Pins the oracle functions that the GPU is compared against but that no closed form
checked before: the velocity correction and pressure update, the velocity boundary
conditions (Mode R ghosts, wall faces, lid), setForNext's p_temp refresh, the Mode R
fluid energy point update, the solid relaxation, and the Mode C boundary values
(inlet, outflow, lid, walls) through a developed duct profile, the global mass
balance and the global energy balance.

Every expected value here is built in numpy from the reference's formulas on
non-trivial data (stretched grids, random fields): a wrong neighbour, index, spacing,
area, face-velocity choice, relaxation, guard or boundary range in the oracle makes
one of them fail.  The RB order lets the energy and solid checks see every cell: the
first colour's cells use only old neighbours, the second colour's only first-colour
values that the test has already computed.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from hsolver_amd import abi, configs

from helpers import oracle_ctx, small_ref_case


def widths(v):
    """Padded per-cell widths of one axis (index 1..n), with the reference's fallbacks
    (fluid_solver.h:1248-1261): w[i] = own width, wm[i] = previous cell's width (own at
    i = 1)."""
    d = np.diff(np.asarray(v, dtype=np.float64))
    n = d.size
    w = np.zeros(n + 2)
    w[1:n + 1] = d
    wm = np.zeros(n + 2)
    wm[1:n + 1] = np.concatenate([[d[0]], d[:-1]])
    return w, wm


def interior(nx, ny, nz):
    return np.meshgrid(np.arange(1, nz + 1), np.arange(1, ny + 1), np.arange(1, nx + 1), indexing="ij")


def fluid_mask(case):
    nx, ny, nz = case.shape
    m = np.zeros((nz + 2, ny + 2, nx + 2), bool)
    m[1:-1, 1:-1, 1:-1] = case.grid.cell_type.reshape(nz, ny, nx) == 0
    return m


# ------------------------------------------------------ velocity correction ---

@pytest.mark.parametrize("mode", [abi.HS_MODE_REF, abi.HS_MODE_CONFIG])
def test_velocity_correction_and_pressure_update(oracle_lib, mode):
    """fluid_solver.h:1459-1571: u += (p'_W - p'_P) A / ap on the faces the loop visits,
    ap = 0 guarded to 1e-10; :1574-1622: p += relax_p p' and p_temp = p on fluid cells,
    p' = 0.  Mode R visits every face whose plus-side cell is fluid (interior faces
    1..n of the padded index); Mode C the faces between two fluid cells."""
    case = small_ref_case()
    case.params.mode = mode
    if mode == abi.HS_MODE_CONFIG:
        case.params.outflow = 0
    ctx = oracle_ctx(oracle_lib, case)
    nx, ny, nz = case.shape
    rng = np.random.default_rng(17)
    fl = fluid_mask(case)
    xp, _ = widths(case.grid.xv)
    yp, _ = widths(case.grid.yv)
    zp, _ = widths(case.grid.zv)
    vel0, ap = {}, {}
    for nm in "uvw":
        vel0[nm] = rng.standard_normal(ctx.get(nm).shape)
        a = rng.uniform(0.5, 2.0, ctx.get(nm + "p").shape)
        ap[nm] = a
    ap["u"][2, 3, 3] = 0.0                      # guarded faces (fluid on both sides), one per component
    ap["v"][2, 3, 2] = 0.0
    ap["w"][3, 2, 3] = 0.0
    pp = np.where(fl, rng.standard_normal(fl.shape), 0.0)
    p0 = rng.standard_normal((nz, ny, nx))
    for nm in "uvw":
        ctx.put(nm, vel0[nm])
        ctx.put(nm + "p", ap[nm])
    ctx.put("pprime", pp)
    ctx.put("p", p0)
    ctx.correct()
    R = mode == abi.HS_MODE_REF
    # expected, built face by face from the rule
    want = {nm: vel0[nm].copy() for nm in "uvw"}
    area = {"u": lambda k, j, i: yp[j] * zp[k], "v": lambda k, j, i: xp[i] * zp[k],
            "w": lambda k, j, i: xp[i] * yp[j]}
    step = {"u": (0, 0, 1), "v": (0, 1, 0), "w": (1, 0, 0)}
    for nm, ext in (("u", (0, 0, 1)), ("v", (0, 1, 0)), ("w", (1, 0, 0))):
        dk, dj, di = step[nm]
        for k in range(1, nz + 1 + (0 if R else ext[0])):
            for j in range(1, ny + 1 + (0 if R else ext[1])):
                for i in range(1, nx + 1 + (0 if R else ext[2])):
                    plus = fl[k, j, i]
                    minus = fl[k - dk, j - dj, i - di]
                    if not (plus if R else (plus and minus)):
                        continue
                    a = ap[nm][k, j, i]
                    a = 1e-10 if a == 0 else a
                    want[nm][k, j, i] += (pp[k - dk, j - dj, i - di] - pp[k, j, i]) * area[nm](k, j, i) / a
    for nm in "uvw":
        np.testing.assert_allclose(ctx.get(nm), want[nm], rtol=1e-14, atol=0, err_msg=nm)
    # the guarded faces moved by (p'_W - p'_P) A / 1e-10
    assert fl[2, 3, 3] and fl[2, 3, 2]
    du = ctx.get("u")[2, 3, 3] - vel0["u"][2, 3, 3]
    assert du == pytest.approx((pp[2, 3, 2] - pp[2, 3, 3]) * yp[3] * zp[2] / 1e-10, rel=1e-14)
    pnew = np.where(fl[1:-1, 1:-1, 1:-1], p0 + case.params.relax_p * pp[1:-1, 1:-1, 1:-1], p0)
    np.testing.assert_allclose(ctx.get("p"), pnew, rtol=1e-15, atol=0)
    pt = ctx.get("ptemp")
    assert np.array_equal(pt[1:-1, 1:-1, 1:-1][fl[1:-1, 1:-1, 1:-1]], pnew[fl[1:-1, 1:-1, 1:-1]])
    assert not np.any(ctx.get("pprime"))


def test_correction_divergence_after_exact_mode_c_solve(oracle_lib):
    """Mode C: with the p' equation solved to round-off, the corrected face velocities
    carry zero net mass out of every fluid cell whose six neighbours are fluid (the
    symmetric operator rho A^2 / a_P is exactly the divergence of the correction).  A
    cell next to a wall, inlet, outflow or solid keeps the Dirichlet p' ghost coupling
    of the reference (Q5), which no velocity face carries: there the divergence left is
    -p'_P / rho times the sum of rho A^2 / a_P over those faces."""
    case = configs.c2(scale=8)
    case.params.order = abi.HS_ORDER_RB
    case.params.p_tol = 1e-14
    case.params.p_maxit = 20000
    ctx = oracle_ctx(oracle_lib, case)
    ctx.init()
    for _ in range(3):
        ctx.step()
    for comp in range(3):
        ctx.momentum(comp)
    st = ctx.pressure_correction()
    assert st["mass_imbalance"] > 1e-6          # the momentum sweeps left a real imbalance
    pp = ctx.get("pprime")
    up, vp, wp = ctx.get("up"), ctx.get("vp"), ctx.get("wp")
    u_star = ctx.get("u")
    ctx.correct()
    nx, ny, nz = case.shape
    rho = case.params.rho
    u, v, w = ctx.get("u"), ctx.get("v"), ctx.get("w")
    xp, _ = widths(case.grid.xv)
    yp, _ = widths(case.grid.yv)
    zp, _ = widths(case.grid.zv)
    K, J, I = interior(nx, ny, nz)
    Ayz, Axz, Axy = yp[J] * zp[K], xp[I] * zp[K], xp[I] * yp[J]
    div = (u[K, J, I + 1] - u[K, J, I]) * Ayz + (v[K, J + 1, I] - v[K, J, I]) * Axz + (w[K + 1, J, I] - w[K, J, I]) * Axy
    fm = fluid_mask(case)
    fl = fm[1:-1, 1:-1, 1:-1]
    ghost = np.zeros_like(div)
    faces = (((K, J, I - 1), Ayz ** 2, up[K, J, I]), ((K, J, I + 1), Ayz ** 2, up[K, J, I + 1]),
             ((K, J - 1, I), Axz ** 2, vp[K, J, I]), ((K, J + 1, I), Axz ** 2, vp[K, J + 1, I]),
             ((K - 1, J, I), Axy ** 2, wp[K, J, I]), ((K + 1, J, I), Axy ** 2, wp[K + 1, J, I]))
    for nb, a2, apf in faces:
        # a face whose other side is not fluid: the operator's ghost coupling rho A^2 / a_P
        # (faces between two solid cells carry a_P = 0 and are not counted)
        off = (~fm[nb]) & fl
        ghost[off] += rho * a2[off] / apf[off]
    want = -pp[1:-1, 1:-1, 1:-1] * ghost / rho
    # the outflow face is a ghost coupling of the operator, then takes the corrected
    # interior face's value (zero-gradient copy after the correction)
    want[:, :, -1] += (u[1:-1, 1:-1, nx + 1] - u_star[1:-1, 1:-1, nx + 1]) * Ayz[:, :, -1]
    scale = np.abs(u).max() * (yp[1] * zp[1])
    inner = fl & (ghost == 0)
    assert inner.sum() > 1000 and (fl & (ghost > 0)).sum() > 500
    assert np.abs(div[inner]).max() < 1e-10 * scale, np.abs(div[inner]).max() / scale
    assert np.abs(div[fl] - want[fl]).max() < 1e-10 * scale, np.abs(div[fl] - want[fl]).max() / scale
    # the zero-gradient outflow face holds the corrected last interior face
    out = fl[:, :, -1]
    assert np.array_equal(u[1:-1, 1:-1, nx + 1][out], u[1:-1, 1:-1, nx][out])
    assert np.any(u[1:-1, 1:-1, nx + 1][out] != u_star[1:-1, 1:-1, nx + 1][out])


# ---------------------------------------------------------- boundary conditions ---

def test_mode_r_velocity_bcs_and_lid(oracle_lib):
    """applyX/Y/ZBoundaryCondition fluid_solver.h:209-322 as setForNext runs them
    (x, then y, then z) on a random field: min-side ghosts zeroed over the loops' own
    ranges (cx+2 / cy+2 / cz+2 entries, so the last ghost row of the longer face
    layouts is left alone), the max-side ghost never touched (the region spans the
    domain: else-if, Q10), the max boundary faces zeroed, and the lid U0 on the top
    interior u row for i < cx (Q9).  Then p_temp holds p on every interior cell
    (:1649-1655)."""
    case = small_ref_case()
    ctx = oracle_ctx(oracle_lib, case)
    nx, ny, nz = case.shape
    rng = np.random.default_rng(23)
    v0 = {nm: rng.standard_normal(ctx.get(nm).shape) for nm in "uvw"}
    for nm in "uvw":
        ctx.put(nm, v0[nm])
    p0 = rng.standard_normal((nz, ny, nx))
    ctx.put("p", p0)
    ctx.set_for_next()
    u, v, w = (v0[nm].copy() for nm in "uvw")
    U0 = case.params.u_lid
    # x (fluid_solver.h:216-247): k < cz+2, j < cy+2
    u[:nz + 2, :ny + 2, 0] = 0
    v[:nz + 2, :ny + 2, 0] = 0
    w[:nz + 2, :ny + 2, 0] = 0
    u[1:nz + 1, 1:ny + 1, nx + 1] = 0
    # y (:288-321): k < cz+2, i < cx+2
    u[:nz + 2, 0, :nx + 2] = 0
    v[:nz + 2, 0, :nx + 2] = 0
    w[:nz + 2, 0, :nx + 2] = 0
    v[1:nz + 1, ny + 1, 1:nx + 1] = 0
    u[1:nz + 1, ny, 1:nx + 1] = U0
    # z (:262-281): j < cy+2, i < cx+2
    u[0, :ny + 2, :nx + 2] = 0
    v[0, :ny + 2, :nx + 2] = 0
    w[0, :ny + 2, :nx + 2] = 0
    w[nz + 1, 1:ny + 1, 1:nx + 1] = 0
    for nm, want in (("u", u), ("v", v), ("w", w)):
        assert np.array_equal(ctx.get(nm), want), nm
    # untouched on purpose: the max-side ghost layers and the extra ghost row of the
    # face layouts (the random values survive)
    assert np.array_equal(ctx.get("u")[1:nz + 1, 1:ny + 1, nx + 2], v0["u"][1:nz + 1, 1:ny + 1, nx + 2])
    assert np.array_equal(ctx.get("v")[1:nz + 1, ny + 2, 1:nx + 1], v0["v"][1:nz + 1, ny + 2, 1:nx + 1])
    pt = ctx.get("ptemp")
    assert np.array_equal(pt[1:-1, 1:-1, 1:-1], ctx.get("p"))


def test_mode_c_boundary_values(oracle_lib):
    """Mode C (this build's configuration physics): the inlet u_in on the x = 0 faces
    of fluid cells, the lid u_lid in the whole y = max u ghost row, the zero-gradient
    outflow copying u, the tangential v / w and T across x = max, and walls whose
    normal velocity a momentum sweep leaves at zero."""
    case = configs.c2(scale=8)
    p = case.params
    p.lid, p.u_lid, p.u_in, p.order = 1, 0.45, 1.25, abi.HS_ORDER_RB
    ctx = oracle_ctx(oracle_lib, case)
    ctx.init()
    nx, ny, nz = case.shape
    fl = fluid_mask(case)
    u = ctx.get("u")
    assert np.array_equal(u[1:nz + 1, 1:ny + 1, 1], np.where(fl[1:nz + 1, 1:ny + 1, 1], 1.25, 0.0))
    assert np.all(u[:, ny + 1, :] == 0.45)
    rng = np.random.default_rng(8)
    for nm in ("u", "v", "w", "tf"):
        a = ctx.get(nm)
        inner = rng.standard_normal(a.shape)
        a[1:-1, 1:-1, 1:-1] = inner[1:-1, 1:-1, 1:-1]
        ctx.put(nm, a)
    ctx.set_for_next()
    u, v, w, t = ctx.get("u"), ctx.get("v"), ctx.get("w"), ctx.get("tf")
    assert np.array_equal(u[1:nz + 1, 1:ny + 1, nx + 1],
                          np.where(fl[1:nz + 1, 1:ny + 1, nx], u[1:nz + 1, 1:ny + 1, nx], 0.0))
    assert np.array_equal(v[1:nz + 1, 2:ny + 1, nx + 1], v[1:nz + 1, 2:ny + 1, nx])
    assert np.array_equal(w[2:nz + 1, 1:ny + 1, nx + 1], w[2:nz + 1, 1:ny + 1, nx])
    assert np.array_equal(t[1:nz + 1, 1:ny + 1, nx + 1], t[1:nz + 1, 1:ny + 1, nx])
    # walls: a sweep from zero wall-normal velocity keeps it zero (those faces are not unknowns)
    for nm, sl in (("v", (slice(None), 1, slice(None))), ("v", (slice(None), ny + 1, slice(None))),
                   ("w", (1, slice(None), slice(None))), ("w", (nz + 1, slice(None), slice(None)))):
        a = ctx.get(nm)
        a[sl] = 0.0
        ctx.put(nm, a)
    for comp in range(3):
        ctx.momentum(comp)
    assert not np.any(ctx.get("v")[:, 1, :]) and not np.any(ctx.get("v")[:, ny + 1, :])
    assert not np.any(ctx.get("w")[1]) and not np.any(ctx.get("w")[nz + 1])


# ------------------------------------------------------------------- energy ---

def test_mode_r_energy_point_update(oracle_lib):
    """solveTemp fluid_solver.h:807-846 on a stretched grid with random T and u, v, w:
    T = (conv + diff) / denom * 0.3 + 0.7 T_old with conv from the cell's west / south /
    bottom face velocity times the neighbour difference (no 1/2 dx, Q11), diff and denom
    from the spacings (dP + dPrev)/2; solid cells skipped.  One red-black sweep: the
    first colour from old values, the second from the first colour's new values."""
    case = small_ref_case(order=abi.HS_ORDER_RB)
    case.params.t_sweeps_max = 1
    ctx = oracle_ctx(oracle_lib, case)
    nx, ny, nz = case.shape
    rng = np.random.default_rng(31)
    for nm in "uvw":
        ctx.put(nm, 0.05 * rng.standard_normal(ctx.get(nm).shape))
    T0 = rng.uniform(280.0, 320.0, ctx.get("tf").shape)
    ctx.put("tf", T0)
    ctx.fluid_energy()
    u, v, w = ctx.get("u"), ctx.get("v"), ctx.get("w")
    xp, xw = widths(case.grid.xv)
    yp, ys = widths(case.grid.yv)
    zp, zf = widths(case.grid.zv)
    a, relax = case.params.alpha_f, case.params.relax_t
    fl = fluid_mask(case)
    T = T0.copy()
    K, J, I = interior(nx, ny, nz)
    for colour in (0, 1):
        m = (((K + J + I) & 1) == colour) & fl[K, J, I]
        k, j, i = K[m], J[m], I[m]
        sx, sy, sz = (xp[i] + xw[i]) / 2.0, (yp[j] + ys[j]) / 2.0, (zp[k] + zf[k]) / 2.0
        conv = (-u[k, j, i] * (T[k, j, i + 1] - T[k, j, i - 1]) - v[k, j, i] * (T[k, j + 1, i] - T[k, j - 1, i])
                - w[k, j, i] * (T[k + 1, j, i] - T[k - 1, j, i]))
        diff = (a * (T[k, j, i + 1] + T[k, j, i - 1]) / sx / sx + a * (T[k, j + 1, i] + T[k, j - 1, i]) / sy / sy
                + a * (T[k + 1, j, i] + T[k - 1, j, i]) / sz / sz)
        denom = 2 * a * (1 / sx / sx + 1 / sy / sy + 1 / sz / sz)
        T[k, j, i] = (conv + diff) / denom * relax + (1 - relax) * T[k, j, i]
    np.testing.assert_allclose(ctx.get("tf"), T, rtol=1e-14, atol=0)
    assert np.array_equal(ctx.get("tf")[~fl], T0[~fl])


def test_solid_relaxation_point_update(oracle_lib):
    """MobSolver::solveTemperature mobs_solver.h:482-488 on a random (non-harmonic)
    field: T = (sum of the six neighbours) / 6 * 0.3 + 0.7 T_old over the solid cells
    (uniform-grid mean regardless of spacing, Q12); everything else untouched."""
    case = small_ref_case(order=abi.HS_ORDER_RB)
    case.params.ts_sweeps_max = 1
    ctx = oracle_ctx(oracle_lib, case)
    ctx.init()
    nx, ny, nz = case.shape
    rng = np.random.default_rng(41)
    T0 = rng.uniform(320.0, 340.0, ctx.get("ts").shape)
    ctx.put("ts", T0)
    ctx.solid_solve()
    so = np.zeros_like(T0, bool)
    so[1:-1, 1:-1, 1:-1] = case.grid.cell_type.reshape(nz, ny, nx) == 1
    relax = case.params.relax_ts
    T = T0.copy()
    K, J, I = interior(nx, ny, nz)
    for colour in (0, 1):
        m = (((K + J + I) & 1) == colour) & so[K, J, I]
        k, j, i = K[m], J[m], I[m]
        s = 1.0 / 6.0 * (T[k, j, i + 1] + T[k, j, i - 1] + T[k, j - 1, i] + T[k, j + 1, i] + T[k - 1, j, i]
                         + T[k + 1, j, i])
        T[k, j, i] = s * relax + (1 - relax) * T[k, j, i]
    assert so.sum() == 27
    np.testing.assert_allclose(ctx.get("ts"), T, rtol=1e-14, atol=0)
    assert np.array_equal(ctx.get("ts")[~so], T0[~so])


def test_mode_c_global_energy_balance(oracle_lib):
    """The conjugate energy equation solved on a frozen divergence-free flow (plug flow
    u = u_in along a duct) with random per-cell conductivity and a heated block: the heat
    put in, sum q V, leaves as net enthalpy flux (outflow rho u A T_cell minus inflow
    rho u A T_wall) plus conduction through the Dirichlet inlet and wall ghosts
    k A (T_cell - T_wall) / width (ghost centre one cell width out); the zero-gradient
    outflow ghost conducts nothing."""
    nx, ny, nz = 16, 6, 5
    xv = np.cumsum(np.concatenate([[0.0], np.linspace(0.8, 1.2, nx)])) / nx * 3
    yv = np.concatenate([[0.0], np.cumsum(np.linspace(1.1, 0.9, ny))]) / ny
    zv = np.linspace(0.0, 1.0, nz + 1) ** 1.2
    ct = np.zeros((nz, ny, nx), np.uint8)
    p = abi.params_default(abi.HS_MODE_CONFIG)
    p.u_in, p.rho, p.t_wall = 0.8, 1.3, 300.0
    case = configs.Case("plug", xv, yv, zv, ct, p)
    ctx = oracle_ctx(oracle_lib, case)
    ctx.init()
    rng = np.random.default_rng(3)
    k = rng.uniform(0.01, 0.05, (nz, ny, nx))
    q = np.zeros((nz, ny, nx))
    q[1:4, 2:5, 4:9] = rng.uniform(0.5, 1.5, (3, 3, 5))
    ctx.put_cells("k", k)
    ctx.put_cells("q", q)
    u = ctx.get("u")
    u[...] = 0.0
    u[1:nz + 1, 1:ny + 1, 1:nx + 2] = p.u_in
    ctx.put("u", u)
    for nm in "vw":
        ctx.put(nm, np.zeros(ctx.get(nm).shape))
    info = ctx.energy_solve(1e-14, 20000)
    assert info["converged"] == 1
    T = ctx.cells("tf")
    dx, dy, dz = np.diff(xv), np.diff(yv), np.diff(zv)
    V = dz[:, None, None] * dy[None, :, None] * dx[None, None, :]
    Ayz = dz[:, None] * dy[None, :]
    Axz = dz[:, None] * dx[None, :]
    Axy = dy[:, None] * dx[None, :]
    heat_in = np.sum(q * V)
    enthalpy = p.rho * p.u_in * np.sum(Ayz * T[:, :, -1]) - p.rho * p.u_in * np.sum(Ayz * p.t_wall)
    cond = np.sum(k[:, :, 0] * Ayz * (T[:, :, 0] - p.t_wall) / dx[0])            # inlet ghost
    cond += np.sum(k[:, 0, :] * Axz * (T[:, 0, :] - p.t_wall) / dy[0])
    cond += np.sum(k[:, -1, :] * Axz * (T[:, -1, :] - p.t_wall) / dy[-1])
    cond += np.sum(k[0] * Axy * (T[0] - p.t_wall) / dz[0])
    cond += np.sum(k[-1] * Axy * (T[-1] - p.t_wall) / dz[-1])
    assert abs(heat_in - (enthalpy + cond)) < 1e-9 * heat_in, (heat_in, enthalpy, cond)
    assert enthalpy > 0 and cond > 0


# ----------------------------------------------------------- developed duct flow ---

def discrete_duct_profile(ny, nz, hy, hz):
    """Fully developed flow in the duct cross-section with the momentum sweep's wall
    treatment: sum over the four neighbours of nu A/d (u_P - u_nb) = A dp, no-slip ghost
    at one cell width (d = the cell width, the reference's fallback); returned shape
    normalised to mean 1 (uniform cells)."""
    Ty = sp.diags([-np.ones(ny - 1), 2 * np.ones(ny), -np.ones(ny - 1)], [-1, 0, 1]) / hy ** 2
    Tz = sp.diags([-np.ones(nz - 1), 2 * np.ones(nz), -np.ones(nz - 1)], [-1, 0, 1]) / hz ** 2
    L = sp.kron(sp.identity(nz), Ty) + sp.kron(Tz, sp.identity(ny))
    phi = spla.spsolve(L.tocsc(), np.ones(ny * nz)).reshape(nz, ny)
    return phi / phi.mean()


def test_mode_c_developed_duct_flow_and_mass_balance(oracle_lib):
    """A duct at Re = 5 from a uniform inlet: converged SIMPLE (mass imbalance 1e-11)
    leaves the discrete fully developed profile downstream (inlet, no-slip walls and
    the zero-gradient outflow together; 2.4e-9 measured), and the outflow carries
    exactly the inlet mass."""
    nx, ny, nz = 32, 8, 6
    L = 4.0
    xv = np.linspace(0.0, L, nx + 1)
    yv = np.linspace(0.0, 1.0, ny + 1)
    zv = np.linspace(0.0, 0.75, nz + 1)
    p = abi.params_default(abi.HS_MODE_CONFIG)
    p.nu = 0.2
    p.order = abi.HS_ORDER_RB
    p.p_tol = 1e-10
    p.mom_sweeps_max = 4
    case = configs.Case("duct", xv, yv, zv, np.zeros((nz, ny, nx), np.uint8), p)
    ctx = oracle_ctx(oracle_lib, case)
    ctx.init()
    for it in range(2000):
        st = ctx.step()
        if st["mass_imbalance"] < 1e-11:
            break
    assert st["mass_imbalance"] < 1e-11, st
    u = ctx.get("u")
    want = discrete_duct_profile(ny, nz, 1.0 / ny, 0.75 / nz) * p.u_in
    for i in (nx - 4, nx - 8):                   # x = 3.5 and 3.0: 3-3.5 duct heights downstream
        prof = u[1:nz + 1, 1:ny + 1, i]
        assert np.abs(prof - want).max() < 1e-7 * want.max(), (i, np.abs(prof - want).max())
    A = (1.0 / ny) * (0.75 / nz)
    inflow = np.sum(u[1:nz + 1, 1:ny + 1, 1]) * A
    outflow = np.sum(u[1:nz + 1, 1:ny + 1, nx + 1]) * A
    assert abs(inflow - p.u_in * 0.75) < 1e-14
    assert abs(outflow - inflow) < 1e-9 * inflow      # the summed cell imbalances (||b|| = 1e-11)
