/*
 * oracle.cpp — CPU FP64 restatement of the reference's SIMPLE conjugate-heat-transfer
 * step (bolyme/hsolver).  TEST INFRASTRUCTURE ONLY (see oracle.h): the product never
 * links, loads or calls this file.
 *
 * What it follows, function by function (reference paths relative to its root):
 *   padded work arrays + initial fills     fluid_solver.h:93-109, mobs_solver.h:50-56
 *   index maps (x fastest)                 fluid_solver.h:184-207, distributed_mesh.h:332-338
 *   cell geometry xp/xw/xe ...             fluid_solver.h:1248-1261 (repeated at :1469-1481 ...)
 *   diffusion coefficients D = 2 nu A/..   fluid_solver.h:1266-1268
 *   boundary conditions                    fluid_solver.h:209-405
 *   x/y/z momentum Gauss-Seidel            fluid_solver.h:577-659 / :491-575 / :407-489
 *   p' operator + right-hand side          fluid_solver.h:1949-2096 (frozen after the first
 *                                          assembly: values1d quirk :2088-2095 vs :2188-2193)
 *   PCG (HYPRE 2.27 ParCSR PCG, identity   fluid_solver.h:1789-1790, settings :1871-1875;
 *        preconditioner, two-norm stop)    algorithm restated from HYPRE's published pcg.c
 *   velocity / pressure correction          fluid_solver.h:1459-1622
 *   fluid energy Gauss-Seidel               fluid_solver.h:794-881
 *   solid conduction Gauss-Seidel           mobs_solver.h:303-371, :467-500
 *   fluid <-> solid interface exchange      solver.h:62-148 with lists fluid_solver.h:983-1006,
 *                                          1095-1117, 1205-1226 and mobs_solver.h:147-166
 *   setForNext                              fluid_solver.h:1634-1682, mobs_solver.h:290-301
 *
 * The configuration physics (HS_MODE_CONFIG) is this build's own extension of the same
 * discretisation; every deviation from the reference is a named switch below.
 *
 * Arithmetic is written in the reference's operation order (and built with
 * -ffp-contract=off) so that a GPU kernel written in the same order reproduces each
 * Gauss-Seidel / Jacobi point update bit for bit.  Reductions are sequential here.
 *
 * Pinning (tests/test_oracle_*.py): closed-form geometry and coefficients on
 * hand-made grids, fixed points of each sweep, the p' matrix against SciPy-built
 * reference matrices and a committed SciPy spsolve golden vector, CG against the
 * textbook recurrence, the boundary-copy loop against a brute-force literal loop,
 * Mode C conduction against the analytic slab solution and the energy balance.
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace {

/* Planted mistakes, test infrastructure of the oracle's own pins: or_test_plant(n)
 * makes the oracle commit mistake n (0 = none) so a test can show that the check meant
 * to pin that function fails.  Never set outside those tests. */
int g_plant = 0;
inline bool planted(int n) { return g_plant == n; }

struct Orc {
    int nx = 0, ny = 0, nz = 0;
    hs_params prm{};
    std::vector<double> X, Y, Z;
    std::vector<uint8_t> type;            // C layout, 0 fluid 1 solid
    // 1D geometry per padded index (reference formulas, valid for 1..n+1)
    std::vector<double> xp, xw, xe, yp, ys, yn, zp, zf, zb;
    // fields
    std::vector<double> u, v, w, up, vp, wp, p, ptemp, pprime, tf, ts, q, k, b;
    std::vector<uint8_t> fl;              // P layout: interior fluid cell
    std::vector<uint8_t> so;              // P layout: interior solid cell
    // interface pairs (fluid P index, solid P index), reference list order
    std::vector<int64_t> if_fluid, if_solid;
    // solid region bounding box (padded, inclusive); Mode R allows one region
    int sb[6] = {0, -1, 0, -1, 0, -1};
    bool has_solid = false;
    // p' operator: Mode R per-row coefficients (P layout); Mode C per-face (U/V/W)
    bool have_matrix = false;
    std::vector<double> ax, ay, az;
    bool inited = false;

    long EXU() const { return nx + 3; }
    long IU(long i, long j, long kk) const { return i + (nx + 3) * (j + (long)(ny + 2) * kk); }
    long IV(long i, long j, long kk) const { return i + (nx + 2) * (j + (long)(ny + 3) * kk); }
    long IW(long i, long j, long kk) const { return i + (nx + 2) * (j + (long)(ny + 2) * kk); }
    long IP(long i, long j, long kk) const { return i + (nx + 2) * (j + (long)(ny + 2) * kk); }
    long IC(long i, long j, long kk) const { return i + nx * (j + (long)ny * kk); }
    long NU_() const { return (long)(nx + 3) * (ny + 2) * (nz + 2); }
    long NV_() const { return (long)(nx + 2) * (ny + 3) * (nz + 2); }
    long NW_() const { return (long)(nx + 2) * (ny + 2) * (nz + 3); }
    long NP_() const { return (long)(nx + 2) * (ny + 2) * (nz + 2); }
    long NC_() const { return (long)nx * ny * nz; }
    bool isfluid(long i, long j, long kk) const {
        return i >= 1 && i <= nx && j >= 1 && j <= ny && kk >= 1 && kk <= nz && fl[IP(i, j, kk)];
    }
    bool modeR() const { return prm.mode == HS_MODE_REF; }
};

/* fluid_solver.h:1248-1261: the reference's cell widths for a padded index i of a
 * region starting at global cell 1, with the fallbacks past either end. */
void axis_geom(const std::vector<double>& V, int n, std::vector<double>& pw,
               std::vector<double>& mw, std::vector<double>& plw) {
    const long S = n + 1;  // ulocation.size()
    pw.assign(n + 3, 0.0); mw.assign(n + 3, 0.0); plw.assign(n + 3, 0.0);
    for (long i = 1; i <= n + 1; ++i) {
        double c = (i < S) ? (V[i] - V[i - 1]) : (V[i - 1] - V[i - 2]);
        double m = (i - 2 >= 0) ? (V[i - 1] - V[i - 2]) : c;
        double p = (i + 1 < S) ? (V[i + 1] - V[i]) : c;
        pw[i] = c; mw[i] = m; plw[i] = p;
    }
}

inline double pos(double a) { return (a < 0.0) ? 0.0 : a; }  /* std::max(a, 0.0) */

/* ---------------------------------------------------------------- BCs ---- */

/* applyXBoundaryCondition fluid_solver.h:209-257 for a region spanning the domain
 * (metadata[6]==1 so only the min-side ghost branch runs, quirk Q10). */
void bc_x_ref(Orc& o) {
    const int cx = o.nx, cy = o.ny, cz = o.nz, fx = cx + 1;
    for (int kk = 0; kk < cz + 2; ++kk)
        for (int j = 0; j < cy + 2; ++j) {
            o.u[o.IU(0, j, kk)] = 0; o.v[o.IV(0, j, kk)] = 0; o.w[o.IW(0, j, kk)] = 0;
        }
    for (int kk = 0; kk < cz; ++kk)
        for (int j = 0; j < cy; ++j) o.u[o.IU(fx, j + 1, kk + 1)] = 0;
    if (planted(4))
        for (int kk = 0; kk < cz + 2; ++kk)
            for (int j = 0; j < cy + 2; ++j) o.u[o.IU(fx + 1, j, kk)] = 0;
}
/* applyYBoundaryCondition fluid_solver.h:284-322 (lid on the top interior u row, Q9). */
void bc_y_ref(Orc& o) {
    const int cx = o.nx, cy = o.ny, cz = o.nz, fy = cy + 1;
    for (int kk = 0; kk < cz + 2; ++kk)
        for (int i = 0; i < cx + 2; ++i) {
            o.u[o.IU(i, 0, kk)] = 0; o.v[o.IV(i, 0, kk)] = 0; o.w[o.IW(i, 0, kk)] = 0;
        }
    for (int kk = 0; kk < cz; ++kk)
        for (int i = 0; i < cx + (planted(5) ? 1 : 0); ++i) {
            if (i < cx) o.v[o.IV(i + 1, fy, kk + 1)] = 0;
            o.u[o.IU(i + 1, cy, kk + 1)] = o.prm.u_lid;
        }
}
/* applyZBoundaryCondition fluid_solver.h:259-282. */
void bc_z_ref(Orc& o) {
    const int cx = o.nx, cy = o.ny, cz = o.nz, fz = cz + 1;
    for (int j = 0; j < cy + 2; ++j)
        for (int i = 0; i < cx + 2; ++i) {
            o.u[o.IU(i, j, 0)] = 0; o.v[o.IV(i, j, 0)] = 0; o.w[o.IW(i, j, 0)] = 0;
        }
    for (int j = 0; j < cy; ++j)
        for (int i = 0; i < cx; ++i) o.w[o.IW(i + 1, j + 1, fz)] = 0;
}

/* applyCellBoundaryCondition fluid_solver.h:324-405, restated literally: one
 * sequential k,j,i pass in which each boundary cell copies its inward neighbour
 * (reading whatever that neighbour holds at that moment), then p(0,0,0)=0.  The
 * output-only t=293 overwrite (:394-404) has no effect on the solve. */
void cell_bc_ref(Orc& o) {
    const int cx = o.nx, cy = o.ny, cz = o.nz;
    const int Nx = o.nx, Ny = o.ny, Nz = o.nz;
    std::vector<double>& P = o.p;
    std::vector<double>& PT = o.ptemp;
    for (int kk = 0; kk < cz; ++kk)
        for (int j = 0; j < cy; ++j)
            for (int i = 0; i < cx; ++i) {
                const int gx = i, gy = j, gz = kk;
                const long c = o.IC(i, j, kk);
                if (gx == 0) P[c] = P[o.IC(i + 1, j, kk)];
                else if (gx == Nx - 1) P[c] = P[o.IC(i - 1, j, kk)];
                else if (gy == 0) P[c] = P[o.IC(i, j + 1, kk)];
                else if (gy == Ny - 1) P[c] = P[o.IC(i, j - 1, kk)];
                else if (gz == 0) P[c] = P[o.IC(i, j, kk + 1)];
                else if (gz == Nz - 1) P[c] = P[o.IC(i, j, kk - 1)];
                const double val = P[c];
                if (gx == 0) {
                    PT[o.IP(i + 1, j + 1, kk + 1)] = val; PT[o.IP(i, j + 1, kk + 1)] = val;
                } else if (gx == Nx - 1) {
                    PT[o.IP(i + 1, j + 1, kk + 1)] = val; PT[o.IP(i + 2, j + 1, kk + 1)] = val;
                } else if (gy == 0) {
                    PT[o.IP(i + 1, j + 1, kk + 1)] = val; PT[o.IP(i + 1, j, kk + 1)] = val;
                } else if (gy == Ny - 1) {
                    PT[o.IP(i + 1, j + 1, kk + 1)] = val; PT[o.IP(i + 1, j + 2, kk + 1)] = val;
                    if (i + 1 == cx) PT[o.IP(i + 2, j + 2, kk + 1)] = val;
                    if (kk + 1 == cz) PT[o.IP(i + 1, j + 2, kk + 2)] = val;
                    if (i + 1 == 1) PT[o.IP(i, j + 2, kk + 1)] = val;
                    if (kk + 1 == 1) PT[o.IP(i + 1, j + 2, kk)] = val;
                    if (i + 1 == cx && kk + 1 == cz) PT[o.IP(i + 2, j + 2, kk + 2)] = val;
                    if (i + 1 == 1 && kk + 1 == 1) PT[o.IP(i, j + 2, kk)] = val;
                } else if (gz == 0) {
                    PT[o.IP(i + 1, j + 1, kk + 1)] = val; PT[o.IP(i + 1, j + 1, kk)] = val;
                } else if (gz == Nz - 1) {
                    PT[o.IP(i + 1, j + 1, kk + 1)] = val; PT[o.IP(i + 1, j + 1, kk + 2)] = val;
                }
            }
    P[o.IC(0, 0, 0)] = 0;
}

/* Mode C boundary values (this build's config physics, no reference counterpart):
 * inlet u on x=min faces of fluid cells, walls hold normal velocity 0, lid ghost on
 * y=max for u, outflow zero-gradient on x=max (faces and tangential ghosts). */
void bc_cfg_init(Orc& o) {
    const int cx = o.nx, cy = o.ny, cz = o.nz;
    for (int kk = 1; kk <= cz; ++kk)
        for (int j = 1; j <= cy; ++j) {
            o.u[o.IU(1, j, kk)] = o.isfluid(1, j, kk) ? o.prm.u_in : 0.0;
        }
    if (o.prm.lid)
        for (int kk = 0; kk < cz + 2; ++kk)
            for (int i = 0; i < cx + 3; ++i) o.u[o.IU(i, cy + 1, kk)] = o.prm.u_lid;
}
void bc_cfg_outflow(Orc& o, int comp) {
    if (!o.prm.outflow) return;
    const int cx = o.nx, cy = o.ny, cz = o.nz;
    if (comp == 0) {
        for (int kk = 1; kk <= cz; ++kk)
            for (int j = 1; j <= cy; ++j)
                o.u[o.IU(cx + 1, j, kk)] = o.isfluid(cx, j, kk) ? o.u[o.IU(cx, j, kk)] : 0.0;
    } else if (comp == 1) {
        for (int kk = 1; kk <= cz; ++kk)
            for (int j = 2; j <= cy; ++j) o.v[o.IV(cx + 1, j, kk)] = o.v[o.IV(cx, j, kk)];
    } else if (comp == 2) {
        for (int kk = 2; kk <= cz; ++kk)
            for (int j = 1; j <= cy; ++j) o.w[o.IW(cx + 1, j, kk)] = o.w[o.IW(cx, j, kk)];
    } else {  // temperature
        if (planted(10)) return;
        for (int kk = 1; kk <= cz; ++kk)
            for (int j = 1; j <= cy; ++j) o.tf[o.IP(cx + 1, j, kk)] = o.tf[o.IP(cx, j, kk)];
    }
}

/* --------------------------------------------------------- momentum ---- */

/* Is the face of component comp at padded (i,j,k) a Mode C unknown?  Both adjacent
 * cells must be interior fluid cells (no-penetration on walls and solid faces). */
bool cfg_solved(const Orc& o, int comp, long i, long j, long kk) {
    if (comp == 0) return o.isfluid(i - 1, j, kk) && o.isfluid(i, j, kk);
    if (comp == 1) return o.isfluid(i, j - 1, kk) && o.isfluid(i, j, kk);
    return o.isfluid(i, j, kk - 1) && o.isfluid(i, j, kk);
}
bool cfg_touch(const Orc& o, int comp, long i, long j, long kk) {
    if (comp == 0) return o.isfluid(i - 1, j, kk) || o.isfluid(i, j, kk);
    if (comp == 1) return o.isfluid(i, j - 1, kk) || o.isfluid(i, j, kk);
    return o.isfluid(i, j, kk - 1) || o.isfluid(i, j, kk);
}

/* One face update of solveX/Y/Zmomentum (fluid_solver.h:588-625, :505-542, :419-457)
 * in the reference's operation order.  `swap` selects the reference's swapped upwind
 * labels (Q7: x and y in y-momentum, z in z-momentum).  Returns the new value;
 * *ap_out receives the diagonal. */
double mom_point(const Orc& o, int comp, long i, long j, long kk, bool write_only_ap,
                 double* ap_out) {
    const double c05 = 0.5 * o.prm.rho;
    const double c2nu = 2 * o.prm.nu;
    const double xp = o.xp[i], xw = o.xw[i], xe = o.xe[i];
    const double yp = o.yp[j], ys = (planted(11) && j == 1) ? 0.0 : o.ys[j], yn = o.yn[j];
    const double zp = o.zp[kk], zf = o.zf[kk], zb = o.zb[kk];
    const double Axy = xp * yp, Axz = xp * zp, Ayz = yp * zp;
    const double Dw = c2nu * Ayz / (xp + xw), De = c2nu * Ayz / (xp + xe);
    const double Ds = c2nu * Axz / (ys + yp), Dn = c2nu * Axz / (yn + yp);
    const double Db = c2nu * Axy / (zb + zp), Df = c2nu * Axy / (zf + zp);
    const std::vector<double>& U = o.u;
    const std::vector<double>& V = o.v;
    const std::vector<double>& W = o.w;
    const std::vector<double>& PT = o.ptemp;
    double Fmx, Fpx, Fmy, Fpy, Fmz, Fpz, b;
    const std::vector<double>* F;
    long s0, sx, sy, sz;  // self index and strides in the component's layout
    if (comp == 0) {
        s0 = o.IU(i, j, kk); sx = 1; sy = o.nx + 3; sz = (long)(o.nx + 3) * (o.ny + 2); F = &U;
        Fmx = c05 * Ayz * (U[s0] + U[s0 - 1]);
        Fpx = c05 * Ayz * (U[s0] + U[s0 + 1]);
        Fmy = c05 * Axz * (V[o.IV(i, j, kk)] + V[o.IV(i - 1, j, kk)]);
        Fpy = c05 * Axz * (V[o.IV(i, j + 1, kk)] + V[o.IV(i - 1, j + 1, kk)]);
        Fpz = c05 * Axy * (W[o.IW(i, j, kk + 1)] + W[o.IW(i - 1, j, kk + 1)]);
        Fmz = c05 * Axy * (W[o.IW(i, j, kk)] + W[o.IW(i - 1, j, kk)]);
        b = Ayz * (PT[o.IP(i - 1, j, kk)] - PT[o.IP(i, j, kk)]);
    } else if (comp == 1) {
        s0 = o.IV(i, j, kk); sx = 1; sy = o.nx + 2; sz = (long)(o.nx + 2) * (o.ny + 3); F = &V;
        Fmy = c05 * Axz * (V[s0] + V[s0 - sy]);
        Fpy = c05 * Axz * (V[s0] + V[s0 + sy]);
        Fmx = c05 * Ayz * (U[o.IU(i, j, kk)] + U[o.IU(i, j - 1, kk)]);
        Fpx = c05 * Ayz * (U[o.IU(i + 1, j, kk)] + U[o.IU(i + 1, j - 1, kk)]);
        Fpz = c05 * Axy * (W[o.IW(i, j, kk + 1)] + W[o.IW(i, j - 1, kk + 1)]);
        Fmz = c05 * Axy * (W[o.IW(i, j, kk)] + W[o.IW(i, j - 1, kk)]);
        b = Axz * (PT[o.IP(i, j - 1, kk)] - PT[o.IP(i, j, kk)]);
    } else {
        s0 = o.IW(i, j, kk); sx = 1; sy = o.nx + 2; sz = (long)(o.nx + 2) * (o.ny + 2); F = &W;
        Fmz = c05 * Axy * (W[s0] + W[s0 - sz]);
        Fpz = c05 * Axy * (W[s0] + W[s0 + sz]);
        Fmx = c05 * Ayz * (U[o.IU(i, j, kk)] + U[o.IU(i, j, kk - 1)]);
        Fpx = c05 * Ayz * (U[o.IU(i + 1, j, kk)] + U[o.IU(i + 1, j, kk - 1)]);
        Fmy = c05 * Axz * (V[o.IV(i, j, kk)] + V[o.IV(i, j, kk - 1)]);
        Fpy = c05 * Axz * (V[o.IV(i, j + 1, kk)] + V[o.IV(i, j + 1, kk - 1)]);
        b = Axy * (PT[o.IP(i, j, kk - 1)] - PT[o.IP(i, j, kk)]);
    }
    const bool R = o.modeR();
    const bool swx = R && comp == 1, swy = R && comp == 1, swz = R && comp == 2;
    const double ae = De + (swx ? pos(-Fmx) : pos(-Fpx));
    const double aw = Dw + (swx ? pos(Fpx) : pos(Fmx));
    const double an = Dn + (swy ? pos(-Fmy) : pos(-Fpy));
    const double as = Ds + (swy ? pos(Fpy) : pos(Fmy));
    const double ab = Db + (swz ? pos(-Fmz) : pos(-Fpz));
    const double af = Df + (swz ? pos(Fpz) : pos(Fmz));
    const double ap = ae + aw + an + as + ab + af;
    *ap_out = ap;
    if (write_only_ap) return 0.0;
    const std::vector<double>& X = *F;
    return (ae * X[s0 + sx] + aw * X[s0 - sx] + an * X[s0 + sy] + as * X[s0 - sy] +
            af * X[s0 - sz] + ab * X[s0 + sz] + b) / ap;
}

struct MomRange { long i0, i1, j0, j1, k0, k1; double N; };

MomRange mom_range(const Orc& o, int comp) {
    // solve() sets xEnd=fx for x, yEnd=fy for y, zEnd=fz for z (fluid_solver.h:1414-1430);
    // the sweep runs to End-1 in the face direction and N is the full box count.
    MomRange r{1, o.nx, 1, o.ny, 1, o.nz, 0};
    const double cx = o.nx, cy = o.ny, cz = o.nz;
    if (comp == 0) r.N = cz * cy * (cx + 1);
    else if (comp == 1) r.N = cz * (cy + 1) * cx;
    else r.N = (cz + 1) * cy * cx;
    if (!o.modeR()) {  // Mode C visits every face that touches a fluid cell
        if (comp == 0) r.i1 = o.nx + 1;
        if (comp == 1) r.j1 = o.ny + 1;
        if (comp == 2) r.k1 = o.nz + 1;
    }
    return r;
}

/* Does the reference's loop update this face?  Mode R: the cell on the plus side
 * (p[getC(i-1,j-1,k-1)] in the padded-index loops) is fluid.  Mode C: both sides. */
inline bool mom_active(const Orc& o, int comp, long i, long j, long kk) {
    if (o.modeR()) return o.isfluid(i, j, kk);
    return cfg_solved(o, comp, i, j, kk);
}

double mom_sweep(Orc& o, int comp) {
    const MomRange r = mom_range(o, comp);
    std::vector<double>& X = comp == 0 ? o.u : (comp == 1 ? o.v : o.w);
    std::vector<double>& AP = comp == 0 ? o.up : (comp == 1 ? o.vp : o.wp);
    const double relax = o.prm.relax_u;
    double sum = 0.0;
    auto idx = [&](long i, long j, long kk) {
        return comp == 0 ? o.IU(i, j, kk) : (comp == 1 ? o.IV(i, j, kk) : o.IW(i, j, kk));
    };
    auto visit = [&](long i, long j, long kk) {
        const long s = idx(i, j, kk);
        if (mom_active(o, comp, i, j, kk)) {
            double ap;
            const double nv = mom_point(o, comp, i, j, kk, false, &ap);
            const double old = X[s];
            X[s] = relax * nv + (1 - relax) * old;
            AP[s] = ap;
            sum += (X[s] - old) * (X[s] - old);
        } else if (!o.modeR() && cfg_touch(o, comp, i, j, kk)) {
            double ap;
            mom_point(o, comp, i, j, kk, true, &ap);
            AP[s] = ap;
        }
    };
    if (o.prm.order == HS_ORDER_RB) {
        for (int colour = 0; colour < 2; ++colour)
            for (long kk = r.k0; kk <= r.k1; ++kk)
                for (long j = r.j0; j <= r.j1; ++j)
                    for (long i = r.i0; i <= r.i1; ++i)
                        if (((i + j + kk) & 1) == colour) visit(i, j, kk);
    } else {
        for (long kk = r.k0; kk <= r.k1; ++kk)
            for (long j = r.j0; j <= r.j1; ++j)
                for (long i = r.i0; i <= r.i1; ++i) visit(i, j, kk);
    }
    return sum;
}

void apply_bc_comp(Orc& o, int comp) {
    if (o.modeR()) {
        if (comp == 0) bc_x_ref(o); else if (comp == 1) bc_y_ref(o); else bc_z_ref(o);
    } else {
        bc_cfg_outflow(o, comp);
    }
}

double cfg_count_solved(const Orc& o, int comp) {
    const MomRange r = mom_range(o, comp);
    double n = 0;
    for (long kk = r.k0; kk <= r.k1; ++kk)
        for (long j = r.j0; j <= r.j1; ++j)
            for (long i = r.i0; i <= r.i1; ++i) n += cfg_solved(o, comp, i, j, kk) ? 1 : 0;
    return n;
}

/* solveX/Y/Zmomentum control loop (fluid_solver.h:580-658): Mode R keeps the
 * residual accumulator that starts at 1 and is divided, never reset (Q6). */
void momentum(Orc& o, int comp, int* sweeps, double* resid) {
    const MomRange r = mom_range(o, comp);
    const bool R = o.modeR();
    const double N = R ? r.N : std::max(1.0, cfg_count_solved(o, comp));
    double residual = 1, global = R ? 1.0 : std::numeric_limits<double>::infinity();
    int step = 0;
    while (step < o.prm.mom_sweeps_max) {
        if (global < o.prm.mom_tol) break;
        const double s = mom_sweep(o, comp);
        step++;
        if (R) { residual = (residual + s) / N; global = std::sqrt(residual); }
        else { global = std::sqrt(s / N); }
        apply_bc_comp(o, comp);
    }
    if (sweeps) *sweeps = step;
    if (resid) *resid = global;
}

/* --------------------------------------------------- p' operator ---- */

/* Mode R coefficients (initColsAndSetValues fluid_solver.h:1966-1996): per fluid
 * row, a_e = a_w = A_yz/up[west face], a_n = a_s = A_xz/vp[south], a_t = a_b =
 * A_xy/wp[bottom]; diagonal = sum of all six (Dirichlet p'=0 ghost, Q5).
 * Mode C (this build): symmetric face coefficient rho*A^2/ap on every face of a
 * fluid cell (area-consistent with the correction u += A/ap * dp'). */
void assemble_matrix(Orc& o) {
    const long NP = o.NP_();
    if (o.modeR()) {
        o.ax.assign(NP, 0.0); o.ay.assign(NP, 0.0); o.az.assign(NP, 0.0);
        for (long kk = 1; kk <= o.nz; ++kk)
            for (long j = 1; j <= o.ny; ++j)
                for (long i = 1; i <= o.nx; ++i) {
                    if (!o.isfluid(i, j, kk)) continue;
                    const double xp = o.xp[i], yp = o.yp[j], zp = o.zp[kk];
                    const double Axy = xp * yp, Axz = xp * zp, Ayz = yp * zp;
                    const long c = o.IP(i, j, kk);
                    o.ax[c] = Ayz / o.up[o.IU(i, j, kk)];
                    o.ay[c] = Axz / o.vp[o.IV(i, j, kk)];
                    o.az[c] = Axy / o.wp[o.IW(i, j, kk)];
                }
    } else {
        o.ax.assign(o.NU_(), 0.0); o.ay.assign(o.NV_(), 0.0); o.az.assign(o.NW_(), 0.0);
        const double rho = o.prm.rho;
        for (long kk = 1; kk <= o.nz; ++kk)
            for (long j = 1; j <= o.ny; ++j)
                for (long i = 1; i <= o.nx + 1; ++i)
                    if (o.isfluid(i - 1, j, kk) || o.isfluid(i, j, kk)) {
                        const double A = o.yp[j] * o.zp[kk];
                        const long f = o.IU(i, j, kk);
                        o.ax[f] = rho * A * A / o.up[f];
                    }
        for (long kk = 1; kk <= o.nz; ++kk)
            for (long j = 1; j <= o.ny + 1; ++j)
                for (long i = 1; i <= o.nx; ++i)
                    if (o.isfluid(i, j - 1, kk) || o.isfluid(i, j, kk)) {
                        const double A = o.xp[i] * o.zp[kk];
                        const long f = o.IV(i, j, kk);
                        o.ay[f] = rho * A * A / o.vp[f];
                    }
        for (long kk = 1; kk <= o.nz + 1; ++kk)
            for (long j = 1; j <= o.ny; ++j)
                for (long i = 1; i <= o.nx; ++i)
                    if (o.isfluid(i, j, kk - 1) || o.isfluid(i, j, kk)) {
                        const double A = o.xp[i] * o.yp[j];
                        const long f = o.IW(i, j, kk);
                        o.az[f] = rho * A * A / o.wp[f];
                    }
    }
    o.have_matrix = true;
}

/* right-hand side: Mode R b = -(du) - (dv) - (dw) (fluid_solver.h:1998-2002, Q4);
 * Mode C b = -rho * sum of area-weighted outward face velocities. */
void assemble_rhs(Orc& o) {
    o.b.assign(o.NP_(), 0.0);
    for (long kk = 1; kk <= o.nz; ++kk)
        for (long j = 1; j <= o.ny; ++j)
            for (long i = 1; i <= o.nx; ++i) {
                if (!o.isfluid(i, j, kk)) continue;
                const double du = o.u[o.IU(i + 1, j, kk)] - o.u[o.IU(i, j, kk)];
                const double dv = o.v[o.IV(i, j + 1, kk)] - o.v[o.IV(i, j, kk)];
                const double dw = o.w[o.IW(i, j, kk + 1)] - o.w[o.IW(i, j, kk)];
                if (o.modeR()) {
                    o.b[o.IP(i, j, kk)] = -du - dv - dw;
                } else {
                    const double xp = o.xp[i], yp = o.yp[j], zp = o.zp[kk];
                    o.b[o.IP(i, j, kk)] = -o.prm.rho * (yp * zp * du + xp * zp * dv + xp * yp * dw);
                }
            }
}

/* y = A x on fluid cells, P layout; entries of x outside fluid cells are treated
 * as zero (Dirichlet ghost).  Row sum order: diagonal, then B, S, W, E, N, T (the
 * CSR column order of fluid_solver.h:2003-2083 with the diagonal first). */
void apply_A(const Orc& o, const std::vector<double>& x, std::vector<double>& y) {
    y.assign(o.NP_(), 0.0);
    const long sy = o.nx + 2, sz = (long)(o.nx + 2) * (o.ny + 2);
    auto X = [&](long i, long j, long kk) { return o.isfluid(i, j, kk) ? x[o.IP(i, j, kk)] : 0.0; };
    for (long kk = 1; kk <= o.nz; ++kk)
        for (long j = 1; j <= o.ny; ++j)
            for (long i = 1; i <= o.nx; ++i) {
                if (!o.isfluid(i, j, kk)) continue;
                const long c = o.IP(i, j, kk);
                double cB, cS, cW, cE, cN, cT;
                if (o.modeR()) {
                    cW = cE = o.ax[c]; cS = cN = o.ay[c]; cB = cT = o.az[c];
                } else {
                    cW = o.ax[o.IU(i, j, kk)]; cE = o.ax[o.IU(i + 1, j, kk)];
                    cS = o.ay[o.IV(i, j, kk)]; cN = o.ay[o.IV(i, j + 1, kk)];
                    cB = o.az[o.IW(i, j, kk)]; cT = o.az[o.IW(i, j, kk + 1)];
                }
                const double d = o.modeR() ? (cE + cW + cN + cS + cT + cB)
                                           : (cW + cE + cS + cN + cB + cT);
                double t = d * x[c];
                t += (-cB) * X(i, j, kk - 1);
                t += (-cS) * X(i, j - 1, kk);
                t += (-cW) * X(i - 1, j, kk);
                t += (-cE) * X(i + 1, j, kk);
                t += (-cN) * X(i, j + 1, kk);
                t += (-cT) * X(i, j, kk + 1);
                y[c] = t;
                (void)sy; (void)sz;
            }
}

double dot_fluid(const Orc& o, const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0.0;
    for (long kk = 1; kk <= o.nz; ++kk)
        for (long j = 1; j <= o.ny; ++j)
            for (long i = 1; i <= o.nx; ++i)
                if (o.isfluid(i, j, kk)) { const long c = o.IP(i, j, kk); s += a[c] * b[c]; }
    return s;
}

/* The CG on several host threads (or_set_threads): the same recurrence as pcg below,
 * with the operator apply, the vector updates and the dot products split into
 * contiguous z-slabs of the padded grid, one per thread, and each dot product summed
 * per slab in the serial loop order, then over the slabs in slab order (deterministic
 * for a given thread count; the rounding differs from the single-thread order).  This
 * is the shape of the reference's z-slab MPI ranks (per-rank partial sums, then
 * MPI_Allreduce) on one host.  Used only for the CPU baseline timing. */
int g_threads = 1;

struct Slabs {
    int n;
    std::vector<long> k0, k1;   // padded planes [k0, k1] of slab t (1..nz split evenly)
};
Slabs make_slabs(const Orc& o, int nt) {
    Slabs s;
    s.n = std::max(1, std::min(nt, o.nz));
    for (int t = 0; t < s.n; ++t) {
        s.k0.push_back(1 + (long)o.nz * t / s.n);
        s.k1.push_back((long)o.nz * (t + 1) / s.n);
    }
    return s;
}

void pcg_threads(const Orc& o, const std::vector<double>& b, std::vector<double>& x, hs_stats* st) {
    const long NP = o.NP_();
    const Slabs sl = make_slabs(o, g_threads);
    const long sz = (long)(o.nx + 2) * (o.ny + 2);
    x.assign(NP, 0.0);
    std::vector<double> r(b), p(b), s(NP, 0.0), part(sl.n, 0.0);
    for (long c = 0; c < NP; ++c) if (!o.fl[c]) { r[c] = 0; p[c] = 0; }
    auto pdot = [&](const std::vector<double>& a, const std::vector<double>& v) {
#pragma omp parallel for num_threads(sl.n) schedule(static, 1)
        for (int t = 0; t < sl.n; ++t) {
            double acc = 0.0;
            for (long kk = sl.k0[t]; kk <= sl.k1[t]; ++kk)
                for (long j = 1; j <= o.ny; ++j)
                    for (long i = 1; i <= o.nx; ++i)
                        if (o.isfluid(i, j, kk)) { const long c = o.IP(i, j, kk); acc += a[c] * v[c]; }
            part[t] = acc;
        }
        double tot = 0.0;
        for (int t = 0; t < sl.n; ++t) tot += part[t];
        return tot;
    };
    auto papply = [&](const std::vector<double>& v, std::vector<double>& y) {
        auto X = [&](long i, long j, long kk) { return o.isfluid(i, j, kk) ? v[o.IP(i, j, kk)] : 0.0; };
#pragma omp parallel for num_threads(sl.n) schedule(static, 1)
        for (int t = 0; t < sl.n; ++t)
            for (long kk = sl.k0[t]; kk <= sl.k1[t]; ++kk)
                for (long j = 1; j <= o.ny; ++j)
                    for (long i = 1; i <= o.nx; ++i) {
                        const long c = o.IP(i, j, kk);
                        if (!o.isfluid(i, j, kk)) { y[c] = 0.0; continue; }
                        double cB, cS, cW, cE, cN, cT;
                        if (o.modeR()) {
                            cW = cE = o.ax[c]; cS = cN = o.ay[c]; cB = cT = o.az[c];
                        } else {
                            cW = o.ax[o.IU(i, j, kk)]; cE = o.ax[o.IU(i + 1, j, kk)];
                            cS = o.ay[o.IV(i, j, kk)]; cN = o.ay[o.IV(i, j + 1, kk)];
                            cB = o.az[o.IW(i, j, kk)]; cT = o.az[o.IW(i, j, kk + 1)];
                        }
                        const double d = o.modeR() ? (cE + cW + cN + cS + cT + cB) : (cW + cE + cS + cN + cB + cT);
                        y[c] = d * v[c] - cB * X(i, j, kk - 1) - cS * X(i, j - 1, kk) - cW * X(i - 1, j, kk) -
                               cE * X(i + 1, j, kk) - cN * X(i, j + 1, kk) - cT * X(i, j, kk + 1);
                    }
    };
    // vector updates over whole slabs of the padded layout
    auto pfor = [&](auto&& body) {
#pragma omp parallel for num_threads(sl.n) schedule(static, 1)
        for (int t = 0; t < sl.n; ++t)
            for (long c = sl.k0[t] * sz; c < (sl.k1[t] + 1) * sz; ++c)
                if (o.fl[c]) body(c);
    };
    const double bb = pdot(b, b);
    if (st) { st->p_iters = 0; st->p_converged = 1; st->p_relres = 0; }
    if (bb == 0.0) return;
    const double eps = o.prm.p_tol * o.prm.p_tol;
    double gamma = pdot(r, p), iprod = gamma;
    int it = 0, conv = 0;
    while (it + 1 <= o.prm.p_maxit) {
        ++it;
        papply(p, s);
        const double sdotp = pdot(s, p);
        if (sdotp == 0.0) break;
        const double alpha = gamma / sdotp;
        if (!(alpha > std::numeric_limits<double>::min())) break;
        const double gamma_old = gamma;
        pfor([&](long c) { x[c] += alpha * p[c]; r[c] -= alpha * s[c]; });
        gamma = pdot(r, r);
        iprod = gamma;
        if (iprod / bb < eps) { conv = 1; break; }
        const double beta = gamma / gamma_old;
        pfor([&](long c) { p[c] = r[c] + beta * p[c]; });
    }
    if (st) { st->p_iters = it; st->p_converged = conv; st->p_relres = std::sqrt(iprod / bb); }
}

/* HYPRE ParCSR PCG as configured at fluid_solver.h:1871-1875 (identity
 * preconditioner, SetTwoNorm(1), SetTol, SetMaxIter, x0 = 0 per :1766/:1862),
 * restated from HYPRE's pcg.c: stop when <r,r>/<b,b> < tol^2 after an update;
 * zero <b,b> returns x = b; zero <s,p> or non-positive alpha ends the loop. */
void pcg(const Orc& o, const std::vector<double>& b, std::vector<double>& x, hs_stats* st) {
    if (g_threads > 1) { pcg_threads(o, b, x, st); return; }
    const long NP = o.NP_();
    x.assign(NP, 0.0);
    const double bb = dot_fluid(o, b, b);
    if (st) { st->p_iters = 0; st->p_converged = 1; st->p_relres = 0; }
    if (bb == 0.0) return;
    const double eps = o.prm.p_tol * o.prm.p_tol;
    std::vector<double> r(b), p(b), s;
    for (long c = 0; c < NP; ++c) if (!o.fl[c]) { r[c] = 0; p[c] = 0; }
    double gamma = dot_fluid(o, r, p);
    double iprod = gamma;
    int it = 0; int conv = 0;
    while (it + 1 <= o.prm.p_maxit) {
        ++it;
        apply_A(o, p, s);
        const double sdotp = dot_fluid(o, s, p);
        if (sdotp == 0.0) break;
        const double alpha = gamma / sdotp;
        if (!(alpha > std::numeric_limits<double>::min())) break;
        const double gamma_old = gamma;
        for (long c = 0; c < NP; ++c) if (o.fl[c]) { x[c] += alpha * p[c]; r[c] -= alpha * s[c]; }
        gamma = dot_fluid(o, r, r);
        iprod = gamma;
        if (iprod / bb < eps) { conv = 1; break; }
        const double beta = gamma / gamma_old;
        for (long c = 0; c < NP; ++c) if (o.fl[c]) p[c] = r[c] + beta * p[c];
    }
    if (st) { st->p_iters = it; st->p_converged = conv; st->p_relres = std::sqrt(iprod / bb); }
}

double true_relres(const Orc& o, const std::vector<double>& b, const std::vector<double>& x) {
    std::vector<double> ax;
    apply_A(o, x, ax);
    double rr = 0, bb = 0;
    for (long kk = 1; kk <= o.nz; ++kk)
        for (long j = 1; j <= o.ny; ++j)
            for (long i = 1; i <= o.nx; ++i)
                if (o.isfluid(i, j, kk)) {
                    const long c = o.IP(i, j, kk);
                    const double r = b[c] - ax[c];
                    rr += r * r; bb += b[c] * b[c];
                }
    return bb > 0 ? std::sqrt(rr / bb) : 0.0;
}

void pressure_correction(Orc& o, hs_stats* st) {
    if (!o.have_matrix || !(o.modeR() && o.prm.freeze_p_matrix)) assemble_matrix(o);
    assemble_rhs(o);
    hs_stats loc{};
    std::vector<double> x;
    pcg(o, o.b, x, &loc);
    for (long c = 0; c < o.NP_(); ++c) o.pprime[c] = o.fl[c] ? x[c] : 0.0;
    if (st) {
        st->p_iters = loc.p_iters; st->p_converged = loc.p_converged; st->p_relres = loc.p_relres;
        st->p_true_relres = true_relres(o, o.b, x);
        st->mass_imbalance = std::sqrt(dot_fluid(o, o.b, o.b));
    }
}

/* velocity correction fluid_solver.h:1459-1571 and pressure update :1574-1598. */
void correct(Orc& o) {
    const bool R = o.modeR();
    const std::vector<double>& PP = o.pprime;
    for (long kk = 1; kk <= o.nz; ++kk)
        for (long j = 1; j <= o.ny; ++j)
            for (long i = 1; i <= o.nx + (R ? 0 : 1); ++i) {
                if (R ? !o.isfluid(i, j, kk) : !cfg_solved(o, 0, i, j, kk)) continue;
                const double faceyz = planted(1) ? o.xp[i] * o.zp[kk] : o.yp[j] * o.zp[kk];
                double ap = o.up[o.IU(i, j, kk)];
                if (ap == 0 && !planted(2)) ap = 1e-10;
                const double corr = (PP[o.IP(i - 1, j, kk)] - PP[o.IP(i, j, kk)]) * faceyz / ap;
                o.u[o.IU(i, j, kk)] = o.u[o.IU(i, j, kk)] + corr;
            }
    for (long kk = 1; kk <= o.nz; ++kk)
        for (long j = 1; j <= o.ny + (R ? 0 : 1); ++j)
            for (long i = 1; i <= o.nx; ++i) {
                if (R ? !o.isfluid(i, j, kk) : !cfg_solved(o, 1, i, j, kk)) continue;
                const double facexz = o.xp[i] * o.zp[kk];
                double ap = o.vp[o.IV(i, j, kk)];
                if (ap == 0) ap = 1e-10;
                const double corr = (PP[o.IP(i, j - 1, kk)] - PP[o.IP(i, j, kk)]) * facexz / ap;
                o.v[o.IV(i, j, kk)] = o.v[o.IV(i, j, kk)] + corr;
            }
    for (long kk = 1; kk <= o.nz + (R ? 0 : 1); ++kk)
        for (long j = 1; j <= o.ny; ++j)
            for (long i = 1; i <= o.nx; ++i) {
                if (R ? !o.isfluid(i, j, kk) : !cfg_solved(o, 2, i, j, kk)) continue;
                const double facexy = o.xp[i] * o.yp[j];
                double ap = o.wp[o.IW(i, j, kk)];
                if (ap == 0) ap = 1e-10;
                const double corr = (PP[o.IP(i, j, kk - 1)] - PP[o.IP(i, j, kk)]) * facexy / ap;
                o.w[o.IW(i, j, kk)] = o.w[o.IW(i, j, kk)] + corr;
            }
    for (long kk = 1; kk <= o.nz; ++kk)
        for (long j = 1; j <= o.ny; ++j)
            for (long i = 1; i <= o.nx; ++i) {
                if (!o.isfluid(i, j, kk)) continue;
                const long c = o.IC(i - 1, j - 1, kk - 1);
                o.p[c] += o.prm.relax_p * PP[o.IP(i - (planted(3) ? 1 : 0), j, kk)];
                o.ptemp[o.IP(i, j, kk)] = o.p[c];
            }
    std::fill(o.pprime.begin(), o.pprime.end(), 0.0);
    if (!R && !planted(12)) bc_cfg_outflow(o, 0);
}

/* ------------------------------------------------------------ energy ---- */

/* solveTemp point update (fluid_solver.h:807-846, Q11 convection without 1/(2 dx)
 * using the west/south/bottom face velocity). */
double temp_point_ref(const Orc& o, long i, long j, long kk) {
    const std::vector<double>& T = o.tf;
    const double xp = o.xp[i], xw = o.xw[i];
    const double yp = o.yp[j], ys = o.ys[j];
    const double zp = o.zp[kk], zf = o.zf[kk];
    const double sx = planted(7) ? (xp + o.xe[i]) / 2.0 : (xp + xw) / 2.0, sy = (yp + ys) / 2.0, sz = (zp + zf) / 2.0;
    const long c = o.IP(i, j, kk), sY = o.nx + 2, sZ = (long)(o.nx + 2) * (o.ny + 2);
    const double a = o.prm.alpha_f;
    const double conv = -o.u[o.IU(i + (planted(6) ? 1 : 0), j, kk)] * (T[c + 1] - T[c - 1])
                        - o.v[o.IV(i, j, kk)] * (T[c + sY] - T[c - sY])
                        - o.w[o.IW(i, j, kk)] * (T[c + sZ] - T[c - sZ]);
    const double diff = a * (T[c + 1] + T[c - 1]) / sx / sx
                        + a * (T[c + sY] + T[c - sY]) / sy / sy
                        + a * (T[c + sZ] + T[c - sZ]) / sz / sz;
    const double denom = 2 * a * (1 / sx / sx + 1 / sy / sy + 1 / sz / sz);
    return (conv + diff) / denom;
}

/* Mode C conjugate energy point update (this build): upwind convection with face
 * mass fluxes rho A u, harmonic-mean face conductance A/(dP/2kP + dN/2kN), source
 * q V; ghosts use dN = dP, kN = kP (the reference's ghost convention). */
double temp_coeffs_cfg(const Orc& o, long i, long j, long kk, double a[6], double* bsrc) {
    const long sY = o.nx + 2, sZ = (long)(o.nx + 2) * (o.ny + 2);
    const long c = o.IP(i, j, kk);
    const double xp = o.xp[i], yp = o.yp[j], zp = o.zp[kk];
    const double Axy = xp * yp, Axz = xp * zp, Ayz = yp * zp;
    const long cc = o.IC(i - 1, j - 1, kk - 1);
    const double kP = o.k[cc];
    auto kof = [&](long ii, long jj, long k2) {
        if (ii < 1 || ii > o.nx || jj < 1 || jj > o.ny || k2 < 1 || k2 > o.nz) return kP;
        return o.k[o.IC(ii - 1, jj - 1, k2 - 1)];
    };
    auto wid = [&](const std::vector<double>& g, long n, long ii, double self) {
        return (ii < 1 || ii > n) ? (planted(9) ? 0.0 : self) : g[ii];
    };
    const double dW = wid(o.xp, o.nx, i - 1, xp), dE = wid(o.xp, o.nx, i + 1, xp);
    const double dS = wid(o.yp, o.ny, j - 1, yp), dN = wid(o.yp, o.ny, j + 1, yp);
    const double dB = wid(o.zp, o.nz, kk - 1, zp), dT = wid(o.zp, o.nz, kk + 1, zp);
    const double DW = Ayz / (0.5 * xp / kP + 0.5 * dW / kof(i - 1, j, kk));
    const double DE = Ayz / (0.5 * xp / kP + 0.5 * dE / kof(i + 1, j, kk));
    const double DS = Axz / (0.5 * yp / kP + 0.5 * dS / kof(i, j - 1, kk));
    const double DN = Axz / (0.5 * yp / kP + 0.5 * dN / kof(i, j + 1, kk));
    const double DB = Axy / (0.5 * zp / kP + 0.5 * dB / kof(i, j, kk - 1));
    const double DT = Axy / (0.5 * zp / kP + 0.5 * dT / kof(i, j, kk + 1));
    const double rho = o.prm.rho;
    const double FW = rho * Ayz * o.u[o.IU(i, j, kk)], FE = rho * Ayz * o.u[o.IU(i + 1, j, kk)];
    const double FS = rho * Axz * o.v[o.IV(i, j, kk)], FN = rho * Axz * o.v[o.IV(i, j + 1, kk)];
    const double FB = rho * Axy * o.w[o.IW(i, j, kk)], FT = rho * Axy * o.w[o.IW(i, j, kk + 1)];
    a[0] = DE + pos(-FE);  // i+1
    a[1] = DW + pos(FW);   // i-1
    a[2] = DN + pos(-FN);  // j+1
    a[3] = DS + pos(FS);   // j-1
    a[4] = DB + pos(FB);   // k-1
    a[5] = DT + pos(-FT);  // k+1
    *bsrc = o.q[cc] * (xp * yp * zp);
    (void)c; (void)sY; (void)sZ;
    return a[0] + a[1] + a[2] + a[3] + a[4] + a[5];
}

double temp_point_cfg(const Orc& o, long i, long j, long kk) {
    double a[6], bs;
    const double ap = temp_coeffs_cfg(o, i, j, kk, a, &bs);
    const long c = o.IP(i, j, kk), sY = o.nx + 2, sZ = (long)(o.nx + 2) * (o.ny + 2);
    const std::vector<double>& T = o.tf;
    return (a[0] * T[c + 1] + a[1] * T[c - 1] + a[2] * T[c + sY] + a[3] * T[c - sY] +
            a[4] * T[c - sZ] + a[5] * T[c + sZ] + bs) / ap;
}

double energy_sweep(Orc& o) {
    const bool R = o.modeR();
    const double relax = o.prm.relax_t;
    double sum = 0.0;
    auto visit = [&](long i, long j, long kk) {
        const long c = o.IP(i, j, kk);
        if (R ? !o.fl[c] : !(o.fl[c] || o.so[c])) return;
        const double nv = R ? temp_point_ref(o, i, j, kk) : temp_point_cfg(o, i, j, kk);
        const double old = o.tf[c];
        o.tf[c] = nv * relax + (1 - relax) * old;
        sum += (o.tf[c] - old) * (o.tf[c] - old);
    };
    if (o.prm.order == HS_ORDER_RB) {
        for (int colour = 0; colour < 2; ++colour)
            for (long kk = 1; kk <= o.nz; ++kk)
                for (long j = 1; j <= o.ny; ++j)
                    for (long i = 1; i <= o.nx; ++i)
                        if (((i + j + kk) & 1) == colour) visit(i, j, kk);
    } else {
        for (long kk = 1; kk <= o.nz; ++kk)
            for (long j = 1; j <= o.ny; ++j)
                for (long i = 1; i <= o.nx; ++i) visit(i, j, kk);
    }
    return sum;
}

/* solveTemp control loop fluid_solver.h:797-880 (Q6 accumulator in Mode R). */
void fluid_energy(Orc& o, int* sweeps, double* resid) {
    const bool R = o.modeR();
    double ncell = 0;
    if (R) ncell = (double)o.nx * o.ny * o.nz;
    else for (long c = 0; c < o.NP_(); ++c) ncell += (o.fl[c] || o.so[c]) ? 1 : 0;
    ncell = std::max(1.0, ncell);
    double residual = 1, global = R ? 1.0 : std::numeric_limits<double>::infinity();
    int step = 0;
    while (step < o.prm.t_sweeps_max) {
        if (global < o.prm.t_tol) break;
        const double s = energy_sweep(o);
        step++;
        if (R) { residual = (residual + s) / ncell; global = std::sqrt(residual); }
        else { global = std::sqrt(s / ncell); bc_cfg_outflow(o, 3); }
    }
    if (sweeps) *sweeps = step;
    if (resid) *resid = global;
}

/* MobSolver::solveTemperature mobs_solver.h:467-500 (uniform 6-neighbour mean, Q12;
 * residual2 = sqrt((residual2 + sum)/N) never reset; stop below CONVERGENCE2 = 1e-3). */
void solid_solve(Orc& o, int* sweeps, double* resid) {
    if (!o.has_solid) { if (sweeps) *sweeps = 0; if (resid) *resid = 0; return; }
    const double N = (double)(o.sb[1] - o.sb[0] + 1) * (o.sb[3] - o.sb[2] + 1) * (o.sb[5] - o.sb[4] + 1);
    const long sY = o.nx + 2, sZ = (long)(o.nx + 2) * (o.ny + 2);
    const double relax = o.prm.relax_ts;
    std::vector<double>& T = o.ts;
    double residual2 = 1;
    int it = 0;
    for (; it < o.prm.ts_sweeps_max; ++it) {
        if (residual2 < o.prm.ts_tol) break;
        auto visit = [&](long i, long j, long kk) {
            const long c = o.IP(i, j, kk);
            if (!o.so[c]) return;
            const double old = T[c];
            double nv = 1.0 / 6.0 * (T[c + 1] + T[c - 1] + T[c - sY] + T[c + sY] + T[c - sZ] + T[c + sZ]);
            nv = planted(8) ? nv * (1 - relax) + relax * old : nv * relax + (1 - relax) * old;
            T[c] = nv;
            residual2 += (nv - old) * (nv - old);
        };
        if (o.prm.order == HS_ORDER_RB) {
            for (int colour = 0; colour < 2; ++colour)
                for (long kk = o.sb[4]; kk <= o.sb[5]; ++kk)
                    for (long j = o.sb[2]; j <= o.sb[3]; ++j)
                        for (long i = o.sb[0]; i <= o.sb[1]; ++i)
                            if (((i + j + kk) & 1) == colour) visit(i, j, kk);
        } else {
            for (long kk = o.sb[4]; kk <= o.sb[5]; ++kk)
                for (long j = o.sb[2]; j <= o.sb[3]; ++j)
                    for (long i = o.sb[0]; i <= o.sb[1]; ++i) visit(i, j, kk);
        }
        residual2 = std::sqrt(residual2 / N);
    }
    if (sweeps) *sweeps = it;
    if (resid) *resid = residual2;
}

/* solver.h:62-148 on one rank with one solid region: the fluid's t_halo (solid cell
 * positions) receives the solid's t_boundary values and vice versa, pairwise by
 * interface face in the order both init() loops produce. */
void exchange(Orc& o) {
    const size_t n = o.if_fluid.size();
    std::vector<double> to_solid(n), to_fluid(n);
    for (size_t a = 0; a < n; ++a) { to_solid[a] = o.tf[o.if_fluid[a]]; to_fluid[a] = o.ts[o.if_solid[a]]; }
    for (size_t a = 0; a < n; ++a) { o.tf[o.if_solid[a]] = to_fluid[a]; o.ts[o.if_fluid[a]] = to_solid[a]; }
}

/* interface lists in init() order: x faces (k,j,i), then y faces, then z faces
 * (fluid_solver.h:900-1229; the solid's loops mobs_solver.h:121-268 visit the same
 * faces in the same order). */
void build_interface(Orc& o) {
    o.if_fluid.clear(); o.if_solid.clear();
    auto add = [&](long fi, long fj, long fk, long si, long sj, long sk) {
        o.if_fluid.push_back(o.IP(fi, fj, fk)); o.if_solid.push_back(o.IP(si, sj, sk));
    };
    for (long kk = 1; kk <= o.nz; ++kk)
        for (long j = 1; j <= o.ny; ++j)
            for (long i = 1; i <= o.nx + 1; ++i) {  // face between cells i-1 and i
                const long a = i - 1, bcell = i;
                if (a < 1 || bcell > o.nx) continue;
                const bool fa = o.fl[o.IP(a, j, kk)], fb = o.fl[o.IP(bcell, j, kk)];
                if (fa && !fb) add(a, j, kk, bcell, j, kk);
                else if (!fa && fb) add(bcell, j, kk, a, j, kk);
            }
    for (long kk = 1; kk <= o.nz; ++kk)
        for (long j = 1; j <= o.ny + 1; ++j)
            for (long i = 1; i <= o.nx; ++i) {
                const long a = j - 1, bcell = j;
                if (a < 1 || bcell > o.ny) continue;
                const bool fa = o.fl[o.IP(i, a, kk)], fb = o.fl[o.IP(i, bcell, kk)];
                if (fa && !fb) add(i, a, kk, i, bcell, kk);
                else if (!fa && fb) add(i, bcell, kk, i, a, kk);
            }
    for (long kk = 1; kk <= o.nz + 1; ++kk)
        for (long j = 1; j <= o.ny; ++j)
            for (long i = 1; i <= o.nx; ++i) {
                const long a = kk - 1, bcell = kk;
                if (a < 1 || bcell > o.nz) continue;
                const bool fa = o.fl[o.IP(i, j, a)], fb = o.fl[o.IP(i, j, bcell)];
                if (fa && !fb) add(i, j, a, i, j, bcell);
                else if (!fa && fb) add(i, j, bcell, i, j, a);
            }
}

int find_solid_regions(Orc& o) {
    // connected components of solid cells (6-connectivity), mesh.cpp:298-454
    const long NC = o.NC_();
    std::vector<int> lab(NC, -1);
    int nreg = 0;
    std::vector<long> stack;
    for (long c0 = 0; c0 < NC; ++c0) {
        if (o.type[c0] != 1 || lab[c0] >= 0) continue;
        int bb[6] = {1 << 30, -1, 1 << 30, -1, 1 << 30, -1};
        stack.assign(1, c0); lab[c0] = nreg;
        while (!stack.empty()) {
            const long c = stack.back(); stack.pop_back();
            const long i = c % o.nx, j = (c / o.nx) % o.ny, kk = c / ((long)o.nx * o.ny);
            bb[0] = std::min<long>(bb[0], i + 1); bb[1] = std::max<long>(bb[1], i + 1);
            bb[2] = std::min<long>(bb[2], j + 1); bb[3] = std::max<long>(bb[3], j + 1);
            bb[4] = std::min<long>(bb[4], kk + 1); bb[5] = std::max<long>(bb[5], kk + 1);
            const long nb[6][3] = {{i - 1, j, kk}, {i + 1, j, kk}, {i, j - 1, kk}, {i, j + 1, kk}, {i, j, kk - 1}, {i, j, kk + 1}};
            for (auto& q : nb) {
                if (q[0] < 0 || q[0] >= o.nx || q[1] < 0 || q[1] >= o.ny || q[2] < 0 || q[2] >= o.nz) continue;
                const long cn = o.IC(q[0], q[1], q[2]);
                if (o.type[cn] == 1 && lab[cn] < 0) { lab[cn] = nreg; stack.push_back(cn); }
            }
        }
        if (nreg == 0) std::memcpy(o.sb, bb, sizeof bb);
        ++nreg;
    }
    return nreg;
}

void set_for_next(Orc& o) {
    if (o.modeR()) {
        // FluidSolver::setForNext fluid_solver.h:1634-1660
        bc_x_ref(o); bc_y_ref(o); bc_z_ref(o);
        cell_bc_ref(o);
        for (long kk = 0; kk < o.nz; ++kk)
            for (long j = 0; j < o.ny; ++j)
                for (long i = 0; i < o.nx; ++i) o.ptemp[o.IP(i + 1, j + 1, kk + 1)] = o.p[o.IC(i, j, kk)];
        // MobSolver::setForNext copies t_temp into the output copy only.
    } else {
        for (int comp = 0; comp < 4; ++comp) bc_cfg_outflow(o, comp);
    }
}

void energy_balance(Orc& o, hs_stats* st) {
    if (!st || o.modeR()) return;
    double heat = 0.0, res = 0.0;
    const long sY = o.nx + 2, sZ = (long)(o.nx + 2) * (o.ny + 2);
    for (long kk = 1; kk <= o.nz; ++kk)
        for (long j = 1; j <= o.ny; ++j)
            for (long i = 1; i <= o.nx; ++i) {
                const long c = o.IP(i, j, kk);
                if (!(o.fl[c] || o.so[c])) continue;
                double a[6], bs;
                const double ap = temp_coeffs_cfg(o, i, j, kk, a, &bs);
                const std::vector<double>& T = o.tf;
                const double r = ap * T[c] - (a[0] * T[c + 1] + a[1] * T[c - 1] + a[2] * T[c + sY] +
                                              a[3] * T[c - sY] + a[4] * T[c - sZ] + a[5] * T[c + sZ]) - bs;
                res += r; heat += bs;
            }
    st->heat_in = heat;
    st->energy_residual = std::fabs(res) / std::max(std::fabs(heat), 1e-300);
}

double* arr(Orc& o, int which) {
    switch (which) {
        case HS_A_U: return o.u.data(); case HS_A_V: return o.v.data(); case HS_A_W: return o.w.data();
        case HS_A_UP: return o.up.data(); case HS_A_VP: return o.vp.data(); case HS_A_WP: return o.wp.data();
        case HS_A_P: return o.p.data(); case HS_A_PTEMP: return o.ptemp.data();
        case HS_A_PPRIME: return o.pprime.data(); case HS_A_TF: return o.tf.data();
        case HS_A_TS: return o.ts.data(); case HS_A_Q: return o.q.data(); case HS_A_K: return o.k.data();
        case HS_A_B: return o.b.data();
    }
    return nullptr;
}

long arr_size(const Orc& o, int which) {
    switch (which) {
        case HS_A_U: case HS_A_UP: return o.NU_();
        case HS_A_V: case HS_A_VP: return o.NV_();
        case HS_A_W: case HS_A_WP: return o.NW_();
        case HS_A_P: case HS_A_Q: case HS_A_K: return o.NC_();
        case HS_A_PTEMP: case HS_A_PPRIME: case HS_A_TF: case HS_A_TS: case HS_A_B: return o.NP_();
    }
    return -1;
}

}  // namespace

struct or_ctx { Orc o; };

/* ------------------------------------------- energy equation as a system ---- */

/* CSR over the interior cells (C layout rows and columns). */
struct Csr {
    std::vector<long> rp, ci;
    std::vector<double> v, b;
};

/* Mode C energy operator on the frozen flow: rows are the sweep's point equations
 * ap T_P = sum a_nb T_nb + q V (temp_coeffs_cfg).  Neighbours outside the domain are
 * ghosts: Dirichlet ghosts move a_nb T_ghost to the right-hand side; the outflow ghost
 * (x = max when outflow is on) equals T_P at the sweep's fixed point, so its a_nb
 * leaves the diagonal instead. */
void energy_csr(const Orc& o, Csr& A) {
    const long n = o.NC_();
    A.rp.assign(n + 1, 0); A.ci.clear(); A.v.clear(); A.b.assign(n, 0.0);
    const long sY = o.nx + 2, sZ = (long)(o.nx + 2) * (o.ny + 2);
    for (long kk = 1; kk <= o.nz; ++kk)
        for (long j = 1; j <= o.ny; ++j)
            for (long i = 1; i <= o.nx; ++i) {
                const long row = o.IC(i - 1, j - 1, kk - 1);
                double a[6], bs;
                double ap = temp_coeffs_cfg(o, i, j, kk, a, &bs);
                const long cc = o.IP(i, j, kk);
                const long nbp[6] = {cc + 1, cc - 1, cc + sY, cc - sY, cc - sZ, cc + sZ};
                const long nbi[6][3] = {{i + 1, j, kk}, {i - 1, j, kk}, {i, j + 1, kk},
                                        {i, j - 1, kk}, {i, j, kk - 1}, {i, j, kk + 1}};
                double r = bs;
                std::vector<std::pair<long, double>> off;
                for (int m = 0; m < 6; ++m) {
                    const long ii = nbi[m][0], jj = nbi[m][1], k2 = nbi[m][2];
                    if (ii < 1 || ii > o.nx || jj < 1 || jj > o.ny || k2 < 1 || k2 > o.nz) {
                        if (m == 0 && o.prm.outflow) ap -= a[m];        // zero-gradient outflow ghost
                        else r += a[m] * o.tf[nbp[m]];                  // Dirichlet ghost
                        continue;
                    }
                    off.push_back({o.IC(ii - 1, jj - 1, k2 - 1), -a[m]});
                }
                A.rp[row] = (long)A.ci.size();
                A.ci.push_back(row); A.v.push_back(ap);
                for (auto& e : off) { A.ci.push_back(e.first); A.v.push_back(e.second); }
                A.b[row] = r;
            }
    A.rp[n] = (long)A.ci.size();
}

void csr_transpose(const Csr& A, Csr& T) {
    const long n = (long)A.rp.size() - 1;
    T.rp.assign(n + 1, 0); T.ci.assign(A.ci.size(), 0); T.v.assign(A.v.size(), 0.0); T.b = A.b;
    for (long e = 0; e < (long)A.ci.size(); ++e) T.rp[A.ci[e] + 1]++;
    for (long i = 0; i < n; ++i) T.rp[i + 1] += T.rp[i];
    std::vector<long> fill(T.rp.begin(), T.rp.end() - 1);
    for (long r = 0; r < n; ++r)
        for (long e = A.rp[r]; e < A.rp[r + 1]; ++e) {
            const long dst = fill[A.ci[e]]++;
            T.ci[dst] = r; T.v[dst] = A.v[e];
        }
}

void csr_mv(const Csr& A, const std::vector<double>& x, std::vector<double>& y) {
    const long n = (long)A.rp.size() - 1;
    y.assign(n, 0.0);
    for (long r = 0; r < n; ++r) {
        double s = 0.0;
        for (long e = A.rp[r]; e < A.rp[r + 1]; ++e) s += A.v[e] * x[A.ci[e]];
        y[r] = s;
    }
}

double vdot(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

/* BiCGStab (van der Vorst 1992), right-preconditioned by the diagonal, x0 = 0, stop
 * when ||r|| <= tol ||b||.  On a (near) breakdown — <r^, v> = 0, omega = 0 or
 * |<r^, r>| <= 1e-10 ||r^|| ||r|| — it restarts with the current residual as the
 * shadow vector (at most 8 times). */
void bicgstab_jacobi(const Csr& A, const std::vector<double>& b, double tol, int maxit, std::vector<double>& x,
                     hs_solve_info& info) {
    const long n = (long)b.size();
    std::vector<double> dinv(n, 1.0);
    for (long r = 0; r < n; ++r)
        for (long e = A.rp[r]; e < A.rp[r + 1]; ++e)
            if (A.ci[e] == r && A.v[e] != 0.0) dinv[r] = 1.0 / A.v[e];
    x.assign(n, 0.0);
    std::vector<double> r = b, rh = b, p(n, 0.0), v(n, 0.0), ph(n), s(n), sh(n), t(n);
    const double bb = vdot(b, b);
    double rho = bb, rho_old = bb, rhn = bb, alpha = 1.0, omega = 1.0, rr = bb;
    info = hs_solve_info{};
    if (bb == 0.0) { info.converged = 1; return; }
    const double eps = tol * tol * bb;
    int it = 0, restarts = 0;
    bool conv = false, brk = false, fresh = true;
    const bool progress = std::getenv("HS_ORACLE_PROGRESS") != nullptr;   // long solves report as they go
    auto restart = [&]() {
        if (restarts >= 8 || !(rr > 0.0)) return false;
        ++restarts;
        rh = r; rho = rr; rhn = rr; fresh = true;
        return true;
    };
    while (it < maxit) {
        const double beta = fresh ? 0.0 : (rho / rho_old) * (alpha / omega);
        fresh = false;
        for (long i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        for (long i = 0; i < n; ++i) ph[i] = dinv[i] * p[i];
        csr_mv(A, ph, v);
        const double sig = vdot(rh, v);
        if (sig == 0.0) { if (restart()) continue; brk = true; break; }
        alpha = rho / sig;
        for (long i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
        ++it;
        if (vdot(s, s) <= eps) {
            for (long i = 0; i < n; ++i) x[i] += alpha * ph[i];
            r = s; rr = vdot(r, r); conv = true; break;
        }
        for (long i = 0; i < n; ++i) sh[i] = dinv[i] * s[i];
        csr_mv(A, sh, t);
        const double tt = vdot(t, t), ts = vdot(t, s);
        const bool wbrk = !(tt > 0.0) || ts == 0.0;
        omega = wbrk ? 0.0 : ts / tt;
        for (long i = 0; i < n; ++i) { x[i] += alpha * ph[i] + omega * sh[i]; r[i] = s[i] - omega * t[i]; }
        rr = vdot(r, r);
        if (progress && it % 100 == 0) std::fprintf(stderr, "[oracle] bicgstab %d: relres %.3e\n", it, std::sqrt(rr / bb));
        if (rr <= eps) { conv = true; break; }
        const double rho_new = vdot(rh, r);
        if (wbrk || !(std::fabs(rho_new) > 1e-10 * std::sqrt(rhn * rr))) {
            if (restart()) continue;
            brk = true;
            break;
        }
        rho_old = rho;
        rho = rho_new;
    }
    std::vector<double> ax;
    csr_mv(A, x, ax);
    double e2 = 0.0;
    for (long i = 0; i < n; ++i) e2 += (b[i] - ax[i]) * (b[i] - ax[i]);
    info.iters = it;
    info.converged = conv ? 1 : 0;
    info.breakdown = brk ? 1 : 0;
    info.restarts = restarts;
    info.relres = std::sqrt(rr / bb);
    info.true_relres = std::sqrt(e2 / bb);
}

extern "C" {

/* test infrastructure: plant mistake n in the oracle (0 = none); see g_plant */
void or_test_plant(int32_t n) { g_plant = n; }

/* host threads of the p' CG (1, the default: the serial recurrence); see pcg_threads */
void or_set_threads(int32_t n) { g_threads = n < 1 ? 1 : n; }

int or_create(const hs_grid* g, const hs_params* p, or_ctx** out) {
    if (!g || !p || !out || g->nx < 3 || g->ny < 3 || g->nz < 3 || !g->xv || !g->yv || !g->zv || !g->cell_type)
        return HS_E_ARG;
    or_ctx* c = new or_ctx();
    Orc& o = c->o;
    o.nx = g->nx; o.ny = g->ny; o.nz = g->nz; o.prm = *p;
    o.X.assign(g->xv, g->xv + o.nx + 1);
    o.Y.assign(g->yv, g->yv + o.ny + 1);
    o.Z.assign(g->zv, g->zv + o.nz + 1);
    o.type.assign(g->cell_type, g->cell_type + o.NC_());
    axis_geom(o.X, o.nx, o.xp, o.xw, o.xe);
    axis_geom(o.Y, o.ny, o.yp, o.ys, o.yn);
    axis_geom(o.Z, o.nz, o.zp, o.zf, o.zb);
    o.fl.assign(o.NP_(), 0); o.so.assign(o.NP_(), 0);
    for (long kk = 1; kk <= o.nz; ++kk)
        for (long j = 1; j <= o.ny; ++j)
            for (long i = 1; i <= o.nx; ++i) {
                const uint8_t t = o.type[o.IC(i - 1, j - 1, kk - 1)];
                o.fl[o.IP(i, j, kk)] = t == 0; o.so[o.IP(i, j, kk)] = t == 1;
            }
    const int nreg = find_solid_regions(o);
    o.has_solid = nreg > 0;
    if (o.modeR()) {
        // one fluid region spanning the domain and at most one solid region (the
        // reference's positional exchange is only defined for that, quirk Q16)
        if (nreg > 1) { delete c; return HS_E_ARG; }
        bool any[6] = {false, false, false, false, false, false};
        for (long kk = 1; kk <= o.nz; ++kk)
            for (long j = 1; j <= o.ny; ++j)
                for (long i = 1; i <= o.nx; ++i)
                    if (o.fl[o.IP(i, j, kk)]) {
                        any[0] |= i == 1; any[1] |= i == o.nx; any[2] |= j == 1;
                        any[3] |= j == o.ny; any[4] |= kk == 1; any[5] |= kk == o.nz;
                    }
        for (bool a : any) if (!a) { delete c; return HS_E_ARG; }
    }
    o.u.assign(o.NU_(), 0.0); o.up.assign(o.NU_(), 0.0);
    o.v.assign(o.NV_(), 0.0); o.vp.assign(o.NV_(), 0.0);
    o.w.assign(o.NW_(), 0.0); o.wp.assign(o.NW_(), 0.0);
    o.p.assign(o.NC_(), 0.0);
    o.ptemp.assign(o.NP_(), 0.0); o.pprime.assign(o.NP_(), 0.0); o.b.assign(o.NP_(), 0.0);
    o.tf.assign(o.NP_(), o.modeR() ? o.prm.t_fluid_init : o.prm.t_wall);
    o.ts.assign(o.NP_(), o.prm.t_solid_init);
    o.q.assign(o.NC_(), 0.0);
    o.k.assign(o.NC_(), 0.0);
    for (long c2 = 0; c2 < o.NC_(); ++c2) o.k[c2] = o.type[c2] == 1 ? o.prm.alpha_f * o.prm.k_ratio : o.prm.alpha_f;
    if (!o.modeR()) {
        // conjugate field: solid cells start at the fluid's initial temperature too
        for (long c2 = 0; c2 < o.NP_(); ++c2) o.tf[c2] = o.prm.t_fluid_init;
        // Dirichlet ghosts at the wall temperature
        for (long kk = 0; kk < o.nz + 2; ++kk)
            for (long j = 0; j < o.ny + 2; ++j)
                for (long i = 0; i < o.nx + 2; ++i)
                    if (i == 0 || j == 0 || kk == 0 || i == o.nx + 1 || j == o.ny + 1 || kk == o.nz + 1)
                        o.tf[o.IP(i, j, kk)] = o.prm.t_wall;
    }
    build_interface(o);
    *out = c;
    return HS_OK;
}

void or_destroy(or_ctx* c) { delete c; }

int64_t or_array_size(const or_ctx* c, int32_t which) { return c ? arr_size(c->o, which) : -1; }

/* FluidSolver::init fluid_solver.h:882-1237 (+ MobSolver::init). */
int or_init(or_ctx* c) {
    if (!c) return HS_E_ARG;
    Orc& o = c->o;
    if (o.modeR()) {
        bc_x_ref(o); bc_y_ref(o); bc_z_ref(o);
        cell_bc_ref(o);
        // interface faces start at zero velocity (:1005, :1116, :1225)
        for (long kk = 1; kk <= o.nz; ++kk)
            for (long j = 1; j <= o.ny; ++j)
                for (long i = 1; i <= o.nx + 1; ++i) {
                    const long a = i - 1, bcell = i;
                    if (a < 1 || bcell > o.nx) continue;
                    if (o.fl[o.IP(a, j, kk)] != o.fl[o.IP(bcell, j, kk)] &&
                        (o.so[o.IP(a, j, kk)] || o.so[o.IP(bcell, j, kk)]))
                        o.u[o.IU(i, j, kk)] = 0;
                }
        for (long kk = 1; kk <= o.nz; ++kk)
            for (long j = 1; j <= o.ny + 1; ++j)
                for (long i = 1; i <= o.nx; ++i) {
                    const long a = j - 1, bcell = j;
                    if (a < 1 || bcell > o.ny) continue;
                    if (o.fl[o.IP(i, a, kk)] != o.fl[o.IP(i, bcell, kk)]) o.v[o.IV(i, j, kk)] = 0;
                }
        for (long kk = 1; kk <= o.nz + 1; ++kk)
            for (long j = 1; j <= o.ny; ++j)
                for (long i = 1; i <= o.nx; ++i) {
                    const long a = kk - 1, bcell = kk;
                    if (a < 1 || bcell > o.nz) continue;
                    if (o.fl[o.IP(i, j, a)] != o.fl[o.IP(i, j, bcell)]) o.w[o.IW(i, j, kk)] = 0;
                }
    } else {
        bc_cfg_init(o);
    }
    o.inited = true;
    return HS_OK;
}

int or_momentum(or_ctx* c, int32_t comp, int32_t* sweeps, double* residual) {
    if (!c || comp < 0 || comp > 2) return HS_E_ARG;
    int s; double r;
    momentum(c->o, comp, &s, &r);
    if (sweeps) *sweeps = s;
    if (residual) *residual = r;
    return HS_OK;
}

int or_pressure_correction(or_ctx* c, hs_stats* st) {
    if (!c) return HS_E_ARG;
    pressure_correction(c->o, st);
    return HS_OK;
}

int or_correct(or_ctx* c) {
    if (!c) return HS_E_ARG;
    correct(c->o);
    return HS_OK;
}

int or_fluid_energy(or_ctx* c, int32_t* sweeps, double* residual) {
    if (!c) return HS_E_ARG;
    int s; double r;
    fluid_energy(c->o, &s, &r);
    if (sweeps) *sweeps = s;
    if (residual) *residual = r;
    return HS_OK;
}

/* FluidSolver::solve fluid_solver.h:1401-1633. */
int or_fluid_solve(or_ctx* c, hs_stats* st) {
    if (!c) return HS_E_ARG;
    if (!c->o.inited) return HS_E_STATE;
    Orc& o = c->o;
    hs_stats loc{};
    for (int comp = 0; comp < 3; ++comp) momentum(o, comp, &loc.mom_sweeps[comp], &loc.mom_residual[comp]);
    pressure_correction(o, &loc);
    correct(o);
    fluid_energy(o, &loc.t_sweeps, &loc.t_residual);
    energy_balance(o, &loc);
    if (st) {
        // keep solid figures the caller may already hold
        loc.ts_sweeps = st->ts_sweeps; loc.ts_residual = st->ts_residual;
        *st = loc;
    }
    return HS_OK;
}

int or_solid_solve(or_ctx* c, hs_stats* st) {
    if (!c) return HS_E_ARG;
    if (!c->o.inited) return HS_E_STATE;
    int s = 0; double r = 0;
    if (c->o.modeR()) solid_solve(c->o, &s, &r);
    if (st) { st->ts_sweeps = s; st->ts_residual = r; }
    return HS_OK;
}

int or_exchange(or_ctx* c) {
    if (!c) return HS_E_ARG;
    if (c->o.modeR()) exchange(c->o);
    return HS_OK;
}

int or_set_for_next(or_ctx* c) {
    if (!c) return HS_E_ARG;
    set_for_next(c->o);
    return HS_OK;
}

int or_cell_bc(or_ctx* c) {
    if (!c) return HS_E_ARG;
    cell_bc_ref(c->o);
    return HS_OK;
}

int or_step(or_ctx* c, hs_stats* st) {
    hs_stats loc{};
    int e = or_fluid_solve(c, &loc);
    if (e) return e;
    if ((e = or_solid_solve(c, &loc))) return e;
    if ((e = or_exchange(c))) return e;
    if ((e = or_set_for_next(c))) return e;
    if (st) *st = loc;
    return HS_OK;
}

int or_energy_balance(or_ctx* c, double* heat_in, double* residual) {
    if (!c || !heat_in || !residual) return HS_E_ARG;
    if (c->o.modeR()) return HS_E_ARG;
    hs_stats loc{};
    energy_balance(c->o, &loc);
    *heat_in = loc.heat_in;
    *residual = loc.energy_residual;
    return HS_OK;
}

int or_run(or_ctx* c, int32_t n, hs_stats* st) {
    for (int32_t it = 0; it < n; ++it) {
        const int e = or_step(c, st ? st + it : nullptr);
        if (e) return e;
    }
    return HS_OK;
}

int or_pressure_apply(or_ctx* c, const double* x, double* y) {
    if (!c || !x || !y) return HS_E_ARG;
    Orc& o = c->o;
    if (!o.have_matrix) return HS_E_STATE;
    std::vector<double> xv(x, x + o.NP_()), yv;
    for (long i = 0; i < o.NP_(); ++i) if (!o.fl[i]) xv[i] = 0;
    apply_A(o, xv, yv);
    std::memcpy(y, yv.data(), sizeof(double) * o.NP_());
    return HS_OK;
}

int or_pressure_solve(or_ctx* c, const double* b, double* x, hs_stats* st) {
    if (!c || !b || !x) return HS_E_ARG;
    Orc& o = c->o;
    if (!o.have_matrix) return HS_E_STATE;
    std::vector<double> bv(b, b + o.NP_()), xv;
    for (long i = 0; i < o.NP_(); ++i) if (!o.fl[i]) bv[i] = 0;
    hs_stats loc{};
    pcg(o, bv, xv, &loc);
    if (st) { *st = loc; st->p_true_relres = true_relres(o, bv, xv); st->mass_imbalance = std::sqrt(dot_fluid(o, bv, bv)); }
    std::memcpy(x, xv.data(), sizeof(double) * o.NP_());
    return HS_OK;
}

int or_get_array(or_ctx* c, int32_t which, double* host, int64_t n) {
    if (!c || !host) return HS_E_ARG;
    const long s = arr_size(c->o, which);
    if (s < 0 || n != s) return HS_E_ARG;
    std::memcpy(host, arr(c->o, which), sizeof(double) * s);
    return HS_OK;
}

int or_put_array(or_ctx* c, int32_t which, const double* host, int64_t n) {
    if (!c || !host) return HS_E_ARG;
    const long s = arr_size(c->o, which);
    if (s < 0 || n != s) return HS_E_ARG;
    std::memcpy(arr(c->o, which), host, sizeof(double) * s);
    return HS_OK;
}

int or_get_fields(or_ctx* c, double* u, double* v, double* w, double* p, double* t) {
    if (!c) return HS_E_ARG;
    Orc& o = c->o;
    const long nx = o.nx, ny = o.ny, nz = o.nz;
    for (long kk = 0; kk < nz; ++kk)
        for (long j = 0; j < ny; ++j) {
            for (long i = 0; i <= nx; ++i) if (u) u[i + (nx + 1) * (j + ny * kk)] = o.u[o.IU(i + 1, j + 1, kk + 1)];
            for (long i = 0; i < nx; ++i) {
                const long cc = o.IC(i, j, kk);
                if (p) p[cc] = o.p[cc];
                if (t) {
                    const long c2 = o.IP(i + 1, j + 1, kk + 1);
                    t[cc] = (o.modeR() && o.so[c2]) ? o.ts[c2] : o.tf[c2];
                }
            }
        }
    for (long kk = 0; kk < nz; ++kk)
        for (long j = 0; j <= ny; ++j)
            for (long i = 0; i < nx; ++i) if (v) v[i + nx * (j + (ny + 1) * kk)] = o.v[o.IV(i + 1, j + 1, kk + 1)];
    for (long kk = 0; kk <= nz; ++kk)
        for (long j = 0; j < ny; ++j)
            for (long i = 0; i < nx; ++i) if (w) w[i + nx * (j + ny * kk)] = o.w[o.IW(i + 1, j + 1, kk + 1)];
    return HS_OK;
}

int or_set_source(or_ctx* c, const double* q) {
    if (!c || !q) return HS_E_ARG;
    c->o.q.assign(q, q + c->o.NC_());
    return HS_OK;
}

int or_set_conductivity(or_ctx* c, const double* k) {
    if (!c || !k) return HS_E_ARG;
    c->o.k.assign(k, k + c->o.NC_());
    return HS_OK;
}

int or_export_csr(or_ctx* c, int64_t* nrows, int64_t* nnz, int64_t* rowptr, int64_t* cols,
                  double* vals, double* rhs, int64_t* row_cell) {
    if (!c) return HS_E_ARG;
    Orc& o = c->o;
    if (!o.have_matrix) return HS_E_STATE;
    std::vector<long> row_of(o.NP_(), -1);
    long nr = 0;
    for (long kk = 1; kk <= o.nz; ++kk)
        for (long j = 1; j <= o.ny; ++j)
            for (long i = 1; i <= o.nx; ++i)
                if (o.isfluid(i, j, kk)) row_of[o.IP(i, j, kk)] = nr++;
    long nz_ = 0, r = 0;
    for (long kk = 1; kk <= o.nz; ++kk)
        for (long j = 1; j <= o.ny; ++j)
            for (long i = 1; i <= o.nx; ++i) {
                if (!o.isfluid(i, j, kk)) continue;
                const long cc = o.IP(i, j, kk);
                double cB, cS, cW, cE, cN, cT;
                if (o.modeR()) { cW = cE = o.ax[cc]; cS = cN = o.ay[cc]; cB = cT = o.az[cc]; }
                else {
                    cW = o.ax[o.IU(i, j, kk)]; cE = o.ax[o.IU(i + 1, j, kk)];
                    cS = o.ay[o.IV(i, j, kk)]; cN = o.ay[o.IV(i, j + 1, kk)];
                    cB = o.az[o.IW(i, j, kk)]; cT = o.az[o.IW(i, j, kk + 1)];
                }
                const double d = o.modeR() ? (cE + cW + cN + cS + cT + cB) : (cW + cE + cS + cN + cB + cT);
                if (rowptr) rowptr[r] = nz_;
                if (rhs) rhs[r] = o.b[cc];
                if (row_cell) row_cell[r] = cc;
                const long nbs[6][3] = {{i, j, kk - 1}, {i, j - 1, kk}, {i - 1, j, kk}, {i + 1, j, kk}, {i, j + 1, kk}, {i, j, kk + 1}};
                const double cf[6] = {cB, cS, cW, cE, cN, cT};
                for (int m = 0; m < 3; ++m)
                    if (o.isfluid(nbs[m][0], nbs[m][1], nbs[m][2])) {
                        if (cols) { cols[nz_] = row_of[o.IP(nbs[m][0], nbs[m][1], nbs[m][2])]; vals[nz_] = -cf[m]; }
                        ++nz_;
                    }
                if (cols) { cols[nz_] = r; vals[nz_] = d; }
                ++nz_;
                for (int m = 3; m < 6; ++m)
                    if (o.isfluid(nbs[m][0], nbs[m][1], nbs[m][2])) {
                        if (cols) { cols[nz_] = row_of[o.IP(nbs[m][0], nbs[m][1], nbs[m][2])]; vals[nz_] = -cf[m]; }
                        ++nz_;
                    }
                ++r;
            }
    if (rowptr) rowptr[r] = nz_;
    if (nrows) *nrows = nr;
    if (nnz) *nnz = nz_;
    return HS_OK;
}

int or_export_energy_csr(or_ctx* c, int64_t* nnz, int64_t* rowptr, int64_t* cols, double* vals, double* rhs) {
    if (!c) return HS_E_ARG;
    Orc& o = c->o;
    if (o.modeR()) return HS_E_STATE;
    Csr A;
    energy_csr(o, A);
    if (nnz) *nnz = (int64_t)A.ci.size();
    if (rowptr) for (size_t i = 0; i < A.rp.size(); ++i) rowptr[i] = A.rp[i];
    if (cols) for (size_t i = 0; i < A.ci.size(); ++i) { cols[i] = A.ci[i]; vals[i] = A.v[i]; }
    if (rhs) for (size_t i = 0; i < A.b.size(); ++i) rhs[i] = A.b[i];
    return HS_OK;
}

int or_energy_solve(or_ctx* c, double tol, int32_t maxit, hs_solve_info* info) {
    if (!c) return HS_E_ARG;
    Orc& o = c->o;
    if (o.modeR()) return HS_E_ARG;
    if (!o.inited) return HS_E_STATE;
    if (!(tol > 0.0) || maxit < 1) return HS_E_ARG;
    Csr A;
    energy_csr(o, A);
    std::vector<double> x;
    hs_solve_info inf{};
    bicgstab_jacobi(A, A.b, tol, maxit, x, inf);
    if (info) *info = inf;
    for (long kk = 1; kk <= o.nz; ++kk)
        for (long j = 1; j <= o.ny; ++j)
            for (long i = 1; i <= o.nx; ++i) o.tf[o.IP(i, j, kk)] = x[o.IC(i - 1, j - 1, kk - 1)];
    bc_cfg_outflow(o, 3);
    if (!std::isfinite(inf.true_relres)) return HS_E_NAN;
    return HS_OK;
}

/* Discrete adjoint: A^T lambda = dJ/dT on the explicit transpose, then
 *   dJ/dq_P = lambda_P V_P
 *   dJ/dk_P = -sum_faces dD/dk_P (lambda_P - lambda_nb)(T_P - T_nb)        (interior)
 *             -sum_bdry  (A/h_P) lambda_P (T_P - T_ghost)                   (Dirichlet)
 * from R(T) = A(k) T - b(q, k), dJ/dp = -lambda^T dR/dp. */
int or_adjoint_energy(or_ctx* c, const double* dJdT, double* dJdq, double* dJdk, double tol, int32_t maxit,
                      int32_t mem, hs_solve_info* info) {
    if (!c || !dJdT) return HS_E_ARG;
    Orc& o = c->o;
    if (o.modeR()) return HS_E_ARG;
    if (!o.inited) return HS_E_STATE;
    if (mem != HS_MEM_HOST || !(tol > 0.0) || maxit < 1) return HS_E_ARG;
    Csr A, At;
    energy_csr(o, A);
    csr_transpose(A, At);
    std::vector<double> g(dJdT, dJdT + o.NC_()), lam;
    hs_solve_info inf{};
    bicgstab_jacobi(At, g, tol, maxit, lam, inf);
    if (info) *info = inf;
    const long sY = o.nx + 2, sZ = (long)(o.nx + 2) * (o.ny + 2);
    for (long kk = 1; kk <= o.nz; ++kk)
        for (long j = 1; j <= o.ny; ++j)
            for (long i = 1; i <= o.nx; ++i) {
                const long ic = o.IC(i - 1, j - 1, kk - 1), cc = o.IP(i, j, kk);
                const double xp = o.xp[i], yp = o.yp[j], zp = o.zp[kk];
                const double lP = lam[ic];
                if (dJdq) dJdq[ic] = lP * (xp * yp * zp);
                if (!dJdk) continue;
                const double kP = o.k[ic], TP = o.tf[cc];
                // the six faces: area, this cell's width, neighbour (in domain?), its width, C and P offsets
                struct Face { double A, h; bool in; double hn; long dc, dp; bool outflow; };
                const Face f[6] = {
                    {yp * zp, xp, i > 1, i > 1 ? o.xp[i - 1] : 0.0, -1, -1, false},
                    {yp * zp, xp, i < o.nx, i < o.nx ? o.xp[i + 1] : 0.0, 1, 1, i == o.nx && o.prm.outflow},
                    {xp * zp, yp, j > 1, j > 1 ? o.yp[j - 1] : 0.0, -(long)o.nx, -sY, false},
                    {xp * zp, yp, j < o.ny, j < o.ny ? o.yp[j + 1] : 0.0, (long)o.nx, sY, false},
                    {xp * yp, zp, kk > 1, kk > 1 ? o.zp[kk - 1] : 0.0, -(long)o.nx * o.ny, -sZ, false},
                    {xp * yp, zp, kk < o.nz, kk < o.nz ? o.zp[kk + 1] : 0.0, (long)o.nx * o.ny, sZ, false}};
                double sum = 0.0;
                for (const Face& F : f) {
                    if (F.in) {
                        const double kN = o.k[ic + F.dc];
                        const double den = 0.5 * F.h / kP + 0.5 * F.hn / kN;
                        const double dD = F.A * (0.5 * F.h) / (kP * kP * den * den);
                        sum += dD * (lP - lam[ic + F.dc]) * (TP - o.tf[cc + F.dp]);
                    } else if (!F.outflow) {
                        sum += (F.A / F.h) * lP * (TP - o.tf[cc + F.dp]);
                    }
                }
                dJdk[ic] = -sum;
            }
    if (!std::isfinite(inf.true_relres)) return HS_E_NAN;
    return HS_OK;
}

int or_energy_solve_rhs(or_ctx* c, int32_t transpose, const double* b, double* x, double tol, int32_t maxit,
                        int32_t mem, hs_solve_info* info) {
    if (!c || !b || !x || mem != HS_MEM_HOST) return HS_E_ARG;
    Orc& o = c->o;
    if (o.modeR()) return HS_E_ARG;
    if (!o.inited) return HS_E_STATE;
    if (!(tol > 0.0) || maxit < 1) return HS_E_ARG;
    Csr A, At;
    energy_csr(o, A);
    const Csr* M = &A;
    if (transpose) { csr_transpose(A, At); M = &At; }
    std::vector<double> bv(b, b + o.NC_()), xv;
    hs_solve_info inf{};
    bicgstab_jacobi(*M, bv, tol, maxit, xv, inf);
    if (info) *info = inf;
    for (long i = 0; i < o.NC_(); ++i) x[i] = xv[i];
    return std::isfinite(inf.true_relres) ? HS_OK : HS_E_NAN;
}

int or_energy_apply(or_ctx* c, int32_t transpose, const double* x, double* y, int32_t mem) {
    if (!c || !x || !y || mem != HS_MEM_HOST) return HS_E_ARG;
    Orc& o = c->o;
    if (o.modeR()) return HS_E_ARG;
    if (!o.inited) return HS_E_STATE;
    Csr A, At;
    energy_csr(o, A);
    const Csr* M = &A;
    if (transpose) { csr_transpose(A, At); M = &At; }
    std::vector<double> xv(x, x + o.NC_()), yv;
    csr_mv(*M, xv, yv);
    for (long i = 0; i < o.NC_(); ++i) y[i] = yv[i];
    return HS_OK;
}

int or_get_cells(or_ctx* c, int32_t which, double* dst, int32_t mem) {
    if (!c || !dst || mem != HS_MEM_HOST) return HS_E_ARG;
    Orc& o = c->o;
    if (which == HS_A_Q || which == HS_A_K) {
        const std::vector<double>& s = which == HS_A_Q ? o.q : o.k;
        for (long i = 0; i < o.NC_(); ++i) dst[i] = s[i];
        return HS_OK;
    }
    if (which != HS_A_TF) return HS_E_ARG;
    for (long kk = 1; kk <= o.nz; ++kk)
        for (long j = 1; j <= o.ny; ++j)
            for (long i = 1; i <= o.nx; ++i) dst[o.IC(i - 1, j - 1, kk - 1)] = o.tf[o.IP(i, j, kk)];
    return HS_OK;
}

int or_put_cells(or_ctx* c, int32_t which, const double* src, int32_t mem) {
    if (!c || !src || mem != HS_MEM_HOST) return HS_E_ARG;
    Orc& o = c->o;
    if (which == HS_A_Q || which == HS_A_K) {
        std::vector<double>& s = which == HS_A_Q ? o.q : o.k;
        for (long i = 0; i < o.NC_(); ++i) s[i] = src[i];
        return HS_OK;
    }
    if (which != HS_A_TF) return HS_E_ARG;
    for (long kk = 1; kk <= o.nz; ++kk)
        for (long j = 1; j <= o.ny; ++j)
            for (long i = 1; i <= o.nx; ++i) o.tf[o.IP(i, j, kk)] = src[o.IC(i - 1, j - 1, kk - 1)];
    return HS_OK;
}

}  // extern "C"
