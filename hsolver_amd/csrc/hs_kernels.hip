// hs_kernels.hip — gfx950 kernels of the SIMPLE conjugate-heat-transfer step.
//
// Every point update is written in the same operation order as the reference loop
// it replaces (built with -ffp-contract=off), so a sweep run in the reference's
// lexicographic order reproduces the reference's values bit for bit:
//   lexicographic Gauss-Seidel on a 7-point stencil == hyperplane (i+j+k = s)
//   wavefront, because every point reads the updated i-1/j-1/k-1 neighbours
//   (plane s-1) and the old i+1/j+1/k+1 neighbours (plane s+1).
// Red-black ordering is the fast path: two fully parallel colour passes per sweep.
//
// Reductions are deterministic: every block stores a partial sum, then a one-block
// finish kernel (next on the stream) adds the partials in a fixed order and runs the
// scalar control step (CG alpha/beta/stop rule, Gauss-Seidel stop rule) on the
// device, so no host round trip is needed per iteration.
#include <cfloat>
#include <cmath>

#include "hs_internal.h"

#include "hs_device.h"

namespace {

using hsd::BX;
using hsd::BY;
using hsd::BT;
using hsd::CBX;
using hsd::CBY;
using hsd::CBT;
using hsd::FLAT;
using hsd::FLAT_BLOCKS;
using hsd::block_sum;
using hsd::put_partial;
using hsd::wave_sum;
using hsd::issued_together;

__device__ __forceinline__ double posd(double a) { return (a < 0.0) ? 0.0 : a; }

// Gauss-Seidel loop control of solveX/Y/Zmomentum :580-658, solveTemp :797-880
// (quirk 1: accumulator starts at 1, divided by N, never reset) and
// MobSolver::solveTemperature mobs_solver.h:470-499 (quirk 2).
__device__ void sweep_control(SweepCtl* w, double s) {
    w->step += 1;
    if (w->quirk == 1) {
        w->residual = (w->residual + s) / w->ncount;
        w->global = sqrt(w->residual);
    } else if (w->quirk == 2) {
        w->residual = sqrt((w->residual + s) / w->ncount);
        w->global = w->residual;
    } else {
        w->residual = s;
        w->global = sqrt(s / w->ncount);
    }
    w->active = (w->step < w->maxsteps) && !(w->global < w->tol) ? 1 : 0;
}

// ------------------------------------------------------------- fills ---
__global__ void k_fill(double* a, long n, double v) {
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) a[t] = v;
}
__global__ void k_copy(double* dst, const double* src, long n) {
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) dst[t] = src[t];
}

// --------------------------------------------------------------- BCs ---
// applyXBoundaryCondition fluid_solver.h:209-257 (region spans the domain: only the
// min-side ghost branch, Q10; east wall face zeroed).
__global__ void k_bc_x_ref(Dims d, double* u, double* v, double* w) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (j < d.ny + 2 && k < d.nz + 2) {
        u[idxU(d, 0, j, k)] = 0; v[idxV(d, 0, j, k)] = 0; w[idxW(d, 0, j, k)] = 0;
        if (j < d.ny && k < d.nz) u[idxU(d, d.nx + 1, j + 1, k + 1)] = 0;
    }
}
// applyYBoundaryCondition fluid_solver.h:284-322 (lid on the top interior u row, Q9).
__global__ void k_bc_y_ref(Dims d, double* u, double* v, double* w, double ulid) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (i < d.nx + 2 && k < d.nz + 2) {
        u[idxU(d, i, 0, k)] = 0; v[idxV(d, i, 0, k)] = 0; w[idxW(d, i, 0, k)] = 0;
    }
    __syncthreads();
    if (i < d.nx && k < d.nz) {
        v[idxV(d, i + 1, d.ny + 1, k + 1)] = 0;
        u[idxU(d, i + 1, d.ny, k + 1)] = ulid;
    }
}
// applyZBoundaryCondition fluid_solver.h:259-282.
__global__ void k_bc_z_ref(Dims d, double* u, double* v, double* w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i < d.nx + 2 && j < d.ny + 2) {
        u[idxU(d, i, j, 0)] = 0; v[idxV(d, i, j, 0)] = 0; w[idxW(d, i, j, 0)] = 0;
        if (i < d.nx && j < d.ny) w[idxW(d, i + 1, j + 1, d.nz + 1)] = 0;
    }
}

// applyCellBoundaryCondition fluid_solver.h:324-393.  The reference runs one sequential
// k,j,i loop in which boundary cell c copies its inward neighbour.  A neighbour later in
// the loop (the +x/+y/+z copies of min-side cells) is read before its own update, a
// neighbour earlier in the loop (the -x/-y/-z copies of max-side cells) after it.  For
// n >= 3 the dependency chains are at most x-max <- y-max <- z-max deep, so the loop is
// exactly: pass 0 every min-side cell and every z-max cell from a snapshot, pass 1 the
// y-max cells, pass 2 the x-max cells (each reading the current array).
__device__ __forceinline__ int cellbc_branch(const Dims& d, int gx, int gy, int gz) {
    if (gx == 0) return 1;
    if (gx == d.nx - 1) return 2;
    if (gy == 0) return 3;
    if (gy == d.ny - 1) return 4;
    if (gz == 0) return 5;
    if (gz == d.nz - 1) return 6;
    return 0;
}
__global__ void k_cellbc_pass(Dims d, double* p, const double* snap, double* pt, int pass) {
    // one thread per cell of the 6 boundary planes: launch over (nx*ny) x 6 planes style
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long nxy = (long)d.nx * d.ny, nxz = (long)d.nx * d.nz, nyz = (long)d.ny * d.nz;
    const long total = 2 * (nxy + nxz + nyz);
    if (t >= total) return;
    int gx, gy, gz;
    long r = t;
    if (r < 2 * nyz) { gx = (r < nyz) ? 0 : d.nx - 1; r %= nyz; gy = r % d.ny; gz = r / d.ny; }
    else if ((r -= 2 * nyz) < 2 * nxz) { gy = (r < nxz) ? 0 : d.ny - 1; r %= nxz; gx = r % d.nx; gz = r / d.nx; }
    else { r -= 2 * nxz; gz = (r < nxy) ? 0 : d.nz - 1; r %= nxy; gx = r % d.nx; gy = r / d.nx; }
    const int br = cellbc_branch(d, gx, gy, gz);
    // a cell appears on up to three planes: handle it on the plane of its branch only
    const bool on_x = (t < 2 * nyz);
    const bool on_y = !on_x && (t < 2 * nyz + 2 * nxz);
    const bool on_z = !on_x && !on_y;
    if ((br == 1 || br == 2) != on_x) return;
    if ((br == 3 || br == 4) != on_y) return;
    if ((br == 5 || br == 6) != on_z) return;
    const int want = (br == 4) ? 1 : (br == 2 ? 2 : 0);
    if (want != pass) return;
    const long c = idxC(d, gx, gy, gz);
    double val;
    switch (br) {
        case 1: val = snap[idxC(d, gx + 1, gy, gz)]; break;
        case 2: val = p[idxC(d, gx - 1, gy, gz)]; break;
        case 3: val = snap[idxC(d, gx, gy + 1, gz)]; break;
        case 4: val = p[idxC(d, gx, gy - 1, gz)]; break;
        case 5: val = snap[idxC(d, gx, gy, gz + 1)]; break;
        default: val = snap[idxC(d, gx, gy, gz - 1)]; break;
    }
    p[c] = val;
    const int i = gx, j = gy, k = gz, cx = d.nx, cz = d.nz;
    switch (br) {
        case 1: pt[idxP(d, i + 1, j + 1, k + 1)] = val; pt[idxP(d, i, j + 1, k + 1)] = val; break;
        case 2: pt[idxP(d, i + 1, j + 1, k + 1)] = val; pt[idxP(d, i + 2, j + 1, k + 1)] = val; break;
        case 3: pt[idxP(d, i + 1, j + 1, k + 1)] = val; pt[idxP(d, i + 1, j, k + 1)] = val; break;
        case 4:
            pt[idxP(d, i + 1, j + 1, k + 1)] = val; pt[idxP(d, i + 1, j + 2, k + 1)] = val;
            if (i + 1 == cx) pt[idxP(d, i + 2, j + 2, k + 1)] = val;
            if (k + 1 == cz) pt[idxP(d, i + 1, j + 2, k + 2)] = val;
            if (i + 1 == 1) pt[idxP(d, i, j + 2, k + 1)] = val;
            if (k + 1 == 1) pt[idxP(d, i + 1, j + 2, k)] = val;
            if (i + 1 == cx && k + 1 == cz) pt[idxP(d, i + 2, j + 2, k + 2)] = val;
            if (i + 1 == 1 && k + 1 == 1) pt[idxP(d, i, j + 2, k)] = val;
            break;
        case 5: pt[idxP(d, i + 1, j + 1, k + 1)] = val; pt[idxP(d, i + 1, j + 1, k)] = val; break;
        default: pt[idxP(d, i + 1, j + 1, k + 1)] = val; pt[idxP(d, i + 1, j + 1, k + 2)] = val; break;
    }
}
__global__ void k_zero_one(double* p) { p[0] = 0.0; }

// p_temp interior = p for every cell (setForNext fluid_solver.h:1655-1660).
__global__ void k_ptemp_all(Dims d, const double* p, double* pt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i < d.nx) pt[idxP(d, i + 1, j + 1, k + 1)] = p[idxC(d, i, j, k)];
}

// Mode C: inlet u on x=min faces of fluid cells, lid ghost row for u on y=max.
__global__ void k_bc_cfg_init(Dims d, double* u, const uint8_t* cls, double uin, int lid, double ulid) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (j >= 1 && j <= d.ny && k >= 1 && k <= d.nz)
        u[idxU(d, 1, j, k)] = (cls[idxP(d, 1, j, k)] == CELL_FLUID) ? uin : 0.0;
    if (lid && k < d.nz + 2)
        for (int i = j; i < d.nx + 3; i += blockDim.x * gridDim.x) u[idxU(d, i, d.ny + 1, k)] = ulid;
}
// Mode C zero-gradient outflow on x = max (u faces, v/w ghosts, temperature ghost).
__global__ void k_bc_cfg_outflow(Dims d, double* u, double* v, double* w, double* t, const uint8_t* cls, int comp) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1, k = blockIdx.y + 1;
    const int cx = d.nx;
    if (comp == 0) {
        if (j <= d.ny && k <= d.nz)
            u[idxU(d, cx + 1, j, k)] = (cls[idxP(d, cx, j, k)] == CELL_FLUID) ? u[idxU(d, cx, j, k)] : 0.0;
    } else if (comp == 1) {
        if (j >= 2 && j <= d.ny && k <= d.nz) v[idxV(d, cx + 1, j, k)] = v[idxV(d, cx, j, k)];
    } else if (comp == 2) {
        if (j <= d.ny && k >= 2 && k <= d.nz) w[idxW(d, cx + 1, j, k)] = w[idxW(d, cx, j, k)];
    } else {
        if (j <= d.ny && k <= d.nz) t[idxP(d, cx + 1, j, k)] = t[idxP(d, cx, j, k)];
    }
}

// interface faces start at zero velocity (fluid_solver.h:1005, :1116, :1225)
__global__ void k_zero_ib(Dims d, double* u, double* v, double* w, const uint8_t* cls) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i < 1 || j < 1 || k < 1) return;
    auto diff = [&](long a, long b) {
        const uint8_t ca = cls[a], cb = cls[b];
        return (ca == CELL_FLUID && cb == CELL_SOLID) || (ca == CELL_SOLID && cb == CELL_FLUID);
    };
    if (i >= 2 && i <= d.nx && j <= d.ny && k <= d.nz && diff(idxP(d, i - 1, j, k), idxP(d, i, j, k))) u[idxU(d, i, j, k)] = 0;
    if (j >= 2 && j <= d.ny && i <= d.nx && k <= d.nz && diff(idxP(d, i, j - 1, k), idxP(d, i, j, k))) v[idxV(d, i, j, k)] = 0;
    if (k >= 2 && k <= d.nz && i <= d.nx && j <= d.ny && diff(idxP(d, i, j, k - 1), idxP(d, i, j, k))) w[idxW(d, i, j, k)] = 0;
}

// ----------------------------------------------------------- momentum ---
struct MomArgs {
    Dims d;
    Geo g;
    double *u, *v, *w, *ap;
    const double* pt;
    const uint8_t* cls;
    double* dres;
    double c05, c2nu, relax;
    int modeR;
};

// One face of solveX/Y/Zmomentum (fluid_solver.h:596-625 / :513-542 / :427-457).
// Returns 1 if the face value was updated, 2 if only its diagonal was written, 0 if
// untouched; *dd receives the squared change.
template <int COMP>
__device__ __forceinline__ int mom_update(const MomArgs& a, long i, long j, long k, double* dd) {
    const Dims& d = a.d;
    // Every value the face can need is loaded before any is used, whatever the cell
    // classes turn out to be: one wait on memory per face instead of four (the classes,
    // the flux velocities, and the neighbours again after the a_P store, which the
    // compiler cannot move across).  Every index of the face box is valid for all of them.
    const long cN = COMP == 0 ? idxP(d, i - 1, j, k) : (COMP == 1 ? idxP(d, i, j - 1, k) : idxP(d, i, j, k - 1));
    const int c1 = a.cls[idxP(d, i, j, k)], c0 = a.cls[cN];
    const double xp = a.g.xp[i], xw = a.g.xw[i], xe = a.g.xe[i];
    const double yp = a.g.yp[j], ys = a.g.ys[j], yn = a.g.yn[j];
    const double zp = a.g.zp[k], zf = a.g.zf[k], zb = a.g.zb[k];
    const double* U = a.u;
    const double* V = a.v;
    const double* W = a.w;
    long s0, sy, sz;
    double* X;
    if (COMP == 0) { s0 = idxU(d, i, j, k); sy = d.sUy; sz = d.sUz; X = a.u; }
    else if (COMP == 1) { s0 = idxV(d, i, j, k); sy = d.sVy; sz = d.sVz; X = a.v; }
    else { s0 = idxW(d, i, j, k); sy = d.sPy; sz = d.sPz; X = a.w; }
    // own component: the face and its six neighbours
    const double x0 = X[s0], xW = X[s0 - 1], xE = X[s0 + 1], xS = X[s0 - sy], xN = X[s0 + sy], xB = X[s0 - sz],
                 xT = X[s0 + sz];
    // the other two components at the four corners each flux needs, and the pressures
    double q0, q1, q2, q3, r0, r1, r2, r3, pm, pp;
    if (COMP == 0) {
        q0 = V[idxV(d, i, j, k)]; q1 = V[idxV(d, i - 1, j, k)]; q2 = V[idxV(d, i, j + 1, k)]; q3 = V[idxV(d, i - 1, j + 1, k)];
        r0 = W[idxW(d, i, j, k + 1)]; r1 = W[idxW(d, i - 1, j, k + 1)]; r2 = W[idxW(d, i, j, k)]; r3 = W[idxW(d, i - 1, j, k)];
        pm = a.pt[idxP(d, i - 1, j, k)];
    } else if (COMP == 1) {
        q0 = U[idxU(d, i, j, k)]; q1 = U[idxU(d, i, j - 1, k)]; q2 = U[idxU(d, i + 1, j, k)]; q3 = U[idxU(d, i + 1, j - 1, k)];
        r0 = W[idxW(d, i, j, k + 1)]; r1 = W[idxW(d, i, j - 1, k + 1)]; r2 = W[idxW(d, i, j, k)]; r3 = W[idxW(d, i, j - 1, k)];
        pm = a.pt[idxP(d, i, j - 1, k)];
    } else {
        q0 = U[idxU(d, i, j, k)]; q1 = U[idxU(d, i, j, k - 1)]; q2 = U[idxU(d, i + 1, j, k)]; q3 = U[idxU(d, i + 1, j, k - 1)];
        r0 = V[idxV(d, i, j, k)]; r1 = V[idxV(d, i, j, k - 1)]; r2 = V[idxV(d, i, j + 1, k)]; r3 = V[idxV(d, i, j + 1, k - 1)];
        pm = a.pt[idxP(d, i, j, k - 1)];
    }
    pp = a.pt[idxP(d, i, j, k)];
    issued_together(c1, c0, xp, xw, xe, yp, ys, yn, zp, zf, zb, x0, xW, xE, xS, xN, xB, xT, q0, q1, q2, q3, r0, r1, r2, r3,
                    pm, pp);

    bool active, touch;
    if (a.modeR) {
        active = c1 == CELL_FLUID;
        touch = false;
    } else {
        const bool f1 = c1 == CELL_FLUID, f0 = c0 == CELL_FLUID;
        active = f0 && f1;
        touch = f0 || f1;
    }
    if (!active && !touch) return 0;
    const double Axy = xp * yp, Axz = xp * zp, Ayz = yp * zp;
    const double Dw = a.c2nu * Ayz / (xp + xw), De = a.c2nu * Ayz / (xp + xe);
    const double Ds = a.c2nu * Axz / (ys + yp), Dn = a.c2nu * Axz / (yn + yp);
    const double Db = a.c2nu * Axy / (zb + zp), Df = a.c2nu * Axy / (zf + zp);
    double Fmx, Fpx, Fmy, Fpy, Fmz, Fpz, b;
    const double c05 = a.c05;
    if (COMP == 0) {
        Fmx = c05 * Ayz * (x0 + xW);
        Fpx = c05 * Ayz * (x0 + xE);
        Fmy = c05 * Axz * (q0 + q1);
        Fpy = c05 * Axz * (q2 + q3);
        Fpz = c05 * Axy * (r0 + r1);
        Fmz = c05 * Axy * (r2 + r3);
        b = Ayz * (pm - pp);
    } else if (COMP == 1) {
        Fmy = c05 * Axz * (x0 + xS);
        Fpy = c05 * Axz * (x0 + xN);
        Fmx = c05 * Ayz * (q0 + q1);
        Fpx = c05 * Ayz * (q2 + q3);
        Fpz = c05 * Axy * (r0 + r1);
        Fmz = c05 * Axy * (r2 + r3);
        b = Axz * (pm - pp);
    } else {
        Fmz = c05 * Axy * (x0 + xB);
        Fpz = c05 * Axy * (x0 + xT);
        Fmx = c05 * Ayz * (q0 + q1);
        Fpx = c05 * Ayz * (q2 + q3);
        Fmy = c05 * Axz * (r0 + r1);
        Fpy = c05 * Axz * (r2 + r3);
        b = Axy * (pm - pp);
    }
    // reference upwind labels: y-momentum swaps x and y, z-momentum swaps z (Q7)
    const bool swx = a.modeR && COMP == 1, swy = a.modeR && COMP == 1, swz = a.modeR && COMP == 2;
    const double ae = De + (swx ? posd(-Fmx) : posd(-Fpx));
    const double aw = Dw + (swx ? posd(Fpx) : posd(Fmx));
    const double an = Dn + (swy ? posd(-Fmy) : posd(-Fpy));
    const double as = Ds + (swy ? posd(Fpy) : posd(Fmy));
    const double ab = Db + (swz ? posd(-Fmz) : posd(-Fpz));
    const double af = Df + (swz ? posd(Fpz) : posd(Fmz));
    const double ap = ae + aw + an + as + ab + af;
    a.ap[s0] = ap;
    if (!active) return 2;
    const double nv = (ae * xE + aw * xW + an * xN + as * xS + af * xB + ab * xT + b) / ap;
    const double val = a.relax * nv + (1 - a.relax) * x0;
    X[s0] = val;
    *dd = (val - x0) * (val - x0);
    return 1;
}

// Red-black colour pass over the face box [1..ni] x [1..nj] x [1..nk], two adjacent faces
// per thread along x: on every plane one of the two has the pass's colour, so every lane
// updates a face at every k (one face per thread idles half the lanes of each plane);
// tiles go to the XCDs in contiguous runs.  Per-block partial of the squared change; the
// finish kernel runs the control step after colour 1.
template <int COMP>
__global__ __launch_bounds__(CBT) void k_mom_rb(MomArgs a, int ni, int nj, int nk, int kc, int colour, int gx, int gy,
                                                int gz, double* partials, SweepCtl* wctl) {
    if (!wctl->active) return;
    const int tid = threadIdx.x + CBX * threadIdx.y;
    const hsd::Tile tl = hsd::tile_of(gx, gy, gz);
    const long ib = 1 + (long)tl.bx * (2 * CBX) + 2 * threadIdx.x, j = 1 + (long)tl.by * CBY + threadIdx.y;
    double acc = 0.0;
    if (ib <= ni && j <= nj) {
        const int k0 = 1 + tl.bz * kc, k1 = min(nk, k0 + kc - 1);
        for (int k = k0; k <= k1; ++k) {
            const long i = ib + ((((ib + j + k) & 1) != colour) ? 1 : 0);
            if (i > ni) continue;
            double dd = 0.0;
            if (mom_update<COMP>(a, i, j, k, &dd) == 1) acc += dd;
        }
    }
    put_partial<CBT>(acc, tid, blockIdx.x, partials + (colour ? gx * gy * gz : 0));
}

// One hyperplane i+j+k = s of the lexicographic sweep; dres gets the squared change.
template <int COMP>
__global__ __launch_bounds__(256) void k_mom_plane(MomArgs a, int ni, int nj, int nk, int s, SweepCtl* wctl, long nlayout) {
    if (!wctl->active) return;
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int j = 1 + (int)(t % nj), k = 1 + (int)(t / nj);
    if (k > nk) return;
    const int i = s - j - k;
    if (i < 1 || i > ni) return;
    double dd = 0.0;
    const int r = mom_update<COMP>(a, i, j, k, &dd);
    if (r == 1) {
        const long s0 = COMP == 0 ? idxU(a.d, i, j, k) : (COMP == 1 ? idxV(a.d, i, j, k) : idxW(a.d, i, j, k));
        a.dres[s0] = dd;
    }
    (void)nlayout;
}

// Partial sums of dres[0..n) (the sweep's squared changes); k_finish runs the control step.
__global__ __launch_bounds__(FLAT) void k_sum_control(const double* x, long n, double* partials, SweepCtl* wctl) {
    if (!wctl->active) return;
    double acc = 0.0;
    for (long t = blockIdx.x * (long)FLAT + threadIdx.x; t < n; t += (long)gridDim.x * FLAT) acc += x[t];
    put_partial<FLAT>(acc, threadIdx.x, blockIdx.x, partials);
}

// ------------------------------------------------------------- energy ---
struct TArgs {
    Dims d;
    Geo g;
    const double *u, *v, *w;
    double* t;
    const double *q, *kc;
    const uint8_t* cls;
    double* dres;
    double alpha, relax, rho;
    int modeR;
    int solid;   // 1: MobSolver::solveTemperature on solid cells (Mode R)
};

// solveTemp point (fluid_solver.h:807-846, Q11).
__device__ __forceinline__ double t_point_ref(const TArgs& a, long i, long j, long k) {
    const Dims& d = a.d;
    const double* T = a.t;
    const double xp = a.g.xp[i], xw = a.g.xw[i];
    const double yp = a.g.yp[j], ys = a.g.ys[j];
    const double zp = a.g.zp[k], zf = a.g.zf[k];
    const double sx = (xp + xw) / 2.0, sy = (yp + ys) / 2.0, sz = (zp + zf) / 2.0;
    const long c = idxP(d, i, j, k), sY = d.sPy, sZ = d.sPz;
    const double al = a.alpha;
    const double conv = -a.u[idxU(d, i, j, k)] * (T[c + 1] - T[c - 1])
                        - a.v[idxV(d, i, j, k)] * (T[c + sY] - T[c - sY])
                        - a.w[idxW(d, i, j, k)] * (T[c + sZ] - T[c - sZ]);
    const double diff = al * (T[c + 1] + T[c - 1]) / sx / sx
                        + al * (T[c + sY] + T[c - sY]) / sy / sy
                        + al * (T[c + sZ] + T[c - sZ]) / sz / sz;
    const double denom = 2 * al * (1 / sx / sx + 1 / sy / sy + 1 / sz / sz);
    return (conv + diff) / denom;
}

// Mode C conjugate energy row at an interior cell (harmonic face conductance from the two
// cells' half widths and conductivities, upwind convection from the face velocities; a face
// on the domain boundary takes the cell's own width and conductivity), together with the
// temperatures it couples.  Every value is loaded before any is used (one wait on memory
// per cell): a neighbour outside the domain is loaded at the clamped index and replaced by
// the cell's own values afterwards.
struct TRowC {
    int cls;
    double ap, cf[6], bs;
    double t0, tE, tW, tN, tS, tB, tT;
};
__device__ __forceinline__ TRowC t_row_cfg(const TArgs& a, long i, long j, long k) {
    const Dims& d = a.d;
    const long c = idxP(d, i, j, k), cc = idxC(d, i - 1, j - 1, k - 1);
    const double* T = a.t;
    const long iw = i > 1 ? i - 1 : i, ie = i < d.nx ? i + 1 : i;
    const long js = j > 1 ? j - 1 : j, jn = j < d.ny ? j + 1 : j;
    const long kb = k > 1 ? k - 1 : k, kt = k < d.nz ? k + 1 : k;
    TRowC r;
    r.cls = a.cls[c];
    const double xp = a.g.xp[i], yp = a.g.yp[j], zp = a.g.zp[k];
    const double gW = a.g.xp[i - 1], gE = a.g.xp[i + 1], gS = a.g.yp[j - 1], gN = a.g.yp[j + 1], gB = a.g.zp[k - 1],
                 gT = a.g.zp[k + 1];
    const double kP = a.kc[cc];
    const double kW0 = a.kc[idxC(d, iw - 1, j - 1, k - 1)], kE0 = a.kc[idxC(d, ie - 1, j - 1, k - 1)];
    const double kS0 = a.kc[idxC(d, i - 1, js - 1, k - 1)], kN0 = a.kc[idxC(d, i - 1, jn - 1, k - 1)];
    const double kB0 = a.kc[idxC(d, i - 1, j - 1, kb - 1)], kT0 = a.kc[idxC(d, i - 1, j - 1, kt - 1)];
    const double uW = a.u[idxU(d, i, j, k)], uE = a.u[idxU(d, i + 1, j, k)];
    const double vS = a.v[idxV(d, i, j, k)], vN = a.v[idxV(d, i, j + 1, k)];
    const double wB = a.w[idxW(d, i, j, k)], wT = a.w[idxW(d, i, j, k + 1)];
    const double q = a.q[cc];
    r.t0 = T[c]; r.tE = T[c + 1]; r.tW = T[c - 1]; r.tN = T[c + d.sPy]; r.tS = T[c - d.sPy];
    r.tB = T[c - d.sPz]; r.tT = T[c + d.sPz];
    issued_together(r.cls, xp, yp, zp, gW, gE, gS, gN, gB, gT, kP, kW0, kE0, kS0, kN0, kB0, kT0, uW, uE, vS, vN, wB, wT,
                    q, r.t0, r.tE, r.tW, r.tN, r.tS, r.tB, r.tT);
    const double Axy = xp * yp, Axz = xp * zp, Ayz = yp * zp;
    const double kW = i - 1 < 1 ? kP : kW0, kE = i + 1 > d.nx ? kP : kE0;
    const double kS = j - 1 < 1 ? kP : kS0, kN = j + 1 > d.ny ? kP : kN0;
    const double kB = k - 1 < 1 ? kP : kB0, kT = k + 1 > d.nz ? kP : kT0;
    const double dW = (i - 1 < 1) ? xp : gW, dE = (i + 1 > d.nx) ? xp : gE;
    const double dS = (j - 1 < 1) ? yp : gS, dN = (j + 1 > d.ny) ? yp : gN;
    const double dB = (k - 1 < 1) ? zp : gB, dT = (k + 1 > d.nz) ? zp : gT;
    const double DW = Ayz / (0.5 * xp / kP + 0.5 * dW / kW);
    const double DE = Ayz / (0.5 * xp / kP + 0.5 * dE / kE);
    const double DS = Axz / (0.5 * yp / kP + 0.5 * dS / kS);
    const double DN = Axz / (0.5 * yp / kP + 0.5 * dN / kN);
    const double DB = Axy / (0.5 * zp / kP + 0.5 * dB / kB);
    const double DT = Axy / (0.5 * zp / kP + 0.5 * dT / kT);
    const double rho = a.rho;
    const double FW = rho * Ayz * uW, FE = rho * Ayz * uE;
    const double FS = rho * Axz * vS, FN = rho * Axz * vN;
    const double FB = rho * Axy * wB, FT = rho * Axy * wT;
    r.cf[0] = DE + posd(-FE);
    r.cf[1] = DW + posd(FW);
    r.cf[2] = DN + posd(-FN);
    r.cf[3] = DS + posd(FS);
    r.cf[4] = DB + posd(FB);
    r.cf[5] = DT + posd(-FT);
    r.bs = q * (xp * yp * zp);
    r.ap = r.cf[0] + r.cf[1] + r.cf[2] + r.cf[3] + r.cf[4] + r.cf[5];
    return r;
}

__device__ __forceinline__ bool t_update(const TArgs& a, long i, long j, long k, double* dd) {
    const Dims& d = a.d;
    const long c = idxP(d, i, j, k);
    double* T = a.t;
    if (!a.solid && !a.modeR) {
        const TRowC r = t_row_cfg(a, i, j, k);
        if (r.cls == CELL_GHOST) return false;
        const double nv = (r.cf[0] * r.tE + r.cf[1] * r.tW + r.cf[2] * r.tN + r.cf[3] * r.tS + r.cf[4] * r.tB +
                           r.cf[5] * r.tT + r.bs) / r.ap;
        const double val = nv * a.relax + (1 - a.relax) * r.t0;
        T[c] = val;
        *dd = (val - r.t0) * (val - r.t0);
        return true;
    }
    const uint8_t cl = a.cls[c];
    double nv;
    if (a.solid) {
        if (cl != CELL_SOLID) return false;
        // MobSolver::solveTemperature mobs_solver.h:484-490 (Q12)
        nv = 1.0 / 6.0 * (T[c + 1] + T[c - 1] + T[c - d.sPy] + T[c + d.sPy] + T[c - d.sPz] + T[c + d.sPz]);
    } else {   // Mode R fluid
        if (cl != CELL_FLUID) return false;
        nv = t_point_ref(a, i, j, k);
    }
    const double old = T[c];
    const double val = nv * a.relax + (1 - a.relax) * old;
    T[c] = val;
    *dd = (val - old) * (val - old);
    return true;
}

// Red-black colour pass of the energy sweep over a box, two adjacent cells per thread
// along x (every lane busy on every plane), tiles in contiguous runs per XCD.
__global__ __launch_bounds__(CBT) void k_t_rb(TArgs a, int i0, int j0, int k0, int ni, int nj, int nk, int kc, int colour,
                                              int gx, int gy, int gz, double* partials, SweepCtl* wctl) {
    if (!wctl->active) return;
    const int tid = threadIdx.x + CBX * threadIdx.y;
    const hsd::Tile tl = hsd::tile_of(gx, gy, gz);
    const long ib = i0 + (long)tl.bx * (2 * CBX) + 2 * threadIdx.x, j = j0 + (long)tl.by * CBY + threadIdx.y;
    double acc = 0.0;
    if (ib < i0 + ni && j < j0 + nj) {
        const int ka = k0 + tl.bz * kc, kb = min(k0 + nk - 1, ka + kc - 1);
        for (int k = ka; k <= kb; ++k) {
            const long i = ib + ((((ib + j + k) & 1) != colour) ? 1 : 0);
            if (i >= i0 + ni) continue;
            double dd = 0.0;
            if (t_update(a, i, j, k, &dd)) acc += dd;
        }
    }
    put_partial<CBT>(acc, tid, blockIdx.x, partials + (colour ? gx * gy * gz : 0));
}

__global__ __launch_bounds__(256) void k_t_plane(TArgs a, int i0, int j0, int k0, int ni, int nj, int nk, int s,
                                                 SweepCtl* wctl) {
    if (!wctl->active) return;
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int jj = (int)(t % nj), kk = (int)(t / nj);
    if (kk >= nk) return;
    const int ii = s - jj - kk;
    if (ii < 0 || ii >= ni) return;
    const long i = i0 + ii, j = j0 + jj, k = k0 + kk;
    double dd = 0.0;
    if (t_update(a, i, j, k, &dd)) a.dres[idxP(a.d, i, j, k)] = dd;
}

// solver.h:62-148: the fluid's solid-side cells take the solid temperature and the
// solid's fluid-side ghosts take the fluid temperature (both from pre-exchange values;
// the written sets are disjoint from the read sets).
__global__ void k_exchange(const long* fi, const long* si, long n, double* tf, double* ts) {
    const long a = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (a >= n) return;
    const long F = fi[a], S = si[a];
    const double from_solid = ts[S], from_fluid = tf[F];
    tf[S] = from_solid;
    ts[F] = from_fluid;
}

// ------------------------------------------------------ p' operator ---
// Mode R row coefficients (initColsAndSetValues fluid_solver.h:1966-1996).
__global__ void k_assemble_ref(Dims d, Geo g, const uint8_t* cls, const double* up, const double* vp, const double* wp,
                               double* ax, double* ay, double* az) {
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y, k = 1 + blockIdx.z;
    if (i > d.nx) return;
    const long c = idxP(d, i, j, k);
    if (cls[c] != CELL_FLUID) { ax[c] = 0; ay[c] = 0; az[c] = 0; return; }
    const double xp = g.xp[i], yp = g.yp[j], zp = g.zp[k];
    const double Axy = xp * yp, Axz = xp * zp, Ayz = yp * zp;
    ax[c] = Ayz / up[idxU(d, i, j, k)];
    ay[c] = Axz / vp[idxV(d, i, j, k)];
    az[c] = Axy / wp[idxW(d, i, j, k)];
}
// Mode C face coefficients rho A^2 / ap on every face touching a fluid cell.
__global__ void k_assemble_cfg(Dims d, Geo g, const uint8_t* cls, const double* up, const double* vp, const double* wp,
                               double* ax, double* ay, double* az, double rho) {
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y, k = 1 + blockIdx.z;
    auto fl = [&](long ii, long jj, long kk) { return cls[idxP(d, ii, jj, kk)] == CELL_FLUID; };
    if (i <= d.nx + 1 && j <= d.ny && k <= d.nz) {
        const long f = idxU(d, i, j, k);
        double v = 0.0;
        if (fl(i - 1, j, k) || fl(i, j, k)) { const double A = g.yp[j] * g.zp[k]; v = rho * A * A / up[f]; }
        ax[f] = v;
    }
    if (i <= d.nx && j <= d.ny + 1 && k <= d.nz) {
        const long f = idxV(d, i, j, k);
        double v = 0.0;
        if (fl(i, j - 1, k) || fl(i, j, k)) { const double A = g.xp[i] * g.zp[k]; v = rho * A * A / vp[f]; }
        ay[f] = v;
    }
    if (i <= d.nx && j <= d.ny && k <= d.nz + 1) {
        const long f = idxW(d, i, j, k);
        double v = 0.0;
        if (fl(i, j, k - 1) || fl(i, j, k)) { const double A = g.xp[i] * g.yp[j]; v = rho * A * A / wp[f]; }
        az[f] = v;
    }
}
// right-hand side (fluid_solver.h:1998-2002 Mode R; area-weighted in Mode C)
__global__ void k_rhs(Dims d, Geo g, const uint8_t* cls, const double* u, const double* v, const double* w,
                      double* b, int modeR, double rho) {
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y, k = 1 + blockIdx.z;
    if (i > d.nx) return;
    const long c = idxP(d, i, j, k);
    if (cls[c] != CELL_FLUID) { b[c] = 0.0; return; }
    const double du = u[idxU(d, i + 1, j, k)] - u[idxU(d, i, j, k)];
    const double dv = v[idxV(d, i, j + 1, k)] - v[idxV(d, i, j, k)];
    const double dw = w[idxW(d, i, j, k + 1)] - w[idxW(d, i, j, k)];
    if (modeR) b[c] = -du - dv - dw;
    else {
        const double xp = g.xp[i], yp = g.yp[j], zp = g.zp[k];
        b[c] = -rho * (yp * zp * du + xp * zp * dv + xp * yp * dw);
    }
}

struct OpArgs {
    Dims d;
    const double *ax, *ay, *az;
    const uint8_t* cls;
    int modeR;
};

// y = A x at one fluid cell: diagonal first, then B, S, W, E, N, T (CSR column order
// of fluid_solver.h:2003-2083).  Entries of x outside fluid cells are zero.  xb/xc/xt
// are x at k-1, k, k+1 and cb/ct the bottom/top z-face coefficients, carried in
// registers by the z-marching callers.
template <typename XA>
__device__ __forceinline__ double apply_row(const OpArgs& o, XA x, long c, long cu, long cv, double xb, double xc,
                                            double xt, double cB, double cT) {
    const Dims& d = o.d;
    double cS, cW, cE, cN, dg;
    if (o.modeR) {
        cW = cE = o.ax[c]; cS = cN = o.ay[c];
        dg = cE + cW + cN + cS + cT + cB;
    } else {
        cW = o.ax[cu]; cE = o.ax[cu + 1];
        cS = o.ay[cv]; cN = o.ay[cv + d.sVy];
        dg = cW + cE + cS + cN + cB + cT;
    }
    double t = dg * xc;
    t += (-cB) * xb;
    t += (-cS) * x(c - d.sPy);
    t += (-cW) * x(c - 1);
    t += (-cE) * x(c + 1);
    t += (-cN) * x(c + d.sPy);
    t += (-cT) * xt;
    return t;
}

// Walks one (i,j) column over planes k0..k1 calling f(c, y, x_c) for fluid cells and
// g(c, x_c) for every cell; x(c) gives the operand at padded index c.
template <typename XA, typename F, typename G>
__device__ __forceinline__ void spmv_column(const OpArgs& o, XA x, long i, long j, int k0, int k1, F f, G g) {
    const Dims& d = o.d;
    long c = idxP(d, i, j, k0), cu = idxU(d, i, j, k0), cv = idxV(d, i, j, k0);
    double xb = x(c - d.sPz), xc = x(c);
    // z-face coefficients: Mode R row coefficient az at the cell (both faces), Mode C the
    // W-layout faces k and k+1 (W shares the P strides)
    double cb = o.az[c];
    for (int k = k0; k <= k1; ++k, c += d.sPz, cu += d.sUz, cv += d.sVz) {
        const double xt = x(c + d.sPz);
        const double ct = o.modeR ? o.az[c] : o.az[c + d.sPz];
        g(c, xc);
        if (o.cls[c] == CELL_FLUID) f(c, apply_row(o, x, c, cu, cv, xb, xc, xt, o.modeR ? ct : cb, ct), xc);
        xb = xc;
        xc = xt;
        cb = o.modeR ? o.az[c + d.sPz] : ct;
    }
}

// Sum of n block partials with the same bits in every block that calls it with the same
// NT: thread t < NT adds entries t, t + NT, ... (eight loads issued together per round),
// each wave adds its lanes, and every thread adds the NT / 64 wave sums in order.  Every
// thread of the block calls it (it holds one barrier); threads t >= NT add nothing.
// *flag: thread 0's value is handed to every thread through the same barrier.
template <int NT>
__device__ __forceinline__ double sum_partials_all(const double* part, int n, int tid, int* flag) {
    __shared__ double sh[NT / 64];
    __shared__ int sh_flag;
    if (tid < NT) {
        double s = 0.0;
        for (int base = 0; base < n; base += 8 * NT) {
            double v[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int t = base + tid + NT * k;
                v[k] = t < n ? part[t] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) s += v[k];
        }
        s = hsd::wave_sum(s);
        if ((tid & 63) == 0) sh[tid >> 6] = s;
        if (tid == 0) sh_flag = *flag;
    }
    __syncthreads();
    *flag = sh_flag;
    double r = 0.0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) r += sh[w];
    return r;
}

// Scalar steps of HYPRE's PCG with the identity preconditioner (fluid_solver.h:1789-1790,
// settings :1871-1875).  gamma_k = <r_k, r_k> is kept in ctl->cg_g[k & 1], so a kernel
// whose blocks all form gamma_{k+1} from the same partials writes the other slot while
// they still read gamma_k.  Each step is a pure function of the control block and the
// reduced value; block 0 of the kernel that takes it (or the one-block finish) records it.
struct CgStep {
    double v;     // alpha or beta
    int stop;     // the solve ends here (breakdown, convergence, maxit, non-finite)
    int nan;
    int conv;
    int iters;
};

// the control-block values a CG step reads, loaded at a kernel's start together with the
// partials (none of them is written while the kernel runs, apart from done)
struct CgIn {
    double bb, eps, g;    // <b,b>, tol^2, gamma_it
    int maxit, done;
};
__device__ __forceinline__ CgIn cg_in(const SolveCtl* ctl, int it) {
    return CgIn{ctl->bb, ctl->eps, ctl->cg_g[it & 1], ctl->maxit, ctl->done};
}

// alpha of CG iteration it from <p, A p>: a zero or non-finite <p, A p>, or alpha not
// above DBL_MIN, ends the solve (HYPRE's breakdown exits)
__device__ __forceinline__ CgStep cg_alpha_step(const CgIn& ci, double pap) {
    CgStep s{0.0, 0, 0, 0, 0};
    if (!isfinite(pap)) { s.stop = 1; s.nan = 1; return s; }
    if (pap == 0.0) { s.stop = 1; return s; }
    s.v = ci.g / pap;
    if (!(s.v > DBL_MIN)) s.stop = 1;
    return s;
}
__device__ __forceinline__ void cg_alpha_record(SolveCtl* ctl, double pap, const CgStep& s) {
    if (s.nan) { ctl->nan = 1; ctl->done = 1; ctl->converged = 0; return; }
    ctl->sdotp = pap;
    if (s.stop) { ctl->done = 1; return; }
    ctl->alpha = s.v;
}

// after CG update it (<r,r> = rr): iteration count it + 1, the relative stop rule
// <r,r> / <b,b> < tol^2, maxit, and beta = gamma_{it+1} / gamma_it
__device__ __forceinline__ CgStep cg_update_step(const CgIn& ci, double rr, int it) {
    CgStep s{0.0, 0, 0, 0, it + 1};
    if (!isfinite(rr)) { s.stop = 1; s.nan = 1; return s; }
    if (rr / ci.bb < ci.eps) { s.stop = 1; s.conv = 1; return s; }
    if (s.iters >= ci.maxit) { s.stop = 1; return s; }
    s.v = rr / ci.g;
    return s;
}
__device__ __forceinline__ void cg_update_record(SolveCtl* ctl, const CgIn& ci, double rr, int it, const CgStep& s) {
    if (s.nan) { ctl->nan = 1; ctl->done = 1; ctl->converged = 0; return; }
    ctl->iters = s.iters;
    ctl->gamma_old = ci.g;
    ctl->cg_g[(it + 1) & 1] = rr;
    ctl->gamma = rr;
    ctl->rr = rr;
    if (s.conv) ctl->converged = 1;
    if (s.stop) { ctl->done = 1; return; }
    ctl->beta = s.v;
}

// s = A p (z-marching columns) with block partials of <s,p> (k_finish: alpha); p is float
// in the mixed-precision flexible CG, the product and the sum FP64.
template <typename P>
__global__ __launch_bounds__(BT) void k_cg_spmv(OpArgs o, const P* p, double* s, int kc, int gx, int gy, int gz,
                                                double* partials, SolveCtl* ctl) {
    if (ctl->done) return;
    const Dims& d = o.d;
    const int tid = threadIdx.x + BX * threadIdx.y;
    const hsd::Tile tl = hsd::tile_of(gx, gy, gz);
    const long i = 1 + tl.bx * BX + threadIdx.x, j = 1 + tl.by * BY + threadIdx.y;
    double acc = 0.0;
    if (i <= d.nx && j <= d.ny) {
        const int k0 = 1 + tl.bz * kc, k1 = min(d.nz, k0 + kc - 1);
        spmv_column(
            o, [p](long q) { return (double)p[q]; }, i, j, k0, k1,
            [&](long c, double y, double pc) {
                s[c] = y;
                acc += y * pc;
            },
            [](long, double) {});
    }
    put_partial<BT>(acc, tid, blockIdx.x, partials);
}

// Flexible CG direction update fused into the operator apply: p_new = z + beta p_old is
// formed at every point the stencil reads (the same expression as a separate update pass,
// so the same values), written to pn for the column's cells, and q = A p_new with block
// partials of <p_new, q>.  pn must not alias p (neighbouring columns still read p_old).
// The directions are stored as T: the mixed-precision solve keeps them in float, formed
// in FP64 and rounded once, and the apply and <p_new, q> use the rounded value, so q is
// the FP64 product of the stored direction.
// With fold (plain CG, z = r), beta is not read from ctl: every block takes the stop rule
// and beta of CG update fold_it itself from that update's nfold block partials of <r,r>
// (cg_update_step), block 0 records them, and a finished solve returns here.
template <typename T>
__global__ __launch_bounds__(BT) void k_fcg_spmv(OpArgs o, const T* __restrict__ z, const T* __restrict__ p,
                                                 T* __restrict__ pn, double* __restrict__ s, int kc, int gx, int gy,
                                                 int gz, double* partials, SolveCtl* ctl, const double* fold, int nfold,
                                                 int fold_it) {
    const int tid = threadIdx.x + BX * threadIdx.y;
    double beta;
    if (fold) {
        // thread 0's done flag is handed on through the sum's barrier, so the block takes
        // one decision
        const CgIn ci = cg_in(ctl, fold_it);
        int done = ci.done;
        const double rr = sum_partials_all<BT>(fold, nfold, tid, &done);
        if (done) return;
        const CgStep st = cg_update_step(ci, rr, fold_it);
        if (blockIdx.x == 0 && tid == 0) cg_update_record(ctl, ci, rr, fold_it, st);
        if (st.stop) return;
        beta = st.v;
    } else {
        if (ctl->done) return;
        beta = ctl->beta;
    }
    const Dims& d = o.d;
    const hsd::Tile tl = hsd::tile_of(gx, gy, gz);
    const long i = 1 + tl.bx * BX + threadIdx.x, j = 1 + tl.by * BY + threadIdx.y;
    double acc = 0.0;
    if (i <= d.nx && j <= d.ny) {
        const int k0 = 1 + tl.bz * kc, k1 = min(d.nz, k0 + kc - 1);
        spmv_column(
            o, [=](long q) { return (double)(T)((double)z[q] + beta * (double)p[q]); }, i, j, k0, k1,
            [&](long c, double y, double pc) {
                s[c] = y;
                acc += y * pc;
            },
            [=](long c, double pc) { pn[c] = (T)pc; });
    }
    put_partial<BT>(acc, tid, blockIdx.x, partials);
}

// x += alpha p, r -= alpha s, block partials of <r,r> (the next apply or k_finish: stop
// rule, beta).  alpha of CG iteration it is formed by every block alike from the apply's
// nsp block partials of <p, A p> (cg_alpha_step); block 0 records it.
__global__ __launch_bounds__(FLAT) void k_cg_update(long n, double* __restrict__ x, double* __restrict__ r,
                                                    const double* __restrict__ p, const double* __restrict__ s,
                                                    const double* sp, int nsp, int it, double* partials,
                                                    SolveCtl* ctl) {
    const CgIn ci = cg_in(ctl, it);
    int done = ci.done;
    const double pap = sum_partials_all<FLAT>(sp, nsp, threadIdx.x, &done);
    if (done) return;
    const CgStep st = cg_alpha_step(ci, pap);
    if (blockIdx.x == 0 && threadIdx.x == 0) cg_alpha_record(ctl, pap, st);
    if (st.stop) return;
    const double alpha = st.v;
    double acc = 0.0;
#pragma unroll 4
    for (long t = blockIdx.x * (long)FLAT + threadIdx.x; t < n; t += (long)gridDim.x * FLAT) {
        x[t] += alpha * p[t];
        const double rv = r[t] - alpha * s[t];
        r[t] = rv;
        acc += rv * rv;
    }
    put_partial<FLAT>(acc, threadIdx.x, blockIdx.x, partials);
}

// block partials of <a,b>
__global__ __launch_bounds__(FLAT) void k_dot(long n, const double* a, const double* b, double* partials) {
    double acc = 0.0;
    for (long t = blockIdx.x * (long)FLAT + threadIdx.x; t < n; t += (long)gridDim.x * FLAT) acc += a[t] * b[t];
    put_partial<FLAT>(acc, threadIdx.x, blockIdx.x, partials);
}

// CG setup: r = p = b, x = 0 (x0 = 0, fluid_solver.h:1766); partials of <b,b>
__global__ __launch_bounds__(FLAT) void k_cg_setup(long n, const double* b, double* x, double* r, double* p,
                                                   double* partials) {
    double acc = 0.0;
    for (long t = blockIdx.x * (long)FLAT + threadIdx.x; t < n; t += (long)gridDim.x * FLAT) {
        const double bv = b[t];
        x[t] = 0.0; r[t] = bv; p[t] = bv;
        acc += bv * bv;
    }
    put_partial<FLAT>(acc, threadIdx.x, blockIdx.x, partials);
}

// One-block finish of a reduction: adds the block partials in a fixed order and runs
// the scalar step of the kernel that produced them.

__global__ __launch_bounds__(1024) void k_finish(const double* partials, int n, int kind, SolveCtl* ctl, SweepCtl* wctl,
                                                 double* out, int it) {
    if (kind == FIN_SWEEP) {
        if (!wctl->active) return;
    } else if (kind == FIN_CG_ALPHA || kind == FIN_CG_UPDATE || kind == FIN_FCG_START || kind == FIN_FCG_UPDATE ||
               kind == FIN_FCG_BETA) {
        if (ctl->done) return;
    }
    const int tid = threadIdx.x;
    // CG's <r,r> is summed in the order the vector kernels use for it (sum_partials_all
    // over 256 threads), so a batch boundary does not change a bit of the iterates
    int unused = 0;
    const double a = kind == FIN_CG_UPDATE ? sum_partials_all<BT>(partials, n, tid, &unused)
                                           : hsd::sum_partials<1024>(partials, n, tid);
    double b = 0.0;
    if (kind == FIN_FCG_BETA || kind == FIN_DOT2 || kind == FIN_FCG_SETUP_WARM)
        b = hsd::sum_partials<1024>(partials + n, n, tid);
    if (tid != 0) return;
    // a Krylov reduction that is not finite ends the solve at once (the host reports
    // HS_E_NAN) instead of iterating on to maxit
    if (kind != FIN_SWEEP && kind != FIN_DOT && kind != FIN_DOT2 && kind != FIN_DIFF) {
        const bool setup = kind == FIN_CG_SETUP || kind == FIN_FCG_SETUP || kind == FIN_FCG_SETUP_WARM;
        if (setup) ctl->nan = 0;
        if (!isfinite(a) || !isfinite(b)) {
            ctl->nan = 1;
            ctl->done = 1;
            ctl->converged = 0;
            if (setup) { ctl->bb = a; ctl->rr = a; ctl->iters = 0; }
            return;
        }
    }
    switch (kind) {
        case FIN_SWEEP: sweep_control(wctl, a); break;
        case FIN_CG_SETUP:
            ctl->bb = a; ctl->gamma = a; ctl->cg_g[0] = a; ctl->rr = a; ctl->iters = 0;
            ctl->converged = (a == 0.0) ? 1 : 0; ctl->done = (a == 0.0) ? 1 : 0;
            break;
        case FIN_CG_ALPHA: {
            ctl->sdotp = a;
            if (a == 0.0) { ctl->done = 1; break; }
            const double alpha = ctl->gamma / a;
            if (!(alpha > DBL_MIN)) { ctl->done = 1; break; }
            ctl->alpha = alpha;
            break;
        }
        case FIN_CG_UPDATE: {
            const CgIn ci = cg_in(ctl, it);
            cg_update_record(ctl, ci, a, it, cg_update_step(ci, a, it));
            break;
        }
        case FIN_DOT: *out = a; break;
        case FIN_DOT2: out[0] = a; out[1] = b; break;
        case FIN_DIFF: out[0] = a - out[1]; break;
        case FIN_FCG_SETUP:
            ctl->bb = a; ctl->rr = a; ctl->iters = 0;
            ctl->converged = (a == 0.0) ? 1 : 0; ctl->done = (a == 0.0) ? 1 : 0;
            break;
        case FIN_FCG_SETUP_WARM: {
            ctl->bb = a; ctl->iters = 0;
            ctl->cold = (b < a) ? 0 : 1;     // no better than the zero start: take that
            ctl->rr = ctl->cold ? a : b;
            const bool conv = a == 0.0 || ctl->rr / a < ctl->eps;
            ctl->converged = conv ? 1 : 0; ctl->done = conv ? 1 : 0;
            break;
        }
        case FIN_FCG_START: ctl->gamma = a; break;
        case FIN_FCG_UPDATE:
            ctl->iters += 1; ctl->rr = a;
            if (a / ctl->bb < ctl->eps) { ctl->converged = 1; ctl->done = 1; break; }
            if (ctl->iters >= ctl->maxit) { ctl->done = 1; break; }
            break;
        case FIN_FCG_BETA:
            ctl->beta = -ctl->alpha * b / ctl->gamma;
            ctl->gamma_old = ctl->gamma;
            ctl->gamma = a;
            break;
    }
}

// plain y = A x (fluid cells; zero elsewhere) without reduction
__global__ __launch_bounds__(BT) void k_spmv(OpArgs o, const double* x, double* y, int kc, int gx, int gy, int gz) {
    const Dims& d = o.d;
    const hsd::Tile tl = hsd::tile_of(gx, gy, gz);
    const long i = 1 + tl.bx * BX + threadIdx.x, j = 1 + tl.by * BY + threadIdx.y;
    if (i > d.nx || j > d.ny) return;
    const int k0 = 1 + tl.bz * kc, k1 = min(d.nz, k0 + kc - 1);
    // every cell of the column is written once: the row's value, or zero off the fluid
    spmv_column(
        o, [x](long q) { return x[q]; }, i, j, k0, k1, [&](long c, double v, double) { y[c] = v; },
        [&](long c, double) {
            if (o.cls[c] != CELL_FLUID) y[c] = 0.0;
        });
}

// ||b - A x||^2 and ||b||^2 in one pass over the fluid rows (block partials, two per
// block).  b is zero outside fluid cells (the assembled right-hand side, or a caller's
// vector after k_mask_fluid), and so are the rows of A there, so the other cells add nothing.
__global__ __launch_bounds__(BT) void k_resid_norms(OpArgs o, const double* x, const double* b, int kc, int gx, int gy,
                                                    int gz, double* partials) {
    const Dims& d = o.d;
    const int tid = threadIdx.x + BX * threadIdx.y;
    const hsd::Tile tl = hsd::tile_of(gx, gy, gz);
    const long i = 1 + tl.bx * BX + threadIdx.x, j = 1 + tl.by * BY + threadIdx.y;
    double rr = 0.0, bb = 0.0;
    if (i <= d.nx && j <= d.ny) {
        const int k0 = 1 + tl.bz * kc, k1 = min(d.nz, k0 + kc - 1);
        spmv_column(
            o, [x](long q) { return x[q]; }, i, j, k0, k1,
            [&](long c, double y, double) {
                const double bc = b[c], r = bc - y;
                rr += r * r;
                bb += bc * bc;
            },
            [](long, double) {});
    }
    hsd::put_partial2<BT>(rr, bb, tid, blockIdx.x, gx * gy * gz, partials);
}

// Residual of a nonzero initial iterate x (flexible CG's warm start) in one pass: r = b - A x
// on every interior cell (b itself off the fluid cells, where b and the rows of A are zero),
// the level-0 right-hand side (T) r, the level-0 iterate zeroed when zero_x0, and block
// partials of <b,b> then <r,r> (two per block).  Padding entries of r, b0 and x0 are not
// written: they hold zeros from allocation and no kernel writes them.
template <typename T>
__global__ __launch_bounds__(BT) void k_warm_resid(OpArgs o, const double* x, const double* b, double* r, T* b0,
                                                   T* x0, int zero_x0, int kc, int gx, int gy, int gz,
                                                   double* partials) {
    const Dims& d = o.d;
    const int tid = threadIdx.x + BX * threadIdx.y;
    const hsd::Tile tl = hsd::tile_of(gx, gy, gz);
    const long i = 1 + tl.bx * BX + threadIdx.x, j = 1 + tl.by * BY + threadIdx.y;
    double rr = 0.0, bb = 0.0;
    if (i <= d.nx && j <= d.ny) {
        const int k0 = 1 + tl.bz * kc, k1 = min(d.nz, k0 + kc - 1);
        spmv_column(
            o, [x](long q) { return x[q]; }, i, j, k0, k1,
            [&](long c, double y, double) {
                const double bc = b[c], rv = bc - y;
                r[c] = rv;
                b0[c] = (T)rv;
                rr += rv * rv;
                bb += bc * bc;
            },
            [&](long c, double) {
                if (zero_x0) x0[c] = T(0);
                if (o.cls[c] != CELL_FLUID) {
                    const double bc = b[c];
                    r[c] = bc;
                    b0[c] = (T)bc;
                }
            });
    }
    hsd::put_partial2<BT>(bb, rr, tid, blockIdx.x, gx * gy * gz, partials);
}

// zero entries outside fluid cells
__global__ void k_mask_fluid(long n, const uint8_t* cls, double* x) {
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        if (cls[t] != CELL_FLUID) x[t] = 0.0;
}
// z = a x + b y (residual check)
__global__ void k_axpby(long n, double a, const double* x, double b, const double* y, double* z) {
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) z[t] = a * x[t] + b * y[t];
}

// ------------------------------------------------------- corrections ---
// velocity correction fluid_solver.h:1459-1571 (Mode R: faces whose plus-side cell is
// fluid; Mode C: faces between two fluid cells), and the pressure update
// fluid_solver.h:1574-1598 of the thread's own cell (p += relax p', p_temp = p on fluid
// cells; p' is only read, so the order of the two does not matter)
__global__ void k_correct(Dims d, Geo g, const uint8_t* cls, const double* pp, double* u, double* v, double* w,
                          const double* up, const double* vp, const double* wp, int modeR, double* p, double* pt,
                          double relax) {
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y, k = 1 + blockIdx.z;
    auto fl = [&](long ii, long jj, long kk) { return cls[idxP(d, ii, jj, kk)] == CELL_FLUID; };
    const int ex = modeR ? 0 : 1;
    if (i <= d.nx + ex && j <= d.ny && k <= d.nz) {
        const bool act = modeR ? fl(i, j, k) : (fl(i - 1, j, k) && fl(i, j, k));
        if (act) {
            const double faceyz = g.yp[j] * g.zp[k];
            double ap = up[idxU(d, i, j, k)];
            if (ap == 0) ap = 1e-10;
            const double corr = (pp[idxP(d, i - 1, j, k)] - pp[idxP(d, i, j, k)]) * faceyz / ap;
            u[idxU(d, i, j, k)] = u[idxU(d, i, j, k)] + corr;
        }
    }
    if (i <= d.nx && j <= d.ny + ex && k <= d.nz) {
        const bool act = modeR ? fl(i, j, k) : (fl(i, j - 1, k) && fl(i, j, k));
        if (act) {
            const double facexz = g.xp[i] * g.zp[k];
            double ap = vp[idxV(d, i, j, k)];
            if (ap == 0) ap = 1e-10;
            const double corr = (pp[idxP(d, i, j - 1, k)] - pp[idxP(d, i, j, k)]) * facexz / ap;
            v[idxV(d, i, j, k)] = v[idxV(d, i, j, k)] + corr;
        }
    }
    if (i <= d.nx && j <= d.ny && k <= d.nz + ex) {
        const bool act = modeR ? fl(i, j, k) : (fl(i, j, k - 1) && fl(i, j, k));
        if (act) {
            const double facexy = g.xp[i] * g.yp[j];
            double ap = wp[idxW(d, i, j, k)];
            if (ap == 0) ap = 1e-10;
            const double corr = (pp[idxP(d, i, j, k - 1)] - pp[idxP(d, i, j, k)]) * facexy / ap;
            w[idxW(d, i, j, k)] = w[idxW(d, i, j, k)] + corr;
        }
    }
    if (i <= d.nx && j <= d.ny && k <= d.nz && fl(i, j, k)) {
        const long c = idxP(d, i, j, k), cc = idxC(d, i - 1, j - 1, k - 1);
        const double nv = p[cc] + relax * pp[c];
        p[cc] = nv;
        pt[c] = nv;
    }
}

// Mode C energy balance: the sum over all cell rows of a_P T_P - sum_nb a_nb T_nb, the rows
// of t_row_cfg (the caller subtracts the heat input sum q V).  Both cells of an interior
// face get the same diffusive part and a_P = sum_nb a_nb, so an interior face with flux F
// (from its minus to its plus cell) adds F (T_plus - T_minus), and the sum telescopes to
//   sum_c [ -T_c m_c + sum over c's boundary faces of (a_face (T_c - T_ghost) - F_in T_c) ]
// with m_c the net outward flux of cell c, F_in a boundary face's flux into c and a_face
// the row coefficient toward its ghost (t_row_cfg's boundary forms).  k_energy_balance
// adds the cells' -T_c m_c (T, class and six face velocities per cell),
// k_energy_balance_faces the boundary faces' terms.
__global__ __launch_bounds__(BT) void k_energy_balance(TArgs a, int kc, double* partials) {
    const Dims& d = a.d;
    const int tid = threadIdx.x + BX * threadIdx.y;
    const long i = 1 + blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
    double acc = 0.0;
    if (i <= d.nx && j <= d.ny) {
        const int k0 = 1 + blockIdx.z * kc, k1 = min(d.nz, k0 + kc - 1);
        const double xp = a.g.xp[i], yp = a.g.yp[j];
        const double Axy = xp * yp;
        const double rho = a.rho;
        double wB = a.w[idxW(d, i, j, k0)];
        for (int k = k0; k <= k1; ++k) {
            const long c = idxP(d, i, j, k);
            const double zp = a.g.zp[k];
            const double uW = a.u[idxU(d, i, j, k)], uE = a.u[idxU(d, i + 1, j, k)];
            const double vS = a.v[idxV(d, i, j, k)], vN = a.v[idxV(d, i, j + 1, k)];
            const double wT = a.w[idxW(d, i, j, k + 1)];
            const double t0 = a.t[c];
            const uint8_t cls = a.cls[c];
            const double Axz = xp * zp, Ayz = yp * zp;
            const double FW = rho * Ayz * uW, FE = rho * Ayz * uE;
            const double FS = rho * Axz * vS, FN = rho * Axz * vN;
            const double FB = rho * Axy * wB, FT = rho * Axy * wT;
            wB = wT;
            if (cls == CELL_GHOST) continue;
            acc -= t0 * ((FE - FW) + (FN - FS) + (FT - FB));
        }
    }
    const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    put_partial<BT>(acc, tid, bid, partials);
}

// The boundary faces' terms a_face (T_c - T_ghost) - F_in T_c, one face per thread: the
// x = min / max faces (ny nz each), then y (nx nz each), then z (nx ny each).  a_face takes
// the cell's own width and conductivity on both halves (t_row_cfg's out-of-domain
// neighbour).
__global__ __launch_bounds__(FLAT) void k_energy_balance_faces(TArgs a, double* partials) {
    const Dims& d = a.d;
    const long nyz = (long)d.ny * d.nz, nxz = (long)d.nx * d.nz, nxy = (long)d.nx * d.ny;
    const long total = 2 * (nyz + nxz + nxy);
    const double rho = a.rho;
    const double* T = a.t;
    double acc = 0.0;
    for (long t = blockIdx.x * (long)FLAT + threadIdx.x; t < total; t += (long)gridDim.x * FLAT) {
        long f = t;
        int axis, hi;
        long i, j, k;
        if (f < 2 * nyz) {
            axis = 0; hi = f >= nyz; f -= hi * nyz;
            j = 1 + f % d.ny; k = 1 + f / d.ny; i = hi ? d.nx : 1;
        } else if ((f -= 2 * nyz) < 2 * nxz) {
            axis = 1; hi = f >= nxz; f -= hi * nxz;
            i = 1 + f % d.nx; k = 1 + f / d.nx; j = hi ? d.ny : 1;
        } else {
            f -= 2 * nxz;
            axis = 2; hi = f >= nxy; f -= hi * nxy;
            i = 1 + f % d.nx; j = 1 + f / d.nx; k = hi ? d.nz : 1;
        }
        const long c = idxP(d, i, j, k);
        if (a.cls[c] == CELL_GHOST) continue;
        const double xp = a.g.xp[i], yp = a.g.yp[j], zp = a.g.zp[k];
        const double kP = a.kc[idxC(d, i - 1, j - 1, k - 1)];
        const double t0 = T[c];
        const double Axy = xp * yp, Axz = xp * zp, Ayz = yp * zp;
        double term;
        if (axis == 0) {
            const double D = Ayz / (0.5 * xp / kP + 0.5 * xp / kP);
            if (!hi) {
                const double FW = rho * Ayz * a.u[idxU(d, i, j, k)];
                term = (D + posd(FW)) * (t0 - T[c - 1]) - FW * t0;
            } else {
                const double FE = rho * Ayz * a.u[idxU(d, i + 1, j, k)];
                term = (D + posd(-FE)) * (t0 - T[c + 1]) + FE * t0;
            }
        } else if (axis == 1) {
            const double D = Axz / (0.5 * yp / kP + 0.5 * yp / kP);
            if (!hi) {
                const double FS = rho * Axz * a.v[idxV(d, i, j, k)];
                term = (D + posd(FS)) * (t0 - T[c - d.sPy]) - FS * t0;
            } else {
                const double FN = rho * Axz * a.v[idxV(d, i, j + 1, k)];
                term = (D + posd(-FN)) * (t0 - T[c + d.sPy]) + FN * t0;
            }
        } else {
            const double D = Axy / (0.5 * zp / kP + 0.5 * zp / kP);
            if (!hi) {
                const double FB = rho * Axy * a.w[idxW(d, i, j, k)];
                term = (D + posd(FB)) * (t0 - T[c - d.sPz]) - FB * t0;
            } else {
                const double FT = rho * Axy * a.w[idxW(d, i, j, k + 1)];
                term = (D + posd(-FT)) * (t0 - T[c + d.sPz]) + FT * t0;
            }
        }
        acc += term;
    }
    put_partial<FLAT>(acc, threadIdx.x, blockIdx.x, partials);
}

__global__ void k_sum_q(Dims d, Geo g, const double* q, const uint8_t* cls, double* partials) {
    double acc = 0.0;
    const long n = d.nC;
    for (long t = blockIdx.x * (long)FLAT + threadIdx.x; t < n; t += (long)gridDim.x * FLAT) {
        const long i = t % d.nx, j = (t / d.nx) % d.ny, k = t / ((long)d.nx * d.ny);
        acc += q[t] * (g.xp[i + 1] * g.yp[j + 1] * g.zp[k + 1]);
    }
    put_partial<FLAT>(acc, threadIdx.x, blockIdx.x, partials);
}

}  // namespace

// =================================================================== launchers ===
namespace {
using hsd::kchunk;
using hsd::grid3;
using hsd::grid1;
MomArgs mom_args(hs_ctx* c, double* ap) {
    MomArgs a;
    a.d = c->d; a.g = c->g; a.u = c->u; a.v = c->v; a.w = c->w; a.ap = ap;
    a.pt = c->ptemp; a.cls = c->cls; a.dres = c->dres;
    a.c05 = 0.5 * c->prm.rho; a.c2nu = 2 * c->prm.nu; a.relax = c->prm.relax_u;
    a.modeR = c->prm.mode == HS_MODE_REF;
    return a;
}
TArgs t_args(hs_ctx* c, bool solid) {
    TArgs a;
    a.d = c->d; a.g = c->g; a.u = c->u; a.v = c->v; a.w = c->w;
    a.t = solid ? c->ts : c->tf; a.q = c->q; a.kc = c->k; a.cls = c->cls; a.dres = c->dres;
    a.alpha = c->prm.alpha_f; a.relax = solid ? c->prm.relax_ts : c->prm.relax_t; a.rho = c->prm.rho;
    a.modeR = c->prm.mode == HS_MODE_REF; a.solid = solid ? 1 : 0;
    return a;
}
OpArgs op_args(hs_ctx* c) {
    OpArgs o;
    o.d = c->d; o.ax = c->ax; o.ay = c->ay; o.az = c->az; o.cls = c->cls;
    o.modeR = c->prm.mode == HS_MODE_REF;
    return o;
}
}  // namespace

int launch_finish(hs_ctx* c, int kind, int n, double* out) {
    k_finish<<<1, 1024, 0, c->stream>>>(c->partials, n, kind, c->sctl, c->wctl, out, 0);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}
int launch_finish_cg(hs_ctx* c, const double* part, int n, int it) {
    k_finish<<<1, 1024, 0, c->stream>>>(part, n, FIN_CG_UPDATE, c->sctl, c->wctl, nullptr, it);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}
int launch_fill(hs_ctx* c, double* a, long n, double v) {
    k_fill<<<FLAT_BLOCKS, 256, 0, c->stream>>>(a, n, v);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}
int launch_copy(hs_ctx* c, double* dst, const double* src, long n) {
    k_copy<<<FLAT_BLOCKS, 256, 0, c->stream>>>(dst, src, n);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}
int bc_x_ref(hs_ctx* c) {
    const Dims& d = c->d;
    k_bc_x_ref<<<dim3((d.ny + 2 + 63) / 64, d.nz + 2), 64, 0, c->stream>>>(d, c->u, c->v, c->w);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}
int bc_y_ref(hs_ctx* c) {
    const Dims& d = c->d;
    k_bc_y_ref<<<dim3((d.nx + 2 + 63) / 64, d.nz + 2), 64, 0, c->stream>>>(d, c->u, c->v, c->w, c->prm.u_lid);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}
int bc_z_ref(hs_ctx* c) {
    const Dims& d = c->d;
    k_bc_z_ref<<<dim3((d.nx + 2 + 63) / 64, d.ny + 2), 64, 0, c->stream>>>(d, c->u, c->v, c->w);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}
int cell_bc_ref(hs_ctx* c) {
    const Dims& d = c->d;
    HS_TRY(launch_copy(c, c->psnap, c->p, d.nC));
    const long total = 2 * ((long)d.nx * d.ny + (long)d.nx * d.nz + (long)d.ny * d.nz);
    const int nb = (int)((total + 255) / 256);
    for (int pass = 0; pass < 3; ++pass) {
        k_cellbc_pass<<<nb, 256, 0, c->stream>>>(d, c->p, c->psnap, c->ptemp, pass);
        HS_CHECK(hipGetLastError());
    }
    k_zero_one<<<1, 1, 0, c->stream>>>(c->p);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}
int ptemp_from_p_all(hs_ctx* c) {
    const Dims& d = c->d;
    k_ptemp_all<<<dim3((d.nx + 63) / 64, d.ny, d.nz), 64, 0, c->stream>>>(d, c->p, c->ptemp);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}
int bc_cfg_init(hs_ctx* c) {
    const Dims& d = c->d;
    k_bc_cfg_init<<<dim3((d.ny + 2 + 63) / 64, d.nz + 2), 64, 0, c->stream>>>(d, c->u, c->cls, c->prm.u_in, c->prm.lid,
                                                                             c->prm.u_lid);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}
int bc_cfg_outflow(hs_ctx* c, int comp) {
    if (!c->prm.outflow) return HS_OK;
    const Dims& d = c->d;
    k_bc_cfg_outflow<<<dim3((d.ny + 63) / 64, d.nz), 64, 0, c->stream>>>(d, c->u, c->v, c->w, c->tf, c->cls, comp);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}
int zero_interface_faces(hs_ctx* c) {
    const Dims& d = c->d;
    k_zero_ib<<<dim3((d.nx + 1 + 63) / 64, d.ny + 1, d.nz + 1), 64, 0, c->stream>>>(d, c->u, c->v, c->w, c->cls);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}

int sweep_begin(hs_ctx* c, double n, double tol, int maxsteps, int quirk) {
    SweepCtl* h = c->h_wctl;
    h->residual = 1.0;
    h->global = (quirk == 0) ? INFINITY : 1.0;
    h->ncount = n;
    h->tol = tol;
    h->step = 0;
    h->maxsteps = maxsteps;
    h->quirk = quirk;
    h->active = (maxsteps > 0 && !(h->global < tol)) ? 1 : 0;
    HS_CHECK(hipMemcpyAsync(c->wctl, h, sizeof(SweepCtl), hipMemcpyHostToDevice, c->stream));
    return HS_OK;
}

int momentum_sweep(hs_ctx* c, int comp) {
    const Dims& d = c->d;
    double* ap = comp == 0 ? c->up : (comp == 1 ? c->vp : c->wp);
    MomArgs a = mom_args(c, ap);
    const bool R = c->prm.mode == HS_MODE_REF;
    int ni = d.nx, nj = d.ny, nk = d.nz;
    if (!R) { if (comp == 0) ni += 1; else if (comp == 1) nj += 1; else nk += 1; }
    if (c->prm.order == HS_ORDER_RB) {
        const int gx = (ni + 2 * CBX - 1) / (2 * CBX), gy = (nj + CBY - 1) / CBY;
        const int kc = kchunk(nk, (long)gx * gy);
        const int gz = (nk + kc - 1) / kc;
        for (int colour = 0; colour < 2; ++colour) {
            const int slot = comp == 0 ? prof_begin(c, HS_K_MOM_RB) : -1;
            if (comp == 0) k_mom_rb<0><<<gx * gy * gz, dim3(CBX, CBY), 0, c->stream>>>(a, ni, nj, nk, kc, colour, gx, gy, gz, c->partials, c->wctl);
            else if (comp == 1) k_mom_rb<1><<<gx * gy * gz, dim3(CBX, CBY), 0, c->stream>>>(a, ni, nj, nk, kc, colour, gx, gy, gz, c->partials, c->wctl);
            else k_mom_rb<2><<<gx * gy * gz, dim3(CBX, CBY), 0, c->stream>>>(a, ni, nj, nk, kc, colour, gx, gy, gz, c->partials, c->wctl);
            HS_CHECK(hipGetLastError());
            HS_TRY(prof_end(c, slot));
        }
        HS_TRY(launch_finish(c, FIN_SWEEP, 2 * gx * gy * gz, nullptr));
    } else {
        const long nl = comp == 0 ? d.nU : (comp == 1 ? d.nV : d.nW);
        HS_TRY(launch_fill(c, c->dres, nl, 0.0));
        const long njk = (long)nj * nk;
        const int nb = (int)((njk + 255) / 256);
        for (int s = 3; s <= ni + nj + nk; ++s) {
            if (comp == 0) k_mom_plane<0><<<nb, 256, 0, c->stream>>>(a, ni, nj, nk, s, c->wctl, nl);
            else if (comp == 1) k_mom_plane<1><<<nb, 256, 0, c->stream>>>(a, ni, nj, nk, s, c->wctl, nl);
            else k_mom_plane<2><<<nb, 256, 0, c->stream>>>(a, ni, nj, nk, s, c->wctl, nl);
        }
        HS_CHECK(hipGetLastError());
        k_sum_control<<<FLAT_BLOCKS, FLAT, 0, c->stream>>>(c->dres, nl, c->partials, c->wctl);
        HS_CHECK(hipGetLastError());
        HS_TRY(launch_finish(c, FIN_SWEEP, FLAT_BLOCKS, nullptr));
    }
    return HS_OK;
}

static int t_sweep_box(hs_ctx* c, bool solid, int i0, int j0, int k0, int ni, int nj, int nk) {
    const Dims& d = c->d;
    TArgs a = t_args(c, solid);
    if (c->prm.order == HS_ORDER_RB) {
        const int gx = (ni + 2 * CBX - 1) / (2 * CBX), gy = (nj + CBY - 1) / CBY;
        const int kc = kchunk(nk, (long)gx * gy);
        const int gz = (nk + kc - 1) / kc;
        for (int colour = 0; colour < 2; ++colour) {
            k_t_rb<<<gx * gy * gz, dim3(CBX, CBY), 0, c->stream>>>(a, i0, j0, k0, ni, nj, nk, kc, colour, gx, gy, gz,
                                                                 c->partials, c->wctl);
            HS_CHECK(hipGetLastError());
        }
        HS_TRY(launch_finish(c, FIN_SWEEP, 2 * gx * gy * gz, nullptr));
    } else {
        HS_TRY(launch_fill(c, c->dres, d.nP, 0.0));
        const long njk = (long)nj * nk;
        const int nb = (int)((njk + 255) / 256);
        for (int s = 0; s <= (ni - 1) + (nj - 1) + (nk - 1); ++s)
            k_t_plane<<<nb, 256, 0, c->stream>>>(a, i0, j0, k0, ni, nj, nk, s, c->wctl);
        HS_CHECK(hipGetLastError());
        k_sum_control<<<FLAT_BLOCKS, FLAT, 0, c->stream>>>(c->dres, d.nP, c->partials, c->wctl);
        HS_CHECK(hipGetLastError());
        HS_TRY(launch_finish(c, FIN_SWEEP, FLAT_BLOCKS, nullptr));
    }
    return HS_OK;
}

int energy_sweep(hs_ctx* c) { return t_sweep_box(c, false, 1, 1, 1, c->d.nx, c->d.ny, c->d.nz); }

int solid_sweep(hs_ctx* c) {
    return t_sweep_box(c, true, c->sb[0], c->sb[2], c->sb[4], c->sb[1] - c->sb[0] + 1, c->sb[3] - c->sb[2] + 1,
                       c->sb[5] - c->sb[4] + 1);
}

int exchange_kernel(hs_ctx* c) {
    if (c->n_if == 0) return HS_OK;
    k_exchange<<<(int)((c->n_if + 255) / 256), 256, 0, c->stream>>>(c->if_fluid, c->if_solid, c->n_if, c->tf, c->ts);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}

int assemble_matrix(hs_ctx* c) {
    const Dims& d = c->d;
    if (c->prm.mode == HS_MODE_REF) {
        k_assemble_ref<<<dim3((d.nx + 63) / 64, d.ny, d.nz), 64, 0, c->stream>>>(d, c->g, c->cls, c->up, c->vp, c->wp, c->ax,
                                                                                  c->ay, c->az);
    } else {
        k_assemble_cfg<<<dim3((d.nx + 1 + 63) / 64, d.ny + 1, d.nz + 1), 64, 0, c->stream>>>(
            d, c->g, c->cls, c->up, c->vp, c->wp, c->ax, c->ay, c->az, c->prm.rho);
    }
    HS_CHECK(hipGetLastError());
    c->have_matrix = true;
    return HS_OK;
}

int assemble_rhs(hs_ctx* c) {
    const Dims& d = c->d;
    k_rhs<<<dim3((d.nx + 63) / 64, d.ny, d.nz), 64, 0, c->stream>>>(d, c->g, c->cls, c->u, c->v, c->w, c->b,
                                                                     c->prm.mode == HS_MODE_REF, c->prm.rho);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}

int spmv(hs_ctx* c, const double* x, double* y) {
    const Dims& d = c->d;
    const long tiles = (long)((d.nx + BX - 1) / BX) * ((d.ny + BY - 1) / BY);
    const int kc = kchunk(d.nz, tiles);
    const dim3 gr = grid3(d.nx, d.ny, d.nz, kc);
    k_spmv<<<grid1(gr), dim3(BX, BY), 0, c->stream>>>(op_args(c), x, y, kc, gr.x, gr.y, gr.z);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}

int dot(hs_ctx* c, const double* a, const double* b, double* out_host) {
    k_dot<<<FLAT_BLOCKS, FLAT, 0, c->stream>>>(c->d.nP, a, b, c->partials);
    HS_CHECK(hipGetLastError());
    HS_TRY(launch_finish(c, FIN_DOT, FLAT_BLOCKS, &c->sctl->red[0]));
    HS_CHECK(hipMemcpyAsync(out_host, &c->sctl->red[0], sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HS_CHECK(hipStreamSynchronize(c->stream));
    return HS_OK;
}

// HYPRE ParCSR PCG with the identity preconditioner (fluid_solver.h:1789-1790, settings
// :1871-1875), x0 = 0.  Iterations are launched in batches; the stop rule is evaluated
// on the device and later launches of a finished solve return at entry.
int cg_solve(hs_ctx* c, const double* b, double* x, hs_stats* st, bool with_resid) {
    const Dims& d = c->d;
    const long n = d.nP;
    SolveCtl* h = c->h_sctl;
    h->eps = c->prm.p_tol * c->prm.p_tol;
    h->maxit = c->prm.p_maxit;
    HS_CHECK(hipMemcpyAsync(&c->sctl->eps, &h->eps, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HS_CHECK(hipMemcpyAsync(&c->sctl->maxit, &h->maxit, sizeof(int), hipMemcpyHostToDevice, c->stream));
    k_cg_setup<<<FLAT_BLOCKS, FLAT, 0, c->stream>>>(n, b, x, c->kr, c->kp, c->partials);
    HS_CHECK(hipGetLastError());
    HS_TRY(launch_finish(c, FIN_CG_SETUP, FLAT_BLOCKS, nullptr));
    const long tiles = (long)((d.nx + BX - 1) / BX) * ((d.ny + BY - 1) / BY);
    const int kc = kchunk(d.nz, tiles);
    const dim3 gr = grid3(d.nx, d.ny, d.nz, kc);
    const OpArgs o = op_args(c);
    int launched = 0;
    int batch = c->last_iters > 8 ? c->last_iters - 4 : 8;
    // p = b for the first direction (k_cg_setup); later ones, p = r + beta p, are formed
    // inside the apply from the old direction and written to the second buffer (the same
    // arithmetic as a separate update pass), and the buffers swap.  The scalar steps run
    // inside the vector kernels: the update forms alpha from the apply's partials, and the
    // next apply forms the stop rule and beta from the update's partials, so an iteration
    // is two launches.  The update's partials sit past the apply's; the batch's last
    // update is finished by k_finish so the host reads a settled control block.
    double *p = c->kp, *p2 = c->kp2;
    const int nsp = (int)(gr.x * gr.y * gr.z);
    double* sp = c->partials;
    double* up = c->partials + std::max(nsp, FLAT_BLOCKS);
    // update blocks: four cells per thread up to 1024 blocks, so the next apply adds at
    // most 1024 partials
    const int nub = (int)std::min<long>(1024, std::max<long>(1, (n + 4 * FLAT - 1) / (4 * FLAT)));
    while (true) {
        int in_batch = 0;
        for (; in_batch < batch && launched < c->prm.p_maxit; ++in_batch, ++launched) {
            const int slot = prof_begin(c, HS_K_SPMV);
            if (launched == 0) {
                k_cg_spmv<double><<<grid1(gr), dim3(BX, BY), 0, c->stream>>>(o, p, c->ks, kc, gr.x, gr.y, gr.z, sp,
                                                                             c->sctl);
            } else {
                // the batch's first apply reads beta from the control block (k_finish)
                k_fcg_spmv<double><<<grid1(gr), dim3(BX, BY), 0, c->stream>>>(
                    o, c->kr, p, p2, c->ks, kc, gr.x, gr.y, gr.z, sp, c->sctl, in_batch == 0 ? nullptr : up, nub,
                    launched - 1);
                std::swap(p, p2);
            }
            HS_TRY(prof_end(c, slot));
            k_cg_update<<<nub, FLAT, 0, c->stream>>>(n, x, c->kr, p, c->ks, sp, nsp, launched, up, c->sctl);
        }
        if (in_batch > 0) HS_TRY(launch_finish_cg(c, up, nub, launched - 1));
        HS_CHECK(hipGetLastError());
        if (with_resid) HS_TRY(launch_resid_norms(c, b, x));
        HS_CHECK(hipMemcpyAsync(h, c->sctl, sizeof(SolveCtl), hipMemcpyDeviceToHost, c->stream));
        HS_CHECK(hipStreamSynchronize(c->stream));
        if (h->done || launched >= c->prm.p_maxit) break;
        batch = 8;
    }
    c->last_iters = h->iters;
    if (st) {
        st->p_iters = h->iters;
        st->p_converged = h->converged;
        st->p_relres = h->bb > 0 ? sqrt(h->rr / h->bb) : 0.0;
        if (with_resid) resid_from_red(h, st);
    }
    if (h->nan || !std::isfinite(h->rr)) return HS_E_NAN;
    return HS_OK;
}

int warm_residual(hs_ctx* c, const double* b, const double* x, double* r, void* b0, void* x0, bool zero_x0,
                  bool f32) {
    const Dims& d = c->d;
    const long tiles = (long)((d.nx + BX - 1) / BX) * ((d.ny + BY - 1) / BY);
    const int kc = kchunk(d.nz, tiles);
    const dim3 gr = grid3(d.nx, d.ny, d.nz, kc);
    const int nb = (int)(gr.x * gr.y * gr.z);
    if (f32)
        k_warm_resid<float><<<nb, dim3(BX, BY), 0, c->stream>>>(op_args(c), x, b, r, (float*)b0, (float*)x0,
                                                                 zero_x0 ? 1 : 0, kc, gr.x, gr.y, gr.z, c->partials);
    else
        k_warm_resid<double><<<nb, dim3(BX, BY), 0, c->stream>>>(op_args(c), x, b, r, (double*)b0, (double*)x0,
                                                                  zero_x0 ? 1 : 0, kc, gr.x, gr.y, gr.z, c->partials);
    HS_CHECK(hipGetLastError());
    return launch_finish(c, FIN_FCG_SETUP_WARM, nb, nullptr);
}

int launch_resid_norms(hs_ctx* c, const double* b, const double* x) {
    const Dims& d = c->d;
    const long tiles = (long)((d.nx + BX - 1) / BX) * ((d.ny + BY - 1) / BY);
    const int kc = kchunk(d.nz, tiles);
    const dim3 gr = grid3(d.nx, d.ny, d.nz, kc);
    const int nb = (int)(gr.x * gr.y * gr.z);
    k_resid_norms<<<nb, dim3(BX, BY), 0, c->stream>>>(op_args(c), x, b, kc, gr.x, gr.y, gr.z, c->partials);
    HS_CHECK(hipGetLastError());
    return launch_finish(c, FIN_DOT2, nb, &c->sctl->red[0]);
}

void resid_from_red(const SolveCtl* h, hs_stats* st) {
    const double rr = h->red[0], bb = h->red[1];
    st->p_true_relres = bb > 0 ? sqrt(rr / bb) : 0.0;
    st->mass_imbalance = sqrt(bb);
}

int true_relres(hs_ctx* c, const double* b, const double* x, double* relres, double* bnorm) {
    HS_TRY(launch_resid_norms(c, b, x));
    double h2[2];
    HS_CHECK(hipMemcpyAsync(h2, &c->sctl->red[0], 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HS_CHECK(hipStreamSynchronize(c->stream));
    const double rr = h2[0], bb = h2[1];
    *relres = bb > 0 ? sqrt(rr / bb) : 0.0;
    *bnorm = sqrt(bb);
    return HS_OK;
}

// velocity correction and pressure update (one pass), then p' = 0
// (fluid_solver.h:1459-1597): when p' is the solve's output in kx the pprime buffer already
// holds zeros, otherwise it is cleared
int correct_and_update(hs_ctx* c) {
    const Dims& d = c->d;
    k_correct<<<dim3((d.nx + 1 + 63) / 64, d.ny + 1, d.nz + 1), 64, 0, c->stream>>>(
        d, c->g, c->cls, pprime_now(c), c->u, c->v, c->w, c->up, c->vp, c->wp, c->prm.mode == HS_MODE_REF, c->p,
        c->ptemp, c->prm.relax_p);
    HS_CHECK(hipGetLastError());
    if (c->pprime_in_kx && !c->pprime_dirty) {
        c->pprime_in_kx = false;
        return HS_OK;
    }
    c->pprime_in_kx = false;
    c->pprime_dirty = false;
    return launch_fill(c, c->pprime, d.nP, 0.0);
}

static int energy_balance_launch(hs_ctx* c) {
    const Dims& d = c->d;
    const long tiles = (long)((d.nx + BX - 1) / BX) * ((d.ny + BY - 1) / BY);
    const int kc = kchunk(d.nz, tiles);
    const dim3 gr = grid3(d.nx, d.ny, d.nz, kc);
    if (c->qsum_version != c->q_version) {   // q changed since the last sum (or never summed)
        k_sum_q<<<FLAT_BLOCKS, FLAT, 0, c->stream>>>(d, c->g, c->q, c->cls, c->partials);
        HS_CHECK(hipGetLastError());
        HS_TRY(launch_finish(c, FIN_DOT, FLAT_BLOCKS, c->qsum));
        c->qsum_version = c->q_version;
    }
    HS_CHECK(hipMemcpyAsync(&c->sctl->red[2], c->qsum, sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    // red[1] = (sum of a_P T_P - sum_nb a_nb T_nb) - red[2]: the cells' partials, then the
    // boundary faces' behind them, finished together
    const int nbc = (int)(gr.x * gr.y * gr.z);
    const long faces = 2 * ((long)d.ny * d.nz + (long)d.nx * d.nz + (long)d.nx * d.ny);
    const int nbf = (int)std::min<long>(1024, (faces + FLAT - 1) / FLAT);
    const TArgs ta = t_args(c, false);
    k_energy_balance<<<gr, dim3(BX, BY), 0, c->stream>>>(ta, kc, c->partials);
    HS_CHECK(hipGetLastError());
    k_energy_balance_faces<<<nbf, FLAT, 0, c->stream>>>(ta, c->partials + nbc);
    HS_CHECK(hipGetLastError());
    HS_TRY(launch_finish(c, FIN_DIFF, nbc + nbf, &c->sctl->red[1]));
    return HS_OK;
}

int energy_balance_async(hs_ctx* c) {
    HS_TRY(energy_balance_launch(c));
    HS_CHECK(hipMemcpyAsync(&c->h_sctl->red[1], &c->sctl->red[1], 2 * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
    return HS_OK;
}

int energy_balance(hs_ctx* c, double* heat, double* resid) {
    HS_TRY(energy_balance_launch(c));
    double h2[2];
    HS_CHECK(hipMemcpyAsync(h2, &c->sctl->red[1], 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HS_CHECK(hipStreamSynchronize(c->stream));
    *resid = h2[0];
    *heat = h2[1];
    return HS_OK;
}

int mask_fluid(hs_ctx* c, double* x) {
    k_mask_fluid<<<1024, 256, 0, c->stream>>>(c->d.nP, c->cls, x);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}

// pn = z + beta p, q = A pn with <pn,q> and alpha (flexible CG, iterations after the first);
// f32: z, p and pn are float
int fcg_spmv_dot(hs_ctx* c, const void* z, bool f32, const void* p, void* pn, double* q) {
    const Dims& d = c->d;
    const long tiles = (long)((d.nx + BX - 1) / BX) * ((d.ny + BY - 1) / BY);
    const int kc = kchunk(d.nz, tiles);
    const dim3 gr = grid3(d.nx, d.ny, d.nz, kc);
    const int slot = prof_begin(c, HS_K_SPMV);
    if (f32)
        k_fcg_spmv<float><<<grid1(gr), dim3(BX, BY), 0, c->stream>>>(
            op_args(c), (const float*)z, (const float*)p, (float*)pn, q, kc, gr.x, gr.y, gr.z, c->partials, c->sctl,
            nullptr, 0, 0);
    else
        k_fcg_spmv<double><<<grid1(gr), dim3(BX, BY), 0, c->stream>>>(
            op_args(c), (const double*)z, (const double*)p, (double*)pn, q, kc, gr.x, gr.y, gr.z, c->partials, c->sctl,
            nullptr, 0, 0);
    HS_CHECK(hipGetLastError());
    HS_TRY(prof_end(c, slot));
    return launch_finish(c, FIN_CG_ALPHA, (int)(gr.x * gr.y * gr.z), nullptr);
}

// q = A p with <p,q> and alpha = gamma / <p,q> in the last block (flexible CG's first
// iteration; f32: p is float).
int cg_spmv_dot(hs_ctx* c, const void* p, bool f32, double* q) {
    const Dims& d = c->d;
    const long tiles = (long)((d.nx + BX - 1) / BX) * ((d.ny + BY - 1) / BY);
    const int kc = kchunk(d.nz, tiles);
    const dim3 gr = grid3(d.nx, d.ny, d.nz, kc);
    const int slot = prof_begin(c, HS_K_SPMV);
    if (f32)
        k_cg_spmv<float><<<grid1(gr), dim3(BX, BY), 0, c->stream>>>(op_args(c), (const float*)p, q, kc, gr.x, gr.y,
                                                                     gr.z, c->partials, c->sctl);
    else
        k_cg_spmv<double><<<grid1(gr), dim3(BX, BY), 0, c->stream>>>(op_args(c), (const double*)p, q, kc, gr.x, gr.y,
                                                                      gr.z, c->partials, c->sctl);
    HS_CHECK(hipGetLastError());
    HS_TRY(prof_end(c, slot));
    return launch_finish(c, FIN_CG_ALPHA, (int)(gr.x * gr.y * gr.z), nullptr);
}
