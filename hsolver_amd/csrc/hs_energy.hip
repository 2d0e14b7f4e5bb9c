// hs_energy.hip — the Mode C conjugate energy equation on a frozen flow as one linear
// system, its multigrid-preconditioned BiCGStab solve, and its discrete adjoint.
//
// Operator (padded cell layout; every interior cell, fluid or solid, is an unknown):
//   (A x)_P = a_P x_P - a_W x_W - a_E x_E - a_S x_S - a_N x_N - a_B x_B - a_T x_T
//   a_W = D_w + max(F_w, 0),  a_E = D_e + max(-F_e, 0)      (likewise S/N and B/T)
//   a_P = a_E + a_W + a_N + a_S + a_B + a_T
// with D the harmonic-mean face conductance A / (h_P/(2 k_P) + h_nb/(2 k_nb)) and F the
// convective flux rho A u through the face (positive along +axis): the coefficients of
// the Mode C energy sweep.  A face on the domain boundary keeps its conductance (ghost
// one cell width out, with the cell's k) and moves a_face T_ghost to the right-hand
// side; an outflow face (zero-gradient ghost) is stored as 0, which folds the ghost
// into the diagonal.  The sweep's fixed point is therefore exactly A T = b.
//
// Only D and F are stored, on the min face of every cell.  The transpose keeps the
// diagonal and reverses F in the off-diagonals:
//   A^T_{P,W} = A_{W,P} = -(D_w + max(-F_w, 0)),
// so one hierarchy serves the forward and the adjoint solve.
//
// Coarsening (2x2x2): D by the series rule of the pressure hierarchy, boundary faces
// included (the missing outer half of a boundary face is taken equal to the face, which
// reproduces the rediscretised boundary conductance on uniform data); F by summing the
// four sub-face fluxes, so a coarse cell keeps its children's net mass balance and the
// coarse operator stays an M-matrix.  Restriction sums the children's residuals,
// prolongation is trilinear, smoothing is red-black Gauss-Seidel.
//
// Krylov: BiCGStab (van der Vorst 1992) in FP64, right-preconditioned by one V-cycle
// (float or double), with the scalar recurrences run on the device by one-block finish
// kernels; the host reads the control block once per batch of iterations.
#include <cfloat>
#include <cmath>

#include "hs_device.h"
#include "hs_eop.h"

using hsd::BT;
using hsd::BX;
using hsd::BY;
using hsd::FLAT;
using hsd::FLAT_BLOCKS;

namespace {

using namespace heo;

// ------------------------------------------------------------ assembly ---
// Face data of the fine operator and the right-hand side, from k, q, the face
// velocities and the ghost temperatures (Mode C energy coefficients).
__global__ void k_e_assemble(Dims d, Geo g, const double* kc, const double* q, const double* u, const double* v,
                             const double* w, const double* tf, double rho, int outflow, EOp op, double* rhs) {
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y, k = 1 + blockIdx.z;
    if (i > d.nx) return;
    const long c = idxP(d, i, j, k);
    const double xp = g.xp[i], yp = g.yp[j], zp = g.zp[k];
    const double Axy = xp * yp, Axz = xp * zp, Ayz = yp * zp;
    const double kP = kc[idxC(d, i - 1, j - 1, k - 1)];
    const bool bw = i == 1, bs = j == 1, bb = k == 1;
    const double hW = bw ? xp : g.xp[i - 1], kW = bw ? kP : kc[idxC(d, i - 2, j - 1, k - 1)];
    const double hS = bs ? yp : g.yp[j - 1], kS = bs ? kP : kc[idxC(d, i - 1, j - 2, k - 1)];
    const double hB = bb ? zp : g.zp[k - 1], kB = bb ? kP : kc[idxC(d, i - 1, j - 1, k - 2)];
    const double DW = Ayz / (0.5 * xp / kP + 0.5 * hW / kW);
    const double DS = Axz / (0.5 * yp / kP + 0.5 * hS / kS);
    const double DB = Axy / (0.5 * zp / kP + 0.5 * hB / kB);
    const double FW = rho * Ayz * u[idxU(d, i, j, k)];
    const double FS = rho * Axz * v[idxV(d, i, j, k)];
    const double FB = rho * Axy * w[idxW(d, i, j, k)];
    op.dx[c] = DW; op.fx[c] = FW;
    op.dy[c] = DS; op.fy[c] = FS;
    op.dz[c] = DB; op.fz[c] = FB;
    double r = q[idxC(d, i - 1, j - 1, k - 1)] * (xp * yp * zp);
    if (bw) r += (DW + pos0(FW)) * tf[c - 1];
    if (bs) r += (DS + pos0(FS)) * tf[c - d.sPy];
    if (bb) r += (DB + pos0(FB)) * tf[c - d.sPz];
    if (i == d.nx) {
        if (outflow) {
            op.dx[c + 1] = 0.0; op.fx[c + 1] = 0.0;
        } else {
            const double DE = Ayz / (0.5 * xp / kP + 0.5 * xp / kP), FE = rho * Ayz * u[idxU(d, i + 1, j, k)];
            op.dx[c + 1] = DE; op.fx[c + 1] = FE;
            r += (DE + pos0(-FE)) * tf[c + 1];
        }
    }
    if (j == d.ny) {
        const double DN = Axz / (0.5 * yp / kP + 0.5 * yp / kP), FN = rho * Axz * v[idxV(d, i, j + 1, k)];
        op.dy[c + d.sPy] = DN; op.fy[c + d.sPy] = FN;
        r += (DN + pos0(-FN)) * tf[c + d.sPy];
    }
    if (k == d.nz) {
        const double DT = Axy / (0.5 * zp / kP + 0.5 * zp / kP), FT = rho * Axy * w[idxW(d, i, j, k + 1)];
        op.dz[c + d.sPz] = DT; op.fz[c + d.sPz] = FT;
        r += (DT + pos0(-FT)) * tf[c + d.sPz];
    }
    rhs[c] = r;
}

template <typename T>
__global__ __launch_bounds__(FLAT) void k_e_convert(long n, EOp op, T* dx, T* dy, T* dz, T* fx, T* fy, T* fz) {
    for (long t = blockIdx.x * (long)FLAT + threadIdx.x; t < n; t += (long)gridDim.x * FLAT) {
        dx[t] = (T)op.dx[t]; dy[t] = (T)op.dy[t]; dz[t] = (T)op.dz[t];
        fx[t] = (T)op.fx[t]; fy[t] = (T)op.fy[t]; fz[t] = (T)op.fz[t];
    }
}

template <typename T>
__device__ __forceinline__ T ser(T ca, T cb, T cc) {
    if (!(cb > T(0))) return T(0);
    if (!(ca > T(0))) ca = cb;
    if (!(cc > T(0))) cc = cb;
    return T(1) / (T(0.5) / ca + T(1) / cb + T(0.5) / cc);
}

template <typename T>
struct EFaces {
    T *dx, *dy, *dz, *fx, *fy, *fz;
};

// Coarse faces of coarse cell (I,J,K): its min faces, plus the max boundary face of
// the last cell along each axis.
template <typename T>
__global__ void k_e_coarsen(EL<T> F, EFaces<T> C, int cnx, int cny, int cnz, long csy, long csz) {
    const int I = 1 + blockIdx.x * blockDim.x + threadIdx.x, J = 1 + blockIdx.y, K = 1 + blockIdx.z;
    if (I > cnx) return;
    const long cc = I + csy * J + csz * K;
    auto fi = [&](long i, long j, long k) { return i + F.sy * j + F.sz * k; };
    const int i0 = 2 * I - 1, j0 = 2 * J - 1, k0 = 2 * K - 1;
    const int ni = (2 * I <= F.nx) ? 2 : 1, nj = (2 * J <= F.ny) ? 2 : 1, nk = (2 * K <= F.nz) ? 2 : 1;
    {   // x faces
        T dm = 0, fm = 0, dp = 0, fp = 0;
        for (int dk = 0; dk < nk; ++dk)
            for (int dj = 0; dj < nj; ++dj) {
                const long f = fi(i0, j0 + dj, k0 + dk);
                dm += ser<T>(F.dx[f - 1], F.dx[f], (ni == 2) ? F.dx[f + 1] : T(0));
                fm += F.fx[f];
                if (I == cnx) {
                    const long e = fi(F.nx + 1, j0 + dj, k0 + dk);
                    dp += ser<T>((ni == 2) ? F.dx[e - 1] : T(0), F.dx[e], T(0));
                    fp += F.fx[e];
                }
            }
        C.dx[cc] = dm; C.fx[cc] = fm;
        if (I == cnx) { C.dx[cc + 1] = dp; C.fx[cc + 1] = fp; }
    }
    {   // y faces
        T dm = 0, fm = 0, dp = 0, fp = 0;
        for (int dk = 0; dk < nk; ++dk)
            for (int di = 0; di < ni; ++di) {
                const long f = fi(i0 + di, j0, k0 + dk);
                dm += ser<T>(F.dy[f - F.sy], F.dy[f], (nj == 2) ? F.dy[f + F.sy] : T(0));
                fm += F.fy[f];
                if (J == cny) {
                    const long e = fi(i0 + di, F.ny + 1, k0 + dk);
                    dp += ser<T>((nj == 2) ? F.dy[e - F.sy] : T(0), F.dy[e], T(0));
                    fp += F.fy[e];
                }
            }
        C.dy[cc] = dm; C.fy[cc] = fm;
        if (J == cny) { C.dy[cc + csy] = dp; C.fy[cc + csy] = fp; }
    }
    {   // z faces
        T dm = 0, fm = 0, dp = 0, fp = 0;
        for (int dj = 0; dj < nj; ++dj)
            for (int di = 0; di < ni; ++di) {
                const long f = fi(i0 + di, j0 + dj, k0);
                dm += ser<T>(F.dz[f - F.sz], F.dz[f], (nk == 2) ? F.dz[f + F.sz] : T(0));
                fm += F.fz[f];
                if (K == cnz) {
                    const long e = fi(i0 + di, j0 + dj, F.nz + 1);
                    dp += ser<T>((nk == 2) ? F.dz[e - F.sz] : T(0), F.dz[e], T(0));
                    fp += F.fz[e];
                }
            }
        C.dz[cc] = dm; C.fz[cc] = fm;
        if (K == cnz) { C.dz[cc + csz] = dp; C.fz[cc + csz] = fp; }
    }
}

// ------------------------------------------------------- V-cycle parts ---
template <typename T>
__global__ __launch_bounds__(BT) void k_e_smooth(EL<T> L, int tr, int kc, int colour, int gx, int gy, int gz,
                                                 T omega, const int* skip) {
    if (*skip) return;
    const hsd::Tile tl = hsd::tile_of(gx, gy, gz);
    const int i = 1 + tl.bx * BX + threadIdx.x, j = 1 + tl.by * BY + threadIdx.y;
    if (i > L.nx || j > L.ny) return;
    const int k0 = 1 + tl.bz * kc, k1 = min(L.nz, k0 + kc - 1);
    for (int k = k0; k <= k1; ++k) {
        if (((i + j + k) & 1) != colour) continue;
        const long c = i + L.sy * j + L.sz * k;
        const Row<T> r = lrow(L, c, tr);
        if (!(r.p > T(0))) continue;
        L.x[c] = hsd::sor(L.x[c], (L.b[c] + offs<T, T>(r, L.x, c, L.sy, L.sz)) / r.p, omega);
    }
}

// One whole red-black sweep of the energy operator (A or A^T) in one pass, out of place
// (x -> xn; the level's two iterate buffers swap afterwards): the two colour passes of
// k_e_smooth, same operations in the same order, so the same values.  The pressure
// smoother's scheme (k_mg_rbsweep): each thread owns a column of a TX x TY tile, at plane
// step k loads its own faces of plane k+1 (x faces W/E, y faces S/N, the z face on top,
// which is the next plane's bottom face; diffusion and flux), forms that plane's row once
// and keeps it for the second colour there at step k+1.  The first colour of plane k+1 is
// relaxed on the tile and a one-cell ring (ring rows by tile rows 0 and TY-1, side cells
// by x threads 0 and 1, from memory); the second colour of plane k reads the first-colour
// planes k-1, k, k+1 from LDS (four slots, one barrier per step).  Cells whose row has no
// positive diagonal keep their value.  Byte offsets on the scalar base addresses.
template <typename T, typename O>
__device__ __forceinline__ T eld(const T* base, O off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}

template <typename T, bool TR>
__device__ __forceinline__ Row<T> row_of(T dW, T dE, T dS, T dN, T dB, T dT, T fW, T fE, T fS, T fN, T fB, T fT) {
    const T aw = dW + pos0(fW), ae = dE + pos0(-fE), as = dS + pos0(fS), an = dN + pos0(-fN);
    const T ab = dB + pos0(fB), at = dT + pos0(-fT);
    Row<T> r;
    r.p = ae + aw + an + as + ab + at;
    if (!TR) {
        r.w = aw; r.e = ae; r.s = as; r.n = an; r.b = ab; r.t = at;
    } else {
        r.w = dW + pos0(-fW); r.e = dE + pos0(fE); r.s = dS + pos0(-fS);
        r.n = dN + pos0(fN); r.b = dB + pos0(-fB); r.t = dT + pos0(fT);
    }
    return r;
}

// first-colour relaxation of one (ring) cell from memory
template <typename T, typename O, bool TR>
__device__ __forceinline__ T e_relax_at(const EL<T>& L, const T* xo, O c, O dy, O dz, T omega) {
    constexpr O E = sizeof(T);
    const Row<T> r = row_of<T, TR>(eld(L.dx, c), eld(L.dx, c + E), eld(L.dy, c), eld(L.dy, c + dy), eld(L.dz, c),
                                   eld(L.dz, c + dz), eld(L.fx, c), eld(L.fx, c + E), eld(L.fy, c), eld(L.fy, c + dy),
                                   eld(L.fz, c), eld(L.fz, c + dz));
    if (!(r.p > T(0))) return eld(xo, c);
    const T off = r.w * eld(xo, c - E) + r.e * eld(xo, c + E) + r.s * eld(xo, c - dy) + r.n * eld(xo, c + dy) +
                  r.b * eld(xo, c - dz) + r.t * eld(xo, c + dz);
    return hsd::sor(eld(xo, c), (eld(L.b, c) + off) / r.p, omega);
}

template <typename T, int TX, int TY, typename O, bool TR>
__global__ __launch_bounds__(TX * TY) void k_e_rbsweep(EL<T> L, const T* __restrict__ xo, T* __restrict__ xn, int c1,
                                                       int kc, int gx, int gy, int gz, T omega, const int* skip) {
    constexpr int RX = TX + 2, RY = TY + 2;
    constexpr O E = sizeof(T);
    if (*skip) return;
    __shared__ T X1[4][RY * RX];
    const hsd::Tile tl = hsd::tile_of(gx, gy, gz);
    const int i0 = 1 + tl.bx * TX, j0 = 1 + tl.by * TY;
    const int k0 = 1 + tl.bz * kc, k1 = min(L.nz, k0 + kc - 1);
    const int lane = threadIdx.x, w = threadIdx.y;
    const int i = i0 + lane, j = j0 + w;
    const bool inside = i <= L.nx && j <= L.ny;
    const int r = (w + 1) * RX + lane + 1;
    const O dy = (O)L.sy * E, dz = (O)L.sz * E;
    const O col = (O)(i + L.sy * j) * E;
    const int rowring = (w == 0) ? j0 - 1 : (w == TY - 1 ? j0 + TY : -1);
    const int rowslot = (w == 0) ? lane + 1 : (TY + 1) * RX + lane + 1;
    const bool rowring_in = rowring >= 1 && rowring <= L.ny && i <= L.nx;
    const int si = lane == 0 ? i0 - 1 : i0 + TX;
    const int sslot = (w + 1) * RX + (lane == 0 ? 0 : TX + 1);
    const bool side_in = lane < 2 && j <= L.ny && si >= 1 && si <= L.nx;

    Row<T> R0{}, R1{};         // rows of planes k and k+1
    T b0 = 0, b1 = 0;          // right-hand side, planes k and k+1
    T dzB = 0, fzB = 0;        // bottom z faces of plane k+1 (the top faces of plane k)
    T xk = 0, xk1 = 0, xk2 = 0;
    T own1 = 0, own1n = 0;
    const int kfirst = k0 - 2;
    if (inside) {
        if (kfirst >= 0) xk = eld(xo, col + (O)kfirst * dz);
        xk1 = eld(xo, col + (O)(k0 - 1) * dz);
        dzB = eld(L.dz, col + (O)(k0 - 1) * dz);
        fzB = eld(L.fz, col + (O)(k0 - 1) * dz);
    }
    for (int k = kfirst; k <= k1; ++k) {
        const int kp = k + 1;   // in [0, nz + 1]
        if (inside) {
            const O c = col + (O)kp * dz;
            const T dzT = kp + 1 <= L.nz + 1 ? eld(L.dz, c + dz) : T(0);
            const T fzT = kp + 1 <= L.nz + 1 ? eld(L.fz, c + dz) : T(0);
            R1 = row_of<T, TR>(eld(L.dx, c), eld(L.dx, c + E), eld(L.dy, c), eld(L.dy, c + dy), dzB, dzT, eld(L.fx, c),
                               eld(L.fx, c + E), eld(L.fy, c), eld(L.fy, c + dy), fzB, fzT);
            b1 = eld(L.b, c);
            if (kp + 1 <= L.nz + 1) xk2 = eld(xo, c + dz);
            dzB = dzT;
            fzB = fzT;
        }
        const bool kin = kp >= 1 && kp <= L.nz;
        if (rowring >= 0) {
            T v = T(0);
            if (rowring_in && kin && ((i + rowring + kp) & 1) == c1)
                v = e_relax_at<T, O, TR>(L, xo, (O)(i + L.sy * rowring) * E + (O)kp * dz, dy, dz, omega);
            X1[kp & 3][rowslot] = v;
        }
        if (lane < 2) {
            T v = T(0);
            if (side_in && kin && ((si + j + kp) & 1) == c1)
                v = e_relax_at<T, O, TR>(L, xo, (O)(si + L.sy * j) * E + (O)kp * dz, dy, dz, omega);
            X1[kp & 3][sslot] = v;
        }
        {
            T v = T(0);
            if (inside && kin && ((i + j + kp) & 1) == c1) {
                if (!(R1.p > T(0))) {
                    v = xk1;
                } else {
                    const O c = col + (O)kp * dz;
                    const T off = R1.w * eld(xo, c - E) + R1.e * eld(xo, c + E) + R1.s * eld(xo, c - dy) +
                                  R1.n * eld(xo, c + dy) + R1.b * xk + R1.t * xk2;
                    v = hsd::sor(xk1, (b1 + off) / R1.p, omega);
                }
            }
            X1[kp & 3][r] = v;
            own1n = v;
        }
        if (k >= k0) {
            __syncthreads();
            if (inside) {
                T out;
                if (((i + j + k) & 1) == c1) {
                    out = own1;
                } else if (!(R0.p > T(0))) {
                    out = xk;
                } else {
                    const T* P0 = X1[k & 3];
                    const T* Pm = X1[(k - 1) & 3];
                    const T* Pp = X1[kp & 3];
                    const T off = R0.w * P0[r - 1] + R0.e * P0[r + 1] + R0.s * P0[r - RX] + R0.n * P0[r + RX] +
                                  R0.b * Pm[r] + R0.t * Pp[r];
                    out = hsd::sor(xk, (b0 + off) / R0.p, omega);
                }
                *reinterpret_cast<T*>(reinterpret_cast<char*>(xn) + col + (O)k * dz) = out;
            }
        }
        R0 = R1;
        b0 = b1;
        xk = xk1;
        xk1 = xk2;
        own1 = own1n;
    }
}

template <typename T>
__device__ __forceinline__ T e_residual_sum(const EL<T>& F, int I, int J, int K, int tr) {
    const int i0 = 2 * I - 1, j0 = 2 * J - 1, k0 = 2 * K - 1;
    const int ni = (2 * I <= F.nx) ? 2 : 1, nj = (2 * J <= F.ny) ? 2 : 1, nk = (2 * K <= F.nz) ? 2 : 1;
    T r = 0;
    for (int dk = 0; dk < nk; ++dk)
        for (int dj = 0; dj < nj; ++dj)
            for (int di = 0; di < ni; ++di) {
                const long f = (i0 + di) + F.sy * (j0 + dj) + F.sz * (k0 + dk);
                const Row<T> rr = lrow(F, f, tr);
                r += F.b[f] - (rr.p * F.x[f] - offs<T, T>(rr, F.x, f, F.sy, F.sz));
            }
    return r;
}

template <typename T>
__global__ __launch_bounds__(BT) void k_e_restrict(EL<T> F, EL<T> C, int tr, int kc, int gx, int gy, int gz,
                                                   const int* skip) {
    // one thread per fine column of a 64 x 4 fine tile (coarse 32 x 2); the 2 x 2 in-plane
    // sums of the children's residuals go through LDS
    if (*skip) return;
    __shared__ T red[BY][BX];
    const hsd::Tile tl = hsd::tile_of(gx, gy, gz);
    const int i = 1 + tl.bx * BX + threadIdx.x, j = 1 + tl.by * BY + threadIdx.y;
    const bool inside = i <= F.nx && j <= F.ny;
    const int t = threadIdx.x + BX * threadIdx.y;
    const int ci = t % (BX / 2), cj = t / (BX / 2);
    const int I = 1 + tl.bx * (BX / 2) + ci, J = 1 + tl.by * (BY / 2) + cj;
    const bool owner = t < (BX / 2) * (BY / 2) && I <= C.nx && J <= C.ny;
    const int K0 = 1 + tl.bz * kc, K1 = min(C.nz, K0 + kc - 1);
    for (int K = K0; K <= K1; ++K) {
        T r = 0;
        if (inside) {
            for (int dk = 0; dk < 2; ++dk) {
                const int k = 2 * K - 1 + dk;
                if (k > F.nz) break;
                const long f = i + F.sy * j + F.sz * k;
                const Row<T> rr = lrow(F, f, tr);
                r += F.b[f] - (rr.p * F.x[f] - offs<T, T>(rr, F.x, f, F.sy, F.sz));
            }
        }
        red[threadIdx.y][threadIdx.x] = r;
        __syncthreads();
        if (owner) {
            const long cc = I + C.sy * J + C.sz * K;
            C.b[cc] = (red[2 * cj][2 * ci] + red[2 * cj][2 * ci + 1]) + (red[2 * cj + 1][2 * ci] + red[2 * cj + 1][2 * ci + 1]);
            C.x[cc] = T(0);
        }
        __syncthreads();
    }
}

// trilinear interpolation of the coarse iterate (27/64, 9/64, 3/64, 1/64 weights);
// ghost coarse cells hold 0
template <typename T>
__device__ __forceinline__ T e_prolong_value(const EL<T>& C, int i, int j, int k) {
    const int I = (i + 1) >> 1, J = (j + 1) >> 1, K = (k + 1) >> 1;
    const int di = (i & 1) ? -1 : 1, dj = (j & 1) ? -1 : 1, dk = (k & 1) ? -1 : 1;
    const long c0 = I + C.sy * J + C.sz * K;
    const long ox = di, oy = dj * C.sy, oz = dk * C.sz;
    const T* x = C.x;
    const T a = T(0.75), b = T(0.25);
    const T zlo = a * (a * x[c0] + b * x[c0 + ox]) + b * (a * x[c0 + oy] + b * x[c0 + ox + oy]);
    const T zhi = a * (a * x[c0 + oz] + b * x[c0 + ox + oz]) + b * (a * x[c0 + oy + oz] + b * x[c0 + ox + oy + oz]);
    return a * zlo + b * zhi;
}

template <typename T>
__global__ __launch_bounds__(BT) void k_e_prolong(EL<T> F, EL<T> C, int kc, int gx, int gy, int gz, const int* skip) {
    if (*skip) return;
    const hsd::Tile tl = hsd::tile_of(gx, gy, gz);
    const int i = 1 + tl.bx * BX + threadIdx.x, j = 1 + tl.by * BY + threadIdx.y;
    if (i > F.nx || j > F.ny) return;
    const int k0 = 1 + tl.bz * kc, k1 = min(F.nz, k0 + kc - 1);
    for (int k = k0; k <= k1; ++k) F.x[i + F.sy * j + F.sz * k] += e_prolong_value(C, i, j, k);
}

// single-workgroup tail for the small levels (see the pressure hierarchy)
constexpr int E_TAIL_THREADS = 1024;
constexpr int E_TAIL_MAX_CELLS = 4096;
constexpr int E_MAX_LEVELS = 16;

template <typename T>
struct ELPack {
    EL<T> l[E_MAX_LEVELS];
};

template <typename T>
__device__ void et_smooth(const EL<T>& L, int tr, int colour) {
    const long n = (long)L.nx * L.ny * L.nz;
    for (long t = threadIdx.x; t < n; t += E_TAIL_THREADS) {
        const int i = 1 + (int)(t % L.nx), j = 1 + (int)((t / L.nx) % L.ny), k = 1 + (int)(t / ((long)L.nx * L.ny));
        if (((i + j + k) & 1) != colour) continue;
        const long c = i + L.sy * j + L.sz * k;
        const Row<T> r = lrow(L, c, tr);
        if (!(r.p > T(0))) continue;
        L.x[c] = (L.b[c] + offs<T, T>(r, L.x, c, L.sy, L.sz)) / r.p;
    }
    __syncthreads();
}

template <typename T>
__device__ void et_restrict(const EL<T>& F, const EL<T>& C, int tr) {
    const long n = (long)C.nx * C.ny * C.nz;
    for (long t = threadIdx.x; t < n; t += E_TAIL_THREADS) {
        const int I = 1 + (int)(t % C.nx), J = 1 + (int)((t / C.nx) % C.ny), K = 1 + (int)(t / ((long)C.nx * C.ny));
        const long cc = I + C.sy * J + C.sz * K;
        C.b[cc] = e_residual_sum(F, I, J, K, tr);
        C.x[cc] = T(0);
    }
    __syncthreads();
}

template <typename T>
__device__ void et_prolong(const EL<T>& F, const EL<T>& C) {
    const long n = (long)F.nx * F.ny * F.nz;
    for (long t = threadIdx.x; t < n; t += E_TAIL_THREADS) {
        const int i = 1 + (int)(t % F.nx), j = 1 + (int)((t / F.nx) % F.ny), k = 1 + (int)(t / ((long)F.nx * F.ny));
        F.x[i + F.sy * j + F.sz * k] += e_prolong_value(C, i, j, k);
    }
    __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(E_TAIL_THREADS) void k_e_tail(ELPack<T> P, int t, int nl, int tr, int pre, int post,
                                                            int coarse, const int* skip) {
    if (*skip) return;
    for (int l = t; l + 1 < nl; ++l) {
        for (int s = 0; s < pre; ++s) { et_smooth(P.l[l], tr, 0); et_smooth(P.l[l], tr, 1); }
        et_restrict(P.l[l], P.l[l + 1], tr);
    }
    for (int s = 0; s < coarse; ++s) { et_smooth(P.l[nl - 1], tr, 0); et_smooth(P.l[nl - 1], tr, 1); }
    for (int s = 0; s < coarse; ++s) { et_smooth(P.l[nl - 1], tr, 1); et_smooth(P.l[nl - 1], tr, 0); }
    for (int l = nl - 2; l >= t; --l) {
        et_prolong(P.l[l], P.l[l + 1]);
        for (int s = 0; s < post; ++s) { et_smooth(P.l[l], tr, 1); et_smooth(P.l[l], tr, 0); }
    }
}

// --------------------------------------------------------------- BiCGStab ---
enum BiKind : int { BI_INIT = 0, BI_ALPHA, BI_S, BI_OMEGA, BI_UPDATE, BI_DOT };

#define E_STRIDE(t, n) for (long t = blockIdx.x * (long)FLAT + threadIdx.x; t < (n); t += (long)gridDim.x * FLAT)

// x = 0, r = r^ = b, p = v = 0; partials of <b,b>
__global__ __launch_bounds__(FLAT) void k_bi_init(long n, const double* b, double* x, double* r, double* rh,
                                                  double* p, double* v, double* partials) {
    double acc = 0.0;
    E_STRIDE(t, n) {
        const double bv = b[t];
        x[t] = 0.0; r[t] = bv; rh[t] = bv; p[t] = 0.0; v[t] = 0.0;
        acc += bv * bv;
    }
    hsd::put_partial2<FLAT>(acc, 0.0, threadIdx.x, blockIdx.x, gridDim.x, partials);
}

// p = r + beta (p - omega v); V-cycle input b0 = p, x0 = 0
template <typename T>
__global__ __launch_bounds__(FLAT) void k_bi_p(long n, const double* r, double* p, const double* v, T* b0, T* x0,
                                               const BiCtl* ctl) {
    if (ctl->skip) return;
    const double beta = ctl->beta, omega = ctl->omega;
    E_STRIDE(t, n) {
        const double pv = r[t] + beta * (p[t] - omega * v[t]);
        p[t] = pv;
        b0[t] = (T)pv;
        x0[t] = T(0);
    }
}

// xd = x0 (the V-cycle output, widened), y = A xd or A^T xd on interior cells;
// partials: which = 0 -> <a, y>, 0;  which = 1 -> <y, a>, <y, y>
template <typename TX>
__global__ __launch_bounds__(BT) void k_bi_spmv(Dims d, EOp op, int tr, const TX* x0, double* xd, double* y,
                                                const double* a, int which, int kc, int gx, int gy, int gz,
                                                double* partials, const BiCtl* ctl) {
    if (ctl && ctl->skip) return;
    const int tid = threadIdx.x + BX * threadIdx.y;
    const hsd::Tile tl = hsd::tile_of(gx, gy, gz);
    const int i = 1 + tl.bx * BX + threadIdx.x, j = 1 + tl.by * BY + threadIdx.y;
    double s1 = 0.0, s2 = 0.0;
    if (i <= d.nx && j <= d.ny) {
        const int k0 = 1 + tl.bz * kc, k1 = min(d.nz, k0 + kc - 1);
        for (int k = k0; k <= k1; ++k) {
            const long c = idxP(d, i, j, k);
            const Row<double> r = row<double>(op.dx, op.dy, op.dz, op.fx, op.fy, op.fz, c, d.sPy, d.sPz, tr);
            const double xc = (double)x0[c];
            const double yv = r.p * xc - offs<double, TX>(r, x0, c, d.sPy, d.sPz);
            if (xd) xd[c] = xc;
            y[c] = yv;
            if (a) {
                if (which == 0) s1 += a[c] * yv;
                else { s1 += yv * a[c]; s2 += yv * yv; }
            }
        }
    }
    if (partials) hsd::put_partial2<BT>(s1, s2, tid, blockIdx.x, gx * gy * gz, partials);
}

// s = r - alpha v (in r); V-cycle input b0 = s, x0 = 0; partials of <s,s>
template <typename T>
__global__ __launch_bounds__(FLAT) void k_bi_s(long n, double* r, const double* v, T* b0, T* x0, double* partials,
                                               const BiCtl* ctl) {
    if (ctl->skip) return;
    const double alpha = ctl->alpha;
    double acc = 0.0;
    E_STRIDE(t, n) {
        const double sv = r[t] - alpha * v[t];
        r[t] = sv;
        b0[t] = (T)sv;
        x0[t] = T(0);
        acc += sv * sv;
    }
    hsd::put_partial2<FLAT>(acc, 0.0, threadIdx.x, blockIdx.x, gridDim.x, partials);
}

// restart after a (near) breakdown: the shadow residual becomes the current residual
__global__ __launch_bounds__(FLAT) void k_bi_restart(long n, const double* r, double* rh, const BiCtl* ctl) {
    if (ctl->done || !ctl->restart) return;
    E_STRIDE(t, n) rh[t] = r[t];
}

// x += alpha p^ + omega s^, r = s - omega t; partials of <r,r>, <r^,r>
__global__ __launch_bounds__(FLAT) void k_bi_update(long n, double* x, double* r, const double* ph, const double* sh,
                                                    const double* tv, const double* rh, double* partials,
                                                    const BiCtl* ctl) {
    if (ctl->done || (ctl->skip && !ctl->half)) return;
    const double alpha = ctl->alpha, omega = ctl->half ? 0.0 : ctl->omega;
    double a = 0.0, b = 0.0;
    E_STRIDE(t, n) {
        x[t] += alpha * ph[t] + omega * sh[t];
        const double rv = r[t] - omega * tv[t];
        r[t] = rv;
        a += rv * rv;
        b += rh[t] * rv;
    }
    hsd::put_partial2<FLAT>(a, b, threadIdx.x, blockIdx.x, gridDim.x, partials);
}

constexpr int BI_MAX_RESTARTS = 8;

// restart bookkeeping: r^ := r (copied by k_bi_restart at the next iteration's start),
// rho = <r, r>, p = r (beta = 0)
__device__ __forceinline__ bool bi_try_restart(BiCtl* c) {
    if (c->restarts >= BI_MAX_RESTARTS || !isfinite(c->rr) || !(c->rr > 0.0)) return false;
    c->restarts += 1;
    c->restart = 1;
    c->rho = c->rr;
    c->rhn = c->rr;
    c->beta = 0.0;
    return true;
}

__global__ __launch_bounds__(1024) void k_bi_finish(const double* partials, int n, int kind, BiCtl* c) {
    if (kind == BI_UPDATE ? c->done : (kind != BI_INIT && kind != BI_DOT && c->skip)) return;
    const int tid = threadIdx.x;
    const double s1 = hsd::sum_partials<1024>(partials, n, tid);
    const double s2 = hsd::sum_partials<1024>(partials + n, n, tid);
    if (tid != 0) return;
    switch (kind) {
        case BI_INIT:
            c->bb = s1; c->rr = s1; c->rho = s1; c->rhn = s1;
            c->alpha = 1.0; c->omega = 1.0; c->beta = 0.0;
            c->iters = 0; c->half = 0; c->breakdown = 0; c->restart = 0; c->restarts = 0; c->wbrk = 0;
            c->converged = (s1 == 0.0) ? 1 : 0;
            c->done = (s1 > 0.0 && isfinite(s1)) ? 0 : 1;
            c->skip = c->done;
            break;
        case BI_ALPHA:
            c->restart = 0;   // this iteration started from the restarted shadow residual
            if (!(s1 != 0.0) || !isfinite(s1)) {
                c->skip = 1;  // abandon the iteration
                if (!bi_try_restart(c)) { c->breakdown = 1; c->done = 1; }
            } else {
                c->alpha = c->rho / s1;
            }
            break;
        case BI_S:
            c->ss = s1;
            if (s1 <= c->eps * c->bb) { c->half = 1; c->skip = 1; }
            break;
        case BI_OMEGA:
            c->ts = s1; c->tt = s2;
            if (s2 > 0.0 && isfinite(s1 / s2) && s1 != 0.0) c->omega = s1 / s2;
            else { c->omega = 0.0; c->wbrk = 1; }
            break;
        case BI_UPDATE: {
            if (c->skip && !c->half) { c->skip = 0; break; }   // abandoned iteration: a restart is pending
            c->iters += 1;
            c->rr = s1;
            const double rho_new = s2;
            if (c->half || s1 <= c->eps * c->bb) { c->converged = 1; c->done = 1; }
            else if (!isfinite(s1)) c->done = 1;
            else if (c->iters >= c->maxit) c->done = 1;
            else if (c->wbrk || !(fabs(rho_new) > 1e-10 * sqrt(c->rhn * s1))) {
                if (!bi_try_restart(c)) { c->breakdown = 1; c->done = 1; }
            } else {
                c->beta = (rho_new / c->rho) * (c->alpha / c->omega);
                c->rho = rho_new;
            }
            c->wbrk = 0;
            c->half = 0;
            c->skip = c->done;
            break;
        }
        case BI_DOT:
            c->red[0] = s1; c->red[1] = s2;
            break;
    }
}

// partials of ||b - A x||^2 and ||b||^2 over interior cells
__global__ __launch_bounds__(BT) void k_e_resid(Dims d, EOp op, int tr, const double* x, const double* b, int kc, int gx,
                                                int gy, int gz, double* partials) {
    const int tid = threadIdx.x + BX * threadIdx.y;
    const hsd::Tile tl = hsd::tile_of(gx, gy, gz);
    const int i = 1 + tl.bx * BX + threadIdx.x, j = 1 + tl.by * BY + threadIdx.y;
    double s1 = 0.0, s2 = 0.0;
    if (i <= d.nx && j <= d.ny) {
        const int k0 = 1 + tl.bz * kc, k1 = min(d.nz, k0 + kc - 1);
        for (int k = k0; k <= k1; ++k) {
            const long c = idxP(d, i, j, k);
            const Row<double> r = row<double>(op.dx, op.dy, op.dz, op.fx, op.fy, op.fz, c, d.sPy, d.sPz, tr);
            const double e = b[c] - (r.p * x[c] - offs<double, double>(r, x, c, d.sPy, d.sPz));
            s1 += e * e;
            s2 += b[c] * b[c];
        }
    }
    hsd::put_partial2<BT>(s1, s2, tid, blockIdx.x, gx * gy * gz, partials);
}

// ----------------------------------------------------- data movement ---
// padded P vector <-> C layout cell array
__global__ void k_e_gather(Dims d, const double* src, double* dst) {   // P -> C
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y, k = 1 + blockIdx.z;
    if (i > d.nx) return;
    dst[idxC(d, i - 1, j - 1, k - 1)] = src[idxP(d, i, j, k)];
}
__global__ void k_e_scatter(Dims d, const double* src, double* dst) {  // C -> P interior
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y, k = 1 + blockIdx.z;
    if (i > d.nx) return;
    dst[idxP(d, i, j, k)] = src[idxC(d, i - 1, j - 1, k - 1)];
}

// ------------------------------------------------------ adjoint gradient ---
// R(T; q, k) = A(k) T - b(q, k) = 0,  A^T lambda = dJ/dT  =>
//   dJ/dq_P = lambda_P V_P
//   dJ/dk_P = -sum over faces of P of dD/dk_P (lambda_P - lambda_nb)(T_P - T_nb)      interior faces
//             -sum over boundary faces    dD/dk_P  lambda_P (T_P - T_ghost)           Dirichlet faces
// with D = A/(h_P/(2k_P) + h_nb/(2k_nb)), dD/dk_P = A (h_P/2) / (k_P^2 den^2); on a
// boundary face D = A k_P / h_P and dD/dk_P = A / h_P; outflow faces carry no D.
__device__ __forceinline__ double dk_face(double A, double hP, double kP, double hN, double kN) {
    const double den = 0.5 * hP / kP + 0.5 * hN / kN;
    return A * (0.5 * hP) / (kP * kP * den * den);
}

__global__ void k_e_grad(Dims d, Geo g, const double* kc, const double* tf, const double* lam, int outflow,
                         double* dJdq, double* dJdk) {
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y, k = 1 + blockIdx.z;
    if (i > d.nx) return;
    const long c = idxP(d, i, j, k);
    const long ic = idxC(d, i - 1, j - 1, k - 1);
    const double xp = g.xp[i], yp = g.yp[j], zp = g.zp[k];
    const double lc = lam[c];
    if (dJdq) dJdq[ic] = lc * (xp * yp * zp);
    if (!dJdk) return;
    const double kP = kc[ic], Tc = tf[c];
    const double Axy = xp * yp, Axz = xp * zp, Ayz = yp * zp;
    double s = 0.0;
    // x-
    if (i > 1) s += dk_face(Ayz, xp, kP, g.xp[i - 1], kc[ic - 1]) * (lc - lam[c - 1]) * (Tc - tf[c - 1]);
    else s += (Ayz / xp) * lc * (Tc - tf[c - 1]);
    // x+
    if (i < d.nx) s += dk_face(Ayz, xp, kP, g.xp[i + 1], kc[ic + 1]) * (lc - lam[c + 1]) * (Tc - tf[c + 1]);
    else if (!outflow) s += (Ayz / xp) * lc * (Tc - tf[c + 1]);
    // y-
    const long cy = d.nx, ky = d.sPy;
    if (j > 1) s += dk_face(Axz, yp, kP, g.yp[j - 1], kc[ic - cy]) * (lc - lam[c - ky]) * (Tc - tf[c - ky]);
    else s += (Axz / yp) * lc * (Tc - tf[c - ky]);
    // y+
    if (j < d.ny) s += dk_face(Axz, yp, kP, g.yp[j + 1], kc[ic + cy]) * (lc - lam[c + ky]) * (Tc - tf[c + ky]);
    else s += (Axz / yp) * lc * (Tc - tf[c + ky]);
    // z-
    const long cz = (long)d.nx * d.ny, kz = d.sPz;
    if (k > 1) s += dk_face(Axy, zp, kP, g.zp[k - 1], kc[ic - cz]) * (lc - lam[c - kz]) * (Tc - tf[c - kz]);
    else s += (Axy / zp) * lc * (Tc - tf[c - kz]);
    // z+
    if (k < d.nz) s += dk_face(Axy, zp, kP, g.zp[k + 1], kc[ic + cz]) * (lc - lam[c + kz]) * (Tc - tf[c + kz]);
    else s += (Axy / zp) * lc * (Tc - tf[c + kz]);
    dJdk[ic] = -s;
}

// ------------------------------------------------------------ host side ---
inline dim3 cells_grid(int nx, int ny, int nz) { return dim3((nx + 63) / 64, ny, nz); }

int e_dalloc(void** p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) return HS_E_NOMEM;
    if (hipMemset(*p, 0, bytes) != hipSuccess) return HS_E_HIP;
    return HS_OK;
}

template <typename T>
int e_alloc_levels(hs_ctx* c) {
    EnergyState* E = c->en;
    const Dims& d = c->d;
    int nx = d.nx, ny = d.ny, nz = d.nz;
    // the aggregation multigrid builds its own levels below the finest one
    const int maxl = c->prm.e_precond == HS_PRECOND_AMG
                         ? 1
                         : std::min(E_MAX_LEVELS, c->prm.mg_levels > 0 ? c->prm.mg_levels : E_MAX_LEVELS);
    for (int l = 0; l < maxl; ++l) {
        ELevel* m = new ELevel();
        m->nx = nx; m->ny = ny; m->nz = nz;
        m->sy = nx + 2; m->sz = (long)(nx + 2) * (ny + 2); m->n = m->sz * (nz + 2);
        m->dx = m->dy = m->dz = m->fx = m->fy = m->fz = m->x = m->b = m->x2 = nullptr;
        m->own_faces = !(l == 0 && sizeof(T) == sizeof(double));
        E->lv.push_back(m);
        if (m->own_faces) {
            void** f[6] = {&m->dx, &m->dy, &m->dz, &m->fx, &m->fy, &m->fz};
            for (void** a : f) HS_TRY(e_dalloc(a, m->n * sizeof(T)));
        } else {
            m->dx = E->dx; m->dy = E->dy; m->dz = E->dz; m->fx = E->fx; m->fy = E->fy; m->fz = E->fz;
        }
        HS_TRY(e_dalloc(&m->x, m->n * sizeof(T)));
        HS_TRY(e_dalloc(&m->x2, m->n * sizeof(T)));
        HS_TRY(e_dalloc(&m->b, m->n * sizeof(T)));
        if (std::min(nx, std::min(ny, nz)) <= 4) break;
        nx = (nx + 1) / 2; ny = (ny + 1) / 2; nz = (nz + 1) / 2;
    }
    return HS_OK;
}

int e_alloc(hs_ctx* c) {
    if (c->en) return HS_OK;
    EnergyState* E = new EnergyState();
    c->en = E;
    E->f32 = c->prm.precision == HS_PREC_MIXED;
    const long n = c->d.nP;
    double** v[] = {&E->dx, &E->dy, &E->dz, &E->fx, &E->fy, &E->fz, &E->rhs, &E->x, &E->r, &E->rh,
                    &E->p,  &E->v,  &E->t,  &E->ph, &E->sh};
    for (double** a : v) HS_TRY(e_dalloc((void**)a, n * sizeof(double)));
    double** s[] = {&E->stage, &E->stage2, &E->stage3};
    for (double** a : s) HS_TRY(e_dalloc((void**)a, c->d.nC * sizeof(double)));
    HS_TRY(e_dalloc((void**)&E->ctl, sizeof(BiCtl)));
    if (hipHostMalloc((void**)&E->h_ctl, sizeof(BiCtl)) != hipSuccess) return HS_E_NOMEM;
    return E->f32 ? e_alloc_levels<float>(c) : e_alloc_levels<double>(c);
}

EOp eop(const EnergyState* E) { return EOp{E->dx, E->dy, E->dz, E->fx, E->fy, E->fz}; }

template <typename T>
int e_build(hs_ctx* c) {
    EnergyState* E = c->en;
    const Dims& d = c->d;
    E->build_id += 1;
    k_e_assemble<<<cells_grid(d.nx, d.ny, d.nz), 64, 0, c->stream>>>(d, c->g, c->k, c->q, c->u, c->v, c->w, c->tf,
                                                                     c->prm.rho, c->prm.outflow, eop(E), E->rhs);
    HS_CHECK(hipGetLastError());
    ELevel* l0 = E->lv[0];
    if (l0->own_faces)
        k_e_convert<T><<<FLAT_BLOCKS, FLAT, 0, c->stream>>>(d.nP, eop(E), (T*)l0->dx, (T*)l0->dy, (T*)l0->dz,
                                                           (T*)l0->fx, (T*)l0->fy, (T*)l0->fz);
    HS_CHECK(hipGetLastError());
    for (size_t l = 0; l + 1 < E->lv.size(); ++l) {
        const ELevel* f = E->lv[l];
        ELevel* m = E->lv[l + 1];
        EFaces<T> C{(T*)m->dx, (T*)m->dy, (T*)m->dz, (T*)m->fx, (T*)m->fy, (T*)m->fz};
        k_e_coarsen<T><<<cells_grid(m->nx, m->ny, m->nz), 64, 0, c->stream>>>(el<T>(f), C, m->nx, m->ny, m->nz, m->sy,
                                                                             m->sz);
        HS_CHECK(hipGetLastError());
    }
    return HS_OK;
}

template <typename T, int TX, int TY>
int launch_e_rbsweep(hs_ctx* c, const EL<T>& L, const T* src, T* dst, int c1, int tr, const int* skip) {
    const int gx = (L.nx + TX - 1) / TX, gy = (L.ny + TY - 1) / TY;
    long want = (1024L * 512 / (TX * TY)) / ((long)gx * gy);
    if (want < 1) want = 1;
    int kc = (int)((L.nz + want - 1) / want);
    if (kc < 1) kc = 1;
    const int gz = (L.nz + kc - 1) / kc;
    const bool small = !c->prm.mg_wide_offsets && (double)L.sz * (L.nz + 2) * sizeof(T) < 2147483648.0;
    const dim3 g(gx * gy * gz), b(TX, TY);
    const T omega = (T)c->prm.e_omega;
    if (small) {
        if (tr) k_e_rbsweep<T, TX, TY, uint32_t, true><<<g, b, 0, c->stream>>>(L, src, dst, c1, kc, gx, gy, gz, omega, skip);
        else k_e_rbsweep<T, TX, TY, uint32_t, false><<<g, b, 0, c->stream>>>(L, src, dst, c1, kc, gx, gy, gz, omega, skip);
    } else {
        if (tr) k_e_rbsweep<T, TX, TY, uint64_t, true><<<g, b, 0, c->stream>>>(L, src, dst, c1, kc, gx, gy, gz, omega, skip);
        else k_e_rbsweep<T, TX, TY, uint64_t, false><<<g, b, 0, c->stream>>>(L, src, dst, c1, kc, gx, gy, gz, omega, skip);
    }
    HS_CHECK(hipGetLastError());
    return HS_OK;
}

template <typename T>
int e_smooth(hs_ctx* c, ELevel* m, int tr, int sweeps, bool forward, const int* skip, bool level0) {
    EL<T> L = el<T>(m);
    const long tiles = (long)((L.nx + BX - 1) / BX) * ((L.ny + BY - 1) / BY);
    const int kc = hsd::kchunk(L.nz, tiles);
    const dim3 gr = hsd::grid3(L.nx, L.ny, L.nz, kc);
    // fused out-of-place sweeps (tile by level width), unless mg_sweep_rows < 0 asks for
    // the two in-place colour passes; after each sweep the level's iterate buffers swap
    if (c->prm.mg_sweep_rows >= 0) {
        const int c1 = forward ? 0 : 1;
        for (; sweeps > 0; --sweeps) {
            const T* src = (const T*)m->x;
            T* dst = (T*)m->x2;
            const int slot = level0 ? prof_begin(c, HS_K_E_SMOOTH0) : -1;
            if (L.nx > 128) HS_TRY((launch_e_rbsweep<T, 256, 4>(c, L, src, dst, c1, tr, skip)));
            else if (L.nx > 64) HS_TRY((launch_e_rbsweep<T, 128, 8>(c, L, src, dst, c1, tr, skip)));
            else HS_TRY((launch_e_rbsweep<T, 64, 8>(c, L, src, dst, c1, tr, skip)));
            HS_TRY(prof_end(c, slot));
            std::swap(m->x, m->x2);
            L.x = (T*)m->x;
        }
    }
    for (int s = 0; s < sweeps; ++s)
        for (int h = 0; h < 2; ++h) {
            const int colour = forward ? h : 1 - h;
            const int slot = level0 ? prof_begin(c, HS_K_E_SMOOTH0) : -1;
            k_e_smooth<T><<<hsd::grid1(gr), dim3(BX, BY), 0, c->stream>>>(L, tr, kc, colour, gr.x, gr.y, gr.z, (T)c->prm.e_omega, skip);
            HS_TRY(prof_end(c, slot));
        }
    HS_CHECK(hipGetLastError());
    return HS_OK;
}

// one V-cycle on level 0: b holds the vector to precondition, x starts at 0
template <typename T>
int e_vcycle(hs_ctx* c, int tr) {
    EnergyState* E = c->en;
    const int* skip = &E->ctl->skip;
    if (c->prm.e_precond == HS_PRECOND_AMG) {
        // red-black sweeps on the finest level, the aggregated levels below it (hs_amg.hip)
        HS_TRY(e_smooth<T>(c, E->lv[0], tr, c->prm.amg_sweeps0, true, skip, true));
        HS_TRY(amg_coarse_e(c, E->lv[0], tr, skip));
        return e_smooth<T>(c, E->lv[0], tr, c->prm.amg_sweeps0, false, skip, true);
    }
    const int nl = (int)E->lv.size();
    int t = 0;
    while (t < nl && (long)E->lv[t]->nx * E->lv[t]->ny * E->lv[t]->nz > E_TAIL_MAX_CELLS) ++t;
    if (t >= nl) t = nl - 1;
    for (int l = 0; l < t; ++l) {
        HS_TRY(e_smooth<T>(c, E->lv[l], tr, c->prm.e_pre, true, skip, l == 0));
        const ELevel* m = E->lv[l + 1];
        const ELevel* f = E->lv[l];
        const long tiles = (long)((f->nx + BX - 1) / BX) * ((f->ny + BY - 1) / BY);
        const int kc = hsd::kchunk(m->nz, tiles);
        const dim3 gr((f->nx + BX - 1) / BX, (f->ny + BY - 1) / BY, (m->nz + kc - 1) / kc);
        k_e_restrict<T><<<hsd::grid1(gr), dim3(BX, BY), 0, c->stream>>>(el<T>(f), el<T>(m), tr, kc, gr.x, gr.y, gr.z,
                                                                        skip);
        HS_CHECK(hipGetLastError());
    }
    ELPack<T> P;
    for (int l = 0; l < nl && l < E_MAX_LEVELS; ++l) P.l[l] = el<T>(E->lv[l]);
    k_e_tail<T><<<1, E_TAIL_THREADS, 0, c->stream>>>(P, t, nl, tr, c->prm.e_pre, c->prm.e_post,
                                                     c->prm.mg_coarse_sweeps, skip);
    HS_CHECK(hipGetLastError());
    for (int l = t - 1; l >= 0; --l) {
        ELevel* f = E->lv[l];
        const long tiles = (long)((f->nx + BX - 1) / BX) * ((f->ny + BY - 1) / BY);
        const int kc = hsd::kchunk(f->nz, tiles);
        const dim3 gr = hsd::grid3(f->nx, f->ny, f->nz, kc);
        k_e_prolong<T><<<hsd::grid1(gr), dim3(BX, BY), 0, c->stream>>>(el<T>(f), el<T>(E->lv[l + 1]), kc, gr.x, gr.y,
                                                                       gr.z, skip);
        HS_CHECK(hipGetLastError());
        HS_TRY(e_smooth<T>(c, f, tr, c->prm.e_post, false, skip, l == 0));
    }
    return HS_OK;
}

int bi_finish(hs_ctx* c, int kind, int n) {
    k_bi_finish<<<1, 1024, 0, c->stream>>>(c->partials, n, kind, c->en->ctl);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}

struct SpmvShape {
    int kc;
    dim3 gr;
    int nb;
};
SpmvShape spmv_shape(const Dims& d) {
    const long tiles = (long)((d.nx + BX - 1) / BX) * ((d.ny + BY - 1) / BY);
    SpmvShape s;
    s.kc = hsd::kchunk(d.nz, tiles);
    s.gr = hsd::grid3(d.nx, d.ny, d.nz, s.kc);
    s.nb = (int)(s.gr.x * s.gr.y * s.gr.z);
    return s;
}

// BiCGStab on A (tr = 0) or A^T (tr = 1); b and the solution (E->x) in the P layout.
template <typename T>
int e_bicgstab(hs_ctx* c, const double* b, int tr, double tol, int maxit, hs_solve_info* info) {
    EnergyState* E = c->en;
    const Dims& d = c->d;
    const long n = d.nP;
    const SpmvShape sh = spmv_shape(d);
    if (2L * sh.nb > 65536 || 2L * FLAT_BLOCKS > 65536) return HS_E_ARG;
    BiCtl* h = E->h_ctl;
    h->eps = tol * tol;
    h->maxit = maxit;
    HS_CHECK(hipMemcpyAsync(&E->ctl->eps, &h->eps, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HS_CHECK(hipMemcpyAsync(&E->ctl->maxit, &h->maxit, sizeof(int), hipMemcpyHostToDevice, c->stream));
    if (c->prm.e_precond == HS_PRECOND_AMG) HS_TRY(amg_setup_e(c, E->lv[0], tr));
    T* b0 = (T*)E->lv[0]->b;
    // the level-0 iterate (the V-cycle's output) moves between two buffers with the fused
    // sweeps: read its current pointer at every use
    auto x0 = [&]() { return (T*)E->lv[0]->x; };
    const int G = FLAT_BLOCKS;
    const EOp op = eop(E);
    k_bi_init<<<G, FLAT, 0, c->stream>>>(n, b, E->x, E->r, E->rh, E->p, E->v, c->partials);
    HS_CHECK(hipGetLastError());
    HS_TRY(bi_finish(c, BI_INIT, G));
    int launched = 0;
    int batch = E->last_iters > 2 ? E->last_iters - 1 : 4;
    while (true) {
        for (int it = 0; it < batch && launched < maxit; ++it, ++launched) {
            k_bi_restart<<<G, FLAT, 0, c->stream>>>(n, E->r, E->rh, E->ctl);
            k_bi_p<T><<<G, FLAT, 0, c->stream>>>(n, E->r, E->p, E->v, b0, x0(), E->ctl);
            HS_TRY(e_vcycle<T>(c, tr));
            k_bi_spmv<T><<<hsd::grid1(sh.gr), dim3(BX, BY), 0, c->stream>>>(d, op, tr, x0(), E->ph, E->v, E->rh, 0, sh.kc,
                                                                            sh.gr.x, sh.gr.y, sh.gr.z, c->partials,
                                                                            E->ctl);
            HS_TRY(bi_finish(c, BI_ALPHA, sh.nb));
            k_bi_s<T><<<G, FLAT, 0, c->stream>>>(n, E->r, E->v, b0, x0(), c->partials, E->ctl);
            HS_TRY(bi_finish(c, BI_S, G));
            HS_TRY(e_vcycle<T>(c, tr));
            k_bi_spmv<T><<<hsd::grid1(sh.gr), dim3(BX, BY), 0, c->stream>>>(d, op, tr, x0(), E->sh, E->t, E->r, 1, sh.kc,
                                                                            sh.gr.x, sh.gr.y, sh.gr.z, c->partials,
                                                                            E->ctl);
            HS_TRY(bi_finish(c, BI_OMEGA, sh.nb));
            k_bi_update<<<G, FLAT, 0, c->stream>>>(n, E->x, E->r, E->ph, E->sh, E->t, E->rh, c->partials, E->ctl);
            HS_TRY(bi_finish(c, BI_UPDATE, G));
        }
        HS_CHECK(hipGetLastError());
        HS_CHECK(hipMemcpyAsync(h, E->ctl, sizeof(BiCtl), hipMemcpyDeviceToHost, c->stream));
        HS_CHECK(hipStreamSynchronize(c->stream));
        if (h->done || launched >= maxit) break;
        batch = 2;
    }
    E->last_iters = h->iters;
    // true residual in FP64
    k_e_resid<<<hsd::grid1(sh.gr), dim3(BX, BY), 0, c->stream>>>(d, op, tr, E->x, b, sh.kc, sh.gr.x, sh.gr.y, sh.gr.z,
                                                                 c->partials);
    HS_CHECK(hipGetLastError());
    HS_TRY(bi_finish(c, BI_DOT, sh.nb));
    double red[2];
    HS_CHECK(hipMemcpyAsync(red, E->ctl->red, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HS_CHECK(hipStreamSynchronize(c->stream));
    if (info) {
        info->iters = h->iters;
        info->converged = h->converged;
        info->breakdown = h->breakdown;
        info->restarts = h->restarts;
        info->relres = h->bb > 0 ? sqrt(h->rr / h->bb) : 0.0;
        info->true_relres = red[1] > 0 ? sqrt(red[0] / red[1]) : 0.0;
    }
    if (!std::isfinite(h->rr) || !std::isfinite(red[0])) return HS_E_NAN;
    if (!h->converged && c->prm.fail_on_noconv) return HS_E_NOCONV;
    return HS_OK;
}

int e_build_any(hs_ctx* c) {
    HS_TRY(e_alloc(c));
    return c->en->f32 ? e_build<float>(c) : e_build<double>(c);
}

int e_solve_any(hs_ctx* c, const double* b, int tr, double tol, int maxit, hs_solve_info* info) {
    return c->en->f32 ? e_bicgstab<float>(c, b, tr, tol, maxit, info) : e_bicgstab<double>(c, b, tr, tol, maxit, info);
}

}  // namespace

int energy_solve(hs_ctx* c, double tol, int maxit, hs_solve_info* info) {
    if (c->prm.mode != HS_MODE_CONFIG) return HS_E_ARG;
    if (!(tol > 0.0) || maxit < 1) return HS_E_ARG;
    HS_TRY(e_build_any(c));
    EnergyState* E = c->en;
    HS_TRY(e_solve_any(c, E->rhs, 0, tol, maxit, info));
    const Dims& d = c->d;
    // interior temperatures <- solution (Dirichlet ghosts untouched), then the outflow ghost
    k_e_gather<<<cells_grid(d.nx, d.ny, d.nz), 64, 0, c->stream>>>(d, E->x, E->stage);
    HS_CHECK(hipGetLastError());
    k_e_scatter<<<cells_grid(d.nx, d.ny, d.nz), 64, 0, c->stream>>>(d, E->stage, c->tf);
    HS_CHECK(hipGetLastError());
    HS_TRY(bc_cfg_outflow(c, 3));
    HS_CHECK(hipStreamSynchronize(c->stream));
    return HS_OK;
}

// A x = b (transpose = 0) or A^T x = b (transpose = 1) for a caller right-hand side;
// b and x are C-layout device vectors.  The solution stays in E->x (P layout).
int energy_solve_rhs(hs_ctx* c, int transpose, const double* b, double* x, double tol, int maxit,
                     hs_solve_info* info) {
    if (c->prm.mode != HS_MODE_CONFIG) return HS_E_ARG;
    if (!(tol > 0.0) || maxit < 1 || !b) return HS_E_ARG;
    HS_TRY(e_build_any(c));
    EnergyState* E = c->en;
    const Dims& d = c->d;
    // right-hand side in the P layout (ghosts 0); E->rhs is rebuilt by every e_build
    HS_TRY(launch_fill(c, E->rhs, d.nP, 0.0));
    k_e_scatter<<<cells_grid(d.nx, d.ny, d.nz), 64, 0, c->stream>>>(d, b, E->rhs);
    HS_CHECK(hipGetLastError());
    HS_TRY(e_solve_any(c, E->rhs, transpose ? 1 : 0, tol, maxit, info));
    if (x) {
        k_e_gather<<<cells_grid(d.nx, d.ny, d.nz), 64, 0, c->stream>>>(d, E->x, x);
        HS_CHECK(hipGetLastError());
    }
    HS_CHECK(hipStreamSynchronize(c->stream));
    return HS_OK;
}

int energy_adjoint(hs_ctx* c, const double* dJdT, double* dJdq, double* dJdk, double tol, int maxit,
                   hs_solve_info* info) {
    if (!dJdT) return HS_E_ARG;
    HS_TRY(energy_solve_rhs(c, 1, dJdT, nullptr, tol, maxit, info));
    EnergyState* E = c->en;
    const Dims& d = c->d;
    k_e_grad<<<cells_grid(d.nx, d.ny, d.nz), 64, 0, c->stream>>>(d, c->g, c->k, c->tf, E->x, c->prm.outflow, dJdq, dJdk);
    HS_CHECK(hipGetLastError());
    HS_CHECK(hipStreamSynchronize(c->stream));
    return HS_OK;
}

int energy_apply(hs_ctx* c, int transpose, const double* x, double* y) {
    if (c->prm.mode != HS_MODE_CONFIG) return HS_E_ARG;
    HS_TRY(e_build_any(c));
    EnergyState* E = c->en;
    const Dims& d = c->d;
    // x (C layout) -> padded work vector with zero ghosts; y = A x on interior cells
    HS_TRY(launch_fill(c, E->p, d.nP, 0.0));
    k_e_scatter<<<cells_grid(d.nx, d.ny, d.nz), 64, 0, c->stream>>>(d, x, E->p);
    HS_CHECK(hipGetLastError());
    const SpmvShape sh = spmv_shape(d);
    k_bi_spmv<double><<<hsd::grid1(sh.gr), dim3(BX, BY), 0, c->stream>>>(d, eop(E), transpose ? 1 : 0, E->p, nullptr,
                                                                         E->v, nullptr, 0, sh.kc, sh.gr.x, sh.gr.y,
                                                                         sh.gr.z, nullptr, nullptr);
    HS_CHECK(hipGetLastError());
    k_e_gather<<<cells_grid(d.nx, d.ny, d.nz), 64, 0, c->stream>>>(d, E->v, y);
    HS_CHECK(hipGetLastError());
    HS_CHECK(hipStreamSynchronize(c->stream));
    return HS_OK;
}

void energy_free(hs_ctx* c) {
    EnergyState* E = c->en;
    if (!E) return;
    amg_free(&E->amg);
    for (ELevel* m : E->lv) {
        if (m->own_faces) {
            void* f[6] = {m->dx, m->dy, m->dz, m->fx, m->fy, m->fz};
            for (void* a : f) if (a) hipFree(a);
        }
        if (m->x) hipFree(m->x);
        if (m->x2) hipFree(m->x2);
        if (m->b) hipFree(m->b);
        delete m;
    }
    double* v[] = {E->dx, E->dy, E->dz, E->fx, E->fy, E->fz, E->rhs, E->x, E->r, E->rh, E->p, E->v, E->t, E->ph, E->sh,
                   E->stage, E->stage2, E->stage3};
    for (double* a : v) if (a) hipFree(a);
    if (E->ctl) hipFree(E->ctl);
    if (E->h_ctl) hipHostFree(E->h_ctl);
    delete E;
    c->en = nullptr;
}

int energy_stage(hs_ctx* c, double** s1, double** s2, double** s3) {
    HS_TRY(e_alloc(c));
    if (s1) *s1 = c->en->stage;
    if (s2) *s2 = c->en->stage2;
    if (s3) *s3 = c->en->stage3;
    return HS_OK;
}

int cells_from_padded(hs_ctx* c, const double* P, double* C) {
    const Dims& d = c->d;
    k_e_gather<<<cells_grid(d.nx, d.ny, d.nz), 64, 0, c->stream>>>(d, P, C);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}

int cells_to_padded(hs_ctx* c, const double* C, double* P) {
    const Dims& d = c->d;
    k_e_scatter<<<cells_grid(d.nx, d.ny, d.nz), 64, 0, c->stream>>>(d, C, P);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}
