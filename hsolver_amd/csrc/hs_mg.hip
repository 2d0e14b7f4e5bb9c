// hs_mg.hip — geometric multigrid preconditioner and flexible CG for the p' operator.
//
// Operator on every level (cell-centred, padded layout (n+2)^3 of that level):
//   (A x)_P = D_P x_P - cw_P x_W - cw_E x_E - cs_P x_S - cs_N x_N - cb_P x_B - cb_T x_T,
//   D_P = g_P + cw_P + cw_E + cs_P + cs_N + cb_P + cb_T,
// with cw/cs/cb the coupling across the west/south/bottom face of P (zero unless both
// cells are active) and g the "ground" conductance of faces that end on a Dirichlet
// p' = 0 ghost (walls, solids, domain boundary: the reference's convention,
// fluid_solver.h:1996).  Inactive cells have D = 0 and x = 0.
//
// Coarsening (2x2x2 agglomeration): the coupling across a coarse face is the sum over
// its four fine sub-faces of the series conductance from coarse centre to coarse centre,
// 1 / (1/(2 c_a) + 1/c_b + 1/(2 c_c)), where c_b is the fine coupling across the coarse
// face and c_a, c_c the fine couplings inside the two coarse cells (harmonic averaging
// across coefficient jumps, Alcouffe-Brandt-Dendy-Painter 1981); grounds add with the
// factor that keeps the Dirichlet point where the finest level puts it (see
// build_hierarchy).  On uniform coefficients both give exactly the rediscretised
// operator on the coarse grid, boundary rows included.  Restriction sums the children's residuals, prolongation
// is trilinear (see prolong_value); smoothing is red-black Gauss-Seidel (symmetric
// order: red, black before the coarse correction, black, red after).
//
// The outer solver is flexible CG (Notay 2000, one-term truncation: beta =
// <z_new, r_new - r_old> / <z_old, r_old>) in FP64 with the stop rule of the CG path;
// the V-cycle runs in float (HS_PREC_MIXED) or double.
#include <cfloat>
#include <cmath>

#include "hs_device.h"
#include "hs_mglevel.h"

using hsd::BT;
using hsd::BX;
using hsd::BY;
using hsd::FLAT;
using hsd::FLAT_BLOCKS;

namespace {

using mgl::Lv;
using mgl::lv;
using mgl::diag_of;
using mgl::offsum;

using hsd::sor;


// fine level from the Mode C face coefficients (rho A^2/ap on U/V/W faces) + fluid mask,
// and the level's active mask: the cell's east / north / top coefficients are the values
// the neighbouring threads store (same faces, same cells), so D > 0 is decided here with
// diag_of's operands in its order
template <typename T>
__global__ void k_mg_fine(Dims d, const uint8_t* cls, const double* ax, const double* ay, const double* az, Lv<T> L,
                          uint8_t* act) {
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y, k = 1 + blockIdx.z;
    if (i > d.nx) return;
    auto fl = [&](long ii, long jj, long kk) { return cls[idxP(d, ii, jj, kk)] == CELL_FLUID; };
    const long c = idxP(d, i, j, k);
    const bool f = fl(i, j, k);
    const double w = ax[idxU(d, i, j, k)], e = ax[idxU(d, i + 1, j, k)];
    const double s = ay[idxV(d, i, j, k)], n = ay[idxV(d, i, j + 1, k)];
    const double b = az[idxW(d, i, j, k)], t = az[idxW(d, i, j, k + 1)];
    const T cw = (T)((f && fl(i - 1, j, k)) ? w : 0.0);
    const T cs = (T)((f && fl(i, j - 1, k)) ? s : 0.0);
    const T cb = (T)((f && fl(i, j, k - 1)) ? b : 0.0);
    L.cw[c] = cw;
    L.cs[c] = cs;
    L.cb[c] = cb;
    double g = 0.0;
    if (f) {
        if (!fl(i - 1, j, k)) g += w;
        if (!fl(i + 1, j, k)) g += e;
        if (!fl(i, j - 1, k)) g += s;
        if (!fl(i, j + 1, k)) g += n;
        if (!fl(i, j, k - 1)) g += b;
        if (!fl(i, j, k + 1)) g += t;
    }
    const T gT = (T)g;
    L.g[c] = gT;
    // the neighbours' west / south / bottom coefficients (their own f and the shared face)
    const T ce = (T)((fl(i + 1, j, k) && f) ? e : 0.0);
    const T cn = (T)((fl(i, j + 1, k) && f) ? n : 0.0);
    const T ct = (T)((fl(i, j, k + 1) && f) ? t : 0.0);
    act[c] = (gT + cw + ce + cs + cn + cb + ct) > T(0) ? 1 : 0;
}

template <typename T>
__device__ __forceinline__ T series(T ca, T cb, T cc) {
    if (!(cb > T(0))) return T(0);
    if (!(ca > T(0))) ca = cb;
    if (!(cc > T(0))) cc = cb;
    return T(1) / (T(0.5) / ca + T(1) / cb + T(0.5) / cc);
}

template <typename T>
__global__ void k_mg_coarsen(Lv<T> F, Lv<T> C, T gscale) {
    const int I = 1 + blockIdx.x * blockDim.x + threadIdx.x, J = 1 + blockIdx.y, K = 1 + blockIdx.z;
    if (I > C.nx) return;
    const long cc = I + C.sy * J + C.sz * K;
    auto fi = [&](long i, long j, long k) { return i + F.sy * j + F.sz * k; };
    const int i0 = 2 * I - 1, j0 = 2 * J - 1, k0 = 2 * K - 1;
    const int ni = (2 * I <= F.nx) ? 2 : 1, nj = (2 * J <= F.ny) ? 2 : 1, nk = (2 * K <= F.nz) ? 2 : 1;
    T sw = 0, ss = 0, sb = 0, g = 0;
    for (int dk = 0; dk < nk; ++dk)
        for (int dj = 0; dj < nj; ++dj) {
            if (I >= 2) {
                const long f = fi(i0, j0 + dj, k0 + dk);
                const T cc2 = (ni == 2) ? F.cw[f + 1] : T(0);
                sw += series<T>(F.cw[f - 1], F.cw[f], cc2);
            }
        }
    for (int dk = 0; dk < nk; ++dk)
        for (int di = 0; di < ni; ++di) {
            if (J >= 2) {
                const long f = fi(i0 + di, j0, k0 + dk);
                const T cc2 = (nj == 2) ? F.cs[f + F.sy] : T(0);
                ss += series<T>(F.cs[f - F.sy], F.cs[f], cc2);
            }
        }
    for (int dj = 0; dj < nj; ++dj)
        for (int di = 0; di < ni; ++di) {
            if (K >= 2) {
                const long f = fi(i0 + di, j0 + dj, k0);
                const T cc2 = (nk == 2) ? F.cb[f + F.sz] : T(0);
                sb += series<T>(F.cb[f - F.sz], F.cb[f], cc2);
            }
        }
    for (int dk = 0; dk < nk; ++dk)
        for (int dj = 0; dj < nj; ++dj)
            for (int di = 0; di < ni; ++di) g += F.g[fi(i0 + di, j0 + dj, k0 + dk)];
    C.cw[cc] = sw;
    C.cs[cc] = ss;
    C.cb[cc] = sb;
    C.g[cc] = gscale * g;
}

// one colour of red-black Gauss-Seidel
template <typename T>
__global__ __launch_bounds__(BT) void k_mg_smooth(Lv<T> L, int kc, int colour, int gx, int gy, int gz, T omega,
                                                  const SolveCtl* ctl) {
    if (ctl->done) return;
    const hsd::Tile tl = hsd::tile_of(gx, gy, gz);
    const int i = 1 + tl.bx * BX + threadIdx.x, j = 1 + tl.by * BY + threadIdx.y;
    if (i > L.nx || j > L.ny) return;
    const int k0 = 1 + tl.bz * kc, k1 = min(L.nz, k0 + kc - 1);
    for (int k = k0; k <= k1; ++k) {
        if (((i + j + k) & 1) != colour) continue;
        const long c = i + L.sy * j + L.sz * k;
        const T D = diag_of(L, c);
        if (!(D > T(0))) continue;
        L.x[c] = sor(L.x[c], (L.b[c] + offsum(L, c)) / D, omega);
    }
}

// One full red-black Gauss-Seidel sweep, out of place (xo -> xn), in one pass over the
// level.  The first colour c1 is updated from xo, the second colour from the new first
// colour values: exactly the two in-place colour passes of k_mg_smooth (same operations
// in the same order, so the results are bit-equal).
//
// A workgroup owns a TX x TY column tile (one thread per column) and marches z through
// its chunk of planes.  At plane step k every thread loads its own column's operator of
// plane k+1 (cw, cw east, cs, cs north, g, b; cb and the iterate of plane k+2) whatever
// its colour, relaxes the first colour of plane k+1 if its cell has that colour, and
// keeps the values in registers until step k+1 relaxes the second colour there: each
// cache line of the level is fetched once per sweep, and only in-plane neighbour reads
// (east / north faces, the iterate's four in-plane neighbours) revisit lines, within the
// same step.  The first colour is also relaxed on the one-cell ring around the tile (row
// j0-1 by tile row 0, row j0+TY by tile row TY-1, the two side cells of each row by x
// threads 0 and 1), from xo, so tiles never read each other's output; the second colour
// of plane k reads the first-colour planes k-1, k, k+1 from LDS.  Four LDS slots and one
// barrier per step: step k writes slot k+1 and reads slots k-1..k+1, so the slot that
// step k+1 writes is idle, and a wave can only write slot k+3 = k-1 after every wave has
// passed step k+1's barrier.
//
// Addresses are byte offsets of type O on the arrays' scalar base addresses (32-bit when
// the level's arrays are under 2 GiB).  EAGER issues the neighbour and ring loads before
// any of them is used (one wait on memory per plane step); otherwise they are read where
// needed, which holds fewer registers (the FP64 form).
template <typename T, typename O>
__device__ __forceinline__ T ldb(const T* base, O off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}

// the old iterate at byte offset off; ZX: the iterate is zero everywhere (first sweep
// from a zero guess), read as the constant and still multiplied, so the arithmetic (and
// every result bit) is that of a sweep over a zero-filled buffer
template <typename T, typename O, bool ZX>
__device__ __forceinline__ T ldx(const T* xo, O off) {
    return ZX ? T(0) : ldb(xo, off);
}

using hsd::issued_together;

template <typename T, typename O, bool EAGER, bool ZX>
__device__ __forceinline__ T relax_at(const Lv<T>& L, const T* xo, O c, O dy, O dz, T omega) {
    constexpr O E = sizeof(T);
    const T g = ldb(L.g, c), cw = ldb(L.cw, c), ce = ldb(L.cw, c + E), cs = ldb(L.cs, c), cn = ldb(L.cs, c + dy),
            cb = ldb(L.cb, c), ct = ldb(L.cb, c + dz), b = ldb(L.b, c);
    const T D = g + cw + ce + cs + cn + cb + ct;
    if (EAGER || ZX) {
        const T x = ldx<T, O, ZX>(xo, c), xw = ldx<T, O, ZX>(xo, c - E), xe = ldx<T, O, ZX>(xo, c + E),
                xs = ldx<T, O, ZX>(xo, c - dy), xn = ldx<T, O, ZX>(xo, c + dy), xb = ldx<T, O, ZX>(xo, c - dz),
                xt = ldx<T, O, ZX>(xo, c + dz);
        // one wait on memory: the iterate loads stay beside the operator loads instead of
        // sinking under the D > 0 test into a second round trip
        if (!ZX) issued_together(x, xw, xe, xs, xn, xb, xt);
        const T off = cw * xw + ce * xe + cs * xs + cn * xn + cb * xb + ct * xt;
        const T rel = (b + off) / D;
        return (D > T(0)) ? sor(x, rel, omega) : x;
    }
    const T x = ldb(xo, c);
    if (!(D > T(0))) return x;
    const T off = cw * ldb(xo, c - E) + ce * ldb(xo, c + E) + cs * ldb(xo, c - dy) + cn * ldb(xo, c + dy) +
                  cb * ldb(xo, c - dz) + ct * ldb(xo, c + dz);
    return sor(x, (b + off) / D, omega);
}

// Float sweeps hold 8 waves per SIMD (at most 64 VGPRs): two 1024-thread workgroups per CU.
template <typename T, int TX, int TY, typename O, bool EAGER, bool ZX>
__global__ __launch_bounds__(TX * TY, sizeof(T) == sizeof(float) ? 8 : 1) void k_mg_rbsweep(
    Lv<T> L, const T* __restrict__ xo, T* __restrict__ xn, int c1, int kc, int gx, int gy, int gz, T omega,
    const SolveCtl* ctl) {
    constexpr int RX = TX + 2, RY = TY + 2;
    constexpr O E = sizeof(T);
    if (ctl->done) return;
    __shared__ T X1[4][RY * RX];
    const hsd::Tile tl = hsd::tile_of(gx, gy, gz);
    const int i0 = 1 + tl.bx * TX, j0 = 1 + tl.by * TY;
    const int k0 = 1 + tl.bz * kc, k1 = min(L.nz, k0 + kc - 1);
    const int lane = threadIdx.x, w = threadIdx.y;
    const int i = i0 + lane, j = j0 + w;
    const bool inside = i <= L.nx && j <= L.ny;
    const int r = (w + 1) * RX + lane + 1;
    const O dy = (O)L.sy * E, dz = (O)L.sz * E;
    const O col = (O)(i + L.sy * j) * E;   // own column, plane 0 (read only when inside)
    const int rowring = (w == 0) ? j0 - 1 : (w == TY - 1 ? j0 + TY : -1);
    const int rowslot = (w == 0) ? lane + 1 : (TY + 1) * RX + lane + 1;
    const bool rowring_in = rowring >= 1 && rowring <= L.ny && i <= L.nx;
    const int si = lane == 0 ? i0 - 1 : i0 + TX;
    const int sslot = (w + 1) * RX + (lane == 0 ? 0 : TX + 1);
    const bool side_in = lane < 2 && j <= L.ny && si >= 1 && si <= L.nx;

    T cw0 = 0, ce0 = 0, cs0 = 0, cn0 = 0, g0 = 0, b0 = 0;   // plane k
    T cw1 = 0, ce1 = 0, cs1 = 0, cn1 = 0, g1 = 0, b1 = 0;   // plane k+1
    T cbk = 0, cbk1 = 0, cbk2 = 0;                          // cb of planes k, k+1, k+2
    T xk = 0, xk1 = 0, xk2 = 0;                             // iterate of planes k, k+1, k+2
    T xw = 0, xe = 0, xs = 0, xnn = 0;                      // iterate's in-plane neighbours, plane k+1
    T own1 = 0, own1n = 0;                                  // first-colour value of the own cell, k / k+1
    const int kfirst = k0 - 2;
    if (inside) {
        if (kfirst >= 0) xk = ldx<T, O, ZX>(xo, col + (O)kfirst * dz);
        xk1 = ldx<T, O, ZX>(xo, col + (O)(k0 - 1) * dz);
        cbk1 = ldb(L.cb, col + (O)(k0 - 1) * dz);
    }
    for (int k = kfirst; k <= k1; ++k) {
        const int kp = k + 1;   // in [0, nz + 1]: planes of the padded level
        if (inside) {
            const O c = col + (O)kp * dz;
            cw1 = ldb(L.cw, c); ce1 = ldb(L.cw, c + E); cs1 = ldb(L.cs, c); cn1 = ldb(L.cs, c + dy);
            g1 = ldb(L.g, c); b1 = ldb(L.b, c);
            if (EAGER || ZX) {
                xw = ldx<T, O, ZX>(xo, c - E); xe = ldx<T, O, ZX>(xo, c + E);
                xs = ldx<T, O, ZX>(xo, c - dy); xnn = ldx<T, O, ZX>(xo, c + dy);
            }
            if (kp + 1 <= L.nz + 1) {
                cbk2 = ldb(L.cb, c + dz);
                xk2 = ldx<T, O, ZX>(xo, c + dz);
            }
        }
        const bool kin = kp >= 1 && kp <= L.nz;
        // first colour of plane k+1 on the ring, then on the own cell
        if (rowring >= 0) {
            T v = T(0);
            if (rowring_in && kin && ((i + rowring + kp) & 1) == c1)
                v = relax_at<T, O, EAGER, ZX>(L, xo, (O)(i + L.sy * rowring) * E + (O)kp * dz, dy, dz, omega);
            X1[kp & 3][rowslot] = v;
        }
        if (lane < 2) {
            T v = T(0);
            if (side_in && kin && ((si + j + kp) & 1) == c1)
                v = relax_at<T, O, EAGER, ZX>(L, xo, (O)(si + L.sy * j) * E + (O)kp * dz, dy, dz, omega);
            X1[kp & 3][sslot] = v;
        }
        {
            const bool on = inside && kin && ((i + j + kp) & 1) == c1;
            if (!EAGER && !ZX && on) {
                const O c = col + (O)kp * dz;
                xw = ldb(xo, c - E); xe = ldb(xo, c + E); xs = ldb(xo, c - dy); xnn = ldb(xo, c + dy);
            }
            const T D = g1 + cw1 + ce1 + cs1 + cn1 + cbk1 + cbk2;
            const T off = cw1 * xw + ce1 * xe + cs1 * xs + cn1 * xnn + cbk1 * xk + cbk2 * xk2;
            const T rel = (b1 + off) / D;
            const T v = on ? ((D > T(0)) ? sor(xk1, rel, omega) : xk1) : T(0);
            X1[kp & 3][r] = v;
            own1n = v;
        }
        // second colour of plane k
        if (k >= k0) {
            __syncthreads();
            if (inside) {
                const T* P0 = X1[k & 3];
                const T* Pm = X1[(k - 1) & 3];
                const T* Pp = X1[kp & 3];
                const T D = g0 + cw0 + ce0 + cs0 + cn0 + cbk + cbk1;
                const T off = cw0 * P0[r - 1] + ce0 * P0[r + 1] + cs0 * P0[r - RX] + cn0 * P0[r + RX] + cbk * Pm[r] +
                              cbk1 * Pp[r];
                const T rel = (b0 + off) / D;
                const T out = (((i + j + k) & 1) == c1) ? own1 : ((D > T(0)) ? sor(xk, rel, omega) : xk);
                *reinterpret_cast<T*>(reinterpret_cast<char*>(xn) + col + (O)k * dz) = out;
            }
        }
        // plane k+1 becomes plane k
        cw0 = cw1; ce0 = ce1; cs0 = cs1; cn0 = cn1; g0 = g1; b0 = b1;
        cbk = cbk1; cbk1 = cbk2;
        xk = xk1; xk1 = xk2;
        own1 = own1n;
    }
}

template <typename T, int TX, int TY, bool ZX>
int launch_rbsweep_z(hs_ctx* c, const Lv<T>& L, const T* src, T* dst, int c1, T omega, bool prof) {
    constexpr bool EAGER = sizeof(T) == sizeof(float);
    const int gx = (L.nx + TX - 1) / TX, gy = (L.ny + TY - 1) / TY;
    // deep z chunks (a chunk reads two planes beyond its own): about 512K threads in
    // flight, i.e. 2 workgroups of 1024 threads per CU
    long want = (1024L * 512 / (TX * TY)) / ((long)gx * gy);
    if (want < 1) want = 1;
    int kc = (int)((L.nz + want - 1) / want);
    if (kc < 1) kc = 1;
    const int gz = (L.nz + kc - 1) / kc;
    const bool small = !c->prm.mg_wide_offsets && (double)L.sz * (L.nz + 2) * sizeof(T) < 2147483648.0;
    // the profiled kernel is the full sweep (28 B per float cell); the zero-guess first
    // sweep of a cycle visit reads no iterate and is left out
    const int slot = (prof && !ZX) ? prof_begin(c, HS_K_MG_SMOOTH0) : -1;
    if (small)
        k_mg_rbsweep<T, TX, TY, uint32_t, EAGER, ZX>
            <<<dim3(gx * gy * gz), dim3(TX, TY), 0, c->stream>>>(L, src, dst, c1, kc, gx, gy, gz, omega, c->sctl);
    else
        k_mg_rbsweep<T, TX, TY, uint64_t, EAGER, ZX>
            <<<dim3(gx * gy * gz), dim3(TX, TY), 0, c->stream>>>(L, src, dst, c1, kc, gx, gy, gz, omega, c->sctl);
    HS_TRY(prof_end(c, slot));
    return HS_OK;
}

// zx: the iterate is zero (first sweep from a zero guess; the buffer is not read)
template <typename T, int TX, int TY>
int launch_rbsweep(hs_ctx* c, const Lv<T>& L, const T* src, T* dst, int c1, T omega, bool prof, bool zx) {
    return zx ? launch_rbsweep_z<T, TX, TY, true>(c, L, src, dst, c1, omega, prof)
              : launch_rbsweep_z<T, TX, TY, false>(c, L, src, dst, c1, omega, prof);
}

// b_c = sum over active children of (b - A x); x_c = 0.
// One thread per fine column of a 64 x 4 fine tile (coarse 32 x 2): each thread forms
// the residuals of its two fine cells of coarse plane K, and the 2 x 2 in-plane sums go
// through LDS.
template <typename T>
__global__ __launch_bounds__(BT) void k_mg_restrict(Lv<T> F, Lv<T> C, int kc, int gx, int gy, int gz, int zero_x,
                                                   const SolveCtl* ctl) {
    if (ctl->done) return;
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
                const T D = diag_of(F, f);
                if (D > T(0)) r += F.b[f] - (D * F.x[f] - offsum(F, f));
            }
        }
        red[threadIdx.y][threadIdx.x] = r;
        __syncthreads();
        if (owner) {
            const T s = (red[2 * cj][2 * ci] + red[2 * cj][2 * ci + 1]) + (red[2 * cj + 1][2 * ci] + red[2 * cj + 1][2 * ci + 1]);
            const long cc = I + C.sy * J + C.sz * K;
            C.b[cc] = s;
            if (zero_x) C.x[cc] = T(0);
        }
        __syncthreads();
    }
}

// The same restriction with the fine column streamed through registers: a thread walks
// its fine column over the chunk's fine planes keeping the iterate of planes k-1, k, k+1
// and cb of planes k, k+1, so every array line is read once; only in-plane neighbour
// reads revisit lines, within the same plane.  TX x TY fine columns per workgroup (TY
// even), the 2 x 2 in-plane sums of each coarse plane through LDS; the arithmetic per
// fine cell is diag_of / offsum's, in their order.
template <typename T, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void k_mg_restrict_s(Lv<T> F, Lv<T> C, int kc, int gx, int gy, int gz,
                                                           int zero_x, const SolveCtl* ctl) {
    if (ctl->done) return;
    __shared__ T red[TY][TX];
    const hsd::Tile tl = hsd::tile_of(gx, gy, gz);
    const int i = 1 + tl.bx * TX + threadIdx.x, j = 1 + tl.by * TY + threadIdx.y;
    const bool inside = i <= F.nx && j <= F.ny;
    const int t = threadIdx.x + TX * threadIdx.y;
    const int ci = t % (TX / 2), cj = t / (TX / 2);
    const int I = 1 + tl.bx * (TX / 2) + ci, J = 1 + tl.by * (TY / 2) + cj;
    const bool owner = t < (TX / 2) * (TY / 2) && I <= C.nx && J <= C.ny;
    const int K0 = 1 + tl.bz * kc, K1 = min(C.nz, K0 + kc - 1);
    const long sy = F.sy, sz = F.sz;
    const long col = i + sy * j;
    T xm = 0, x0 = 0, cb0 = 0;
    if (inside) {
        const long c = col + sz * (2 * K0 - 1);
        xm = F.x[c - sz];
        x0 = F.x[c];
        cb0 = F.cb[c];
    }
    for (int K = K0; K <= K1; ++K) {
        T r = 0;
        for (int dk = 0; dk < 2; ++dk) {
            const int k = 2 * K - 1 + dk;
            if (k > F.nz) break;
            const long f = col + sz * k;
            T xp = 0, cb1 = 0;
            if (inside) {
                xp = F.x[f + sz];
                cb1 = F.cb[f + sz];
                const T cw = F.cw[f], ce = F.cw[f + 1], cs = F.cs[f], cn = F.cs[f + sy];
                const T D = F.g[f] + cw + ce + cs + cn + cb0 + cb1;
                const T off = cw * F.x[f - 1] + ce * F.x[f + 1] + cs * F.x[f - sy] + cn * F.x[f + sy] + cb0 * xm + cb1 * xp;
                if (D > T(0)) r += F.b[f] - (D * x0 - off);
            }
            xm = x0;
            x0 = xp;
            cb0 = cb1;
        }
        red[threadIdx.y][threadIdx.x] = r;
        __syncthreads();
        if (owner) {
            const T s = (red[2 * cj][2 * ci] + red[2 * cj][2 * ci + 1]) + (red[2 * cj + 1][2 * ci] + red[2 * cj + 1][2 * ci + 1]);
            const long cc = I + C.sy * J + C.sz * K;
            C.b[cc] = s;
            if (zero_x) C.x[cc] = T(0);
        }
        __syncthreads();
    }
}

// Trilinear interpolation of the coarse correction to fine cell (i,j,k) (1-based): the
// fine centre sits a quarter coarse cell from its parent's centre, so each axis weighs
// the parent 3/4 and the coarse neighbour on the fine cell's side 1/4 (27/64, 9/64,
// 3/64, 1/64 in 3-D).  Ghost and inactive coarse cells hold 0, which is the p' = 0
// Dirichlet value of walls and solids.  With summed (constant) restriction the
// transfer orders satisfy m_P + m_R > 2m for this second-order operator.
template <typename T>
__device__ __forceinline__ T prolong_value(const Lv<T>& C, int i, int j, int k, const T* __restrict__ x) {
    const int I = (i + 1) >> 1, J = (j + 1) >> 1, K = (k + 1) >> 1;
    const int di = (i & 1) ? -1 : 1, dj = (j & 1) ? -1 : 1, dk = (k & 1) ? -1 : 1;
    const long c0 = I + C.sy * J + C.sz * K;
    const long ox = di, oy = dj * C.sy, oz = dk * C.sz;
    const T a = T(0.75), b = T(0.25);
    const T zlo = a * (a * x[c0] + b * x[c0 + ox]) + b * (a * x[c0 + oy] + b * x[c0 + ox + oy]);
    const T zhi = a * (a * x[c0 + oz] + b * x[c0 + ox + oz]) + b * (a * x[c0 + oy + oz] + b * x[c0 + ox + oy + oz]);
    return a * zlo + b * zhi;
}

// x_f += P x_c on active fine cells (64 x 4 fine columns per workgroup, z chunks).  The
// planes are independent: four at a time, with the mask, the fine iterate and the eight
// coarse values of all four loaded before any store (whatever the mask says; a group
// past the chunk's last plane loads that plane again and does not store it), so four
// planes cost one memory round trip.
template <typename T>
__global__ __launch_bounds__(BT) void k_mg_prolong(Lv<T> F, Lv<T> C, int kc, int gx, int gy, int gz,
                                                  const SolveCtl* ctl) {
    constexpr int U = 4;
    if (ctl->done) return;
    const hsd::Tile tl = hsd::tile_of(gx, gy, gz);
    const int i = 1 + tl.bx * BX + threadIdx.x, j = 1 + tl.by * BY + threadIdx.y;
    if (i > F.nx || j > F.ny) return;
    const int k0 = 1 + tl.bz * kc, k1 = min(F.nz, k0 + kc - 1);
    T* __restrict__ xf = F.x;
    const T* __restrict__ xc = C.x;
    const uint8_t* __restrict__ act = F.act;
    const long col = i + F.sy * j;
    for (int k = k0; k <= k1; k += U) {
        int on[U];
        T xv[U], pv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = min(k + u, k1);
            const long f = col + F.sz * kk;
            on[u] = act[f];
            xv[u] = xf[f];
            pv[u] = prolong_value(C, i, j, kk, xc);
        }
        hsd::issued_together(on[0], on[1], on[2], on[3], xv[0], xv[1], xv[2], xv[3], pv[0], pv[1], pv[2], pv[3]);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (on[u] && k + u <= k1) xf[col + F.sz * (k + u)] = xv[u] + pv[u];
    }
}

// launch shapes of the tiled transfer kernels over a level of nx x ny x nz cells
struct TileShape {
    int kc;
    dim3 gr;
};
inline TileShape tile_shape(int nx, int ny, int nz) {
    const long tiles = (long)((nx + BX - 1) / BX) * ((ny + BY - 1) / BY);
    TileShape s;
    s.kc = hsd::kchunk(nz, tiles);
    s.gr = hsd::grid3(nx, ny, nz, s.kc);
    return s;
}

// zero_x: also clear the coarse iterate (not needed when its first sweep reads it as zero)
template <typename T, int TX, int TY>
int launch_restrict_s(hs_ctx* c, const Lv<T>& F, const Lv<T>& C, bool zero_x) {
    const int gx = (F.nx + TX - 1) / TX, gy = (F.ny + TY - 1) / TY;
    // about 512K threads in flight, z chunks of coarse planes
    long want = (1024L * 512 / (TX * TY)) / ((long)gx * gy);
    if (want < 1) want = 1;
    int kc = (int)((C.nz + want - 1) / want);
    if (kc < 1) kc = 1;
    const int gz = (C.nz + kc - 1) / kc;
    k_mg_restrict_s<T, TX, TY><<<gx * gy * gz, dim3(TX, TY), 0, c->stream>>>(F, C, kc, gx, gy, gz, zero_x ? 1 : 0,
                                                                          c->sctl);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}

template <typename T>
int launch_restrict(hs_ctx* c, const Lv<T>& F, const Lv<T>& C, bool zero_x) {
    // fine-cell tiles in x and y, coarse planes in z
    const long tiles = (long)((F.nx + BX - 1) / BX) * ((F.ny + BY - 1) / BY);
    TileShape s;
    s.kc = hsd::kchunk(C.nz, tiles);
    s.gr = dim3((F.nx + BX - 1) / BX, (F.ny + BY - 1) / BY, (C.nz + s.kc - 1) / s.kc);
    // wide levels: the register-streamed form (256 x 4 or 128 x 8 fine columns)
    if (F.nx > 128) return launch_restrict_s<T, 256, 4>(c, F, C, zero_x);
    if (F.nx > 64) return launch_restrict_s<T, 128, 8>(c, F, C, zero_x);
    k_mg_restrict<T><<<hsd::grid1(s.gr), dim3(BX, BY), 0, c->stream>>>(F, C, s.kc, s.gr.x, s.gr.y, s.gr.z, zero_x ? 1 : 0,
                                                                      c->sctl);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}

template <typename T>
int launch_prolong(hs_ctx* c, const Lv<T>& F, const Lv<T>& C) {
    const TileShape s = tile_shape(F.nx, F.ny, F.nz);
    k_mg_prolong<T><<<hsd::grid1(s.gr), dim3(BX, BY), 0, c->stream>>>(F, C, s.kc, s.gr.x, s.gr.y, s.gr.z, c->sctl);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}

// active-row mask of a level (one byte per cell instead of seven coefficient reads)
template <typename T>
__global__ void k_mg_mask(Lv<T> L, uint8_t* act) {
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y, k = 1 + blockIdx.z;
    if (i > L.nx) return;
    const long c = i + L.sy * j + L.sz * k;
    act[c] = diag_of(L, c) > T(0) ? 1 : 0;
}

// ---- single-workgroup tail: every level from `t` down, in one launch ----
// Small levels are launch-bound as separate kernels; one workgroup runs their
// pre-smoothing, restriction, coarsest sweeps, prolongation and post-smoothing with
// workgroup barriers between stages (the data stays in L2).
constexpr int TAIL_THREADS = 1024;
constexpr int TAIL_MAX_CELLS = 4096;   // levels at or below 16^3
constexpr int MAX_LEVELS = 16;

template <typename T>
struct LvPack {
    Lv<T> l[MAX_LEVELS];
};

template <typename T>
__device__ void tail_smooth(const Lv<T>& L, int colour) {
    const long ncol = (long)L.nx * L.ny * L.nz;
    for (long t = threadIdx.x; t < ncol; t += TAIL_THREADS) {
        const int i = 1 + (int)(t % L.nx), j = 1 + (int)((t / L.nx) % L.ny), k = 1 + (int)(t / ((long)L.nx * L.ny));
        if (((i + j + k) & 1) != colour) continue;
        const long c = i + L.sy * j + L.sz * k;
        const T D = diag_of(L, c);
        if (!(D > T(0))) continue;
        L.x[c] = (L.b[c] + offsum(L, c)) / D;
    }
    __syncthreads();
}

template <typename T>
__device__ void tail_restrict(const Lv<T>& F, const Lv<T>& C) {
    const long nc = (long)C.nx * C.ny * C.nz;
    for (long t = threadIdx.x; t < nc; t += TAIL_THREADS) {
        const int I = 1 + (int)(t % C.nx), J = 1 + (int)((t / C.nx) % C.ny), K = 1 + (int)(t / ((long)C.nx * C.ny));
        const int i0 = 2 * I - 1, j0 = 2 * J - 1, k0 = 2 * K - 1;
        const int ni = (2 * I <= F.nx) ? 2 : 1, nj = (2 * J <= F.ny) ? 2 : 1, nk = (2 * K <= F.nz) ? 2 : 1;
        T r = 0;
        for (int dk = 0; dk < nk; ++dk)
            for (int dj = 0; dj < nj; ++dj)
                for (int di = 0; di < ni; ++di) {
                    const long f = (i0 + di) + F.sy * (j0 + dj) + F.sz * (k0 + dk);
                    const T D = diag_of(F, f);
                    if (!(D > T(0))) continue;
                    r += F.b[f] - (D * F.x[f] - offsum(F, f));
                }
        const long cc = I + C.sy * J + C.sz * K;
        C.b[cc] = r;
        C.x[cc] = T(0);
    }
    __syncthreads();
}

template <typename T>
__device__ void tail_prolong(const Lv<T>& F, const Lv<T>& C) {
    const long nf = (long)F.nx * F.ny * F.nz;
    for (long t = threadIdx.x; t < nf; t += TAIL_THREADS) {
        const int i = 1 + (int)(t % F.nx), j = 1 + (int)((t / F.nx) % F.ny), k = 1 + (int)(t / ((long)F.nx * F.ny));
        const long f = i + F.sy * j + F.sz * k;
        if (!F.act[f]) continue;
        F.x[f] += prolong_value(C, i, j, k, C.x);
    }
    __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(TAIL_THREADS) void k_mg_tail(LvPack<T> P, int t, int nl, int pre, int post,
                                                           int coarse, const SolveCtl* ctl) {
    if (ctl->done) return;
    for (int l = t; l + 1 < nl; ++l) {
        for (int s = 0; s < pre; ++s) { tail_smooth(P.l[l], 0); tail_smooth(P.l[l], 1); }
        tail_restrict(P.l[l], P.l[l + 1]);
    }
    for (int s = 0; s < coarse; ++s) { tail_smooth(P.l[nl - 1], 0); tail_smooth(P.l[nl - 1], 1); }
    for (int s = 0; s < coarse; ++s) { tail_smooth(P.l[nl - 1], 1); tail_smooth(P.l[nl - 1], 0); }
    for (int l = nl - 2; l >= t; --l) {
        tail_prolong(P.l[l], P.l[l + 1]);
        for (int s = 0; s < post; ++s) { tail_smooth(P.l[l], 1); tail_smooth(P.l[l], 0); }
    }
}

// ---- flexible CG pieces (FP64 outer vectors; the level-0 V-cycle arrays are T) ----
// Each reduction kernel stores block partials; k_finish (hs_kernels.hip) adds them.
#define GRID_STRIDE(t, n) for (long t = blockIdx.x * (long)FLAT + threadIdx.x; t < (n); t += (long)gridDim.x * FLAT)

// x = 0, r = b, level-0 right-hand side = (T) b, level-0 iterate = 0; partials of <b,b>
template <typename T>
__global__ __launch_bounds__(FLAT) void k_fcg_setup(long n, const double* b, double* x, double* r, T* b0, T* x0,
                                                    int zero_x0, double* partials) {
    double acc = 0.0;
    GRID_STRIDE(t, n) {
        const double bv = b[t];
        x[t] = 0.0;
        r[t] = bv;
        b0[t] = (T)bv;
        if (zero_x0) x0[t] = T(0);
        acc += bv * bv;
    }
    hsd::put_partial<FLAT>(acc, threadIdx.x, blockIdx.x, partials);
}

// the zero start after a warm one that was no better (ctl->cold): x = 0, r = b
template <typename T>
__global__ __launch_bounds__(FLAT) void k_fcg_cold(long n, const double* b, double* x, double* r, T* b0, T* x0,
                                                   int zero_x0, const SolveCtl* ctl) {
    if (!ctl->cold) return;
    GRID_STRIDE(t, n) {
        const double bv = b[t];
        x[t] = 0.0;
        r[t] = bv;
        b0[t] = (T)bv;
        if (zero_x0) x0[t] = T(0);
    }
}

// out = the polynomial extrapolation of the last ord solutions x1 (newest), x2, x3:
// ord 1: x1; 2: 2 x1 - x2; 3: 3 x1 - 3 x2 + x3 (out may be x3: each element is read
// before it is written)
__global__ __launch_bounds__(FLAT) void k_extrapolate(long n, int ord, const double* x1, const double* x2,
                                                      const double* x3, double* out) {
    GRID_STRIDE(t, n) {
        double v = x1[t];
        if (ord == 2) v = 2.0 * v - x2[t];
        else if (ord == 3) v = 3.0 * v - 3.0 * x2[t] + x3[t];
        out[t] = v;
    }
}

// p = z (the level-0 iterate; the direction is stored as T); partials of <r, z>
template <typename T>
__global__ __launch_bounds__(FLAT) void k_fcg_start(long n, const double* r, const T* z, T* p, double* partials,
                                                    const SolveCtl* ctl) {
    if (ctl->done) return;
    double acc = 0.0;
    GRID_STRIDE(t, n) {
        const T zt = z[t];
        p[t] = zt;
        acc += r[t] * (double)zt;
    }
    hsd::put_partial<FLAT>(acc, threadIdx.x, blockIdx.x, partials);
}

// Two cells per thread: 16-byte loads and stores of the FP64 vectors (hipMalloc
// arrays, 256-byte aligned); thread 0 of block 0 takes the odd last cell.
template <typename T> struct Pair;
template <> struct Pair<float> { using V = float2; };
template <> struct Pair<double> { using V = double2; };

// x += alpha p, r -= alpha q; level-0 right-hand side = (T) r, iterate = 0 (unless the
// first sweep reads it as zero); partials of <r,r>.  The direction p is stored as T.
template <typename T>
__global__ __launch_bounds__(FLAT) void k_fcg_update(long n, double* x, double* r, const T* p, const double* q,
                                                     T* b0, T* x0, int zero_x0, double* partials, const SolveCtl* ctl) {
    using V = typename Pair<T>::V;
    if (ctl->done) return;
    const double alpha = ctl->alpha;
    double acc = 0.0;
    GRID_STRIDE(t, n >> 1) {
        const double2 xv = ((const double2*)x)[t];
        const V pv = ((const V*)p)[t];
        const double2 rv = ((const double2*)r)[t], qv = ((const double2*)q)[t];
        issued_together(xv.x, xv.y, pv.x, pv.y, rv.x, rv.y, qv.x, qv.y);
        const double r0 = rv.x - alpha * qv.x, r1 = rv.y - alpha * qv.y;
        ((double2*)x)[t] = double2{xv.x + alpha * (double)pv.x, xv.y + alpha * (double)pv.y};
        ((double2*)r)[t] = double2{r0, r1};
        acc += r0 * r0 + r1 * r1;
        ((V*)b0)[t] = V{(T)r0, (T)r1};
        if (zero_x0) ((V*)x0)[t] = V{T(0), T(0)};
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        const long t = n - 1;
        x[t] += alpha * (double)p[t];
        const double rv = r[t] - alpha * q[t];
        r[t] = rv;
        acc += rv * rv;
        b0[t] = (T)rv;
        if (zero_x0) x0[t] = T(0);
    }
    hsd::put_partial<FLAT>(acc, threadIdx.x, blockIdx.x, partials);
}

// partials of <z, r> and <z, q> (z = level-0 iterate)
template <typename T>
__global__ __launch_bounds__(FLAT) void k_fcg_beta(long n, const T* z, const double* r, const double* q,
                                                   double* partials, const SolveCtl* ctl) {
    using V = typename Pair<T>::V;
    if (ctl->done) return;
    double a = 0.0, b = 0.0;
    GRID_STRIDE(t, n >> 1) {
        const V zv = ((const V*)z)[t];
        const double2 rv = ((const double2*)r)[t], qv = ((const double2*)q)[t];
        issued_together(zv.x, zv.y, rv.x, rv.y, qv.x, qv.y);
        const double z0 = (double)zv.x, z1 = (double)zv.y;
        a += z0 * rv.x + z1 * rv.y;
        b += z0 * qv.x + z1 * qv.y;
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        const double zv = (double)z[n - 1];
        a += zv * r[n - 1];
        b += zv * q[n - 1];
    }
    hsd::put_partial2<FLAT>(a, b, threadIdx.x, blockIdx.x, gridDim.x, partials);
}

// ------------------------------------------------------------ host side ---
template <typename T>
int alloc_level(MGLevel* m) {
    void** arrs[7] = {&m->cw, &m->cs, &m->cb, &m->g, &m->x, &m->b, &m->x2};
    for (void** a : arrs) {
        if (hipMalloc(a, m->n * sizeof(T)) != hipSuccess) return HS_E_NOMEM;
        if (hipMemset(*a, 0, m->n * sizeof(T)) != hipSuccess) return HS_E_HIP;
    }
    if (hipMalloc(&m->act, m->n) != hipSuccess) return HS_E_NOMEM;
    if (hipMemset(m->act, 0, m->n) != hipSuccess) return HS_E_HIP;
    return HS_OK;
}

// zx: the level's iterate is zero on entry and the first fused sweep does not read it
// (the callers then skip writing those zeros)
template <typename T>
int smooth(hs_ctx* c, MGLevel* m, int sweeps, bool forward, bool zx = false, float omega_set = 0.0f) {
    Lv<T> L = lv<T>(m);
    const long tiles = (long)((L.nx + BX - 1) / BX) * ((L.ny + BY - 1) / BY);
    const int kc = hsd::kchunk(L.nz, tiles);
    const dim3 gr = hsd::grid3(L.nx, L.ny, L.nz, kc);
    // fused out-of-place sweeps x -> x2; after each the level's two iterate buffers swap,
    // so m->x always holds the latest iterate (every user reads it at launch time)
    const int c1 = forward ? 0 : 1;
    // tile of the fused sweep: 0 picks one by the level's width (256 x 4, 128 x 8 or
    // 64 x 8 columns); 4, 8 or 16 force 64-wide tiles of that many rows; below 0 the two
    // in-place colour passes run instead
    const int rows = c->prm.mg_sweep_rows;
    // over-relaxation: the finest level's, or that of the levels below it
    const T omega = (T)(omega_set > 0.0f ? omega_set : m->level == 0 ? c->prm.mg_omega : c->prm.mg_omega_coarse);
    if (sweeps > 0 && rows >= 0) {
        for (; sweeps > 0; --sweeps) {
            const T* src = (T*)m->x;
            T* dst = (T*)m->x2;
            const bool prof = m->level == 0;
            if (rows == 4) HS_TRY((launch_rbsweep<T, 64, 4>(c, L, src, dst, c1, omega, prof, zx)));
            else if (rows == 16) HS_TRY((launch_rbsweep<T, 64, 16>(c, L, src, dst, c1, omega, prof, zx)));
            else if (rows > 0 || L.nx <= 64) HS_TRY((launch_rbsweep<T, 64, 8>(c, L, src, dst, c1, omega, prof, zx)));
            else if (L.nx <= 128) HS_TRY((launch_rbsweep<T, 128, 8>(c, L, src, dst, c1, omega, prof, zx)));
            else HS_TRY((launch_rbsweep<T, 256, 4>(c, L, src, dst, c1, omega, prof, zx)));
            zx = false;
            std::swap(m->x, m->x2);
            L.x = (T*)m->x;
        }
    }
    for (int s = 0; s < sweeps; ++s)
        for (int h = 0; h < 2; ++h) {
            const int colour = forward ? h : 1 - h;
            const int slot = m->level == 0 ? prof_begin(c, HS_K_MG_SMOOTH0) : -1;
            k_mg_smooth<T><<<hsd::grid1(gr), dim3(BX, BY), 0, c->stream>>>(L, kc, colour, gr.x, gr.y, gr.z, omega,
                                                                        c->sctl);
            HS_TRY(prof_end(c, slot));
        }
    HS_CHECK(hipGetLastError());
    return HS_OK;
}

// sweeps before / after the coarse correction on level l (the finest level may differ)
inline int pre_sweeps(const hs_ctx* c, int l) { return (l == 0 && c->prm.mg_pre0 >= 0) ? c->prm.mg_pre0 : c->prm.mg_pre; }
// a level smoothed by the fused sweep starts its cycle visit from a zero guess that the
// first sweep does not read, so nothing needs to write those zeros
inline bool zero_guess(const hs_ctx* c, int l) { return c->prm.mg_sweep_rows >= 0 && pre_sweeps(c, l) > 0; }
inline int post_sweeps(const hs_ctx* c, int l) {
    return (l == 0 && c->prm.mg_post0 >= 0) ? c->prm.mg_post0 : c->prm.mg_post;
}

// One V-cycle on the level-0 arrays: b holds the residual, x starts at 0 and leaves
// holding the preconditioned residual (both prepared / consumed by the FCG kernels).
// Multigrid cycle from level l down to the level `lend` that is solved directly (the dense
// inverse, or the single-workgroup tail kernel for every level from `lend` down).
// Level 0 is visited once; each coarser level recurses mg_cycle times (1: V-cycle,
// 2: W-cycle below the finest level).  The second visit starts from the first's
// iterate: restriction forms the residual of the current iterate, the coarse correction
// is added by prolongation.
// zx: level l's iterate is zero on entry (its first sweep does not read it)
template <typename T>
int cycle(hs_ctx* c, int l, int lend, bool dense, bool zx) {
    const int nl = (int)c->mg.size();
    if (l == lend) {
        if (dense) return dense_solve(c, c->mg[l], sizeof(T) == sizeof(float));
        LvPack<T> P;
        for (int q = 0; q < nl && q < MAX_LEVELS; ++q) P.l[q] = lv<T>(c->mg[q]);
        k_mg_tail<T><<<1, TAIL_THREADS, 0, c->stream>>>(P, l, nl, c->prm.mg_pre, c->prm.mg_post,
                                                        c->prm.mg_coarse_sweeps, c->sctl);
        HS_CHECK(hipGetLastError());
        return HS_OK;
    }
    HS_TRY(smooth<T>(c, c->mg[l], pre_sweeps(c, l), true, zx));
    const bool zc = l + 1 != lend && zero_guess(c, l + 1);
    HS_TRY(launch_restrict<T>(c, lv<T>(c->mg[l]), lv<T>(c->mg[l + 1]), !zc));
    const int reps = (l == 0 || l + 1 == lend) ? 1 : std::max(1, c->prm.mg_cycle);
    for (int r = 0; r < reps; ++r) HS_TRY(cycle<T>(c, l + 1, lend, dense, zc && r == 0));
    HS_TRY(launch_prolong<T>(c, lv<T>(c->mg[l]), lv<T>(c->mg[l + 1])));
    return smooth<T>(c, c->mg[l], post_sweeps(c, l), false);
}

// One cycle on the level-0 arrays: b holds the residual, x starts at 0 and leaves holding
// the preconditioned residual (both prepared / consumed by the FCG kernels).
// level solved directly by the cycle: the dense level, or the first one small enough for
// the single-workgroup tail
inline int cycle_end(const hs_ctx* c, bool* dense) {
    const int nl = (int)c->mg.size();
    *dense = c->prm.mg_dense_coarse > 0 && c->dense;
    if (*dense) return nl - 1;
    int t = 0;
    while (t < nl && (long)c->mg[t]->nx * c->mg[t]->ny * c->mg[t]->nz > TAIL_MAX_CELLS) ++t;
    return t >= nl ? nl - 1 : t;
}
// the cycle's level-0 iterate is read as zero by its first sweep (nothing clears it)
inline bool cycle_zero_guess(const hs_ctx* c) {
    bool dense;
    return cycle_end(c, &dense) > 0 && zero_guess(c, 0);
}

// Aggregation multigrid cycle: red-black sweeps on the finest level (the level-0 iterate
// is zero on entry), the aggregated levels below it (hs_amg.hip), sweeps in the
// opposite colour order
template <typename T>
int amg_vcycle(hs_ctx* c) {
    MGLevel* m = c->mg[0];
    HS_TRY(smooth<T>(c, m, c->prm.amg_sweeps0, true, false, c->prm.amg_omega));
    HS_TRY(amg_coarse_p(c, &c->sctl->done));
    return smooth<T>(c, m, c->prm.amg_sweeps0, false, false, c->prm.amg_omega);
}

template <typename T>
int vcycle(hs_ctx* c) {
    if (c->prm.precond == HS_PRECOND_AMG) return amg_vcycle<T>(c);
    bool dense;
    const int t = cycle_end(c, &dense);
    return cycle<T>(c, 0, t, dense, cycle_zero_guess(c));
}

template <typename T>
int build_hierarchy(hs_ctx* c) {
    const Dims& d = c->d;
    if (c->mg.empty()) {
        int nx = d.nx, ny = d.ny, nz = d.nz;
        const int maxl = std::min(MAX_LEVELS, c->prm.mg_levels > 0 ? c->prm.mg_levels : MAX_LEVELS);
        for (int l = 0; l < maxl; ++l) {
            MGLevel* m = new MGLevel();
            m->level = l;
            m->nx = nx; m->ny = ny; m->nz = nz;
            m->sy = nx + 2; m->sz = (long)(nx + 2) * (ny + 2); m->n = m->sz * (nz + 2);
            m->f32 = sizeof(T) == 4;
            m->cw = m->cs = m->cb = m->g = m->x = m->b = m->x2 = nullptr;
            m->act = nullptr;
            c->mg.push_back(m);
            HS_TRY(alloc_level<T>(m));
            // the aggregation multigrid builds its own levels below the finest one
            if (c->prm.precond == HS_PRECOND_AMG) break;
            if (std::min(nx, std::min(ny, nz)) <= 4) break;
            // with the dense coarse solve the hierarchy ends at the first small enough level
            if (c->prm.mg_dense_coarse > 0 && (long)nx * ny * nz <= dense_max_cells()) break;
            nx = (nx + 1) / 2; ny = (ny + 1) / 2; nz = (nz + 1) / 2;
        }
    }
    Lv<T> L0 = lv<T>(c->mg[0]);
    k_mg_fine<T><<<dim3((d.nx + 63) / 64, d.ny, d.nz), 64, 0, c->stream>>>(d, c->cls, c->ax, c->ay, c->az, L0,
                                                                          c->mg[0]->act);
    HS_CHECK(hipGetLastError());
    if (c->prm.precond == HS_PRECOND_AMG) return amg_setup_p(c);
    for (size_t l = 0; l + 1 < c->mg.size(); ++l) {
        Lv<T> F = lv<T>(c->mg[l]), C = lv<T>(c->mg[l + 1]);
        // ground conductance: the Dirichlet point sits half a finest-level cell outside the
        // boundary face, i.e. (h_l + h_0)/2 from a level-l boundary centre, so a coarse
        // ground is (2^l + 1)/(2^(l+1) + 1) of the summed fine grounds (2/3, 3/5, 5/9, ...)
        const double pw = std::ldexp(1.0, (int)l);
        const T gscale = (T)((pw + 1.0) / (2.0 * pw + 1.0));
        k_mg_coarsen<T><<<dim3((C.nx + 63) / 64, C.ny, C.nz), 64, 0, c->stream>>>(F, C, gscale);
        HS_CHECK(hipGetLastError());
    }
    for (MGLevel* m : c->mg) {
        if (m->level == 0) continue;   // written by k_mg_fine
        k_mg_mask<T><<<dim3((m->nx + 63) / 64, m->ny, m->nz), 64, 0, c->stream>>>(lv<T>(m), m->act);
        HS_CHECK(hipGetLastError());
    }
    if (c->prm.mg_dense_coarse > 0 && (long)c->mg.back()->nx * c->mg.back()->ny * c->mg.back()->nz <= dense_max_cells())
        HS_TRY(dense_setup(c, c->mg.back(), sizeof(T) == sizeof(float)));
    return HS_OK;
}

template <typename T>
int fcg(hs_ctx* c, const double* b, double* x, hs_stats* st, bool with_resid) {
    const long n = c->d.nP;
    HS_TRY(build_hierarchy<T>(c));
    SolveCtl* h = c->h_sctl;
    h->eps = c->prm.p_tol * c->prm.p_tol;
    h->maxit = c->prm.p_maxit;
    HS_CHECK(hipMemcpyAsync(&c->sctl->eps, &h->eps, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HS_CHECK(hipMemcpyAsync(&c->sctl->maxit, &h->maxit, sizeof(int), hipMemcpyHostToDevice, c->stream));
    // directions in T (float in the mixed-precision solve), in the FP64 buffers' storage
    double *r = c->kr, *q = c->ks;
    T *p = (T*)c->kp, *p2 = (T*)c->kp2;
    T* b0 = (T*)c->mg[0]->b;
    // the level-0 iterate (the V-cycle's output z) moves between two buffers with the
    // smoother's sweeps: read its current pointer at every use
    auto z = [&]() { return (T*)c->mg[0]->x; };
    const int G = FLAT_BLOCKS;
    const int zero_x0 = cycle_zero_guess(c) ? 0 : 1;
    // warm start (p_warm = m): the initial iterate is the polynomial through the last
    // min(m, 3, held) SIMPLE p' solutions (1: the last one, still in x; 2: 2 x1 - x2;
    // 3: 3 x1 - 3 x2 + x3).  For m >= 2 it is formed in the oldest history buffer, which
    // becomes the solve's output, and the history moves down one place.  A start whose
    // residual is not below ||b|| falls back to zero.
    const int ord = x == c->kx ? std::min(std::min(c->prm.p_warm, 3), c->kx_hist) : 0;
    if (ord > 0) {
        if (c->prm.p_warm >= 2) {
            for (double*& h : c->khist)
                if (!h && hipMalloc(&h, n * sizeof(double)) != hipSuccess) return HS_E_NOMEM;
            k_extrapolate<<<G, FLAT, 0, c->stream>>>(n, ord, c->kx, c->khist[0], c->khist[1], c->khist[1]);
            double* next = c->khist[1];
            c->khist[1] = c->khist[0];
            c->khist[0] = c->kx;
            c->kx = x = next;
        }
        HS_TRY(warm_residual(c, b, x, r, b0, z(), zero_x0 != 0, sizeof(T) == sizeof(float)));
        k_fcg_cold<T><<<G, FLAT, 0, c->stream>>>(n, b, x, r, b0, z(), zero_x0, c->sctl);
    } else {
        k_fcg_setup<T><<<G, FLAT, 0, c->stream>>>(n, b, x, r, b0, z(), zero_x0, c->partials);
        HS_TRY(launch_finish(c, FIN_FCG_SETUP, G, nullptr));
    }
    HS_TRY(vcycle<T>(c));
    k_fcg_start<T><<<G, FLAT, 0, c->stream>>>(n, r, z(), p, c->partials, c->sctl);
    HS_TRY(launch_finish(c, FIN_FCG_START, G, nullptr));
    HS_CHECK(hipGetLastError());
    int launched = 0;
    // first batch: as many iterations as the last solve took (an iteration queued after
    // convergence exits at once on ctl->done), then two at a time
    int batch = c->last_iters > 0 ? c->last_iters : 4;
    while (true) {
        for (int it = 0; it < batch && launched < c->prm.p_maxit; ++it, ++launched) {
            // the first direction is p = z (k_fcg_start); later ones are formed by the
            // fused update + apply into the second buffer, and the buffers swap
            if (launched == 0) {
                HS_TRY(cg_spmv_dot(c, p, sizeof(T) == sizeof(float), q));
            } else {
                HS_TRY(fcg_spmv_dot(c, z(), sizeof(T) == sizeof(float), p, p2, q));
                std::swap(p, p2);
            }
            k_fcg_update<T><<<G, FLAT, 0, c->stream>>>(n, x, r, p, q, b0, z(), zero_x0, c->partials, c->sctl);
            HS_TRY(launch_finish(c, FIN_FCG_UPDATE, G, nullptr));
            HS_TRY(vcycle<T>(c));
            k_fcg_beta<T><<<G, FLAT, 0, c->stream>>>(n, z(), r, q, c->partials, c->sctl);
            HS_TRY(launch_finish(c, FIN_FCG_BETA, G, nullptr));
        }
        HS_CHECK(hipGetLastError());
        if (with_resid) HS_TRY(launch_resid_norms(c, b, x));
        HS_CHECK(hipMemcpyAsync(h, c->sctl, sizeof(SolveCtl), hipMemcpyDeviceToHost, c->stream));
        HS_CHECK(hipStreamSynchronize(c->stream));
        if (h->done || launched >= c->prm.p_maxit) break;
        batch = 2;
    }
    c->last_iters = h->iters;
    if (c->dense && c->dense->iters_ref < 0) c->dense->iters_ref = h->iters;
    if (c->amgp && c->amgp->iters_ref < 0) c->amgp->iters_ref = h->iters;
    if (st) {
        st->p_iters = h->iters;
        st->p_converged = h->converged;
        st->p_relres = h->bb > 0 ? sqrt(h->rr / h->bb) : 0.0;
        if (with_resid) resid_from_red(h, st);
    }
    if (h->nan || !std::isfinite(h->rr)) return HS_E_NAN;
    return HS_OK;
}

}  // namespace

int gmg_setup(hs_ctx* c) {
    if (c->prm.mode != HS_MODE_CONFIG) return HS_E_ARG;
    return c->prm.precision == HS_PREC_MIXED ? build_hierarchy<float>(c) : build_hierarchy<double>(c);
}

int fcg_solve(hs_ctx* c, const double* b, double* x, hs_stats* st, bool with_resid) {
    // the multigrid operator form needs the symmetric face-coefficient p' operator
    if (c->prm.mode != HS_MODE_CONFIG) return HS_E_ARG;
    return c->prm.precision == HS_PREC_MIXED ? fcg<float>(c, b, x, st, with_resid)
                                             : fcg<double>(c, b, x, st, with_resid);
}

void gmg_free(hs_ctx* c) {
    dense_free(c);
    amg_free(&c->amgp);
    for (MGLevel* m : c->mg) {
        void* arrs[8] = {m->cw, m->cs, m->cb, m->g, m->x, m->b, m->x2, m->act};
        for (void* a : arrs) if (a) (void)hipFree(a);
        delete m;
    }
    c->mg.clear();
}
