// hs_eop.h — device view of the Mode C energy operator on one level of the energy
// hierarchy (padded cell layout of the level), shared by the energy multigrid, its
// BiCGStab and the aggregation multigrid built on its finest level.
//
//   (A x)_P = a_P x_P - a_W x_W - a_E x_E - a_S x_S - a_N x_N - a_B x_B - a_T x_T
//   a_W = D_w + max(F_w, 0),  a_E = D_e + max(-F_e, 0)      (likewise S/N and B/T)
//   a_P = a_E + a_W + a_N + a_S + a_B + a_T
// with D the face conductance and F the convective flux through the face (positive along
// +axis), both stored on the min face of every cell.  The transpose keeps the diagonal
// and reverses F in the off-diagonals.
#pragma once

#include "hs_internal.h"

namespace heo {

template <typename T>
__device__ __forceinline__ T pos0(T a) {
    return a < T(0) ? T(0) : a;
}

template <typename T>
struct EL {
    int nx, ny, nz;
    long sy, sz;
    const T *dx, *dy, *dz, *fx, *fy, *fz;
    T *x, *b;
};

template <typename T>
EL<T> el(const ELevel* m) {
    return EL<T>{m->nx, m->ny, m->nz, m->sy, m->sz, (const T*)m->dx, (const T*)m->dy, (const T*)m->dz,
                 (const T*)m->fx, (const T*)m->fy, (const T*)m->fz, (T*)m->x, (T*)m->b};
}

struct EOp {
    double *dx, *dy, *dz, *fx, *fy, *fz;
};

// Row P: off-diagonal magnitudes of A (tr = 0) or A^T (tr = 1) and the diagonal of A.
template <typename T>
struct Row {
    T w, e, s, n, b, t, p;
};

template <typename T>
__device__ __forceinline__ Row<T> row(const T* dx, const T* dy, const T* dz, const T* fx, const T* fy, const T* fz,
                                      long c, long sy, long sz, int tr) {
    const T dW = dx[c], dE = dx[c + 1], dS = dy[c], dN = dy[c + sy], dB = dz[c], dT = dz[c + sz];
    const T fW = fx[c], fE = fx[c + 1], fS = fy[c], fN = fy[c + sy], fB = fz[c], fT = fz[c + sz];
    const T aw = dW + pos0(fW), ae = dE + pos0(-fE), as = dS + pos0(fS), an = dN + pos0(-fN);
    const T ab = dB + pos0(fB), at = dT + pos0(-fT);
    Row<T> r;
    r.p = ae + aw + an + as + ab + at;
    if (!tr) {
        r.w = aw; r.e = ae; r.s = as; r.n = an; r.b = ab; r.t = at;
    } else {
        r.w = dW + pos0(-fW); r.e = dE + pos0(fE); r.s = dS + pos0(-fS);
        r.n = dN + pos0(fN); r.b = dB + pos0(-fB); r.t = dT + pos0(fT);
    }
    return r;
}

template <typename T>
__device__ __forceinline__ Row<T> lrow(const EL<T>& L, long c, int tr) {
    return row<T>(L.dx, L.dy, L.dz, L.fx, L.fy, L.fz, c, L.sy, L.sz, tr);
}

template <typename T, typename TX>
__device__ __forceinline__ T offs(const Row<T>& r, const TX* x, long c, long sy, long sz) {
    return r.w * (T)x[c - 1] + r.e * (T)x[c + 1] + r.s * (T)x[c - sy] + r.n * (T)x[c + sy] + r.b * (T)x[c - sz] +
           r.t * (T)x[c + sz];
}

}  // namespace heo
