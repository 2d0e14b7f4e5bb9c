// hs_mglevel.h — device view of one level of the pressure multigrid hierarchy (shared
// by the smoothers / transfers and the dense coarsest-level solve).
//
// Operator on a level (cell-centred, padded layout (n+2)^3 of that level):
//   (A x)_P = D_P x_P - cw_P x_W - cw_E x_E - cs_P x_S - cs_N x_N - cb_P x_B - cb_T x_T,
//   D_P = g_P + cw_P + cw_E + cs_P + cs_N + cb_P + cb_T;  inactive cells have D = 0.
#pragma once

#include "hs_internal.h"

namespace mgl {

template <typename T>
struct Lv {
    int nx, ny, nz;
    long sy, sz;
    T *cw, *cs, *cb, *g, *x, *b;
    const uint8_t* act;   // active-row mask (D > 0)
};

template <typename T>
inline Lv<T> lv(const MGLevel* m) {
    return Lv<T>{m->nx, m->ny, m->nz, m->sy, m->sz, (T*)m->cw, (T*)m->cs, (T*)m->cb, (T*)m->g, (T*)m->x, (T*)m->b,
                 m->act};
}

template <typename T>
__device__ __forceinline__ T diag_of(const Lv<T>& L, long c) {
    return L.g[c] + L.cw[c] + L.cw[c + 1] + L.cs[c] + L.cs[c + L.sy] + L.cb[c] + L.cb[c + L.sz];
}

template <typename T>
__device__ __forceinline__ T offsum(const Lv<T>& L, long c) {
    const T* x = L.x;
    return L.cw[c] * x[c - 1] + L.cw[c + 1] * x[c + 1] + L.cs[c] * x[c - L.sy] + L.cs[c + L.sy] * x[c + L.sy] +
           L.cb[c] * x[c - L.sz] + L.cb[c + L.sz] * x[c + L.sz];
}

}  // namespace mgl
