// hs_device.h — device helpers shared by the kernel files: wave64 / block sums and the
// deterministic two-kernel reduction (block partials + one-block finish).
#pragma once

#include "hs_internal.h"

namespace hsd {

constexpr int BX = 64, BY = 4, BT = BX * BY;
// red-black colour passes of the momentum and energy sweeps: two faces per thread along x,
// so a tile is 64 faces by 16 rows (taller tiles re-read fewer neighbour rows: 230 us
// against 243 us for 128 x 4 at 256^3)
constexpr int CBX = 32, CBY = 16, CBT = CBX * CBY;
constexpr int FLAT = 256;
constexpr int FLAT_BLOCKS = 2048;

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Block sum over NT threads (multiple of 64); result valid in tid 0.
template <int NT>
__device__ __forceinline__ double block_sum(double v, int tid) {
    __shared__ double sh[NT / 64];
    v = wave_sum(v);
    const int lane = tid & 63, wid = tid >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (wid == 0) {
        r = (lane < NT / 64) ? sh[lane] : 0.0;
        r = wave_sum(r);
    }
    __syncthreads();
    return r;
}

// Two sums at once (same fixed order as block_sum).
template <int NT>
__device__ __forceinline__ void block_sum2(double& a, double& b, int tid) {
    __shared__ double sh[2][NT / 64];
    a = wave_sum(a);
    b = wave_sum(b);
    const int lane = tid & 63, wid = tid >> 6;
    if (lane == 0) { sh[0][wid] = a; sh[1][wid] = b; }
    __syncthreads();
    double ra = 0.0, rb = 0.0;
    if (wid == 0) {
        ra = (lane < NT / 64) ? sh[0][lane] : 0.0;
        rb = (lane < NT / 64) ? sh[1][lane] : 0.0;
        ra = wave_sum(ra);
        rb = wave_sum(rb);
    }
    __syncthreads();
    a = ra;
    b = rb;
}

// Reductions: every block stores its partial sum with a plain store; a separate
// one-block finish kernel, launched next on the same stream, adds the partials in a
// fixed order and runs the scalar control step.  Stream order makes the partials
// visible without per-block device-scope release/acquire (on a multi-XCD device each
// such release writes back the XCD's L2), and the fixed order keeps results
// deterministic run to run.
template <int NT>
__device__ __forceinline__ void put_partial(double v, int tid, int bid, double* partials) {
    const double bs = block_sum<NT>(v, tid);
    if (tid == 0) partials[bid] = bs;
}

template <int NT>
__device__ __forceinline__ void put_partial2(double a, double b, int tid, int bid, int nblocks, double* partials) {
    block_sum2<NT>(a, b, tid);
    if (tid == 0) {
        partials[bid] = a;
        partials[nblocks + bid] = b;
    }
}

template <int NT>
__device__ __forceinline__ double sum_partials(const double* partials, int n, int tid) {
    double s = 0.0;
    for (int i = tid; i < n; i += NT) s += partials[i];
    return block_sum<NT>(s, tid);
}

// XCD-aware tile order.  Workgroups are dispatched round-robin over the 8 XCDs (each
// with its own L2), so launch-order neighbours sit on different L2s.  A 1-D launch of
// nb blocks is remapped so that XCD x owns one contiguous run of logical tiles (x fastest,
// then y, then z): tiles that share halo rows / planes share an L2.
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int xcd_tile(int bid, int nb) {
    const int q = nb / NUM_XCD, rem = nb % NUM_XCD;
    const int x = bid % NUM_XCD, r = bid / NUM_XCD;
    return (x < rem) ? x * (q + 1) + r : rem * (q + 1) + (x - rem) * q + r;
}
struct Tile {
    int bx, by, bz;
};
__device__ __forceinline__ Tile tile_of(int gx, int gy, int gz) {
    const int nb = gx * gy * gz;
    const int L = xcd_tile((int)blockIdx.x, nb);
    return Tile{L % gx, (L / gx) % gy, L / (gx * gy)};
}
inline dim3 grid1(const dim3& g3) { return dim3(g3.x * g3.y * g3.z); }

// an empty asm statement taking each value as an operand: the loads that produce them are
// all issued before it, so the compiler cannot sink some of them under a later branch
template <typename V>
__device__ __forceinline__ int held(const V& v) {
    asm volatile("" ::"v"(v));
    return 0;
}
template <typename... V>
__device__ __forceinline__ void issued_together(const V&... v) {
    const int order[] = {held(v)...};
    (void)order;
}

// over-relaxed Gauss-Seidel value x + omega (gs - x); omega == 1 gives gs itself (the
// exact Gauss-Seidel expression, not x + 1 (gs - x))
template <typename T>
__device__ __forceinline__ T sor(T x, T gs, T omega) {
    return omega == T(1) ? gs : x + omega * (gs - x);
}

inline int kchunk(int nk, long tiles) {
    long want = 4096 / (tiles > 0 ? tiles : 1);
    if (want < 1) want = 1;
    int kc = (int)((nk + want - 1) / want);
    return kc < 1 ? 1 : kc;
}
inline dim3 grid3(int ni, int nj, int nk, int kc) {
    return dim3((ni + BX - 1) / BX, (nj + BY - 1) / BY, (nk + kc - 1) / kc);
}

// block partials one p' Krylov solve writes at most: the flat kernels' two sums per block
// and the two sums per block of the z-marching residual pass over the grid's columns
inline long krylov_partials(const Dims& d) {
    const long tiles = (long)((d.nx + BX - 1) / BX) * ((d.ny + BY - 1) / BY);
    const dim3 gr = grid3(d.nx, d.ny, d.nz, kchunk(d.nz, tiles));
    const long nb = (long)gr.x * gr.y * gr.z;
    return 2 * nb > 2L * FLAT_BLOCKS ? 2 * nb : 2L * FLAT_BLOCKS;
}

}  // namespace hsd
