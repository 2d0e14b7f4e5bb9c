// hs_device.h — device helpers shared by the kernel files: wave64 / block sums and the
// deterministic "last block finishes" reduction.
#pragma once

#include "hs_internal.h"

namespace hsd {

constexpr int BX = 64, BY = 4, BT = BX * BY;
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

// Writes the block partial (agent-scope store), bumps the counter with acq_rel; the
// block that finishes last sums all partials in a fixed order.  Returns true in every
// thread of that block; *out valid in tid 0.  The counter is reset for the next use.
template <int NT>
__device__ bool reduce_last(double v, int tid, int bid, int nblocks, double* partials, unsigned int* counter,
                            double* out) {
    __shared__ int am_last;
    const double bs = block_sum<NT>(v, tid);
    if (tid == 0) {
        __hip_atomic_store(&partials[bid], bs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        am_last = (prev == (unsigned)(nblocks - 1));
    }
    __syncthreads();
    if (!am_last) return false;
    __threadfence();
    double s = 0.0;
    for (int i = tid; i < nblocks; i += NT) s += __hip_atomic_load(&partials[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s = block_sum<NT>(s, tid);
    if (tid == 0) {
        *out = s;
        __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return true;
}

// Same for two sums (partials laid out [2][nblocks]).
template <int NT>
__device__ bool reduce2_last(double a, double b, int tid, int bid, int nblocks, double* partials,
                             unsigned int* counter, double* outa, double* outb) {
    __shared__ int am_last;
    block_sum2<NT>(a, b, tid);
    if (tid == 0) {
        __hip_atomic_store(&partials[bid], a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&partials[nblocks + bid], b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        am_last = (prev == (unsigned)(nblocks - 1));
    }
    __syncthreads();
    if (!am_last) return false;
    __threadfence();
    double sa = 0.0, sb = 0.0;
    for (int i = tid; i < nblocks; i += NT) {
        sa += __hip_atomic_load(&partials[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        sb += __hip_atomic_load(&partials[nblocks + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    block_sum2<NT>(sa, sb, tid);
    if (tid == 0) {
        *outa = sa;
        *outb = sb;
        __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return true;
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

}  // namespace hsd
