// hs_dense.hip — direct solve of the coarsest multigrid level on the matrix cores.
//
// The coarsest level (at most DENSE_MAX_CELLS cells, 16^3 under a 256^3 grid) is
// assembled as a dense matrix in natural cell order, in the hierarchy's scalar (FP32 for
// the mixed-precision V-cycle, FP64 for the FP64 one; rows of inactive cells and of the
// padding to a multiple of 64 are identity rows), and inverted in place by blocked
// Gauss-Jordan elimination without pivoting (the level operator is symmetric positive
// definite).  For every 64-wide block pivot k:
//   P    = A_kk^-1                         one workgroup, the block in LDS
//   C_I  = A_Ik P              (I != k)    MFMA tile products
//   A_IJ = A_IJ - C_I A_kJ     (I, J != k) MFMA rank-64 update of the whole matrix
//   A_kJ = P A_kJ,  A_Ik = -C_I,  A_kk = P
// All products run on v_mfma_f32_16x16x4f32 or v_mfma_f64_16x16x4f64 (64x64 output tile
// per workgroup, four wavefronts in a 2x2 arrangement of 32x32, K = 64 from LDS).  The
// inverse is kept between hierarchy builds and refreshed at most every `mg_dense_coarse`
// builds (sooner when the flexible CG's iteration count shows the inverse has gone stale).
//
// Each V-cycle then solves the coarse system with the inverse and FP64 iterative
// refinement against the level's current stencil:
//   z = A^-1 r;  repeat mg_dense_refine times: z += A^-1 (r - A z)      (residual in FP64)
// The inverse apply is a full-chip GEMV (matrix and input in the matrix scalar, FP64 sums
// across lanes).  A refinement step shrinks the error by the inverse's accuracy against
// the current operator: the FP32 rounding of the mixed path's inverse, and for either
// precision the change of the coarse operator since the inverse was built.
#include <cmath>

#include "hs_device.h"
#include "hs_mglevel.h"

using mgl::Lv;

namespace {

constexpr int DB = 64;          // Gauss-Jordan block and MFMA output tile
constexpr int TP = DB + 1;      // LDS pitch (odd: column reads of 16 rows hit 16 banks)
constexpr int DENSE_MAX_CELLS = 4096;

// ---- 16x16x4 matrix-core products -------------------------------------------------------
// Both forms take A[i][k] from lane i + 16 k and B[k][j] from lane j + 16 k.  The FP32
// form returns D[4 (lane/16) + v][lane % 16] in element v of the accumulator, the FP64
// form D[lane/16 + 4 v][lane % 16] (both checked element by element on gfx950).
template <typename S>
struct Mma;
template <>
struct Mma<float> {
    typedef float acc4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc4 op(float a, float b, acc4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int v) { return 4 * (l >> 4) + v; }
};
template <>
struct Mma<double> {
    typedef double acc4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc4 op(double a, double b, acc4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int v) { return (l >> 4) + 4 * v; }
};

// acc += As (64x64, LDS pitch TP) * Bs (64x64, LDS pitch TP).  Wave w owns the 32x32
// output block at rows 32*(w>>1), cols 32*(w&1), as 2x2 tiles of 16x16.
template <typename S>
__device__ __forceinline__ void tile_mma(const S* As, const S* Bs, typename Mma<S>::acc4 acc[2][2]) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int r0 = 32 * (w >> 1), c0 = 32 * (w & 1);
    const int lr = l & 15, lk = l >> 4;
#pragma unroll 4
    for (int k0 = 0; k0 < DB; k0 += 4) {
        const S a0 = As[(r0 + lr) * TP + k0 + lk];
        const S a1 = As[(r0 + 16 + lr) * TP + k0 + lk];
        const S b0 = Bs[(k0 + lk) * TP + c0 + lr];
        const S b1 = Bs[(k0 + lk) * TP + c0 + 16 + lr];
        acc[0][0] = Mma<S>::op(a0, b0, acc[0][0]);
        acc[0][1] = Mma<S>::op(a0, b1, acc[0][1]);
        acc[1][0] = Mma<S>::op(a1, b0, acc[1][0]);
        acc[1][1] = Mma<S>::op(a1, b1, acc[1][1]);
    }
}

// (row, col) inside the 64x64 tile of accumulator element (ti, tj, v) of this lane
template <typename S>
__device__ __forceinline__ void acc_pos(int ti, int tj, int v, int* row, int* col) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    *row = 32 * (w >> 1) + 16 * ti + Mma<S>::row(l, v);
    *col = 32 * (w & 1) + 16 * tj + (l & 15);
}

// 64x64 block at (r0, c0) of a row-major matrix with leading dimension ld -> LDS, with
// 16-byte loads (4 floats or 2 doubles)
template <typename S>
__device__ __forceinline__ void load_tile(S* Sm, const S* G, long ld, long r0, long c0) {
    constexpr int V = 16 / sizeof(S), PER_ROW = DB / V;
    for (int e = threadIdx.x; e < DB * PER_ROW; e += 256) {
        const int r = e / PER_ROW, q = (e % PER_ROW) * V;
        const S* src = G + (r0 + r) * ld + c0 + q;
        if (V == 4) {
            const float4 v = *reinterpret_cast<const float4*>(src);
            Sm[r * TP + q] = v.x; Sm[r * TP + q + 1] = v.y; Sm[r * TP + q + 2] = v.z; Sm[r * TP + q + 3] = v.w;
        } else {
            const double2 v = *reinterpret_cast<const double2*>(src);
            Sm[r * TP + q] = v.x; Sm[r * TP + q + 1] = v.y;
        }
    }
}

template <typename S>
__device__ __forceinline__ void zero_acc(typename Mma<S>::acc4 acc[2][2]) {
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = typename Mma<S>::acc4{S(0), S(0), S(0), S(0)};
}

// ---- assembly ---------------------------------------------------------------------
template <typename S>
__global__ void k_dense_fill(S* M, long n) {
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) M[t] = S(0);
}

// row r = cell (i,j,k) of the level in natural order; identity for inactive / padding rows
template <typename T>
__global__ void k_dense_assemble(Lv<T> L, T* M, int np) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= np) return;
    T* row = M + (long)r * np;
    const int n = L.nx * L.ny * L.nz;
    if (r >= n) { row[r] = T(1); return; }
    const int i = 1 + r % L.nx, j = 1 + (r / L.nx) % L.ny, k = 1 + r / (L.nx * L.ny);
    const long c = i + L.sy * j + L.sz * k;
    const T D = mgl::diag_of(L, c);
    if (!(D > T(0))) { row[r] = T(1); return; }
    row[r] = D;
    const int sx = 1, sy = L.nx, sz = L.nx * L.ny;
    if (i > 1) row[r - sx] = -L.cw[c];
    if (i < L.nx) row[r + sx] = -L.cw[c + 1];
    if (j > 1) row[r - sy] = -L.cs[c];
    if (j < L.ny) row[r + sy] = -L.cs[c + L.sy];
    if (k > 1) row[r - sz] = -L.cb[c];
    if (k < L.nz) row[r + sz] = -L.cb[c + L.sz];
}

// ---- blocked Gauss-Jordan steps ------------------------------------------------------
// P = inverse of the 64x64 diagonal block kb, Gauss-Jordan on the block in LDS by four
// wavefronts: thread t owns row t % 64, columns 16 (t / 64) .. +15.  Every pivot step
// reads the pivot row, the pivot and the row's pivot-column entry into registers, then
// (after a barrier: another wave may still be reading what this step rewrites) writes
// the row's 16 new entries; the owner of the pivot column also writes the row's
// multiplier there.  A second barrier orders the writes before the next step's reads.
constexpr int TQ = DB + 4;      // row pitch of the pivot block in LDS
template <typename S>
__global__ __launch_bounds__(256) void k_gj_diag(const S* M, int np, int kb, S* P) {
    __shared__ S Sm[DB * TQ];
    const int i = threadIdx.x & 63, h = threadIdx.x >> 6, q0 = 16 * h;
    const long o = (long)kb * DB;
#pragma unroll
    for (int q = 0; q < 16; ++q) Sm[i * TQ + q0 + q] = M[(o + i) * np + o + q0 + q];
    __syncthreads();
#pragma clang loop unroll(disable)
    for (int p = 0; p < DB; ++p) {
        const S piv = S(1) / Sm[p * TQ + p];
        const S f = Sm[i * TQ + p] * piv;       // a_ip / a_pp (this thread's row)
        const bool own = i == p;
        S pv[16], mv[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            pv[q] = Sm[p * TQ + q0 + q];
            mv[q] = Sm[i * TQ + q0 + q];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; ++q) Sm[i * TQ + q0 + q] = own ? pv[q] * piv : mv[q] - f * pv[q];
        if (p >= q0 && p < q0 + 16) Sm[i * TQ + p] = own ? piv : -f;
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) P[i * DB + q0 + q] = Sm[i * TQ + q0 + q];
}

// C_I = A_{I,kb} P for every row block I != kb (C stored as np/64 blocks of 64x64)
template <typename S>
__global__ __launch_bounds__(256) void k_gj_panel(const S* M, int np, int kb, const S* P, S* C) {
    const int I = blockIdx.x;
    if (I == kb) return;
    __shared__ S As[DB * TP], Bs[DB * TP];
    load_tile(As, M, np, (long)I * DB, (long)kb * DB);
    load_tile(Bs, P, DB, 0, 0);
    __syncthreads();
    typename Mma<S>::acc4 acc[2][2];
    zero_acc<S>(acc);
    tile_mma(As, Bs, acc);
    S* Cb = C + (long)I * DB * DB;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int v = 0; v < 4; ++v) {
                int r, cc;
                acc_pos<S>(a, b, v, &r, &cc);
                Cb[r * DB + cc] = acc[a][b][v];
            }
}

// A_IJ -= C_I A_{kb,J} for I, J != kb
template <typename S>
__global__ __launch_bounds__(256) void k_gj_update(S* M, int np, int kb, const S* C) {
    const int nb = np / DB;
    const int I = blockIdx.x % nb, J = blockIdx.x / nb;
    if (I == kb || J == kb) return;
    __shared__ S As[DB * TP], Bs[DB * TP];
    load_tile(As, C + (long)I * DB * DB, DB, 0, 0);
    load_tile(Bs, M, np, (long)kb * DB, (long)J * DB);
    __syncthreads();
    typename Mma<S>::acc4 acc[2][2];
    zero_acc<S>(acc);
    tile_mma(As, Bs, acc);
    S* T = M + (long)I * DB * np + (long)J * DB;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int v = 0; v < 4; ++v) {
                int r, cc;
                acc_pos<S>(a, b, v, &r, &cc);
                T[(long)r * np + cc] -= acc[a][b][v];
            }
}

// A_{kb,J} = P A_{kb,J},  A_{J,kb} = -C_J  (J != kb);  A_{kb,kb} = P
template <typename S>
__global__ __launch_bounds__(256) void k_gj_finish(S* M, int np, int kb, const S* P, const S* C) {
    const int J = blockIdx.x;
    const long o = (long)kb * DB;
    if (J == kb) {
        for (int e = threadIdx.x; e < DB * DB; e += 256) M[(o + (e >> 6)) * np + o + (e & 63)] = P[e];
        return;
    }
    __shared__ S As[DB * TP], Bs[DB * TP];
    load_tile(As, P, DB, 0, 0);
    load_tile(Bs, M, np, o, (long)J * DB);
    __syncthreads();
    typename Mma<S>::acc4 acc[2][2];
    zero_acc<S>(acc);
    tile_mma(As, Bs, acc);
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int v = 0; v < 4; ++v) {
                int r, cc;
                acc_pos<S>(a, b, v, &r, &cc);
                M[(o + r) * np + (long)J * DB + cc] = acc[a][b][v];
            }
    const S* Cb = C + (long)J * DB * DB;
    for (int e = threadIdx.x; e < DB * DB; e += 256) M[((long)J * DB + (e >> 6)) * np + o + (e & 63)] = -Cb[e];
}

// ---- the per-V-cycle solve --------------------------------------------------------------
// y = M x (acc = 0) or y += M x (acc = 1): one row per wavefront, 16-byte loads of the
// matrix and of x (both in the matrix scalar), partial sums per lane in that scalar,
// FP64 across lanes.
template <typename S>
__global__ __launch_bounds__(256) void k_dense_gemv(const S* M, int np, const S* x, double* y, int acc,
                                                    const int* stop) {
    if (*stop) return;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + w;
    S s = S(0);
    if (sizeof(S) == 4) {
        const float4* row = reinterpret_cast<const float4*>(M + (long)r * np);
        const float4* xv = reinterpret_cast<const float4*>(x);
        for (int q = l; q < np / 4; q += 64) {
            const float4 a = row[q], b = xv[q];
            s += a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
        }
    } else {
        const double2* row = reinterpret_cast<const double2*>(M + (long)r * np);
        const double2* xv = reinterpret_cast<const double2*>(x);
        for (int q = l; q < np / 2; q += 64) {
            const double2 a = row[q], b = xv[q];
            s += a.x * b.x + a.y * b.y;
        }
    }
    const double d = hsd::wave_sum((double)s);
    if (l == 0) y[r] = acc ? y[r] + d : d;
}

// r = level right-hand side in natural order (inactive / padding rows 0)
template <typename T>
__global__ void k_dense_gather(Lv<T> L, double* r, T* xf, int np, const SolveCtl* ctl) {
    if (ctl->done) return;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const int n = L.nx * L.ny * L.nz;
    double v = 0.0;
    if (q < n) {
        const int i = 1 + q % L.nx, j = 1 + (q / L.nx) % L.ny, k = 1 + q / (L.nx * L.ny);
        const long c = i + L.sy * j + L.sz * k;
        if (mgl::diag_of(L, c) > T(0)) v = (double)L.b[c];
    }
    r[q] = v;
    xf[q] = (T)v;
}

// e = r - A z in FP64 with the level's coefficients (inactive / padding rows 0), stored
// as the input of the next inverse apply
template <typename T>
__global__ void k_dense_resid(Lv<T> L, const double* z, const double* r, T* e, int np, const SolveCtl* ctl) {
    if (ctl->done) return;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const int n = L.nx * L.ny * L.nz;
    double v = 0.0;
    if (q < n) {
        const int i = 1 + q % L.nx, j = 1 + (q / L.nx) % L.ny, k = 1 + q / (L.nx * L.ny);
        const long c = i + L.sy * j + L.sz * k;
        const double D = (double)mgl::diag_of(L, c);
        if (D > 0.0) {
            const int sy = L.nx, sz = L.nx * L.ny;
            double off = 0.0;
            if (i > 1) off += (double)L.cw[c] * z[q - 1];
            if (i < L.nx) off += (double)L.cw[c + 1] * z[q + 1];
            if (j > 1) off += (double)L.cs[c] * z[q - sy];
            if (j < L.ny) off += (double)L.cs[c + L.sy] * z[q + sy];
            if (k > 1) off += (double)L.cb[c] * z[q - sz];
            if (k < L.nz) off += (double)L.cb[c + L.sz] * z[q + sz];
            v = r[q] - (D * z[q] - off);
        }
    }
    e[q] = (T)v;
}

// level iterate = z on active cells (the V-cycle's coarse correction)
template <typename T>
__global__ void k_dense_scatter(Lv<T> L, const double* z, const SolveCtl* ctl) {
    if (ctl->done) return;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = L.nx * L.ny * L.nz;
    if (q >= n) return;
    const int i = 1 + q % L.nx, j = 1 + (q / L.nx) % L.ny, k = 1 + q / (L.nx * L.ny);
    const long c = i + L.sy * j + L.sz * k;
    L.x[c] = (mgl::diag_of(L, c) > T(0)) ? (T)z[q] : T(0);
}

// in-place blocked Gauss-Jordan inverse of the assembled matrix D->M
template <typename T>
int gj_invert(hs_ctx* c, DenseCoarse* D) {
    const int np = D->np;
    T* M = static_cast<T*>(D->M);
    T* P = static_cast<T*>(D->P);
    T* C = static_cast<T*>(D->C);
    const int nb = np / DB;
    for (int kb = 0; kb < nb; ++kb) {
        k_gj_diag<T><<<1, 256, 0, c->stream>>>(M, np, kb, P);
        k_gj_panel<T><<<nb, 256, 0, c->stream>>>(M, np, kb, P, C);
        k_gj_update<T><<<nb * nb, 256, 0, c->stream>>>(M, np, kb, C);
        k_gj_finish<T><<<nb, 256, 0, c->stream>>>(M, np, kb, P, C);
    }
    HS_CHECK(hipGetLastError());
    return HS_OK;
}

// the inverse in the level's own scalar T
template <typename T>
int dense_factor_t(hs_ctx* c, const MGLevel* m) {
    DenseCoarse* D = c->dense;
    const int np = D->np;
    const long nn = (long)np * np;
    T* M = static_cast<T*>(D->M);
    k_dense_fill<T><<<2048, 256, 0, c->stream>>>(M, nn);
    HS_CHECK(hipGetLastError());
    k_dense_assemble<T><<<(np + 255) / 256, 256, 0, c->stream>>>(mgl::lv<T>(m), M, np);
    HS_CHECK(hipGetLastError());
    return gj_invert<T>(c, D);
}

template <typename T>
int dense_solve_t(hs_ctx* c, const MGLevel* m) {
    DenseCoarse* D = c->dense;
    const int np = D->np;
    const Lv<T> L = mgl::lv<T>(m);
    const T* M = static_cast<const T*>(D->M);
    T* xf = static_cast<T*>(D->xf);
    const int gb = (np + 255) / 256;
    k_dense_gather<T><<<gb, 256, 0, c->stream>>>(L, D->r, xf, np, c->sctl);
    k_dense_gemv<T><<<np / 4, 256, 0, c->stream>>>(M, np, xf, D->z, 0, &c->sctl->done);
    for (int s = 0; s < c->prm.mg_dense_refine; ++s) {
        k_dense_resid<T><<<gb, 256, 0, c->stream>>>(L, D->z, D->r, xf, np, c->sctl);
        k_dense_gemv<T><<<np / 4, 256, 0, c->stream>>>(M, np, xf, D->z, 1, &c->sctl->done);
    }
    k_dense_scatter<T><<<gb, 256, 0, c->stream>>>(L, D->z, c->sctl);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}


// ---- CSR levels (aggregation multigrid) -------------------------------------------------
// row r: diagonal dd[r], off-diagonal entries vd[rp[r] .. rp[r+1]) in columns ci; rows n..np
// of the padding are identity rows
template <typename T>
__global__ void k_dense_assemble_csr(int n, const int* rp, const int* ci, const double* vd, const double* dd, T* M,
                                     int np) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= np) return;
    T* row = M + (long)r * np;
    if (r >= n) { row[r] = T(1); return; }
    row[r] = (T)dd[r];
    for (int t = rp[r]; t < rp[r + 1]; ++t) row[ci[t]] += (T)vd[t];
}

// r = b (FP64), xf = b in the inverse's scalar; padding rows 0
template <typename T>
__global__ void k_dense_gather_csr(int n, const T* b, double* r, T* xf, int np, const int* skip) {
    if (*skip) return;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const double v = q < n ? (double)b[q] : 0.0;
    r[q] = v;
    xf[q] = (T)v;
}

// e = r - A z in FP64 with the level's FP64 entries
template <typename T>
__global__ void k_dense_resid_csr(int n, const int* rp, const int* ci, const double* vd, const double* dd,
                                  const double* z, const double* r, T* e, int np, const int* skip) {
    if (*skip) return;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    double v = 0.0;
    if (q < n) {
        double az = dd[q] * z[q];
        for (int t = rp[q]; t < rp[q + 1]; ++t) az += vd[t] * z[ci[t]];
        v = r[q] - az;
    }
    e[q] = (T)v;
}

template <typename T>
__global__ void k_dense_scatter_csr(int n, const double* z, T* x, const int* skip) {
    if (*skip) return;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) x[q] = (T)z[q];
}

int dense_alloc(DenseCoarse** Dp, int np, bool f32) {
    const size_t es = f32 ? sizeof(float) : sizeof(double);
    DenseCoarse* D = *Dp;
    if (D && (D->np != np || D->f64 == f32)) dense_release(Dp), D = nullptr;
    if (D) return HS_OK;
    D = new DenseCoarse();
    *Dp = D;
    D->np = np;
    D->f64 = !f32;
    if (hipMalloc(&D->M, (size_t)np * np * es) != hipSuccess) return HS_E_NOMEM;
    if (hipMalloc(&D->C, (size_t)np * DB * es) != hipSuccess) return HS_E_NOMEM;
    if (hipMalloc(&D->P, (size_t)DB * DB * es) != hipSuccess) return HS_E_NOMEM;
    double** v[2] = {&D->r, &D->z};
    for (double** a : v)
        if (hipMalloc(a, (size_t)np * sizeof(double)) != hipSuccess) return HS_E_NOMEM;
    if (hipMalloc(&D->xf, (size_t)np * es) != hipSuccess) return HS_E_NOMEM;
    return HS_OK;
}

template <typename T>
int dense_factor_csr_t(hs_ctx* c, DenseCoarse* D, int n, const int* rp, const int* ci, const double* vd,
                       const double* dd) {
    const int np = D->np;
    T* M = static_cast<T*>(D->M);
    k_dense_fill<T><<<2048, 256, 0, c->stream>>>(M, (long)np * np);
    HS_CHECK(hipGetLastError());
    k_dense_assemble_csr<T><<<(np + 255) / 256, 256, 0, c->stream>>>(n, rp, ci, vd, dd, M, np);
    HS_CHECK(hipGetLastError());
    return gj_invert<T>(c, D);
}

template <typename T>
int dense_solve_csr_t(hs_ctx* c, DenseCoarse* D, int n, const int* rp, const int* ci, const double* vd,
                      const double* dd, const T* b, T* x, int refine, const int* skip) {
    const int np = D->np;
    const T* M = static_cast<const T*>(D->M);
    T* xf = static_cast<T*>(D->xf);
    const int gb = (np + 255) / 256;
    k_dense_gather_csr<T><<<gb, 256, 0, c->stream>>>(n, b, D->r, xf, np, skip);
    k_dense_gemv<T><<<np / 4, 256, 0, c->stream>>>(M, np, xf, D->z, 0, skip);
    for (int s = 0; s < refine; ++s) {
        k_dense_resid_csr<T><<<gb, 256, 0, c->stream>>>(n, rp, ci, vd, dd, D->z, D->r, xf, np, skip);
        k_dense_gemv<T><<<np / 4, 256, 0, c->stream>>>(M, np, xf, D->z, 1, skip);
    }
    k_dense_scatter_csr<T><<<gb, 256, 0, c->stream>>>(n, D->z, x, skip);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}

}  // namespace

int dense_factor_csr(hs_ctx* c, DenseCoarse** D, int n, const int* rp, const int* ci, const double* vd,
                     const double* dd, bool f32) {
    if (n < 1 || n > DENSE_MAX_CELLS) return HS_E_ARG;
    HS_TRY(dense_alloc(D, (n + DB - 1) / DB * DB, f32));
    (*D)->factorizations += 1;
    return f32 ? dense_factor_csr_t<float>(c, *D, n, rp, ci, vd, dd) : dense_factor_csr_t<double>(c, *D, n, rp, ci, vd, dd);
}

int dense_solve_csr(hs_ctx* c, DenseCoarse* D, int n, const int* rp, const int* ci, const double* vd,
                    const double* dd, const void* b, void* x, bool f32, int refine, const int* skip) {
    if (!D) return HS_E_STATE;
    return f32 ? dense_solve_csr_t<float>(c, D, n, rp, ci, vd, dd, (const float*)b, (float*)x, refine, skip)
               : dense_solve_csr_t<double>(c, D, n, rp, ci, vd, dd, (const double*)b, (double*)x, refine, skip);
}

void dense_release(DenseCoarse** Dp) {
    DenseCoarse* D = *Dp;
    if (!D) return;
    void* a[6] = {D->M, D->C, D->P, D->r, D->z, D->xf};
    for (void* p : a) if (p) (void)hipFree(p);
    delete D;
    *Dp = nullptr;
}

int dense_max_cells() { return DENSE_MAX_CELLS; }

// (re)build the inverse of the coarsest level when it is due
int dense_setup(hs_ctx* c, const MGLevel* m, bool f32) {
    const long n = (long)m->nx * m->ny * m->nz;
    if (n > DENSE_MAX_CELLS) return HS_E_ARG;
    const int np = (int)((n + DB - 1) / DB) * DB;
    const size_t es = f32 ? sizeof(float) : sizeof(double);
    if (!c->dense) {
        DenseCoarse* D = new DenseCoarse();
        c->dense = D;
        D->np = np;
        D->f64 = !f32;
        if (hipMalloc(&D->M, (size_t)np * np * es) != hipSuccess) return HS_E_NOMEM;
        if (hipMalloc(&D->C, (size_t)np * DB * es) != hipSuccess) return HS_E_NOMEM;
        if (hipMalloc(&D->P, (size_t)DB * DB * es) != hipSuccess) return HS_E_NOMEM;
        double** v[2] = {&D->r, &D->z};
        for (double** a : v)
            if (hipMalloc(a, (size_t)np * sizeof(double)) != hipSuccess) return HS_E_NOMEM;
        if (hipMalloc(&D->xf, (size_t)np * es) != hipSuccess) return HS_E_NOMEM;
    }
    DenseCoarse* D = c->dense;
    if (D->np != np || D->f64 == f32) return HS_E_STATE;
    // Refactor on the first build, once `mg_dense_coarse` builds have passed, or sooner
    // when the last solve took 2 or more FCG iterations more than the first solve after
    // the last factorization.  Between factorizations the refinement (if any) runs against
    // the current level operator, so a stale inverse only weakens the coarse correction,
    // and the iteration count shows when it does.
    const int every = std::max(1, c->prm.mg_dense_coarse);
    const bool stale = D->iters_ref >= 0 && c->last_iters >= D->iters_ref + 2;
    if (D->factorizations == 0 || D->since >= every || stale) {
        HS_TRY(f32 ? dense_factor_t<float>(c, m) : dense_factor_t<double>(c, m));
        D->factorizations += 1;
        D->since = 0;
        D->iters_ref = -1;
    }
    D->since += 1;
    D->builds += 1;
    return HS_OK;
}

int dense_solve(hs_ctx* c, const MGLevel* m, bool f32) {
    return f32 ? dense_solve_t<float>(c, m) : dense_solve_t<double>(c, m);
}

void dense_free(hs_ctx* c) { dense_release(&c->dense); }
