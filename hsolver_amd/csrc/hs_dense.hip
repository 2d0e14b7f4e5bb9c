// hs_dense.hip — direct solve of the coarsest multigrid level on the matrix cores.
//
// The coarsest level (at most DENSE_MAX_CELLS cells, 16^3 under a 256^3 grid) is
// assembled as a dense FP32 matrix in natural cell order (rows of inactive cells and of
// the padding to a multiple of 64 are identity rows) and inverted in place by blocked
// Gauss-Jordan elimination without pivoting (the level operator is symmetric positive
// definite).  For every 64-wide block pivot k:
//   P    = A_kk^-1                         one wavefront, rows in registers
//   C_I  = A_Ik P              (I != k)    MFMA tile products
//   A_IJ = A_IJ - C_I A_kJ     (I, J != k) MFMA rank-64 update of the whole matrix
//   A_kJ = P A_kJ,  A_Ik = -C_I,  A_kk = P
// All products run on v_mfma_f32_16x16x4f32 (64x64 output tile per workgroup, four
// wavefronts in a 2x2 arrangement of 32x32, K = 64 from LDS).  The inverse is kept
// between hierarchy builds and refreshed at most every `mg_dense_coarse` builds (sooner
// when the flexible CG's iteration count shows the inverse has gone stale).
//
// Each V-cycle then solves the coarse system with the inverse and FP64 iterative
// refinement against the level's own stencil:
//   z = A^-1 r;  repeat: z += A^-1 (r - A z)      (residual in FP64)
// The inverse apply is a full-chip GEMV (FP32 matrix and input, FP64 sums across lanes):
// each refinement step shrinks the error by the inverse's FP32 accuracy.
#include <cmath>

#include "hs_device.h"
#include "hs_mglevel.h"

using mgl::Lv;

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int DB = 64;          // Gauss-Jordan block and MFMA output tile
constexpr int TP = DB + 1;      // LDS pitch (odd: column reads of 16 rows hit 16 banks)
constexpr int DENSE_MAX_CELLS = 4096;

// ---- 64x64x64 tile product on the matrix cores -------------------------------------
// acc += As (64x64, LDS pitch TP) * Bs (64x64, LDS pitch TP).  Wave w owns the 32x32
// output block at rows 32*(w>>1), cols 32*(w&1), as 2x2 tiles of 16x16.  Lane l of a
// v_mfma_f32_16x16x4f32 supplies A[l%16][l/16] and B[l/16][l%16] and receives
// D[4*(l/16)+v][l%16] in element v of its accumulator.
__device__ __forceinline__ void tile_mma(const float* As, const float* Bs, f4 acc[2][2]) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int r0 = 32 * (w >> 1), c0 = 32 * (w & 1);
    const int lr = l & 15, lk = l >> 4;
#pragma unroll 4
    for (int k0 = 0; k0 < DB; k0 += 4) {
        const float a0 = As[(r0 + lr) * TP + k0 + lk];
        const float a1 = As[(r0 + 16 + lr) * TP + k0 + lk];
        const float b0 = Bs[(k0 + lk) * TP + c0 + lr];
        const float b1 = Bs[(k0 + lk) * TP + c0 + 16 + lr];
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
    }
}

// (row, col) inside the 64x64 tile of accumulator element (ti, tj, v) of this lane
__device__ __forceinline__ void acc_pos(int ti, int tj, int v, int* row, int* col) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    *row = 32 * (w >> 1) + 16 * ti + 4 * (l >> 4) + v;
    *col = 32 * (w & 1) + 16 * tj + (l & 15);
}

// 64x64 block at (r0, c0) of a row-major matrix with leading dimension ld -> LDS
__device__ __forceinline__ void load_tile(float* S, const float* G, long ld, long r0, long c0) {
    for (int e = threadIdx.x; e < DB * DB / 4; e += 256) {
        const int r = e >> 4, q = (e & 15) * 4;
        const float4 v = *reinterpret_cast<const float4*>(G + (r0 + r) * ld + c0 + q);
        S[r * TP + q] = v.x; S[r * TP + q + 1] = v.y; S[r * TP + q + 2] = v.z; S[r * TP + q + 3] = v.w;
    }
}

__device__ __forceinline__ void zero_acc(f4 acc[2][2]) {
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};
}

// ---- assembly ---------------------------------------------------------------------
__global__ void k_dense_fill(float* M, long n) {
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) M[t] = 0.f;
}

// row r = cell (i,j,k) of the level in natural order; identity for inactive / padding rows
template <typename T>
__global__ void k_dense_assemble(Lv<T> L, float* M, int np) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= np) return;
    float* row = M + (long)r * np;
    const int n = L.nx * L.ny * L.nz;
    if (r >= n) { row[r] = 1.f; return; }
    const int i = 1 + r % L.nx, j = 1 + (r / L.nx) % L.ny, k = 1 + r / (L.nx * L.ny);
    const long c = i + L.sy * j + L.sz * k;
    const T D = mgl::diag_of(L, c);
    if (!(D > T(0))) { row[r] = 1.f; return; }
    row[r] = (float)D;
    const int sx = 1, sy = L.nx, sz = L.nx * L.ny;
    if (i > 1) row[r - sx] = -(float)L.cw[c];
    if (i < L.nx) row[r + sx] = -(float)L.cw[c + 1];
    if (j > 1) row[r - sy] = -(float)L.cs[c];
    if (j < L.ny) row[r + sy] = -(float)L.cs[c + L.sy];
    if (k > 1) row[r - sz] = -(float)L.cb[c];
    if (k < L.nz) row[r + sz] = -(float)L.cb[c + L.sz];
}

// ---- blocked Gauss-Jordan steps ------------------------------------------------------
// P = inverse of the 64x64 diagonal block kb: Gauss-Jordan by one wavefront on the block
// in LDS, lane i owning row i (16-byte aligned rows, float4 accesses, the sweep over a
// row unrolled so its LDS traffic is issued back to back).  A wavefront's LDS reads and
// writes complete in program order, so the pivot row (read by every lane, rewritten by
// lane p in the same sweep) and each lane's own row need no barrier between pivots.
constexpr int TQ = DB + 4;      // row pitch of the pivot block in LDS (floats)
__global__ __launch_bounds__(64) void k_gj_diag(const float* M, int np, int kb, float* P) {
    __shared__ float4 S4[DB * TQ / 4];
    float* S = reinterpret_cast<float*>(S4);
    const int i = threadIdx.x;
    const long o = (long)kb * DB;
    const float4* src = reinterpret_cast<const float4*>(M + (o + i) * np + o);
    float4* my = S4 + i * (TQ / 4);
#pragma unroll
    for (int q = 0; q < DB / 4; ++q) my[q] = src[q];
    __syncthreads();
#pragma clang loop unroll(disable)
    for (int p = 0; p < DB; ++p) {
        const float4* pr = S4 + p * (TQ / 4);
        const float piv = 1.f / S[p * TQ + p];
        const float f = S[i * TQ + p] * piv;     // a_ip / a_pp (this lane's row)
        const bool own = i == p;
        float4 pv[DB / 4], mv[DB / 4];
#pragma unroll
        for (int q = 0; q < DB / 4; ++q) {
            pv[q] = pr[q];
            mv[q] = my[q];
        }
#pragma unroll
        for (int q = 0; q < DB / 4; ++q) {
            float4 v;
            v.x = own ? pv[q].x * piv : mv[q].x - f * pv[q].x;
            v.y = own ? pv[q].y * piv : mv[q].y - f * pv[q].y;
            v.z = own ? pv[q].z * piv : mv[q].z - f * pv[q].z;
            v.w = own ? pv[q].w * piv : mv[q].w - f * pv[q].w;
            my[q] = v;
        }
        S[i * TQ + p] = own ? piv : -f;
    }
    __syncthreads();
    float4* dst = reinterpret_cast<float4*>(P + i * DB);
#pragma unroll
    for (int q = 0; q < DB / 4; ++q) dst[q] = my[q];
}

// C_I = A_{I,kb} P for every row block I != kb (C stored as np/64 blocks of 64x64)
__global__ __launch_bounds__(256) void k_gj_panel(const float* M, int np, int kb, const float* P, float* C) {
    const int I = blockIdx.x;
    if (I == kb) return;
    __shared__ float As[DB * TP], Bs[DB * TP];
    load_tile(As, M, np, (long)I * DB, (long)kb * DB);
    load_tile(Bs, P, DB, 0, 0);
    __syncthreads();
    f4 acc[2][2];
    zero_acc(acc);
    tile_mma(As, Bs, acc);
    float* Cb = C + (long)I * DB * DB;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int v = 0; v < 4; ++v) {
                int r, cc;
                acc_pos(a, b, v, &r, &cc);
                Cb[r * DB + cc] = acc[a][b][v];
            }
}

// A_IJ -= C_I A_{kb,J} for I, J != kb
__global__ __launch_bounds__(256) void k_gj_update(float* M, int np, int kb, const float* C) {
    const int nb = np / DB;
    const int I = blockIdx.x % nb, J = blockIdx.x / nb;
    if (I == kb || J == kb) return;
    __shared__ float As[DB * TP], Bs[DB * TP];
    load_tile(As, C + (long)I * DB * DB, DB, 0, 0);
    load_tile(Bs, M, np, (long)kb * DB, (long)J * DB);
    __syncthreads();
    f4 acc[2][2];
    zero_acc(acc);
    tile_mma(As, Bs, acc);
    float* T = M + (long)I * DB * np + (long)J * DB;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int v = 0; v < 4; ++v) {
                int r, cc;
                acc_pos(a, b, v, &r, &cc);
                T[(long)r * np + cc] -= acc[a][b][v];
            }
}

// A_{kb,J} = P A_{kb,J},  A_{J,kb} = -C_J  (J != kb);  A_{kb,kb} = P
__global__ __launch_bounds__(256) void k_gj_finish(float* M, int np, int kb, const float* P, const float* C) {
    const int J = blockIdx.x;
    const long o = (long)kb * DB;
    if (J == kb) {
        for (int e = threadIdx.x; e < DB * DB; e += 256) M[(o + (e >> 6)) * np + o + (e & 63)] = P[e];
        return;
    }
    __shared__ float As[DB * TP], Bs[DB * TP];
    load_tile(As, P, DB, 0, 0);
    load_tile(Bs, M, np, o, (long)J * DB);
    __syncthreads();
    f4 acc[2][2];
    zero_acc(acc);
    tile_mma(As, Bs, acc);
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int v = 0; v < 4; ++v) {
                int r, cc;
                acc_pos(a, b, v, &r, &cc);
                M[(o + r) * np + (long)J * DB + cc] = acc[a][b][v];
            }
    const float* Cb = C + (long)J * DB * DB;
    for (int e = threadIdx.x; e < DB * DB; e += 256) M[((long)J * DB + (e >> 6)) * np + o + (e & 63)] = -Cb[e];
}

// ---- the per-V-cycle solve --------------------------------------------------------------
// y = M x (acc = 0) or y += M x (acc = 1): one row per wavefront, float4 loads of the
// FP32 matrix and of the FP32 copy of x, FP32 partial sums per lane, FP64 across lanes.
__global__ __launch_bounds__(256) void k_dense_gemv(const float* M, int np, const float* x, double* y, int acc,
                                                    const SolveCtl* ctl) {
    if (ctl->done) return;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + w;
    const float4* row = reinterpret_cast<const float4*>(M + (long)r * np);
    const float4* xv = reinterpret_cast<const float4*>(x);
    float s = 0.f;
    for (int q = l; q < np / 4; q += 64) {
        const float4 a = row[q], b = xv[q];
        s += a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
    }
    const double d = hsd::wave_sum((double)s);
    if (l == 0) y[r] = acc ? y[r] + d : d;
}

// r = level right-hand side in natural order (inactive / padding rows 0)
template <typename T>
__global__ void k_dense_gather(Lv<T> L, double* r, float* xf, int np, const SolveCtl* ctl) {
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
    xf[q] = (float)v;
}

// e = r - A z in FP64 with the level's coefficients (inactive / padding rows 0), stored
// as the FP32 input of the next inverse apply
template <typename T>
__global__ void k_dense_resid(Lv<T> L, const double* z, const double* r, float* e, int np, const SolveCtl* ctl) {
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
    e[q] = (float)v;
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

template <typename T>
int dense_factor_t(hs_ctx* c, const MGLevel* m) {
    DenseCoarse* D = c->dense;
    const int np = D->np;
    const long nn = (long)np * np;
    k_dense_fill<<<2048, 256, 0, c->stream>>>(D->M, nn);
    HS_CHECK(hipGetLastError());
    k_dense_assemble<T><<<(np + 255) / 256, 256, 0, c->stream>>>(mgl::lv<T>(m), D->M, np);
    HS_CHECK(hipGetLastError());
    const int nb = np / DB;
    for (int kb = 0; kb < nb; ++kb) {
        k_gj_diag<<<1, 64, 0, c->stream>>>(D->M, np, kb, D->P);
        k_gj_panel<<<nb, 256, 0, c->stream>>>(D->M, np, kb, D->P, D->C);
        k_gj_update<<<nb * nb, 256, 0, c->stream>>>(D->M, np, kb, D->C);
        k_gj_finish<<<nb, 256, 0, c->stream>>>(D->M, np, kb, D->P, D->C);
    }
    HS_CHECK(hipGetLastError());
    return HS_OK;
}

template <typename T>
int dense_solve_t(hs_ctx* c, const MGLevel* m) {
    DenseCoarse* D = c->dense;
    const int np = D->np;
    const Lv<T> L = mgl::lv<T>(m);
    const int gb = (np + 255) / 256;
    k_dense_gather<T><<<gb, 256, 0, c->stream>>>(L, D->r, D->xf, np, c->sctl);
    k_dense_gemv<<<np / 4, 256, 0, c->stream>>>(D->M, np, D->xf, D->z, 0, c->sctl);
    for (int s = 0; s < c->prm.mg_dense_refine; ++s) {
        k_dense_resid<T><<<gb, 256, 0, c->stream>>>(L, D->z, D->r, D->xf, np, c->sctl);
        k_dense_gemv<<<np / 4, 256, 0, c->stream>>>(D->M, np, D->xf, D->z, 1, c->sctl);
    }
    k_dense_scatter<T><<<gb, 256, 0, c->stream>>>(L, D->z, c->sctl);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}

}  // namespace

int dense_max_cells() { return DENSE_MAX_CELLS; }

// (re)build the inverse of the coarsest level when it is due
int dense_setup(hs_ctx* c, const MGLevel* m, bool f32) {
    const long n = (long)m->nx * m->ny * m->nz;
    if (n > DENSE_MAX_CELLS) return HS_E_ARG;
    const int np = (int)((n + DB - 1) / DB) * DB;
    if (!c->dense) {
        DenseCoarse* D = new DenseCoarse();
        c->dense = D;
        D->np = np;
        if (hipMalloc(&D->M, (size_t)np * np * sizeof(float)) != hipSuccess) return HS_E_NOMEM;
        if (hipMalloc(&D->C, (size_t)np * DB * sizeof(float)) != hipSuccess) return HS_E_NOMEM;
        if (hipMalloc(&D->P, (size_t)DB * DB * sizeof(float)) != hipSuccess) return HS_E_NOMEM;
        double** v[2] = {&D->r, &D->z};
        for (double** a : v)
            if (hipMalloc(a, (size_t)np * sizeof(double)) != hipSuccess) return HS_E_NOMEM;
        if (hipMalloc(&D->xf, (size_t)np * sizeof(float)) != hipSuccess) return HS_E_NOMEM;
    }
    DenseCoarse* D = c->dense;
    if (D->np != np) return HS_E_STATE;
    // Refactor on the first build, once `mg_dense_coarse` builds have passed, or sooner
    // when the last solve took 2 or more FCG iterations more than the first solve after
    // the last factorization.  Between factorizations the FP64 refinement runs against
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

void dense_free(hs_ctx* c) {
    DenseCoarse* D = c->dense;
    if (!D) return;
    void* a[6] = {D->M, D->C, D->P, D->r, D->z, D->xf};
    for (void* p : a) if (p) (void)hipFree(p);
    delete D;
    c->dense = nullptr;
}
