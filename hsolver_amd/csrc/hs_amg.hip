// hs_amg.hip — aggregation (algebraic) multigrid for the p' solve and the energy solves.
//
// The reference wires HYPRE BoomerAMG as the PCG preconditioner and leaves it commented
// out (fluid_solver.h:16, 1775-1787, 1876-1888).  This fills that slot with an algebraic
// multigrid designed for one GPU:
//
//   finest level   the structured 7-point operator the geometric hierarchies already
//                  hold (the pressure multigrid's level 0, or the energy hierarchy's
//                  level 0 as A or A^T), smoothed by their red-black Gauss-Seidel sweeps;
//   aggregation    two passes of pairwise matching per level, so an aggregate holds up
//                  to four rows of the level above.  A pass runs eight handshake rounds
//                  in which hashed colours split the unmatched rows into proposers and
//                  acceptors: a proposer asks its strongest unmatched acceptor neighbour,
//                  an acceptor takes the strongest proposer (rows left over stay single).
//                  Strength of the pair (i, j): the symmetrised coupling -(a_ij + a_ji)/2
//                  over sqrt(a_ii a_jj), and only couplings of at least theta times the
//                  row's largest are candidates.  Scaling by the diagonals keeps large
//                  aggregates on the coarse levels from attracting every neighbour;
//   coarse levels  Galerkin operators for piecewise-constant prolongation, A_c = P^T A P:
//                  the sums of the entries between two aggregates, those inside an
//                  aggregate added to its diagonal.  One thread forms one coarse row from
//                  its members' rows (columns in order of first appearance), so the
//                  hierarchy is the same on every run.  CSR, FP64 for the setup and the
//                  coarsest solve's refinement, the cycle's scalar for the cycle;
//   cycle          V-cycle: Jacobi (weight 2/3) on the aggregated levels, the prolongated
//                  correction scaled by alpha, and once a level has at most 4096 rows the
//                  dense inverse of hs_dense.hip (matrix cores) with FP64 refinement.
//
// Why it helps where the geometric hierarchy struggles: a highly conducting region (a
// fin at k_s/k_f = 1000) couples about a thousand times more strongly inside itself than
// to the fluid around it, so its rows pair only with each other and the region stays a
// coarse unknown of its own down to the coarse levels; 2x2x2 coarsening merges thin fins
// with the fluid between them and loses the modes in which neighbouring fins differ.
// Pairs also follow the strong direction of stretched cells.
//
// The matching, numbering and Galerkin products all run on the GPU; the host reads back
// sizes (a few ints per pass) to allocate the next level.  The p' operator changes every
// SIMPLE iteration: its aggregates (and the coarsest inverse) are kept and only the
// Galerkin values are recomputed, until the dense-inverse policy of the geometric path
// (mg_dense_coarse builds, or an FCG iteration rise of 2) asks for a fresh aggregation.
#include <cmath>
#include <cstring>

#include "hs_device.h"
#include "hs_eop.h"
#include "hs_mglevel.h"

using hsd::FLAT;
using hsd::FLAT_BLOCKS;

namespace {

constexpr int ROUNDS = 8;          // handshake rounds of one pairwise pass
constexpr int MAXW = 128;          // distinct coarse columns of one Galerkin row
constexpr int SL = 64;             // rows per slice of the sliced ELL format
constexpr int GAL_T = 64;          // threads of a Galerkin block (one row each)
constexpr int AMG_MAX_LEVELS = 24;
constexpr int SCAN_T = 256, SCAN_ITEMS = 4, SCAN_BLOCK = SCAN_T * SCAN_ITEMS;

#define G_STRIDE(i, n) for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < (n); i += (long)gridDim.x * blockDim.x)

// ---------------------------------------------------------- finest-level views ---
// Entries of row c: diagonal diag(c) and couplings a[d] >= 0 to the neighbour c + off(d)
// (the matrix entry is -a[d]), d = W, E, S, N, B, T; opposite directions are d ^ 1.
template <typename T>
struct PView {   // p' operator on the pressure multigrid's level 0 (symmetric)
    mgl::Lv<T> L;
    __device__ bool active(long c) const { return L.act[c] != 0; }
    __device__ long off(int d) const {
        const long s = d < 2 ? 1 : (d < 4 ? L.sy : L.sz);
        return (d & 1) ? s : -s;
    }
    __device__ double diag(long c) const { return (double)mgl::diag_of(L, c); }
    __device__ void coup(long c, double a[6]) const {
        a[0] = L.cw[c]; a[1] = L.cw[c + 1]; a[2] = L.cs[c]; a[3] = L.cs[c + L.sy]; a[4] = L.cb[c]; a[5] = L.cb[c + L.sz];
    }
    __device__ const T* x() const { return L.x; }
    __device__ const T* b() const { return L.b; }
    // the same tests for a cell (i, j, k) of the interior and its neighbour in direction d
    __device__ bool active_at(long c, int, int, int) const { return L.act[c] != 0; }
    __device__ bool nb_active(long nb, int, int, int, int) const { return L.act[nb] != 0; }
    __device__ int nx() const { return L.nx; }
    __device__ int ny() const { return L.ny; }
    __device__ int nz() const { return L.nz; }
};

template <typename T>
struct EView {   // energy operator A (tr = 0) or A^T (tr = 1) on the energy level 0
    heo::EL<T> L;
    int tr;
    __device__ bool active(long c) const {
        // 32-bit decoding (the padded layout has fewer than 2^32 cells, checked at setup)
        const unsigned cu = (unsigned)c, sy = (unsigned)L.sy, ny2 = (unsigned)(L.ny + 2);
        const unsigned q = cu / sy, i = cu - q * sy, k = q / ny2, j = q - k * ny2;
        return i >= 1 && i <= (unsigned)L.nx && j >= 1 && j <= (unsigned)L.ny && k >= 1 && k <= (unsigned)L.nz;
    }
    __device__ bool active_at(long, int, int, int) const { return true; }   // every interior cell
    __device__ bool nb_active(long, int d, int i, int j, int k) const {
        switch (d) {
            case 0: return i > 1;
            case 1: return i < L.nx;
            case 2: return j > 1;
            case 3: return j < L.ny;
            case 4: return k > 1;
            default: return k < L.nz;
        }
    }
    __device__ int nx() const { return L.nx; }
    __device__ int ny() const { return L.ny; }
    __device__ int nz() const { return L.nz; }
    __device__ long off(int d) const {
        const long s = d < 2 ? 1 : (d < 4 ? L.sy : L.sz);
        return (d & 1) ? s : -s;
    }
    __device__ double diag(long c) const { return (double)heo::lrow(L, c, tr).p; }
    __device__ void coup(long c, double a[6]) const {
        const heo::Row<T> r = heo::lrow(L, c, tr);
        a[0] = r.w; a[1] = r.e; a[2] = r.s; a[3] = r.n; a[4] = r.b; a[5] = r.t;
    }
    __device__ const T* x() const { return L.x; }
    __device__ const T* b() const { return L.b; }
};

// ------------------------------------------------------------ pairwise matching ---
// One round: every unmatched row gets a colour from a hash of (row, round).  A blue row
// proposes to its strongest unmatched red neighbour among its strong ones; a red row
// accepts the strongest blue row that proposed to it, and writes both ends of the pair.
// A blue row proposes to one red row only, so no two red rows write the same entry.
// (Proposals between rows of one colour class would follow the strength gradient of a
// stretched grid row by row and leave long proposal chains unmatched.)
__device__ __forceinline__ bool is_blue(long i, int round) {
    unsigned long long x = (unsigned long long)i * 0x9E3779B97F4A7C15ull ^ (0xD1B54A32D192ED03ull * (unsigned)(round + 1));
    x ^= x >> 31;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 29;
    return (x >> 40) & 1ull;
}

// strengths of row c's six couplings (0 where the neighbour is not an unknown) and their max
template <typename V>
__device__ __forceinline__ double strengths0(const V& v, long c, double s[6]) {
    double a[6];
    v.coup(c, a);
    const double dc = v.diag(c);
    double smax = 0.0;
    for (int d = 0; d < 6; ++d) {
        s[d] = 0.0;
        const long nb = c + v.off(d);
        if (!v.active(nb)) continue;
        double an[6];
        v.coup(nb, an);
        const double sym = 0.5 * (a[d] + an[d ^ 1]), dn = v.diag(nb);
        if (sym > 0.0 && dc > 0.0 && dn > 0.0) s[d] = sym / sqrt(dc * dn);
        smax = fmax(smax, s[d]);
    }
    return smax;
}

template <typename V>
__global__ __launch_bounds__(FLAT) void k_propose0(V v, long n, const int* mate, int* prop, float theta, int round) {
    G_STRIDE(c, n) {
        int best = -1;
        if (v.active(c) && mate[c] < 0 && is_blue(c, round)) {
            double s[6];
            const double smax = strengths0(v, c, s);
            double bs = 0.0;
            for (int d = 0; d < 6; ++d) {
                const long nb = c + v.off(d);
                if (s[d] > 0.0 && s[d] >= (double)theta * smax && mate[nb] < 0 && !is_blue(nb, round) &&
                    (s[d] > bs || (s[d] == bs && nb < best))) {
                    bs = s[d];
                    best = (int)nb;
                }
            }
        }
        prop[c] = best;
    }
}

template <typename V>
__global__ __launch_bounds__(FLAT) void k_respond0(V v, long n, int* mate, const int* prop, int round) {
    G_STRIDE(j, n) {
        if (!v.active(j) || mate[j] >= 0 || is_blue(j, round)) continue;
        double s[6];
        strengths0(v, j, s);
        int best = -1;
        double bs = 0.0;
        for (int d = 0; d < 6; ++d) {
            const long i = j + v.off(d);
            if (s[d] > 0.0 && prop[i] == (int)j && (s[d] > bs || (s[d] == bs && i < best))) {
                bs = s[d];
                best = (int)i;
            }
        }
        if (best >= 0) {
            mate[j] = best;
            mate[best] = (int)j;
        }
    }
}

__device__ __forceinline__ double csr_strength(int i, int t, const int* rp, const int* ci, const double* vd,
                                               const double* dd) {
    const int j = ci[t];
    double aji = 0.0;
    for (int u = rp[j]; u < rp[j + 1]; ++u)
        if (ci[u] == i) { aji = vd[u]; break; }
    const double sym = -0.5 * (vd[t] + aji), di = dd[i], dj = dd[j];
    return (sym > 0.0 && di > 0.0 && dj > 0.0) ? sym / sqrt(di * dj) : 0.0;
}

// strength of every stored entry, once per pass (the rounds then read it); entries of one
// row sit at rp[i] .. rp[i+1] in the order of the matrix
__global__ __launch_bounds__(FLAT) void k_strength_csr(long n, const int* rp, const int* ci, const double* vd,
                                                       const double* dd, double* st) {
    G_STRIDE(i, n) {
        for (int t = rp[i]; t < rp[i + 1]; ++t) st[t] = csr_strength((int)i, t, rp, ci, vd, dd);
    }
}

__global__ __launch_bounds__(FLAT) void k_propose_csr(long n, const int* rp, const int* ci, const double* st,
                                                      const int* mate, int* prop, float theta, int round) {
    G_STRIDE(i, n) {
        int best = -1;
        if (mate[i] < 0 && is_blue(i, round)) {
            double smax = 0.0;
            for (int t = rp[i]; t < rp[i + 1]; ++t) smax = fmax(smax, st[t]);
            double bs = 0.0;
            for (int t = rp[i]; t < rp[i + 1]; ++t) {
                const int j = ci[t];
                if (mate[j] >= 0 || is_blue(j, round)) continue;
                const double s = st[t];
                if (s > 0.0 && s >= (double)theta * smax && (s > bs || (s == bs && j < best))) {
                    bs = s;
                    best = j;
                }
            }
        }
        prop[i] = best;
    }
}

// the matrix pattern is symmetric, so every row that proposed to j is in j's row
__global__ __launch_bounds__(FLAT) void k_respond_csr(long n, const int* rp, const int* ci, const double* st, int* mate,
                                                      const int* prop, int round) {
    G_STRIDE(j, n) {
        if (mate[j] >= 0 || is_blue(j, round)) continue;
        int best = -1;
        double bs = 0.0;
        for (int t = rp[j]; t < rp[j + 1]; ++t) {
            const int i = ci[t];
            if (prop[i] != (int)j) continue;
            const double s = st[t];
            if (s > 0.0 && (s > bs || (s == bs && i < best))) {
                bs = s;
                best = i;
            }
        }
        if (best >= 0) {
            mate[j] = best;
            mate[best] = (int)j;
        }
    }
}

// leaders: rows that are unknowns and either single or the lower row of their pair
template <typename V>
__global__ __launch_bounds__(FLAT) void k_lead0(V v, long n, const int* mate, int* flag) {
    G_STRIDE(c, n) flag[c] = (v.active(c) && (mate[c] < 0 || c < mate[c])) ? 1 : 0;
}
__global__ __launch_bounds__(FLAT) void k_lead_csr(long n, const int* mate, int* flag) {
    G_STRIDE(i, n) flag[i] = (mate[i] < 0 || i < mate[i]) ? 1 : 0;
}

// ------------------------------------------------------------------- prefix sum ---
// exclusive scan of int[n]: per-block scans of 1024 items, one block scans the block
// sums, a third pass adds them; the total goes to *total
__global__ __launch_bounds__(SCAN_T) void k_scan_local(long n, const int* in, int* out, int* bsum) {
    __shared__ int sh[SCAN_T];
    const long base = (long)blockIdx.x * SCAN_BLOCK + (long)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    int tot = 0;
    for (int q = 0; q < SCAN_ITEMS; ++q) {
        v[q] = base + q < n ? in[base + q] : 0;
        tot += v[q];
    }
    sh[threadIdx.x] = tot;
    __syncthreads();
    for (int off = 1; off < SCAN_T; off <<= 1) {
        const int t = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    int ex = sh[threadIdx.x] - tot;
    for (int q = 0; q < SCAN_ITEMS; ++q) {
        if (base + q < n) out[base + q] = ex;
        ex += v[q];
    }
    if (threadIdx.x == SCAN_T - 1) bsum[blockIdx.x] = sh[SCAN_T - 1];
}

__global__ __launch_bounds__(1024) void k_scan_top(int nb, int* bsum, long long* total) {
    __shared__ int sh[1024];
    __shared__ long long carry;   // 64-bit, so that a total beyond the int range is seen
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 1024) {
        const int i = base + threadIdx.x;
        const int v = i < nb ? bsum[i] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int t = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) bsum[i] = (int)(carry + sh[threadIdx.x] - v);
        __syncthreads();
        if (threadIdx.x == 1023) carry += sh[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(FLAT) void k_scan_add(long n, int* out, const int* bsum) {
    G_STRIDE(i, n) out[i] += bsum[i / SCAN_BLOCK];
}

// ---------------------------------------------------------------- numbering ---
// aggregate number of a leader = its position among the leaders; members (2 per aggregate)
__global__ __launch_bounds__(FLAT) void k_number(long n, const int* flag, const int* pos, const int* mate, int* agg,
                                                 int* memb2) {
    G_STRIDE(i, n) {
        if (flag[i]) {
            const int a = pos[i];
            agg[i] = a;
            memb2[2L * a] = (int)i;
            memb2[2L * a + 1] = mate[i];
        }
    }
}
__global__ __launch_bounds__(FLAT) void k_follow(long n, const int* flag, const int* mate, int* agg) {
    G_STRIDE(i, n) {
        if (!flag[i]) agg[i] = mate[i] >= 0 ? agg[mate[i]] : -1;
    }
}

// two pairwise passes composed: rows of the level above -> rows of the new level, and the
// (up to four) members of each new row
__global__ __launch_bounds__(FLAT) void k_compose_agg(long n, const int* aggA, const int* aggB, int* agg) {
    G_STRIDE(r, n) {
        const int a = aggA[r];
        agg[r] = a >= 0 ? aggB[a] : -1;
    }
}
__global__ __launch_bounds__(FLAT) void k_compose_memb(long nc, const int* membB, const int* membA, int* memb4) {
    G_STRIDE(I, nc) {
        for (int q = 0; q < 2; ++q) {
            const int m = membB[2 * I + q];
            memb4[4 * I + 2 * q] = m >= 0 ? membA[2L * m] : -1;
            memb4[4 * I + 2 * q + 1] = m >= 0 ? membA[2L * m + 1] : -1;
        }
    }
}

// ---------------------------------------------------------------- Galerkin rows ---
// Coarse row I from its (up to two) members: count pass (rp == nullptr: cnt[I] = distinct
// columns, flag[0] = 1 if a row would exceed MAXW) or fill pass (columns, entries,
// diagonal at rp[I]).  The distinct columns and their sums sit in LDS, column-major so
// that the 64 threads of a block hit 64 different banks.
struct GalOut {
    int* cnt;
    const int* rp;
    int* ci;
    double* vd;
    double* dd;
    int* flag;
};

template <int W>
__device__ __forceinline__ void gal_add(int (*cols)[GAL_T], double (*vals)[GAL_T], int& m, bool& over, int J,
                                        double e) {
    const int t0 = threadIdx.x;
    int t = 0;
    while (t < m && cols[t][t0] != J) ++t;
    if (t < m) vals[t][t0] += e;
    else if (m < W) { cols[m][t0] = J; vals[m][t0] = e; ++m; }
    else over = true;
}

__device__ __forceinline__ void gal_store(const GalOut& o, long I, int (*cols)[GAL_T], double (*vals)[GAL_T], int m,
                                          bool over, double diag) {
    const int t0 = threadIdx.x;
    if (!o.rp) {
        o.cnt[I] = m;
        if (over) o.flag[0] = 1;
        return;
    }
    // the refills of kept aggregates meet the same columns as the count pass (the pattern
    // follows the aggregation-time map, not the values); never write past the row's slots
    const int at = o.rp[I], cap = o.rp[I + 1] - at;
    for (int t = 0; t < m && t < cap; ++t) {
        o.ci[at + t] = cols[t][t0];
        o.vd[at + t] = vals[t][t0];
    }
    for (int t = m; t < cap; ++t) {   // (only after a change of the pattern) zero entries on the diagonal
        o.ci[at + t] = (int)I;
        o.vd[at + t] = 0.0;
    }
    o.dd[I] = diag;
}

template <typename V, int W>
__global__ __launch_bounds__(GAL_T) void k_gal0(V v, long nc, const int* memb2, const int* agg, GalOut o) {
    __shared__ int cols[W][GAL_T];
    __shared__ double vals[W][GAL_T];
    const long I = blockIdx.x * (long)GAL_T + threadIdx.x;
    if (I >= nc) return;
    int m = 0;
    bool over = false;
    double diag = 0.0;
    for (int q = 0; q < 2; ++q) {
        const int r = memb2[2 * I + q];
        if (r < 0) continue;
        diag += v.diag(r);
        double a[6];
        v.coup(r, a);
        for (int d = 0; d < 6; ++d) {
            // a neighbour is an unknown of the hierarchy iff the aggregation gave it an
            // aggregate: the finest level's own test (D > 0 in the level's scalar) depends on
            // the values, which change between the refills of kept aggregates
            const int J = agg[r + v.off(d)];
            if (J < 0) continue;
            if (J == (int)I) diag -= a[d];
            else gal_add<W>(cols, vals, m, over, J, -a[d]);
        }
    }
    gal_store(o, I, cols, vals, m, over, diag);
}

template <int W>
__global__ __launch_bounds__(GAL_T) void k_gal_csr(long nc, const int* memb2, const int* agg, const int* frp,
                                                   const int* fci, const double* fvd, const double* fdd, GalOut o) {
    __shared__ int cols[W][GAL_T];
    __shared__ double vals[W][GAL_T];
    const long I = blockIdx.x * (long)GAL_T + threadIdx.x;
    if (I >= nc) return;
    int m = 0;
    bool over = false;
    double diag = 0.0;
    for (int q = 0; q < 2; ++q) {
        const int r = memb2[2 * I + q];
        if (r < 0) continue;
        diag += fdd[r];
        for (int t = frp[r]; t < frp[r + 1]; ++t) {
            const int J = agg[fci[t]];
            if (J == (int)I) diag += fvd[t];
            else gal_add<W>(cols, vals, m, over, J, fvd[t]);
        }
    }
    gal_store(o, I, cols, vals, m, over, diag);
}

template <typename T>
__global__ __launch_bounds__(FLAT) void k_narrow(long n, const double* src, T* dst) {
    G_STRIDE(i, n) dst[i] = (T)src[i];
}

// ---------------------------------------------------------------------- cycle ---
// Jacobi from a zero iterate: x = w b / d
template <typename T>
__global__ __launch_bounds__(FLAT) void k_jac0(long n, const T* dg, const T* b, T* x, T w, const int* skip) {
    if (*skip) return;
    G_STRIDE(i, n) x[i] = w * b[i] / dg[i];
}

// sliced ELL: slice widths (the longest row of each 64) and their entry counts
__global__ __launch_bounds__(FLAT) void k_sell_width(long n, const int* rp, int* sw, int* cnt) {
    const long ns = (n + SL - 1) / SL;
    G_STRIDE(s, ns) {
        int w = 0;
        const long i1 = min(n, (s + 1) * SL);
        for (long i = s * SL; i < i1; ++i) w = max(w, rp[i + 1] - rp[i]);
        sw[s] = w;
        cnt[s] = w * SL;
    }
}

// entries of row i at so[s] + k SL + (i mod SL); a row shorter than its slice is padded
// with entries of value 0 on its own column
template <typename T>
__global__ __launch_bounds__(FLAT) void k_sell_fill(long n, const int* rp, const int* ci, const double* vd, const int* so,
                                                    const int* sw, int* sc, T* sv, int pattern) {
    G_STRIDE(i, n) {
        const long s = i / SL;
        const long base = so[s] + (i % SL);
        const int r0 = rp[i], len = rp[i + 1] - r0, w = sw[s];
        for (int k = 0; k < w; ++k) {
            if (pattern) sc[base + (long)k * SL] = k < len ? ci[r0 + k] : (int)i;
            sv[base + (long)k * SL] = k < len ? (T)vd[r0 + k] : T(0);
        }
    }
}

template <typename T>
__device__ __forceinline__ T sell_offsum(long i, const int* so, const int* sw, const int* sc, const T* sv, const T* x) {
    const long s = i / SL;
    const long base = so[s] + (i % SL);
    const int w = sw[s];
    T acc = T(0);
    for (int k = 0; k < w; ++k) acc += sv[base + (long)k * SL] * x[sc[base + (long)k * SL]];
    return acc;
}

// xo = x + w (b - A x) / d
template <typename T>
__global__ __launch_bounds__(FLAT) void k_jac(long n, const int* so, const int* sw, const int* sc, const T* sv,
                                              const T* dg, const T* b, const T* x, T* xo, T w, const int* skip) {
    if (*skip) return;
    G_STRIDE(i, n) {
        const T xi = x[i];
        const T s = b[i] - dg[i] * xi - sell_offsum(i, so, sw, sc, sv, x);
        xo[i] = xi + w * s / dg[i];
    }
}

// r = b - A x
template <typename T>
__global__ __launch_bounds__(FLAT) void k_resid(long n, const int* so, const int* sw, const int* sc, const T* sv,
                                                const T* dg, const T* b, const T* x, T* r, const int* skip) {
    if (*skip) return;
    G_STRIDE(i, n) r[i] = b[i] - dg[i] * x[i] - sell_offsum(i, so, sw, sc, sv, x);
}

// r = b - A x on the interior of the structured finest level (0 on cells that are not
// unknowns; the ghost layer is not written: only aggregate members are read back), one
// thread per cell of an x row, rows by block (y, z)
template <typename V, typename T>
__global__ __launch_bounds__(64) void k_resid0(V v, long sy, long sz, T* r, const int* skip) {
    if (*skip) return;
    const int i = 1 + blockIdx.x * 64 + threadIdx.x, j = 1 + blockIdx.y, k = 1 + blockIdx.z;
    if (i > v.nx()) return;
    const long c = i + sy * j + sz * k;
    const T* x = v.x();
    const T* b = v.b();
    T e = T(0);
    if (v.active_at(c, i, j, k)) {
        double a[6];
        v.coup(c, a);
        e = b[c] - (T)v.diag(c) * x[c];
        for (int d = 0; d < 6; ++d) {
            const long nb = c + v.off(d);
            if (v.nb_active(nb, d, i, j, k)) e += (T)a[d] * x[nb];
        }
    }
    r[c] = e;
}

// bc[I] = sum of r over the members of I
template <typename T>
__global__ __launch_bounds__(FLAT) void k_gather4(long nc, const int* memb4, const T* r, T* bc, const int* skip) {
    if (*skip) return;
    G_STRIDE(I, nc) {
        const int4 m = reinterpret_cast<const int4*>(memb4)[I];
        T s = r[m.x];
        if (m.y >= 0) s += r[m.y];
        if (m.z >= 0) s += r[m.z];
        if (m.w >= 0) s += r[m.w];
        bc[I] = s;
    }
}

// x += alpha xc[agg] (rows without an aggregate unchanged)
template <typename T>
__global__ __launch_bounds__(FLAT) void k_prolong(long n, const int* agg, const T* xc, T* x, T alpha, const int* skip) {
    if (*skip) return;
    G_STRIDE(i, n) {
        const int a = agg[i];
        if (a >= 0) x[i] += alpha * xc[a];
    }
}

// ---------------------------------------------------------------- host side ---
inline int blocks_for(long n) {
    const long b = (n + FLAT - 1) / FLAT;
    return (int)std::max(1L, std::min(b, (long)FLAT_BLOCKS * 4));
}

int amalloc(void** p, size_t bytes) {
    if (bytes == 0) bytes = 4;
    if (hipMalloc(p, bytes) != hipSuccess) return HS_E_NOMEM;
    return HS_OK;
}

void level_free(AmgLevel* L) {
    void* a[] = {L->rp, L->ci, L->vd, L->dd, L->x, L->x2, L->b, L->r, L->agg, L->memb, L->rpi, L->cii, L->membi,
                 L->aggi, L->vdi, L->ddi, L->membf, L->aggb, L->so, L->sw, L->sc, L->sv};
    for (void* p : a) if (p) (void)hipFree(p);
    if (L->dg && L->dg != (void*)L->dd) (void)hipFree(L->dg);
    delete L;
}

void levels_free(AmgHier* H) {
    for (AmgLevel* L : H->lv) level_free(L);
    H->lv.clear();
}

// work ints: mate, prop, flag, pos (n each) and the block sums of the scan
int ensure_scratch(AmgHier* H, long n) {
    const long need = 4 * n + (n + SCAN_BLOCK - 1) / SCAN_BLOCK + 64;
    if (H->scratch_n >= need) return HS_OK;
    if (H->scratch) (void)hipFree(H->scratch);
    H->scratch = nullptr;
    H->scratch_n = 0;
    HS_TRY(amalloc((void**)&H->scratch, need * sizeof(int)));
    H->scratch_n = need;
    return HS_OK;
}

// exclusive scan of in[n] into out[n]; returns the total (synchronises)
int scan(hs_ctx* c, AmgHier* H, long n, const int* in, int* out, int* bsum, long* total) {
    if (n <= 0) {
        *total = 0;
        return HS_OK;
    }
    const int nb = (int)((n + SCAN_BLOCK - 1) / SCAN_BLOCK);
    k_scan_local<<<nb, SCAN_T, 0, c->stream>>>(n, in, out, bsum);
    k_scan_top<<<1, 1024, 0, c->stream>>>(nb, bsum, reinterpret_cast<long long*>(H->dflag + 2));
    k_scan_add<<<blocks_for(n), FLAT, 0, c->stream>>>(n, out, bsum);
    HS_CHECK(hipGetLastError());
    HS_CHECK(hipMemcpyAsync(H->h_flag, H->dflag, 4 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HS_CHECK(hipStreamSynchronize(c->stream));
    long long t;
    std::memcpy(&t, H->h_flag + 2, sizeof t);
    if (t > 2147483647LL) return HS_E_ARG;   // offsets are ints: a level this large is not built
    *total = (long)t;
    return HS_OK;
}

// One pairwise pass over n rows (the finest level through its view V, or a CSR matrix):
// matching, leaders numbered in row order.  Out: nagg, agg[n], memb2[2 nagg] (allocated).
template <typename V>
int pair_pass0(hs_ctx* c, AmgHier* H, const V& v, long n, float theta, long* nagg, int** agg, int** memb2) {
    HS_TRY(ensure_scratch(H, n));
    int *mate = H->scratch, *prop = mate + n, *flag = prop + n, *pos = flag + n, *bsum = pos + n;
    const int nb = blocks_for(n);
    HS_CHECK(hipMemsetAsync(mate, 0xff, n * sizeof(int), c->stream));
    for (int r = 0; r < ROUNDS; ++r) {
        k_propose0<V><<<nb, FLAT, 0, c->stream>>>(v, n, mate, prop, theta, r);
        k_respond0<V><<<nb, FLAT, 0, c->stream>>>(v, n, mate, prop, r);
    }
    k_lead0<V><<<nb, FLAT, 0, c->stream>>>(v, n, mate, flag);
    HS_CHECK(hipGetLastError());
    HS_TRY(scan(c, H, n, flag, pos, bsum, nagg));
    HS_TRY(amalloc((void**)agg, n * sizeof(int)));
    HS_TRY(amalloc((void**)memb2, 2 * std::max(1L, *nagg) * sizeof(int)));
    k_number<<<nb, FLAT, 0, c->stream>>>(n, flag, pos, mate, *agg, *memb2);
    k_follow<<<nb, FLAT, 0, c->stream>>>(n, flag, mate, *agg);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}

int pair_pass_csr(hs_ctx* c, AmgHier* H, long n, long nnz, const int* rp, const int* ci, const double* vd,
                  const double* dd, float theta, long* nagg, int** agg, int** memb2) {
    HS_TRY(ensure_scratch(H, n));
    int *mate = H->scratch, *prop = mate + n, *flag = prop + n, *pos = flag + n, *bsum = pos + n;
    const int nb = blocks_for(n);
    struct Owned {   // the strengths, freed on every return
        double* p = nullptr;
        ~Owned() { if (p) (void)hipFree(p); }
    } own;
    HS_TRY(amalloc((void**)&own.p, std::max(1L, nnz) * sizeof(double)));
    double* st = own.p;
    k_strength_csr<<<nb, FLAT, 0, c->stream>>>(n, rp, ci, vd, dd, st);
    HS_CHECK(hipMemsetAsync(mate, 0xff, n * sizeof(int), c->stream));
    for (int r = 0; r < ROUNDS; ++r) {
        k_propose_csr<<<nb, FLAT, 0, c->stream>>>(n, rp, ci, st, mate, prop, theta, r);
        k_respond_csr<<<nb, FLAT, 0, c->stream>>>(n, rp, ci, st, mate, prop, r);
    }
    k_lead_csr<<<nb, FLAT, 0, c->stream>>>(n, mate, flag);
    HS_CHECK(hipGetLastError());
    HS_TRY(scan(c, H, n, flag, pos, bsum, nagg));   // synchronises: the strengths may go
    (void)hipFree(own.p);
    own.p = nullptr;
    HS_TRY(amalloc((void**)agg, n * sizeof(int)));
    HS_TRY(amalloc((void**)memb2, 2 * std::max(1L, *nagg) * sizeof(int)));
    k_number<<<nb, FLAT, 0, c->stream>>>(n, flag, pos, mate, *agg, *memb2);
    k_follow<<<nb, FLAT, 0, c->stream>>>(n, flag, mate, *agg);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}

// Galerkin matrix of nc aggregates.  launch(GalOut) runs the count or the fill kernel.
// Pattern (rp, ci, vd, dd) allocated on the first call (*rp == nullptr); later calls
// refill the values.  Returns HS_E_ARG (no allocation kept) if a row exceeds MAXW.
template <typename F>
int galerkin(hs_ctx* c, AmgHier* H, long nc, F launch, int** rp, int** ci, double** vd, double** dd, long* nnz) {
    if (!*rp) {
        HS_TRY(ensure_scratch(H, nc + 1));
        int* cnt = H->scratch;
        int* bsum = H->scratch + 2 * (nc + 1);
        HS_CHECK(hipMemsetAsync(H->dflag, 0, sizeof(int), c->stream));
        launch(GalOut{cnt, nullptr, nullptr, nullptr, nullptr, H->dflag});
        HS_CHECK(hipGetLastError());
        HS_CHECK(hipMemsetAsync(cnt + nc, 0, sizeof(int), c->stream));
        HS_TRY(amalloc((void**)rp, (nc + 1) * sizeof(int)));
        long total = 0;
        HS_TRY(scan(c, H, nc + 1, cnt, *rp, bsum, &total));
        if (H->h_flag[0]) {
            (void)hipFree(*rp);
            *rp = nullptr;
            return HS_E_ARG;
        }
        *nnz = total;
        HS_TRY(amalloc((void**)ci, std::max(1L, total) * sizeof(int)));
        HS_TRY(amalloc((void**)vd, std::max(1L, total) * sizeof(double)));
        HS_TRY(amalloc((void**)dd, std::max(1L, nc) * sizeof(double)));
    }
    launch(GalOut{nullptr, *rp, *ci, *vd, *dd, H->dflag});
    HS_CHECK(hipGetLastError());
    return HS_OK;
}

inline int gal_blocks(long nc) { return (int)std::max(1L, (nc + GAL_T - 1) / GAL_T); }

// the cycle's copies of a level's entries: the sliced ELL form (pattern built on the first
// call, values refilled on later ones) and the diagonal (an alias in an FP64 hierarchy)
template <typename T>
int sell_level(hs_ctx* c, AmgHier* H, AmgLevel* L) {
    const long n = L->n, ns = (n + SL - 1) / SL;
    const bool first = !L->so;
    if (first) {
        HS_TRY(ensure_scratch(H, ns + 1));
        int* cnt = H->scratch;
        int* bsum = H->scratch + 2 * (ns + 1);
        HS_TRY(amalloc((void**)&L->sw, std::max(1L, ns) * sizeof(int)));
        HS_TRY(amalloc((void**)&L->so, (ns + 1) * sizeof(int)));
        k_sell_width<<<blocks_for(ns), FLAT, 0, c->stream>>>(n, L->rp, L->sw, cnt);
        HS_CHECK(hipGetLastError());
        HS_CHECK(hipMemsetAsync(cnt + ns, 0, sizeof(int), c->stream));
        HS_TRY(scan(c, H, ns + 1, cnt, L->so, bsum, &L->nsell));
        HS_TRY(amalloc((void**)&L->sc, std::max(1L, L->nsell) * sizeof(int)));
        HS_TRY(amalloc(&L->sv, std::max(1L, L->nsell) * sizeof(T)));
    }
    k_sell_fill<T><<<blocks_for(n), FLAT, 0, c->stream>>>(n, L->rp, L->ci, L->vd, L->so, L->sw, L->sc, (T*)L->sv,
                                                          first ? 1 : 0);
    if (sizeof(T) == sizeof(double)) {
        L->dg = L->dd;
    } else {
        if (!L->dg) HS_TRY(amalloc(&L->dg, std::max(1L, n) * sizeof(T)));
        k_narrow<T><<<blocks_for(n), FLAT, 0, c->stream>>>(n, L->dd, (T*)L->dg);
    }
    HS_CHECK(hipGetLastError());
    return HS_OK;
}

template <typename T>
int level_vectors(AmgLevel* L) {
    const size_t b = std::max(1L, L->n) * sizeof(T);
    HS_TRY(amalloc(&L->x, b));
    HS_TRY(amalloc(&L->x2, b));
    HS_TRY(amalloc(&L->b, b));
    HS_TRY(amalloc(&L->r, b));
    return HS_OK;
}

// Galerkin kernels for rows of at most W distinct columns (W from the bound of the row
// lengths: twice the longest row of the level above, two members per aggregate)
inline int width_class(int bound) { return bound <= 16 ? 16 : bound <= 32 ? 32 : bound <= 64 ? 64 : MAXW; }

template <typename V>
void launch_gal0(hs_ctx* c, const V& v, long nc, const int* memb2, const int* agg, const GalOut& o, int W) {
    const int g = gal_blocks(nc);
    if (W <= 16) k_gal0<V, 16><<<g, GAL_T, 0, c->stream>>>(v, nc, memb2, agg, o);
    else if (W <= 32) k_gal0<V, 32><<<g, GAL_T, 0, c->stream>>>(v, nc, memb2, agg, o);
    else if (W <= 64) k_gal0<V, 64><<<g, GAL_T, 0, c->stream>>>(v, nc, memb2, agg, o);
    else k_gal0<V, MAXW><<<g, GAL_T, 0, c->stream>>>(v, nc, memb2, agg, o);
}

void launch_gal_csr(hs_ctx* c, long nc, const int* memb2, const int* agg, const int* frp, const int* fci,
                    const double* fvd, const double* fdd, const GalOut& o, int W) {
    const int g = gal_blocks(nc);
    if (W <= 16) k_gal_csr<16><<<g, GAL_T, 0, c->stream>>>(nc, memb2, agg, frp, fci, fvd, fdd, o);
    else if (W <= 32) k_gal_csr<32><<<g, GAL_T, 0, c->stream>>>(nc, memb2, agg, frp, fci, fvd, fdd, o);
    else if (W <= 64) k_gal_csr<64><<<g, GAL_T, 0, c->stream>>>(nc, memb2, agg, frp, fci, fvd, fdd, o);
    else k_gal_csr<MAXW><<<g, GAL_T, 0, c->stream>>>(nc, memb2, agg, frp, fci, fvd, fdd, o);
}

// Galerkin matrix of the intermediate (pass A) aggregates of L from the level above:
// pattern on the first call, values on later ones
template <typename V>
int refill_intermediate(hs_ctx* c, AmgHier* H, const V* v0, const AmgLevel* fine, AmgLevel* L) {
    const long ni = L->ni;
    const int* membf = L->membf;
    const int* aggi = L->aggi;
    L->wbi = std::min(MAXW, 2 * (fine ? fine->wb : 6));
    const int W = width_class(L->wbi);
    if (fine)
        return galerkin(c, H, ni, [&](GalOut o) {
            launch_gal_csr(c, ni, membf, aggi, fine->rp, fine->ci, fine->vd, fine->dd, o, W);
        }, &L->rpi, &L->cii, &L->vdi, &L->ddi, &L->nnzi);
    return galerkin(c, H, ni, [&](GalOut o) { launch_gal0<V>(c, *v0, ni, membf, aggi, o, W); }, &L->rpi, &L->cii,
                    &L->vdi, &L->ddi, &L->nnzi);
}

// Galerkin matrix of L's rows from its intermediate matrix
int refill_level(hs_ctx* c, AmgHier* H, AmgLevel* L) {
    const long n = L->n;
    const int* membi = L->membi;
    const int* ab = L->aggb;
    L->wb = std::min(MAXW, 2 * L->wbi);
    const int W = width_class(L->wb);
    return galerkin(c, H, n, [&](GalOut o) {
        launch_gal_csr(c, n, membi, ab, L->rpi, L->cii, L->vdi, L->ddi, o, W);
    }, &L->rp, &L->ci, &L->vd, &L->dd, &L->nnz);
}

// New level below `fine` (the finest level's view v0 when fine == nullptr): two pairwise
// passes and their Galerkin matrices.  *out = nullptr (nothing kept) when a Galerkin row
// would exceed MAXW or the passes stall (fewer than 10% fewer rows): the level above
// then ends the hierarchy.
template <typename T, typename V>
int make_level(hs_ctx* c, AmgHier* H, const V* v0, const AmgLevel* fine, float theta, AmgLevel** out) {
    *out = nullptr;
    AmgLevel* L = new AmgLevel();
    const long nf = fine ? fine->n : H->n0;
    // below 64K rows every coupling is a candidate: the strong sets of the small levels'
    // large aggregates leave too many rows single for the hierarchy to reach the dense
    // inverse's 4096 rows
    if (nf < 65536) theta = 0.0f;
    int e = HS_OK;
    // pass A over the level above -> intermediate matrix
    if (fine)
        e = pair_pass_csr(c, H, nf, fine->nnz, fine->rp, fine->ci, fine->vd, fine->dd, theta, &L->ni, &L->aggi,
                          &L->membf);
    else e = pair_pass0<V>(c, H, *v0, nf, theta, &L->ni, &L->aggi, &L->membf);
    if (e == HS_OK && L->ni == 0) {   // no unknowns on the level above: nothing to coarsen
        level_free(L);
        return HS_OK;
    }
    if (e == HS_OK) e = refill_intermediate<V>(c, H, v0, fine, L);
    // pass B over the intermediate matrix -> the new level
    if (e == HS_OK)
        e = pair_pass_csr(c, H, L->ni, L->nnzi, L->rpi, L->cii, L->vdi, L->ddi, theta, &L->n, &L->aggb, &L->membi);
    if (e == HS_OK) e = refill_level(c, H, L);
    if (e == HS_OK) {
        e = amalloc((void**)&L->agg, std::max(1L, nf) * sizeof(int));
        if (e == HS_OK) e = amalloc((void**)&L->memb, 4 * std::max(1L, L->n) * sizeof(int));
    }
    if (e == HS_OK) {
        k_compose_agg<<<blocks_for(nf), FLAT, 0, c->stream>>>(nf, L->aggi, L->aggb, L->agg);
        k_compose_memb<<<blocks_for(L->n), FLAT, 0, c->stream>>>(L->n, L->membi, L->membf, L->memb);
        if (hipGetLastError() != hipSuccess) e = HS_E_HIP;
    }
    if (e == HS_OK) e = sell_level<T>(c, H, L);
    if (e == HS_OK) e = level_vectors<T>(L);
    if (e != HS_OK || L->n >= (long)(0.9 * (double)nf)) {
        level_free(L);
        return e == HS_E_ARG ? HS_OK : e;
    }
    *out = L;
    return HS_OK;
}

}  // namespace

namespace {

// Options of a hierarchy, resolved from the parameters.
struct AmgOpt {
    float theta;
    double alpha;
    int sweeps;      // Jacobi sweeps before / after on the aggregated levels
    int refine;      // FP64 refinement steps of the dense coarsest solve
    int coarse;      // Jacobi sweeps on a coarsest level too large for the dense inverse
};

AmgOpt options(const hs_ctx* c, bool energy) {
    const hs_params& p = c->prm;
    AmgOpt o;
    o.theta = p.amg_theta;     // chosen at hs_create when left <= 0
    o.alpha = p.amg_alpha > 0.0f ? (double)p.amg_alpha : (energy ? 1.6 : 1.4);
    o.sweeps = p.amg_sweeps;   // likewise
    o.refine = std::max(0, p.mg_dense_refine);
    o.coarse = std::max(1, p.mg_coarse_sweeps);
    return o;
}

int hier_alloc(AmgHier** Hp, bool f32, int kind, long n0) {
    if (n0 >= 2147483647L) return HS_E_ARG;   // row numbers are ints
    AmgHier* H = *Hp;
    if (!H) {
        H = new AmgHier();
        *Hp = H;
        if (hipMalloc((void**)&H->dflag, 4 * sizeof(int)) != hipSuccess) return HS_E_NOMEM;
        if (hipHostMalloc((void**)&H->h_flag, 4 * sizeof(int), 0) != hipSuccess) return HS_E_NOMEM;
    }
    if (H->f32 != f32 || H->n0 != n0 || H->kind != kind) {
        levels_free(H);
        dense_release(&H->dense);
        H->aggregations = 0;
        H->built_id = -1;
    }
    H->f32 = f32;
    H->kind = kind;
    H->n0 = n0;
    return HS_OK;
}

// Fresh aggregation: levels below the finest until one has at most dense_max_cells()
// rows (dense inverse) or coarsening ends (Jacobi sweeps there).
template <typename T, typename V>
int aggregate(hs_ctx* c, AmgHier* H, const V& v0, const AmgOpt& o) {
    levels_free(H);
    dense_release(&H->dense);
    const AmgLevel* fine = nullptr;
    while ((int)H->lv.size() < AMG_MAX_LEVELS) {
        AmgLevel* L = nullptr;
        HS_TRY((make_level<T, V>(c, H, &v0, fine, o.theta, &L)));
        if (!L) break;
        H->lv.push_back(L);
        fine = L;
        if (L->n <= dense_max_cells()) break;
    }
    // an empty hierarchy (no unknowns, or no level coarsens) leaves the cycle to the finest
    // level's sweeps
    H->aggregations += 1;
    H->since = 0;
    H->iters_ref = -1;
    return HS_OK;
}

// New values on the kept aggregates: every Galerkin matrix again, top down.
template <typename T, typename V>
int refill(hs_ctx* c, AmgHier* H, const V& v0) {
    const AmgLevel* fine = nullptr;
    for (AmgLevel* L : H->lv) {
        HS_TRY(refill_intermediate<V>(c, H, &v0, fine, L));
        HS_TRY(refill_level(c, H, L));
        HS_TRY(sell_level<T>(c, H, L));
        fine = L;
    }
    return HS_OK;
}

int factor_coarsest(hs_ctx* c, AmgHier* H) {
    if (H->lv.empty()) return HS_OK;
    AmgLevel* L = H->lv.back();
    if (L->n > dense_max_cells()) return HS_OK;
    return dense_factor_csr(c, &H->dense, (int)L->n, L->rp, L->ci, L->vd, L->dd, H->f32);
}

// V-cycle over the aggregated levels from lv[l]: its b holds the right-hand side, its x
// leaves with the correction
template <typename T>
int cycle_csr(hs_ctx* c, AmgHier* H, size_t l, const AmgOpt& o, const int* skip) {
    AmgLevel* L = H->lv[l];
    const int nb = blocks_for(L->n);
    const T w = (T)(2.0 / 3.0);
    auto jacobi = [&](int sweeps) -> int {
        for (int s = 0; s < sweeps; ++s) {
            k_jac<T><<<nb, FLAT, 0, c->stream>>>(L->n, L->so, L->sw, L->sc, (const T*)L->sv, (const T*)L->dg,
                                                 (const T*)L->b, (const T*)L->x, (T*)L->x2, w, skip);
            std::swap(L->x, L->x2);
        }
        HS_CHECK(hipGetLastError());
        return HS_OK;
    };
    if (l + 1 == H->lv.size()) {
        if (H->dense)
            return dense_solve_csr(c, H->dense, (int)L->n, L->rp, L->ci, L->vd, L->dd, L->b, L->x, H->f32, o.refine,
                                   skip);
        k_jac0<T><<<nb, FLAT, 0, c->stream>>>(L->n, (const T*)L->dg, (const T*)L->b, (T*)L->x, w, skip);
        return jacobi(o.coarse - 1);
    }
    k_jac0<T><<<nb, FLAT, 0, c->stream>>>(L->n, (const T*)L->dg, (const T*)L->b, (T*)L->x, w, skip);
    HS_TRY(jacobi(o.sweeps - 1));
    AmgLevel* C = H->lv[l + 1];
    k_resid<T><<<nb, FLAT, 0, c->stream>>>(L->n, L->so, L->sw, L->sc, (const T*)L->sv, (const T*)L->dg,
                                           (const T*)L->b, (const T*)L->x, (T*)L->r, skip);
    k_gather4<T><<<blocks_for(C->n), FLAT, 0, c->stream>>>(C->n, C->memb, (const T*)L->r, (T*)C->b, skip);
    HS_CHECK(hipGetLastError());
    HS_TRY(cycle_csr<T>(c, H, l + 1, o, skip));
    k_prolong<T><<<nb, FLAT, 0, c->stream>>>(L->n, C->agg, (const T*)C->x, (T*)L->x, (T)o.alpha, skip);
    HS_CHECK(hipGetLastError());
    return jacobi(o.sweeps);
}

// below the finest level: its residual into r0 (the finest level's spare iterate buffer,
// free between the smoother's sweeps), gathered onto the first aggregated level, the
// cycle there, the scaled correction added to x0
template <typename T, typename V>
int coarse_part(hs_ctx* c, AmgHier* H, const V& v0, int nx, int ny, int nz, long sy, long sz, T* x0, T* r0,
                const AmgOpt& o, const int* skip) {
    AmgLevel* L1 = H->lv[0];
    k_resid0<V, T><<<dim3((nx + 63) / 64, ny, nz), 64, 0, c->stream>>>(v0, sy, sz, r0, skip);
    k_gather4<T><<<blocks_for(L1->n), FLAT, 0, c->stream>>>(L1->n, L1->memb, r0, (T*)L1->b, skip);
    HS_CHECK(hipGetLastError());
    HS_TRY(cycle_csr<T>(c, H, 0, o, skip));
    k_prolong<T><<<blocks_for(H->n0), FLAT, 0, c->stream>>>(H->n0, L1->agg, (const T*)L1->x, x0, (T)o.alpha, skip);
    HS_CHECK(hipGetLastError());
    return HS_OK;
}

// p': fresh aggregation on the first build and whenever the dense-inverse policy of the
// geometric path would refactor (mg_dense_coarse builds, or 2 more FCG iterations than
// the first solve after the last aggregation); otherwise new Galerkin values on the kept
// aggregates, with the coarsest inverse refined against them
template <typename T>
int setup_p(hs_ctx* c) {
    const MGLevel* m = c->mg[0];
    HS_TRY(hier_alloc(&c->amgp, sizeof(T) == sizeof(float), 0, m->n));
    AmgHier* H = c->amgp;
    const PView<T> v{mgl::lv<T>(m)};
    const AmgOpt o = options(c, false);
    const int every = c->prm.mg_dense_coarse > 0 ? c->prm.mg_dense_coarse : 1;
    const bool stale = H->iters_ref >= 0 && c->last_iters >= H->iters_ref + 2;
    if (H->aggregations == 0 || H->since >= every || stale) {
        HS_TRY((aggregate<T, PView<T>>(c, H, v, o)));
        HS_TRY(factor_coarsest(c, H));
    } else {
        HS_TRY((refill<T, PView<T>>(c, H, v)));
    }
    H->since += 1;
    H->builds += 1;
    return HS_OK;
}

}  // namespace

int amg_setup_p(hs_ctx* c) {
    if (c->mg.empty()) return HS_E_STATE;
    return c->mg[0]->f32 ? setup_p<float>(c) : setup_p<double>(c);
}

int amg_coarse_p(hs_ctx* c, const int* skip) {
    AmgHier* H = c->amgp;
    if (!H || c->mg.empty()) return HS_E_STATE;
    if (H->lv.empty()) return HS_OK;
    const MGLevel* m = c->mg[0];
    const AmgOpt o = options(c, false);
    if (H->f32) {
        const PView<float> v{mgl::lv<float>(m)};
        return coarse_part<float, PView<float>>(c, H, v, m->nx, m->ny, m->nz, m->sy, m->sz, (float*)m->x, (float*)m->x2, o, skip);
    }
    const PView<double> v{mgl::lv<double>(m)};
    return coarse_part<double, PView<double>>(c, H, v, m->nx, m->ny, m->nz, m->sy, m->sz, (double*)m->x, (double*)m->x2, o, skip);
}

// energy: one aggregation per operator build and direction (the operator is frozen for
// the solves that follow it)
int amg_setup_e(hs_ctx* c, ELevel* l0, int tr) {
    EnergyState* E = c->en;
    if (!E) return HS_E_STATE;
    const int kind = tr ? 2 : 1;
    HS_TRY(hier_alloc(&E->amg, E->f32, kind, l0->n));
    AmgHier* H = E->amg;
    if (H->built_id == E->build_id) return HS_OK;
    const AmgOpt o = options(c, true);
    if (E->f32) {
        const EView<float> v{heo::el<float>(l0), tr};
        HS_TRY((aggregate<float, EView<float>>(c, H, v, o)));
    } else {
        const EView<double> v{heo::el<double>(l0), tr};
        HS_TRY((aggregate<double, EView<double>>(c, H, v, o)));
    }
    HS_TRY(factor_coarsest(c, H));
    H->built_id = E->build_id;
    H->builds += 1;
    return HS_OK;
}

int amg_coarse_e(hs_ctx* c, ELevel* l0, int tr, const int* skip) {
    EnergyState* E = c->en;
    AmgHier* H = E ? E->amg : nullptr;
    if (!H || H->kind != (tr ? 2 : 1)) return HS_E_STATE;
    if (H->lv.empty()) return HS_OK;
    const AmgOpt o = options(c, true);
    if (H->f32) {
        const EView<float> v{heo::el<float>(l0), tr};
        return coarse_part<float, EView<float>>(c, H, v, l0->nx, l0->ny, l0->nz, l0->sy, l0->sz, (float*)l0->x, (float*)l0->x2, o, skip);
    }
    const EView<double> v{heo::el<double>(l0), tr};
    return coarse_part<double, EView<double>>(c, H, v, l0->nx, l0->ny, l0->nz, l0->sy, l0->sz, (double*)l0->x, (double*)l0->x2, o, skip);
}

void amg_free(AmgHier** Hp) {
    AmgHier* H = *Hp;
    if (!H) return;
    levels_free(H);
    dense_release(&H->dense);
    if (H->scratch) (void)hipFree(H->scratch);
    if (H->dflag) (void)hipFree(H->dflag);
    if (H->h_flag) (void)hipHostFree(H->h_flag);
    delete H;
    *Hp = nullptr;
}
