// hs_static.hip — hs_pressure_solve_static: the p' solve of hs_pressure_solve for callers
// that cannot allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// The solve's scratch, meaning the Krylov control block (alpha, beta, stop flags) and the
// block partials of every reduction, lives in `scratch`, a fixed file-scope device array
// (STATIC_CTL_BYTES + STATIC_PARTIALS doubles, 256 KiB + 512 B) that each GPU holds
// until the process ends.  Everything else the solve touches was allocated by hs_create
// (Krylov vectors, staging) or by an earlier multigrid solve on the context (the
// hierarchy, the dense coarse inverse), so the entry point itself calls no hipMalloc,
// hipFree or allocating library routine.  Grids whose reductions need more partials than
// the array holds are refused with HS_E_ARG.
//
// Every call on one GPU shares that GPU's array: the control block is cleared at the
// start of each call (a fill kernel), and the partials are written before they are read
// within each reduction.  Calls that overlap on one GPU (other host threads, other
// contexts' streams) race on the array and must be serialised by the caller.
#include <mutex>

#include "hs_device.h"

namespace {
constexpr long STATIC_CTL_DOUBLES = 64;        // SolveCtl, padded to 512 bytes
constexpr long STATIC_PARTIALS = 32768;        // block partials, 256 KiB
constexpr long STATIC_DOUBLES = STATIC_CTL_DOUBLES + STATIC_PARTIALS;
constexpr int MAX_DEVICES = 64;
static_assert(sizeof(SolveCtl) <= STATIC_CTL_DOUBLES * sizeof(double), "the control block must fit its slot");
}  // namespace

// one copy on every GPU the library runs on; host code must use the address from
// hipGetSymbolAddress (the name itself is a host-side stand-in, not a device address)
static __device__ __align__(256) double scratch[STATIC_DOUBLES];

namespace {
std::mutex g_scratch_mu;
double* g_scratch_dev[MAX_DEVICES] = {};

// device address of the current GPU's copy of `scratch`, looked up on the first call on
// that GPU (with it current) and kept per device
int scratch_address(int dev, double** out) {
    if (dev < 0 || dev >= MAX_DEVICES) return HS_E_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "hsolver: hipGetSymbolAddress(scratch) on device %d: %s\n", dev, hipGetErrorString(e));
            return HS_E_HIP;
        }
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_dev[dev];
    return HS_OK;
}

// puts the context's own partials / control block back when the call returns
struct ScratchSwap {
    hs_ctx* c;
    double* partials;
    SolveCtl* sctl;
    ~ScratchSwap() {
        c->partials = partials;
        c->sctl = sctl;
    }
};
}  // namespace

extern "C" int hs_pressure_solve_static(hs_ctx* c, const double* b_dev, double* x_dev, hs_stats* st) {
    if (!c || !b_dev || !x_dev) return HS_E_ARG;
    const long need = hsd::krylov_partials(c->d);
    if (need > STATIC_PARTIALS) {
        fprintf(stderr,
                "hsolver: hs_pressure_solve_static: a %dx%dx%d grid needs %ld reduction partials, the fixed "
                "scratch holds %ld\n",
                c->d.nx, c->d.ny, c->d.nz, need, STATIC_PARTIALS);
        return HS_E_ARG;
    }
    if (!c->have_matrix) return HS_E_STATE;
    // the aggregation multigrid allocates when it aggregates again: not available here
    if (c->prm.precond == HS_PRECOND_AMG) return HS_E_ARG;
    // the multigrid solve allocates its hierarchy (and dense coarse inverse) on first use
    if (c->prm.precond == HS_PRECOND_GMG && (c->mg.empty() || (c->prm.mg_dense_coarse > 0 && !c->dense)))
        return HS_E_STATE;
    int dev = -1;
    HS_CHECK(hipGetDevice(&dev));
    if (dev != c->device) return HS_E_STATE;
    double* s = nullptr;
    HS_TRY(scratch_address(dev, &s));
    HS_TRY(launch_fill(c, s, STATIC_CTL_DOUBLES, 0.0));
    ScratchSwap keep{c, c->partials, c->sctl};
    c->sctl = reinterpret_cast<SolveCtl*>(s);
    c->partials = s + STATIC_CTL_DOUBLES;
    const size_t bytes = (size_t)c->d.nP * sizeof(double);
    HS_CHECK(hipMemcpyAsync(c->kb, b_dev, bytes, hipMemcpyDeviceToDevice, c->stream));
    hs_stats loc{};
    HS_TRY(pressure_solve_staged(c, &loc));
    HS_CHECK(hipMemcpyAsync(x_dev, c->kx, bytes, hipMemcpyDeviceToDevice, c->stream));
    HS_CHECK(hipStreamSynchronize(c->stream));
    if (st) *st = loc;
    return HS_OK;
}
