// noalloc.hip -- the standalone multigrid solve without device allocation.
//
// vdi_mg_solve_noalloc is vdi_mg_create + vdi_mg_solve on caller-owned device
// arrays for callers that cannot call hipMalloc (code under HIP graph
// capture, tight serving loops, harnesses that forbid allocation).  Every work
// buffer the solve needs beyond b and z (coarse-level iterates and right-hand
// sides, the pre-smoothed level-0 iterate, per-workgroup partial sums, the
// solve state and the factors of the coarsest level's exact solve) is carved
// out of one file-scope device array, `scratch`, of kScratchBytes.  Each GPU
// holds its own copy of that array for the life of the process, and every
// call on a GPU shares it: calls that overlap on one GPU race.
#include <mutex>

#include "mg.hpp"

using namespace vdi;

namespace {

constexpr size_t kScratchBytes = (size_t)VDI_NOALLOC_SCRATCH_MB << 20;
constexpr int kMaxDevices = 64;

// The array's name is a host-side stand-in in host code: kernels get the
// device address that hipGetSymbolAddress returns, kept per device.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void *g_scratch_dev[kMaxDevices] = {};

// Device address of this GPU's scratch array (looked up once per device, with
// that device current).
int scratch_address(void **out) {
    int dev = 0;
    VDI_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return VDI_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        VDI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return VDI_ERR_HIP;
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return VDI_OK;
}

}  // namespace

extern "C" int vdi_mg_solve_noalloc(int kind, int m, int n, double cx, double cy, int x_sym, const double *b_dev,
                                    double *z_dev, double rel_tol, int max_cycles, int *vcycles) {
    if (kind < VDI_MAC || kind > VDI_VY || m < 2 || n < 2 || !b_dev || !z_dev || !(rel_tol > 0) || max_cycles < 1)
        return VDI_ERR_ARG;
    void *base = nullptr;
    VDI_TRY(scratch_address(&base));
    Arena arena;
    arena.base = static_cast<char *>(base);
    arena.cap = kScratchBytes;
    const hipStream_t st = nullptr;
    MG g;
    // The hierarchy is laid out in the arena (init also writes the factors of
    // the coarsest level's exact solve there); a size that does not fit is
    // rejected here.
    VDI_TRY(g.init(kind, m, n, cx, cy, x_sym, z_dev, const_cast<double *>(b_dev), st, &arena));
    // The array keeps the previous call's contents: the solve state and the
    // partial sums start from zero (fill kernels, no memset).  Every other
    // piece is written before it is read.
    VDI_TRY(g.clear_solve_state());
    int cyc = 0;
    const int rc = g.solve(rel_tol, max_cycles, &cyc);
    VDI_HIP(hipStreamSynchronize(st));
    if (vcycles) *vcycles = cyc;
    return rc;
}
