// vdi_internal.hpp -- shared definitions of the gfx950 fluid-step library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/vdi.h"

namespace vdi {

// Last HIP error text, reported through vdi_last_error().
void set_error(const char *what, hipError_t e, const char *file, int line);

#define VDI_HIP(call)                                                     \
    do {                                                                  \
        hipError_t e_ = (call);                                           \
        if (e_ != hipSuccess) {                                           \
            ::vdi::set_error(#call, e_, __FILE__, __LINE__);              \
            return VDI_ERR_HIP;                                           \
        }                                                                 \
    } while (0)

#define VDI_TRY(call)                      \
    do {                                   \
        int rc_ = (call);                  \
        if (rc_ != VDI_OK) return rc_;     \
    } while (0)

// Launch check: kernel launches report errors through hipGetLastError.
#define VDI_LAUNCHED() VDI_HIP(hipGetLastError())

inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }

// Device buffer owned by the library.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int alloc(size_t nbytes) {
        release();
        bytes = nbytes;
        if (nbytes == 0) return VDI_OK;
        VDI_HIP(hipMalloc(&p, nbytes));
        VDI_HIP(hipMemset(p, 0, nbytes));
        return VDI_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
    ~DevBuf() { release(); }
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
};

// Boundary type of a cell-centred operator side.
//   N : Neumann, ghost = interior value        (diagonal +0)
//   DF: Dirichlet on the face, ghost = -value  (diagonal +2c)
//   DC: Dirichlet at the ghost centre, ghost=0 (diagonal +c)
enum { BC_N = 0, BC_DF = 1, BC_DC = 2 };

// Cell-centred operator A = sigma I + s * Lap with per-side boundary types
// (MAC: sigma=0, s=+1; visco: sigma=1, s=-1).
struct CCOp {
    double sigma, s;
    int bcl, bcr, bcb, bct;
};

// Nodal Q1 finite-element operator (mgs_fem.cc:8-11); the same stencil holds
// on every level because the bilinear stiffness only depends on dx/dy.
struct FemOp {
    double fm, fex, fey, fc, hex, hey;
};

// One multigrid level on the device.
struct Level {
    int m, n;       // unknowns
    double cx, cy;  // cell-centred couplings
    double *z, *b;  // solution / right-hand side (level 0: owned elsewhere)
};

}  // namespace vdi
