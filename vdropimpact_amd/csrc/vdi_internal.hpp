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

// One multigrid level on the device, in FP64 or FP32.
// A level kernel never reads and writes the same array (its workgroups read
// halos that neighbouring workgroups would otherwise already have updated):
// the down kernel reads z and writes zalt, the up kernel reads zalt and
// writes z.
template <class T>
struct LevelT {
    int m, n;    // unknowns
    T cx, cy;    // cell-centred couplings
    T inv_int;   // reciprocal of the interior diagonal
    T *z, *b;    // solution / right-hand side (level 0: owned elsewhere)
    T *zalt;     // pre-smoothed iterate between the down and up kernels
};
using Level = LevelT<double>;
using LevelF = LevelT<float>;

// Times every launch of one kernel class with HIP events recorded on the
// stream that runs it, and keeps the algorithmic bytes of each launch (arrays
// read + arrays written, each counted once).  Works for direct launches and
// for captured graphs: a graph owns the event pairs recorded while it was
// captured and the host harvests them after every replay.
struct KernelTimer {
    struct Slot {
        int pair;
        double bytes;
    };
    int cls = -1;  // class being timed (VDI_KCLASS_*), -1 = off
    std::vector<hipEvent_t> ev;
    size_t next = 0, reserved = 0;  // pairs [0, reserved) belong to captured graphs
    std::vector<Slot> pending;      // recorded since the last harvest / capture start
    std::vector<float> dur_ms;
    std::vector<double> bytes;
    int begin(int c, hipStream_t st) {
        if (c != cls || 2 * next + 2 > ev.size()) return -1;
        (void)hipEventRecord(ev[2 * next], st);
        return (int)next++;
    }
    void end(int pair, hipStream_t st, double alg_bytes) {
        if (pair < 0) return;
        (void)hipEventRecord(ev[2 * pair + 1], st);
        pending.push_back(Slot{pair, alg_bytes});
    }
    // Reads the durations of the given (completed) launches.
    void harvest(const std::vector<Slot> &slots) {
        for (const Slot &s : slots) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, ev[2 * s.pair], ev[2 * s.pair + 1]) == hipSuccess) {
                dur_ms.push_back(ms);
                bytes.push_back(s.bytes);
            }
        }
    }
    // After a direct (uncaptured) step has completed.
    void harvest_direct() {
        harvest(pending);
        pending.clear();
        next = reserved;
    }
    void release() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        ev.clear();
        pending.clear();
        dur_ms.clear();
        bytes.clear();
        next = reserved = 0;
    }
    ~KernelTimer() { release(); }
};

}  // namespace vdi
