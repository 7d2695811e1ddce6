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

// Device memory handed out from a region the caller owns (no allocation):
// 256-byte aligned pieces, in order, until the region is used up.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
};

// Device buffer: owned by the library (alloc), or a piece of an Arena
// (borrowed, never freed here).
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    bool owned = false;
    int alloc(size_t nbytes) {
        release();
        bytes = nbytes;
        if (nbytes == 0) return VDI_OK;
        VDI_HIP(hipMalloc(&p, nbytes));
        owned = true;
        VDI_HIP(hipMemset(p, 0, nbytes));
        return VDI_OK;
    }
    // A piece of the arena when one is given, otherwise an allocation.
    int alloc(size_t nbytes, Arena *arena) {
        if (!arena) return alloc(nbytes);
        release();
        const size_t off = (arena->used + 255) & ~(size_t)255;
        if (off + nbytes > arena->cap) return VDI_ERR_ARG;
        p = arena->base + off;
        bytes = nbytes;
        arena->used = off + nbytes;
        return VDI_OK;
    }
    void release() {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        owned = false;
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

// Workgroup timing stamp: the wall-clock ticks at which one workgroup of a
// timed launch started and finished, stored by the workgroup itself (plain
// stores into its own stamp: atomics of every workgroup on one stamp
// serialised across the XCDs and lengthened the launch by a quarter).  A
// launch's duration runs from its earliest start to its latest end; a
// workgroup that returned at once (a converged solve) leaves t0 = ~0 and is
// not counted.
struct Stamp {
    unsigned long long t0, t1;
};

#ifdef __HIPCC__
// ts: the launch's stamps, one per workgroup of its (1-D) grid.
__device__ __forceinline__ void stamp_begin(Stamp *ts) {
    if (ts && threadIdx.x == 0 && threadIdx.y == 0) ts[blockIdx.x].t0 = (unsigned long long)wall_clock64();
}
__device__ __forceinline__ void stamp_end(Stamp *ts) {
    if (!ts) return;
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) ts[blockIdx.x].t1 = (unsigned long long)wall_clock64();
}
// Per-simulation clocks of a batched step, three counters per simulation:
// busy[3 s] the workgroup busy ticks summed over its work items, busy[3 s + 1]
// the start of its first work item and busy[3 s + 2] the end of its last one
// (its step's span inside the batched step).
__device__ __forceinline__ unsigned long long busy_begin(const unsigned long long *busy) {
    return busy ? (unsigned long long)wall_clock64() : 0ull;
}
__device__ __forceinline__ void busy_end(unsigned long long *busy, int sim, unsigned long long t0) {
    if (busy && threadIdx.x == 0 && threadIdx.y == 0) {
        const unsigned long long t1 = (unsigned long long)wall_clock64();
        atomicAdd(&busy[3 * sim], t1 - t0);
        atomicMin(&busy[3 * sim + 1], t0);
        atomicMax(&busy[3 * sim + 2], t1);
    }
}
#endif

// Times every launch of one kernel class with stamps written by the kernels
// themselves, and keeps the algorithmic bytes of each launch (arrays read +
// arrays written, each counted once).  Works for direct launches and for
// captured graphs: a graph owns the contiguous slot range handed out while it
// was captured, clears it at its start and the host reads it after every
// replay.
struct KernelTimer {
    int cls = -1;       // class being timed (VDI_KCLASS_*), -1 = off
    DevBuf stamps;      // ncap workgroup stamps
    Stamp *host = nullptr;  // pinned mirror
    int cap = 0;            // launch slots
    long ncap = 0;          // workgroup stamps
    int next = 0, reserved = 0;  // slots [0, reserved) belong to captured graphs
    double ticks_per_ms = 1e5;
    std::vector<double> slot_bytes;
    std::vector<long> off;  // launch slot k owns stamps [off[k], off[k + 1])
    std::vector<float> dur_ms;
    std::vector<double> bytes;
    // A launch of nwg workgroups belongs to a level class c and (FP64 level
    // kernels) to the class of its kernel function c2 (-1: none).
    Stamp *slot(int c, double alg_bytes, int c2 = -1, int nwg = 1) {
        if ((c != cls && c2 != cls) || next >= cap || off[next] + nwg > ncap) return nullptr;
        slot_bytes[next] = alg_bytes;
        off[next + 1] = off[next] + nwg;
        return stamps.as<Stamp>() + off[next++];
    }
    int setup(int c, int capacity);
    // Enqueues the reset of the stamps of slots [a, b) (b = cap: every stamp
    // from slot a on; a kernel, so it can be captured).
    int reset(int a, int b, hipStream_t st);
    // Reads slots [a, b) after their launches have completed.
    int harvest(int a, int b, hipStream_t st);
    void release();
    ~KernelTimer() { release(); }
};

}  // namespace vdi
