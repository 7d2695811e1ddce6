// mg.hip -- multigrid kernels and driver (see mg.hpp for the scheme).
//
// Kernel structure of one V-cycle on an L-level hierarchy whose levels
// lc..L-1 hold at most kCoarsePoints unknowns:
//   down(l), l = 0..lc-1 : nu1 smoothing sweeps of level l, residual,
//                          restriction into level l+1   (one kernel per level)
//   coarse               : a complete V-cycle on levels lc..L-1 inside one
//                          workgroup's LDS
//   up(l), l = lc-1..0   : prolongation + correction, nu2 sweeps; at level 0
//                          also ||b - A z||^2, reduced by the last workgroup,
//                          which raises the device-side convergence flag.
// Every kernel returns at once when the flag is set, so a solve enqueues the
// predicted number of cycles and synchronises with the host once.
//
// The down/up kernels are temporally blocked: a workgroup loads its
// kTX x kTY output tile plus a halo of H cells into LDS, performs the sweeps
// on a region that shrinks by one cell per colour pass (so every update sees
// exactly the values a global sweep would), and writes back only the tile.
// Halo: cell-centred red-black H = 2 nu (+1 for a residual); nodal four-
// colour H = 4 nu (+1 residual, +1 more for full-weighting restriction).
// Values outside the domain are zero in LDS, which is exactly how both
// operators drop missing neighbours.
#include "mg.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace vdi {

constexpr int kTX = 64, kTY = 32;   // output tile of the level kernels
constexpr int kBX = 32, kBY = 8;    // 256 threads
constexpr int kCoarsePoints = 3072; // levels this small go to the single-workgroup kernel
constexpr int kCoarseThreads = 256;

// ------------------------------------------------------------------------
// Operators on plain arrays (mgs_mac.hh:37-50, visco_impl.hh:32-53,
// mgs_fem.hh:42-56 + mgs_fem.cc:34-50), in the level's value type T
// ------------------------------------------------------------------------

__device__ __forceinline__ double bc_diag_factor(int bc) {
    return bc == BC_DF ? 2.0 : (bc == BC_DC ? 1.0 : 0.0);
}
__device__ __forceinline__ double bc_ghost_factor(int bc) {
    return bc == BC_DF ? -1.0 : (bc == BC_DC ? 0.0 : 1.0);
}

template <class T>
__device__ __forceinline__ T cc_diag(const CCOp &op, int m, int n, T cx, T cy, int i, int j) {
    T d = 0;
    d += T(i > 0 ? 1.0 : bc_diag_factor(op.bcl)) * cx;
    d += T(i < m - 1 ? 1.0 : bc_diag_factor(op.bcr)) * cx;
    d += T(j > 0 ? 1.0 : bc_diag_factor(op.bcb)) * cy;
    d += T(j < n - 1 ? 1.0 : bc_diag_factor(op.bct)) * cy;
    return T(op.sigma) - T(op.s) * d;
}

template <class T>
__device__ __forceinline__ T cc_off(int m, int n, T cx, T cy, const T *z, int i, int j) {
    const T *w = z + i + (size_t)m * j;
    T sx = 0, sy = 0;
    if (i > 0) sx += w[-1];
    if (i < m - 1) sx += w[1];
    if (j > 0) sy += w[-m];
    if (j < n - 1) sy += w[m];
    return cx * sx + cy * sy;
}

template <class T>
__device__ __forceinline__ T fem_acc(const FemOp &op, int M, int N, int i, int j) {
    return T((i > 0 && i < M - 1 ? 1.0 : 0.5) * (j > 0 && j < N - 1 ? 1.0 : 0.5) * op.fm);
}

template <class T>
__device__ __forceinline__ T fem_off(const FemOp &op, int M, int N, const T *z, int i, int j) {
    const T *w = z + i + (size_t)M * j, *wm = w - M;
    const T fc = T(op.fc), fex = T(op.fex), fey = T(op.fey), hex = T(op.hex), hey = T(op.hey);
    if (j == 0) return 0;
    if (j == N - 1) {
        if (i == 0) return fc * wm[1] + hey * wm[0] + hex * w[1];
        if (i == M - 1) return fc * wm[-1] + hey * wm[0] + hex * w[-1];
        return fc * (wm[-1] + wm[1]) + fey * wm[0] + hex * (w[-1] + w[1]);
    }
    const T *wp = w + M;
    if (i == 0) return fc * (wm[1] + wp[1]) + hey * (wm[0] + wp[0]) + fex * w[1];
    if (i == M - 1) return fc * (wm[-1] + wp[-1]) + hey * (wm[0] + wp[0]) + fex * w[-1];
    return fc * (wm[-1] + wm[1] + wp[-1] + wp[1]) + fey * (wm[0] + wp[0]) + fex * (w[-1] + w[1]);
}

template <class T>
__device__ __forceinline__ T Az_plain(bool fem, const CCOp &cop, const FemOp &fop, int m, int n, T cx, T cy,
                                     const T *z, int i, int j) {
    if (fem) return fem_acc<T>(fop, m, n, i, j) * z[i + (size_t)m * j] + fem_off<T>(fop, m, n, z, i, j);
    return T(cop.s) * cc_off<T>(m, n, cx, cy, z, i, j) + cc_diag<T>(cop, m, n, cx, cy, i, j) * z[i + (size_t)m * j];
}

// ------------------------------------------------------------------------
// Operators on an LDS tile with a zero-filled outside (row pitch RX)
// ------------------------------------------------------------------------

template <class T>
__device__ __forceinline__ T cc_off_t(T cx, T cy, const T *s, int l, int RX) {
    return cx * (s[l - 1] + s[l + 1]) + cy * (s[l - RX] + s[l + RX]);
}

// 9-point Q1 stencil with missing neighbours zero: edge columns halve the
// vertical coupling (hey), the top row halves the horizontal one (hex).
template <class T>
__device__ __forceinline__ T fem_off_t(const FemOp &op, int M, int N, const T *s, int l, int RX, int i, int j) {
    const T wi = (i == 0 || i == M - 1) ? T(0.5) : T(1);
    const T wj = (j == N - 1) ? T(0.5) : T(1);
    return T(op.fc) * (s[l - RX - 1] + s[l - RX + 1] + s[l + RX - 1] + s[l + RX + 1]) +
           wi * T(op.fey) * (s[l - RX] + s[l + RX]) + wj * T(op.fex) * (s[l - 1] + s[l + 1]);
}

template <class T>
__device__ __forceinline__ T fem_inv_acc(const LevelT<T> &L, int i, int j) {
    return T(((i == 0 || i == L.m - 1) ? 2.0 : 1.0) * ((j == 0 || j == L.n - 1) ? 2.0 : 1.0)) * L.inv_int;
}

struct Tile {
    int x0, y0, gx0, gy0, RX, RY, H;
};

__device__ __forceinline__ Tile make_tile(int bx, int by, int H) {
    Tile t;
    t.H = H;
    t.x0 = bx * kTX;
    t.y0 = by * kTY;
    t.gx0 = t.x0 - H;
    t.gy0 = t.y0 - H;
    t.RX = kTX + 2 * H;
    t.RY = kTY + 2 * H;
    return t;
}

// Bilinear cell-centred prolongation of the coarse correction at fine (i,j).
template <class T>
__device__ __forceinline__ T cc_prolong_at(const CCOp &op, int Mc, int Nc, const T *e, int i, int j) {
    int I = i >> 1, J = j >> 1;
    int I2 = (i & 1) ? I + 1 : I - 1, J2 = (j & 1) ? J + 1 : J - 1;
    T fx = 1, fy = 1;
    if (I2 < 0) { fx = T(bc_ghost_factor(op.bcl)); I2 = I; }
    else if (I2 >= Mc) { fx = T(bc_ghost_factor(op.bcr)); I2 = I; }
    if (J2 < 0) { fy = T(bc_ghost_factor(op.bcb)); J2 = J; }
    else if (J2 >= Nc) { fy = T(bc_ghost_factor(op.bct)); J2 = J; }
    T e00 = e[I + (size_t)Mc * J], e10 = fx * e[I2 + (size_t)Mc * J];
    T e01 = fy * e[I + (size_t)Mc * J2], e11 = fx * fy * e[I2 + (size_t)Mc * J2];
    return T(0.5625) * e00 + T(0.1875) * (e10 + e01) + T(0.0625) * e11;
}

// Nodal bilinear prolongation at fine node (i,j), j >= 1.
template <class T>
__device__ __forceinline__ T fem_prolong_at(const T *ez, int Mc, int i, int j) {
    const T *q = ez + (i >> 1) + (size_t)Mc * (j >> 1);
    if (!(i & 1) && !(j & 1)) return q[0];
    if (!(j & 1)) return T(0.5) * (q[0] + q[1]);
    if (!(i & 1)) return T(0.5) * (q[0] + q[Mc]);
    return T(0.25) * (q[0] + q[1] + q[Mc] + q[Mc + 1]);
}

// Loads zsrc (+ prolongated coarse correction when C is given; zero when
// zsrc is null) and b over the tile region; outside the domain both are zero.
template <class T, bool FEM>
__device__ __forceinline__ void tile_load(const LevelT<T> &F, const T *zsrc, const LevelT<T> *C, const CCOp &op,
                                          T *sz, T *sb, const Tile &t) {
    for (int ly = threadIdx.y; ly < t.RY; ly += kBY) {
        const int gy = t.gy0 + ly;
        const bool yin = gy >= 0 && gy < F.n;
        for (int lx = threadIdx.x; lx < t.RX; lx += kBX) {
            const int gx = t.gx0 + lx;
            const bool in = yin && gx >= 0 && gx < F.m;
            T zv = 0, bv = 0;
            if (in) {
                const size_t g = gx + (size_t)F.m * gy;
                if (zsrc) zv = zsrc[g];
                bv = F.b[g];
                if (C) {
                    if (FEM) { if (gy > 0) zv += fem_prolong_at<T>(C->z, C->m, gx, gy); }
                    else zv += cc_prolong_at<T>(op, C->m, C->n, C->z, gx, gy);
                }
            }
            sz[ly * t.RX + lx] = zv;
            sb[ly * t.RX + lx] = bv;
        }
    }
}

template <class T>
__device__ __forceinline__ void tile_store(const LevelT<T> &F, T *dst, const T *sz, const Tile &t) {
    for (int y = threadIdx.y; y < kTY; y += kBY) {
        const int gy = t.y0 + y;
        if (gy >= F.n) break;
        for (int x = threadIdx.x; x < kTX; x += kBX) {
            const int gx = t.x0 + x;
            if (gx >= F.m) break;
            dst[gx + (size_t)F.m * gy] = sz[(y + t.H) * t.RX + x + t.H];
        }
    }
}

// Red-black half sweep of colour c on the region shrunk by s cells.
template <class T>
__device__ __forceinline__ void cc_tile_sweep(const LevelT<T> &L, const CCOp &op, T *sz, const T *sb, const Tile &t,
                                              int s, int c) {
    const int X0 = max(0, t.gx0 + s), X1 = min(L.m, t.x0 + kTX + t.H - s);
    const int Y0 = max(0, t.gy0 + s), Y1 = min(L.n, t.y0 + kTY + t.H - s);
    const T sgn = T(op.s);
    for (int gy = Y0 + threadIdx.y; gy < Y1; gy += kBY) {
        const int xs = X0 + (((X0 + gy) & 1) ^ c);
        const int row = (gy - t.gy0) * t.RX - t.gx0;
        const bool yint = gy > 0 && gy < L.n - 1;
        for (int gx = xs + 2 * threadIdx.x; gx < X1; gx += 2 * kBX) {
            const int l = row + gx;
            const T rhs = sb[l] - sgn * cc_off_t<T>(L.cx, L.cy, sz, l, t.RX);
            const bool in = yint && gx > 0 && gx < L.m - 1;
            sz[l] = in ? rhs * L.inv_int : rhs / cc_diag<T>(op, L.m, L.n, L.cx, L.cy, gx, gy);
        }
    }
}

// Four-colour sweep of colour (ci,cj) on the region shrunk by s; row 0 fixed.
template <class T>
__device__ __forceinline__ void fem_tile_sweep(const LevelT<T> &L, const FemOp &op, T *sz, const T *sb, const Tile &t,
                                               int s, int ci, int cj) {
    const int X0 = max(0, t.gx0 + s), X1 = min(L.m, t.x0 + kTX + t.H - s);
    const int Y0 = max(1, t.gy0 + s), Y1 = min(L.n, t.y0 + kTY + t.H - s);
    const int ys = Y0 + ((Y0 ^ cj) & 1), xs = X0 + ((X0 ^ ci) & 1);
    for (int gy = ys + 2 * threadIdx.y; gy < Y1; gy += 2 * kBY) {
        const int row = (gy - t.gy0) * t.RX - t.gx0;
        for (int gx = xs + 2 * threadIdx.x; gx < X1; gx += 2 * kBX) {
            const int l = row + gx;
            sz[l] = (sb[l] - fem_off_t<T>(op, L.m, L.n, sz, l, t.RX, gx, gy)) * fem_inv_acc<T>(L, gx, gy);
        }
    }
}

template <class T>
__device__ __forceinline__ T cc_resid_t(const LevelT<T> &L, const CCOp &op, const T *sz, const T *sb, int RX, int l,
                                        int gx, int gy) {
    return sb[l] - (T(op.s) * cc_off_t<T>(L.cx, L.cy, sz, l, RX) + cc_diag<T>(op, L.m, L.n, L.cx, L.cy, gx, gy) * sz[l]);
}

template <class T>
__device__ __forceinline__ T fem_resid_t(const FemOp &op, int M, int N, const T *sz, const T *sb, int RX, int l,
                                         int gx, int gy) {
    const T acc = fem_acc<T>(op, M, N, gx, gy);
    if (gy == 0) return sb[l] - acc * sz[l];
    return sb[l] - (acc * sz[l] + fem_off_t<T>(op, M, N, sz, l, RX, gx, gy));
}

// Workgroup sum (fixed order); result valid in thread 0.
__device__ __forceinline__ double block_sum(double a, double *sh) {
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o);
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    const int w = tid >> 6, nw = (blockDim.x * blockDim.y + 63) >> 6;
    if ((tid & 63) == 0) sh[w] = a;
    __syncthreads();
    if (tid == 0) {
        a = 0;
        for (int k = 0; k < nw; k++) a += sh[k];
    }
    return a;
}

// Convergence bookkeeping: the last workgroup to finish sums the per-
// workgroup partials in index order and sets the flag.
__device__ __forceinline__ void finish_norm(const MGCtl &ctl, double part, int nblocks, double tol2, double *sh) {
    __shared__ int last;
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    const int bid = blockIdx.x + gridDim.x * blockIdx.y;
    part = block_sum(part, sh);
    if (tid == 0) {
        ctl.partial[bid] = part;
        __threadfence();
        unsigned tk = atomicAdd(&ctl.state->ticket, 1u);
        last = (tk == (unsigned)nblocks - 1);
    }
    __syncthreads();
    if (last && tid == 0) {
        __threadfence();
        double rr = 0;
        for (int k = 0; k < nblocks; k++) rr += __hip_atomic_load(&ctl.partial[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        MGState *st = ctl.state;
        st->rr = rr;
        st->ticket = 0;
        st->cycles += 1;
        st->done = !(rr == rr) ? 2 : (rr <= tol2 * st->bb ? 1 : 0);
    }
}

// ------------------------------------------------------------------------
// Level kernels (T = double for FP64 hierarchies, float for the inner
// cycles of the mixed form)
// ------------------------------------------------------------------------

template <class T>
__global__ void __launch_bounds__(256) k_cc_down(LevelT<T> F, LevelT<T> C, CCOp op, int nu, int zero_z,
                                                 const MGState *state) {
    extern __shared__ double lds_d[];
    T *lds = reinterpret_cast<T *>(lds_d);
    if (state->done) return;
    const Tile t = make_tile(blockIdx.x, blockIdx.y, 2 * nu + 1);
    T *sz = lds, *sb = lds + t.RX * t.RY;
    tile_load<T, false>(F, zero_z ? nullptr : F.z, nullptr, op, sz, sb, t);
    __syncthreads();
    for (int s = 1; s <= 2 * nu; s++) {
        cc_tile_sweep<T>(F, op, sz, sb, t, s, (s - 1) & 1);
        __syncthreads();
    }
    tile_store<T>(F, F.zalt, sz, t);
    const int I0 = t.x0 >> 1, J0 = t.y0 >> 1;
    const int I1 = min(C.m, I0 + kTX / 2), J1 = min(C.n, J0 + kTY / 2);
    for (int J = J0 + threadIdx.y; J < J1; J += kBY)
        for (int I = I0 + threadIdx.x; I < I1; I += kBX) {
            T r = 0;
            for (int b2 = 0; b2 < 2; b2++)
                for (int a2 = 0; a2 < 2; a2++) {
                    const int gx = 2 * I + a2, gy = 2 * J + b2;
                    const int l = (gy - t.gy0) * t.RX + gx - t.gx0;
                    r += cc_resid_t<T>(F, op, sz, sb, t.RX, l, gx, gy);
                }
            C.b[I + (size_t)C.m * J] = T(0.25) * r;
        }
}

template <class T>
__global__ void __launch_bounds__(256) k_cc_up(LevelT<T> F, LevelT<T> C, CCOp op, int nu, int norm, MGCtl ctl,
                                               double tol2) {
    extern __shared__ double lds_d[];
    T *lds = reinterpret_cast<T *>(lds_d);
    __shared__ double sh[8];
    if (ctl.state->done) return;
    const Tile t = make_tile(blockIdx.x, blockIdx.y, 2 * nu + norm);
    T *sz = lds, *sb = lds + t.RX * t.RY;
    tile_load<T, false>(F, F.zalt, &C, op, sz, sb, t);
    __syncthreads();
    for (int s = 1; s <= 2 * nu; s++) {
        cc_tile_sweep<T>(F, op, sz, sb, t, s, (s - 1) & 1);
        __syncthreads();
    }
    tile_store<T>(F, F.z, sz, t);
    if (!norm) return;
    double rr = 0;
    for (int y = threadIdx.y; y < kTY; y += kBY) {
        const int gy = t.y0 + y;
        if (gy >= F.n) break;
        for (int x = threadIdx.x; x < kTX; x += kBX) {
            const int gx = t.x0 + x;
            if (gx >= F.m) break;
            const double r = cc_resid_t<T>(F, op, sz, sb, t.RX, (y + t.H) * t.RX + x + t.H, gx, gy);
            rr += r * r;
        }
    }
    finish_norm(ctl, rr, gridDim.x * gridDim.y, tol2, sh);
}

template <class T>
__global__ void __launch_bounds__(256) k_fem_down(LevelT<T> F, LevelT<T> C, FemOp op, int nu, int zero_z,
                                                  const MGState *state) {
    extern __shared__ double lds_d[];
    T *lds = reinterpret_cast<T *>(lds_d);
    if (state->done) return;
    const Tile t = make_tile(blockIdx.x, blockIdx.y, 4 * nu + 2);
    T *sz = lds, *sb = lds + t.RX * t.RY;
    tile_load<T, true>(F, zero_z ? nullptr : F.z, nullptr, CCOp{}, sz, sb, t);
    __syncthreads();
    for (int s = 1; s <= 4 * nu; s++) {
        const int c = (s - 1) & 3;
        fem_tile_sweep<T>(F, op, sz, sb, t, s, c & 1, c >> 1);
        __syncthreads();
    }
    tile_store<T>(F, F.zalt, sz, t);
    const int I0 = t.x0 >> 1, J0 = t.y0 >> 1;
    const int I1 = min(C.m, I0 + kTX / 2), J1 = min(C.n, J0 + kTY / 2);
    for (int J = J0 + threadIdx.y; J < J1; J += kBY)
        for (int I = I0 + threadIdx.x; I < I1; I += kBX) {
            T v = 0;
            if (J > 0) {
                for (int b2 = -1; b2 <= 1; b2++) {
                    const int gy = 2 * J + b2;
                    if (gy >= F.n) continue;
                    T sx = 0;
                    for (int a2 = -1; a2 <= 1; a2++) {
                        const int gx = 2 * I + a2;
                        if (gx < 0 || gx >= F.m) continue;
                        const int l = (gy - t.gy0) * t.RX + gx - t.gx0;
                        sx += (a2 ? T(0.5) : T(1)) * fem_resid_t<T>(op, F.m, F.n, sz, sb, t.RX, l, gx, gy);
                    }
                    v += (b2 ? T(0.5) : T(1)) * sx;
                }
            }
            C.b[I + (size_t)C.m * J] = v;
        }
}

template <class T>
__global__ void __launch_bounds__(256) k_fem_up(LevelT<T> F, LevelT<T> C, FemOp op, int nu, int norm, MGCtl ctl,
                                                double tol2) {
    extern __shared__ double lds_d[];
    T *lds = reinterpret_cast<T *>(lds_d);
    __shared__ double sh[8];
    if (ctl.state->done) return;
    const Tile t = make_tile(blockIdx.x, blockIdx.y, 4 * nu + norm);
    T *sz = lds, *sb = lds + t.RX * t.RY;
    tile_load<T, true>(F, F.zalt, &C, CCOp{}, sz, sb, t);
    __syncthreads();
    for (int s = 1; s <= 4 * nu; s++) {
        const int c = (s - 1) & 3;
        fem_tile_sweep<T>(F, op, sz, sb, t, s, c & 1, c >> 1);
        __syncthreads();
    }
    tile_store<T>(F, F.z, sz, t);
    if (!norm) return;
    double rr = 0;
    for (int y = threadIdx.y; y < kTY; y += kBY) {
        const int gy = t.y0 + y;
        if (gy >= F.n) break;
        for (int x = threadIdx.x; x < kTX; x += kBX) {
            const int gx = t.x0 + x;
            if (gx >= F.m) break;
            const double r = fem_resid_t<T>(op, F.m, F.n, sz, sb, t.RX, (y + t.H) * t.RX + x + t.H, gx, gy);
            rr += r * r;
        }
    }
    finish_norm(ctl, rr, gridDim.x * gridDim.y, tol2, sh);
}

// Mixed form, level 0, end of an inner FP32 cycle: prolongate and smooth the
// FP32 correction e (pre-smoothed iterate in E.zalt, right-hand side r32 in
// E.b), then z_out = z_in + e and the new FP64 residual r = b - A z_out,
// which is reduced for the convergence test and stored in r_out as the next
// cycle's right-hand side.  Inputs and outputs are distinct arrays.
template <bool FEM>
__global__ void __launch_bounds__(256) k_mx_up0(Level Z, const double *z_in, double *z_out, LevelF E, float *r_out,
                                                LevelF C, CCOp cop, FemOp fop, int nu, MGCtl ctl, double tol2) {
    extern __shared__ double lds_d[];
    __shared__ double sh[8];
    if (ctl.state->done) return;
    const Tile t = make_tile(blockIdx.x, blockIdx.y, (FEM ? 4 : 2) * nu + 1);
    const int R = t.RX * t.RY;
    double *zd = lds_d;
    float *se = reinterpret_cast<float *>(lds_d + R), *sr = se + R;
    tile_load<float, FEM>(E, E.zalt, &C, cop, se, sr, t);
    __syncthreads();
    for (int s = 1; s <= (FEM ? 4 : 2) * nu; s++) {
        if (FEM) {
            const int c = (s - 1) & 3;
            fem_tile_sweep<float>(E, fop, se, sr, t, s, c & 1, c >> 1);
        } else {
            cc_tile_sweep<float>(E, cop, se, sr, t, s, (s - 1) & 1);
        }
        __syncthreads();
    }
    // z + e on the tile and a one-cell ring (zero outside the domain)
    for (int ly = t.H - 1 + threadIdx.y; ly < t.H + kTY + 1; ly += kBY) {
        const int gy = t.gy0 + ly;
        for (int lx = t.H - 1 + threadIdx.x; lx < t.H + kTX + 1; lx += kBX) {
            const int gx = t.gx0 + lx;
            const bool in = gy >= 0 && gy < Z.n && gx >= 0 && gx < Z.m;
            const int l = ly * t.RX + lx;
            zd[l] = in ? z_in[gx + (size_t)Z.m * gy] + (double)se[l] : 0.0;
        }
    }
    __syncthreads();
    double rr = 0;
    for (int y = threadIdx.y; y < kTY; y += kBY) {
        const int gy = t.y0 + y;
        if (gy >= Z.n) break;
        for (int x = threadIdx.x; x < kTX; x += kBX) {
            const int gx = t.x0 + x;
            if (gx >= Z.m) break;
            const int l = (y + t.H) * t.RX + x + t.H;
            const size_t g = gx + (size_t)Z.m * gy;
            double az;
            if (FEM) {
                const double acc = fem_acc<double>(fop, Z.m, Z.n, gx, gy);
                az = gy == 0 ? acc * zd[l] : acc * zd[l] + fem_off_t<double>(fop, Z.m, Z.n, zd, l, t.RX, gx, gy);
            } else {
                az = cop.s * cc_off_t<double>(Z.cx, Z.cy, zd, l, t.RX) +
                     cc_diag<double>(cop, Z.m, Z.n, Z.cx, Z.cy, gx, gy) * zd[l];
            }
            const double r = Z.b[g] - az;
            z_out[g] = zd[l];
            r_out[g] = (float)r;
            rr += r * r;
        }
    }
    finish_norm(ctl, rr, gridDim.x * gridDim.y, tol2, sh);
}

// ------------------------------------------------------------------------
// Coarse levels: one workgroup, all levels in LDS
// ------------------------------------------------------------------------

template <class T>
__device__ void coarse_sweeps(bool fem, const CoarseDesc &d, const CCOp &cop, const FemOp &fop, int l, T *lds,
                              int count) {
    const int m = d.m[l], n = d.n[l], N = m * n;
    T *z = lds + d.off[l], *b = z + N;
    const T cx = T(d.cx[l]), cy = T(d.cy[l]);
    const int ncol = fem ? 4 : 2;
    for (int k = 0; k < count; k++)
        for (int c = 0; c < ncol; c++) {
            for (int ij = threadIdx.x; ij < N; ij += blockDim.x) {
                const int i = ij % m, j = ij / m;
                if (fem) {
                    if ((i & 1) != (c & 1) || (j & 1) != (c >> 1) || j == 0) continue;
                    z[ij] = (b[ij] - fem_off<T>(fop, m, n, z, i, j)) / fem_acc<T>(fop, m, n, i, j);
                } else {
                    if (((i + j) & 1) != c) continue;
                    z[ij] = (b[ij] - T(cop.s) * cc_off<T>(m, n, cx, cy, z, i, j)) / cc_diag<T>(cop, m, n, cx, cy, i, j);
                }
            }
            __syncthreads();
        }
}

template <class T>
__global__ void __launch_bounds__(kCoarseThreads) k_coarse(int fem, CoarseDesc d, CCOp cop, FemOp fop, int nu1,
                                                            int nu2, int ncs, MGCtl ctl, double tol2) {
    extern __shared__ double lds_d[];
    T *lds = reinterpret_cast<T *>(lds_d);
    __shared__ double sh[8];
    if (ctl.state->done) return;
    const int Tn = blockDim.x;
    const T *gb = reinterpret_cast<const T *>(d.gb);
    T *gz = reinterpret_cast<T *>(d.gz);
    {
        const int N = d.m[0] * d.n[0];
        T *z = lds + d.off[0], *b = z + N;
        for (int k = threadIdx.x; k < N; k += Tn) {
            b[k] = gb[k];
            z[k] = d.top ? gz[k] : T(0);
        }
    }
    __syncthreads();
    for (int l = 0; l + 1 < d.nl; l++) {
        coarse_sweeps<T>(fem, d, cop, fop, l, lds, nu1);
        const int m = d.m[l], n = d.n[l], Mc = d.m[l + 1], Nc = d.n[l + 1];
        const T *z = lds + d.off[l], *b = z + m * n;
        T *zc = lds + d.off[l + 1], *bc = zc + Mc * Nc;
        const T cx = T(d.cx[l]), cy = T(d.cy[l]);
        for (int k = threadIdx.x; k < Mc * Nc; k += Tn) {
            const int I = k % Mc, J = k / Mc;
            T v = 0;
            if (!fem) {
                for (int b2 = 0; b2 < 2; b2++)
                    for (int a2 = 0; a2 < 2; a2++) {
                        const int i = 2 * I + a2, j = 2 * J + b2;
                        v += b[i + m * j] - Az_plain<T>(false, cop, fop, m, n, cx, cy, z, i, j);
                    }
                v *= T(0.25);
            } else if (J > 0) {
                for (int b2 = -1; b2 <= 1; b2++) {
                    const int j = 2 * J + b2;
                    if (j >= n) continue;
                    T sx = 0;
                    for (int a2 = -1; a2 <= 1; a2++) {
                        const int i = 2 * I + a2;
                        if (i < 0 || i >= m) continue;
                        sx += (a2 ? T(0.5) : T(1)) * (b[i + m * j] - Az_plain<T>(true, cop, fop, m, n, T(0), T(0), z, i, j));
                    }
                    v += (b2 ? T(0.5) : T(1)) * sx;
                }
            }
            bc[k] = v;
            zc[k] = 0;
        }
        __syncthreads();
    }
    coarse_sweeps<T>(fem, d, cop, fop, d.nl - 1, lds, ncs);
    for (int l = d.nl - 2; l >= 0; l--) {
        const int m = d.m[l], n = d.n[l], Mc = d.m[l + 1], Nc = d.n[l + 1];
        T *z = lds + d.off[l];
        const T *ec = lds + d.off[l + 1];
        for (int k = threadIdx.x; k < m * n; k += Tn) {
            const int i = k % m, j = k / m;
            if (fem) { if (j > 0) z[k] += fem_prolong_at<T>(ec, Mc, i, j); }
            else z[k] += cc_prolong_at<T>(cop, Mc, Nc, ec, i, j);
        }
        __syncthreads();
        coarse_sweeps<T>(fem, d, cop, fop, l, lds, nu2);
    }
    {
        const int N = d.m[0] * d.n[0];
        const T *z = lds + d.off[0], *b = z + N;
        for (int k = threadIdx.x; k < N; k += Tn) gz[k] = z[k];
        if (!d.top) return;
        double rr = 0;
        for (int k = threadIdx.x; k < N; k += Tn) {
            const int i = k % d.m[0], j = k / d.m[0];
            const double r = b[k] - Az_plain<T>(fem, cop, fop, d.m[0], d.n[0], T(d.cx[0]), T(d.cy[0]), z, i, j);
            rr += r * r;
        }
        finish_norm(ctl, rr, 1, tol2, sh);
    }
}

// ------------------------------------------------------------------------
// Norms, Dirichlet row, extrapolation, operator application
// ------------------------------------------------------------------------

// ||b - A z||^2 and ||b||^2 of level 0; the last workgroup stores them and
// sets the flag: 1 converged, 3 zero right-hand side, 0 otherwise.  With r32
// non-null (mixed form) the residual is also stored in FP32.
__global__ void __launch_bounds__(256) k_begin(int fem, Level L, CCOp cop, FemOp fop, MGCtl ctl, double tol2,
                                               float *r32) {
    __shared__ double sh[8];
    __shared__ int last;
    double rr = 0, bb = 0;
    const long N = (long)L.m * L.n;
    for (long ij = blockIdx.x * (long)blockDim.x + threadIdx.x; ij < N; ij += (long)gridDim.x * blockDim.x) {
        const int i = (int)(ij % L.m), j = (int)(ij / L.m);
        const double bv = L.b[ij], r = bv - Az_plain<double>(fem, cop, fop, L.m, L.n, L.cx, L.cy, L.z, i, j);
        if (r32) r32[ij] = (float)r;
        rr += r * r;
        bb += bv * bv;
    }
    rr = block_sum(rr, sh);
    __syncthreads();
    bb = block_sum(bb, sh);
    if (threadIdx.x == 0) {
        ctl.partial[2 * blockIdx.x] = rr;
        ctl.partial[2 * blockIdx.x + 1] = bb;
        __threadfence();
        unsigned tk = atomicAdd(&ctl.state->ticket, 1u);
        last = (tk == gridDim.x - 1);
    }
    __syncthreads();
    if (last && threadIdx.x == 0) {
        __threadfence();
        double R = 0, B = 0;
        for (int k = 0; k < (int)gridDim.x; k++) {
            R += __hip_atomic_load(&ctl.partial[2 * k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            B += __hip_atomic_load(&ctl.partial[2 * k + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        MGState *st = ctl.state;
        st->ticket = 0;
        st->rr = R;
        st->bb = B;
        st->cycles = 0;
        st->done = (B == 0) ? 3 : (!(R == R) ? 2 : (R <= tol2 * B ? 1 : 0));
    }
}

__global__ void k_fem_dirichlet(Level L, FemOp op) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.m) return;
    L.z[i] = L.b[i] / fem_acc<double>(op, L.m, L.n, i, 0);
}

__global__ void k_apply(int fem, Level L, CCOp cop, FemOp fop, double *out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= L.m || j >= L.n) return;
    out[i + (size_t)L.m * j] = Az_plain<double>(fem, cop, fop, L.m, L.n, L.cx, L.cy, L.z, i, j);
}

// mgs_common.cc:37-44: z <- b z - c z_prev, z_prev <- old z.
__global__ void k_extrapolate(double *z, double *zp, long N, double bb, double c) {
    long ij = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (ij >= N) return;
    double zz = z[ij];
    z[ij] = bb * zz - c * zp[ij];
    zp[ij] = zz;
}

// Same update with the coefficients (b, c) read from device memory, so a
// captured step graph can be replayed with new times.
__global__ void k_extrapolate_dev(double *z, double *zp, long N, const double *coef) {
    long ij = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (ij >= N) return;
    const double bb = coef[0], c = coef[1];
    double zz = z[ij];
    z[ij] = bb * zz - c * zp[ij];
    zp[ij] = zz;
}

// End of a solve whose cycle count is only known on the device: the mixed
// form leaves its iterate in zalt after an odd number of cycles, and a zero
// right-hand side means z = 0.
__global__ void k_solve_end(const MGState *st, double *z, const double *zalt, long N, int mixed) {
    long ij = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (ij >= N) return;
    if (st->done == 3) z[ij] = 0;
    else if (mixed && (st->cycles & 1)) z[ij] = zalt[ij];
}

__global__ void k_zero(double *z, long N) {
    long ij = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (ij < N) z[ij] = 0;
}

__global__ void k_reset_state(MGState *st) {
    st->done = 0;
    st->cycles = 0;
    st->ticket = 0;
}

// ------------------------------------------------------------------------
// Driver
// ------------------------------------------------------------------------

static size_t tile_lds_bytes(int H, size_t elem) { return 2 * elem * (size_t)(kTX + 2 * H) * (kTY + 2 * H); }

int MG::init(int kind_, int m, int n, double cx, double cy, int x_sym, double *z0, double *b0, hipStream_t stream) {
    kind = kind_;
    fem = kind == VDI_FEM;
    st = stream;
    if (m < 2 || n < 2) return VDI_ERR_ARG;
    if (kind == VDI_MAC) {
        cc = CCOp{0.0, 1.0, BC_N, BC_N, BC_DF, BC_N};
    } else if (kind == VDI_VX || kind == VDI_VY) {
        int bl = x_sym ? (kind == VDI_VX ? BC_DF : BC_N) : BC_DC;
        cc = CCOp{1.0, -1.0, bl, BC_DC, BC_N, BC_DC};
        vfx = cx;
        vfy = cy;
    } else if (kind == VDI_FEM) {
        double dxdy = cx / cy, dydx = cy / cx;
        fe.fm = 4. / 3. * (dxdy + dydx);
        fe.fey = 1. / 3. * (-2 * dxdy + dydx);
        fe.hey = 0.5 * fe.fey;
        fe.fex = 1. / 3. * (-2 * dydx + dxdy);
        fe.hex = 0.5 * fe.fex;
        fe.fc = -1. / 6. * (dxdy + dydx);
    } else {
        return VDI_ERR_ARG;
    }
    lm.clear();
    ln.clear();
    int mc = fem ? m - 1 : m, nc = fem ? n - 1 : n;
    for (;;) {
        lm.push_back(fem ? mc + 1 : mc);
        ln.push_back(fem ? nc + 1 : nc);
        if ((int)lm.size() >= 24) break;
        if (mc % 2 || nc % 2 || mc / 2 < 2 || nc / 2 < 2) break;
        mc /= 2;
        nc /= 2;
    }
    nl = (int)lm.size();
    lc = nl - 1;
    while (lc > 0 && (long)lm[lc - 1] * ln[lc - 1] <= kCoarsePoints) lc--;
    if (nl - lc > kMaxCoarse) return VDI_ERR_ARG;
    size_t coarse = 0;
    for (int l = 1; l < nl; l++) coarse += (size_t)lm[l] * ln[l];
    const long N0 = (long)m * n;
    // FP64: zalt on level 0; z, b, zalt on every coarser level
    VDI_TRY(store_.alloc((N0 + 3 * coarse) * sizeof(double)));
    // FP32 (mixed form): level 0 zalt and two r32 buffers; z, b, zalt below
    VDI_TRY(store_f_.alloc((3 * (size_t)N0 + 3 * coarse) * sizeof(float)));
    if (!z0) { VDI_TRY(own_z_.alloc(N0 * sizeof(double))); z0 = own_z_.as<double>(); }
    if (!b0) { VDI_TRY(own_b_.alloc(N0 * sizeof(double))); b0 = own_b_.as<double>(); }
    VDI_TRY(zprev_.alloc(N0 * sizeof(double)));
    L.assign(nl, Level{});
    LF.assign(nl, LevelF{});
    double *p = store_.as<double>();
    float *q = store_f_.as<float>();
    for (int l = 0; l < nl; l++) {
        L[l].m = LF[l].m = lm[l];
        L[l].n = LF[l].n = ln[l];
        const size_t Nl = (size_t)lm[l] * ln[l];
        if (l == 0) {
            L[l].z = z0;
            L[l].b = b0;
            L[l].zalt = p; p += Nl;
            LF[l].z = nullptr;
            LF[l].zalt = q; q += Nl;
            r32_[0] = q; q += Nl;
            r32_[1] = q; q += Nl;
            LF[l].b = r32_[0];
        } else {
            L[l].z = p; p += Nl;
            L[l].b = p; p += Nl;
            L[l].zalt = p; p += Nl;
            LF[l].z = q; q += Nl;
            LF[l].b = q; q += Nl;
            LF[l].zalt = q; q += Nl;
        }
    }
    set_coeffs(cx, cy);
    long maxblk = std::max<long>(1024, (long)cdiv(lm[0], kTX) * cdiv(ln[0], kTY));
    VDI_TRY(partial_.alloc(2 * maxblk * sizeof(double)));
    VDI_TRY(state_.alloc(sizeof(MGState)));
    if (!host_state_) VDI_HIP(hipHostMalloc((void **)&host_state_, sizeof(MGState), hipHostMallocDefault));
    coarse_lds_ = 0;
    for (int l = lc; l < nl; l++) coarse_lds_ += 2 * sizeof(double) * (size_t)lm[l] * ln[l];
    static bool attr_set = false;
    if (!attr_set) {
        int maxlds = 0;
        VDI_HIP(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, 0));
        const size_t big = std::max<size_t>(tile_lds_bytes(4 * kMaxNu + 2, sizeof(double)),
                                            2 * sizeof(double) * kCoarsePoints * 2);
        if ((size_t)maxlds < big) return VDI_ERR_ARG;
        const void *ks[] = {(const void *)k_cc_down<double>, (const void *)k_cc_up<double>,
                            (const void *)k_fem_down<double>, (const void *)k_fem_up<double>,
                            (const void *)k_coarse<double>, (const void *)k_cc_down<float>,
                            (const void *)k_cc_up<float>, (const void *)k_fem_down<float>,
                            (const void *)k_fem_up<float>, (const void *)k_coarse<float>,
                            (const void *)k_mx_up0<false>, (const void *)k_mx_up0<true>};
        for (const void *k : ks) VDI_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)big));
        attr_set = true;
    }
    return VDI_OK;
}

MG::~MG() {
    if (host_state_) (void)hipHostFree(host_state_);
}

void MG::set_coeffs(double cx, double cy) {
    for (int l = 0; l < nl; l++) {
        L[l].cx = cx;
        L[l].cy = cy;
        LF[l].cx = (float)cx;
        LF[l].cy = (float)cy;
        if (fem) {
            L[l].inv_int = 1.0 / fe.fm;
        } else {
            L[l].inv_int = 1.0 / (cc.sigma - cc.s * 2.0 * (cx + cy));
        }
        LF[l].inv_int = (float)L[l].inv_int;
        cx *= 0.25;
        cy *= 0.25;
    }
}

MGCtl MG::ctl() const {
    MGCtl c;
    c.state = state_.as<MGState>();
    c.partial = partial_.as<double>();
    return c;
}

template <class T>
static CoarseDesc coarse_desc(const std::vector<LevelT<T>> &Lv, int lc, int nl) {
    CoarseDesc d{};
    d.nl = nl - lc;
    int off = 0;
    for (int k = 0; k < d.nl; k++) {
        const LevelT<T> &v = Lv[lc + k];
        d.m[k] = v.m;
        d.n[k] = v.n;
        d.cx[k] = (double)v.cx;
        d.cy[k] = (double)v.cy;
        d.off[k] = off;
        off += 2 * v.m * v.n;
    }
    d.gz = reinterpret_cast<double *>(Lv[lc].z);
    d.gb = reinterpret_cast<double *>(Lv[lc].b);
    d.top = lc == 0;
    return d;
}

// Enqueues the down kernel of level l of hierarchy H (FP64 or FP32).
template <class T>
int MG::launch_down(const std::vector<LevelT<T>> &H, int l, int zero_z) {
    const LevelT<T> &F = H[l], &C = H[l + 1];
    const dim3 blk(kBX, kBY);
    const int slot = timer ? timer->begin(l == 0 ? VDI_KCLASS_MG_DOWN0 : VDI_KCLASS_MG_MID, st) : -1;
    if (!fem) {
        dim3 g(cdiv(F.m, kTX), cdiv(F.n, kTY));
        hipLaunchKernelGGL(k_cc_down<T>, g, blk, tile_lds_bytes(2 * nu1 + 1, sizeof(T)), st, F, C, cc, nu1, zero_z,
                           (const MGState *)state_.as<MGState>());
    } else {
        dim3 g(cdiv(C.m, kTX / 2), cdiv(C.n, kTY / 2));
        hipLaunchKernelGGL(k_fem_down<T>, g, blk, tile_lds_bytes(4 * nu1 + 2, sizeof(T)), st, F, C, fe, nu1, zero_z,
                           (const MGState *)state_.as<MGState>());
    }
    VDI_LAUNCHED();
    // reads z (unless the level starts from zero) and b; writes zalt and coarse b
    if (timer) timer->end(slot, st, sizeof(T) * ((zero_z ? 2.0 : 3.0) * F.m * F.n + 1.0 * C.m * C.n));
    return VDI_OK;
}

template <class T>
int MG::launch_up(const std::vector<LevelT<T>> &H, int l, int norm, double tol2) {
    const LevelT<T> &F = H[l], &C = H[l + 1];
    const dim3 blk(kBX, kBY);
    const int slot = timer ? timer->begin(l == 0 ? VDI_KCLASS_MG_UP0 : VDI_KCLASS_MG_MID, st) : -1;
    dim3 g(cdiv(F.m, kTX), cdiv(F.n, kTY));
    if (!fem)
        hipLaunchKernelGGL(k_cc_up<T>, g, blk, tile_lds_bytes(2 * nu2 + norm, sizeof(T)), st, F, C, cc, nu2, norm,
                           ctl(), tol2);
    else
        hipLaunchKernelGGL(k_fem_up<T>, g, blk, tile_lds_bytes(4 * nu2 + norm, sizeof(T)), st, F, C, fe, nu2, norm,
                           ctl(), tol2);
    VDI_LAUNCHED();
    // reads zalt, b, coarse correction; writes z
    if (timer) timer->end(slot, st, sizeof(T) * (3.0 * F.m * F.n + 1.0 * C.m * C.n));
    return VDI_OK;
}

template <class T>
int MG::launch_coarse(const std::vector<LevelT<T>> &H, double tol2) {
    CoarseDesc d = coarse_desc<T>(H, lc, nl);
    const int slot = timer ? timer->begin(VDI_KCLASS_MG_COARSE, st) : -1;
    hipLaunchKernelGGL(k_coarse<T>, dim3(1), dim3(kCoarseThreads), coarse_lds_ / sizeof(double) * sizeof(T), st,
                       (int)fem, d, cc, fe, nu1, nu2, ncs, ctl(), tol2);
    VDI_LAUNCHED();
    if (timer) timer->end(slot, st, sizeof(T) * (d.top ? 3.0 : 2.0) * d.m[0] * d.n[0]);
    return VDI_OK;
}

bool MG::mixed_active() const { return form == VDI_FORM_MIXED && lc > 0; }

int MG::launch_cycle(double tol2) {
    if (!mixed_active()) {
        // coarser levels hold corrections, which start from zero
        for (int l = 0; l < lc; l++) VDI_TRY(launch_down<double>(L, l, l > 0));
        VDI_TRY(launch_coarse<double>(L, tol2));
        for (int l = lc - 1; l >= 0; l--) VDI_TRY(launch_up<double>(L, l, l == 0, tol2));
        return VDI_OK;
    }
    // Mixed form: one FP32 V-cycle on A e = r32 from e = 0, then z += e and
    // the FP64 residual (k_mx_up0).  z and r32 alternate between two
    // buffers from one cycle to the next.
    const int par = mx_enqueued_ & 1;
    double *z_in = par ? L[0].zalt : L[0].z, *z_out = par ? L[0].z : L[0].zalt;
    float *r_in = r32_[par], *r_out = r32_[par ^ 1];
    LF[0].b = r_in;
    for (int l = 0; l < lc; l++) VDI_TRY(launch_down<float>(LF, l, 1));
    VDI_TRY(launch_coarse<float>(LF, tol2));
    for (int l = lc - 1; l >= 1; l--) VDI_TRY(launch_up<float>(LF, l, 0, tol2));
    const dim3 blk(kBX, kBY);
    const int H = (fem ? 4 : 2) * nu2 + 1;
    const size_t lds = (size_t)(kTX + 2 * H) * (kTY + 2 * H) * (sizeof(double) + 2 * sizeof(float));
    const int slot = timer ? timer->begin(VDI_KCLASS_MG_UP0, st) : -1;
    dim3 g(cdiv(L[0].m, kTX), cdiv(L[0].n, kTY));
    if (fem)
        hipLaunchKernelGGL(k_mx_up0<true>, g, blk, lds, st, L[0], (const double *)z_in, z_out, LF[0], r_out, LF[1], cc,
                           fe, nu2, ctl(), tol2);
    else
        hipLaunchKernelGGL(k_mx_up0<false>, g, blk, lds, st, L[0], (const double *)z_in, z_out, LF[0], r_out, LF[1], cc,
                           fe, nu2, ctl(), tol2);
    VDI_LAUNCHED();
    // reads e (zalt), r32, coarse e, z64, b64; writes z64, r32
    const double N0 = (double)L[0].m * L[0].n;
    if (timer) timer->end(slot, st, 4.0 * (3.0 * N0 + (double)LF[1].m * LF[1].n) + 8.0 * 3.0 * N0);
    mx_enqueued_++;
    return VDI_OK;
}

// After a mixed-form solve, the solution is in L[0].zalt when an odd number
// of cycles ran; bring it back to L[0].z.
int MG::mixed_finish(int executed) {
    if (!mixed_active() || !(executed & 1)) return VDI_OK;
    VDI_HIP(hipMemcpyAsync(L[0].z, L[0].zalt, sizeof(double) * unknowns(), hipMemcpyDeviceToDevice, st));
    return VDI_OK;
}

double MG::cycle_bytes() const {
    auto N = [&](int l) { return (double)lm[l] * ln[l]; };
    double b = 0;
    if (!mixed_active()) {
        for (int l = 0; l < lc; l++) b += 8.0 * ((l == 0 ? 3.0 : 2.0) * N(l) + N(l + 1));
        b += 8.0 * (lc == 0 ? 3.0 : 2.0) * N(lc);
        for (int l = 0; l < lc; l++) b += 8.0 * (3.0 * N(l) + N(l + 1));
    } else {
        for (int l = 0; l < lc; l++) b += 4.0 * (2.0 * N(l) + N(l + 1));
        b += 4.0 * 2.0 * N(lc);
        for (int l = 1; l < lc; l++) b += 4.0 * (3.0 * N(l) + N(l + 1));
        b += 4.0 * (3.0 * N(0) + N(1)) + 8.0 * 3.0 * N(0);
    }
    return b;
}

int MG::enforce_dirichlet() {
    if (!fem) return VDI_OK;
    hipLaunchKernelGGL(k_fem_dirichlet, dim3(cdiv(L[0].m, 256)), dim3(256), 0, st, L[0], fe);
    VDI_LAUNCHED();
    return VDI_OK;
}

int MG::begin(double tol) {
    VDI_TRY(enforce_dirichlet());
    mx_enqueued_ = 0;
    const long N = (long)L[0].m * L[0].n;
    const int nb = (int)std::min<long>(1024, (N + 255) / 256);
    hipLaunchKernelGGL(k_begin, dim3(nb), dim3(256), 0, st, (int)fem, L[0], cc, fe, ctl(), tol * tol,
                       mixed_active() ? r32_[0] : (float *)nullptr);
    VDI_LAUNCHED();
    return VDI_OK;
}

int MG::read_state(MGState *out) {
    VDI_HIP(hipMemcpyAsync(host_state_, state_.p, sizeof(MGState), hipMemcpyDeviceToHost, st));
    VDI_HIP(hipStreamSynchronize(st));
    *out = *host_state_;
    return VDI_OK;
}

int MG::residual(double *rnorm, double *bnorm) {
    VDI_TRY(begin(0.0));
    MGState s;
    VDI_TRY(read_state(&s));
    *rnorm = std::sqrt(s.rr);
    *bnorm = std::sqrt(s.bb);
    hipLaunchKernelGGL(k_reset_state, dim3(1), dim3(1), 0, st, state_.as<MGState>());
    VDI_LAUNCHED();
    return VDI_OK;
}

int MG::vcycles(int k) {
    // begin() also prepares the FP32 residual of the mixed form; the flag it
    // raises is cleared so that exactly k cycles run.
    VDI_TRY(begin(0.0));
    hipLaunchKernelGGL(k_reset_state, dim3(1), dim3(1), 0, st, state_.as<MGState>());
    VDI_LAUNCHED();
    for (int c = 0; c < k; c++) VDI_TRY(launch_cycle(-1.0));
    return mixed_finish(k);
}

int MG::solve(double tol, int max_cycles, int *cycles) {
    const double tol2 = tol * tol;
    VDI_TRY(begin(tol));
    const int pred = std::max(1, std::min(max_cycles, last_cycles));
    for (int c = 0; c < pred; c++) VDI_TRY(launch_cycle(tol2));
    MGState s;
    VDI_TRY(read_state(&s));
    while (s.done == 0) {
        if (s.cycles >= max_cycles) break;
        VDI_TRY(launch_cycle(tol2));
        VDI_TRY(read_state(&s));
    }
    last_cycles = s.cycles;
    if (cycles) *cycles = s.cycles;
    VDI_TRY(mixed_finish(s.cycles));
    if (s.done == 3) {
        const long N = (long)L[0].m * L[0].n;
        hipLaunchKernelGGL(k_zero, dim3(cdiv(N, 256)), dim3(256), 0, st, L[0].z, N);
        VDI_LAUNCHED();
        last_cycles = 0;
        if (cycles) *cycles = 0;
    } else if (s.done != 1) {
        return VDI_ERR_MG_NOCONV;
    }
    total_cycles += last_cycles;
    solves++;
    return VDI_OK;
}

int MG::initial_guess(double t) {
    long N = unknowns();
    if (scount < 2) {
        if (scount == 1) {
            VDI_HIP(hipMemcpyAsync(zprev_.p, L[0].z, N * sizeof(double), hipMemcpyDeviceToDevice, st));
            t0 = t;
        } else {
            t_prev = t;
        }
        scount++;
        return VDI_OK;
    }
    if (std::fabs(t0 - t_prev) < std::numeric_limits<double>::epsilon() * 10) {
        fputs("Multigrid routine called for duplicate times\n", stderr);
        return VDI_OK;
    }
    double a = 1. / (t0 - t_prev), bb = (t - t_prev) * a, c = (t - t0) * a;
    hipLaunchKernelGGL(k_extrapolate, dim3(cdiv(N, 256)), dim3(256), 0, st, L[0].z, zprev_.as<double>(), N, bb, c);
    VDI_LAUNCHED();
    t_prev = t0;
    t0 = t;
    return VDI_OK;
}

int MG::guess_coeffs(double t, double *bb, double *c) {
    if (scount < 2) {
        if (scount == 1) t0 = t;
        else t_prev = t;
        scount++;
        return 0;  // the caller copies z to z_prev when scount was 1
    }
    if (std::fabs(t0 - t_prev) < std::numeric_limits<double>::epsilon() * 10) {
        fputs("Multigrid routine called for duplicate times\n", stderr);
        return -1;
    }
    double a = 1. / (t0 - t_prev);
    *bb = (t - t_prev) * a;
    *c = (t - t0) * a;
    t_prev = t0;
    t0 = t;
    return 1;
}

int MG::enqueue_extrapolate(const double *coef) {
    const long N = unknowns();
    hipLaunchKernelGGL(k_extrapolate_dev, dim3(cdiv(N, 256)), dim3(256), 0, st, L[0].z, zprev_.as<double>(), N, coef);
    VDI_LAUNCHED();
    return VDI_OK;
}

int MG::enqueue_fixed(double tol, int budget) {
    VDI_TRY(begin(tol));
    for (int c = 0; c < budget; c++) VDI_TRY(launch_cycle(tol * tol));
    const long N = unknowns();
    hipLaunchKernelGGL(k_solve_end, dim3(cdiv(N, 256)), dim3(256), 0, st, (const MGState *)state_.as<MGState>(),
                       L[0].z, (const double *)L[0].zalt, N, (int)mixed_active());
    VDI_LAUNCHED();
    return VDI_OK;
}

int MG::apply(const double *z, double *out) {
    Level Lz = L[0];
    Lz.z = const_cast<double *>(z);
    dim3 g(cdiv(Lz.m, 64), cdiv(Lz.n, 4));
    hipLaunchKernelGGL(k_apply, g, dim3(64, 4), 0, st, (int)fem, Lz, cc, fe, out);
    VDI_LAUNCHED();
    return VDI_OK;
}

}  // namespace vdi
