// mg.hip -- multigrid kernels and driver (see mg.hpp for the scheme).
//
// Kernel structure of one V-cycle on an L-level hierarchy whose coarsest
// level lc = L-1 holds at most kExactPoints unknowns:
//   down(l), l = 0..lc-1 : nu1 smoothing sweeps of level l, residual,
//                          restriction into level l+1   (one kernel per level)
//   coarse               : the exact solve of level lc (separation of
//                          variables) inside one workgroup's LDS
//   up(l), l = lc-1..0   : prolongation + correction, nu2 sweeps; at level 0
//                          also per-workgroup partials of ||b - A z||^2
//                          (level 0 takes V(nu1, nu2), every restricted level
//                          V(nu1s, nu2s))
//   norm                 : one workgroup sums the partials in a fixed order
//                          and raises the device-side convergence flag.
// Every kernel returns at once when the flag is set, so a solve enqueues the
// predicted number of cycles and synchronises with the host once.
//
// The down/up kernels are temporally blocked: a workgroup loads its
// tx x ty output tile plus a halo of H cells into LDS, performs the sweeps
// on a region that shrinks by one cell per colour pass (so every update sees
// exactly the values a global sweep would), and writes back only the tile.
// Halo: cell-centred red-black H = 2 nu (+1 for a residual); nodal four-
// colour H = 4 nu (+1 residual, +1 more for full-weighting restriction).
// Values outside the domain are zero in LDS, which is exactly how both
// operators drop missing neighbours.  The tile shape is a launch argument
// chosen per level by the autotuner.
//
// Workgroup-to-tile order is XCD-aware: gfx950 hands consecutive workgroups
// of a launch to its 8 XCDs in turn, so tile = xcd_tile(workgroup) gives each
// XCD one contiguous band of tiles, whose shared halos then hit that XCD's
// own L2.
#include "mg.hpp"

#include "coarse_sep.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace vdi {

// The last five fit the nodal grids (2^k + 1 nodes a side), whose extra
// node adds a column and a row of nearly empty tiles to the power-of-two
// shapes: at 1025 x 513, 56 x 40 gives 247 tiles (one per CU) where 64 x 32
// gives 289 and 64 x 64 gives 153; at 513 x 257, 28 x 20 gives 247.
// The last eight are streaming launches (segments of 16 .. 128 rows, then 20,
// 40, 12 and 10).  A launch is fastest when its wavefronts come to just under a
// whole number per SIMD: at 2048 x 1024, 16 rows give a few more wavefronts
// than the SIMDs and 20 rows fewer; at 4096 x 2048, 32 rows give 2.25 per
// SIMD and 40 rows 1.8.
const TileShape kTileShapes[kNumTileShapes] = {{64, 32, 0},  {32, 8, 0},   {32, 16, 0},  {64, 8, 0},  {64, 16, 0},
                                               {32, 32, 0},  {128, 16, 0}, {128, 32, 0}, {64, 64, 0}, {56, 40, 0},
                                               {72, 36, 0},  {48, 24, 0},  {28, 20, 0},  {36, 16, 0}, {128, 16, 1},
                                               {128, 32, 1}, {128, 64, 1}, {128, 128, 1}, {128, 20, 1}, {128, 40, 1}, {128, 12, 1}, {128, 10, 1}};

// Position b of n work items -> item index such that the items of one XCD
// (b mod 8) are contiguous.
__device__ __forceinline__ int xcd_tile(int b, int n) {
    const int q = n >> 3, r = n & 7, x = b & 7, k = b >> 3;
    return x * q + min(x, r) + k;
}

// ------------------------------------------------------------------------
// Operators on plain arrays (mgs_mac.hh:37-50, visco_impl.hh:32-53,
// mgs_fem.hh:42-56 + mgs_fem.cc:34-50), in the level's value type T
// ------------------------------------------------------------------------

__host__ __device__ __forceinline__ double bc_diag_factor(int bc) {
    return bc == BC_DF ? 2.0 : (bc == BC_DC ? 1.0 : 0.0);
}
__host__ __device__ __forceinline__ double bc_ghost_factor(int bc) {
    return bc == BC_DF ? -1.0 : (bc == BC_DC ? 0.0 : 1.0);
}

// The cell-centred tile formulas below fix every rounding in the source (no
// contraction left to the compiler): the compiler's choice of which product
// to fuse into an add can depend on the context a formula is inlined into,
// and the level kernels, their batched forms and the coarse kernel must give
// the same bits for the same values.
template <class T>
__host__ __device__ __forceinline__ T cc_diag(const CCOp &op, int m, int n, T cx, T cy, int i, int j) {
#pragma clang fp contract(off)
    T d = 0;
    d += T(i > 0 ? 1.0 : bc_diag_factor(op.bcl)) * cx;
    d += T(i < m - 1 ? 1.0 : bc_diag_factor(op.bcr)) * cx;
    d += T(j > 0 ? 1.0 : bc_diag_factor(op.bcb)) * cy;
    d += T(j < n - 1 ? 1.0 : bc_diag_factor(op.bct)) * cy;
    return T(op.sigma) - T(op.s) * d;
}

template <class T>
__host__ __device__ __forceinline__ T cc_off(int m, int n, T cx, T cy, const T *z, int i, int j) {
    const T *w = z + i + (size_t)m * j;
    T sx = 0, sy = 0;
    if (i > 0) sx += w[-1];
    if (i < m - 1) sx += w[1];
    if (j > 0) sy += w[-m];
    if (j < n - 1) sy += w[m];
    return cx * sx + cy * sy;
}

template <class T>
__host__ __device__ __forceinline__ T fem_acc(const FemOp &op, int M, int N, int i, int j) {
    return T((i > 0 && i < M - 1 ? 1.0 : 0.5) * (j > 0 && j < N - 1 ? 1.0 : 0.5) * op.fm);
}

template <class T>
__host__ __device__ __forceinline__ T fem_off(const FemOp &op, int M, int N, const T *z, int i, int j) {
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
__host__ __device__ __forceinline__ T Az_plain(bool fem, const CCOp &cop, const FemOp &fop, int m, int n, T cx, T cy,
                                     const T *z, int i, int j) {
    if (fem) return fem_acc<T>(fop, m, n, i, j) * z[i + (size_t)m * j] + fem_off<T>(fop, m, n, z, i, j);
    return T(cop.s) * cc_off<T>(m, n, cx, cy, z, i, j) + cc_diag<T>(cop, m, n, cx, cy, i, j) * z[i + (size_t)m * j];
}

// ------------------------------------------------------------------------
// Operators on an LDS tile with a zero-filled outside (row pitch RX)
// ------------------------------------------------------------------------

template <class T>
__device__ __forceinline__ T cc_off_t(T cx, T cy, const T *s, int l, int RX) {
#pragma clang fp contract(off)
    return fma(cx, s[l - 1] + s[l + 1], cy * (s[l - RX] + s[l + RX]));
}

// 9-point Q1 stencil with missing neighbours zero: edge columns halve the
// vertical coupling (hey), the top row halves the horizontal one (hex).
// The rounding is fixed in the source (fem_off_v), like the cell-centred
// formulas', so that the tiled and the streaming kernels agree bit for bit:
// no fused multiply-adds, every product and sum rounded in the order of the
// oracle's fem_off (oracle/vdo.c, built without contraction), which it then
// matches bit for bit for the same iterate.
template <class T>
__device__ __forceinline__ T fem_off_v(const FemOp &op, T wi, T wj, T dl, T d, T dr, T l, T r, T ul, T u, T ur) {
#pragma clang fp contract(off)
    const T diag = ((dl + dr) + ul) + ur;
    return (T(op.fc) * diag + (wi * T(op.fey)) * (d + u)) + (wj * T(op.fex)) * (l + r);
}
template <class T>
__device__ __forceinline__ T fem_off_t(const FemOp &op, int M, int N, const T *s, int l, int RX, int i, int j) {
    const T wi = (i == 0 || i == M - 1) ? T(0.5) : T(1);
    const T wj = (j == N - 1) ? T(0.5) : T(1);
    return fem_off_v<T>(op, wi, wj, s[l - RX - 1], s[l - RX], s[l - RX + 1], s[l - 1], s[l + 1], s[l + RX - 1],
                        s[l + RX], s[l + RX + 1]);
}

// Boundary class of a cell-centred point: xc + 3 yc with xc, yc = 0 inside,
// 1 on the first and 2 on the last column / row.
template <class T>
__device__ __forceinline__ int cc_class(const LevelT<T> &L, int i, int j) {
    return (i == 0 ? 1 : (i == L.m - 1 ? 2 : 0)) + 3 * (j == 0 ? 1 : (j == L.n - 1 ? 2 : 0));
}

// Reciprocal of the cell-centred diagonal per boundary class, written to
// tab[0..8] (LDS) by threads 0..8 (one division each, once per launch);
// tab[0] is the level's stored interior reciprocal.  Sweeps then multiply
// every point by tab[class] instead of dividing at boundary points: an FP64
// division behind a branch taken by every wave with a boundary lane was an
// eighth of a cell-centred V-cycle at 1024 x 512.
template <class T>
__device__ __forceinline__ void cc_inv_table(const LevelT<T> &L, const CCOp &op, T *tab, int tid) {
    if (tid < 9) {
        const int xc = tid % 3, yc = tid / 3;
        const int i = xc == 1 ? 0 : (xc == 2 ? L.m - 1 : min(1, L.m - 1));
        const int j = yc == 1 ? 0 : (yc == 2 ? L.n - 1 : min(1, L.n - 1));
        tab[tid] = tid == 0 ? L.inv_int : T(1) / cc_diag<T>(op, L.m, L.n, L.cx, L.cy, i, j);
    }
}
constexpr int kInvTab = 16;  // LDS entries reserved for the table

template <class T>
__device__ __forceinline__ T fem_inv_acc(const LevelT<T> &L, int i, int j) {
    return T(((i == 0 || i == L.m - 1) ? 2.0 : 1.0) * ((j == 0 || j == L.n - 1) ? 2.0 : 1.0)) * L.inv_int;
}

// A tile of TX x TY points at (x0, y0) with a halo of H columns and HY rows
// on each side (the nodal sweeps need fewer rows than columns).
struct Tile {
    int x0, y0, gx0, gy0, RX, RY, H, HY, TX, TY;
};

__device__ __forceinline__ Tile make_tile(int bx, int by, int H, int TX, int TY, int HY) {
    Tile t;
    t.H = H;
    t.HY = HY;
    t.TX = TX;
    t.TY = TY;
    t.x0 = bx * TX;
    t.y0 = by * TY;
    t.gx0 = t.x0 - H;
    t.gy0 = t.y0 - HY;
    t.RX = TX + 2 * H;
    t.RY = TY + 2 * HY;
    return t;
}

// Halo of a level launch with nu sweeps: columns, and rows (HY).  Cell-centred
// red-black passes shrink the region whose values equal a global sweep's by
// one point per pass both ways, plus one for the residual (norm or
// restriction).  The nodal four-colour passes shrink it by one column per
// pass, but by one row per pair of passes (a pair updates rows of one
// parity): ceil(p / 2) rows after pass p, with the tile's first row even.
// Nodal down launches need residuals one row and column beyond the tile (full
// weighting), up launches with the norm the tile's own residuals; the row
// halo is kept even so that the tile's first row stays even.
__host__ __device__ __forceinline__ int halo_cols(bool fem, int nu, int extra) { return (fem ? 4 : 2) * nu + extra; }
__host__ __device__ __forceinline__ int halo_rows(bool fem, int nu, int extra) {
    return fem ? 2 * nu + 2 * extra : 2 * nu + extra;
}

// Bilinear weights of a cell-centred prolongation: the coarse value q00 of
// the fine point's own coarse cell, q10 / q01 / q11 of its neighbours across
// x / y / both (read at boundary-clamped coarse points), fx and fy the ghost
// factors of a neighbour outside the coarse level (1 inside).
template <class T>
__device__ __forceinline__ T cc_prolong_combine(T fx, T fy, T q00, T q10, T q01, T q11) {
#pragma clang fp contract(off)
    const T e00 = q00, e10 = fx * q10, e01 = fy * q01, e11 = fx * fy * q11;
    return fma(T(0.0625), e11, fma(T(0.5625), e00, T(0.1875) * (e10 + e01)));
}

// Bilinear cell-centred prolongation of the coarse correction at fine (i,j).
template <class T>
__device__ __forceinline__ T cc_prolong_at(const CCOp &op, int Mc, int Nc, const T *e, int i, int j,
                                           size_t pitch = 0) {
#pragma clang fp contract(off)
    if (!pitch) pitch = Mc;
    int I = i >> 1, J = j >> 1;
    int I2 = (i & 1) ? I + 1 : I - 1, J2 = (j & 1) ? J + 1 : J - 1;
    T fx = 1, fy = 1;
    if (I2 < 0) { fx = T(bc_ghost_factor(op.bcl)); I2 = I; }
    else if (I2 >= Mc) { fx = T(bc_ghost_factor(op.bcr)); I2 = I; }
    if (J2 < 0) { fy = T(bc_ghost_factor(op.bcb)); J2 = J; }
    else if (J2 >= Nc) { fy = T(bc_ghost_factor(op.bct)); J2 = J; }
    return cc_prolong_combine<T>(fx, fy, e[I + pitch * J], e[I2 + pitch * J], e[I + pitch * J2], e[I2 + pitch * J2]);
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

// Weights of the nodal bilinear prolongation at a fine node of column /
// row parity io / jo from the coarse values q00 (own), q10 (next column),
// q01 (next row), q11.  Power-of-two weights by parity: (1,0,0,0),
// (1/2,1/2,0,0), (1/2,0,1/2,0) or (1/4,1/4,1/4,1/4).  Scaling by a power of
// two is exact, so the weighted sum is bit for bit the parity formula (up to
// the sign of a zero) whatever the compiler fuses.
template <class T>
__device__ __forceinline__ T fem_prolong_combine(int io, int jo, T q00, T q10, T q01, T q11) {
    const T wx = io ? T(0.5) : T(1), wy = jo ? T(0.5) : T(1);
    const T w10 = io ? wx * wy : T(0), w01 = jo ? wx * wy : T(0), w11 = (io && jo) ? T(0.25) : T(0);
    return wx * wy * q00 + w10 * q10 + w01 * q01 + w11 * q11;
}

// Nodal bilinear prolongation at fine node (i,j) with every load made
// unconditionally (the neighbours a parity does not use are clamped into the
// level and ignored), so that a round of tile loads issues all its loads
// before the first one is waited for.  Same value as fem_prolong_at.
template <class T>
__device__ __forceinline__ T fem_prolong_all(const T *ez, int Mc, int Nc, int i, int j, size_t pitch = 0) {
    if (!pitch) pitch = Mc;
    const int I = i >> 1, J = j >> 1, I1 = min(I + 1, Mc - 1), J1 = min(J + 1, Nc - 1);
    const T q00 = ez[I + pitch * J], q10 = ez[I1 + pitch * J];
    const T q01 = ez[I + pitch * J1], q11 = ez[I1 + pitch * J1];
    return fem_prolong_combine<T>(i & 1, j & 1, q00, q10, q01, q11);
}

// q = a / b, r = a % b for 0 <= a < 2^22, 1 <= b < 2^16, from the FP32
// reciprocal: a * rcp(b) is within one of a / b, and one correction each way
// makes it exact (an integer division is a long dependent chain).
__device__ __forceinline__ void small_divmod(int a, int b, int &q, int &r) {
    q = (int)((float)a * __builtin_amdgcn_rcpf((float)b));
    r = a - q * b;
    if (r < 0) {
        r += b;
        q--;
    }
    if (r >= b) {
        r -= b;
        q++;
    }
}

// (P, r) = (q / W, q % W) for q = q0, q0 + step, ..., advanced without a
// division per point.  1 <= W < 2^16, q < 2^22.
struct Walk {
    int q, P, r, W, dP, dr;
    __device__ __forceinline__ Walk(int q0, int w, int step) {
        W = w;
        q = q0;
        small_divmod(q0, w, P, r);
        small_divmod(step, w, dP, dr);
    }
    __device__ __forceinline__ void next(int step) {
        q += step;
        P += dP;
        r += dr;
        if (r >= W) {
            r -= W;
            P++;
        }
    }
};

// Loads zsrc (+ prolongated coarse correction when C is given; zero when
// zsrc is null) and b over the tile region; outside the domain both are zero.
// The region is walked as one flat index range, kLoadUnroll elements per
// thread and round.  Every load is unconditional -- points outside the domain
// read the nearest domain point and the value is dropped -- so the compiler
// issues all loads of a round before waiting for any of them, and a round
// costs one memory round trip instead of one per element.
constexpr int kLoadUnroll = 4;

template <class T, bool FEM>
__device__ __forceinline__ void tile_load(const LevelT<T> &F, const T *zsrc, const LevelT<T> *C, const CCOp &op,
                                          T *sz, T *sb, const Tile &t) {
    const int nel = t.RX * t.RY;
    Walk w(threadIdx.x + kBX * threadIdx.y, t.RX, kLvThreads);
    while (w.q < nel) {
        T zv[kLoadUnroll], bv[kLoadUnroll];
        int li[kLoadUnroll];
#pragma unroll
        for (int u = 0; u < kLoadUnroll; u++) {
            const int gx = t.gx0 + w.r, gy = t.gy0 + w.P;
            const bool in = gy >= 0 && gy < F.n && gx >= 0 && gx < F.m;
            const int cx = min(max(gx, 0), F.m - 1), cy = min(max(gy, 0), F.n - 1);
            const size_t g = cx + (size_t)F.m * cy;
            T z = zsrc ? zsrc[g] : T(0);
            const T b = F.b[g];
            if (C) {
                if (FEM) {
                    const T e = fem_prolong_all<T>(C->z, C->m, C->n, cx, cy);
                    if (cy > 0) z += e;
                } else {
                    z += cc_prolong_at<T>(op, C->m, C->n, C->z, cx, cy);
                }
            }
            zv[u] = in ? z : T(0);
            bv[u] = in ? b : T(0);
            li[u] = w.q < nel ? w.q : -1;
            w.next(kLvThreads);
        }
#pragma unroll
        for (int u = 0; u < kLoadUnroll; u++)
            if (li[u] >= 0) {
                sz[li[u]] = zv[u];
                sb[li[u]] = bv[u];
            }
    }
}

template <class T>
__device__ __forceinline__ void tile_store(const LevelT<T> &F, T *dst, const T *sz, const Tile &t) {
    for (int y = threadIdx.y; y < t.TY; y += kBY) {
        const int gy = t.y0 + y;
        if (gy >= F.n) break;
        for (int x = threadIdx.x; x < t.TX; x += kBX) {
            const int gx = t.x0 + x;
            if (gx >= F.m) break;
            dst[gx + (size_t)F.m * gy] = sz[(y + t.HY) * t.RX + x + t.H];
        }
    }
}

// Red-black half sweep of colour c on the region shrunk by s cells.
template <class T, int BX = kBX, int BY = kBY>
__device__ __forceinline__ void cc_tile_sweep(const LevelT<T> &L, const CCOp &op, T *sz, const T *sb, const Tile &t,
                                              int s, int c, const T *itab) {
#pragma clang fp contract(off)
    const int X0 = max(0, t.gx0 + s), X1 = min(L.m, t.x0 + t.TX + t.H - s);
    const int Y0 = max(0, t.gy0 + s), Y1 = min(L.n, t.y0 + t.TY + t.HY - s);
    const T sgn = T(op.s);
    for (int gy = Y0 + threadIdx.y; gy < Y1; gy += BY) {
        const int xs = X0 + (((X0 + gy) & 1) ^ c);
        const int row = (gy - t.gy0) * t.RX - t.gx0;
        for (int gx = xs + 2 * threadIdx.x; gx < X1; gx += 2 * BX) {
            const int l = row + gx;
            const T rhs = fma(-sgn, cc_off_t<T>(L.cx, L.cy, sz, l, t.RX), sb[l]);
            sz[l] = rhs * itab[cc_class<T>(L, gx, gy)];
        }
    }
}

// Four-colour sweep of colour (ci,cj) on the region shrunk by s; row 0 fixed.
// The colour's nx x ny points are dealt to the threads as one flat range, so
// no thread takes more than ceil(nx ny / threads) of them (a 2-D thread grid
// gave some threads four points of a 33 x 17 colour while most had one; the
// nodal V-cycle at 1025 x 513 went from 125.5 to 119 us).
template <class T, int BX = kBX, int BY = kBY>
__device__ __forceinline__ void fem_tile_sweep(const LevelT<T> &L, const FemOp &op, T *sz, const T *sb, const Tile &t,
                                               int s, int ci, int cj) {
    const int sy = (s + 1) >> 1;  // rows lost after pass s (halo_rows)
    const int X0 = max(0, t.gx0 + s), X1 = min(L.m, t.x0 + t.TX + t.H - s);
    const int Y0 = max(1, t.gy0 + sy), Y1 = min(L.n, t.y0 + t.TY + t.HY - sy);
    const int ys = Y0 + ((Y0 ^ cj) & 1), xs = X0 + ((X0 ^ ci) & 1);
    const int nx = (X1 - xs + 1) >> 1, ny = (Y1 - ys + 1) >> 1;
    if (nx <= 0 || ny <= 0) return;
    const int n = nx * ny;
    for (int k = threadIdx.x + BX * threadIdx.y; k < n; k += BX * BY) {
        int r, q;
        small_divmod(k, nx, r, q);
        const int gy = ys + 2 * r, gx = xs + 2 * q;
        const int l = (gy - t.gy0) * t.RX - t.gx0 + gx;
        sz[l] = (sb[l] - fem_off_t<T>(op, L.m, L.n, sz, l, t.RX, gx, gy)) * fem_inv_acc<T>(L, gx, gy);
    }
}

template <class T>
__device__ __forceinline__ T cc_resid_t(const LevelT<T> &L, const CCOp &op, const T *sz, const T *sb, int RX, int l,
                                        int gx, int gy) {
#pragma clang fp contract(off)
    return sb[l] - fma(T(op.s), cc_off_t<T>(L.cx, L.cy, sz, l, RX), cc_diag<T>(op, L.m, L.n, L.cx, L.cy, gx, gy) * sz[l]);
}

template <class T>
__device__ __forceinline__ T fem_resid_t(const FemOp &op, int M, int N, const T *sz, const T *sb, int RX, int l,
                                         int gx, int gy) {
#pragma clang fp contract(off)
    const T acc = fem_acc<T>(op, M, N, gx, gy);
    if (gy == 0) return sb[l] - acc * sz[l];
    return sb[l] - (acc * sz[l] + fem_off_t<T>(op, M, N, sz, l, RX, gx, gy));
}

// fem_resid_t off row 0, without the branch (the restriction never weights
// row 0: coarse row 0 is Dirichlet and rows 2J-1.. of J >= 1 start at 1).
template <class T>
__device__ __forceinline__ T fem_resid_nd(const FemOp &op, int M, int N, const T *sz, const T *sb, int RX, int l,
                                          int gx, int gy) {
#pragma clang fp contract(off)
    const T acc = fem_acc<T>(op, M, N, gx, gy);
    return sb[l] - (acc * sz[l] + fem_off_t<T>(op, M, N, sz, l, RX, gx, gy));
}

// Restricted residual at coarse point (I, J) from a tile holding the fine
// level: average of 4 (cell-centred) or full weighting (nodal; coarse row 0
// is Dirichlet and gets zero).  A fine node outside the domain gets weight
// zero instead of a branch (it lies in the tile's halo, so its residual is
// readable), which adds an exact zero.
template <class T, bool FEM>
__device__ __forceinline__ T restrict_at(const LevelT<T> &F, const CCOp &cop, const FemOp &fop, const T *sz,
                                         const T *sb, const Tile &t, int I, int J) {
    T v = 0;
    if (!FEM) {
#pragma clang fp contract(off)
        for (int b2 = 0; b2 < 2; b2++)
            for (int a2 = 0; a2 < 2; a2++) {
                const int gx = 2 * I + a2, gy = 2 * J + b2;
                const int l = (gy - t.gy0) * t.RX + gx - t.gx0;
                v += cc_resid_t<T>(F, cop, sz, sb, t.RX, l, gx, gy);
            }
        v *= T(0.25);
    } else {
        for (int b2 = -1; b2 <= 1; b2++) {
            const int gy = 2 * J + b2;
            T sx = 0;
#pragma unroll
            for (int a2 = -1; a2 <= 1; a2++) {
                const int gx = 2 * I + a2;
                const int l = (gy - t.gy0) * t.RX + gx - t.gx0;
                const T wa = (gx < 0 || gx >= F.m) ? T(0) : (a2 ? T(0.5) : T(1));
                sx += wa * fem_resid_nd<T>(fop, F.m, F.n, sz, sb, t.RX, l, gx, gy);
            }
            v += (gy >= F.n ? T(0) : (b2 ? T(0.5) : T(1))) * sx;
        }
        if (J == 0) v = 0;
    }
    return v;
}

// sh holds one partial per wave: 8 entries in every kernel that calls this.
static_assert(kLvThreads <= 8 * 64 && kCoarseThreads <= 8 * 64, "block_sum takes at most eight waves");

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
    __syncthreads();
    return a;
}

// Convergence bookkeeping, first half: each workgroup stores its partial of
// ||b - A z||^2 (plain store; the one-workgroup k_norm_finish launched next
// sums them in index order and raises the flag).  Kernel boundaries order
// the stores, so no device-scope fence is needed here.
__device__ __forceinline__ void store_partial(const MGCtl &ctl, double part, int bid, double *sh) {
    part = block_sum(part, sh);
    if (threadIdx.x + blockDim.x * threadIdx.y == 0) ctl.partial[bid] = part;
}

// Single-workgroup variant (the coarse kernel when it owns level 0): the
// workgroup's own sum is final.
__device__ __forceinline__ void finish_norm_single(const MGCtl &ctl, double part, double tol2, double *sh) {
    const double rr = block_sum(part, sh);
    if (threadIdx.x + blockDim.x * threadIdx.y == 0) {
        MGState *st = ctl.state;
        st->rr = rr;
        st->cycles += 1;
        st->done = !(rr == rr) ? 2 : (rr <= tol2 * st->bb ? 1 : 0);
    }
    __syncthreads();
}

// ------------------------------------------------------------------------
// Level kernel bodies (T = double for FP64 hierarchies, float for the inner
// cycles of the mixed form).  sh/last: workgroup scratch of the reduction.
// ------------------------------------------------------------------------

// The previous V-cycle's norm, finished by every workgroup of the next
// level-0 down launch (LvArgs::fin): threads 0..255 take the partials in
// k_norm_finish's order (thread t: t, t + 256, ...) and the workgroup sum is
// k_norm_finish's, so every workgroup gets the bits that kernel computes;
// workgroup 0 makes its state update.  pf holds the loads, issued before the
// tile's own so that the two round trips overlap.  Returns false when the
// solve stops here (all workgroups agree; one that already sees workgroup
// 0's new flag stops too, which is what it would have computed).
constexpr int kFinRegs = kFinMax / 256;

// With pairs (k_begin's partials: ||b - A z||^2, ||b||^2 per workgroup) the
// same holds for k_norm_finish's begin branch: pf holds 2 kFinRegs values.
__device__ __forceinline__ void fin_load(const MGCtl &ctl, int np, int pairs, double *pf) {
    const int tid = threadIdx.x + kBX * threadIdx.y;
#pragma unroll
    for (int u = 0; u < kFinRegs; u++) {
        const int k = tid + 256 * u;
        const bool in = tid < 256 && k < np;
        pf[u] = in ? ctl.partial[pairs ? 2 * k : k] : 0.0;
        pf[kFinRegs + u] = in && pairs ? ctl.partial[2 * k + 1] : 0.0;
    }
}

__device__ __forceinline__ int fin_norm(const MGCtl &ctl, int np, int pairs, double tol2, const double *pf,
                                        double *sh) {
    const int tid = threadIdx.x + kBX * threadIdx.y;
    MGState *st = ctl.state;
    const double bb_state = pairs ? 0.0 : st->bb;
    double r = 0, b = 0;
#pragma unroll
    for (int u = 0; u < kFinRegs; u++)
        if (tid < 256 && tid + 256 * u < np) {
            r += pf[u];
            b += pf[kFinRegs + u];
        }
    r = block_sum(r, sh);
    if (pairs) b = block_sum(b, sh);
    if (tid == 0) {
        sh[8] = r;
        sh[9] = b;
    }
    __syncthreads();
    const double R = sh[8], B = pairs ? sh[9] : bb_state;
    const int nd = pairs ? ((B == 0) ? 3 : (!(R == R) ? 2 : (R <= tol2 * B ? 1 : 0)))
                         : (!(R == R) ? 2 : (R <= tol2 * B ? 1 : 0));
    if (tid == 0 && blockIdx.x == 0 && !st->done) {
        st->rr = R;
        if (pairs) {
            st->bb = B;
            st->cycles = 0;
        } else {
            st->cycles += 1;
        }
        st->done = nd;
    }
    return nd;
}

template <class T, bool FEM>
__device__ __forceinline__ void down_body(const LvArgs<T> &a, int tile, T *lds, double *sh) {
    const LevelT<T> &F = a.F, &C = a.C;
    const Tile t = make_tile(tile % a.nbx, tile / a.nbx, halo_cols(FEM, a.nu, FEM ? 2 : 1), a.tx, a.ty,
                             halo_rows(FEM, a.nu, 1));
    T *sz = lds, *sb = lds + t.RX * t.RY, *itab = sb + t.RX * t.RY;
    double pf[2 * kFinRegs];
    if (a.fin) fin_load(a.ctl, a.fin, a.fin_begin, pf);
    tile_load<T, FEM>(F, a.flag ? nullptr : F.z, nullptr, a.cop, sz, sb, t);
    if (!FEM) cc_inv_table<T>(F, a.cop, itab, threadIdx.x + kBX * threadIdx.y);
    if (a.fin) {
        const int nd = fin_norm(a.ctl, a.fin, a.fin_begin, a.tol2, pf, sh);
        if (nd == 3 && a.fin_begin && sizeof(T) == sizeof(double)) {
            // zero right-hand side: the solution is z = 0 (each workgroup its
            // tile), which an FP64 solve's end launch would otherwise write
            for (int y = threadIdx.y; y < t.TY && t.y0 + y < F.n; y += kBY)
                for (int x = threadIdx.x; x < t.TX && t.x0 + x < F.m; x += kBX)
                    F.z[t.x0 + x + (size_t)F.m * (t.y0 + y)] = T(0);
        }
        if (nd) return;
    }
    __syncthreads();
    if (FEM) {
        for (int s = 1; s <= 4 * a.nu; s++) {
            const int c = (s - 1) & 3;
            fem_tile_sweep<T>(F, a.fop, sz, sb, t, s, c & 1, c >> 1);
            __syncthreads();
        }
    } else {
        for (int s = 1; s <= 2 * a.nu; s++) {
            cc_tile_sweep<T>(F, a.cop, sz, sb, t, s, (s - 1) & 1, itab);
            __syncthreads();
        }
    }
    tile_store<T>(F, F.zalt, sz, t);
    const int I0 = t.x0 >> 1, J0 = t.y0 >> 1;
    const int I1 = min(C.m, I0 + t.TX / 2), J1 = min(C.n, J0 + t.TY / 2);
    for (int J = J0 + threadIdx.y; J < J1; J += kBY)
        for (int I = I0 + threadIdx.x; I < I1; I += kBX)
            C.b[I + (size_t)C.m * J] = restrict_at<T, FEM>(F, a.cop, a.fop, sz, sb, t, I, J);
}

template <class T, bool FEM>
__device__ __forceinline__ void up_body(const LvArgs<T> &a, int tile, T *lds, double *sh, int *last) {
    const LevelT<T> &F = a.F;
    const Tile t = make_tile(tile % a.nbx, tile / a.nbx, halo_cols(FEM, a.nu, a.flag), a.tx, a.ty,
                             halo_rows(FEM, a.nu, a.flag));
    T *sz = lds, *sb = lds + t.RX * t.RY, *itab = sb + t.RX * t.RY;
    tile_load<T, FEM>(F, F.zalt, &a.C, a.cop, sz, sb, t);
    if (!FEM) cc_inv_table<T>(F, a.cop, itab, threadIdx.x + kBX * threadIdx.y);
    __syncthreads();
    if (FEM) {
        for (int s = 1; s <= 4 * a.nu; s++) {
            const int c = (s - 1) & 3;
            fem_tile_sweep<T>(F, a.fop, sz, sb, t, s, c & 1, c >> 1);
            __syncthreads();
        }
    } else {
        for (int s = 1; s <= 2 * a.nu; s++) {
            cc_tile_sweep<T>(F, a.cop, sz, sb, t, s, (s - 1) & 1, itab);
            __syncthreads();
        }
    }
    tile_store<T>(F, F.z, sz, t);
    if (!a.flag) return;
    double rr = 0;
    for (int y = threadIdx.y; y < t.TY; y += kBY) {
        const int gy = t.y0 + y;
        if (gy >= F.n) break;
        for (int x = threadIdx.x; x < t.TX; x += kBX) {
            const int gx = t.x0 + x;
            if (gx >= F.m) break;
            const int l = (y + t.HY) * t.RX + x + t.H;
            const double r = FEM ? fem_resid_t<T>(a.fop, F.m, F.n, sz, sb, t.RX, l, gx, gy)
                                 : cc_resid_t<T>(F, a.cop, sz, sb, t.RX, l, gx, gy);
            rr += r * r;
        }
    }
    store_partial(a.ctl, rr, tile, sh);
    (void)last;
}

// Mixed form, level 0, end of an inner FP32 cycle: prolongate and smooth the
// FP32 correction e (pre-smoothed iterate in E.zalt, right-hand side r32 in
// E.b), then z_out = z_in + e and the new FP64 residual r = b - A z_out,
// which is reduced for the convergence test and stored in r_out as the next
// cycle's right-hand side.  Inputs and outputs are distinct arrays.
// A z of the mixed form's FP64 residual at (gx, gy), z on a tile of row pitch
// RX (the rounding fixed in the source, shared with the streaming form).
template <bool FEM>
__device__ __forceinline__ double mx_az(const MxArgs &a, const double *zd, int l, int RX, int gx, int gy) {
#pragma clang fp contract(off)
    const Level &Z = a.Z;
    if (FEM) {
        const double acc = fem_acc<double>(a.fop, Z.m, Z.n, gx, gy);
        return gy == 0 ? acc * zd[l] : acc * zd[l] + fem_off_t<double>(a.fop, Z.m, Z.n, zd, l, RX, gx, gy);
    }
    return fma(a.cop.s, cc_off_t<double>(Z.cx, Z.cy, zd, l, RX), cc_diag<double>(a.cop, Z.m, Z.n, Z.cx, Z.cy, gx, gy) * zd[l]);
}

template <bool FEM>
__device__ __forceinline__ void mx_body(const MxArgs &a, int tile, double *lds_d, double *sh, int *last) {
    const Level &Z = a.Z;
    const Tile t = make_tile(tile % a.nbx, tile / a.nbx, halo_cols(FEM, a.nu, 1), a.tx, a.ty, halo_rows(FEM, a.nu, 1));
    const int R = t.RX * t.RY;
    double *zd = lds_d;
    float *se = reinterpret_cast<float *>(lds_d + R), *sr = se + R, *itab = sr + R;
    tile_load<float, FEM>(a.E, a.E.zalt, &a.C, a.cop, se, sr, t);
    if (!FEM) cc_inv_table<float>(a.E, a.cop, itab, threadIdx.x + kBX * threadIdx.y);
    __syncthreads();
    for (int s = 1; s <= (FEM ? 4 : 2) * a.nu; s++) {
        if (FEM) {
            const int c = (s - 1) & 3;
            fem_tile_sweep<float>(a.E, a.fop, se, sr, t, s, c & 1, c >> 1);
        } else {
            cc_tile_sweep<float>(a.E, a.cop, se, sr, t, s, (s - 1) & 1, itab);
        }
        __syncthreads();
    }
    // z + e on the tile and a one-cell ring (zero outside the domain)
    const int tid = threadIdx.x + kBX * threadIdx.y;
    {
        const int RW = t.TX + 2, nel = RW * (t.TY + 2);
        for (int k0 = tid; k0 < nel; k0 += kLvThreads * kLoadUnroll) {
            double zz[kLoadUnroll];
#pragma unroll
            for (int u = 0; u < kLoadUnroll; u++) {
                const int k = k0 + kLvThreads * u;
                zz[u] = 0;
                if (k < nel) {
                    const int ry = k / RW, rx = k - ry * RW;
                    const int gx = t.x0 - 1 + rx, gy = t.y0 - 1 + ry;
                    if (gy >= 0 && gy < Z.n && gx >= 0 && gx < Z.m) zz[u] = a.z_in[gx + (size_t)Z.m * gy];
                }
            }
#pragma unroll
            for (int u = 0; u < kLoadUnroll; u++) {
                const int k = k0 + kLvThreads * u;
                if (k < nel) {
                    const int ry = k / RW, rx = k - ry * RW;
                    const int gx = t.x0 - 1 + rx, gy = t.y0 - 1 + ry;
                    const bool in = gy >= 0 && gy < Z.n && gx >= 0 && gx < Z.m;
                    const int l = (ry + t.HY - 1) * t.RX + rx + t.H - 1;
                    zd[l] = in ? zz[u] + (double)se[l] : 0.0;
                }
            }
        }
    }
    __syncthreads();
    double rr = 0;
    const int nel = t.TX * t.TY;
    for (int k0 = tid; k0 < nel; k0 += kLvThreads * kLoadUnroll) {
        double bv[kLoadUnroll];
#pragma unroll
        for (int u = 0; u < kLoadUnroll; u++) {
            const int k = k0 + kLvThreads * u;
            bv[u] = 0;
            if (k < nel) {
                const int y = k / t.TX, x = k - y * t.TX;
                const int gx = t.x0 + x, gy = t.y0 + y;
                if (gx < Z.m && gy < Z.n) bv[u] = Z.b[gx + (size_t)Z.m * gy];
            }
        }
#pragma unroll
        for (int u = 0; u < kLoadUnroll; u++) {
            const int k = k0 + kLvThreads * u;
            if (k >= nel) continue;
            const int y = k / t.TX, x = k - y * t.TX;
            const int gx = t.x0 + x, gy = t.y0 + y;
            if (gx >= Z.m || gy >= Z.n) continue;
            const int l = (y + t.HY) * t.RX + x + t.H;
            const size_t g = gx + (size_t)Z.m * gy;
            const double r = bv[u] - mx_az<FEM>(a, zd, l, t.RX, gx, gy);
            a.z_out[g] = zd[l];
            a.r_out[g] = (float)r;
            rr += r * r;
        }
    }
    store_partial(a.ctl, rr, tile, sh);
    (void)last;
}

// ------------------------------------------------------------------------
// Level kernels: direct (one record, grid = tiles) and batched (records in
// device memory, a persistent grid walking the work list)
// ------------------------------------------------------------------------

template <class T, bool FEM>
__global__ void __launch_bounds__(kLvThreads) k_down(LvArgs<T> a) {
    extern __shared__ double lds_d[];
    __shared__ double sh[10];
    // A launch that finishes the start-of-solve norms runs fin_norm in every
    // workgroup, also in those dispatched after workgroup 0 has raised the
    // flag: each must zero its own tile when the right-hand side is zero.
    if (a.ctl.state->done && !a.fin_begin) return;
    stamp_begin(a.ts);
    down_body<T, FEM>(a, xcd_tile(blockIdx.x, a.nb), reinterpret_cast<T *>(lds_d), sh);
    stamp_end(a.ts);
}

template <class T, bool FEM>
__global__ void __launch_bounds__(kLvThreads) k_down_b(const LvArgs<T> *A, const WorkItem *W, int nw, BatchAux aux) {
    extern __shared__ double lds_d[];
    __shared__ double sh[10];  // batched records never fold a norm (fin = 0)
    stamp_begin(aux.ts);
    for (int p = blockIdx.x; p < nw; p += gridDim.x) {
        const WorkItem w = W[xcd_tile(p, nw)];
        const LvArgs<T> a = A[w.rec];
        if (a.ctl.state->done) continue;
        const unsigned long long t0 = busy_begin(aux.busy);
        down_body<T, FEM>(a, w.tile, reinterpret_cast<T *>(lds_d), sh);
        __syncthreads();
        busy_end(aux.busy, w.sim, t0);
    }
    stamp_end(aux.ts);
}

template <class T, bool FEM>
__global__ void __launch_bounds__(kLvThreads) k_up(LvArgs<T> a) {
    extern __shared__ double lds_d[];
    __shared__ double sh[8];
    __shared__ int last;
    if (a.ctl.state->done) return;
    stamp_begin(a.ts);
    up_body<T, FEM>(a, xcd_tile(blockIdx.x, a.nb), reinterpret_cast<T *>(lds_d), sh, &last);
    stamp_end(a.ts);
}

template <class T, bool FEM>
__global__ void __launch_bounds__(kLvThreads) k_up_b(const LvArgs<T> *A, const WorkItem *W, int nw, BatchAux aux) {
    extern __shared__ double lds_d[];
    __shared__ double sh[8];
    __shared__ int last;
    stamp_begin(aux.ts);
    for (int p = blockIdx.x; p < nw; p += gridDim.x) {
        const WorkItem w = W[xcd_tile(p, nw)];
        const LvArgs<T> a = A[w.rec];
        if (a.ctl.state->done) continue;
        const unsigned long long t0 = busy_begin(aux.busy);
        up_body<T, FEM>(a, w.tile, reinterpret_cast<T *>(lds_d), sh, &last);
        __syncthreads();
        busy_end(aux.busy, w.sim, t0);
    }
    stamp_end(aux.ts);
}

template <bool FEM>
__global__ void __launch_bounds__(kLvThreads) k_mx_up0(MxArgs a) {
    extern __shared__ double lds_d[];
    __shared__ double sh[8];
    __shared__ int last;
    if (a.ctl.state->done) return;
    stamp_begin(a.ts);
    mx_body<FEM>(a, xcd_tile(blockIdx.x, a.nb), lds_d, sh, &last);
    stamp_end(a.ts);
}

template <bool FEM>
__global__ void __launch_bounds__(kLvThreads) k_mx_up0_b(const MxArgs *A, const WorkItem *W, int nw, BatchAux aux) {
    extern __shared__ double lds_d[];
    __shared__ double sh[8];
    __shared__ int last;
    stamp_begin(aux.ts);
    for (int p = blockIdx.x; p < nw; p += gridDim.x) {
        const WorkItem w = W[xcd_tile(p, nw)];
        const MxArgs a = A[w.rec];
        if (a.ctl.state->done) continue;
        const unsigned long long t0 = busy_begin(aux.busy);
        mx_body<FEM>(a, w.tile, lds_d, sh, &last);
        __syncthreads();
        busy_end(aux.busy, w.sim, t0);
    }
    stamp_end(aux.ts);
}

}  // namespace vdi

#include "mg_stream.hpp"

namespace vdi {

// ------------------------------------------------------------------------
// The coarsest level on one workgroup.  The level sits in LDS as an
// (m + 2) x (n + 2) tile with a zero border, for z and for b, so that its
// residual (and, where the level is not separable, its sweeps) is the tiled
// kernels' own code on a tile covering the whole level, with no boundary
// test in front of an LDS read.  Threads are laid out 32 x 16.
__device__ __forceinline__ Tile coarse_tile(int m, int n) {
    Tile t;
    t.H = 1;
    t.HY = 1;
    t.TX = m;
    t.TY = n;
    t.x0 = t.y0 = 0;
    t.gx0 = t.gy0 = -1;
    t.RX = m + 2;
    t.RY = n + 2;
    return t;
}

template <class T>
__device__ __forceinline__ LevelT<T> coarse_level(const CoarseDesc &d) {
    LevelT<T> L{};
    L.m = d.m;
    L.n = d.n;
    L.cx = T(d.cx);
    L.cy = T(d.cy);
    L.inv_int = T(d.inv);
    return L;
}

// Exact solve of the level by separation of variables (the factors of
// MG::coarse_factors, staged in the double workspace ws): bh = (V^T (x) I) b,
// one Thomas solve per mode along the kept direction, z = (V (x) I) zh.
// Unknown (p, t) sits at tile column p, row t + yoff (tdir 0) or column t,
// row p + yoff (tdir 1); yoff = 1 for the nodal system, whose row 0 is the
// Dirichlet row z = b / a_cc: its coupling to row 1 goes to the right-hand
// side, kept in z's row 1 until the solve overwrites it.
template <class T, bool FEM>
__device__ __forceinline__ void coarse_exact(const CoarseArgs<T> &a, T *z, const T *b, double *ws) {
    const CoarseDesc &d = a.d;
    const int tid = threadIdx.x + kCoarseBX * threadIdx.y, Tn = kCoarseThreads;
    const int RX = d.m + 2, Tt = d.T, P = d.P, PP = P + 1;
    double *hat = ws;
    const double *V = hat + (size_t)Tt * PP, *iv = V + (size_t)Tt * Tt, *cp = iv + (size_t)Tt * PP;
    if (FEM) {
        for (int i = tid; i < d.m; i += Tn) z[RX + i + 1] = b[RX + i + 1] / fem_acc<T>(a.fop, d.m, d.n, i, 0);
        __syncthreads();
        for (int i = tid; i < d.m; i += Tn) {
            const int l = 2 * RX + i + 1;
            const T wi = (i == 0 || i == d.m - 1) ? T(0.5) : T(1);
            z[l] = b[l] - (T(a.fop.fc) * (z[l - RX - 1] + z[l - RX + 1]) + wi * T(a.fop.fey) * z[l - RX]);
        }
        __syncthreads();
    }
    const int yoff = FEM ? 1 : 0;
    // tile index of (p, t = 0) and the step per t
    const int tstep = d.tdir == 0 ? RX : 1;
    // Both transforms are blocked kSepModes x kSepCols per thread: modes
    // (rows) k0.., columns p, p + PG, ...; a value read from LDS serves
    // kSepModes or kSepCols products.
    const int KG = (Tt + kSepModes - 1) / kSepModes, PG = (P + kSepCols - 1) / kSepCols;
    for (int o = tid; o < KG * PG; o += Tn) {
        const int kg = o / PG, pg = o - kg * PG, k0 = kg * kSepModes;
        int l[kSepCols];
#pragma unroll
        for (int j = 0; j < kSepCols; j++) {
            const int p = min(pg + j * PG, P - 1);
            l[j] = d.tdir == 0 ? (yoff + 1) * RX + p + 1 : (p + yoff + 1) * RX + 1;
        }
        double acc[kSepModes][kSepCols];
#pragma unroll
        for (int u = 0; u < kSepModes; u++)
#pragma unroll
            for (int j = 0; j < kSepCols; j++) acc[u][j] = 0;
#pragma unroll 4
        for (int t = 0; t < Tt; t++) {
            double bv[kSepCols], vv[kSepModes];
#pragma unroll
            for (int j = 0; j < kSepCols; j++) {
                const bool row1 = FEM && l[j] < 3 * RX;  // the corrected right-hand side of row 1
                bv[j] = double(row1 ? z[l[j]] : b[l[j]]);
                l[j] += tstep;
            }
#pragma unroll
            for (int u = 0; u < kSepModes; u++) vv[u] = V[t * Tt + min(k0 + u, Tt - 1)];
#pragma unroll
            for (int u = 0; u < kSepModes; u++)
#pragma unroll
                for (int j = 0; j < kSepCols; j++) acc[u][j] = fma(vv[u], bv[j], acc[u][j]);
        }
#pragma unroll
        for (int u = 0; u < kSepModes; u++)
#pragma unroll
            for (int j = 0; j < kSepCols; j++) {
                const int p = pg + j * PG;
                if (k0 + u < Tt && p < P) hat[(k0 + u) * PP + p] = acc[u][j];
            }
    }
    __syncthreads();
    // one Thomas solve per mode, y_p = r_p iv_p - cp_p y_(p-1) (cp_p =
    // e iv_p), so that each step is one dependent multiply-add; the next four
    // values are loaded while the current four are eliminated
    for (int k = tid; k < Tt; k += Tn) {
        double *h = hat + k * PP;
        const double *v = iv + k * PP, *c = cp + k * PP;
        double y = 0;
        const int P4 = P & ~3;
        double hn[4], wn[4], cn[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            hn[u] = h[min(u, P - 1)];
            wn[u] = v[min(u, P - 1)];
            cn[u] = c[min(u, P - 1)];
        }
        for (int p = 0; p < P4; p += 4) {
            double rc[4], cc[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                rc[u] = hn[u] * wn[u];
                cc[u] = cn[u];
                hn[u] = h[min(p + 4 + u, P - 1)];
                wn[u] = v[min(p + 4 + u, P - 1)];
                cn[u] = c[min(p + 4 + u, P - 1)];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                y = fma(-cc[u], y, rc[u]);
                h[p + u] = y;
            }
        }
        for (int p = P4; p < P; p++) h[p] = y = fma(-c[p], y, h[p] * v[p]);
        // back substitution x_p = y_p - cp_p x_(p+1) from p = P - 2 down, four at a time
        int p = P - 2;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            hn[u] = h[max(p - u, 0)];
            wn[u] = c[max(p - u, 0)];
        }
        for (; p >= 3; p -= 4) {
            double hc[4], wc[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                hc[u] = hn[u];
                wc[u] = wn[u];
                hn[u] = h[max(p - 4 - u, 0)];
                wn[u] = c[max(p - 4 - u, 0)];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                y = fma(-wc[u], y, hc[u]);
                h[p - u] = y;
            }
        }
        for (; p >= 0; p--) h[p] = y = fma(-c[p], y, h[p]);
    }
    __syncthreads();
    // the back transform, blocked the same way (rows t0.., columns p, p + PG, ...)
    for (int o = tid; o < KG * PG; o += Tn) {
        const int tg = o / PG, pg = o - tg * PG, t0 = tg * kSepModes;
        double acc[kSepModes][kSepCols];
#pragma unroll
        for (int u = 0; u < kSepModes; u++)
#pragma unroll
            for (int j = 0; j < kSepCols; j++) acc[u][j] = 0;
#pragma unroll 4
        for (int k = 0; k < Tt; k++) {
            double hv[kSepCols], vv[kSepModes];
#pragma unroll
            for (int j = 0; j < kSepCols; j++) hv[j] = hat[k * PP + min(pg + j * PG, P - 1)];
#pragma unroll
            for (int u = 0; u < kSepModes; u++) vv[u] = V[min(t0 + u, Tt - 1) * Tt + k];
#pragma unroll
            for (int u = 0; u < kSepModes; u++)
#pragma unroll
                for (int j = 0; j < kSepCols; j++) acc[u][j] = fma(vv[u], hv[j], acc[u][j]);
        }
#pragma unroll
        for (int u = 0; u < kSepModes; u++)
#pragma unroll
            for (int j = 0; j < kSepCols; j++) {
                const int t = t0 + u, p = pg + j * PG;
                if (t >= Tt || p >= P) continue;
                const int l = d.tdir == 0 ? (t + yoff + 1) * RX + p + 1 : (p + yoff + 1) * RX + t + 1;
                z[l] = T(acc[u][j]);
            }
    }
    __syncthreads();
}

// ncs smoothing sweeps of a level that is not separable (coarsening stopped
// early on odd sizes with a long shorter side).
template <class T>
__device__ __forceinline__ void coarse_sweeps(bool fem, const CoarseDesc &d, const CCOp &cop, const FemOp &fop,
                                              T *z, const T *b, const T *itab, int count) {
    const LevelT<T> L = coarse_level<T>(d);
    const Tile t = coarse_tile(L.m, L.n);
    for (int c = 0; c < count; c++) {
        if (fem) {
            for (int q = 0; q < 4; q++) {
                fem_tile_sweep<T, kCoarseBX, kCoarseBY>(L, fop, z, b, t, 1, q & 1, q >> 1);
                __syncthreads();
            }
        } else {
            for (int q = 0; q < 2; q++) {
                cc_tile_sweep<T, kCoarseBX, kCoarseBY>(L, cop, z, b, t, 1, q, itab);
                __syncthreads();
            }
        }
    }
}

template <class T, bool FEM>
__device__ __forceinline__ void coarse_body_t(const CoarseArgs<T> &a, T *lds, double *sh) {
    const CoarseDesc &d = a.d;
    const int tid = threadIdx.x + kCoarseBX * threadIdx.y, Tn = kCoarseThreads;
    const T *gb = reinterpret_cast<const T *>(d.gb);
    T *gz = reinterpret_cast<T *>(d.gz);
    const int m = d.m, N = m * d.n, RX = m + 2, NT = RX * (d.n + 2);
    T *z = lds, *b = z + NT, *itab = b + NT;
    double *ws = reinterpret_cast<double *>(reinterpret_cast<char *>(lds) + d.ws);
    // the sweeps start from the incoming iterate on a top level, from zero
    // below it; the exact solve needs no start
    const bool load_z = !d.sep && d.top;
    // load the level into the padded tiles; the first round of loads is
    // issued before the tiles are cleared (the borders stay zero) and the
    // factors are staged, so that its round trip overlaps them
    {
        Walk w(tid, m, Tn);
        T zv[kLoadUnroll], bv[kLoadUnroll];
        int li[kLoadUnroll];
        auto load_round = [&]() {
#pragma unroll
            for (int u = 0; u < kLoadUnroll; u++) {
                const int k = min(w.q, N - 1);
                li[u] = w.q < N ? (w.P + 1) * RX + w.r + 1 : -1;
                bv[u] = gb[k];
                zv[u] = load_z ? gz[k] : T(0);
                w.next(Tn);
            }
        };
        auto store_round = [&]() {
#pragma unroll
            for (int u = 0; u < kLoadUnroll; u++)
                if (li[u] >= 0) {
                    b[li[u]] = bv[u];
                    z[li[u]] = zv[u];
                }
        };
        // V, the Thomas factors and the off-diagonals, in their LDS layout:
        // kSepLoads values per thread and round, issued with the level's
        const int nf = d.sep ? d.T * d.T + 2 * d.T * (d.P + 1) + d.T : 0;
        double *dst = ws + d.T * (d.P + 1);
        double fv[kSepLoads];
        int fbase = tid;
        auto fac_round = [&]() {
#pragma unroll
            for (int u = 0; u < kSepLoads; u++) fv[u] = d.fac[min(fbase + u * Tn, nf - 1)];
        };
        auto fac_store = [&]() {
#pragma unroll
            for (int u = 0; u < kSepLoads; u++)
                if (fbase + u * Tn < nf) dst[fbase + u * Tn] = fv[u];
            fbase += kSepLoads * Tn;
        };
        load_round();
        if (fbase < nf) fac_round();
        for (int k = tid; k < 2 * NT; k += Tn) lds[k] = T(0);
        if (!d.sep && !FEM && tid < 9) cc_inv_table<T>(coarse_level<T>(d), a.cop, itab, tid);
        __syncthreads();
        store_round();
        if (fbase < nf) fac_store();
        while (fbase < nf) {
            fac_round();
            fac_store();
        }
        while (w.q < N) {
            load_round();
            store_round();
        }
    }
    __syncthreads();
    if (d.sep) coarse_exact<T, FEM>(a, z, b, ws);
    else coarse_sweeps<T>(FEM, d, a.cop, a.fop, z, b, itab, a.ncs);
    const LevelT<T> F = coarse_level<T>(d);
    double rr = 0;
    for (int j = threadIdx.y; j < F.n; j += kCoarseBY)
        for (int i = threadIdx.x; i < F.m; i += kCoarseBX) {
            const int l0 = (j + 1) * RX + i + 1;
            gz[i + (size_t)F.m * j] = z[l0];
            if (d.top) {
                const double r = FEM ? fem_resid_t<T>(a.fop, F.m, F.n, z, b, RX, l0, i, j)
                                     : cc_resid_t<T>(F, a.cop, z, b, RX, l0, i, j);
                rr += r * r;
            }
        }
    if (d.top) finish_norm_single(a.ctl, rr, a.tol2, sh);
}

template <class T>
__device__ __forceinline__ void coarse_body(const CoarseArgs<T> &a, T *lds, double *sh, int *last) {
    if (a.fem) coarse_body_t<T, true>(a, lds, sh);
    else coarse_body_t<T, false>(a, lds, sh);
    (void)last;
}

template <class T>
__global__ void __launch_bounds__(kCoarseThreads) k_coarse(CoarseArgs<T> a) {
    extern __shared__ double lds_d[];
    __shared__ double sh[8];
    __shared__ int last;
    if (a.ctl.state->done) return;
    stamp_begin(a.ts);
    coarse_body<T>(a, reinterpret_cast<T *>(lds_d), sh, &last);
    stamp_end(a.ts);
}

// One workgroup per system.
template <class T>
__global__ void __launch_bounds__(kCoarseThreads) k_coarse_b(const CoarseArgs<T> *A, const WorkItem *W, int nw,
                                                              BatchAux aux) {
    extern __shared__ double lds_d[];
    __shared__ double sh[8];
    __shared__ int last;
    stamp_begin(aux.ts);
    for (int p = blockIdx.x; p < nw; p += gridDim.x) {
        const WorkItem w = W[p];
        const CoarseArgs<T> &a = A[w.rec];  // read in place: a copy would go to scratch
        if (a.ctl.state->done) continue;
        const unsigned long long t0 = busy_begin(aux.busy);
        coarse_body<T>(a, reinterpret_cast<T *>(lds_d), sh, &last);
        __syncthreads();
        busy_end(aux.busy, w.sim, t0);
    }
    stamp_end(aux.ts);
}

// ------------------------------------------------------------------------
// Multi-workgroup exact solve of a larger coarsest level (both directions
// diagonalised, MG::coarse_factors).  Phases 1 and 3: out = in M over rows
// (M = U, then U^T), a workgroup owning kXRows x kXCols outputs with its
// input rows and M's columns staged in LDS; phase 2: kXColsV columns of w
// through V^T, the diagonal 1 / (lam + mu) and V.  Arithmetic in FP64; each
// output's dot product is shared by four lanes (x_dot_quad).
// ------------------------------------------------------------------------

// Right-hand side of unknown row t (level row t + yoff) at column p: the
// nodal row 1 loses its coupling to the Dirichlet row z0 = b0 / a_cc.
template <class T>
__device__ __forceinline__ double x_rhs(const XArgs<T> &a, int t, int p) {
    const int y = t + a.yoff, m = a.m;
    double v = double(a.b[(size_t)y * m + p]);
    if (a.yoff && y == 1) {
        auto z0 = [&](int q) {
            return (q < 0 || q >= m) ? 0.0 : double(a.b[q]) / fem_acc<double>(a.fop, m, a.n, q, 0);
        };
        const double wi = (p == 0 || p == m - 1) ? 0.5 : 1.0;
        v -= a.fop.fc * (z0(p - 1) + z0(p + 1)) + wi * a.fop.fey * z0(p);
    }
    return v;
}

// Staging of two arrays into LDS in shared rounds: each thread loads up to
// kXStage values of the first and kXStage2 of the second before it stores any
// of them, so that a round costs one memory round trip.
constexpr int kXStage = 16, kXStage2 = 4;
template <class F1, class F2>
__device__ __forceinline__ void x_stage2(double *d1, int n1, F1 f1, double *d2, int n2, F2 f2) {
    for (int b1 = threadIdx.x, b2 = threadIdx.x; b1 < n1 || b2 < n2;
         b1 += kXStage * kXThreads, b2 += kXStage2 * kXThreads) {
        double v1[kXStage], v2[kXStage2];
#pragma unroll
        for (int u = 0; u < kXStage; u++) v1[u] = f1(min(b1 + u * kXThreads, n1 - 1));
#pragma unroll
        for (int u = 0; u < kXStage2; u++) v2[u] = f2(min(b2 + u * kXThreads, n2 - 1));
#pragma unroll
        for (int u = 0; u < kXStage; u++)
            if (b1 + u * kXThreads < n1) d1[b1 + u * kXThreads] = v1[u];
#pragma unroll
        for (int u = 0; u < kXStage2; u++)
            if (b2 + u * kXThreads < n2) d2[b2 + u * kXThreads] = v2[u];
    }
}

// sum_q x[q * sx] y[q * sy], q < n, with four independent partial sums
__device__ __forceinline__ double x_dot(const double *x, int sx, const double *y, int sy, int n) {
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int q = 0;
    for (; q + 4 <= n; q += 4) {
        s0 = fma(x[q * sx], y[q * sy], s0);
        s1 = fma(x[(q + 1) * sx], y[(q + 1) * sy], s1);
        s2 = fma(x[(q + 2) * sx], y[(q + 2) * sy], s2);
        s3 = fma(x[(q + 3) * sx], y[(q + 3) * sy], s3);
    }
    for (; q < n; q++) s0 = fma(x[q * sx], y[q * sy], s0);
    return (s0 + s1) + (s2 + s3);
}

// x_dot on the four lanes k = 0..3 of a quad (consecutive lanes): lane k
// accumulates x_dot's partial sum s_k (terms q = k, k + 4, ... below n - n % 4
// in order; lane 0 then the remaining terms), and the quad combines them as
// x_dot does, (s0 + s1) + (s2 + s3), so the result is x_dot's bit for bit.
// Every lane of the wavefront must call it.
__device__ __forceinline__ double x_dot_quad(const double *x, int sx, const double *y, int sy, int n, int k) {
    const int n4 = n & ~3;
    double s = 0;
    for (int q = k; q < n4; q += 4) s = fma(x[q * sx], y[q * sy], s);
    if (k == 0)
        for (int q = n4; q < n; q++) s = fma(x[q * sx], y[q * sy], s);
    const double s2 = s + __shfl_xor(s, 1);  // lanes 0, 1: s0 + s1; lanes 2, 3: s2 + s3
    return s2 + __shfl_xor(s2, 2);
}

template <class T>
__device__ void xsolve_body(const XArgs<T> &a, int blk, double *lds) {
    const int tid = threadIdx.x, P = a.P, Tn = a.Tn;
    const double *V = a.fac, *U = V + (size_t)Tn * Tn, *Ut = U + (size_t)P * P, *D = Ut + (size_t)P * P;
    if (a.phase != 2) {
        const int R = a.quad ? kXRowsQ : kXRows;
        const int rb = blk / a.nbc, cb = blk - rb * a.nbc, r0 = rb * R, c0 = cb * kXCols;
        double *in = lds, *mx = lds + R * P;
        const double *M = a.phase == 1 ? U : Ut;
        x_stage2(
            mx, P * kXCols,
            [&](int i) {
                const int q = i >> 5, c = i & (kXCols - 1);
                return M[(size_t)q * P + min(c0 + c, P - 1)];
            },
            in, R * P,
            [&](int i) {
                const int r = i / P, p = i - r * P, t = min(r0 + r, Tn - 1);
                return a.phase == 1 ? x_rhs<T>(a, t, p) : double(a.w[(size_t)t * P + p]);
            });
        __syncthreads();
        // output o (quad layout: quarter k of its dot product)
        const int o = a.quad ? tid >> 2 : tid, k = a.quad ? tid & 3 : 0;
        const int r = o / kXCols, c = o - r * kXCols, t = r0 + r, j = c0 + c;
        const double acc = a.quad ? x_dot_quad(in + r * P, 1, mx + c, kXCols, P, k) : x_dot(in + r * P, 1, mx + c, kXCols, P);
        if (k == 0 && t < Tn && j < P) {
            if (a.phase == 1) a.w[(size_t)t * P + j] = T(acc);
            else a.z[(size_t)(t + a.yoff) * a.m + j] = T(acc);
        }
        // the nodal Dirichlet row z0 = b0 / a_cc (row-block 0 of phase 1)
        if (a.phase == 1 && a.yoff && rb == 0 && tid < kXCols && c0 + tid < P)
            a.z[c0 + tid] = T(double(a.b[c0 + tid]) / fem_acc<double>(a.fop, a.m, a.n, c0 + tid, 0));
        return;
    }
    const int CV = a.quad ? kXColsVQ : kXColsV;
    const int j0 = blk * CV;
    double *Vl = lds, *wl = Vl + Tn * Tn, *hl = wl + Tn * CV;
    x_stage2(
        Vl, Tn * Tn, [&](int i) { return V[i]; }, wl, Tn * CV,
        [&](int i) {
            const int t = i / CV, c = i - t * CV;
            return double(a.w[(size_t)t * P + min(j0 + c, P - 1)]);
        });
    __syncthreads();
    if (a.quad) {
        // outputs o < Tn CV (at most kXThreads / 4: Tn <= 64), quarter q4
        const int no = Tn * CV, o = min(tid >> 2, no - 1), q4 = tid & 3;
        {
            const int k = o / CV, c = o - k * CV;
            const double v = x_dot_quad(Vl + k, Tn, wl + c, CV, Tn, q4) * D[(size_t)k * P + min(j0 + c, P - 1)];
            if (q4 == 0 && (tid >> 2) < no) hl[o] = v;
        }
        __syncthreads();
        const int t = o / CV, c = o - t * CV;
        const double acc = x_dot_quad(Vl + t * Tn, 1, hl + c, CV, Tn, q4);
        if (q4 == 0 && (tid >> 2) < no && j0 + c < P) a.w[(size_t)t * P + j0 + c] = T(acc);
        return;
    }
    for (int i = tid; i < Tn * CV; i += kXThreads) {
        const int k = i / CV, c = i - k * CV;
        hl[i] = x_dot(Vl + k, Tn, wl + c, CV, Tn) * D[(size_t)k * P + min(j0 + c, P - 1)];
    }
    __syncthreads();
    for (int i = tid; i < Tn * CV; i += kXThreads) {
        const int t = i / CV, c = i - t * CV;
        const double acc = x_dot(Vl + t * Tn, 1, hl + c, CV, Tn);
        if (j0 + c < P) a.w[(size_t)t * P + j0 + c] = T(acc);
    }
}

template <class T>
__global__ void __launch_bounds__(kXThreads) k_xsolve(XArgs<T> a) {
    extern __shared__ double lds_d[];
    if (a.ctl.state->done) return;
    stamp_begin(a.ts);
    xsolve_body<T>(a, blockIdx.x, lds_d);
    stamp_end(a.ts);
}

template <class T>
__global__ void __launch_bounds__(kXThreads) k_xsolve_b(const XArgs<T> *A, const WorkItem *W, int nw, BatchAux aux) {
    extern __shared__ double lds_d[];
    stamp_begin(aux.ts);
    for (int p = blockIdx.x; p < nw; p += gridDim.x) {
        const WorkItem w = W[p];
        const XArgs<T> &a = A[w.rec];  // read in place: a copy would go to scratch
        if (a.ctl.state->done) continue;
        const unsigned long long t0 = busy_begin(aux.busy);
        xsolve_body<T>(a, w.tile, lds_d);
        __syncthreads();
        busy_end(aux.busy, w.sim, t0);
    }
    stamp_end(aux.ts);
}

// ------------------------------------------------------------------------
// Start of a solve, element-wise operations, operator application
// ------------------------------------------------------------------------

// Partials of ||b - A z||^2 and ||b||^2 of level 0 over blocks bid, bid + nb,
// ...; k_norm_finish (begin = 1) sums them and sets the flag: 1 converged,
// 2 NaN, 3 zero right-hand side, 0 otherwise.  With r32 non-null (mixed form)
// the residual is also stored in FP32.
__device__ void begin_body(const BeginArgs &a, int bid, double *sh, int *last) {
    const Level &L = a.L;
    double rr = 0, bb = 0;
    const int N = L.m * L.n;
    // row-major walk without 64-bit division: (i, j) advance with ij
    int ij = bid * 256 + threadIdx.x;
    const int stride = a.nb * 256;
    int j = ij / L.m, i = ij - j * L.m;
    const int sj = stride / L.m, si = stride - sj * L.m;
    for (; ij < N; ij += stride) {
        const double bv = L.b[ij], r = bv - Az_plain<double>(a.fem, a.cop, a.fop, L.m, L.n, L.cx, L.cy, L.z, i, j);
        if (a.r32) a.r32[ij] = (float)r;
        rr += r * r;
        bb += bv * bv;
        i += si;
        j += sj;
        if (i >= L.m) { i -= L.m; j++; }
    }
    rr = block_sum(rr, sh);
    bb = block_sum(bb, sh);
    if (threadIdx.x == 0) {
        a.ctl.partial[2 * bid] = rr;
        a.ctl.partial[2 * bid + 1] = bb;
    }
    (void)last;
}

// Second half of a reduction: one workgroup sums the partials of the
// previous launch in a fixed order and updates the solve state.
//   begin = 1: partials are (||r||^2, ||b||^2) pairs; sets rr, bb, cycles = 0
//              and the flag (1 converged, 2 NaN, 3 zero right-hand side);
//   begin = 0: partials are ||r||^2 after a V-cycle; cycles += 1.
__device__ void norm_finish_body(const NormArgs &a, double *sh) {
    const int T = blockDim.x * blockDim.y;
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    if (a.begin) {
        double r = 0, b = 0;
        for (int k = tid; k < a.nb; k += T) { r += a.ctl.partial[2 * k]; b += a.ctl.partial[2 * k + 1]; }
        const double R = block_sum(r, sh), B = block_sum(b, sh);
        if (tid == 0) {
            MGState *st = a.ctl.state;
            st->rr = R;
            st->bb = B;
            st->cycles = 0;
            st->done = (B == 0) ? 3 : (!(R == R) ? 2 : (R <= a.tol2 * B ? 1 : 0));
        }
    } else {
        // the partials are read before the flag (not after it) so that the
        // two memory round trips overlap; a solve that had already converged
        // leaves its state alone
        double r = 0;
        for (int k = tid; k < a.nb; k += T) r += a.ctl.partial[k];
        const double R = block_sum(r, sh);
        if (tid == 0) {
            MGState *st = a.ctl.state;
            if (!st->done) {
                st->rr = R;
                st->cycles += 1;
                st->done = !(R == R) ? 2 : (R <= a.tol2 * st->bb ? 1 : 0);
            }
        }
    }
}

__global__ void __launch_bounds__(256) k_norm_finish(NormArgs a) {
    __shared__ double sh[8];
    norm_finish_body(a, sh);
}

__global__ void __launch_bounds__(256) k_norm_finish_b(const NormArgs *A, int n) {
    __shared__ double sh[8];
    for (int p = blockIdx.x; p < n; p += gridDim.x) {
        norm_finish_body(A[p], sh);
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_begin(BeginArgs a) {
    __shared__ double sh[8];
    __shared__ int last;
    stamp_begin(a.ts);
    begin_body(a, blockIdx.x, sh, &last);
    if (a.reset && blockIdx.x == 0 && threadIdx.x == 0) {
        a.ctl.state->done = 0;
        a.ctl.state->cycles = 0;
    }
    stamp_end(a.ts);
}

__global__ void __launch_bounds__(256) k_begin_b(const BeginArgs *A, const WorkItem *W, int nw, BatchAux aux) {
    __shared__ double sh[8];
    __shared__ int last;
    stamp_begin(aux.ts);
    for (int p = blockIdx.x; p < nw; p += gridDim.x) {
        const WorkItem w = W[p];
        const unsigned long long t0 = busy_begin(aux.busy);
        begin_body(A[w.rec], w.tile, sh, &last);
        busy_end(aux.busy, w.sim, t0);
    }
    stamp_end(aux.ts);
}

// Element-wise work of block bid (256 entries per block and round):
//   VEC_EXTRAP    mgs_common.cc:37-44: z <- b z - c z_prev, z_prev <- old z,
//                 with (b, c) read from device memory so that a captured
//                 graph can be replayed with new times;
//   VEC_SOLVE_END end of a solve whose cycle count is only known on the
//                 device: the mixed form leaves its iterate in zalt after an
//                 odd number of cycles, and a zero right-hand side means z = 0;
//   VEC_DIRICHLET FEM Dirichlet row z = b / a_cc (row 0, M entries).
__device__ __forceinline__ void vec_body(const VecArgs &a, int bid) {
    const long stride = (long)a.nb * 256;
    if (a.op == VEC_EXTRAP) {
        const double bb = a.coef[0], c = a.coef[1];
        for (long ij = bid * 256L + threadIdx.x; ij < a.N; ij += stride) {
            const double zz = a.z[ij];
            a.z[ij] = bb * zz - c * a.zp[ij];
            a.zp[ij] = zz;
        }
    } else if (a.op == VEC_SOLVE_END) {
        const int done = a.state->done, odd = a.mixed && (a.state->cycles & 1);
        if (done != 3 && !odd) return;
        for (long ij = bid * 256L + threadIdx.x; ij < a.N; ij += stride) a.z[ij] = done == 3 ? 0.0 : a.zalt[ij];
    } else {
        for (long i = bid * 256L + threadIdx.x; i < a.M; i += stride)
            a.z[i] = a.b[i] / ((i > 0 && i < a.M - 1 ? 0.5 : 0.25) * a.fm);
    }
}

__global__ void __launch_bounds__(256) k_vec(VecArgs a) { vec_body(a, blockIdx.x); }

__global__ void __launch_bounds__(256) k_vec_b(const VecArgs *A, const WorkItem *W, int nw, BatchAux aux) {
    stamp_begin(aux.ts);
    for (int p = blockIdx.x; p < nw; p += gridDim.x) {
        const WorkItem w = W[p];
        vec_body(A[w.rec], w.tile);
    }
    stamp_end(aux.ts);
}

__global__ void k_apply(int fem, Level L, CCOp cop, FemOp fop, double *out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= L.m || j >= L.n) return;
    out[i + (size_t)L.m * j] = Az_plain<double>(fem, cop, fop, L.m, L.n, L.cx, L.cy, L.z, i, j);
}

// mgs_common.cc:37-44 with the coefficients by value (direct path).
__global__ void k_extrapolate(double *z, double *zp, long N, double bb, double c) {
    long ij = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (ij >= N) return;
    const double zz = z[ij];
    z[ij] = bb * zz - c * zp[ij];
    zp[ij] = zz;
}

// r = b - A z for a level-0 iterate z (any array), and per-block partial
// sums of r^2 (block order, summed later in index order).
__global__ void __launch_bounds__(256) k_resid_out(int fem, Level L, CCOp cop, FemOp fop, const double *z, double *r,
                                                   double *part) {
    __shared__ double sh[8];
    const int N = L.m * L.n;
    double rr = 0;
    for (int ij = blockIdx.x * 256 + threadIdx.x; ij < N; ij += gridDim.x * 256) {
        const int j = ij / L.m, i = ij - j * L.m;
        const double v = L.b[ij] - Az_plain<double>(fem, cop, fop, L.m, L.n, L.cx, L.cy, z, i, j);
        if (r) r[ij] = v;
        rr += v * v;
    }
    rr = block_sum(rr, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = rr;
}

__global__ void k_zero(double *z, long N) {
    long ij = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (ij < N) z[ij] = 0;
}

__global__ void k_copy_d(const double *a, double *b, long N) {
    long ij = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (ij < N) b[ij] = a[ij];
}

__global__ void k_reset_state(MGState *st) {
    st->done = 0;
    st->cycles = 0;
}

// ------------------------------------------------------------------------
// Launchers
// ------------------------------------------------------------------------

static size_t tile_lds_bytes(int H, int HY, size_t elem, int tx, int ty) {
    return 2 * elem * (size_t)(tx + 2 * H) * (ty + 2 * HY) + kInvTab * elem;
}

namespace mgk {

static int g_cus = 0, g_max_lds = 0;

int set_attributes() {
    static bool done = false;
    if (done) return VDI_OK;
    int dev = 0;
    VDI_HIP(hipGetDevice(&dev));
    VDI_HIP(hipDeviceGetAttribute(&g_max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    VDI_HIP(hipDeviceGetAttribute(&g_cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (g_max_lds < 65536) return VDI_ERR_ARG;
    const void *ks[] = {
        (const void *)k_down<double, false>, (const void *)k_down<double, true>, (const void *)k_down<float, false>,
        (const void *)k_down<float, true>, (const void *)k_up<double, false>, (const void *)k_up<double, true>,
        (const void *)k_up<float, false>, (const void *)k_up<float, true>, (const void *)k_down_b<double, false>,
        (const void *)k_down_b<double, true>, (const void *)k_down_b<float, false>, (const void *)k_down_b<float, true>,
        (const void *)k_up_b<double, false>, (const void *)k_up_b<double, true>, (const void *)k_up_b<float, false>,
        (const void *)k_up_b<float, true>, (const void *)k_coarse<double>, (const void *)k_coarse<float>,
        (const void *)k_coarse_b<double>, (const void *)k_coarse_b<float>, (const void *)k_mx_up0<false>,
        (const void *)k_mx_up0<true>, (const void *)k_mx_up0_b<false>, (const void *)k_mx_up0_b<true>,
        (const void *)k_xsolve<double>, (const void *)k_xsolve<float>, (const void *)k_xsolve_b<double>,
        (const void *)k_xsolve_b<float>};
    // dynamic LDS may use what the kernel's own static LDS leaves
    int dyn = g_max_lds;
    for (const void *k : ks) {
        hipFuncAttributes fa;
        VDI_HIP(hipFuncGetAttributes(&fa, k));
        dyn = std::min(dyn, g_max_lds - (int)fa.sharedSizeBytes);
    }
    for (const void *k : ks) VDI_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, dyn));
    g_max_lds = dyn;
    done = true;
    return VDI_OK;
}

int max_lds() { return g_max_lds; }

int persistent_grid(const void *kernel, size_t lds, int nw) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, lds) != hipSuccess || per_cu < 1) per_cu = 1;
    const int full = per_cu * std::max(1, g_cus);
    const int g = std::min(full, nw);
    return std::max(8, (g + 7) / 8 * 8);
}

const void *down_kernel(int prec, bool fem, bool b) {
    if (prec == 0) return fem ? (b ? (const void *)k_down_b<double, true> : (const void *)k_down<double, true>)
                              : (b ? (const void *)k_down_b<double, false> : (const void *)k_down<double, false>);
    return fem ? (b ? (const void *)k_down_b<float, true> : (const void *)k_down<float, true>)
               : (b ? (const void *)k_down_b<float, false> : (const void *)k_down<float, false>);
}
const void *up_kernel(int prec, bool fem, bool b) {
    if (prec == 0) return fem ? (b ? (const void *)k_up_b<double, true> : (const void *)k_up<double, true>)
                              : (b ? (const void *)k_up_b<double, false> : (const void *)k_up<double, false>);
    return fem ? (b ? (const void *)k_up_b<float, true> : (const void *)k_up<float, true>)
               : (b ? (const void *)k_up_b<float, false> : (const void *)k_up<float, false>);
}
const void *coarse_kernel(int prec, bool b) {
    if (prec == 0) return b ? (const void *)k_coarse_b<double> : (const void *)k_coarse<double>;
    return b ? (const void *)k_coarse_b<float> : (const void *)k_coarse<float>;
}
const void *mx_kernel(bool fem, bool b) {
    return fem ? (b ? (const void *)k_mx_up0_b<true> : (const void *)k_mx_up0<true>)
               : (b ? (const void *)k_mx_up0_b<false> : (const void *)k_mx_up0<false>);
}

template <class T, template <class, int> class K>
static void launch_stream(const LvArgs<T> &a, hipStream_t st) {
    switch (a.nu) {
        case 0: hipLaunchKernelGGL((K<T, 0>::fn), dim3(a.nb), dim3(kSThreads), 0, st, a); break;
        case 1: hipLaunchKernelGGL((K<T, 1>::fn), dim3(a.nb), dim3(kSThreads), 0, st, a); break;
        case 2: hipLaunchKernelGGL((K<T, 2>::fn), dim3(a.nb), dim3(kSThreads), 0, st, a); break;
        case 3: hipLaunchKernelGGL((K<T, 3>::fn), dim3(a.nb), dim3(kSThreads), 0, st, a); break;
        default: hipLaunchKernelGGL((K<T, 4>::fn), dim3(a.nb), dim3(kSThreads), 0, st, a); break;
    }
}
template <class T, int NU> struct SDown { static constexpr auto fn = k_sdown<T, NU>; };
template <class T, int NU> struct SUp { static constexpr auto fn = k_sup<T, NU>; };
template <class T, int NU> struct SNDown { static constexpr auto fn = k_sndown<T, NU>; };
template <class T, int NU> struct SNUp { static constexpr auto fn = k_snup<T, NU>; };

template <class T> int down(const LvArgs<T> &a, bool fem, size_t lds, hipStream_t st) {
    if (a.items && fem) launch_stream<T, SNDown>(a, st);
    else if (a.items) launch_stream<T, SDown>(a, st);
    else if (fem) hipLaunchKernelGGL((k_down<T, true>), dim3(a.nb), dim3(kBX, kBY), lds, st, a);
    else hipLaunchKernelGGL((k_down<T, false>), dim3(a.nb), dim3(kBX, kBY), lds, st, a);
    VDI_LAUNCHED();
    return VDI_OK;
}
template <class T> int up(const LvArgs<T> &a, bool fem, size_t lds, hipStream_t st) {
    if (a.items && fem) launch_stream<T, SNUp>(a, st);
    else if (a.items) launch_stream<T, SUp>(a, st);
    else if (fem) hipLaunchKernelGGL((k_up<T, true>), dim3(a.nb), dim3(kBX, kBY), lds, st, a);
    else hipLaunchKernelGGL((k_up<T, false>), dim3(a.nb), dim3(kBX, kBY), lds, st, a);
    VDI_LAUNCHED();
    return VDI_OK;
}
template <class T> int coarse(const CoarseArgs<T> &a, size_t lds, hipStream_t st) {
    hipLaunchKernelGGL(k_coarse<T>, dim3(1), dim3(kCoarseBX, kCoarseBY), lds, st, a);
    VDI_LAUNCHED();
    return VDI_OK;
}
template <bool FEM>
static void launch_smx(const MxArgs &a, hipStream_t st) {
    switch (a.nu) {
        case 0: hipLaunchKernelGGL((k_smx<FEM, 0>), dim3(a.nb), dim3(kSThreads), 0, st, a); break;
        case 1: hipLaunchKernelGGL((k_smx<FEM, 1>), dim3(a.nb), dim3(kSThreads), 0, st, a); break;
        case 2: hipLaunchKernelGGL((k_smx<FEM, 2>), dim3(a.nb), dim3(kSThreads), 0, st, a); break;
        case 3: hipLaunchKernelGGL((k_smx<FEM, 3>), dim3(a.nb), dim3(kSThreads), 0, st, a); break;
        default: hipLaunchKernelGGL((k_smx<FEM, 4>), dim3(a.nb), dim3(kSThreads), 0, st, a); break;
    }
}
int mx(const MxArgs &a, bool fem, size_t lds, hipStream_t st) {
    if (a.items && fem) launch_smx<true>(a, st);
    else if (a.items) launch_smx<false>(a, st);
    else if (fem) hipLaunchKernelGGL(k_mx_up0<true>, dim3(a.nb), dim3(kBX, kBY), lds, st, a);
    else hipLaunchKernelGGL(k_mx_up0<false>, dim3(a.nb), dim3(kBX, kBY), lds, st, a);
    VDI_LAUNCHED();
    return VDI_OK;
}
int begin(const BeginArgs &a, hipStream_t st) {
    hipLaunchKernelGGL(k_begin, dim3(a.nb), dim3(256), 0, st, a);
    VDI_LAUNCHED();
    return VDI_OK;
}
int norm(const NormArgs &a, hipStream_t st) {
    hipLaunchKernelGGL(k_norm_finish, dim3(1), dim3(256), 0, st, a);
    VDI_LAUNCHED();
    return VDI_OK;
}
int norm_b(const NormArgs *A, int n, hipStream_t st) {
    if (n <= 0) return VDI_OK;
    hipLaunchKernelGGL(k_norm_finish_b, dim3(n), dim3(256), 0, st, A, n);
    VDI_LAUNCHED();
    return VDI_OK;
}
int vec(const VecArgs &a, hipStream_t st) {
    hipLaunchKernelGGL(k_vec, dim3(a.nb), dim3(256), 0, st, a);
    VDI_LAUNCHED();
    return VDI_OK;
}

template <class T> int down_b(const LvArgs<T> *A, const WorkItem *w, int nw, bool fem, size_t lds, int grid,
                              BatchAux aux, hipStream_t st) {
    if (fem) hipLaunchKernelGGL((k_down_b<T, true>), dim3(grid), dim3(kBX, kBY), lds, st, A, w, nw, aux);
    else hipLaunchKernelGGL((k_down_b<T, false>), dim3(grid), dim3(kBX, kBY), lds, st, A, w, nw, aux);
    VDI_LAUNCHED();
    return VDI_OK;
}
template <class T> int up_b(const LvArgs<T> *A, const WorkItem *w, int nw, bool fem, size_t lds, int grid,
                            BatchAux aux, hipStream_t st) {
    if (fem) hipLaunchKernelGGL((k_up_b<T, true>), dim3(grid), dim3(kBX, kBY), lds, st, A, w, nw, aux);
    else hipLaunchKernelGGL((k_up_b<T, false>), dim3(grid), dim3(kBX, kBY), lds, st, A, w, nw, aux);
    VDI_LAUNCHED();
    return VDI_OK;
}
template <class T> int xsolve_grid(const XArgs<T> &a) {
    const int cv = a.quad ? kXColsVQ : kXColsV;
    return a.phase == 2 ? (a.P + cv - 1) / cv : a.nbr * a.nbc;
}
template int xsolve_grid<double>(const XArgs<double> &);
template int xsolve_grid<float>(const XArgs<float> &);
template <class T> int xsolve(const XArgs<T> &a, size_t lds, hipStream_t st) {
    const int nb = xsolve_grid(a);
    hipLaunchKernelGGL(k_xsolve<T>, dim3(nb), dim3(kXThreads), lds, st, a);
    VDI_LAUNCHED();
    return VDI_OK;
}
template <class T> int xsolve_b(const XArgs<T> *A, const WorkItem *w, int nw, size_t lds, int grid, BatchAux aux,
                                hipStream_t st) {
    hipLaunchKernelGGL(k_xsolve_b<T>, dim3(grid), dim3(kXThreads), lds, st, A, w, nw, aux);
    VDI_LAUNCHED();
    return VDI_OK;
}
const void *xsolve_kernel(int prec, bool b) {
    if (prec == 0) return b ? (const void *)k_xsolve_b<double> : (const void *)k_xsolve<double>;
    return b ? (const void *)k_xsolve_b<float> : (const void *)k_xsolve<float>;
}
template <class T> int coarse_b(const CoarseArgs<T> *A, const WorkItem *w, int nw, size_t lds, int grid,
                                BatchAux aux, hipStream_t st) {
    hipLaunchKernelGGL(k_coarse_b<T>, dim3(grid), dim3(kCoarseBX, kCoarseBY), lds, st, A, w, nw, aux);
    VDI_LAUNCHED();
    return VDI_OK;
}
int mx_b(const MxArgs *A, const WorkItem *w, int nw, bool fem, size_t lds, int grid, BatchAux aux, hipStream_t st) {
    if (fem) hipLaunchKernelGGL(k_mx_up0_b<true>, dim3(grid), dim3(kBX, kBY), lds, st, A, w, nw, aux);
    else hipLaunchKernelGGL(k_mx_up0_b<false>, dim3(grid), dim3(kBX, kBY), lds, st, A, w, nw, aux);
    VDI_LAUNCHED();
    return VDI_OK;
}
int begin_b(const BeginArgs *A, const WorkItem *w, int nw, int grid, BatchAux aux, hipStream_t st) {
    hipLaunchKernelGGL(k_begin_b, dim3(grid), dim3(256), 0, st, A, w, nw, aux);
    VDI_LAUNCHED();
    return VDI_OK;
}
int vec_b(const VecArgs *A, const WorkItem *w, int nw, int grid, BatchAux aux, hipStream_t st) {
    hipLaunchKernelGGL(k_vec_b, dim3(grid), dim3(256), 0, st, A, w, nw, aux);
    VDI_LAUNCHED();
    return VDI_OK;
}

template int down<double>(const LvArgs<double> &, bool, size_t, hipStream_t);
template int down<float>(const LvArgs<float> &, bool, size_t, hipStream_t);
template int up<double>(const LvArgs<double> &, bool, size_t, hipStream_t);
template int up<float>(const LvArgs<float> &, bool, size_t, hipStream_t);
template int coarse<double>(const CoarseArgs<double> &, size_t, hipStream_t);
template int coarse<float>(const CoarseArgs<float> &, size_t, hipStream_t);
template int down_b<double>(const LvArgs<double> *, const WorkItem *, int, bool, size_t, int, BatchAux, hipStream_t);
template int down_b<float>(const LvArgs<float> *, const WorkItem *, int, bool, size_t, int, BatchAux, hipStream_t);
template int up_b<double>(const LvArgs<double> *, const WorkItem *, int, bool, size_t, int, BatchAux, hipStream_t);
template int up_b<float>(const LvArgs<float> *, const WorkItem *, int, bool, size_t, int, BatchAux, hipStream_t);
template int coarse_b<double>(const CoarseArgs<double> *, const WorkItem *, int, size_t, int, BatchAux, hipStream_t);
template int xsolve<double>(const XArgs<double> &, size_t, hipStream_t);
template int xsolve<float>(const XArgs<float> &, size_t, hipStream_t);
template int xsolve_b<double>(const XArgs<double> *, const WorkItem *, int, size_t, int, BatchAux, hipStream_t);
template int xsolve_b<float>(const XArgs<float> *, const WorkItem *, int, size_t, int, BatchAux, hipStream_t);
template int coarse_b<float>(const CoarseArgs<float> *, const WorkItem *, int, size_t, int, BatchAux, hipStream_t);

}  // namespace mgk

// ------------------------------------------------------------------------
// KernelTimer (declared in vdi_internal.hpp)
// ------------------------------------------------------------------------

__global__ void k_stamp_reset(Stamp *s, int n) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        s[k].t0 = ~0ull;
        s[k].t1 = 0;
    }
}

int KernelTimer::setup(int c, int capacity) {
    release();
    cls = c;
    if (c < 0) return VDI_OK;
    cap = capacity;
    // workgroup stamps: up to 512 per launch on average, at most 4 Mi (64 MiB)
    ncap = std::min<long>(4L << 20, std::max<long>(4096, 512L * cap));
    VDI_TRY(stamps.alloc(sizeof(Stamp) * (size_t)ncap));
    VDI_HIP(hipHostMalloc((void **)&host, sizeof(Stamp) * (size_t)ncap, hipHostMallocDefault));
    slot_bytes.assign(cap, 0.0);
    off.assign(cap + 1, 0);
    int dev = 0, khz = 0;
    VDI_HIP(hipGetDevice(&dev));
    VDI_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    ticks_per_ms = khz > 0 ? (double)khz : 1e5;
    return VDI_OK;
}

int KernelTimer::reset(int a, int b, hipStream_t st) {
    if (b <= a || a > next) return VDI_OK;
    const long s0 = off[a], s1 = b >= cap ? ncap : off[std::min(b, next)];
    if (s1 <= s0) return VDI_OK;
    hipLaunchKernelGGL(k_stamp_reset, dim3((unsigned)((s1 - s0 + 255) / 256)), dim3(256), 0, st, stamps.as<Stamp>() + s0,
                       (int)(s1 - s0));
    VDI_LAUNCHED();
    return VDI_OK;
}

int KernelTimer::harvest(int a, int b, hipStream_t st) {
    b = std::min(b, next);
    if (b <= a) return VDI_OK;
    const long s0 = off[a], s1 = off[b];
    VDI_HIP(hipMemcpyAsync(host + s0, stamps.as<Stamp>() + s0, sizeof(Stamp) * (size_t)(s1 - s0), hipMemcpyDeviceToHost,
                           st));
    VDI_HIP(hipStreamSynchronize(st));
    for (int k = a; k < b; k++) {
        unsigned long long t0 = ~0ull, t1 = 0;
        for (long i = off[k]; i < off[k + 1]; i++) {
            const Stamp &s = host[i];
            if (s.t0 == ~0ull || s.t1 < s.t0) continue;  // returned at once
            t0 = std::min(t0, s.t0);
            t1 = std::max(t1, s.t1);
        }
        if (t0 == ~0ull) continue;  // every workgroup returned at once
        dur_ms.push_back((float)((double)(t1 - t0) / ticks_per_ms));
        bytes.push_back(slot_bytes[k]);
    }
    return VDI_OK;
}

void KernelTimer::release() {
    stamps.release();
    if (host) (void)hipHostFree(host);
    host = nullptr;
    cap = next = reserved = 0;
    ncap = 0;
    slot_bytes.clear();
    off.clear();
    dur_ms.clear();
    bytes.clear();
    cls = -1;
}

// ------------------------------------------------------------------------
// Driver
// ------------------------------------------------------------------------



// Largest default shape that still gives every CU a tile.
static int default_tile(int m, int n) {
    const int prefs[] = {0, 4, 2, 1};
    for (int s : prefs) {
        const TileShape &t = kTileShapes[s];
        if ((long)cdiv(m, t.tx) * cdiv(n, t.ty) >= 256) return s;
    }
    return 1;
}

int MG::init(int kind_, int m, int n, double cx, double cy, int x_sym, double *z0, double *b0, hipStream_t stream,
             Arena *arena) {
    VDI_TRY(mgk::set_attributes());
    arena_ = arena;
    if (arena && (!z0 || !b0)) return VDI_ERR_ARG;
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
        {
            // a restricted level small enough for the multi-workgroup exact
            // solve is the coarsest one
            const int l = (int)lm.size() - 1, rows = fem ? ln.back() - 1 : ln.back();
            if (l >= 1 && (long)lm.back() * ln.back() <= kExactBig && std::max(lm.back(), ln.back()) <= kExactBigSide &&
                rows >= 1 && rows <= kSepMaxT)
                break;
        }
        if ((long)lm.back() * ln.back() <= kExactPoints) break;
        if (mc % 2 || nc % 2 || mc / 2 < 2 || nc / 2 < 2) break;
        mc /= 2;
        nc /= 2;
    }
    nl = (int)lm.size();
    lc = nl - 1;
    size_t coarse = 0;
    for (int l = 1; l < nl; l++) coarse += (size_t)lm[l] * ln[l];
    const long N0 = (long)m * n;
    // FP64: zalt on level 0; z, b, zalt on every coarser level
    VDI_TRY(store_.alloc((N0 + 3 * coarse) * sizeof(double), arena));
    // FP32 (mixed form): level 0 zalt and two r32 buffers; z, b, zalt below
    if (!arena) VDI_TRY(store_f_.alloc((3 * (size_t)N0 + 3 * coarse) * sizeof(float)));
    if (!z0) { VDI_TRY(own_z_.alloc(N0 * sizeof(double))); z0 = own_z_.as<double>(); }
    if (!b0) { VDI_TRY(own_b_.alloc(N0 * sizeof(double))); b0 = own_b_.as<double>(); }
    if (!arena) VDI_TRY(zprev_.alloc(N0 * sizeof(double)));
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
            if (q) {
                LF[l].zalt = q; q += Nl;
                r32_[0] = q; q += Nl;
                r32_[1] = q; q += Nl;
            }
            LF[l].b = r32_[0];
        } else {
            L[l].z = p; p += Nl;
            L[l].b = p; p += Nl;
            L[l].zalt = p; p += Nl;
            if (q) {
                LF[l].z = q; q += Nl;
                LF[l].b = q; q += Nl;
                LF[l].zalt = q; q += Nl;
            }
        }
    }
    {
        // the coarsest level: solved exactly when its shorter side (without
        // the nodal Dirichlet row) has at most kSepMaxT unknowns
        const int Mc = lm[lc], Nc = ln[lc], rows = fem ? Nc - 1 : Nc;
        fac_.reset();
        big_ = lc >= 1 && (long)Mc * Nc <= kExactBig && std::max(Mc, Nc) <= kExactBigSide && rows >= 1 &&
               rows <= kSepMaxT;
        sep_tdir_ = big_ || rows <= Mc ? 0 : 1;
        sep_T_ = sep_tdir_ == 0 ? rows : Mc;
        sep_P_ = sep_tdir_ == 0 ? Mc : rows;
        sep_t0_ = fem ? 1 : 0;
        if (!big_ && (sep_T_ > kSepMaxT || rows < 1)) sep_T_ = sep_P_ = 0;
        // LDS: z and b tiles and the reciprocal table, then the double workspace
        for (int pr = 0; pr < 2; pr++) {
            const size_t elem = pr ? sizeof(float) : sizeof(double);
            size_t bytes = (2 * (size_t)(Mc + 2) * (Nc + 2) + kInvTab) * elem;
            bytes = (bytes + 7) & ~(size_t)7;
            if (sep_T_ > 0) bytes += (sep::factor_doubles(sep_T_, sep_P_) + (size_t)sep_T_ * (sep_P_ + 1)) * sizeof(double);
            coarse_lds_[pr] = big_ ? 0 : bytes;
        }
        if (coarse_lds_[0] > (size_t)mgk::max_lds()) return VDI_ERR_ARG;
        if (big_ && std::max(x_lds(1), x_lds(2)) > (size_t)mgk::max_lds()) return VDI_ERR_ARG;
    }
    VDI_TRY(set_coeffs(cx, cy));
    for (int pr = 0; pr < 2; pr++)
        for (int w = 0; w < 2; w++) {
            tile[pr][w].assign(nl, 0);
            for (int l = 0; l < nl; l++) tile[pr][w][l] = default_tile(lm[l], ln[l]);
        }
    tile_mx = tile[1][1][0];
    // partial sums: one per tile of the smallest shape, two per begin block
    const TileShape &small = kTileShapes[1];
    long maxblk = std::max<long>(2048, (long)cdiv(lm[0], small.tx) * cdiv(ln[0], small.ty));
    VDI_TRY(partial_.alloc(2 * maxblk * sizeof(double), arena));
    VDI_TRY(state_.alloc(sizeof(MGState), arena));
    if (!host_state_ && !arena)
        VDI_HIP(hipHostMalloc((void **)&host_state_, sizeof(MGState), hipHostMallocDefault));
    return VDI_OK;
}

MG::~MG() {
    if (host_state_) (void)hipHostFree(host_state_);
}

void MG::level_coeffs(double cx, double cy) {
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

int MG::set_coeffs(double cx, double cy) {
    level_coeffs(cx, cy);
    return coarse_factors();
}

void MG::adopt_coeffs(double cx, double cy, const std::shared_ptr<CoarseFactors> &fac) {
    level_coeffs(cx, cy);
    if (kind == VDI_VX || kind == VDI_VY) {
        vfx = cx;
        vfy = cy;
    }
    fac_ = fac;
}

// Factors of the coarsest level's exact solve (coarse_sep.hpp), from the
// level's couplings and boundary types.  Checked once on the host: the solve
// must return a known solution of the level operator (Az_plain) to 1e-9.
int MG::coarse_factors() {
    if (sep_T_ == 0) return VDI_OK;
    const int l = lc, M = lm[l], Nn = ln[l];
    sep::Tri Pt, Rt, Sp, Qp;
    auto solve_host = [&](const std::vector<double> &f, const std::vector<double> &b, std::vector<double> &z) {
        if (big_) sep::solve2(sep_T_, sep_P_, f, b, z);
        else sep::solve(sep_T_, sep_P_, f, b, z);
    };
    if (fem) {
        if (!sep::fem_pencil(fe.fm, fe.fex, fe.fey, fe.fc, M, Nn, sep_tdir_, Pt, Rt, Sp, Qp)) return VDI_ERR_ARG;
    } else {
        sep::cc_pencil(cc.sigma, cc.s, cc.bcl, cc.bcr, cc.bcb, cc.bct, M, Nn, L[l].cx, L[l].cy, sep_tdir_, Pt, Rt, Sp,
                       Qp);
    }
    std::vector<double> f;
    if (!(big_ ? sep::factor2(Pt, Rt, Sp, Qp, f) : sep::factor(Pt, Rt, Sp, Qp, f))) return VDI_ERR_ARG;
    {
        const int T = sep_T_, P = sep_P_, y0 = sep_t0_;
        std::vector<double> zt((size_t)M * Nn, 0.0), bfull((size_t)M * Nn), bs((size_t)T * P), zs;
        for (int j = y0; j < Nn; j++)
            for (int i = 0; i < M; i++) zt[i + (size_t)M * j] = std::cos(0.37 * i + 1.3 * j) + 0.01 * ((i * 7 + j * 3) % 11);
        for (int j = 0; j < Nn; j++)
            for (int i = 0; i < M; i++)
                bfull[i + (size_t)M * j] = Az_plain<double>(fem, cc, fe, M, Nn, L[l].cx, L[l].cy, zt.data(), i, j);
        for (int t = 0; t < T; t++)
            for (int p = 0; p < P; p++) {
                const int i = sep_tdir_ == 0 ? p : t, j = (sep_tdir_ == 0 ? t : p) + y0;
                bs[p + (size_t)P * t] = bfull[i + (size_t)M * j];
            }
        solve_host(f, bs, zs);
        double e2 = 0, n2 = 0;
        for (int t = 0; t < T; t++)
            for (int p = 0; p < P; p++) {
                const int i = sep_tdir_ == 0 ? p : t, j = (sep_tdir_ == 0 ? t : p) + y0;
                const double d = zs[p + (size_t)P * t] - zt[i + (size_t)M * j];
                e2 += d * d;
                n2 += zt[i + (size_t)M * j] * zt[i + (size_t)M * j];
            }
        if (!(e2 <= 1e-18 * n2)) return VDI_ERR_ARG;
    }
    if (!fac_ || fac_.use_count() > 1) {
        // a captured step still holds the current buffer: write a new one
        auto fresh = std::make_shared<CoarseFactors>();
        VDI_TRY(fresh->f.alloc(f.size() * sizeof(double), arena_));
        fac_ = fresh;
    }
    VDI_HIP(hipMemcpy(fac_->f.p, f.data(), f.size() * sizeof(double), hipMemcpyHostToDevice));
    return VDI_OK;
}

MGCtl MG::ctl() const {
    MGCtl c;
    c.state = state_.as<MGState>();
    c.partial = partial_.as<double>();
    return c;
}


size_t MG::down_lds(int prec, int l) const {
    const TileShape &t = kTileShapes[tile[prec][0][l]];
    if (t.stream) return 0;
    const int nu = l == 0 ? nu1 : nu1s;
    return tile_lds_bytes(halo_cols(fem, nu, fem ? 2 : 1), halo_rows(fem, nu, 1), prec ? sizeof(float) : sizeof(double),
                          t.tx, t.ty);
}
size_t MG::up_lds(int prec, int l) const {
    const TileShape &t = kTileShapes[tile[prec][1][l]];
    if (t.stream) return 0;
    const int nu = l == 0 ? nu2 : nu2s, norm = l == 0;
    return tile_lds_bytes(halo_cols(fem, nu, norm), halo_rows(fem, nu, norm), prec ? sizeof(float) : sizeof(double),
                          t.tx, t.ty);
}
size_t MG::mx_lds() const {
    const TileShape &t = kTileShapes[tile_mx];
    if (t.stream) return 0;
    const int H = halo_cols(fem, nu2, 1), HY = halo_rows(fem, nu2, 1);
    return (size_t)(t.tx + 2 * H) * (t.ty + 2 * HY) * (sizeof(double) + 2 * sizeof(float)) + kInvTab * sizeof(float);
}

// Streaming launch of a level with nu sweeps (a.nu), segments of a.ty rows:
// strips of stream_wout columns (128 less the halo on both sides: 2 nu + 1
// rounded up to even for the cell-centred levels, 4 nu + 2 for the nodal
// ones), four wavefront items per workgroup.
static int stream_wout(bool fem, int nu) { return kSW - 2 * (fem ? 4 * nu + 2 : (2 * nu + 2) & ~1); }

template <class T>
static void stream_geometry(LvArgs<T> &a, bool fem) {
    a.nbx = cdiv(a.F.m, stream_wout(fem, a.nu));
    a.items = a.nbx * cdiv(a.F.n, a.ty);
    a.nb = cdiv(a.items, kSWaves);
}

template <class T>
LvArgs<T> MG::down_args(const std::vector<LevelT<T>> &H, int l, int zero_z) {
    const int prec = sizeof(T) == 4;
    LvArgs<T> a{};
    a.F = H[l];
    a.C = H[l + 1];
    a.cop = cc;
    a.fop = fe;
    a.nu = l == 0 ? nu1 : nu1s;
    a.flag = zero_z;
    const TileShape &t = kTileShapes[tile[prec][0][l]];
    a.tx = t.tx;
    a.ty = t.ty;
    if (t.stream) {
        stream_geometry(a, fem);
    } else if (fem) {
        a.nbx = cdiv(a.C.m, t.tx / 2);
        a.nb = a.nbx * cdiv(a.C.n, t.ty / 2);
    } else {
        a.nbx = cdiv(a.F.m, t.tx);
        a.nb = a.nbx * cdiv(a.F.n, t.ty);
    }
    a.ctl = ctl();
    a.tol2 = 0;
    a.ts = nullptr;
    return a;
}

template <class T>
LvArgs<T> MG::up_args(const std::vector<LevelT<T>> &H, int l, int norm, double tol2) {
    const int prec = sizeof(T) == 4;
    LvArgs<T> a{};
    a.F = H[l];
    a.C = H[l + 1];
    a.cop = cc;
    a.fop = fe;
    a.nu = l == 0 ? nu2 : nu2s;
    a.flag = norm;
    const TileShape &t = kTileShapes[tile[prec][1][l]];
    a.tx = t.tx;
    a.ty = t.ty;
    if (t.stream) {
        stream_geometry(a, fem);
    } else {
        a.nbx = cdiv(a.F.m, t.tx);
        a.nb = a.nbx * cdiv(a.F.n, t.ty);
    }
    a.ctl = ctl();
    a.tol2 = tol2;
    a.ts = nullptr;
    return a;
}

template <class T>
CoarseArgs<T> MG::coarse_args(const std::vector<LevelT<T>> &H, double tol2) {
    CoarseArgs<T> a{};
    CoarseDesc &d = a.d;
    const LevelT<T> &v = H[lc];
    d.m = v.m;
    d.n = v.n;
    d.cx = (double)v.cx;
    d.cy = (double)v.cy;
    d.inv = (double)v.inv_int;
    d.gz = reinterpret_cast<double *>(v.z);
    d.gb = reinterpret_cast<double *>(v.b);
    d.top = lc == 0;
    d.sep = fac_ ? 1 : 0;
    d.tdir = sep_tdir_;
    d.T = sep_T_;
    d.P = sep_P_;
    d.t0 = sep_t0_;
    const size_t tiles = (2 * (size_t)(v.m + 2) * (v.n + 2) + kInvTab) * sizeof(T);
    d.ws = (int)((tiles + 7) & ~(size_t)7);
    d.fac = fac_ ? fac_->f.as<double>() : nullptr;
    a.cop = cc;
    a.fop = fe;
    a.fem = fem;
    a.nu1 = nu1;
    a.nu2 = nu2;
    a.nu1s = nu1s;
    a.nu2s = nu2s;
    a.ncs = ncs;
    a.ctl = ctl();
    a.tol2 = tol2;
    a.ts = nullptr;
    return a;
}

// Mixed form: one FP32 V-cycle on A e = r32 from e = 0, then z += e and the
// FP64 residual (k_mx_up0).  z and r32 alternate between two buffers from
// one cycle to the next; parity = cycles enqueued so far mod 2.
LvArgs<float> MG::mx_down0_args(int parity) {
    LvArgs<float> a = down_args<float>(LF, 0, 1);
    a.F.b = r32_[parity & 1];
    return a;
}

MxArgs MG::mx_args(int parity, double tol2) {
    const int par = parity & 1;
    MxArgs a{};
    a.Z = L[0];
    a.z_in = par ? L[0].zalt : L[0].z;
    a.z_out = par ? L[0].z : L[0].zalt;
    a.E = LF[0];
    a.E.b = r32_[par];
    a.r_out = r32_[par ^ 1];
    a.C = LF[1];
    a.cop = cc;
    a.fop = fe;
    a.nu = nu2;
    const TileShape &t = kTileShapes[tile_mx];
    a.tx = t.tx;
    a.ty = t.ty;
    if (t.stream) {
        a.nbx = cdiv(L[0].m, stream_wout(fem, nu2));
        a.items = a.nbx * cdiv(L[0].n, t.ty);
        a.nb = cdiv(a.items, kSWaves);
    } else {
        a.nbx = cdiv(L[0].m, t.tx);
        a.nb = a.nbx * cdiv(L[0].n, t.ty);
    }
    a.ctl = ctl();
    a.tol2 = tol2;
    a.ts = nullptr;
    return a;
}

BeginArgs MG::begin_args(double tol2) {
    BeginArgs a{};
    a.fem = fem;
    a.L = L[0];
    a.cop = cc;
    a.fop = fe;
    a.ctl = ctl();
    a.tol2 = tol2;
    a.r32 = mixed_active() ? r32_[0] : nullptr;
    a.nb = (int)std::min<long>(1024, cdiv((long)L[0].m * L[0].n, 256));
    a.ts = nullptr;
    return a;
}

NormArgs MG::norm_args(int begin, double tol2) {
    NormArgs a{};
    a.ctl = ctl();
    a.begin = begin;
    a.tol2 = tol2;
    if (begin) a.nb = begin_args(tol2).nb;
    else a.nb = (int)up0_partials();
    return a;
}

// Partial sums of ||b - A z||^2 the level-0 up launch stores (one per
// workgroup).
long MG::up0_partials() const {
    if (lc <= 0) return 0;
    const TileShape &t = kTileShapes[mixed_active() ? tile_mx : tile[0][1][0]];
    if (t.stream) return cdiv((long)cdiv(lm[0], stream_wout(fem, nu2)) * cdiv(ln[0], t.ty), (long)kSWaves);
    return (long)cdiv(lm[0], t.tx) * cdiv(ln[0], t.ty);
}

VecArgs MG::vec_args(int op, const double *coef) {
    VecArgs a{};
    a.op = op;
    a.mixed = mixed_active();
    a.N = (long)L[0].m * L[0].n;
    a.z = L[0].z;
    a.zp = zprev_.as<double>();
    a.zalt = L[0].zalt;
    a.b = L[0].b;
    a.coef = coef;
    a.state = state_.as<MGState>();
    a.fm = fe.fm;
    a.M = L[0].m;
    const long work = op == VEC_DIRICHLET ? a.M : a.N;
    a.nb = (int)std::min<long>(2048, std::max<long>(1, cdiv(work, 256)));
    return a;
}

double MG::down_bytes(int prec, int l, int zero_z) const {
    const double e = prec ? 4.0 : 8.0;
    return e * ((zero_z ? 2.0 : 3.0) * lm[l] * ln[l] + 1.0 * lm[l + 1] * ln[l + 1]);
}
double MG::up_bytes(int prec, int l) const {
    const double e = prec ? 4.0 : 8.0;
    return e * (3.0 * lm[l] * ln[l] + 1.0 * lm[l + 1] * ln[l + 1]);
}
// reads b (and z on a top level without the exact solve), writes z; the
// exact solve also reads its FP64 factors
double MG::coarse_bytes(int prec) const {
    const double e = prec ? 4.0 : 8.0;
    if (big_) {
        // three launches: b -> w, w -> w, w -> z, and the factors once each
        return 3.0 * (2.0 * e * lm[lc] * ln[lc]) + 8.0 * (double)sep::factor2_doubles(sep_T_, sep_P_);
    }
    const double f = sep_T_ > 0 ? 8.0 * (double)sep::factor_doubles(sep_T_, sep_P_) : 0.0;
    return e * (lc == 0 && sep_T_ == 0 ? 3.0 : 2.0) * lm[lc] * ln[lc] + f;
}

template <class T>
XArgs<T> MG::x_args(const std::vector<LevelT<T>> &H, int phase, int quad) {
    XArgs<T> a{};
    a.quad = quad;
    const LevelT<T> &v = H[lc];
    a.phase = phase;
    a.m = v.m;
    a.n = v.n;
    a.yoff = sep_t0_;
    a.Tn = sep_T_;
    a.P = sep_P_;
    const int R = quad ? kXRowsQ : kXRows;
    a.nbr = (sep_T_ + R - 1) / R;
    a.nbc = (sep_P_ + kXCols - 1) / kXCols;
    a.b = v.b;
    a.w = v.zalt;
    a.z = v.z;
    a.fac = fac_ ? fac_->f.as<double>() : nullptr;
    a.fop = fe;
    a.ctl = ctl();
    a.ts = nullptr;
    return a;
}

size_t MG::x_lds(int phase) const {
    if (phase == 2) return ((size_t)sep_T_ * sep_T_ + 2 * (size_t)sep_T_ * kXColsV) * sizeof(double);
    return ((size_t)kXRows * sep_P_ + (size_t)sep_P_ * kXCols) * sizeof(double);
}
// reads e (zalt), r32, coarse e, z64, b64; writes z64, r32
double MG::mx_bytes() const {
    const double N0 = (double)lm[0] * ln[0];
    return 4.0 * (3.0 * N0 + (double)lm[1] * ln[1]) + 8.0 * 3.0 * N0;
}

// Kernel-function class of an FP64 level launch (VDI_KCLASS_*).
static int func_class(bool fem, bool up, bool stream) {
    if (stream) return fem ? (up ? VDI_KCLASS_SFEM_UP : VDI_KCLASS_SFEM_DOWN) : (up ? VDI_KCLASS_SCC_UP : VDI_KCLASS_SCC_DOWN);
    return fem ? (up ? VDI_KCLASS_FEM_UP : VDI_KCLASS_FEM_DOWN) : (up ? VDI_KCLASS_CC_UP : VDI_KCLASS_CC_DOWN);
}

// Enqueues the down kernel of level l of hierarchy H (FP64 or FP32).
template <class T>
int MG::launch_down(const std::vector<LevelT<T>> &H, int l, int zero_z, int fin, double tol2, int fin_begin) {
    const int prec = sizeof(T) == 4;
    LvArgs<T> a = down_args<T>(H, l, zero_z);
    a.fin = fin;
    a.fin_begin = fin_begin;
    a.tol2 = tol2;
    if (timer)
        a.ts = timer->slot(l == 0 ? VDI_KCLASS_MG_DOWN0 : VDI_KCLASS_MG_MID, down_bytes(prec, l, zero_z),
                           prec ? -1 : func_class(fem, false, kTileShapes[tile[prec][0][l]].stream), a.nb);
    return mgk::down<T>(a, fem, down_lds(prec, l), st);
}

template <class T>
int MG::launch_up(const std::vector<LevelT<T>> &H, int l, int norm, double tol2) {
    const int prec = sizeof(T) == 4;
    LvArgs<T> a = up_args<T>(H, l, norm, tol2);
    if (timer)
        a.ts = timer->slot(l == 0 ? VDI_KCLASS_MG_UP0 : VDI_KCLASS_MG_MID, up_bytes(prec, l),
                           prec ? -1 : func_class(fem, true, kTileShapes[tile[prec][1][l]].stream), a.nb);
    return mgk::up<T>(a, fem, up_lds(prec, l), st);
}

template <class T>
int MG::launch_coarse(const std::vector<LevelT<T>> &H, double tol2) {
    const int prec = sizeof(T) == 4;
    if (big_) {
        for (int ph = 1; ph <= 3; ph++) {
            XArgs<T> x = x_args<T>(H, ph);
            if (timer)
                x.ts = timer->slot(prec ? VDI_KCLASS_MG_COARSE : VDI_KCLASS_XSOLVE, coarse_bytes(prec) / 3.0, -1,
                                   mgk::xsolve_grid(x));
            VDI_TRY(mgk::xsolve<T>(x, x_lds(ph), st));
        }
        return VDI_OK;
    }
    CoarseArgs<T> a = coarse_args<T>(H, tol2);
    if (timer) a.ts = timer->slot(VDI_KCLASS_MG_COARSE, coarse_bytes(prec));
    return mgk::coarse<T>(a, coarse_lds(prec), st);
}

int MG::launch_vec(const VecArgs &a) { return mgk::vec(a, st); }

bool MG::mixed_active() const { return form == VDI_FORM_MIXED && lc > 0; }

int MG::launch_cycle(double tol2) { return launch_cycle(tol2, 0, 1); }

bool MG::can_fold_norm() const {
    if (lc <= 0) return false;
    return up0_partials() <= kFinMax;
}

int MG::begin_folded(double tol) {
    if (!can_fold_norm()) return begin(tol);
    VDI_TRY(enforce_dirichlet());
    mx_enqueued_ = 0;
    BeginArgs a = begin_args(tol * tol);
    a.reset = 1;
    return mgk::begin(a, st);
}

int MG::launch_cycle(double tol2, int fin_prev, int norm) {
    // partials the level-0 down kernel finishes: the previous cycle's
    // (fin_prev = 1) or k_begin's pairs (fin_prev = 2)
    const int fin = !fin_prev || !can_fold_norm() ? 0 : (fin_prev == 2 ? norm_args(1, tol2).nb : norm_args(0, tol2).nb);
    const int fin_begin = fin_prev == 2 && fin > 0;
    if (!mixed_active()) {
        // coarser levels hold corrections, which start from zero
        for (int l = 0; l < lc; l++) VDI_TRY(launch_down<double>(L, l, l > 0, l == 0 ? fin : 0, tol2, fin_begin));
        VDI_TRY(launch_coarse<double>(L, tol2));
        for (int l = lc - 1; l >= 0; l--) VDI_TRY(launch_up<double>(L, l, l == 0, tol2));
        // the level-0 norm is reduced by its own one-workgroup kernel, or by
        // the next cycle's level-0 down kernel (the coarse kernel reduces it
        // itself when it owns level 0)
        if (lc > 0 && (norm || !can_fold_norm())) VDI_TRY(mgk::norm(norm_args(0, tol2), st));
        return VDI_OK;
    }
    const int par = mx_enqueued_ & 1;
    {
        LvArgs<float> a = mx_down0_args(par);
        a.fin = fin;
        a.fin_begin = fin_begin;
        a.tol2 = tol2;
        if (timer) a.ts = timer->slot(VDI_KCLASS_MG_DOWN0, down_bytes(1, 0, 1), -1, a.nb);
        VDI_TRY(mgk::down<float>(a, fem, down_lds(1, 0), st));
    }
    for (int l = 1; l < lc; l++) VDI_TRY(launch_down<float>(LF, l, 1));
    VDI_TRY(launch_coarse<float>(LF, tol2));
    for (int l = lc - 1; l >= 1; l--) VDI_TRY(launch_up<float>(LF, l, 0, tol2));
    MxArgs a = mx_args(par, tol2);
    if (timer) a.ts = timer->slot(VDI_KCLASS_MG_UP0, mx_bytes(), -1, a.nb);
    VDI_TRY(mgk::mx(a, fem, mx_lds(), st));
    if (norm || !can_fold_norm()) VDI_TRY(mgk::norm(norm_args(0, tol2), st));
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
    double b = 0;
    if (!mixed_active()) {
        for (int l = 0; l < lc; l++) b += down_bytes(0, l, l > 0) + up_bytes(0, l);
        b += coarse_bytes(0);
    } else {
        for (int l = 0; l < lc; l++) b += down_bytes(1, l, 1);
        for (int l = 1; l < lc; l++) b += up_bytes(1, l);
        b += coarse_bytes(1) + mx_bytes();
    }
    return b;
}

int MG::enforce_dirichlet() {
    if (!fem) return VDI_OK;
    return launch_vec(vec_args(VEC_DIRICHLET, nullptr));
}

int MG::begin(double tol) {
    VDI_TRY(enforce_dirichlet());
    mx_enqueued_ = 0;
    VDI_TRY(mgk::begin(begin_args(tol * tol), st));
    return mgk::norm(norm_args(1, tol * tol), st);
}

int MG::clear_solve_state() {
    hipLaunchKernelGGL(k_zero, dim3(cdiv(sizeof(MGState) / sizeof(double), 256)), dim3(256), 0, st,
                       state_.as<double>(), (long)(sizeof(MGState) / sizeof(double)));
    VDI_LAUNCHED();
    const long np = (long)(partial_.bytes / sizeof(double));
    hipLaunchKernelGGL(k_zero, dim3(cdiv(np, 256)), dim3(256), 0, st, partial_.as<double>(), np);
    VDI_LAUNCHED();
    return VDI_OK;
}

int MG::reset_history() {
    scount = 0;
    t0 = t_prev = 0;
    last_cycles = 0;
    const long N = unknowns();
    hipLaunchKernelGGL(k_zero, dim3(cdiv(N, 256)), dim3(256), 0, st, L[0].z, N);
    VDI_LAUNCHED();
    if (zprev_.p) {
        hipLaunchKernelGGL(k_zero, dim3(cdiv(N, 256)), dim3(256), 0, st, zprev_.as<double>(), N);
        VDI_LAUNCHED();
    }
    return clear_solve_state();
}

int MG::read_state(MGState *out) {
    if (!host_state_) {
        // arena mode: no pinned buffer, a pageable copy
        VDI_HIP(hipMemcpyAsync(out, state_.p, sizeof(MGState), hipMemcpyDeviceToHost, st));
        VDI_HIP(hipStreamSynchronize(st));
        return VDI_OK;
    }
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
    VDI_TRY(begin_folded(tol));
    const int pred = std::max(1, std::min(max_cycles, last_cycles));
    // no host read inside this run: the start's norms and each cycle's are
    // finished by the next level-0 down kernel
    for (int c = 0; c < pred; c++) VDI_TRY(launch_cycle(tol2, c > 0 ? 1 : 2, c == pred - 1));
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
    const long N = unknowns();
    const int was = scount;
    double bb = 1, c = 0;
    const int r = guess_coeffs(t, &bb, &c);
    if (r == 0 && was == 1)
        VDI_HIP(hipMemcpyAsync(zprev_.p, L[0].z, N * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (r != 1) return VDI_OK;
    hipLaunchKernelGGL(k_extrapolate, dim3(cdiv(N, 256)), dim3(256), 0, st, L[0].z, zprev_.as<double>(), N, bb, c);
    VDI_LAUNCHED();
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

int MG::enqueue_extrapolate(const double *coef) { return launch_vec(vec_args(VEC_EXTRAP, coef)); }

int MG::enqueue_fixed(double tol, int budget) {
    // the start's norms and each cycle's are finished by the next level-0
    // down kernel, the last cycle's by its own launch
    if (budget < 1) {
        VDI_TRY(begin(tol));
        return launch_vec(vec_args(VEC_SOLVE_END, nullptr));
    }
    VDI_TRY(begin_folded(tol));
    for (int c = 0; c < budget; c++) VDI_TRY(launch_cycle(tol * tol, c > 0 ? 1 : 2, c == budget - 1));
    // FP64 with folded norms: the first level-0 down kernel wrote z = 0 for a
    // zero right-hand side, and there is no odd-cycle buffer to bring back
    if (!mixed_active() && can_fold_norm()) return VDI_OK;
    return launch_vec(vec_args(VEC_SOLVE_END, nullptr));
}

int MG::residual_out(const double *z, double *r, double *part, int nb) {
    hipLaunchKernelGGL(k_resid_out, dim3(nb), dim3(256), 0, st, (int)fem, L[0], cc, fe, z, r, part);
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

int MG::set_tile(int prec, int which, int level, int shape) {
    if (prec < 0 || prec > 1 || which < 0 || which > 2 || shape < 0 || shape >= kNumTileShapes) return VDI_ERR_ARG;
    // streaming launches: on levels whose arrays a buffer resource's 32-bit
    // byte range covers (the mixed form's level-0 up reads FP64 level 0)
    if (kTileShapes[shape].stream) {
        const int elem = (which == 2 || !prec) ? 8 : 4;
        if (level >= 0 && level < lc && (long)lm[level] * ln[level] * elem >= (1L << 31) - (1L << 20))
            return VDI_ERR_ARG;
    }
    if (which == 2) {
        if (level != 0) return VDI_ERR_ARG;
        const int old = tile_mx;
        tile_mx = shape;
        if (mx_lds() > (size_t)mgk::max_lds()) { tile_mx = old; return VDI_ERR_ARG; }
        return VDI_OK;
    }
    if (level < 0 || level >= lc) return VDI_ERR_ARG;
    const int old = tile[prec][which][level];
    tile[prec][which][level] = shape;
    const size_t need = which == 0 ? down_lds(prec, level) : up_lds(prec, level);
    if (need > (size_t)mgk::max_lds()) { tile[prec][which][level] = old; return VDI_ERR_ARG; }
    return VDI_OK;
}

int MG::set_schedule(int n1, int n2, int n1s, int n2s) {
    if (n1 < 0 || n2 < 0 || n1 + n2 == 0 || n1 > kMaxNu || n2 > kMaxNu) return VDI_ERR_ARG;
    if (n1s < 0 || n2s < 0 || n1s + n2s == 0 || n1s > kMaxNu || n2s > kMaxNu) return VDI_ERR_ARG;
    // applied to the settings only when every tile fits: a rejected schedule
    // leaves sweeps and tile shapes as they were
    const int keep[4] = {nu1, nu2, nu1s, nu2s};
    std::vector<int> keep_tile[2][2];
    for (int pr = 0; pr < 2; pr++)
        for (int w = 0; w < 2; w++) keep_tile[pr][w] = tile[pr][w];
    const int keep_mx = tile_mx;
    const int rc = apply_schedule(n1, n2, n1s, n2s);
    if (rc != VDI_OK) {
        nu1 = keep[0];
        nu2 = keep[1];
        nu1s = keep[2];
        nu2s = keep[3];
        for (int pr = 0; pr < 2; pr++)
            for (int w = 0; w < 2; w++) tile[pr][w] = keep_tile[pr][w];
        tile_mx = keep_mx;
    }
    return rc;
}

int MG::apply_schedule(int n1, int n2, int n1s, int n2s) {
    nu1 = n1;
    nu2 = n2;
    nu1s = n1s;
    nu2s = n2s;
    const int fallback[] = {0, 4, 2, 1};
    for (int prec = 0; prec < 2; prec++)
        for (int which = 0; which < 2; which++)
            for (int l = 0; l < (int)tile[prec][which].size(); l++) {
                auto need = [&]() { return which == 0 ? down_lds(prec, l) : up_lds(prec, l); };
                if (need() <= (size_t)mgk::max_lds()) continue;
                for (int s : fallback) {
                    tile[prec][which][l] = s;
                    if (need() <= (size_t)mgk::max_lds()) break;
                }
                if (need() > (size_t)mgk::max_lds()) return VDI_ERR_ARG;
            }
    if (mx_lds() > (size_t)mgk::max_lds()) {
        for (int s : fallback) {
            tile_mx = s;
            if (mx_lds() <= (size_t)mgk::max_lds()) break;
        }
        if (mx_lds() > (size_t)mgk::max_lds()) return VDI_ERR_ARG;
    }
    return VDI_OK;
}

// Times reps launches of one kernel (after one untimed launch) with events.
template <class F>
static int time_launches(hipStream_t st, int reps, F &&launch, float *ms) {
    hipEvent_t e0, e1;
    VDI_HIP(hipEventCreate(&e0));
    VDI_HIP(hipEventCreate(&e1));
    int rc = launch();
    if (rc == VDI_OK) {
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < reps && rc == VDI_OK; r++) rc = launch();
        (void)hipEventRecord(e1, st);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(ms, e0, e1);
        *ms /= reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipGetLastError();
    return rc;
}

int MG::autotune(int reps) {
    if (reps < 1) reps = 1;
    const long N0 = (long)lm[0] * ln[0];
    DevBuf save;
    VDI_TRY(save.alloc(2 * N0 * sizeof(double)));
    VDI_HIP(hipMemcpyAsync(save.p, L[0].z, N0 * sizeof(double), hipMemcpyDeviceToDevice, st));
    VDI_HIP(hipMemcpyAsync(save.as<double>() + N0, L[0].b, N0 * sizeof(double), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_reset_state, dim3(1), dim3(1), 0, st, state_.as<MGState>());
    VDI_LAUNCHED();
    KernelTimer *tsave = timer;
    timer = nullptr;
    int rc = VDI_OK;
    for (int prec = 0; prec < 2 && rc == VDI_OK; prec++) {
        for (int which = 0; which < 3 && rc == VDI_OK; which++) {
            for (int l = 0; l < lc && rc == VDI_OK; l++) {
                if (which == 2 && (prec == 0 || l > 0)) continue;
                if (which == 1 && prec == 1 && l == 0) continue;  // the mixed form's level-0 up is k_mx_up0
                int best = -1;
                float best_ms = 0;
                for (int s = 0; s < kNumTileShapes && rc == VDI_OK; s++) {
                    if (set_tile(prec, which, l, s) != VDI_OK) continue;
                    float ms = 0;
                    auto launch = [&]() -> int {
                        if (which == 2) return mgk::mx(mx_args(0, -1.0), fem, mx_lds(), st);
                        if (prec == 0) {
                            if (which == 0) return mgk::down<double>(down_args<double>(L, l, l > 0), fem, down_lds(0, l), st);
                            return mgk::up<double>(up_args<double>(L, l, l == 0, -1.0), fem, up_lds(0, l), st);
                        }
                        if (which == 0) {
                            LvArgs<float> a = l == 0 ? mx_down0_args(0) : down_args<float>(LF, l, 1);
                            return mgk::down<float>(a, fem, down_lds(1, l), st);
                        }
                        return mgk::up<float>(up_args<float>(LF, l, 0, -1.0), fem, up_lds(1, l), st);
                    };
                    rc = time_launches(st, reps, launch, &ms);
                    if (rc == VDI_OK && (best < 0 || ms < best_ms)) { best = s; best_ms = ms; }
                }
                if (best >= 0) (void)set_tile(prec, which, l, best);
            }
        }
    }
    timer = tsave;
    VDI_HIP(hipMemcpyAsync(L[0].z, save.p, N0 * sizeof(double), hipMemcpyDeviceToDevice, st));
    VDI_HIP(hipMemcpyAsync(L[0].b, save.as<double>() + N0, N0 * sizeof(double), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_reset_state, dim3(1), dim3(1), 0, st, state_.as<MGState>());
    VDI_LAUNCHED();
    VDI_HIP(hipStreamSynchronize(st));
    return rc;
}

template LvArgs<double> MG::down_args<double>(const std::vector<Level> &, int, int);
template LvArgs<float> MG::down_args<float>(const std::vector<LevelF> &, int, int);
template LvArgs<double> MG::up_args<double>(const std::vector<Level> &, int, int, double);
template LvArgs<float> MG::up_args<float>(const std::vector<LevelF> &, int, int, double);
template CoarseArgs<double> MG::coarse_args<double>(const std::vector<Level> &, double);
template XArgs<double> MG::x_args<double>(const std::vector<Level> &, int, int);
template XArgs<float> MG::x_args<float>(const std::vector<LevelF> &, int, int);
template CoarseArgs<float> MG::coarse_args<float>(const std::vector<LevelF> &, double);

}  // namespace vdi
