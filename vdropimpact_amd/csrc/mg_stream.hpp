// mg_stream.hpp -- streaming level kernels for large levels: cell-centred
// (k_sdown, k_sup), nodal (k_sndown, k_snup) and the mixed form's level-0 up
// launch (k_smx).
//
// Included by mg.hip after the tiled level kernels, whose operator helpers
// (cc_diag, cc_prolong_combine, fem_off_v, fem_prolong_combine) and norm
// bookkeeping (fin_load, fin_norm, store_partial) these kernels share, so
// that they compute the tiled kernels' bits: the same Gauss-Seidel sweeps,
// residual, restriction and prolongation as down_body / up_body / mx_body
// (whose FP64 residual, mx_az, k_smx restates in the same order), with every
// rounding in the same order.
//
// Layout.  One wavefront owns a strip of kSW = 128 columns -- lane k holds
// the even column c0 + 2k and the odd column c0 + 2k + 1 -- and a segment of
// S rows, and walks up the rows of its strip.  The rows it is working on sit
// in a rolling window of registers: z and b of both columns for the last
// P + 4 rows and kSAhead rows loaded ahead.  A point's vertical neighbours
// are in the lane's own registers; of its two horizontal neighbours one is
// the lane's other column and the other comes from the next lane with one
// DPP whole-wave shift (wave_shr:1 / wave_shl:1).  No LDS, no barriers.
//
// Schedule.  When row y has been loaded, pass p (p = 1..P, colour (p-1)&1)
// runs on row y - p, the passes of a step in order.  A point of row r in
// pass p then reads row r+1 after pass p-1 (run earlier in the same step)
// and not yet pass p+1, and row r-1 after pass p-1 and before pass p+1 --
// exactly the values a global sweep gives.  The residual, the stores and
// (down) the restriction run on row y - P - 2.  (The code also supports a
// lag of 2, pass p on row y - 2p, whose P updates per step are independent;
// its P more steps per wavefront cost more than the chained updates.)
// Each pass shrinks the region whose values equal the global sweep's by one
// row and column at each end, so a strip carries a halo of P + 2 columns
// (even, so that the restriction's 2 x 2 blocks lie within one lane) and a
// segment a halo of P + 1 rows; the halos are recomputed by the neighbouring
// wavefronts.  Points outside the level are zero and never updated, as in
// the tiled kernels.
#pragma once

#include <utility>

namespace vdi {

constexpr int kSW = 128;       // columns of a strip (two per lane)
constexpr int kSWaves = 4;     // independent wavefronts per workgroup
constexpr int kSThreads = 64 * kSWaves;
constexpr int kSAhead = 4;     // rows of z and b loaded ahead of the step's row
// nodal up launches load two rows ahead: with four they need more than 256
// registers, one wavefront per SIMD
// (P > 0: the coarse ring of a window needs four rows)
template <bool UP, int P> constexpr int kSNAhead = (UP && P > 0) ? 2 : kSAhead;
constexpr int kSQAhead = 2;    // fine rows by which a coarse row (up launches) is loaded ahead

// Lane k receives lane k-1's value (lane 0: zero) / lane k+1's (lane 63: zero).
__device__ __forceinline__ int lane_from_left_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x138 /* wave_shr:1 */, 0xf, 0xf, true);
}
__device__ __forceinline__ int lane_from_right_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x130 /* wave_shl:1 */, 0xf, 0xf, true);
}
__device__ __forceinline__ float lane_from_left(float v) { return __int_as_float(lane_from_left_i(__float_as_int(v))); }
__device__ __forceinline__ float lane_from_right(float v) { return __int_as_float(lane_from_right_i(__float_as_int(v))); }
__device__ __forceinline__ double lane_from_left(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = lane_from_left_i((int)b), hi = lane_from_left_i((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double lane_from_right(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = lane_from_right_i((int)b), hi = lane_from_right_i((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Rolling window slot of the row k rows below the step's row, in a step
// whose position in the unrolled window is U.
template <int W>
__device__ __forceinline__ constexpr int sslot(int U, int k) {
    return ((U - k) % W + W) % W;
}

// Per-wavefront state of a streaming launch (level F, lane columns xa, xb).
template <class T>
struct SLane {
    int xa, xb;          // the lane's even and odd column (may lie outside the level)
    int ca, cb;          // the same clamped into the level (unconditional loads)
    bool ina, inb;       // inside the level
    bool outa, outb;     // inside the strip's output columns
    // reciprocal diagonals and diagonals (residual) of the lane's columns
    // for an interior (0), the first (1) and the last (2) row; scalars, so
    // that a per-row choice stays a register select
    T ia0, ia1, ia2, ib0, ib1, ib2;
    T da0, da1, da2, db0, db1, db2;
    // up launches: coarse column of both columns, the coarse columns across
    // x of the even / odd one (clamped) and their ghost factors
    int ci, cia, cib;
    T fxa, fxb;
    T fyb, fyt;  // ghost factors of a coarse row below / above the coarse level
};

template <class T>
__device__ __forceinline__ void slane_init(const LevelT<T> &F, const LevelT<T> &C, const CCOp &op, int x0, int wout,
                                           int hh, SLane<T> &s) {
    const int lane = threadIdx.x & 63;
    s.xa = x0 - hh + 2 * lane;
    s.xb = s.xa + 1;
    s.ina = s.xa >= 0 && s.xa < F.m;
    s.inb = s.xb >= 0 && s.xb < F.m;
    s.outa = s.ina && s.xa >= x0 && s.xa < x0 + wout;
    s.outb = s.inb && s.xb >= x0 && s.xb < x0 + wout;
    s.ca = min(max(s.xa, 0), F.m - 1);
    s.cb = min(max(s.xb, 0), F.m - 1);
    // a row of each class (cc_inv_table's representatives); the interior
    // point's reciprocal is the level's stored one
    const int j0 = min(1, F.n - 1), j2 = F.n - 1;
    s.da0 = cc_diag<T>(op, F.m, F.n, F.cx, F.cy, s.ca, j0);
    s.da1 = cc_diag<T>(op, F.m, F.n, F.cx, F.cy, s.ca, 0);
    s.da2 = cc_diag<T>(op, F.m, F.n, F.cx, F.cy, s.ca, j2);
    s.db0 = cc_diag<T>(op, F.m, F.n, F.cx, F.cy, s.cb, j0);
    s.db1 = cc_diag<T>(op, F.m, F.n, F.cx, F.cy, s.cb, 0);
    s.db2 = cc_diag<T>(op, F.m, F.n, F.cx, F.cy, s.cb, j2);
    s.ia0 = (s.ca > 0 && s.ca < F.m - 1) ? F.inv_int : T(1) / s.da0;
    s.ib0 = (s.cb > 0 && s.cb < F.m - 1) ? F.inv_int : T(1) / s.db0;
    s.ia1 = T(1) / s.da1;
    s.ia2 = T(1) / s.da2;
    s.ib1 = T(1) / s.db1;
    s.ib2 = T(1) / s.db2;
    // cc_prolong_at's coarse columns: for an even fine column 2I the
    // neighbour across x is I - 1, for the odd one I + 1, each replaced by I
    // with the side's ghost factor outside the coarse level
    s.ci = min(s.ca >> 1, max(C.m - 1, 0));
    s.cia = s.ci - 1;
    s.cib = s.ci + 1;
    s.fxa = T(1);
    s.fxb = T(1);
    if (s.cia < 0) { s.fxa = T(bc_ghost_factor(op.bcl)); s.cia = s.ci; }
    if (s.cib >= C.m) { s.fxb = T(bc_ghost_factor(op.bcr)); s.cib = s.ci; }
    // kept in vector registers: as scalars they are reloaded from the kernel
    // arguments (and waited for) at every step
    s.fyb = T(bc_ghost_factor(op.bcb));
    s.fyt = T(bc_ghost_factor(op.bct));
    asm volatile("" : "+v"(s.fyb), "+v"(s.fyt));
}

// Row class: 0 interior, 1 first, 2 last row.  INT: the row is known to be
// an interior one.
template <bool INT, class T>
__device__ __forceinline__ T by_row_class(T v0, T v1, T v2, int r, int n) {
    if constexpr (INT) return v0;
    else return r == 0 ? v1 : (r == n - 1 ? v2 : v0);
}

// Buffer access with range checking: a load at an offset outside the
// buffer returns 0 and a store there is dropped (MUBUF out-of-range
// behaviour), so points outside the level or the strip need no branch --
// the level's zero padding and the output guards are offsets.  Every array
// a streaming launch touches is below 2 GiB (MG::set_tile checks).
using SRsrc = __amdgpu_buffer_rsrc_t;
constexpr int kSOut = (int)0x80000000u;  // an offset outside every buffer

__device__ __forceinline__ SRsrc srsrc(const void *p, long bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
typedef int sint2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double sload(double, SRsrc r, int off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ float sload(float, SRsrc r, int off) {
    return __int_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ void sstore(SRsrc r, int off, double v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(sint2, v), r, off, 0, 0);
}
__device__ __forceinline__ void sstore(SRsrc r, int off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_int(v), r, off, 0, 0);
}

// Buffers of one launch: the level's z source (z for down, zalt for up; a
// zero-length buffer when the level starts from zero), b, the output (zalt
// for down, z for up) and the coarse level's array (b for down, z for up).
struct SBufs {
    SRsrc zsrc, b, dst, coarse;
};

// The rolling window: z and b of both columns per row slot and, for up
// launches, a ring of W / 2 coarse rows (the three coarse values a lane's
// prolongation reads from each: its own coarse column and the neighbours
// across x of its even and odd column).  A coarse row serves four fine rows
// and is loaded once.
template <class T, int W>
struct SWin {
    T za[W], zb[W], ba[W], bb[W];
    T qc[W / 2], qa[W / 2], qb[W / 2];
};

// Coarse rows of cc_prolong_at for fine row y (inside the level): J and the
// neighbour across y, J2, with its ghost factor (fyb / fyt below / above the
// coarse level).
template <class T>
__device__ __forceinline__ void sprolong_rows(const LevelT<T> &C, T fyb, T fyt, int y, int &J, int &J2, T &fy) {
    J = y >> 1;
    J2 = (y & 1) ? J + 1 : J - 1;
    const bool lo = J2 < 0, hi = J2 >= C.n;
    fy = lo ? fyb : (hi ? fyt : T(1));
    J2 = (lo || hi) ? J : J2;
}

// Stage 1 of a row, kSAhead steps before it is needed: the loads of z and
// b of row y into slot K, each returning zero for a point outside the level
// or the segment's rows.  No value is used here, so the loads stay in
// flight.
template <class T, int W, int K, class LN>
__device__ __forceinline__ void srow_issue(const LevelT<T> &F, const SBufs &bf, const LN &s, int y, int ylast,
                                           SWin<T, W> &w) {
    const bool row = y >= 0 && y < F.n && y <= ylast;
    const int base = F.m * y;
    const int oa = (row && s.ina) ? (int)sizeof(T) * (base + s.xa) : kSOut;
    const int ob = (row && s.inb) ? (int)sizeof(T) * (base + s.xb) : kSOut;
    w.za[K] = sload(T(), bf.zsrc, oa);
    w.zb[K] = sload(T(), bf.zsrc, ob);
    w.ba[K] = sload(T(), bf.b, oa);
    w.bb[K] = sload(T(), bf.b, ob);
}

// Up launches: the three coarse values of coarse row J (clamped into the
// coarse level) into ring slot K.
template <class T, int W, int K>
__device__ __forceinline__ void sq_issue(const LevelT<T> &C, const SBufs &bf, const SLane<T> &s, int J,
                                         SWin<T, W> &w) {
    const int r = C.m * min(max(J, 0), C.n - 1);
    const bool in = s.ina || s.inb;
    w.qc[K] = sload(T(), bf.coarse, in ? (int)sizeof(T) * (r + s.ci) : kSOut);
    w.qa[K] = sload(T(), bf.coarse, in ? (int)sizeof(T) * (r + s.cia) : kSOut);
    w.qb[K] = sload(T(), bf.coarse, in ? (int)sizeof(T) * (r + s.cib) : kSOut);
}

// Stage 2 (up launches), when row y reaches the step front: add the
// prolongated correction from ring slots KJ (coarse row J = y >> 1) and KJ2
// (J2 = J -+ 1; cc_prolong_at's ghost rule when J2 lies outside the coarse
// level), and zero what lies outside the level or the segment's rows
// (tile_load's z += cc_prolong_at, zero outside).
template <class T, int W, int K, int KJ, int KJ2>
__device__ __forceinline__ void srow_fix(const LevelT<T> &F, const LevelT<T> &C, const CCOp &op, const SLane<T> &s,
                                         int y, int ylast, SWin<T, W> &w) {
    const bool row = y >= 0 && y < F.n && y <= ylast;
    int J, J2;
    T fy;
    sprolong_rows<T>(C, s.fyb, s.fyt, min(max(y, 0), F.n - 1), J, J2, fy);
    const bool same = J2 == J;  // J2 outside the coarse level: row J with the ghost factor
    const T c1 = same ? w.qc[KJ] : w.qc[KJ2], a1 = same ? w.qa[KJ] : w.qa[KJ2], b1 = same ? w.qb[KJ] : w.qb[KJ2];
    const T va = w.za[K] + cc_prolong_combine<T>(s.fxa, fy, w.qc[KJ], w.qa[KJ], c1, a1);
    const T vb = w.zb[K] + cc_prolong_combine<T>(s.fxb, fy, w.qc[KJ], w.qb[KJ], c1, b1);
    w.za[K] = (row && s.ina) ? va : T(0);
    w.zb[K] = (row && s.inb) ? vb : T(0);
    // keep the sum here: sunk to the row's first use, it would keep coarse
    // loads in flight (and their registers live) for longer
    asm volatile("" : "+v"(w.za[K]), "+v"(w.zb[K]));
}

// Off-diagonal part of the cell-centred operator at the lane's even (A) or
// odd (B) column of the row in slot c, rows below / above in slots d / u:
// cc_off_t's rounding.
template <class T>
__device__ __forceinline__ T soff_a(T cx, T cy, const T *za, const T *zb, int c, int d, int u) {
#pragma clang fp contract(off)
    return fma(cx, lane_from_left(zb[c]) + zb[c], cy * (za[d] + za[u]));
}
template <class T>
__device__ __forceinline__ T soff_b(T cx, T cy, const T *za, const T *zb, int c, int d, int u) {
#pragma clang fp contract(off)
    return fma(cx, za[c] + lane_from_right(za[c]), cy * (zb[d] + zb[u]));
}

// Pass p of one step (step position U in the unrolled window, step row y):
// row y - LAG p, points of colour (p-1)&1, inside the level and inside the
// region still exact after p passes (a register select, not a branch).  Y is
// the parity of the step's row, known at compile time (segments start on
// even rows, W is even), so the column a pass updates is too: a runtime
// choice of za or zb would merge the two stores into one through a selected
// pointer and keep the window out of registers.
//   LAG = 2: a point of row r in pass p reads row r+1 after pass p-1 (run
// one step earlier) and row r-1 after pass p-1 and before pass p+1, and the
// P updates of a step are independent.  LAG = 1: row r+1's pass p-1 and row
// r-1's pass p+1 fall in the same step, before and after pass p since the
// passes of a step run in order; the window is P rows shorter, the updates
// of a step a chain.
template <bool INT, class T, int W, int U, int Y, int LAG, int p>
__device__ __forceinline__ void spass(const LevelT<T> &F, const CCOp &op, const SLane<T> &s, int y, int ylo, int yhi,
                                      SWin<T, W> &w) {
#pragma clang fp contract(off)
    constexpr int c = sslot<W>(U, LAG * p), d = sslot<W>(U, LAG * p + 1), u = sslot<W>(U, LAG * p - 1);
    const int r = y - LAG * p;
    const bool ok = r >= max(ylo + p, 0) && r <= min(yhi - p, F.n - 1);
    const T sgn = T(op.s);
    if constexpr (((Y - LAG * p) & 1) == ((p - 1) & 1)) {
        const T rhs = fma(-sgn, soff_a<T>(F.cx, F.cy, w.za, w.zb, c, d, u), w.ba[c]);
        const T v = rhs * by_row_class<INT, T>(s.ia0, s.ia1, s.ia2, r, F.n);
        w.za[c] = (ok && s.ina) ? v : w.za[c];
    } else {
        const T rhs = fma(-sgn, soff_b<T>(F.cx, F.cy, w.za, w.zb, c, d, u), w.bb[c]);
        const T v = rhs * by_row_class<INT, T>(s.ib0, s.ib1, s.ib2, r, F.n);
        w.zb[c] = (ok && s.inb) ? v : w.zb[c];
    }
}

// The P colour passes of one step (p = 1..P as template constants, so that
// every window slot is a compile-time register).
template <bool INT, class T, int W, int U, int Y, int LAG, int... Q>
__device__ __forceinline__ void spasses(const LevelT<T> &F, const CCOp &op, const SLane<T> &s, int y, int ylo,
                                        int yhi, SWin<T, W> &w, std::integer_sequence<int, Q...>) {
    (spass<INT, T, W, U, Y, LAG, Q + 1>(F, op, s, y, ylo, yhi, w), ...);
}

// Residual b - A z at both columns of the row in slot c (cc_resid_t's
// rounding).
template <bool INT, class T, int W>
__device__ __forceinline__ void sresid(const LevelT<T> &F, const CCOp &op, const SLane<T> &s, int r,
                                       const SWin<T, W> &w, int c, int d, int u, T &ra, T &rb) {
#pragma clang fp contract(off)
    const T sgn = T(op.s);
    ra = w.ba[c] - fma(sgn, soff_a<T>(F.cx, F.cy, w.za, w.zb, c, d, u),
                       by_row_class<INT, T>(s.da0, s.da1, s.da2, r, F.n) * w.za[c]);
    rb = w.bb[c] - fma(sgn, soff_b<T>(F.cx, F.cy, w.za, w.zb, c, d, u),
                       by_row_class<INT, T>(s.db0, s.db1, s.db2, r, F.n) * w.zb[c]);
}

// The arithmetic of one step: the passes and the residual of row q in slot
// c (rows below / above in d / u).
template <bool INT, class T, int P, int LAG, int W, int U, int Y>
__device__ __forceinline__ void scompute(const LevelT<T> &F, const CCOp &op, const SLane<T> &s, int y, int ylo,
                                         int yhi, int q, int c, int d, int u, SWin<T, W> &w, T &ra, T &rb) {
    spasses<INT, T, W, U, Y, LAG>(F, op, s, y, ylo, yhi, w, std::make_integer_sequence<int, P>{});
    sresid<INT, T, W>(F, op, s, q, w, c, d, u, ra, rb);
}

// One step (row y, window position U, row parity Y), without branches
// around memory operations:
// issue the loads of row y + kSAhead, finish row y (up), run the passes,
// then store z of row q = y - LAG P - 2 (zalt for down, z for up; dropped
// outside the segment's output rows and the strip's output columns) and add
// its residual to the restriction (down, stored after the odd row of each
// pair) or to ||r||^2 (up with the norm).
template <class T, int P, bool UP, int LAG, int W, int U, int Y>
__device__ __forceinline__ void sstep(const LvArgs<T> &a, const SLane<T> &s, const SBufs &bf, int y, int ylo, int yhi,
                                      int y0, int y1, int ck, int tb, SWin<T, W> &w, T &acc, double &rr) {
#pragma clang fp contract(off)
    const LevelT<T> &F = a.F, &C = a.C;
    // a step is a scheduling region of its own: the loads stay kSAhead steps
    // ahead (hoisting a window's loads to its top would keep every value of
    // the window live at once)
    __builtin_amdgcn_sched_barrier(0);
    srow_issue<T, W, sslot<W>(U, -kSAhead)>(F, bf, s, y + kSAhead, yhi, w);
    if constexpr (UP) {
        // with the window's first row ys = 2k + 1, step t = tb + U finishes a
        // row whose coarse rows are k + (t >> 1) and the next one (J and J2
        // in the order of the row's parity); an even step loads coarse row
        // k + (t + kSQAhead) / 2 + 1, first needed kSQAhead + 1 steps later
        constexpr int R = W / 2, A = U >> 1;
        if constexpr (!(U & 1))
            sq_issue<T, W, ((U + kSQAhead) / 2 + 1) % R>(C, bf, s, ck + (tb + U + kSQAhead) / 2 + 1, w);
        if constexpr (!(U & 1)) srow_fix<T, W, sslot<W>(U, 0), A % R, (A + 1) % R>(F, C, a.cop, s, y, yhi, w);
        else srow_fix<T, W, sslot<W>(U, 0), (A + 1) % R, A % R>(F, C, a.cop, s, y, yhi, w);
    }
    // the passes and the residual: rows y - LAG P - 2 .. y - LAG; away from
    // the first and last row of the level every point has the interior
    // row coefficients (a branch without memory operations)
    constexpr int c = sslot<W>(U, LAG * P + 2), d = sslot<W>(U, LAG * P + 3), u = sslot<W>(U, LAG * P + 1);
    const int q = y - LAG * P - 2;
    T ra, rb;
    if (q > 0 && y - LAG < F.n - 1) scompute<true, T, P, LAG, W, U, Y>(F, a.cop, s, y, ylo, yhi, q, c, d, u, w, ra, rb);
    else scompute<false, T, P, LAG, W, U, Y>(F, a.cop, s, y, ylo, yhi, q, c, d, u, w, ra, rb);
    const bool qrow = q >= y0 && q < y1;
    const int base = F.m * q;
    const bool va = qrow && s.outa, vb = qrow && s.outb;
    sstore(bf.dst, va ? (int)sizeof(T) * (base + s.xa) : kSOut, w.za[c]);
    sstore(bf.dst, vb ? (int)sizeof(T) * (base + s.xb) : kSOut, w.zb[c]);
    if constexpr (UP) {
        const double da = va ? (double)ra : 0.0, db = vb ? (double)rb : 0.0;
        rr += da * da;
        rr += db * db;
    } else if constexpr (!(Y & 1)) {
        acc = T(0);
        acc += ra;
        acc += rb;
    } else {
        acc += ra;
        acc += rb;
        acc *= T(0.25);
        sstore(bf.coarse, va ? (int)sizeof(T) * ((s.xa >> 1) + C.m * (q >> 1)) : kSOut, acc);
    }
}

// W consecutive steps (U = 0..W-1), so that every window slot is a
// compile-time register; the first row ys has parity YS.  The steps stop at
// step nsteps (a scalar branch per step; the && fold evaluates in order and
// stops at the first false).
template <class T, int P, bool UP, int LAG, int W, int YS, int... U>
__device__ __forceinline__ void sstep_window(const LvArgs<T> &a, const SLane<T> &s, const SBufs &bf, int ys, int tb,
                                             int nsteps, int ylo, int yhi, int y0, int y1, int ck, SWin<T, W> &w,
                                             T &acc, double &rr, std::integer_sequence<int, U...>) {
    (void)((tb + U < nsteps &&
            (sstep<T, P, UP, LAG, W, U, (YS + U) & 1>(a, s, bf, ys + tb + U, ylo, yhi, y0, y1, ck, tb, w, acc, rr),
             true)) &&
           ...);
}

// Item (strip, segment) of wavefront w of workgroup b.
// (readfirstlane: the item is the same for the whole wavefront, so every row
// test below is scalar)
__device__ __forceinline__ int stream_item(int nb) {
    return __builtin_amdgcn_readfirstlane(kSWaves * xcd_tile(blockIdx.x, nb) + (int)(threadIdx.x >> 6));
}

// The rows loaded before the first step: z and b of rows ys..ys+kSAhead-1,
// and (up) coarse rows ck .. ck + kSQAhead / 2.
template <class T, int W, bool UP, int... K>
__device__ __forceinline__ void sprologue(const LevelT<T> &F, const LevelT<T> &C, const SBufs &bf, const SLane<T> &s,
                                          int ys, int ylast, int ck, SWin<T, W> &w, std::integer_sequence<int, K...>) {
    (srow_issue<T, W, K>(F, bf, s, ys + K, ylast, w), ...);
    if constexpr (UP) ((K <= kSQAhead / 2 ? sq_issue<T, W, K>(C, bf, s, ck + K, w) : void()), ...);
}

template <class T, int NU, bool UP>
__device__ __forceinline__ void stream_body(const LvArgs<T> &a, int item, double &rr) {
    constexpr int P = 2 * NU;
    constexpr int HH = (P + 2) & ~1;   // columns of halo (even)
    constexpr int HV = P + 1;          // rows of halo
    constexpr int WOUT = kSW - 2 * HH;
    // pass lag 1 (see the schedule above)
    constexpr int LAG = 1;
    constexpr int W = LAG * P + 4 + kSAhead;  // even: a window's rows keep their parities
    static_assert(W % 2 == 0 && (P + 1) % 2 == 1, "even window, odd first row");
    static_assert(kSQAhead % 2 == 0 && W / 2 >= kSQAhead / 2 + 3, "the coarse ring holds the rows in use");
    const LevelT<T> &F = a.F, &C = a.C;
    const int S = a.ty;
    const int sx = item % a.nbx, sy = item / a.nbx;
    const int x0 = sx * WOUT, y0 = sy * S, y1 = min(y0 + S, F.n);
    SLane<T> s;
    slane_init<T>(F, C, a.cop, x0, WOUT, HH, s);
    const long nf = (long)F.m * F.n * sizeof(T), nc = (long)C.m * C.n * sizeof(T);
    SBufs bf;
    bf.zsrc = srsrc(UP ? F.zalt : F.z, (!UP && a.flag) ? 0 : nf);  // zero start: every z load is out of range
    bf.b = srsrc(F.b, nf);
    bf.dst = srsrc(UP ? F.z : F.zalt, nf);
    bf.coarse = srsrc(UP ? C.z : C.b, nc);
    SWin<T, W> w = {};
    // rows ys..ylast are loaded (those inside the level); the region exact
    // after pass p is [ys + p, ylast - p] clamped to the level
    const int ys = y0 - HV, ylast = y0 + S + HV - 1;
    const int ck = (ys - 1) >> 1;  // ys is odd: coarse row of the window's first row pair
    sprologue<T, W, UP>(F, C, bf, s, ys, ylast, ck, w, std::make_integer_sequence<int, kSAhead>{});
    // steps until the last output row has had its residual
    const int nsteps = (y1 - 1) + LAG * P + 2 - ys + 1;
    T acc = T(0);
    for (int tb = 0; tb < nsteps; tb += W)
        sstep_window<T, P, UP, LAG, W, HV & 1>(a, s, bf, ys, tb, nsteps, ys, ylast, y0, y1, ck, w, acc, rr,
                                          std::make_integer_sequence<int, W>{});
}

// Zero right-hand side found by the folded start-of-solve norm: z = 0 on the
// wavefront's output region (strips of wout columns).
template <class T>
__device__ __forceinline__ void stream_zero(const LvArgs<T> &a, int item, int wout) {
    const int sx = item % a.nbx, sy = item / a.nbx;
    const int x0 = sx * wout, y0 = sy * a.ty;
    const int lane = threadIdx.x & 63;
    for (int y = y0; y < min(y0 + a.ty, a.F.n); y++)
        for (int x = x0 + lane; x < min(x0 + wout, a.F.m); x += 64) a.F.z[x + (size_t)a.F.m * y] = T(0);
}

template <class T, int NU>
__global__ void __launch_bounds__(kSThreads) k_sdown(LvArgs<T> a) {
    __shared__ double sh[10];
    if (a.ctl.state->done && !a.fin_begin) return;
    stamp_begin(a.ts);
    const int item = stream_item(a.nb);
    double pf[2 * kFinRegs];
    if (a.fin) {
        fin_load(a.ctl, a.fin, a.fin_begin, pf);
        const int nd = fin_norm(a.ctl, a.fin, a.fin_begin, a.tol2, pf, sh);
        if (nd == 3 && a.fin_begin && sizeof(T) == sizeof(double) && item < a.items) stream_zero<T>(a, item, kSW - 2 * ((2 * NU + 2) & ~1));
        if (nd) return;
    }
    double rr = 0;
    if (item < a.items) stream_body<T, NU, false>(a, item, rr);
    stamp_end(a.ts);
}

template <class T, int NU>
__global__ void __launch_bounds__(kSThreads) k_sup(LvArgs<T> a) {
    __shared__ double sh[8];
    if (a.ctl.state->done) return;
    stamp_begin(a.ts);
    const int bid = xcd_tile(blockIdx.x, a.nb);
    const int item = stream_item(a.nb);
    double rr = 0;
    if (item < a.items) stream_body<T, NU, true>(a, item, rr);
    if (a.flag) store_partial(a.ctl, rr, bid, sh);
    stamp_end(a.ts);
}

// ------------------------------------------------------------------------
// Nodal (FEM) levels
// ------------------------------------------------------------------------
// The same scheme for the four-colour Gauss-Seidel of the 9-point Q1
// operator (fem_tile_sweep).  Pass p has colour (p-1)&3: columns of parity
// (p-1)&1 -- the lane's even (A) or odd (B) column -- on rows of parity
// ((p-1)>>1)&1; on a row of the other parity it updates nothing.  It runs one
// row behind the previous pass (LAG 1); a point of row r then reads rows r-1
// and r+1 between the passes a global sweep would give it, as for the
// cell-centred levels.  An update reads the other column's three rows, one
// of them from the neighbouring lane.  Row 0 is Dirichlet and never updated.
// The halo is 4 nu + 2 columns and 2 nu + 2 rows.  Along a row each pass
// shrinks the exact region by one column (a pass's points read the columns
// the pass before updated).  Across rows the passes come in pairs on rows of
// one parity, so after pass p the region has lost ceil(p / 2) rows at each
// end; the residual needs one row more, and the full-weighting restriction
// reads residuals one node beyond the output.

template <class T>
struct SNLane {
    int xa, xb;            // the lane's even and odd column (may lie outside the level)
    bool ina, inb, outa, outb;
    T wia, wib;            // edge-column factor of the vertical coupling (fem_off_t's wi)
    T ia0, ia2, ib0, ib2;  // fem_inv_acc for an interior row (0) and the last row (2)
    T aa0, aa1, ab0, ab1;  // fem_acc for an interior row (0) and the first or last row (1)
    T wla, wrb;            // restriction weight of the column left of A / of B (zero outside)
    int ci, ci1;           // up launches: coarse columns I = xa >> 1 and min(I + 1, Mc - 1)
};

template <class T>
__device__ __forceinline__ void snlane_init(const LevelT<T> &F, const LevelT<T> &C, int x0, int wout, int hh,
                                            const FemOp &op, SNLane<T> &s) {
    const int lane = threadIdx.x & 63;
    s.xa = x0 - hh + 2 * lane;
    s.xb = s.xa + 1;
    s.ina = s.xa >= 0 && s.xa < F.m;
    s.inb = s.xb >= 0 && s.xb < F.m;
    s.outa = s.ina && s.xa >= x0 && s.xa < x0 + wout;
    s.outb = s.inb && s.xb >= x0 && s.xb < x0 + wout;
    const int ca = min(max(s.xa, 0), F.m - 1), cb = min(max(s.xb, 0), F.m - 1);
    s.wia = (ca == 0 || ca == F.m - 1) ? T(0.5) : T(1);
    s.wib = (cb == 0 || cb == F.m - 1) ? T(0.5) : T(1);
    const int j0 = min(1, F.n - 1), j2 = F.n - 1;
    s.ia0 = fem_inv_acc<T>(F, ca, j0);
    s.ia2 = fem_inv_acc<T>(F, ca, j2);
    s.ib0 = fem_inv_acc<T>(F, cb, j0);
    s.ib2 = fem_inv_acc<T>(F, cb, j2);
    s.aa0 = fem_acc<T>(op, F.m, F.n, ca, j0);
    s.aa1 = fem_acc<T>(op, F.m, F.n, ca, 0);
    s.ab0 = fem_acc<T>(op, F.m, F.n, cb, j0);
    s.ab1 = fem_acc<T>(op, F.m, F.n, cb, 0);
    s.wla = (s.xa - 1 < 0 || s.xa - 1 >= F.m) ? T(0) : T(0.5);
    s.wrb = (s.xb >= F.m) ? T(0) : T(0.5);
    s.ci = min(max(s.xa, 0) >> 1, max(C.m - 1, 0));
    s.ci1 = min(s.ci + 1, max(C.m - 1, 0));
}

// Off-diagonal part of the nodal operator at the even (A) or odd (B) column
// of the row in slot c (rows below / above in d / u), fem_off_t's rounding.
template <class T>
__device__ __forceinline__ T snoff_a(const FemOp &op, T wi, T wj, const T *za, const T *zb, int c, int d, int u) {
    return fem_off_v<T>(op, wi, wj, lane_from_left(zb[d]), za[d], zb[d], lane_from_left(zb[c]), zb[c],
                        lane_from_left(zb[u]), za[u], zb[u]);
}
template <class T>
__device__ __forceinline__ T snoff_b(const FemOp &op, T wi, T wj, const T *za, const T *zb, int c, int d, int u) {
    return fem_off_v<T>(op, wi, wj, za[d], zb[d], lane_from_right(za[d]), za[c], lane_from_right(za[c]), za[u], zb[u],
                        lane_from_right(za[u]));
}

// Pass p of a step (row y - p, LAG 1).  Y: parity of the step's row.
template <bool INT, class T, int W, int U, int Y, int p>
__device__ __forceinline__ void snpass(const LevelT<T> &F, const FemOp &op, const SNLane<T> &s, int y, int ylo,
                                       int yhi, SWin<T, W> &w) {
#pragma clang fp contract(off)
    constexpr int cj = ((p - 1) >> 1) & 1, ci = (p - 1) & 1;
    if constexpr (((Y - p) & 1) == cj) {
        constexpr int c = sslot<W>(U, p), d = sslot<W>(U, p + 1), u = sslot<W>(U, p - 1);
        const int r = y - p;
        const bool ok = r >= max(ylo + (p + 1) / 2, 1) && r <= min(yhi - (p + 1) / 2, F.n - 1);
        const T wj = INT ? T(1) : (r == F.n - 1 ? T(0.5) : T(1));
        if constexpr (ci == 0) {
            const T v = (w.ba[c] - snoff_a<T>(op, s.wia, wj, w.za, w.zb, c, d, u)) *
                        (INT ? s.ia0 : (r == F.n - 1 ? s.ia2 : s.ia0));
            w.za[c] = (ok && s.ina) ? v : w.za[c];
        } else {
            const T v = (w.bb[c] - snoff_b<T>(op, s.wib, wj, w.za, w.zb, c, d, u)) *
                        (INT ? s.ib0 : (r == F.n - 1 ? s.ib2 : s.ib0));
            w.zb[c] = (ok && s.inb) ? v : w.zb[c];
        }
    }
}

template <bool INT, class T, int W, int U, int Y, int... Q>
__device__ __forceinline__ void snpasses(const LevelT<T> &F, const FemOp &op, const SNLane<T> &s, int y, int ylo,
                                         int yhi, SWin<T, W> &w, std::integer_sequence<int, Q...>) {
    (snpass<INT, T, W, U, Y, Q + 1>(F, op, s, y, ylo, yhi, w), ...);
}

// Residual b - A z at both columns of row q in slot c (fem_resid_t's
// rounding; row 0 has only the diagonal).
template <bool INT, class T, int W>
__device__ __forceinline__ void snresid(const LevelT<T> &F, const FemOp &op, const SNLane<T> &s, int q,
                                        const SWin<T, W> &w, int c, int d, int u, T &ra, T &rb) {
#pragma clang fp contract(off)
    const bool edge = !INT && (q == 0 || q == F.n - 1);
    const T wj = INT ? T(1) : (q == F.n - 1 ? T(0.5) : T(1));
    const T acca = edge ? s.aa1 : s.aa0, accb = edge ? s.ab1 : s.ab0;
    const T fa = w.ba[c] - (acca * w.za[c] + snoff_a<T>(op, s.wia, wj, w.za, w.zb, c, d, u));
    const T fb = w.bb[c] - (accb * w.zb[c] + snoff_b<T>(op, s.wib, wj, w.za, w.zb, c, d, u));
    ra = (!INT && q == 0) ? w.ba[c] - acca * w.za[c] : fa;
    rb = (!INT && q == 0) ? w.bb[c] - accb * w.zb[c] : fb;
}

template <bool INT, class T, int P, int W, int U, int Y>
__device__ __forceinline__ void sncompute(const LevelT<T> &F, const FemOp &op, const SNLane<T> &s, int y, int ylo,
                                          int yhi, int q, int c, int d, int u, SWin<T, W> &w, T &ra, T &rb) {
    snpasses<INT, T, W, U, Y>(F, op, s, y, ylo, yhi, w, std::make_integer_sequence<int, P>{});
    snresid<INT, T, W>(F, op, s, q, w, c, d, u, ra, rb);
}

// Up launches: the two coarse values (columns I, I1) of coarse row J into
// ring slot K.
template <class T, int W, int K>
__device__ __forceinline__ void snq_issue(const LevelT<T> &C, const SBufs &bf, const SNLane<T> &s, int J,
                                          SWin<T, W> &w) {
    const int r = C.m * min(max(J, 0), C.n - 1);
    const bool in = s.ina || s.inb;
    w.qc[K] = sload(T(), bf.coarse, in ? (int)sizeof(T) * (r + s.ci) : kSOut);
    w.qa[K] = sload(T(), bf.coarse, in ? (int)sizeof(T) * (r + s.ci1) : kSOut);
}

// Up launches, when row y reaches the step front: z = zalt + the
// prolongated correction (fem_prolong_all, none on row 0) from ring slots
// KJ (coarse row J = y >> 1) and KJ1 (J + 1, or J itself on the last coarse
// row), zero outside the level or the segment's rows.
template <class T, int W, int K, int KJ, int KJ1>
__device__ __forceinline__ void snrow_fix(const LevelT<T> &F, const LevelT<T> &C, const SNLane<T> &s, int y,
                                          int ylast, SWin<T, W> &w) {
    const bool row = y >= 0 && y < F.n && y <= ylast;
    const bool same = min(max(y, 0), F.n - 1) / 2 + 1 > C.n - 1;
    const T c1 = same ? w.qc[KJ] : w.qc[KJ1], a1 = same ? w.qa[KJ] : w.qa[KJ1];
    const int jo = y & 1;
    const T ea = fem_prolong_combine<T>(0, jo, w.qc[KJ], w.qa[KJ], c1, a1);
    const T eb = fem_prolong_combine<T>(1, jo, w.qc[KJ], w.qa[KJ], c1, a1);
    const T va = y > 0 ? w.za[K] + ea : w.za[K], vb = y > 0 ? w.zb[K] + eb : w.zb[K];
    w.za[K] = (row && s.ina) ? va : T(0);
    w.zb[K] = (row && s.inb) ? vb : T(0);
    asm volatile("" : "+v"(w.za[K]), "+v"(w.zb[K]));
}

// One nodal step (row y, window position U, row parity Y): loads of row
// y + kSAhead, finish row y (up), passes, residual of row q = y - P - 2;
// down: zalt of row q for the segment's output rows [y0, y1), and the
// full-weighting restriction (coarse row J collects rows 2J-1, 2J, 2J+1
// in that order, weights 1/2, 1, 1/2, zero outside the level; coarse row 0
// is zero) over the residual rows y0 - 1 .. y0 + S - 1; up: z of row q and
// ||r||^2.
template <class T, int P, bool UP, int W, int U, int Y>
__device__ __forceinline__ void snstep(const LvArgs<T> &a, const SNLane<T> &s, const SBufs &bf, int y, int ylo,
                                       int yhi, int y0, int y1, int ck, int tb, SWin<T, W> &w, T &acc, T &acc_next,
                                       double &rr) {
#pragma clang fp contract(off)
    const LevelT<T> &F = a.F, &C = a.C;
    __builtin_amdgcn_sched_barrier(0);
    srow_issue<T, W, sslot<W>(U, -kSNAhead<UP, P>)>(F, bf, s, y + kSNAhead<UP, P>, yhi, w);
    if constexpr (UP) {
        constexpr int R = W / 2, A = U >> 1;
        if constexpr (!(U & 1))
            snq_issue<T, W, ((U + kSQAhead) / 2 + 1) % R>(C, bf, s, ck + (tb + U + kSQAhead) / 2 + 1, w);
        snrow_fix<T, W, sslot<W>(U, 0), A % R, (A + 1) % R>(F, C, s, y, yhi, w);
    }
    constexpr int c = sslot<W>(U, P + 2), d = sslot<W>(U, P + 3), u = sslot<W>(U, P + 1);
    const int q = y - P - 2;
    T ra, rb;
    if (q > 0 && y - 1 < F.n - 1) sncompute<true, T, P, W, U, Y>(F, a.fop, s, y, ylo, yhi, q, c, d, u, w, ra, rb);
    else sncompute<false, T, P, W, U, Y>(F, a.fop, s, y, ylo, yhi, q, c, d, u, w, ra, rb);
    const bool qrow = q >= y0 && q < y1;
    const int base = F.m * q;
    const bool va = qrow && s.outa, vb = qrow && s.outb;
    sstore(bf.dst, va ? (int)sizeof(T) * (base + s.xa) : kSOut, w.za[c]);
    sstore(bf.dst, vb ? (int)sizeof(T) * (base + s.xb) : kSOut, w.zb[c]);
    if constexpr (UP) {
        const double da = va ? (double)ra : 0.0, db = vb ? (double)rb : 0.0;
        rr += da * da;
        rr += db * db;
    } else {
        // restrict_at's row sum: columns 2I-1 (the left lane's odd one), 2I, 2I+1
        T sx = T(0);
        sx += s.wla * lane_from_left(rb);
        sx += ra;
        sx += s.wrb * rb;
        const T wrow = q >= F.n ? T(0) : T(0.5);
        if constexpr (!(Y & 1)) {
            acc += (q >= F.n ? T(0) : T(1)) * sx;  // row 2J
        } else {
            acc += wrow * sx;  // row 2J + 1 closes coarse row J ...
            const int J = q >> 1;
            const bool st = J >= (y0 >> 1) && J < min((y0 + a.ty) >> 1, C.n) && s.outa;
            sstore(bf.coarse, st ? (int)sizeof(T) * ((s.xa >> 1) + C.m * J) : kSOut, J == 0 ? T(0) : acc);
            acc = T(0);  // ... and opens coarse row J + 1
            acc += wrow * sx;
        }
    }
    (void)acc_next;
}

// The rows loaded before the first step: z and b of rows ys..ys+kSAhead-1,
// and (up) coarse rows ck .. ck + kSQAhead / 2.
template <class T, int W, bool UP, int... K>
__device__ __forceinline__ void snprologue(const LevelT<T> &F, const LevelT<T> &C, const SBufs &bf,
                                           const SNLane<T> &s, int ys, int ylast, int ck, SWin<T, W> &w,
                                           std::integer_sequence<int, K...>) {
    (srow_issue<T, W, K>(F, bf, s, ys + K, ylast, w), ...);
    if constexpr (UP) ((K <= kSQAhead / 2 ? snq_issue<T, W, K>(C, bf, s, ck + K, w) : void()), ...);
}

template <class T, int P, bool UP, int W, int YS, int... U>
__device__ __forceinline__ void snstep_window(const LvArgs<T> &a, const SNLane<T> &s, const SBufs &bf, int ys, int tb,
                                              int nsteps, int ylo, int yhi, int y0, int y1, int ck, SWin<T, W> &w,
                                              T &acc, double &rr, std::integer_sequence<int, U...>) {
    T dummy = T(0);
    (void)((tb + U < nsteps && (snstep<T, P, UP, W, U, (YS + U) & 1>(a, s, bf, ys + tb + U, ylo, yhi, y0, y1, ck, tb,
                                                                     w, acc, dummy, rr),
                                true)) &&
           ...);
}

template <class T, int NU, bool UP>
__device__ __forceinline__ void nstream_body(const LvArgs<T> &a, int item, double &rr) {
    constexpr int P = 4 * NU;
    constexpr int HH = 4 * NU + 2;  // columns of halo (even)
    constexpr int HV = 2 * NU + 2;  // rows of halo (even)
    constexpr int WOUT = kSW - 2 * HH;
    constexpr int W = P + 4 + kSNAhead<UP, P>;
    static_assert(W % 2 == 0 && HH % 2 == 0 && HV % 2 == 0, "even window, even first row");
    static_assert(kSQAhead % 2 == 0 && W / 2 >= kSQAhead / 2 + 3, "the coarse ring holds the rows in use");
    const LevelT<T> &F = a.F, &C = a.C;
    const int S = a.ty;
    const int sx = item % a.nbx, sy = item / a.nbx;
    const int x0 = sx * WOUT, y0 = sy * S, y1 = min(y0 + S, F.n);
    SNLane<T> s;
    snlane_init<T>(F, C, x0, WOUT, HH, a.fop, s);
    const long nf = (long)F.m * F.n * sizeof(T), nc = (long)C.m * C.n * sizeof(T);
    SBufs bf;
    bf.zsrc = srsrc(UP ? F.zalt : F.z, (!UP && a.flag) ? 0 : nf);
    bf.b = srsrc(F.b, nf);
    bf.dst = srsrc(UP ? F.z : F.zalt, nf);
    bf.coarse = srsrc(UP ? C.z : C.b, nc);
    SWin<T, W> w = {};
    const int ys = y0 - HV, ylast = y0 + S + HV - 1;
    const int ck = ys >> 1;  // ys is even: coarse row of the window's first row
    snprologue<T, W, UP>(F, C, bf, s, ys, ylast, ck, w, std::make_integer_sequence<int, kSNAhead<UP, P>>{});
    // residual rows up to the segment's last output row (down: row 2J + 1
    // of its last coarse row J, outside the level for the top one)
    const int qlast = UP ? y1 - 1 : 2 * min((y0 + S) >> 1, C.n) - 1;
    const int nsteps = qlast + P + 2 - ys + 1;
    T acc = T(0);
    for (int tb = 0; tb < nsteps; tb += W)
        snstep_window<T, P, UP, W, 0>(a, s, bf, ys, tb, nsteps, ys, ylast, y0, y1, ck, w, acc, rr,
                                      std::make_integer_sequence<int, W>{});
}

template <class T, int NU>
__global__ void __launch_bounds__(kSThreads) k_sndown(LvArgs<T> a) {
    __shared__ double sh[10];
    if (a.ctl.state->done && !a.fin_begin) return;
    stamp_begin(a.ts);
    const int item = stream_item(a.nb);
    double pf[2 * kFinRegs];
    if (a.fin) {
        fin_load(a.ctl, a.fin, a.fin_begin, pf);
        const int nd = fin_norm(a.ctl, a.fin, a.fin_begin, a.tol2, pf, sh);
        if (nd == 3 && a.fin_begin && sizeof(T) == sizeof(double) && item < a.items)
            stream_zero<T>(a, item, kSW - 2 * (4 * NU + 2));
        if (nd) return;
    }
    double rr = 0;
    if (item < a.items) nstream_body<T, NU, false>(a, item, rr);
    stamp_end(a.ts);
}

template <class T, int NU>
__global__ void __launch_bounds__(kSThreads) k_snup(LvArgs<T> a) {
    __shared__ double sh[8];
    if (a.ctl.state->done) return;
    stamp_begin(a.ts);
    const int bid = xcd_tile(blockIdx.x, a.nb);
    const int item = stream_item(a.nb);
    double rr = 0;
    if (item < a.items) nstream_body<T, NU, true>(a, item, rr);
    if (a.flag) store_partial(a.ctl, rr, bid, sh);
    stamp_end(a.ts);
}

// ------------------------------------------------------------------------
// The mixed form's level-0 up launch (mx_body) as a streaming launch
// ------------------------------------------------------------------------
// The FP32 correction e is prolongated and post-smoothed as in an up launch
// (lag 1, with the FP32 level's own coefficients).  Behind the last pass an
// FP64 stage forms z_out = z_in + e on a three-row ring (zero outside the
// level), the FP64 residual r = b - A z_out with the level-0 operator
// (mx_az's rounding), and stores z_out, r (FP32: the next cycle's
// right-hand side) and ||r||^2.

constexpr int kSMxAhead = 2;  // rows by which z_in and b (FP64) are loaded ahead
// z_in's lookahead: its first row, y0 - 1, is loaded at step y0 + P - KZ,
// which must not lie before the first step y0 - HV (P + HV = 1 for the
// cell-centred levels without post-smoothing)
template <int P, int HV> constexpr int kSMxZ = P + HV < kSMxAhead ? P + HV : kSMxAhead;

template <int W>
struct SMxWin {
    double zia[W], zib[W], bza[W], bzb[W], zoa[W], zob[W];
};

struct SMxBufs {
    SRsrc zin, b, zout, rout;
};

// FP64 residual of row q at both columns (mx_az), z_out rows q-1, q, q+1 in
// slots d, c, u.
template <bool FEM, bool INT, int W, class LN>
__device__ __forceinline__ void smx_resid(const MxArgs &m, const LN &s, int q, const SMxWin<W> &x, int c, int d,
                                          int u, double &ra, double &rb) {
#pragma clang fp contract(off)
    const Level &Z = m.Z;
    if constexpr (FEM) {
        const bool edge = !INT && (q == 0 || q == Z.n - 1);
        const double wj = INT ? 1.0 : (q == Z.n - 1 ? 0.5 : 1.0);
        const double acca = edge ? s.aa1 : s.aa0, accb = edge ? s.ab1 : s.ab0;
        const double fa = acca * x.zoa[c] + snoff_a<double>(m.fop, s.wia, wj, x.zoa, x.zob, c, d, u);
        const double fb = accb * x.zob[c] + snoff_b<double>(m.fop, s.wib, wj, x.zoa, x.zob, c, d, u);
        ra = x.bza[c] - ((!INT && q == 0) ? acca * x.zoa[c] : fa);
        rb = x.bzb[c] - ((!INT && q == 0) ? accb * x.zob[c] : fb);
    } else {
        ra = x.bza[c] - fma(m.cop.s, soff_a<double>(Z.cx, Z.cy, x.zoa, x.zob, c, d, u),
                            by_row_class<INT, double>(s.da0, s.da1, s.da2, q, Z.n) * x.zoa[c]);
        rb = x.bzb[c] - fma(m.cop.s, soff_b<double>(Z.cx, Z.cy, x.zoa, x.zob, c, d, u),
                            by_row_class<INT, double>(s.db0, s.db1, s.db2, q, Z.n) * x.zob[c]);
    }
}

// One step of the streaming mixed level-0 up launch (row y, window position
// U, row parity Y); LE / LZ: the lane state for the FP32 level E and the FP64
// level Z.
template <bool FEM, int P, int KA, int KZ, int W, int U, int Y, class LE, class LZ>
__device__ __forceinline__ void smx_step(const MxArgs &m, const LE &s, const LZ &sz, const SBufs &bf,
                                         const SMxBufs &bd, int y, int ylo, int yhi, int y0, int y1, int ck, int tb,
                                         SWin<float, W> &w, SMxWin<W> &x, double &rr) {
#pragma clang fp contract(off)
    const LevelF &E = m.E, &C = m.C;
    const Level &Z = m.Z;
    __builtin_amdgcn_sched_barrier(0);
    // FP32 stage: the correction's loads, prolongation and passes (lag 1)
    srow_issue<float, W, sslot<W>(U, -KA)>(E, bf, s, y + KA, yhi, w);
    constexpr int R = W / 2, A = U >> 1;
    const bool interior = y - P - 2 > 0 && y - 1 < E.n - 1;
    if constexpr (FEM) {
        if constexpr (!(U & 1))
            snq_issue<float, W, ((U + kSQAhead) / 2 + 1) % R>(C, bf, s, ck + (tb + U + kSQAhead) / 2 + 1, w);
        snrow_fix<float, W, sslot<W>(U, 0), A % R, (A + 1) % R>(E, C, s, y, yhi, w);
        if (interior) snpasses<true, float, W, U, Y>(E, m.fop, s, y, ylo, yhi, w, std::make_integer_sequence<int, P>{});
        else snpasses<false, float, W, U, Y>(E, m.fop, s, y, ylo, yhi, w, std::make_integer_sequence<int, P>{});
    } else {
        if constexpr (!(U & 1))
            sq_issue<float, W, ((U + kSQAhead) / 2 + 1) % R>(C, bf, s, ck + (tb + U + kSQAhead) / 2 + 1, w);
        if constexpr (!(U & 1)) srow_fix<float, W, sslot<W>(U, 0), A % R, (A + 1) % R>(E, C, m.cop, s, y, yhi, w);
        else srow_fix<float, W, sslot<W>(U, 0), (A + 1) % R, A % R>(E, C, m.cop, s, y, yhi, w);
        if (interior) spasses<true, float, W, U, Y, 1>(E, m.cop, s, y, ylo, yhi, w, std::make_integer_sequence<int, P>{});
        else spasses<false, float, W, U, Y, 1>(E, m.cop, s, y, ylo, yhi, w, std::make_integer_sequence<int, P>{});
    }
    // FP64 stage: z_in and b loaded KZ and kSMxAhead rows ahead of their
    // use, only on the rows the segment's residuals read (z_out rows y0 - 1
    // .. y1, b rows y0 .. y1 - 1)
    {
        const int ri = y - P - 1 + KZ, rb = y - P - 2 + kSMxAhead;
        constexpr int ki = sslot<W>(U, P + 1 - KZ), kb = sslot<W>(U, P + 2 - kSMxAhead);
        const bool vi = ri >= max(y0 - 1, 0) && ri <= min(y1, Z.n - 1), vb = rb >= y0 && rb < y1;
        x.zia[ki] = sload(0.0, bd.zin, (vi && s.ina) ? 8 * (Z.m * ri + s.xa) : kSOut);
        x.zib[ki] = sload(0.0, bd.zin, (vi && s.inb) ? 8 * (Z.m * ri + s.xb) : kSOut);
        x.bza[kb] = sload(0.0, bd.b, (vb && s.ina) ? 8 * (Z.m * rb + s.xa) : kSOut);
        x.bzb[kb] = sload(0.0, bd.b, (vb && s.inb) ? 8 * (Z.m * rb + s.xb) : kSOut);
    }
    // z_out of row r1 = y - P - 1, whose correction is final (mx_body's zd)
    {
        const int r1 = y - P - 1;
        constexpr int k1 = sslot<W>(U, P + 1);
        const bool row = r1 >= 0 && r1 < Z.n;
        x.zoa[k1] = (row && s.ina) ? x.zia[k1] + (double)w.za[k1] : 0.0;
        x.zob[k1] = (row && s.inb) ? x.zib[k1] + (double)w.zb[k1] : 0.0;
    }
    // the residual of row q = y - P - 2 and the stores
    constexpr int c = sslot<W>(U, P + 2), d = sslot<W>(U, P + 3), u = sslot<W>(U, P + 1);
    const int q = y - P - 2;
    double ra, rb;
    if (q > 0 && q < Z.n - 1) smx_resid<FEM, true, W>(m, sz, q, x, c, d, u, ra, rb);
    else smx_resid<FEM, false, W>(m, sz, q, x, c, d, u, ra, rb);
    const bool qrow = q >= y0 && q < y1;
    const int base = Z.m * q;
    const bool va = qrow && s.outa, vb = qrow && s.outb;
    sstore(bd.zout, va ? 8 * (base + s.xa) : kSOut, x.zoa[c]);
    sstore(bd.zout, vb ? 8 * (base + s.xb) : kSOut, x.zob[c]);
    sstore(bd.rout, va ? 4 * (base + s.xa) : kSOut, (float)ra);
    sstore(bd.rout, vb ? 4 * (base + s.xb) : kSOut, (float)rb);
    const double qa = va ? ra : 0.0, qb = vb ? rb : 0.0;
    rr += qa * qa;
    rr += qb * qb;
}

template <bool FEM, int P, int KA, int KZ, int W, int YS, class LE, class LZ, int... U>
__device__ __forceinline__ void smx_window(const MxArgs &m, const LE &s, const LZ &sz, const SBufs &bf,
                                           const SMxBufs &bd, int ys, int tb, int nsteps, int ylo, int yhi, int y0,
                                           int y1, int ck, SWin<float, W> &w, SMxWin<W> &x, double &rr,
                                           std::integer_sequence<int, U...>) {
    (void)((tb + U < nsteps && (smx_step<FEM, P, KA, KZ, W, U, (YS + U) & 1>(m, s, sz, bf, bd, ys + tb + U, ylo, yhi,
                                                                             y0, y1, ck, tb, w, x, rr),
                                true)) &&
           ...);
}

template <bool FEM, int NU>
__device__ __forceinline__ void mx_stream_body(const MxArgs &m, int item, double &rr) {
    constexpr int P = FEM ? 4 * NU : 2 * NU;
    constexpr int HH = FEM ? 4 * NU + 2 : (2 * NU + 2) & ~1;  // columns of halo (even)
    constexpr int HV = FEM ? 2 * NU + 2 : 2 * NU + 1;          // rows of halo
    constexpr int WOUT = kSW - 2 * HH;
    constexpr int KA = FEM ? kSNAhead<true, P> : kSAhead;
    constexpr int KZ = kSMxZ<P, HV>;
    constexpr int W = P + 4 + KA;
    static_assert(W % 2 == 0 && W / 2 >= kSQAhead / 2 + 3, "even window holding the coarse rows in use");
    const LevelF &E = m.E, &C = m.C;
    const int S = m.ty;
    const int sx = item % m.nbx, sy = item / m.nbx;
    const int x0 = sx * WOUT, y0 = sy * S, y1 = min(y0 + S, E.n);
    const long nf = (long)E.m * E.n;
    SBufs bf;
    bf.zsrc = srsrc(E.zalt, nf * 4);
    bf.b = srsrc(E.b, nf * 4);
    bf.dst = srsrc(E.z, 0);
    bf.coarse = srsrc(C.z, (long)C.m * C.n * 4);
    SMxBufs bd;
    bd.zin = srsrc(m.z_in, nf * 8);
    bd.b = srsrc(m.Z.b, nf * 8);
    bd.zout = srsrc(m.z_out, nf * 8);
    bd.rout = srsrc(m.r_out, nf * 4);
    SWin<float, W> w = {};
    SMxWin<W> x = {};
    const int ys = y0 - HV, ylast = y0 + S + HV - 1;
    const int nsteps = (y1 - 1) + P + 2 - ys + 1;
    if constexpr (FEM) {
        SNLane<float> s;
        snlane_init<float>(E, C, x0, WOUT, HH, m.fop, s);
        SNLane<double> sz;
        snlane_init<double>(m.Z, m.Z, x0, WOUT, HH, m.fop, sz);
        const int ck = ys >> 1;
        snprologue<float, W, true>(E, C, bf, s, ys, ylast, ck, w, std::make_integer_sequence<int, KA>{});
        for (int tb = 0; tb < nsteps; tb += W)
            smx_window<true, P, KA, KZ, W, 0>(m, s, sz, bf, bd, ys, tb, nsteps, ys, ylast, y0, y1, ck, w, x, rr,
                                          std::make_integer_sequence<int, W>{});
    } else {
        SLane<float> s;
        slane_init<float>(E, C, m.cop, x0, WOUT, HH, s);
        SLane<double> sz;
        slane_init<double>(m.Z, m.Z, m.cop, x0, WOUT, HH, sz);
        const int ck = (ys - 1) >> 1;
        sprologue<float, W, true>(E, C, bf, s, ys, ylast, ck, w, std::make_integer_sequence<int, KA>{});
        for (int tb = 0; tb < nsteps; tb += W)
            smx_window<false, P, KA, KZ, W, HV & 1>(m, s, sz, bf, bd, ys, tb, nsteps, ys, ylast, y0, y1, ck, w, x, rr,
                                                std::make_integer_sequence<int, W>{});
    }
}

template <bool FEM, int NU>
__global__ void __launch_bounds__(kSThreads) k_smx(MxArgs a) {
    __shared__ double sh[8];
    if (a.ctl.state->done) return;
    stamp_begin(a.ts);
    const int bid = xcd_tile(blockIdx.x, a.nb);
    const int item = stream_item(a.nb);
    double rr = 0;
    if (item < a.items) mx_stream_body<FEM, NU>(a, item, rr);
    store_partial(a.ctl, rr, bid, sh);
    stamp_end(a.ts);
}

}  // namespace vdi
