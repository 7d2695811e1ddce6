// mg.hpp -- geometric multigrid for the four linear systems of the step.
//
// Replaces the external TGMG library that the reference drives through
// tgmg<S,double,double>::solve_v_cycle (mgs_mac.hh:57, mgs_fem.hh:64,
// visco_impl.cc:31).  The operators are the reference's (mgs_mac.hh:25-50,
// mgs_fem.hh:42-56 + mgs_fem.cc:34-50, visco_impl.hh:28-53); the cycle is this
// project's own:
//   cell-centred (MAC, visco): red-black Gauss-Seidel, average-of-4
//     restriction, bilinear prolongation with boundary-type ghosts,
//     rediscretised coarse operators (couplings / 4 per level);
//   nodal FEM: four-colour Gauss-Seidel, full-weighting restriction (the
//     transpose of bilinear prolongation) so that the coarse Q1 stiffness is
//     the Galerkin one, Dirichlet row kept at zero correction;
//   coarsest level (the first with at most kExactPoints unknowns): solved
//     exactly by separation of variables -- every operator here is a sum of
//     two Kronecker products of 1-D tridiagonal matrices, so a generalized
//     eigenbasis of the shorter direction turns it into independent
//     tridiagonal systems along the longer one (see MG::coarse_factors).
//
// Every kernel takes one argument record.  A direct launch passes it by
// value; a batched launch passes an array of records (one per simulation)
// and a work list of (record, tile) pairs that a persistent grid walks, so
// one launch advances the same stage of many independent systems.
#pragma once

#include <memory>

#include "vdi_internal.hpp"

namespace vdi {

// Factors of the coarsest level's exact solve (FP64): the eigenbasis V of the
// transform direction and the Thomas factors of every mode's tridiagonal
// system.  Shared: a captured step keeps the factors its coarse kernels were
// recorded with, and a coefficient change while they are held writes a fresh
// buffer instead.
struct CoarseFactors {
    DevBuf f;
};

constexpr int kMaxNu = 4;       // largest nu1/nu2 the tiled kernels accept
// Level-kernel workgroups: 32 x 16 = 512 threads (eight waves).  With the
// 64 x 32 tiles of the finest level there is about one tile per CU, and 512
// threads instead of 256 made the cell-centred V-cycle at 1024 x 512 6 %
// faster (two waves per SIMD to overlap loads and sweeps); 1024 threads
// exceed the register budget of the nodal kernels.
constexpr int kBX = 32, kBY = 16;
constexpr int kLvThreads = kBX * kBY;
constexpr int kCoarseBX = 32, kCoarseBY = 16;  // coarse-kernel threads, 2-D
constexpr int kCoarseThreads = kCoarseBX * kCoarseBY;
// Coarsening stops at the first level with at most kExactPoints unknowns
// (or earlier, on odd sizes); the single-workgroup kernel solves that level,
// exactly when its shorter direction has at most kSepMaxT unknowns, else by
// ncs smoothing sweeps.
constexpr int kExactPoints = 3072;
constexpr int kSepMaxT = 64;
// A restricted level (not level 0) of at most kExactBig unknowns and at most
// kExactBigSide a side is the coarsest one already, and is solved exactly by
// three multi-workgroup launches (both directions diagonalised): rows times
// U, columns through V^T, 1 / (lam + mu) and V, rows times U^T.
constexpr int kExactBig = 9000;
constexpr int kExactBigSide = 256;
// Two block layouts of the transforms, with the same bits.  Throughput
// (batched launches, where every CU is busy): one output per thread, row
// transforms kXRows x kXCols outputs per workgroup, the column transform
// kXColsV columns (of at most 64 unknowns).  Latency (direct launches, a
// single system leaving most CUs idle): every output's dot product shared by
// four lanes (x_dot_quad) on four times the workgroups, kXRowsQ x kXCols and
// kXColsVQ.
constexpr int kXRows = 8, kXCols = 32, kXColsV = 4;
constexpr int kXRowsQ = 2, kXColsVQ = 1;
static_assert(kXCols == 32, "the row transform's staging indexes columns by shift and mask");
constexpr int kXThreads = 256;
static_assert(kXRows * kXCols == kXThreads && 4 * kXRowsQ * kXCols == kXThreads && 4 * 64 * kXColsVQ == kXThreads,
              "one output per thread / four lanes per output");
// The exact solve's transforms take a block of kSepModes modes (rows) by
// kSepCols columns per thread; its factors are staged kSepLoads values per
// thread and round (one round for the 64 x 32 levels of the 1024 x 512
// workload).
constexpr int kSepModes = 2;
constexpr int kSepCols = 2;
constexpr int kSepLoads = 12;

// Device-side solve state: convergence flag (0 running, 1 converged,
// 2 NaN, 3 zero right-hand side), V-cycles run, ||b-Az||^2 and ||b||^2.
struct MGState {
    int done;
    int cycles;
    int pad[2];
    double rr, bb;
};

struct MGCtl {
    MGState *state;
    double *partial;  // one (or two) doubles per workgroup
};

// Output tile of a level kernel (cells or nodes); both even.  stream = 1:
// a streaming launch (cell-centred levels only) -- one wavefront per strip
// of tx = 128 columns and ty rows (mg_stream.hpp); its output strips are
// narrower by the halo.
struct TileShape {
    int tx, ty;
    int stream;
};

// Tile shapes the autotuner chooses from.
constexpr int kNumTileShapes = 22;
extern const TileShape kTileShapes[kNumTileShapes];

// Down / up kernel of one level.
template <class T>
struct LvArgs {
    LevelT<T> F, C;
    CCOp cop;
    FemOp fop;
    int nu;
    int flag;          // down: start from z = 0; up: compute ||b - A z||^2 (level 0)
    int tx, ty, nbx, nb;  // tile shape, tiles per row, tiles in the launch
    MGCtl ctl;
    double tol2;
    Stamp *ts;
    // level-0 down: number of ||b - A z||^2 partials of the previous V-cycle
    // to finish (sum, convergence test, state update) before smoothing, in
    // place of a k_norm_finish launch; 0: none (at most kFinMax)
    int fin;
    // 1: those partials are k_begin's (||b - A z||^2, ||b||^2) pairs
    int fin_begin;
    // streaming launch (mg_stream.hpp): wavefront items (strips x segments);
    // nbx is then the number of strips and nb the number of workgroups.
    // 0: a tiled launch.
    int items;
};
constexpr int kFinMax = 1024;

// Mixed form, level 0 up kernel (see k_mx_up0).
struct MxArgs {
    Level Z;
    const double *z_in;
    double *z_out;
    LevelF E;
    float *r_out;
    LevelF C;
    CCOp cop;
    FemOp fop;
    int nu, tx, ty, nbx, nb;
    MGCtl ctl;
    double tol2;
    Stamp *ts;
    int items;  // streaming launch (mg_stream.hpp): wavefront items; 0: tiled
};

// The coarsest level for the single-workgroup kernel.  LDS holds the level
// as two (m + 2) x (n + 2) tiles with a zero border (z, then b, in the
// kernel's precision), then a double workspace at byte offset ws: the
// transformed right-hand side (T rows of P + 1), V (T x T), the two Thomas
// factors (T rows of P + 1 each) and the T off-diagonals.
struct CoarseDesc {
    int top, m, n;
    double cx, cy, inv;
    double *gz, *gb;
    int sep;       // 1: exact separable solve; 0: ncs smoothing sweeps
    int tdir;      // transform direction: 0 along y (rows), 1 along x (columns)
    int T, P, t0;  // transform unknowns, tridiagonal length, first transformed row/column
    int ws;        // byte offset of the double workspace
    const double *fac;  // V ([t][k], T x T), inv and cp ([k][p], T x P each), off-diagonals (T)
};
// One phase of the multi-workgroup exact solve of the coarsest level
// (phase 1: w = B U over rows; 2: each column of w through V^T, the
// diagonal and V; 3: z = w U^T over rows).  w is the level's zalt array
// ([t * P + j], T x P), b and z the level's own; unknown row t is level row
// t + yoff (yoff = 1 for the nodal system's Dirichlet row 0).
template <class T>
struct XArgs {
    int phase, m, n, yoff, Tn, P, nbr, nbc;
    int quad;  // 1: latency layout (kXRowsQ, kXColsVQ, four lanes per output)
    const T *b;
    T *w, *z;
    const double *fac;  // V (Tn x Tn), U (P x P), U^T, 1 / (lam + mu) (Tn x P)
    FemOp fop;
    MGCtl ctl;
    Stamp *ts;
};

template <class T>
struct CoarseArgs {
    CoarseDesc d;
    CCOp cop;
    FemOp fop;
    int fem, nu1, nu2, ncs;
    int nu1s, nu2s;  // sweeps on the restricted levels (nu1, nu2 on a top level held here)
    MGCtl ctl;
    double tol2;
    Stamp *ts;
};

// Start of a solve: ||b - A z||^2, ||b||^2 and the flag.
struct BeginArgs {
    int fem;
    Level L;
    CCOp cop;
    FemOp fop;
    MGCtl ctl;
    double tol2;
    float *r32;  // mixed form: FP32 copy of the residual
    int nb;
    Stamp *ts;
    // 1: workgroup 0 clears the flag and the cycle count (the first level-0
    // down kernel finishes this launch's norms, LvArgs::fin_begin)
    int reset;
};

// Element-wise operations on level-0 arrays.
enum { VEC_EXTRAP = 0, VEC_SOLVE_END = 1, VEC_DIRICHLET = 2 };
struct VecArgs {
    int op;
    int mixed;
    long N;
    double *z, *zp;
    const double *zalt, *b;
    const double *coef;     // VEC_EXTRAP: (b, c) in device memory
    const MGState *state;   // VEC_SOLVE_END
    double fm;              // VEC_DIRICHLET: FEM diagonal scale (a_cc = fm/4 at corners, fm/2 along the row)
    int M;
    int nb;
};

// Second half of the norm reductions (k_norm_finish).
struct NormArgs {
    MGCtl ctl;
    int nb, begin;
    double tol2;
};

// Work item of a batched launch: record index, tile (or block) index and
// the simulation it belongs to (for per-simulation busy time).
struct WorkItem {
    int rec, tile, sim;
};

// Per-launch extras of a batched launch: the launch stamp and, optionally,
// per-simulation workgroup busy ticks (atomically accumulated).
struct BatchAux {
    Stamp *ts;
    unsigned long long *busy;
};

class MG {
  public:
    int kind = VDI_MAC;
    bool fem = false;
    int nl = 0;   // levels
    int lc = 0;   // the coarsest level (nl - 1), solved by the single-workgroup kernel
    std::vector<Level> L;    // FP64 hierarchy
    std::vector<LevelF> LF;  // FP32 hierarchy of the mixed form (level 0: e32, r32)
    std::vector<int> lm, ln;
    CCOp cc{};
    FemOp fe{};
    int nu1 = 2, nu2 = 2, ncs = 8;
    int nu1s = 2, nu2s = 2;  // V(nu1s, nu2s) on levels 1 .. nl - 2
    int form = VDI_FORM_FP64;
    hipStream_t st = nullptr;
    // tile shape index (into kTileShapes) per level: [precision][down/up][level]
    std::vector<int> tile[2][2];
    int tile_mx = 0;

    // time-extrapolation history (mgs_common.hh:21-38, mgs_common.cc:17-50)
    int scount = 0;
    double t0 = 0, t_prev = 0;
    double vfx = 0, vfy = 0;  // visco couplings set up (visco_impl.cc:26-30)

    int last_cycles = 0;
    long total_cycles = 0, solves = 0;
    KernelTimer *timer = nullptr;  // optional launch timing

    // Sets up the hierarchy.  z0/b0 are the level-0 solution and RHS buffers
    // (owned by the caller); when null the MG allocates its own.  With an
    // arena every device buffer is a piece of it and nothing is allocated
    // (FP64 form only: no FP32 hierarchy, no time-extrapolation history, no
    // pinned readback).
    int init(int kind, int m, int n, double cx, double cy, int x_sym, double *z0, double *b0,
             hipStream_t stream, Arena *arena = nullptr);
    // Sets the level couplings and recomputes the coarse factors; returns their status.
    int set_coeffs(double cx, double cy);
    // The coarse factors in use, and re-adopting ones recorded earlier for the
    // couplings (cx, cy) (no recomputation; the level couplings follow).
    std::shared_ptr<CoarseFactors> coarse_factors_in_use() const { return fac_; }
    void adopt_coeffs(double cx, double cy, const std::shared_ptr<CoarseFactors> &fac);
    int unknowns() const { return lm[0] * ln[0]; }
    double *z() const { return L[0].z; }
    double *b() const { return L[0].b; }
    double *z_prev() const { return zprev_.as<double>(); }
    MGCtl ctl() const;

    // Argument records (shared by direct and batched launches).
    template <class T> LvArgs<T> down_args(const std::vector<LevelT<T>> &H, int l, int zero_z);
    template <class T> LvArgs<T> up_args(const std::vector<LevelT<T>> &H, int l, int norm, double tol2);
    template <class T> CoarseArgs<T> coarse_args(const std::vector<LevelT<T>> &H, double tol2);
    MxArgs mx_args(int parity, double tol2);
    LvArgs<float> mx_down0_args(int parity);
    BeginArgs begin_args(double tol2);
    VecArgs vec_args(int op, const double *coef);
    // begin = 1: after k_begin; 0: after a V-cycle's level-0 up kernel.
    NormArgs norm_args(int begin, double tol2);
    // Byte counts of the launches above (for timing).
    double down_bytes(int prec, int l, int zero_z) const;
    double up_bytes(int prec, int l) const;
    double coarse_bytes(int prec) const;
    double mx_bytes() const;

    // FEM Dirichlet row z = b / a_cc on level 0 (no-op for other kinds).
    int enforce_dirichlet();
    // Enqueues the start of a solve: Dirichlet row, ||b||, ||b - A z||.
    int begin(double tol);
    // Enqueues one V-cycle; kernels return at once when the device flag says
    // the solve has converged (tol2 < 0: never converge).
    int launch_cycle(double tol2);
    // fin_prev: the level-0 down kernel finishes the previous cycle's norm
    // (which that cycle did not launch, norm = 0); only when lc > 0.
    int launch_cycle(double tol2, int fin_prev, int norm);
    // begin() whose norms the first cycle's level-0 down kernel finishes
    // (launch_cycle with fin_prev = 2); only when can_fold_norm().
    int begin_folded(double tol);
    // Whether a run of cycles can fold each cycle's norm into the next one.
    bool can_fold_norm() const;
    long up0_partials() const;
    // Mixed form in use (needs at least one tiled level).
    bool mixed_active() const;
    int mixed_finish(int executed);
    // Algorithmic bytes of one V-cycle (arrays read + written per kernel).
    double cycle_bytes() const;
    int read_state(MGState *out);
    // Synchronous residual norms.
    int residual(double *rnorm, double *bnorm);
    // V-cycles until ||b-Az|| <= tol ||b||; *cycles = V-cycles run.
    int solve(double tol, int max_cycles, int *cycles);
    // Exactly k V-cycles.
    int vcycles(int k);
    // mgs_common::initial_guess(t)
    int initial_guess(double t);
    // Host half of initial_guess for captured steps: updates the time
    // history and returns 1 with the extrapolation coefficients, 0 while the
    // history is filling (scount < 2), -1 for duplicate times.
    int guess_coeffs(double t, double *bb, double *c);
    // Device half: extrapolation with coefficients read from device memory.
    int enqueue_extrapolate(const double *coef);
    // A solve with a fixed cycle budget and no host synchronisation; the
    // outcome is in state() afterwards.
    int enqueue_fixed(double tol, int budget);
    MGState *state() const { return state_.as<MGState>(); }
    // out = A z for level-0 device arrays
    int apply(const double *z, double *out);
    // r = b - A z (r may be null) and nb per-block partial sums of r^2 in
    // part[0..nb), on this system's stream (capturable).
    int residual_out(const double *z, double *r, double *part, int nb);
    // Times every candidate tile shape of every tiled level kernel on this
    // system's arrays and keeps the fastest; level-0 z and b are restored.
    int autotune(int reps);
    // Sets one tile choice: prec 0 FP64 / 1 FP32, which 0 down / 1 up /
    // 2 mixed level-0 up; shape indexes kTileShapes.
    int set_tile(int prec, int which, int level, int shape);
    // Smoothing sweeps per level (0..kMaxNu, not both 0).  A tile shape whose
    // halo no longer fits LDS moves to the largest default shape that does.
    int set_schedule(int nu1, int nu2) { return set_schedule(nu1, nu2, nu1, nu2); }
    int set_schedule(int nu1, int nu2, int nu1s, int nu2s);
    // LDS bytes of a launch.
    size_t down_lds(int prec, int l) const;
    size_t up_lds(int prec, int l) const;
    size_t mx_lds() const;
    size_t coarse_lds(int prec) const { return coarse_lds_[prec ? 1 : 0]; }
    // Recomputes the factors of the coarsest level's exact solve (after a
    // coefficient change).
    int coarse_factors();
    // Zeroes the solve state and the partial sums (fill kernels on st).
    int clear_solve_state();
    // A fresh solver for a new simulation on the same grid and couplings:
    // empty time history, z = z_prev = 0 (as a newly constructed one).
    int reset_history();

  private:
    int apply_schedule(int nu1, int nu2, int nu1s, int nu2s);
    template <class T> int launch_down(const std::vector<LevelT<T>> &H, int l, int zero_z, int fin = 0,
                                       double tol2 = 0, int fin_begin = 0);
    template <class T> int launch_up(const std::vector<LevelT<T>> &H, int l, int norm, double tol2);
    template <class T> int launch_coarse(const std::vector<LevelT<T>> &H, double tol2);
  public:
    template <class T> XArgs<T> x_args(const std::vector<LevelT<T>> &H, int phase, int quad = 1);
    size_t x_lds(int phase) const;
    bool big_exact() const { return big_; }
  private:
    int launch_vec(const VecArgs &a);
    DevBuf store_;  // coarse z/b
    DevBuf store_f_;
    float *r32_[2] = {nullptr, nullptr};  // FP32 residual, ping-pong
    int mx_enqueued_ = 0;
    DevBuf zprev_;
    DevBuf partial_;
    DevBuf state_;
    DevBuf own_z_, own_b_;
    Arena *arena_ = nullptr;              // set: buffers come from it (see init)
    std::shared_ptr<CoarseFactors> fac_;  // exact solve of the coarsest level (null: sweeps)
    int sep_tdir_ = 0, sep_T_ = 0, sep_P_ = 0, sep_t0_ = 0;
    bool big_ = false;  // the multi-workgroup exact solve (kExactBig)
    void level_coeffs(double cx, double cy);
    MGState *host_state_ = nullptr;  // pinned
    size_t coarse_lds_[2] = {0, 0};

  public:
    ~MG();
};

// Kernel launchers shared with the batched step (mg.hip).
namespace mgk {
// Direct launches.
template <class T> int down(const LvArgs<T> &a, bool fem, size_t lds, hipStream_t st);
template <class T> int up(const LvArgs<T> &a, bool fem, size_t lds, hipStream_t st);
template <class T> int coarse(const CoarseArgs<T> &a, size_t lds, hipStream_t st);
template <class T> int xsolve(const XArgs<T> &a, size_t lds, hipStream_t st);
template <class T> int xsolve_grid(const XArgs<T> &a);
int mx(const MxArgs &a, bool fem, size_t lds, hipStream_t st);
int begin(const BeginArgs &a, hipStream_t st);
int vec(const VecArgs &a, hipStream_t st);
int norm(const NormArgs &a, hipStream_t st);
int norm_b(const NormArgs *args, int n, hipStream_t st);
// Batched launches: args[k] are records in device memory, work items pick
// (record, tile); grid = persistent workgroup count (a multiple of 8).
template <class T> int down_b(const LvArgs<T> *args, const WorkItem *w, int nw, bool fem, size_t lds, int grid,
                              BatchAux aux, hipStream_t st);
template <class T> int up_b(const LvArgs<T> *args, const WorkItem *w, int nw, bool fem, size_t lds, int grid,
                            BatchAux aux, hipStream_t st);
template <class T> int xsolve_b(const XArgs<T> *args, const WorkItem *w, int nw, size_t lds, int grid, BatchAux aux,
                                hipStream_t st);
const void *xsolve_kernel(int prec, bool batched);
template <class T> int coarse_b(const CoarseArgs<T> *args, const WorkItem *w, int nw, size_t lds, int grid,
                                BatchAux aux, hipStream_t st);
int mx_b(const MxArgs *args, const WorkItem *w, int nw, bool fem, size_t lds, int grid, BatchAux aux, hipStream_t st);
int begin_b(const BeginArgs *args, const WorkItem *w, int nw, int grid, BatchAux aux, hipStream_t st);
int vec_b(const VecArgs *args, const WorkItem *w, int nw, int grid, BatchAux aux, hipStream_t st);
const void *coarse_kernel(int prec, bool batched);
// Workgroups of a persistent grid for a kernel (occupancy x CUs, multiple of 8).
int persistent_grid(const void *kernel, size_t lds, int nw);
const void *down_kernel(int prec, bool fem, bool batched);
const void *up_kernel(int prec, bool fem, bool batched);
const void *mx_kernel(bool fem, bool batched);
// One-time kernel attributes (dynamic LDS up to the device limit).
int set_attributes();
int max_lds();
}  // namespace mgk

}  // namespace vdi
