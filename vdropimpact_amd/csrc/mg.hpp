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
//   coarsest level: a fixed number of smoothing sweeps.
#pragma once

#include "vdi_internal.hpp"

namespace vdi {

constexpr int kMaxCoarse = 12;  // levels handled by the single-workgroup kernel
constexpr int kMaxNu = 3;       // largest nu1/nu2 the tiled kernels accept

// Device-side solve state: convergence flag (0 running, 1 converged,
// 2 NaN, 3 zero right-hand side), V-cycles run, ticket of the last-
// workgroup reduction, ||b-Az||^2 and ||b||^2.
struct MGState {
    int done;
    int cycles;
    unsigned ticket;
    int pad;
    double rr, bb;
};

struct MGCtl {
    MGState *state;
    double *partial;  // one (or two) doubles per workgroup
};

// Levels lc..L-1 for the single-workgroup kernel (LDS offsets in doubles;
// z at off, b at off + m*n).
struct CoarseDesc {
    int nl, top;
    int m[kMaxCoarse], n[kMaxCoarse], off[kMaxCoarse];
    double cx[kMaxCoarse], cy[kMaxCoarse];
    double *gz, *gb;
};

class MG {
  public:
    int kind = VDI_MAC;
    bool fem = false;
    int nl = 0;   // levels
    int lc = 0;   // first level of the single-workgroup kernel
    std::vector<Level> L;    // FP64 hierarchy
    std::vector<LevelF> LF;  // FP32 hierarchy of the mixed form (level 0: e32, r32)
    std::vector<int> lm, ln;
    CCOp cc{};
    FemOp fe{};
    int nu1 = 2, nu2 = 2, ncs = 8;
    int form = VDI_FORM_FP64;
    hipStream_t st = nullptr;

    // time-extrapolation history (mgs_common.hh:21-38, mgs_common.cc:17-50)
    int scount = 0;
    double t0 = 0, t_prev = 0;
    double vfx = 0, vfy = 0;  // visco couplings set up (visco_impl.cc:26-30)

    int last_cycles = 0;
    long total_cycles = 0, solves = 0;
    KernelTimer *timer = nullptr;  // optional launch timing

    // Sets up the hierarchy.  z0/b0 are the level-0 solution and RHS buffers
    // (owned by the caller); when null the MG allocates its own.
    int init(int kind, int m, int n, double cx, double cy, int x_sym, double *z0, double *b0,
             hipStream_t stream);
    void set_coeffs(double cx, double cy);
    int unknowns() const { return lm[0] * ln[0]; }
    double *z() const { return L[0].z; }
    double *b() const { return L[0].b; }
    double *z_prev() const { return zprev_.as<double>(); }
    MGCtl ctl() const;

    // FEM Dirichlet row z = b / a_cc on level 0 (no-op for other kinds).
    int enforce_dirichlet();
    // Enqueues the start of a solve: Dirichlet row, ||b||, ||b - A z||.
    int begin(double tol);
    // Enqueues one V-cycle; kernels return at once when the device flag says
    // the solve has converged (tol2 < 0: never converge).
    int launch_cycle(double tol2);
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

  private:
    template <class T> int launch_down(const std::vector<LevelT<T>> &H, int l, int zero_z);
    template <class T> int launch_up(const std::vector<LevelT<T>> &H, int l, int norm, double tol2);
    template <class T> int launch_coarse(const std::vector<LevelT<T>> &H, double tol2);
    DevBuf store_;  // coarse z/b
    DevBuf store_f_;
    float *r32_[2] = {nullptr, nullptr};  // FP32 residual, ping-pong
    int mx_enqueued_ = 0;
    DevBuf zprev_;
    DevBuf partial_;
    DevBuf state_;
    DevBuf own_z_, own_b_;
    MGState *host_state_ = nullptr;  // pinned
    size_t coarse_lds_ = 0;

  public:
    ~MG();
};

}  // namespace vdi
