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
//     the Galerkin one, Dirichlet row kept at zero correction.
//   coarsest level: a fixed number of smoothing sweeps in one workgroup.
#pragma once

#include "vdi_internal.hpp"

namespace vdi {

class MG {
  public:
    int kind = VDI_MAC;
    bool fem = false;
    int nl = 0;
    std::vector<Level> L;      // FP64 hierarchy
    std::vector<int> lm, ln;   // level sizes
    CCOp cc{};
    FemOp fe{};
    int nu1 = 2, nu2 = 2, ncs = 64;
    int form = VDI_FORM_FP64;
    hipStream_t st = nullptr;

    // time-extrapolation history (mgs_common.hh:21-38, mgs_common.cc:17-50)
    int scount = 0;
    double t0 = 0, t_prev = 0;
    double vfx = 0, vfy = 0;   // visco couplings set up (visco_impl.cc:26-30)

    int last_cycles = 0;
    long total_cycles = 0, solves = 0;

    // Sets up the hierarchy.  z0/b0 are the level-0 solution and RHS buffers
    // (owned by the caller); when null the MG allocates its own.
    int init(int kind, int m, int n, double cx, double cy, int x_sym, double *z0, double *b0,
             hipStream_t stream);
    void set_coeffs(double cx, double cy);
    int unknowns() const { return lm[0] * ln[0]; }
    double *z() const { return L[0].z; }
    double *b() const { return L[0].b; }
    double *z_prev() const { return zprev_.as<double>(); }

    int vcycle();
    // Enqueues ||b-Az||^2 and ||b||^2 into the device pair norms_dev();
    int enqueue_norms();
    // Synchronous residual norms.
    int residual(double *rnorm, double *bnorm);
    // FEM Dirichlet row z = b / a_cc on level 0 (no-op for other kinds).
    int enforce_dirichlet();
    // V-cycles until ||b-Az|| <= tol ||b||; *cycles = V-cycles run.
    int solve(double tol, int max_cycles, int *cycles);
    int vcycles(int k);
    // mgs_common::initial_guess(t)
    int initial_guess(double t);
    // out = A z for level-0 device arrays
    int apply(const double *z, double *out);

  private:
    DevBuf store_;   // coarse z/b
    DevBuf zprev_;
    DevBuf partial_;
    DevBuf norms_;   // 2 doubles
    DevBuf own_z_, own_b_;
    double *host_norms_ = nullptr;  // pinned
    int nblk_norm_ = 0;
    int vcycle_level(int l);
    int sweep(int l, int count);
  public:
    ~MG();
};

}  // namespace vdi
