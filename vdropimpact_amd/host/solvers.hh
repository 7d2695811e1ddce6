// solvers.hh -- the four linear solvers of a simulation, on the GPU.
//
// mgs_mac::solve_v_cycle(t) (mgs_mac.hh:54-59), mgs_fem::solve_v_cycle(t)
// (mgs_fem.hh:61-66) and visco_impl::setup_and_solve(vfx, vfy, t)
// (visco_impl.cc:21-39) with the reference's meaning: a time-extrapolated
// initial guess, then V-cycles on the simulation's shared source array until
// the solve converges; failure stops the program with the reference's
// message.  The V-cycle counts feed the frame timing line (fluid_2d.cc:66-86).
#pragma once

#include "vdi.h"

class solver_base {
  public:
    solver_base(int which) : which_(which) {}
    void attach(vdi_ctx *c) { ctx_ = c; }
    // V-cycles and solves since the last reset_counters (per frame).
    long vcycles() const;
    double avg_iters() const;
    void reset_counters();

  protected:
    vdi_ctx *ctx_ = nullptr;
    int which_;
    long base_cycles_ = 0, base_solves_ = 0;
};

class mgs_mac : public solver_base {
  public:
    mgs_mac() : solver_base(VDI_MAC) {}
    void solve_v_cycle(double t);
};

class mgs_fem : public solver_base {
  public:
    mgs_fem() : solver_base(VDI_FEM) {}
    void solve_v_cycle(double t);
};

class visco_impl : public solver_base {
  public:
    explicit visco_impl(bool horizontal_vel) : solver_base(horizontal_vel ? VDI_VX : VDI_VY) {}
    void setup_and_solve(double vfx, double vfy, double t);
};
