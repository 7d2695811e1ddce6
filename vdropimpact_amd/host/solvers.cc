#include "solvers.hh"

#include <cstdio>
#include <cstdlib>

#include "common.hh"

long solver_base::vcycles() const {
    vdi_stats s;
    gpu_check(vdi_get_stats(ctx_, &s), "vdi_get_stats");
    return s.total_cycles[which_] - base_cycles_;
}

double solver_base::avg_iters() const {
    vdi_stats s;
    gpu_check(vdi_get_stats(ctx_, &s), "vdi_get_stats");
    const long n = s.solves[which_] - base_solves_;
    return n > 0 ? double(s.total_cycles[which_] - base_cycles_) / n : 0.0;
}

void solver_base::reset_counters() {
    vdi_stats s;
    gpu_check(vdi_get_stats(ctx_, &s), "vdi_get_stats");
    base_cycles_ = s.total_cycles[which_];
    base_solves_ = s.solves[which_];
}

// mgs_mac.hh:66-70, mgs_fem.hh:73-77 (fatal_error), visco_impl.cc:34-37 (fputs)
static void solve_failed(int st, const char *msg, bool fatal) {
    if (st == VDI_ERR_MG_NOCONV) {
        if (fatal) fatal_error(msg, 1);
        fprintf(stderr, "%s\n", msg);
        exit(1);
    }
    gpu_check(st, "solve");
}

void mgs_mac::solve_v_cycle(double t) {
    int c = 0;
    const int st = vdi_mg_solve_v_cycle(ctx_, VDI_MAC, t, &c);
    if (st != VDI_OK) solve_failed(st, "Multigrid routine for MAC projection failed", true);
}

void mgs_fem::solve_v_cycle(double t) {
    int c = 0;
    const int st = vdi_mg_solve_v_cycle(ctx_, VDI_FEM, t, &c);
    if (st != VDI_OK) solve_failed(st, "Multigrid routine for finite-element problem failed", true);
}

void visco_impl::setup_and_solve(double vfx, double vfy, double t) {
    int c = 0;
    const int st = vdi_visco_setup_and_solve(ctx_, which_, vfx, vfy, t, &c);
    if (st != VDI_OK) solve_failed(st, "Multigrid routine for implicit viscosity failed", false);
}
