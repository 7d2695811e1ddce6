// fluid_2d.hh -- the drop-impact time integrator with its state on the GPU.
//
// The reference's fluid_2d (fluid_2d.hh:21-140) with the same public
// interface: construction from a .cfg file, the frame loop solve(), the step
// step_forward(dt), initial fields, restart files and the output writers.
// The step and the solvers run in libvdi_hip.so through the C ABI of
// include/vdi.h; this class keeps the frame loop, the files and a host copy of
// what a frame writes.  The solver wrappers (mgs_mac, mgs_fem, visco_impl)
// are in solvers.hh.
#pragma once

#include <cstdio>
#include <vector>

#include "fileinfo.hh"
#include "solvers.hh"
#include "vdi.h"

class fluid_2d : public fileinfo {
  public:
    const int mn, me, ne, ml;
    const double dx, dy, xsp, ysp, xxsp, yysp;
    const double dt_reg;       // the regular time step (fluid_2d.cc:17)
    std::vector<double> tm;    // tracer positions x0, y0, x1, y1, ... (host copy)
    double time = 0;
    int f_num = 0;
    explicit fluid_2d(const char *infile);
    ~fluid_2d();
    fluid_2d(const fluid_2d &) = delete;
    fluid_2d &operator=(const fluid_2d &) = delete;
    // fluid_2d.cc:40-89: frames 1..nframes of l steps each, with output.
    void solve();
    // fluid_2d.cc:93-405 on the GPU (tracers, gas layer, the whole step).
    void step_forward(double dt);
    // fluid_2d_sio.cc:14-35
    void init_fields();
    // fluid_2d_sio.cc:37-119: res.<sn> in the output directory.
    void load_restart(int sn);
    void load_last_restart();
    void save_restart(int sn);
    // fluid_2d_sio.cc:149-166: the fields chosen by `output`, tracers, restart.
    void write_files(int k);
    // fluid_2d.cc:495-505: random positions in the domain (C rand()).
    void init_tracers();
    // Gnuplot matrix of u (mode 0), v (1), p (2), p on the bottom row (3) or
    // the vorticity (4) (fluid_2d_sio.cc:185-224).
    void output(const char *prefix, int mode, int sn, bool ghost = false);
    void output_pbase_full(int sn);
    void output_tracers(int sn);
    // The largest step not above dt_reg that divides interval (fluid_2d.hh:82-86).
    int timestep_select(double interval, double &adt) const {
        const int l = static_cast<int>(interval / dt_reg) + 1;
        adt = interval / l;
        return l;
    }
    FILE *odir_open(const char *prefix, int sn, const char *mode) const;
    vdi_ctx *gpu() const { return ctx; }

  private:
    vdi_ctx *ctx = nullptr;
    mgs_mac ms_mac;
    mgs_fem ms_fem;
    visco_impl vix, viy;
    std::vector<float> buf;
    std::vector<double> hu, hv, hp;   // host copies of u, v, p for a frame
    // rand() draws for init_tracers, taken before the GPU runtime starts (it
    // calls rand() itself), so the positions are the ones the reference's
    // process draws from the C library's default seed
    std::vector<int> tracer_draws;
    bool restart_exists(int sn) const;
    void fetch_fields();
    double vorticity(int i, int j) const;
};
