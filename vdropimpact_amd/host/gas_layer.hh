// gas_layer.hh -- the gas layer of one simulation, held on the GPU.
//
// Host face of gas_layer / gas_layer_solve (gas_layer.hh:10-60,
// gas_layer_solve.hh:20-80): the physical parameters (gl_params), the update
// call of the step (gas_layer::update, done on the device by
// vdi_gas_layer_update) and the profile writers of the frame output
// (gas_layer_solve.cc:258-308), which read the device arrays.
#pragma once

#include <cstdio>

#include "vdi.h"

// gl_params (gas_layer_solve.hh:20-41).
struct gl_params {
    double alpha;  // 1 / (exponent of the pressure-density relation)
    double mu;     // gas viscosity
    double Pamb;   // ambient pressure
    double h0;     // initial height of the drop
    double R;      // drop radius
    double V;      // initial drop speed
    double sigma;  // surface tension
};

class gas_layer : public gl_params {
  public:
    const int m, me, ml;
    const double ax, bx, dx;
    const bool x_sym;
    gas_layer(int m_, double ax_, double bx_, const gl_params &p, bool x_sym_);
    // The simulation whose device arrays this layer reads (set by fluid_2d).
    void attach(vdi_ctx *c) { ctx = c; }
    // gas_layer::update (gas_layer.cc:36-46), on the device.
    void update(double t, double dt);
    // Gnuplot binary matrix profile: fld 0 height, 1 flow across the
    // boundary, 2 gas pressure (gas_layer_solve.cc:258-282).
    void output_profile(const char *filename, float *buf, int sn, double ay, int fld);
    // The same profiles at full precision (gas_layer_solve.cc:291-308).
    void output_profile_double(const char *filename, float *buf, int sn, int fld);

  private:
    vdi_ctx *ctx = nullptr;
    void fetch(int fld, double *out, int *len);
};
