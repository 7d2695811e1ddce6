// fileinfo.hh -- the drop-impact configuration file (.cfg) reader.
//
// Same keywords, defaults, checks, derived quantities and messages as the
// reference's fileinfo (fileinfo.hh:17-116, fileinfo.cc:11-193), so the
// reference's sims/*.cfg files are read unchanged.  Two of the reference's
// options have no GPU step behind them and stop with a message: the gas-layer
// data files (gas_layer_data, a testing model) and the non-inertial frame
// (nif_center, nif_range).
#pragma once

#include <string>

#include "gas_layer.hh"

// Output flags of the `output` keyword (fileinfo.cc:96-111).
enum {
    OUT_U = 1, OUT_V = 2, OUT_P = 4, OUT_PB = 8, OUT_PBFULL = 16, OUT_W = 32, OUT_H = 64, OUT_FBD = 128,
    OUT_PG = 256, OUT_HFULL = 512, OUT_FBDFULL = 1024, OUT_PGFULL = 2048
};

struct fileinfo {
    int m = -1, n = -1;           // grid cells
    unsigned int fflags = 0;      // OUT_* bits
    int ntrace = 0;               // tracers
    int nframes = -1;             // output frames
    int restart_freq = 0;         // restart file every this many frames (0: none)
    int nif_mode = 0;             // non-inertial frame (0 only)
    bool implicit_visc = false;
    bool x_sym = false;
    bool mr_time_output = false;  // machine-readable timing lines
    double ax = 0, bx = 0, by = 0;
    const double ay = 0;
    double mu = -1, muinv = 0, rho = -1, rhoinv = 0;
    double tmult = -1;            // time step padding factor
    double t_end = 0;             // simulation duration (s)
    std::string filename;         // output directory: the .cfg name with .odr
    gas_layer *gl = nullptr;      // the gas layer (gas_layer_model)
    explicit fileinfo(const char *infile);
    ~fileinfo();
    fileinfo(const fileinfo &) = delete;
    fileinfo &operator=(const fileinfo &) = delete;
};
