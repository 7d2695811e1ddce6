// cfg_dump -- reads a .cfg file as the simulation would and prints what it
// derived, one key=value per line (the parser's test and a user's check of a
// configuration before a run).  Touches no GPU.
#include <cstdio>

#include "fileinfo.hh"

int main(int argc, char **argv) {
    if (argc != 2) {
        fputs("Syntax: ./cfg_dump <input_file>\n", stderr);
        return 1;
    }
    fileinfo fi(argv[1]);
    const double dx = (fi.bx - fi.ax) / fi.m;
    const double dt_reg = fi.implicit_visc ? dx * fi.tmult : dx * dx * fi.rho * fi.muinv * fi.tmult * 0.25;
    printf("m=%d\nn=%d\nfflags=%u\nntrace=%d\nnframes=%d\nrestart_freq=%d\n", fi.m, fi.n, fi.fflags, fi.ntrace,
           fi.nframes, fi.restart_freq);
    printf("implicit_visc=%d\nx_sym=%d\nmr_time_output=%d\n", (int)fi.implicit_visc, (int)fi.x_sym,
           (int)fi.mr_time_output);
    printf("ax=%.17g\nbx=%.17g\nay=%.17g\nby=%.17g\n", fi.ax, fi.bx, fi.ay, fi.by);
    printf("mu=%.17g\nrho=%.17g\ntmult=%.17g\nt_end=%.17g\ndt_reg=%.17g\n", fi.mu, fi.rho, fi.tmult, fi.t_end, dt_reg);
    printf("alpha=%.17g\nmug=%.17g\nPamb=%.17g\nh0=%.17g\nR=%.17g\nV=%.17g\nsigma=%.17g\n", fi.gl->alpha, fi.gl->mu,
           fi.gl->Pamb, fi.gl->h0, fi.gl->R, fi.gl->V, fi.gl->sigma);
    printf("filename=%s\n", fi.filename.c_str());
    return 0;
}
