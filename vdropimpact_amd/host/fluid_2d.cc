// fluid_2d.cc -- frame loop, initial state, restart files and output writers
// of the drop-impact simulation whose step runs on the GPU (the reference's
// fluid_2d.cc:14-89, 495-505 and fluid_2d_sio.cc:14-224 formats).
#include "fluid_2d.hh"

#include <sys/stat.h>

#include <cstdlib>
#include <cstring>
#include <string>

#include "common.hh"

fluid_2d::fluid_2d(const char *infile)
    : fileinfo(infile), mn(m * n), me(m + 1), ne(n + 1), ml(m + 4), dx((bx - ax) / m), dy((by - ay) / n),
      xsp(1. / dx), ysp(1. / dy), xxsp(xsp * xsp), yysp(ysp * ysp),
      dt_reg(implicit_visc ? dx * tmult : dx * dx * rho * muinv * tmult * 0.25), vix(true), viy(false),
      buf(m > 123 ? m + 5 : 128) {
    tracer_draws.resize(2 * (size_t)ntrace);
    for (int &r : tracer_draws) r = rand();
    vdi_params P;
    P.m = m;
    P.n = n;
    P.x_sym = x_sym;
    P.implicit_visc = implicit_visc;
    P.ax = ax;
    P.bx = bx;
    P.ay = ay;
    P.by = by;
    P.rho = rho;
    P.mu = mu;
    P.dt_reg = dt_reg;
    P.alpha = gl->alpha;
    P.mug = gl->mu;
    P.Pamb = gl->Pamb;
    P.h0 = gl->h0;
    P.R = gl->R;
    P.V = gl->V;
    P.sigma = gl->sigma;
    gpu_check(vdi_create(&P, &ctx), "vdi_create");
    gl->attach(ctx);
    ms_mac.attach(ctx);
    ms_fem.attach(ctx);
    vix.attach(ctx);
    viy.attach(ctx);
}

fluid_2d::~fluid_2d() {
    if (ctx) vdi_destroy(ctx);
}

void fluid_2d::solve() {
    double adt;
    const int l = timestep_select(t_end / nframes, adt);
    if (f_num == 0) {
        write_files(0);
        puts("# Output frame 0");
    }
    ms_mac.reset_counters();
    ms_fem.reset_counters();
    vix.reset_counters();
    viy.reset_counters();
    double t0 = wtime();
    for (int k = f_num + 1; k <= nframes; k++) {
        for (int j = 0; j < l; j++) step_forward(adt);
        gpu_check(vdi_synchronize(ctx), "vdi_synchronize");
        const double t1 = wtime();
        write_files(k);
        const double t2 = wtime();
        // The solves run inside the step on the GPU (partly side by side), so
        // no wall time of their own is kept: the solver and gas-layer time
        // columns of the machine-readable line are 0; the V-cycle counts are
        // the frame's.
        if (mr_time_output) {
            printf("%d %d %.8g %.8g  %d %.8g %d %.8g", k, l, t1 - t0, t2 - t1, (int)ms_mac.vcycles(), 0.0,
                   (int)ms_fem.vcycles(), 0.0);
            if (!implicit_visc) printf("  %.8g %.8g\n", 0.0, 0.0);
            else printf(" %d %.8g %d %.8g  %.8g %.8g\n", (int)vix.vcycles(), 0.0, (int)viy.vcycles(), 0.0, 0.0, 0.0);
        } else {
            printf("# Output frame %d [%d, %.8g s, %.8g s] {MAC %.2f, FEM %.2f", k, l, t1 - t0, t2 - t1,
                   ms_mac.avg_iters(), ms_fem.avg_iters());
            if (!implicit_visc) puts("}");
            else printf(", VX %.2f, VY %.2f}\n", vix.avg_iters(), viy.avg_iters());
        }
        fflush(stdout);
        ms_mac.reset_counters();
        ms_fem.reset_counters();
        vix.reset_counters();
        viy.reset_counters();
        t0 = t2;
    }
}

void fluid_2d::step_forward(double dt) {
    gpu_check(vdi_step_forward(ctx, dt), "step_forward");
    gpu_check(vdi_get_time(ctx, &time), "vdi_get_time");
}

void fluid_2d::init_fields() {
    f_num = 0;
    time = 0;
    gpu_check(vdi_init_fields(ctx), "init_fields");
    if (ntrace > 0) init_tracers();
}

void fluid_2d::init_tracers() {
    tm.assign(2 * (size_t)ntrace, 0.0);
    // the first call uses the draws taken at construction, later calls draw anew
    std::vector<int> r;
    r.swap(tracer_draws);
    if (r.size() != tm.size()) {
        r.resize(tm.size());
        for (int &v : r) v = rand();
    }
    for (size_t k = 0; k < tm.size(); k += 2) {
        tm[k] = ax + (bx - ax) / RAND_MAX * static_cast<double>(r[k]);
        tm[k + 1] = ay + (by - ay) / RAND_MAX * static_cast<double>(r[k + 1]);
    }
    gpu_check(vdi_set_tracers(ctx, tm.data(), ntrace), "vdi_set_tracers");
}

FILE *fluid_2d::odir_open(const char *prefix, int sn, const char *mode) const {
    const std::string path = filename + "/" + prefix + "." + std::to_string(sn);
    return safe_fopen(path.c_str(), mode);
}

bool fluid_2d::restart_exists(int sn) const {
    const std::string path = filename + "/res." + std::to_string(sn);
    struct stat st;
    return stat(path.c_str(), &st) == 0;
}

// Restart file (fluid_2d_sio.cc:97-119, gas_layer.hh:40-50,
// gas_layer_solve.cc:326-332): the time; u, v, p of every cell, ghosts
// included, row by row; the tracers; then the gas layer: frame velocity,
// liquid pressure p[me], shear tau[2..m+1], height h[m], its rate ht[m] and the
// gas pressure pg[me].  The library's state body is the same without the
// tracers.
void fluid_2d::save_restart(int sn) {
    std::vector<double> st(vdi_state_len(ctx));
    gpu_check(vdi_download_state(ctx, st.data(), (long)st.size()), "vdi_download_state");
    const size_t nf = 1 + 3 * (size_t)ml * (n + 4);
    FILE *f = odir_open("res", sn, "wb");
    fwrite(st.data(), sizeof(double), nf, f);
    if (ntrace > 0) {
        tm.resize(2 * (size_t)ntrace);
        gpu_check(vdi_get_tracers(ctx, tm.data(), ntrace), "vdi_get_tracers");
        fwrite(tm.data(), sizeof(double), tm.size(), f);
    }
    fwrite(st.data() + nf, sizeof(double), st.size() - nf, f);
    fclose(f);
}

void fluid_2d::load_restart(int sn) {
    std::vector<double> st(vdi_state_len(ctx));
    FILE *f = odir_open("res", sn, "rb");
    safe_fread(st.data(), sizeof(double), 1, f, "simulation time");
    f_num = sn;
    double *o = st.data() + 1;
    for (int j = 0; j < n + 4; j++, o += 3 * ml) safe_fread(o, sizeof(double), 3 * (size_t)ml, f, "primary fields");
    tm.assign(2 * (size_t)ntrace, 0.0);
    safe_fread(tm.data(), sizeof(double), tm.size(), f, "tracer information");
    safe_fread(o++, sizeof(double), 1, f, "Vframe");
    safe_fread(o, sizeof(double), me, f, "p array");
    o += me;
    safe_fread(o, sizeof(double), m, f, "tau array");
    o += m;
    safe_fread(o, sizeof(double), m, f, "h array");
    o += m;
    safe_fread(o, sizeof(double), m, f, "ht array");
    o += m;
    safe_fread(o, sizeof(double), me, f, "pg array");
    fclose(f);
    time = st[0];
    gpu_check(vdi_upload_state(ctx, st.data(), (long)st.size()), "vdi_upload_state");
    if (ntrace > 0) gpu_check(vdi_set_tracers(ctx, tm.data(), ntrace), "vdi_set_tracers");
}

void fluid_2d::load_last_restart() {
    if (restart_freq == 0) puts("# Restart frequency not set, starting from frame 0");
    if (restart_exists(restart_freq)) {
        int lr = restart_freq;
        while (restart_exists(lr + restart_freq)) {
            lr += restart_freq;
            if (lr >= nframes) {
                fputs("# Final restart file found; nothing to do\n", stderr);
                exit(1);
            }
        }
        printf("# Starting from restart file at frame %d\n", lr);
        load_restart(lr);
        return;
    }
    puts("# No restart file found, starting from frame 0");
    init_fields();
}

void fluid_2d::fetch_fields() {
    const size_t np = (size_t)ml * (n + 4);
    hu.resize(np);
    hv.resize(np);
    hp.resize(np);
    gpu_check(vdi_get_array(ctx, VDI_U, hu.data(), (long)np), "u");
    gpu_check(vdi_get_array(ctx, VDI_V, hv.data(), (long)np), "v");
    gpu_check(vdi_get_array(ctx, VDI_P, hp.data(), (long)np), "p");
}

void fluid_2d::write_files(int k) {
    if (fflags & (OUT_U | OUT_V | OUT_P | OUT_PB | OUT_PBFULL | OUT_W)) fetch_fields();
    if (fflags & OUT_U) output("u", 0, k, true);
    if (fflags & OUT_V) output("v", 1, k, true);
    if (fflags & OUT_P) output("p", 2, k);
    if (fflags & OUT_PB) output("pb", 3, k);
    if (fflags & OUT_PBFULL) output_pbase_full(k);
    if (fflags & OUT_W) output("w", 4, k);
    const char *dir = filename.c_str();
    if (fflags & OUT_H) gl->output_profile(dir, buf.data(), k, ay, 0);
    if (fflags & OUT_FBD) gl->output_profile(dir, buf.data(), k, ay, 1);
    if (fflags & OUT_PG) gl->output_profile(dir, buf.data(), k, ay, 2);
    if (fflags & OUT_HFULL) gl->output_profile_double(dir, buf.data(), k, 0);
    if (fflags & OUT_FBDFULL) gl->output_profile_double(dir, buf.data(), k, 1);
    if (fflags & OUT_PGFULL) gl->output_profile_double(dir, buf.data(), k, 2);
    if (ntrace > 0) output_tracers(k);
    if (restart_freq > 0 && k > 0 && k % restart_freq == 0) save_restart(k);
}

// Vorticity at node (i, j), the lower left corner of cell (i, j), from the
// centred differences of the four surrounding cells (fluid_2d.hh:128-133).
double fluid_2d::vorticity(int i, int j) const {
    auto at = [&](const std::vector<double> &a, int ci, int cj) { return a[(size_t)(cj + 2) * ml + (ci + 2)]; };
    return 0.5 * (xsp * (at(hv, i, j) - at(hv, i - 1, j) + at(hv, i, j - 1) - at(hv, i - 1, j - 1)) -
                  ysp * (at(hu, i, j) + at(hu, i - 1, j) - at(hu, i, j - 1) - at(hu, i - 1, j - 1)));
}

// Gnuplot binary matrix (fluid_2d_sio.cc:185-224): a header record with the
// column count and each column's x, then one record per row with its y and
// the values, as 4-byte floats.  Cell fields (modes 0, 1) sit at cell centres
// and, with ghost, cover the ghost cells too; the pressure and vorticity
// (modes 2, 4) at the nodes; mode 3 is the pressure's bottom row alone.
void fluid_2d::output(const char *prefix, int mode, int sn, bool ghost) {
    const bool cen = mode == 0 || mode == 1;
    const int li = ghost ? ml : (cen ? m : m + 1);
    const int lj = mode == 3 ? 1 : (ghost ? n + 4 : (cen ? n : n + 1));
    const double disp = (cen ? 0.5 : 0) - (ghost ? 2 : 0);
    const int c0 = ghost ? -2 : 0;   // index of the first row and column
    std::vector<float> rec(li + 1);
    rec[0] = (float)li;
    for (int i = 0; i < li; i++) rec[i + 1] = (float)(ax + (i + disp) * dx);
    FILE *f = odir_open(prefix, sn, "wb");
    fwrite(rec.data(), sizeof(float), li + 1, f);
    for (int j = 0; j < lj; j++) {
        rec[0] = (float)(ay + (j + disp) * dy);
        const int cj = c0 + j;
        for (int i = 0; i < li; i++) {
            const int ci = c0 + i;
            const size_t q = (size_t)(cj + 2) * ml + (ci + 2);
            double val;
            switch (mode) {
            case 0: val = hu[q]; break;
            case 1: val = hv[q]; break;
            case 4: val = vorticity(ci, cj); break;
            default: val = hp[q];
            }
            rec[i + 1] = (float)val;
        }
        fwrite(rec.data(), sizeof(float), li + 1, f);
    }
    fclose(f);
}

// The pressure of the bottom node row in double precision (fluid_2d_sio.cc:
// 168-183): the node count, ax, bx, then the values.
void fluid_2d::output_pbase_full(int sn) {
    FILE *f = odir_open("pbfull", sn, "wb");
    fwrite(&me, sizeof(int), 1, f);
    fwrite(&ax, sizeof(double), 1, f);
    fwrite(&bx, sizeof(double), 1, f);
    std::vector<double> pr(me);
    for (int i = 0; i < me; i++) pr[i] = hp[(size_t)2 * ml + (i + 2)];
    fwrite(pr.data(), sizeof(double), me, f);
    fclose(f);
}

// Tracer positions as floats, x0 y0 x1 y1 ... (fluid_2d_sio.cc:123-147).
void fluid_2d::output_tracers(int sn) {
    tm.resize(2 * (size_t)ntrace);
    gpu_check(vdi_get_tracers(ctx, tm.data(), ntrace), "vdi_get_tracers");
    std::vector<float> out(tm.begin(), tm.end());
    FILE *f = odir_open("trace", sn, "wb");
    fwrite(out.data(), sizeof(float), out.size(), f);
    fclose(f);
}
