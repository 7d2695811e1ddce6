// tt_analysis -- lift-off analysis of a finished simulation (the reference's
// tt_analysis.cc:59-172): for every frame's height profile, the lowest point
// of the interface (after Gaussian smoothing and with a cubic Hermite
// interpolant between grid points, or on the raw grid values with -r) is
// written to <name>.tta (.ttb with -r) as "frame time x_min h_min"; then the
// minimum over the approach (the search stops once the height has risen 0.5 %
// above its minimum) gives the lift-off line: viscosity in cSt, minimum
// height, its time from the nominal contact time h0/V and from the time the
// drop crosses the evanescent-field height h_ev, and that frame as a
// percentage of the run.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "common.hh"
#include "fileinfo.hh"
#include "gp_matrix.hh"

namespace {

const int files_max = 16384;
const float h_ev = 500e-9;       // height of the evanescent field (m)
const float glen = 1.2e-6;       // length scale of the smoothing Gaussian (m)
const float gaussian_fac = 7.;   // standard deviations in the smoothing window

// Cubic Hermite interpolant on [0, 1] with end values t3 (at 0) and t0 (at 1)
// and end slopes t2 (at 0) and t1 (at 1).
double hermite(double x, double t0, double t1, double t2, double t3) {
    return t0 * x * x * (3 - 2 * x) + x * (1 - x) * (-t1 * (1 - x) + t2 * x) + t3 * (2 * x * x * x - 3 * x * x + 1);
}

// Gaussian smoothing of the first row, mirrored at both ends; q gets m values.
void smooth(const gp_matrix &gp, double *q) {
    const int m = gp.m;
    const float sls = glen / gp.dx, fac = 0.5 / (sls * sls);
    const int cut = int(sls * gaussian_fac) + 1;
    std::vector<double> w(2 * cut + 1);
    double *wc = w.data() + cut;
    float tot = 0.5;
    for (int i = 1; i <= cut; i++) tot += wc[i] = std::exp(-fac * i * i);
    tot = 0.5 / tot;
    wc[0] = tot;
    for (int i = 1; i <= cut; i++) wc[-i] = wc[i] *= tot;
    for (int i = 0; i < m; i++) {
        float s = 0;
        for (int j = i - cut; j <= i + cut; j++) s += wc[j - i] * gp.f[j < 0 ? -1 - j : (j >= m ? 2 * m - 1 - j : j)];
        q[i] = s;
    }
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2 || (argc > 3 && strcmp(argv[1], "-r") != 0)) {
        fputs("Syntax: ./tt_analysis [-r] <output_dir>\n\n"
              "The output directory should have a '.odr' suffix. The '-r' flag uses the\n"
              "raw values without smoothing and cubic interpolation.\n",
              stderr);
        return 1;
    }
    const bool raw = argc == 3;
    const char *fn = argv[argc - 1];
    const size_t l = strlen(fn);
    if (l < 4) fatal_error("Filename is too short", 1);
    if (l > 224) fatal_error("Filename is too long", 1);
    if (strcmp(fn + l - 4, ".odr") != 0) fatal_error("Filename must end in '.odr'", 1);
    const std::string stem(fn, l - 3);
    fileinfo fi((stem + "cfg").c_str());
    const gas_layer &gls = *fi.gl;
    const double fdt = fi.t_end / fi.nframes, tmin0 = gls.h0 / gls.V;
    double tmin0_ev = 0;
    FILE *of = safe_fopen((stem + (raw ? "ttb" : "tta")).c_str(), "w");

    int gimin = -1, m = 0, i;
    double gfmin = DBL_MAX;
    std::vector<double> q;
    bool efield = true, hsearch = true;
    for (i = 0; i <= files_max; i++) {
        const std::string hname = std::string(fn) + "/height." + std::to_string(i);
        struct stat st;
        if (stat(hname.c_str(), &st) != 0) break;
        gp_matrix gp(hname.c_str());
        if (i == 0) m = gp.m;
        else if (m != gp.m) fatal_error("Grids change size", 1);
        // one value past the end continues the last (the last interval's
        // right slope then uses a one-sided difference)
        q.assign(m + 1, 0.0);
        if (raw) for (int j = 0; j < m; j++) q[j] = gp.f[j];
        else smooth(gp, q.data());
        q[m] = q[m - 1];

        double fmin = gp.f[0], xmin = gp.x[0];
        for (int k = 0; k < m - 1; k++) {
            const double *f = q.data() + k;
            const double t0 = f[1], t1 = 0.5 * (f[2] - f[0]), t2 = k == 0 ? f[1] - f[0] : 0.5 * (f[1] - f[-1]),
                         t3 = f[0];
            const double a = 6 * (t3 - t0) + 3 * (t1 + t2), b = 6 * t0 - 2 * t1 - 4 * t2 - 6 * t3, c = t2;
            double det = b * b - 4 * a * c;
            if (f[1] < fmin) fmin = f[1], xmin = gp.x[k + 1];
            // interior stationary points of the cubic on this interval
            if (!raw && det >= 0) {
                det = std::sqrt(det);
                for (double xs : {0.5 * (-b + det), 0.5 * (-b - det)}) {
                    if (0 < xs && xs < a) {
                        xs /= a;
                        const double fv = hermite(xs, t0, t1, t2, t3);
                        if (fv < fmin) {
                            fmin = fv;
                            xmin = (1 - xs) * gp.x[k] + xs * gp.x[k + 1];
                        }
                    }
                }
            }
        }
        if (hsearch && fmin < gfmin) {
            gfmin = fmin;
            gimin = i;
            // time origin of Kolinski et al.: when the drop becomes visible
            // at the evanescent-field height
            if (efield && fmin < h_ev) {
                tmin0_ev = i * fdt - xmin * xmin / (2 * gls.R * gls.V) + h_ev / gls.V;
                efield = false;
            }
        }
        fprintf(of, "%d %g %.10g %.10g\n", i, i * fdt, xmin, fmin);
        if (fmin > gfmin * 1.005) hsearch = false;
    }
    fclose(of);
    const double gtmin = gimin * fdt;
    printf("# %s: %d files processed\n%g %g %g %g %g\n", fn, i, fi.mu * fi.rhoinv * 1e6, gfmin, gtmin - tmin0,
           gtmin - tmin0_ev, 100. * float(gimin) / float(i - 1));
    return 0;
}
