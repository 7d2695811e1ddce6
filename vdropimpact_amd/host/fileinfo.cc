// fileinfo.cc -- reading a .cfg file and deriving the simulation constants
// (fileinfo.cc:11-193 of the reference: same keywords, checks and messages).
#include "fileinfo.hh"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "common.hh"

namespace {

// The words of one line after its comment, split at spaces, tabs and line
// ends; reading past the last word or leaving words over stops the program
// with the reference's messages (fileinfo.cc:158-176).
class Words {
  public:
    Words(const char *line, int ln) : ln_(ln) {
        std::string s(line);
        const size_t hash = s.find('#');
        if (hash != std::string::npos) s.resize(hash);
        size_t i = 0;
        while (i < s.size()) {
            while (i < s.size() && is_sep(s[i])) i++;
            size_t j = i;
            while (j < s.size() && !is_sep(s[j])) j++;
            if (j > i) w_.push_back(s.substr(i, j - i));
            i = j;
        }
    }
    bool empty() const { return w_.empty(); }
    const std::string &key() const { return w_[0]; }
    bool more() const { return next_ < w_.size(); }
    const char *next() {
        if (!more()) {
            fprintf(stderr, "Not enough arguments at input line %d\n", ln_);
            exit(1);
        }
        return w_[next_++].c_str();
    }
    void done() const {
        if (more()) {
            fprintf(stderr, "Too many arguments at input line %d\n", ln_);
            exit(1);
        }
    }
    double last_double() {
        const double v = atof(next());
        done();
        return v;
    }
    int last_int() {
        const int v = atoi(next());
        done();
        return v;
    }

  private:
    static bool is_sep(char c) { return c == ' ' || c == '\t' || c == '\n'; }
    std::vector<std::string> w_;
    size_t next_ = 1;
    int ln_;
};

void check_invalid(double val, const char *name) {
    if (val < 0) {
        fprintf(stderr, "Value of %s either invalid or not set\n", name);
        exit(1);
    }
}

}  // namespace

fileinfo::fileinfo(const char *infile) {
    double L_nd = -1, h0_nd = -1, t_end_nd = -1, nul = -1, R = -1, V = -1, alpha = -1, sigma = -1, mug = -1,
           Pamb = -1;
    int gas_mode = 0;  // 0 unset, 1 gas_layer_data, 2 gas_layer_model

    // the output directory is the file name with .cfg replaced by .odr
    const size_t l = strlen(infile);
    if (l < 4) fatal_error("Filename is too short", 1);
    if (strcmp(infile + l - 4, ".cfg") != 0) fatal_error("Filename must end in '.cfg'", 1);
    filename.assign(infile, l - 3);
    filename += "odr";

    FILE *f = safe_fopen(infile, "r");
    std::vector<char> buf(512);
    for (int ln = 1; fgets(buf.data(), (int)buf.size(), f); ln++) {
        Words w(buf.data(), ln);
        if (w.empty()) continue;
        const std::string &k = w.key();
        if (k == "grid_points") {
            m = atoi(w.next());
            n = w.last_int();
            if (m <= 0 || n <= 0) fatal_error("Grid dimensions must be positive", 1);
        } else if (k == "L_nd") L_nd = w.last_double();
        else if (k == "h0_nd") h0_nd = w.last_double();
        else if (k == "t_end_nd") t_end_nd = w.last_double();
        else if (k == "tracers") ntrace = w.last_int();
        else if (k == "frames") nframes = w.last_int();
        else if (k == "tmult") tmult = w.last_double();
        else if (k == "restart_freq") restart_freq = w.last_int();
        else if (k == "nul") nul = w.last_double();
        else if (k == "nul_cSt") nul = 1e-6 * w.last_double();
        else if (k == "rhol") rho = w.last_double();
        else if (k == "R") R = w.last_double();
        else if (k == "V") V = w.last_double();
        else if (k == "gamma") alpha = 1. / w.last_double();
        else if (k == "alpha") alpha = w.last_double();
        else if (k == "sigma") sigma = w.last_double();
        else if (k == "mug") mug = w.last_double();
        else if (k == "Pamb") Pamb = w.last_double();
        else if (k == "x_sym") { x_sym = true; w.done(); }
        else if (k == "implicit_visc") { implicit_visc = true; w.done(); }
        else if (k == "mr_time_output") { mr_time_output = true; w.done(); }
        else if (k == "gas_layer_data") {
            gas_mode = 1;
            for (int q = 0; q < 4; q++) (void)w.next();   // m_dat n_dat pressure-file tau-file
            w.done();
        } else if (k == "gas_layer_model") { gas_mode = 2; w.done(); }
        else if (k == "nif_center") { nif_mode = 1; w.done(); }
        else if (k == "nif_range") { nif_mode = 2; (void)w.last_double(); }
        else if (k == "output") {
            static const struct { const char *name; unsigned bit; } kinds[] = {
                {"u", OUT_U}, {"v", OUT_V}, {"p", OUT_P}, {"pb", OUT_PB}, {"pbfull", OUT_PBFULL}, {"w", OUT_W},
                {"h", OUT_H}, {"fbd", OUT_FBD}, {"pg", OUT_PG}, {"hfull", OUT_HFULL}, {"fbdfull", OUT_FBDFULL},
                {"pgfull", OUT_PGFULL}};
            fflags = 0;
            const char *t = w.next();
            for (;;) {
                bool known = false;
                for (const auto &kd : kinds)
                    if (strcmp(t, kd.name) == 0) { fflags |= kd.bit; known = true; }
                if (!known) fatal_error("Output type not understood", 1);
                if (!w.more()) break;
                t = w.next();
            }
        } else {
            fprintf(stderr, "Keyword '%s' not understood on line %d\n", k.c_str(), ln);
            exit(1);
        }
    }
    fclose(f);

    if (m == -1) fatal_error("Grid dimensions not set", 1);
    check_invalid(L_nd, "L_nd");
    check_invalid(h0_nd, "h0_nd");
    check_invalid(t_end_nd, "t_end_nd");
    check_invalid(tmult, "tmult");
    check_invalid(nframes, "frames");
    check_invalid(restart_freq, "restart_freq");
    check_invalid(nul, "nul");
    check_invalid(rho, "rhol");
    mu = nul * rho;
    muinv = 1. / mu;
    rhoinv = 1. / rho;

    // Scalings of the lubrication problem (fileinfo.cc:141-147): with the
    // Stokes number st = mu_g / (rho V R), lengths across the layer scale as
    // R st^(2/3), along it as R st^(1/3), and times as R st^(2/3) / V.
    const double st = mug / (rho * V * R), st3 = std::pow(st, 1. / 3.), st23 = st3 * st3;
    const double L = L_nd * R * st3, h0 = h0_nd * R * st23;
    t_end = t_end_nd * R * st23 / V;
    bx = L;
    by = L * n / m;
    if (x_sym) ax = 0;
    else {
        ax = -L;
        by *= 2;
    }

    if (gas_mode == 0) fatal_error("Either gas_layer_data or gas_layer_model must be set", 1);
    if (gas_mode == 1) fatal_error("gas_layer_data is not available with the GPU step; use gas_layer_model", 1);
    check_invalid(alpha, "alpha");
    check_invalid(mug, "mug");
    check_invalid(Pamb, "Pamb");
    check_invalid(R, "R");
    check_invalid(V, "V");
    check_invalid(sigma, "sigma");
    if (nif_mode != 0) fatal_error("The non-inertial frame is not available with the GPU step", 1);
    gl_params glp{alpha, mug, Pamb, h0, R, V, sigma};
    gl = new gas_layer(m, ax, bx, glp, x_sym);
}

fileinfo::~fileinfo() { delete gl; }
