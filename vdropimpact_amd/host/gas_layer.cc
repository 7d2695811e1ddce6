// gas_layer.cc -- host face of the device gas layer: the update call and the
// profile writers (gas_layer_solve.cc:258-308 formats).
#include "gas_layer.hh"

#include <cstring>
#include <string>
#include <vector>

#include "common.hh"

gas_layer::gas_layer(int m_, double ax_, double bx_, const gl_params &p, bool x_sym_)
    : gl_params(p), m(m_), me(m_ + 1), ml(m_ + 4), ax(ax_), bx(bx_), dx((bx_ - ax_) / m_), x_sym(x_sym_) {}

void gas_layer::update(double t, double dt) { gpu_check(vdi_gas_layer_update(ctx, t, dt), "gas_layer::update"); }

// fld 0: h (cells), 1: fbd (cells), 2: pg (nodes)
void gas_layer::fetch(int fld, double *out, int *len) {
    const int id = fld == 0 ? VDI_H : fld == 1 ? VDI_FBD : VDI_PG;
    *len = fld == 2 ? me : m;
    gpu_check(vdi_get_array(ctx, id, out, *len), "gas layer profile");
}

static std::string profile_name(const char *dir, const char *stem, int sn) {
    return std::string(dir) + "/" + stem + "." + std::to_string(sn);
}

void gas_layer::output_profile(const char *filename, float *buf, int sn, double ay, int fld) {
    (void)buf;
    static const char *stems[] = {"height", "fbd", "pg"};
    std::vector<double> val(me);
    int l = 0;
    fetch(fld, val.data(), &l);
    // Gnuplot binary matrix: a header record (count, then the x of each
    // column: cell centres, or nodes for the pressure), then one record per
    // row (its y, then the values), all as 4-byte floats
    std::vector<float> rec(l + 1);
    rec[0] = (float)l;
    const double off = fld == 2 ? 0 : 0.5;
    for (int i = 0; i < l; i++) rec[i + 1] = (float)(ax + (i + off) * dx);
    FILE *f = safe_fopen(profile_name(filename, stems[fld], sn).c_str(), "wb");
    fwrite(rec.data(), sizeof(float), l + 1, f);
    rec[0] = (float)ay;
    for (int i = 0; i < l; i++) rec[i + 1] = (float)val[i];
    fwrite(rec.data(), sizeof(float), l + 1, f);
    fclose(f);
}

void gas_layer::output_profile_double(const char *filename, float *buf, int sn, int fld) {
    (void)buf;
    static const char *stems[] = {"hfull", "fbdfull", "pgfull"};
    std::vector<double> val(me);
    int l = 0;
    fetch(fld, val.data(), &l);
    // the count, the x range [ax, bx] and the values in double precision
    FILE *f = safe_fopen(profile_name(filename, stems[fld], sn).c_str(), "wb");
    const double range[2] = {ax, bx};
    fwrite(&l, sizeof(int), 1, f);
    fwrite(range, sizeof(double), 2, f);
    fwrite(val.data(), sizeof(double), l, f);
    fclose(f);
}
