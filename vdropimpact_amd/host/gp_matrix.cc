#include "gp_matrix.hh"

#include <cstdio>
#include <cstdlib>

#include "common.hh"

gp_matrix::gp_matrix(const char *filename) {
    FILE *fp = safe_fopen(filename, "rb");
    float head;
    safe_fread(&head, sizeof(float), 1, fp, "matrix header");
    m = (int)head;
    if (m < 1) fatal_error("Matrix file has no columns", 1);
    std::vector<float> rec(m + 1);
    safe_fread(rec.data(), sizeof(float), m, fp, "matrix header");
    x.assign(rec.begin(), rec.begin() + m);
    while (fread(rec.data(), sizeof(float), m + 1, fp) == (size_t)(m + 1)) {
        y.push_back(rec[0]);
        f.insert(f.end(), rec.begin() + 1, rec.end());
        n++;
    }
    fclose(fp);
    if (n < 1) fatal_error("Matrix file has no rows", 1);
    dx = m > 1 ? x[1] - x[0] : 0;
}
