// gp_matrix.hh -- reader of the Gnuplot binary matrix files the simulation
// writes (a header record with the column count and the columns' x, then one
// record per row with its y and the row's values, all 4-byte floats).  Stands
// in for the matrix reader of the absent utils-gp library that
// tt_analysis.cc:10 includes; values are kept in double precision.
#pragma once

#include <vector>

struct gp_matrix {
    int m = 0, n = 0;        // columns, rows
    std::vector<double> x;   // column positions
    std::vector<double> y;   // row positions
    std::vector<double> f;   // values, row after row (f[i + m j])
    double dx = 0;           // column spacing (x[1] - x[0])
    explicit gp_matrix(const char *filename);
};
