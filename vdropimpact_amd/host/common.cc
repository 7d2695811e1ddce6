#include "common.hh"

#include <chrono>
#include <cstdlib>

#include "vdi.h"

void fatal_error(const char *msg, int code) {
    fprintf(stderr, "Error: %s\n", msg);
    exit(code);
}

FILE *safe_fopen(const char *filename, const char *mode) {
    FILE *f = fopen(filename, mode);
    if (!f) {
        fprintf(stderr, "Error opening file \"%s\"\n", filename);
        exit(1);
    }
    return f;
}

void safe_fread(void *ptr, size_t size, size_t count, FILE *fp, const char *what) {
    if (fread(ptr, size, count, fp) != count) {
        fprintf(stderr, "Can't read %s from file\n", what);
        exit(1);
    }
}

double wtime() {
    using clk = std::chrono::steady_clock;
    return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

void gpu_check(int status, const char *where) {
    if (status == VDI_OK) return;
    const char *detail = status == VDI_ERR_HIP ? vdi_last_error() : "";
    fprintf(stderr, "Error: %s (%s)%s%s\n", vdi_status_string(status), where, *detail ? ": " : "", detail);
    exit(1);
}
