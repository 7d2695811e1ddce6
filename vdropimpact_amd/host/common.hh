// common.hh -- fail-stop helpers of the host layer.
//
// The reference stops on every error with a message and exit status
// (fatal_error, safe_fopen, safe_fread in common.cc:11-47); the host layer
// keeps that behaviour and those messages, so scripts that watch the
// reference's output read this one's the same way.
#pragma once

#include <cstddef>
#include <cstdio>

void fatal_error(const char *msg, int code);
FILE *safe_fopen(const char *filename, const char *mode);
void safe_fread(void *ptr, size_t size, size_t count, FILE *fp, const char *what);
// Wall-clock seconds (the reference's wtime, common.hh:15-23).
double wtime();
// Exits with the library's message when a C-ABI call fails.
void gpu_check(int status, const char *where);
