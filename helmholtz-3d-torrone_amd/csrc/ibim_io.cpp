// ibim_io.cpp -- the reference's text file formats, for interoperating with its
// driver (host-only code).
//
//   distance<N>.txt        (N+1)^3 values, raster [i=y][j=x][k=z], whitespace
//                          separated (read at large3Dexample2.c:405-424)
//   Rel_bN<N>k<k>.txt      "<size>\n" then one "%.15lf\n" per value, real and
//   Rel_bimgN<N>k<k>.txt   imaginary part of the RHS (print_f, :1235-1335;
//                          read back at :3108-3127)
//   Rel_DensityN<N>k<k>.txt, Rel_DensityimgN<N>k<k>.txt
//                          "%.15lf " per value, no header (:3446-3474; read at
//                          :858-866)
#include <cstdio>

#include "../../include/ibim.h"

namespace {

int read_values(FILE* f, long n, double* out) {
    for (long i = 0; i < n; ++i)
        if (fscanf(f, "%lf", &out[i]) != 1) return IBIM_ERR_SIZE;
    return IBIM_OK;
}

}  // namespace

extern "C" {

int ibim_read_distance_file(const char* path, int n, double* level) {
    FILE* f = fopen(path, "r");
    if (!f) return IBIM_ERR_ARG;
    const long total = (long)(n + 1) * (n + 1) * (n + 1);
    int rc = read_values(f, total, level);
    fclose(f);
    return rc;
}

int ibim_write_distance_file(const char* path, int n, const double* level) {
    FILE* f = fopen(path, "w");
    if (!f) return IBIM_ERR_ARG;
    const long S = n + 1;
    for (long i = 0; i < S; ++i)
        for (long j = 0; j < S; ++j) {
            for (long k = 0; k < S; ++k) fprintf(f, "%.17g ", level[(i * S + j) * S + k]);
            fputc('\n', f);
        }
    return fclose(f) == 0 ? IBIM_OK : IBIM_ERR_ARG;
}

int ibim_write_rhs_text(const char* path_re, const char* path_im, int count, const double* bre, const double* bim) {
    const char* paths[2] = {path_re, path_im};
    const double* vals[2] = {bre, bim};
    for (int q = 0; q < 2; ++q) {
        FILE* f = fopen(paths[q], "w");
        if (!f) return IBIM_ERR_ARG;
        fprintf(f, "%d\n", count);
        for (int i = 0; i < count; ++i) fprintf(f, "%.15lf\n", vals[q][i]);
        if (fclose(f)) return IBIM_ERR_ARG;
    }
    return IBIM_OK;
}

int ibim_read_rhs_text(const char* path_re, const char* path_im, int count, double* bre, double* bim) {
    const char* paths[2] = {path_re, path_im};
    double* vals[2] = {bre, bim};
    for (int q = 0; q < 2; ++q) {
        FILE* f = fopen(paths[q], "r");
        if (!f) return IBIM_ERR_ARG;
        int size = -1;
        int rc = fscanf(f, "%d", &size) == 1 && size == count ? read_values(f, count, vals[q]) : IBIM_ERR_SIZE;
        fclose(f);
        if (rc) return rc;
    }
    return IBIM_OK;
}

int ibim_write_density_text(const char* path_re, const char* path_im, int count, const double* re, const double* im) {
    const char* paths[2] = {path_re, path_im};
    const double* vals[2] = {re, im};
    for (int q = 0; q < 2; ++q) {
        FILE* f = fopen(paths[q], "w");
        if (!f) return IBIM_ERR_ARG;
        for (int i = 0; i < count; ++i) fprintf(f, "%.15lf ", vals[q][i]);
        if (fclose(f)) return IBIM_ERR_ARG;
    }
    return IBIM_OK;
}

int ibim_read_density_text(const char* path_re, const char* path_im, int count, double* re, double* im) {
    const char* paths[2] = {path_re, path_im};
    double* vals[2] = {re, im};
    for (int q = 0; q < 2; ++q) {
        FILE* f = fopen(paths[q], "r");
        if (!f) return IBIM_ERR_ARG;
        int rc = read_values(f, count, vals[q]);
        fclose(f);
        if (rc) return rc;
    }
    return IBIM_OK;
}

}  // extern "C"
