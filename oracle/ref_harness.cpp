// Test-only harness around the reference build (oracle/_ref/libibimref.so).
//
// This file is test infrastructure for the oracle: it is compiled together with
// the reference's own C sources (compiled in place from the reference checkout by
// oracle/Makefile, target `ref`) and exposes plain extern "C" entry points that
// tests/golden/make_golden.py drives through ctypes to write golden fixtures.
// Nothing in the product library links or loads it.
//
// The reference keeps its state in globals defined by its driver
// (large3Dexample2.c:48-74, compiled here with main renamed), so this harness
// only declares them and fills them in before each call.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <unistd.h>
#include <fcntl.h>

#include "BICGSTAB.h"
#include "HelmholtzKernel.h"
#include "distance_application3D.h"
#include "interpolate3D.h"
#include "weno.h"
#include "functions.h"

// Globals defined by large3Dexample2.c:46-74.
extern int N, SIZE;
extern double H, HCUBED, WAVE, G_DELTA;
extern double ***DIST, ***DIST_DX, ***DIST_DY, ***DIST_DZ, ***MCURV, ***GCURV;
extern double *Density, *Densityimg;

// Driver functions with external linkage (prototypes at large3Dexample2.c:76-95).
double omp_get_Ccomp(int size, int pi, int pj, double epsilon, double tempC[]);
int print_f(int size, char *filename, char *filenameimg, double delta, double epsilon, int mode);

namespace {

double ***alloc_grid(int s) {
    double ***g = (double ***)malloc(sizeof(double **) * s);
    for (int i = 0; i < s; ++i) {
        g[i] = (double **)malloc(sizeof(double *) * s);
        for (int j = 0; j < s; ++j) g[i][j] = (double *)calloc(s, sizeof(double));
    }
    return g;
}

void to_flat(double ***g, int s, double *out) {
    for (int i = 0; i < s; ++i)
        for (int j = 0; j < s; ++j) memcpy(out + ((size_t)i * s + j) * s, g[i][j], sizeof(double) * s);
}

void from_flat(double ***g, int s, const double *in) {
    for (int i = 0; i < s; ++i)
        for (int j = 0; j < s; ++j) memcpy(g[i][j], in + ((size_t)i * s + j) * s, sizeof(double) * s);
}

// stdout capture: the reference reports iteration counts only through printf.
int g_saved_fd = -1;
char g_cap_path[256];

void capture_begin() {
    fflush(stdout);
    snprintf(g_cap_path, sizeof g_cap_path, "/tmp/ibimref_cap_%d.txt", (int)getpid());
    int fd = open(g_cap_path, O_CREAT | O_TRUNC | O_WRONLY, 0600);
    g_saved_fd = dup(1);
    dup2(fd, 1);
    close(fd);
}

void capture_end(char *buf, int cap) {
    fflush(stdout);
    dup2(g_saved_fd, 1);
    close(g_saved_fd);
    FILE *f = fopen(g_cap_path, "r");
    int n = 0;
    if (f) {
        n = (int)fread(buf, 1, cap - 1, f);
        fclose(f);
    }
    buf[n > 0 ? n : 0] = 0;
    unlink(g_cap_path);
}

double **rows_of(double *flat, int n) {
    double **r = (double **)malloc(sizeof(double *) * n);
    for (int i = 0; i < n; ++i) r[i] = flat + (size_t)i * n;
    return r;
}

}  // namespace

extern "C" {

// Grid setup: N cells per axis on [-2,2]^3 (frame.h:6-7), wavenumber k.
void ref_setup(int n, double k, const double *level) {
    N = n;
    SIZE = n + 1;
    H = (INT_U - INT_L) / (double)n;  // large3Dexample2.c:174
    HCUBED = H * H * H;
    G_DELTA = H;
    WAVE = k;
    DIST = alloc_grid(SIZE);
    DIST_DX = alloc_grid(SIZE);
    DIST_DY = alloc_grid(SIZE);
    DIST_DZ = alloc_grid(SIZE);
    MCURV = alloc_grid(SIZE);
    GCURV = alloc_grid(SIZE);
    from_flat(DIST, SIZE, level);
}

void ref_set_gradient(const double *gx, const double *gy, const double *gz) {
    from_flat(DIST_DX, SIZE, gx);
    from_flat(DIST_DY, SIZE, gy);
    from_flat(DIST_DZ, SIZE, gz);
}

void ref_set_curvatures(const double *mc, const double *gc) {
    from_flat(MCURV, SIZE, mc);
    from_flat(GCURV, SIZE, gc);
}

// ASSIGN_GRADIENT then BothCurvs on every node (large3Dexample2.c:443-455).
void ref_geometry(int weno_order, double *gx, double *gy, double *gz, double *mc, double *gc) {
    ASSIGN_GRADIENT(N, H, DIST, DIST_DX, DIST_DY, DIST_DZ, weno_order);
    for (int i = 0; i < SIZE; ++i)
        for (int j = 0; j < SIZE; ++j)
            for (int k = 0; k < SIZE; ++k) BothCurvs(DIST, i, j, k, N, H, &MCURV[i][j][k], &GCURV[i][j][k]);
    to_flat(DIST_DX, SIZE, gx);
    to_flat(DIST_DY, SIZE, gy);
    to_flat(DIST_DZ, SIZE, gz);
    to_flat(MCURV, SIZE, mc);
    to_flat(GCURV, SIZE, gc);
}

// Single-node second derivatives: out = {xx, yy, zz, xy, xz, yz}.
void ref_second_derivs(int i, int j, int k, double *out) {
    out[0] = XDER2(DIST, i, j, k, N, H);
    out[1] = YDER2(DIST, i, j, k, N, H);
    out[2] = ZDER2(DIST, i, j, k, N, H);
    out[3] = secondpartialxy(DIST, i, j, k, N, H);
    out[4] = secondpartialxz(DIST, i, j, k, N, H);
    out[5] = secondpartialyz(DIST, i, j, k, N, H);
}

void ref_weno(const double *f, double dx, int order, double *out2) {
    double fi[7];
    memcpy(fi, f, sizeof(double) * (order == 5 ? 7 : 5));
    weno(&out2[0], &out2[1], fi, dx, order);
}

void ref_weno_axis(int axis, int i, int j, int k, int order, double *out2) {
    if (axis == 0) WENO_DX(DIST, i, j, k, N, H, &out2[0], &out2[1], order);
    else if (axis == 1) WENO_DY(DIST, i, j, k, N, H, &out2[0], &out2[1], order);
    else WENO_DZ(DIST, i, j, k, N, H, &out2[0], &out2[1], order);
}

int ref_interpgrad(double x, double y, double z, double *out3) {
    return cal_interpgrad(N, H, x, y, z, DIST_DX, DIST_DY, DIST_DZ, &out3[0], &out3[1], &out3[2]);
}

void ref_bothcurvs_at(double x, double y, double z, double *out2) {
    cal_BothCurvs(x, y, z, H, MCURV, GCURV, &out2[0], &out2[1]);
}

double ref_cal_dist(double x, double y, double z) { return cal_dist(DIST, x, y, z, N, H); }
double ref_delta(double eps, double x) { return cal_delta(eps, x); }
double ref_sine_weight(double t, double tau) { return SineWeight(t, tau); }
double ref_poly_weight(double t, double tau) { return PolyWeight(t, tau); }
double ref_cal_phi(const double *x, const double *y) { return cal_phi(x[0], x[1], x[2], y[0], y[1], y[2]); }
double ref_phiimg(const double *x, const double *y) { return phiimg(x[0], x[1], x[2], y[0], y[1], y[2]); }

// Traditional kernel entries through omp_get_Ccomp (large3Dexample2.c:4241-4576).
// The first call builds the tube from the globals, so one process serves one grid.
void ref_trad_entries(int size, double eps, int npairs, const int *pi, const int *pj, double *re, double *im) {
    double t[2];
    for (int p = 0; p < npairs; ++p) {
        omp_get_Ccomp(size, pi[p], pj[p], eps, t);
        re[p] = t[0];
        im[p] = t[1];
    }
}

// Traditional tube arrays through DoubleLayerKernel (HelmholtzKernel.c:613-761),
// which builds the same |d|<=eps tube; the size x size FD matrix is written too.
int ref_double_layer(int size, double eps, double *zx, double *zy, double *zz, double *sx, double *sy, double *sz,
                     double *dist, double *gx, double *gy, double *gz, double *delta, double *jac, double *reg,
                     double *regimg, double *result, double *Kre, double *Kim) {
    double **kr = rows_of(Kre, size), **ki = rows_of(Kim, size);
    int rc = DoubleLayerKernel(size, eps, DIST, DIST_DX, DIST_DY, DIST_DZ, MCURV, GCURV, zx, zy, zz, sx, sy, sz, dist,
                               gx, gy, gz, delta, jac, reg, regimg, result, kr, ki, WAVE);
    free(kr);
    free(ki);
    return rc;
}

// Combined-field kernel (HelmholtzKernel.c:173-610).
int ref_combo(int size, double delta, double eps, double wavenum, double eta, double *zx, double *zy, double *zz,
              double *sx, double *sy, double *sz, double *dist, double *gx, double *gy, double *gz, double *jac,
              double *reg, double *regimg, double *sinew, double *siner, double *Kre, double *Kim) {
    double **kr = rows_of(Kre, size), **ki = rows_of(Kim, size);
    char buf[4096];
    capture_begin();
    int rc = new_HelmholtzKernelCombo(size, delta, eps, DIST, DIST_DX, DIST_DY, DIST_DZ, MCURV, GCURV, zx, zy, zz, sx,
                                      sy, sz, dist, gx, gy, gz, jac, reg, regimg, nullptr, sinew, nullptr, siner,
                                      nullptr, nullptr, kr, ki, wavenum, eta);
    capture_end(buf, sizeof buf);
    free(kr);
    free(ki);
    return rc;
}

// Sphere-only analytic kernel (HelmholtzKernel.c:4-169).
int ref_new_kernel(int size, double delta, double eps, double *zx, double *zy, double *zz, double *sx, double *sy,
                   double *sz, double *dist, double *gx, double *gy, double *gz, double *curv, double *jac,
                   double *sinew, double *siner, double *Kre, double *Kim) {
    double **kr = rows_of(Kre, size), **ki = rows_of(Kim, size);
    int rc = new_HelmholtzKernel(size, delta, eps, zx, zy, zz, sx, sy, sz, dist, gx, gy, gz, curv, jac, nullptr, sinew,
                                 nullptr, siner, nullptr, nullptr, kr, ki);
    free(kr);
    free(ki);
    return rc;
}

// Complex BiCG (BICGSTAB.c:5-339 serial, 344-765 OpenMP). Returns the iteration
// count parsed from the reference's own report; *err receives its final error.
int ref_compbicg(int size, double *Are, double *Aim, double *b, double *bimg, double *x, double *ximg, int use_omp,
                 double *err) {
    double **kr = rows_of(Are, size), **ki = rows_of(Aim, size);
    char buf[1 << 14];
    capture_begin();
    if (use_omp) omp_CompBiCG(size, x, ximg, kr, ki, b, bimg, 16);
    else CompBiCG(size, x, ximg, kr, ki, b, bimg);
    capture_end(buf, sizeof buf);
    free(kr);
    free(ki);
    int runs = -1;
    *err = NAN;
    const char *p = strstr(buf, "After ");
    if (p) sscanf(p, "After %d runs, the error is %lf", &runs, err);
    if (strstr(buf, "Zero BC")) runs = 0;
    return runs;
}

// RHS through print_f (large3Dexample2.c:1217-1352) into two %.15lf text files
// written in the current directory; mode 0 = traditional tube.
int ref_print_f(int size, double eps, const char *fre, const char *fim) {
    char a[256], b[256];
    snprintf(a, sizeof a, "%s", fre);
    snprintf(b, sizeof b, "%s", fim);
    char buf[1 << 14];
    capture_begin();
    int rc = print_f(size, a, b, -1.0, eps, 0);
    capture_end(buf, sizeof buf);
    return rc;
}

}  // extern "C"
