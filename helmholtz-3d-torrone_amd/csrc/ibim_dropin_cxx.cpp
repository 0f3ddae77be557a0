// ibim_dropin_cxx.cpp -- C++-linkage definitions of the reference's entry
// points. The reference compiles its .c files as C++ (Makefile:1,10-11) and its
// headers carry no extern "C", so a program built from them references mangled
// names (e.g. _Z8CompBiCGiPdS_PS_S0_S_S_). The prototypes below repeat the
// reference headers' argument lists exactly (BICGSTAB.h:15-17,
// HelmholtzKernel.h:20-39, weno.h:26-31, distance_application3D.h:9-12) so the
// mangled names match; each forwards to the shared implementation.
#include "ibim_dropin.hpp"

using namespace ibim;

void weno(double *dfp, double *dfn, double fi[], const double dx, const int order) {
    dropin_weno(dfp, dfn, fi, dx, order);
}
void WENO_DX(double ***uvalue, const int i, const int j, const int k, const int N, const double H, double *ux_p,
             double *ux_n, const int order) {
    dropin_weno_axis(0, uvalue, i, j, k, N, H, ux_p, ux_n, order);
}
void WENO_DY(double ***uvalue, const int i, const int j, const int k, const int N, const double H, double *ux_p,
             double *ux_n, const int order) {
    dropin_weno_axis(1, uvalue, i, j, k, N, H, ux_p, ux_n, order);
}
void WENO_DZ(double ***uvalue, const int i, const int j, const int k, const int N, const double H, double *ux_p,
             double *ux_n, const int order) {
    dropin_weno_axis(2, uvalue, i, j, k, N, H, ux_p, ux_n, order);
}

double GaussianCurv(double ***dist, int i, int j, int k, int N, double H) {
    double d[6], m, g;
    dropin_second(dist, i, j, k, N, H, d);
    dropin_curvatures(d, &m, &g);
    return g;
}
double MeanCurv(double ***dist, int i, int j, int k, int N, double H) {
    double d[6], m, g;
    dropin_second(dist, i, j, k, N, H, d);
    dropin_curvatures(d, &m, &g);
    return m;
}
int BothCurvs(double ***dist, int i, int j, int k, int N, double H, double *meancurv, double *gausscurv) {
    double d[6];
    dropin_second(dist, i, j, k, N, H, d);
    dropin_curvatures(d, meancurv, gausscurv);
    return 0;
}
int ASSIGN_GRADIENT(int N, double H, double ***dist, double ***graddx, double ***graddy, double ***graddz,
                    int weno_mode) {
    return dropin_assign_gradient(N, H, dist, graddx, graddy, graddz, weno_mode);
}

int CompBiCG(int size, double *density, double *densityimg, double **Kernel, double **Kernelimg, double *b,
             double *bimg) {
    return dropin_compbicg(size, density, densityimg, Kernel, Kernelimg, b, bimg, 1000);
}
int omp_CompBiCG(int size, double *density, double *densityimg, double **Kernel, double **Kernelimg, double *b,
                 double *bimg, int chunk) {
    (void)chunk;
    return dropin_compbicg(size, density, densityimg, Kernel, Kernelimg, b, bimg, 2000);
}

int new_HelmholtzKernel(int size, double delta, double epsilon, double *zxlist, double *zylist, double *zzlist,
                        double *zxstarlist, double *zystarlist, double *zzstarlist, double *dist, double *gradx,
                        double *grady, double *gradz, double *curv, double *Jacobian, double *PolyW, double *SineW,
                        double *PolyResult, double *SineResult, double **PolyKernelH, double **PolyKernelHimg,
                        double **SineKernelH, double **SineKernelHimg) {
    (void)PolyW;
    (void)PolyResult;
    (void)PolyKernelH;
    (void)PolyKernelHimg;
    return dropin_new_kernel(size, delta, epsilon, zxlist, zylist, zzlist, zxstarlist, zystarlist, zzstarlist, dist,
                             gradx, grady, gradz, curv, Jacobian, SineW, SineResult, SineKernelH, SineKernelHimg);
}

int new_HelmholtzKernelCombo(int size, double delta, double epsilon, double ***level, double ***leveldx,
                             double ***leveldy, double ***leveldz, double ***levelmcurv, double ***levelgcurv,
                             double *zxlist, double *zylist, double *zzlist, double *zxstarlist, double *zystarlist,
                             double *zzstarlist, double *dist, double *gradx, double *grady, double *gradz,
                             double *Jacobian, double *regfactor, double *regfactorimg, double *PolyW, double *SineW,
                             double *PolyResult, double *SineResult, double **PolyKernelH, double **PolyKernelHimg,
                             double **SineKernelH, double **SineKernelHimg, double wavenum, double eta) {
    (void)PolyW;
    (void)PolyResult;
    (void)PolyKernelH;
    (void)PolyKernelHimg;
    return dropin_combo(size, delta, epsilon, level, leveldx, leveldy, leveldz, levelmcurv, levelgcurv, zxlist,
                        zylist, zzlist, zxstarlist, zystarlist, zzstarlist, dist, gradx, grady, gradz, Jacobian,
                        regfactor, regfactorimg, SineW, SineResult, SineKernelH, SineKernelHimg, wavenum, eta);
}

int DoubleLayerKernel(int size, double epsilon, double ***level, double ***leveldx, double ***leveldy,
                      double ***leveldz, double ***levelmcurv, double ***levelgcurv, double *zxlist, double *zylist,
                      double *zzlist, double *zxstarlist, double *zystarlist, double *zzstarlist, double *dist,
                      double *gradx, double *grady, double *gradz, double *delta, double *Jacobian, double *regfactor,
                      double *regfactorimg, double *result, double **Kernel, double **Kernelimg, double wavenum) {
    (void)wavenum;
    return dropin_double_layer(size, epsilon, level, leveldx, leveldy, leveldz, levelmcurv, levelgcurv, zxlist,
                               zylist, zzlist, zxstarlist, zystarlist, zzstarlist, dist, gradx, grady, gradz, delta,
                               Jacobian, regfactor, regfactorimg, result, Kernel, Kernelimg);
}
