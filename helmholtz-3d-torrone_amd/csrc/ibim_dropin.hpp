// ibim_dropin.hpp -- implementations behind the reference-named entry points,
// shared by the extern "C" definitions (ibim_dropin.hip) and the C++-linkage
// definitions (ibim_dropin_cxx.cpp). Deliberately free of HIP and of ibim.h so
// that the C++-linkage translation unit never sees the extern "C" declarations.
#pragma once

namespace ibim {

void dropin_weno(double* dfp, double* dfn, double* fi, double dx, int order);
void dropin_weno_axis(int axis, double*** u, int i, int j, int k, int n, double h, double* up, double* un, int order);
void dropin_second(double*** u, int i, int j, int k, int n, double h, double* o6);
int dropin_assign_gradient(int n, double h, double*** dist, double*** gx, double*** gy, double*** gz, int mode);
int dropin_compbicg(int size, double* density, double* densityimg, double** K, double** Ki, double* b, double* bimg,
                    int maxit);
int dropin_combo(int size, double delta, double epsilon, double*** level, double*** ldx, double*** ldy,
                 double*** ldz, double*** lmc, double*** lgc, double* zx, double* zy, double* zz, double* sx,
                 double* sy, double* sz, double* dist, double* gx, double* gy, double* gz, double* jac, double* reg,
                 double* regimg, double* sinew, double* siner, double** Kre, double** Kim, double wavenum,
                 double eta);
int dropin_double_layer(int size, double epsilon, double*** level, double*** ldx, double*** ldy, double*** ldz,
                        double*** lmc, double*** lgc, double* zx, double* zy, double* zz, double* sx, double* sy,
                        double* sz, double* dist, double* gx, double* gy, double* gz, double* delta, double* jac,
                        double* reg, double* regimg, double* result, double** K, double** Ki);
int dropin_new_kernel(int size, double delta, double epsilon, double* zxl, double* zyl, double* zzl, double* sxl,
                      double* syl, double* szl, double* distl, double* gxl, double* gyl, double* gzl, double* curvl,
                      double* jacl, double* sinew, double* siner, double** Kre, double** Kim);
// Six second derivatives -> mean / Gaussian curvature (distance_application3D.c:35-54).
void dropin_curvatures(const double* d6, double* mean, double* gauss);

}  // namespace ibim
