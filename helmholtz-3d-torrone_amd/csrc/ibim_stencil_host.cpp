// ibim_stencil_host.cpp -- the reference's per-node scalar entries (weno,
// WENO_D*, MeanCurv / GaussianCurv / BothCurvs; weno.h:26-31,
// distance_application3D.h:9-11) on the host, with the stencil math the GPU
// kernels use (ibim_math.hpp) read through the caller's double*** grid.
// Host-only (built with g++), so the same code also builds under the CPU
// sanitizers.
#include "ibim_dropin.hpp"
#include "ibim_math.hpp"

namespace ibim {

namespace {

// double*** accessors for the shared stencil math.
struct PtrLine {
    double*** u;
    int i, j, k, axis;  // axis 0 = x (j), 1 = y (i), 2 = z (k)
    double operator()(int m) const { return axis == 0 ? u[i][m][k] : (axis == 1 ? u[m][j][k] : u[i][j][m]); }
};

struct PtrPlane {
    double*** u;
    int i, j, k, pair;  // 0 = xy (i, j), 1 = xz (j, k), 2 = yz (i, k)
    double operator()(int p, int q) const {
        return pair == 0 ? u[i + p][j + q][k] : (pair == 1 ? u[i][j + p][k + q] : u[i + p][j][k + q]);
    }
};

}  // namespace

// ----------------------------------------------------------------- weno.h ---
void dropin_weno(double* dfp, double* dfn, double* fi, double dx, int order) { weno_1d(fi, dx, order, dfp, dfn); }

void dropin_weno_axis(int axis, double*** u, int i, int j, int k, int n, double h, double* up, double* un, int order) {
    const int c = axis == 0 ? j : (axis == 1 ? i : k);
    weno_line(PtrLine{u, i, j, k, axis}, c, n, h, order, up, un);
}

// --------------------------------------------------- distance_application3D.h
void dropin_curvatures(const double* d6, double* mean, double* gauss) { curvatures_from(d6, mean, gauss); }

void dropin_second(double*** u, int i, int j, int k, int n, double h, double* o6) {
    o6[0] = d2_line(PtrLine{u, i, j, k, 0}, j, n, h, 0);
    o6[1] = d2_line(PtrLine{u, i, j, k, 1}, i, n, h, 1);
    o6[2] = d2_line(PtrLine{u, i, j, k, 2}, k, n, h, 1);
    o6[3] = dmixed(PtrPlane{u, i, j, k, 0}, i, j, n, h);
    o6[4] = dmixed(PtrPlane{u, i, j, k, 1}, j, k, n, h);
    o6[5] = dmixed(PtrPlane{u, i, j, k, 2}, i, k, n, h);
}

}  // namespace ibim
