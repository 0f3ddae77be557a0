// ibim_eval.hip -- right-hand sides at the closest points and the field at
// receivers (one workgroup per receiver, fixed-order FP64 block reduction).
#include "ibim_entries.hpp"
#include "ibim_internal.hpp"

namespace ibim {

namespace {

constexpr int kThreads = 256;

__global__ void __launch_bounds__(kThreads) k_rhs(const double* __restrict__ tube, int count, int kind, double k,
                                                  double x0, double y0, double z0, double amp, double dx, double dy,
                                                  double dz, double2* __restrict__ b) {
    int p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= count) return;
    const size_t c = (size_t)count;
    double sx = tube[IBIM_F_SX * c + p], sy = tube[IBIM_F_SY * c + p], sz = tube[IBIM_F_SZ * c + p];
    double re, im;
    if (kind == IBIM_RHS_POINT_SOURCE) {
        // e^{ik r} / (4 pi r): the exterior field of a source inside the surface
        // (the reference's cal_u is this times -4 pi / k, large3Dexample2.c:4642)
        double a = sx - x0, bb = sy - y0, cc = sz - z0;
        double r = sqrt(a * a + bb * bb + cc * cc);
        double sv, cv;
        sincos(k * r, &sv, &cv);
        re = cv / (4. * kPi * r);
        im = sv / (4. * kPi * r);
    } else if (kind == IBIM_RHS_PLANE_NEUMANN) {
        // d/dn of amp e^{ik dir.x} along the tube normal: cal_partialu /
        // cal_partialuimg with INCIDENTALWAVE (large3Dexample2.c:4735-4746, 4792-4810)
        double nx = tube[IBIM_F_NX * c + p], ny = tube[IBIM_F_NY * c + p], nz = tube[IBIM_F_NZ * c + p];
        double ph = sx * dx + sy * dy + sz * dz;
        double dn = -1. * (dx * nx + dy * ny + dz * nz);
        re = amp * k * dn * sin(k * ph);
        im = -amp * k * dn * cos(k * ph);
    } else {
        // sound-soft data of an incident plane wave: -amp e^{ik dir.x*}
        double ph = k * (sx * dx + sy * dy + sz * dz);
        double sv, cv;
        sincos(ph, &sv, &cv);
        re = -amp * cv;
        im = -amp * sv;
    }
    b[p] = make_double2(re, im);
}

// Field at receiver blockIdx.x. mode IBIM_EVAL_SINGLE: NeumannTest
// (large3Dexample2.c:1810-1824); IBIM_EVAL_DOUBLE: the layer of the Dirichlet
// kernel, w_j grad_x Phi(x, y*_j) . n_j; IBIM_EVAL_COMBINED: that layer minus
// i eta w_j Phi(x, y*_j), the combined-field potential; IBIM_EVAL_COMBO: Combo_NeumannTest's
// sample kernel (large3Dexample2.c:1988-2058) with the near field from reg_j.
__global__ void __launch_bounds__(kThreads) k_eval(const double* __restrict__ tube, int count, ibim_params prm,
                                                   int mode, const double2* __restrict__ phi,
                                                   const double* __restrict__ xyz, double2* __restrict__ out) {
    __shared__ double sh[2][kThreads / 64];
    const size_t c = (size_t)count;
    const double X = xyz[3 * blockIdx.x], Y = xyz[3 * blockIdx.x + 1], Z = xyz[3 * blockIdx.x + 2];
    const double k = prm.k;
    double sr = 0.0, si = 0.0;
    for (int j = threadIdx.x; j < count; j += kThreads) {
        double a = X - tube[IBIM_F_SX * c + j], b = Y - tube[IBIM_F_SY * c + j], cc = Z - tube[IBIM_F_SZ * c + j];
        double rho = sqrt(a * a + b * b + cc * cc);
        double kr = 0.0, ki = 0.0;
        if (mode == IBIM_EVAL_SINGLE) {
            if (rho > prm.tau) {
                double sv, cv;
                sincos(k * rho, &sv, &cv);
                kr = cv / (-4. * kPi * rho);
                ki = sv / (-4. * kPi * rho);
            }
        } else if (mode == IBIM_EVAL_DOUBLE || mode == IBIM_EVAL_COMBINED) {
            if (rho > prm.tau) {
                double c1, c2;
                grad_phi_coeffs(k, rho, &c1, &c2);
                double ip = a * tube[IBIM_F_NX * c + j] + b * tube[IBIM_F_NY * c + j] + cc * tube[IBIM_F_NZ * c + j];
                kr = c1 * ip;
                ki = c2 * ip;
                if (mode == IBIM_EVAL_COMBINED) {  // - i eta Phi
                    double gr, gi;
                    phi_value(k, rho, &gr, &gi);
                    kr = kr + prm.eta * gi;
                    ki = ki - prm.eta * gr;
                }
            }
        } else {
            double sv, cv;
            sincos(k * rho, &sv, &cv);
            double ip = a * tube[IBIM_F_NX * c + j] + b * tube[IBIM_F_NY * c + j] + cc * tube[IBIM_F_NZ * c + j];
            double f1 = 4. * kPi * rho, f2 = f1 * rho;
            if (rho > prm.tau) {
                kr = (cv / rho + k * sv) * ip / f2;
                ki = (sv / rho - k * cv) * ip / f2;
            } else {
                kr = tube[IBIM_F_REG * c + j];
            }
            kr += prm.eta * sv / f1;
            ki += -1. * prm.eta * cv / f1;
        }
        double2 ph = phi[j];
        double w = tube[IBIM_F_W * c + j];
        sr += (kr * ph.x - ki * ph.y) * w;
        si += (kr * ph.y + ki * ph.x) * w;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sr += __shfl_xor(sr, off);
        si += __shfl_xor(si, off);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = sr;
        sh[1][threadIdx.x >> 6] = si;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        out[blockIdx.x] = make_double2((sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]),
                                       (sh[1][0] + sh[1][1]) + (sh[1][2] + sh[1][3]));
}

}  // namespace

int launch_rhs(Context* ctx, int kind, const double* x0, double amp, const double* dir, int col) {
    const int count = ctx->count;
    double2* b = ctx->rhs.as<double2>() + (size_t)col * count;
    double X0[3] = {0, 0, 0}, D[3] = {0, 0, 0};
    if (x0)
        for (int i = 0; i < 3; ++i) X0[i] = x0[i];
    if (dir)
        for (int i = 0; i < 3; ++i) D[i] = dir[i];
    TimedScope ts(ctx, "rhs");
    hipLaunchKernelGGL(k_rhs, dim3((count + kThreads - 1) / kThreads), dim3(kThreads), 0, ctx->stream,
                       ctx->tube.as<double>(), count, kind, ctx->prm.k, X0[0], X0[1], X0[2], amp, D[0], D[1], D[2], b);
    IBIM_HIP(hipGetLastError());
    return IBIM_OK;
}

int launch_eval(Context* ctx, int mode, int col, int nrecv, const double* xyz_dev, double* out_dev) {
    const double2* phi = ctx->density.as<double2>() + (size_t)col * ctx->count;
    TimedScope ts(ctx, "evaluate");
    hipLaunchKernelGGL(k_eval, dim3(nrecv), dim3(kThreads), 0, ctx->stream, ctx->tube.as<double>(), ctx->count,
                       ctx->prm, mode, phi, xyz_dev, (double2*)out_dev);
    IBIM_HIP(hipGetLastError());
    return IBIM_OK;
}

}  // namespace ibim
