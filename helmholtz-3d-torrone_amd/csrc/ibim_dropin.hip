// ibim_dropin.hip -- the reference's entry points (HelmholtzKernel.h:20-39,
// BICGSTAB.h:15-17, weno.h:26-31, distance_application3D.h:9-12) on top of the
// device-resident API. Grid-sized and N^2 work runs on the GPU; the per-node
// scalar entries (weno, WENO_D*, *Curv) run on the host with the same math the
// kernels use. This file defines the extern "C" symbols; ibim_dropin_cxx.cpp
// defines the C++-linkage twins that a program compiled against the
// reference's own headers links to.
//
// The drop-ins share one lazily created handle on the current device; like the
// reference functions they are not reentrant.
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ibim_dropin.hpp"
#include "ibim_entries.hpp"
#include "ibim_internal.hpp"

// The reference's caller defines these (large3Dexample2.c:48-55) and its
// callees read them (HelmholtzKernel.h:16-18); weak definitions here let a
// library-only caller (ctypes) link, and a caller's own definitions override.
__attribute__((weak)) int N = 20;
__attribute__((weak)) int SIZE = 21;
__attribute__((weak)) double H = 0.2;
__attribute__((weak)) double HCUBED = 0.008;
__attribute__((weak)) double WAVE = 15.0;

namespace ibim {

namespace {

constexpr double kLo = -2.0;  // INT_L (frame.h:7)

ibim_ctx* g_ctx = nullptr;

ibim_ctx* dropin_ctx() {
    if (!g_ctx && ibim_create(&g_ctx) != IBIM_OK) g_ctx = nullptr;
    return g_ctx;
}

void gather(double*** g, int S, double* flat) {
    for (int i = 0; i < S; ++i)
        for (int j = 0; j < S; ++j) memcpy(flat + ((size_t)i * S + j) * S, g[i][j], sizeof(double) * S);
}

void scatter(const double* flat, int S, double*** g) {
    for (int i = 0; i < S; ++i)
        for (int j = 0; j < S; ++j) memcpy(g[i][j], flat + ((size_t)i * S + j) * S, sizeof(double) * S);
}

// Copy rows of the assembled operator into a double** pair.
int rows_out(ibim_ctx* ctx, int size, double** re, double** im) {
    if (ctx->mat_precision < 0 || size != ctx->count) return IBIM_ERR_ARG;
    HostRows dst;
    dst.re_rows = re;
    dst.im_rows = im;
    return download_rows(ctx, 0, size, dst);
}

// N^2 fill with finite-difference gradients of the reference's Phi
// (cal_partial, functions.c:121-155: central differences with step H in the
// second argument, of cal_phi = cos(k rho)/(-4 pi rho) and phiimg).
// sign +1 / diag +1/2 / near test rho^2 < tau^2 with reg_i: DoubleLayerKernel
// (HelmholtzKernel.c:726-758); sign -1 / diag -1/2 / rho^2 < thr with a
// constant: new_HelmholtzKernel (HelmholtzKernel.c:119-166).
__device__ __forceinline__ void phi_ref(double k, double dx, double dy, double dz, double* re, double* im) {
    double nrm = sqrt(fabs(dx * dx + dy * dy + dz * dz));
    if (nrm == 0.0) {
        *re = 0.0;
        *im = 0.0;
        return;
    }
    *re = cos(k * nrm) / (-4.0 * kPi * nrm);
    *im = sin(k * nrm) / (-4.0 * kPi * nrm);
}

__global__ void __launch_bounds__(256) k_assemble_fd(const double* __restrict__ tube, int count, double k, double hstep,
                                                     double sign, double diag, double thr2, int const_reg,
                                                     double reg_const, double2* __restrict__ mat, size_t ld) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= count) return;
    const size_t c = (size_t)count;
    double xi = tube[IBIM_F_SX * c + i], yi = tube[IBIM_F_SY * c + i], zi = tube[IBIM_F_SZ * c + i];
    double xj = tube[IBIM_F_SX * c + j], yj = tube[IBIM_F_SY * c + j], zj = tube[IBIM_F_SZ * c + j];
    double nx = tube[IBIM_F_NX * c + i], ny = tube[IBIM_F_NY * c + i], nz = tube[IBIM_F_NZ * c + i];
    double g[6];
    const double off[3][3] = {{hstep, 0, 0}, {0, hstep, 0}, {0, 0, hstep}};
    for (int m = 0; m < 3; ++m) {
        double ar, ai, br, bi;
        phi_ref(k, xi - (xj + off[m][0]), yi - (yj + off[m][1]), zi - (zj + off[m][2]), &ar, &ai);
        phi_ref(k, xi - (xj - off[m][0]), yi - (yj - off[m][1]), zi - (zj - off[m][2]), &br, &bi);
        g[m] = (ar - br) / (2.0 * hstep);
        g[3 + m] = (ai - bi) / (2.0 * hstep);
    }
    double pr = sign * (g[0] * nx + g[1] * ny + g[2] * nz);
    double pim = sign * (g[3] * nx + g[4] * ny + g[5] * nz);
    double d2 = (xi - xj) * (xi - xj) + (yi - yj) * (yi - yj) + (zi - zj) * (zi - zj);
    if (d2 < thr2) {
        pr = const_reg ? reg_const : tube[IBIM_F_REG * c + i];
        pim = 0.0;
    }
    double w = tube[IBIM_F_W * c + j];
    double re = pr * w, im = pim * w;
    if (i == j) re += diag;
    mat[(size_t)i * ld + j] = make_double2(re, im);
}

int assemble_fd(ibim_ctx* ctx, double sign, double diag, double thr2, int const_reg, double reg_const) {
    const int count = ctx->count;
    ctx->ld = ((size_t)count + 63) / 64 * 64;
    if (ctx->mat.ensure(sizeof(double2) * ctx->ld * (size_t)count)) return IBIM_ERR_NOMEM;
    IBIM_HIP(hipMemsetAsync(ctx->mat.ptr, 0, sizeof(double2) * ctx->ld * (size_t)count, ctx->stream));
    hipLaunchKernelGGL(k_assemble_fd, dim3((count + 255) / 256, count), dim3(256), 0, ctx->stream,
                       ctx->tube.as<double>(), count, ctx->prm.k, H, sign, diag, thr2, const_reg, reg_const,
                       ctx->mat.as<double2>(), ctx->ld);
    IBIM_HIP(hipGetLastError());
    ctx->mat_precision = IBIM_C128;
    ctx->mat_pair = ctx->mat_pair64 = ctx->pair_aux = ctx->free_cache_valid = false;
    return IBIM_OK;
}

// Grid upload shared by the kernel drop-ins: level set plus the caller's
// gradient and curvature grids (the reference's new_HelmholtzKernelCombo and
// DoubleLayerKernel take them as arguments instead of recomputing them).
int upload_grids(ibim_ctx* ctx, double*** level, double*** dx, double*** dy, double*** dz, double*** mc, double*** gc) {
    const int S = SIZE;
    const size_t nodes = (size_t)S * S * S;
    std::vector<double> f[6];
    double*** src[6] = {level, dx, dy, dz, mc, gc};
    for (int q = 0; q < 6; ++q) {
        f[q].resize(nodes);
        gather(src[q], S, f[q].data());
    }
    int rc = ibim_set_level_host(ctx, N, kLo, H, f[0].data());
    if (!rc) rc = ibim_set_grid_fields_host(ctx, f[1].data(), f[2].data(), f[3].data(), f[4].data(), f[5].data());
    return rc;
}

void report(const char* who, int rc) {
    if (rc) fprintf(stderr, "%s: %s (code %d)\n", who, ibim_last_error(), rc);
}

}  // namespace

// The host scalar entries (weno, WENO_D*, *Curv) are in ibim_stencil_host.cpp.

int dropin_assign_gradient(int n, double h, double*** dist, double*** gx, double*** gy, double*** gz, int mode) {
    ibim_ctx* ctx = dropin_ctx();
    if (!ctx) return IBIM_ERR_NODEVICE;
    if (mode != 3 && mode != 5) {
        set_error("ASSIGN_GRADIENT: weno_mode %d is not a WENO order (3 or 5)", mode);
        return IBIM_ERR_ARG;
    }
    const int S = n + 1;
    std::vector<double> flat((size_t)S * S * S), a(flat.size()), b(flat.size()), c(flat.size());
    gather(dist, S, flat.data());
    int rc = ibim_set_level_host(ctx, n, kLo, h, flat.data());
    if (!rc) rc = launch_gradient(ctx, mode);
    if (!rc) rc = ibim_get_grid(ctx, a.data(), b.data(), c.data(), nullptr, nullptr);
    if (rc) return rc;
    scatter(a.data(), S, gx);
    scatter(b.data(), S, gy);
    scatter(c.data(), S, gz);
    return 0;
}

// -------------------------------------------------------------- BICGSTAB.h --
int dropin_compbicg(int size, double* density, double* densityimg, double** K, double** Ki, double* b, double* bimg,
                    int maxit) {
    ibim_ctx* ctx = dropin_ctx();
    if (!ctx) return IBIM_ERR_NODEVICE;
    if (size <= 0 || !K || !Ki) return IBIM_ERR_ARG;
    // the rows are read in place (double**) and interleaved on the way to HBM
    ctx->count = size;
    ctx->have_tube = false;  // a host operator without points
    ctx->ld = ((size_t)size + 63) / 64 * 64;
    ctx->mat_precision = -1;
    ctx->mat_pair = ctx->mat_pair64 = ctx->pair_aux = ctx->free_cache_valid = false;
    if (ctx->mat.ensure(sizeof(double2) * ctx->ld * (size_t)size)) return IBIM_ERR_NOMEM;
    HostRows src;
    src.re_rows = K;
    src.im_rows = Ki;
    int rc = upload_rows(ctx, src, IBIM_C128);
    if (rc) return rc;
    ctx->mat_precision = IBIM_C128;
    rc = ibim_set_nrhs(ctx, 1);
    if (!rc) rc = ibim_set_rhs_host(ctx, 0, b, bimg);
    if (rc) return rc;
    ibim_solve_opts o{};
    o.method = IBIM_SOLVE_BICG;
    o.maxit = maxit;
    o.tol = 1e-7;  // DEFAULTACC (BICGSTAB.h:12)
    ibim_solve_info info{};
    ctx->nrhs = 1;
    rc = ibim_solve(ctx, &o, &info, nullptr);
    if (!rc) rc = ibim_get_density(ctx, 0, density, densityimg);
    return rc;
}

// ---------------------------------------------------------- HelmholtzKernel.h
int dropin_combo(int size, double delta, double epsilon, double*** level, double*** ldx, double*** ldy,
                 double*** ldz, double*** lmc, double*** lgc, double* zx, double* zy, double* zz, double* sx,
                 double* sy, double* sz, double* dist, double* gx, double* gy, double* gz, double* jac, double* reg,
                 double* regimg, double* sinew, double* siner, double** Kre, double** Kim, double wavenum,
                 double eta) {
    ibim_ctx* ctx = dropin_ctx();
    if (!ctx) return IBIM_ERR_NODEVICE;
    (void)wavenum;  // the reference passes WAVE here (large3Dexample2.c:1028); WAVE is used throughout
    int rc = upload_grids(ctx, level, ldx, ldy, ldz, lmc, lgc);
    if (rc) return rc;
    ibim_params prm{};
    prm.k = WAVE;
    prm.eps = epsilon;
    prm.tau = 2.0 * H;  // TAU_FACTOR (PDE.h:29)
    prm.delta_in = delta;
    prm.eta = eta;
    prm.sine_root = 0.1;  // functions.c:864
    int count = 0;
    rc = ibim_tube(ctx, &prm, IBIM_TUBE_COMBO, &count);
    if (rc) return rc;
    if (count != size) {
        set_error("Counter = %d, specified size = %d (HelmholtzKernel.c:403-408)", count, size);
        return IBIM_ERR_SIZE;
    }
    rc = ibim_assemble(ctx, IBIM_C128);
    if (rc) return rc;
    std::vector<double> t((size_t)IBIM_NF * count);
    rc = ibim_get_tube(ctx, nullptr, t.data());
    if (rc) return rc;
    auto F = [&](int f, int p) { return t[(size_t)f * count + p]; };
    for (int p = 0; p < count; ++p) {
        zx[p] = F(IBIM_F_ZX, p);
        zy[p] = F(IBIM_F_ZY, p);
        zz[p] = F(IBIM_F_ZZ, p);
        sx[p] = F(IBIM_F_SX, p);
        sy[p] = F(IBIM_F_SY, p);
        sz[p] = F(IBIM_F_SZ, p);
        dist[p] = F(IBIM_F_DIST, p);
        gx[p] = F(IBIM_F_NX, p);
        gy[p] = F(IBIM_F_NY, p);
        gz[p] = F(IBIM_F_NZ, p);
        jac[p] = F(IBIM_F_JAC, p);
        reg[p] = F(IBIM_F_REG, p);
        regimg[p] = 0.0;
        if (sinew) sinew[p] = F(IBIM_F_DELTA, p);
        if (siner) siner[p] = F(IBIM_F_W, p);
    }
    rc = rows_out(ctx, count, Kre, Kim);
    return rc ? rc : ibim_regcount(ctx);
}

int dropin_double_layer(int size, double epsilon, double*** level, double*** ldx, double*** ldy, double*** ldz,
                        double*** lmc, double*** lgc, double* zx, double* zy, double* zz, double* sx, double* sy,
                        double* sz, double* dist, double* gx, double* gy, double* gz, double* delta, double* jac,
                        double* reg, double* regimg, double* result, double** K, double** Ki) {
    ibim_ctx* ctx = dropin_ctx();
    if (!ctx) return IBIM_ERR_NODEVICE;
    int rc = upload_grids(ctx, level, ldx, ldy, ldz, lmc, lgc);
    if (rc) return rc;
    ibim_params prm{};
    prm.k = WAVE;
    prm.eps = epsilon;
    prm.tau = 2.0 * H;
    prm.interior = 1;  // DoubleLayerKernel keeps the interior sign of reg (HelmholtzKernel.c:710)
    int count = 0;
    rc = ibim_tube(ctx, &prm, IBIM_TUBE_TRADITIONAL, &count);
    if (rc) return rc;
    if (count != size) {
        set_error("size= %d, counti = %d (HelmholtzKernel.c:720-724)", size, count);
        return IBIM_ERR_SIZE;
    }
    rc = assemble_fd(ctx, 1.0, 0.5, prm.tau * prm.tau, 0, 0.0);
    if (rc) return rc;
    std::vector<double> t((size_t)IBIM_NF * count);
    rc = ibim_get_tube(ctx, nullptr, t.data());
    if (rc) return rc;
    auto F = [&](int f, int p) { return t[(size_t)f * count + p]; };
    for (int p = 0; p < count; ++p) {
        zx[p] = F(IBIM_F_ZX, p);
        zy[p] = F(IBIM_F_ZY, p);
        zz[p] = F(IBIM_F_ZZ, p);
        sx[p] = F(IBIM_F_SX, p);
        sy[p] = F(IBIM_F_SY, p);
        sz[p] = F(IBIM_F_SZ, p);
        dist[p] = F(IBIM_F_DIST, p);
        gx[p] = F(IBIM_F_NX, p);
        gy[p] = F(IBIM_F_NY, p);
        gz[p] = F(IBIM_F_NZ, p);
        delta[p] = F(IBIM_F_DELTA, p);
        jac[p] = F(IBIM_F_JAC, p);
        reg[p] = F(IBIM_F_REG, p);
        regimg[p] = 0.0;
        result[p] = F(IBIM_F_W, p);
    }
    return rows_out(ctx, count, K, Ki);
}

// Sphere-only analytic variant (HelmholtzKernel.c:4-169): the unit sphere at
// the origin (frame.h:19-22), analytic normals and curvature, the Sine weight,
// a constant near-field value and FD gradients of Phi.
int dropin_new_kernel(int size, double delta, double epsilon, double* zxl, double* zyl, double* zzl, double* sxl,
                      double* syl, double* szl, double* distl, double* gxl, double* gyl, double* gzl, double* curvl,
                      double* jacl, double* sinew, double* siner, double** Kre, double** Kim) {
    ibim_ctx* ctx = dropin_ctx();
    if (!ctx) return IBIM_ERR_NODEVICE;
    const double R = 1.0;  // radius, centre (0, 0, 0)
    const double tau = 0.1 * H, thr = 0.01 * H * H;
    const double t1 = 1.0 / (4.0 * kPi * tau * R), t2 = (-5.0 * tau) / (256.0 * pow(R, 3.0) * kPi);
    const double t3 = (-50.0 * tau) / (1536.0 * pow(R, 3.0) * kPi), t4 = (pow(WAVE, 2.0) * tau) / (24.0 * kPi * R);
    const double regc = -1.0 * (t1 + t2 + t3 + t4);
    std::vector<double> t;
    int count = 0;
    for (int pass = 0; pass < 2; ++pass) {
        int c = 0;
        for (int i = 0; i < SIZE; ++i) {
            double y = kLo + (double)i * H;
            for (int j = 0; j < SIZE; ++j) {
                double x = kLo + (double)j * H;
                for (int k = 0; k < SIZE; ++k) {
                    double z = kLo + (double)k * H;
                    double r = sqrt(x * x + y * y + z * z);
                    double d = R - r;
                    if (fabs(d) > epsilon || fabs(d) < delta) continue;
                    if (pass == 1) {
                        double gx = (0.0 - x) / r, gy = (0.0 - y) / r, gz = (0.0 - z) / r;
                        double px = x - d * gx, py = y - d * gy, pz = z - d * gz;
                        double pr = sqrt(px * px + py * py + pz * pz);
                        double cv = 1. / (R - d), jv = 1. + d * cv;
                        double sw = sine_weight(fabs(d / epsilon), delta / epsilon, 0.1) / (2.0 * epsilon);
                        double v[IBIM_NF] = {x, y, z, px, py, pz, d, (0.0 - px) / pr, (0.0 - py) / pr,
                                             (0.0 - pz) / pr, sw, jv, HCUBED * jv * sw, regc, cv, 0.0};
                        for (int f = 0; f < IBIM_NF; ++f) t[(size_t)f * count + c] = v[f];
                    }
                    ++c;
                }
            }
        }
        if (pass == 0) {
            count = c;
            if (count != size) {
                set_error("Size different for new Kernel (HelmholtzKernel.c:113-117): %d vs %d", count, size);
                return IBIM_ERR_SIZE;
            }
            t.assign((size_t)IBIM_NF * count, 0.0);
        }
    }
    ctx->count = count;
    ctx->have_tube = false;  // the sphere's analytic points serve only this dense operator
    ctx->prm = ibim_params{};
    ctx->prm.k = WAVE;
    if (ctx->tube.ensure(sizeof(double) * t.size())) return IBIM_ERR_NOMEM;
    IBIM_HIP(hipMemcpyAsync(ctx->tube.ptr, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice, ctx->stream));
    int rc = assemble_fd(ctx, -1.0, -0.5, thr, 1, regc);
    if (rc) return rc;
    for (int p = 0; p < count; ++p) {
        auto F = [&](int f) { return t[(size_t)f * count + p]; };
        zxl[p] = F(IBIM_F_ZX);
        zyl[p] = F(IBIM_F_ZY);
        zzl[p] = F(IBIM_F_ZZ);
        sxl[p] = F(IBIM_F_SX);
        syl[p] = F(IBIM_F_SY);
        szl[p] = F(IBIM_F_SZ);
        distl[p] = F(IBIM_F_DIST);
        gxl[p] = F(IBIM_F_NX);
        gyl[p] = F(IBIM_F_NY);
        gzl[p] = F(IBIM_F_NZ);
        curvl[p] = F(IBIM_F_HM);
        jacl[p] = F(IBIM_F_JAC);
        if (sinew) sinew[p] = F(IBIM_F_DELTA);
        if (siner) siner[p] = F(IBIM_F_W);
    }
    return rows_out(ctx, count, Kre, Kim);
}

}  // namespace ibim

// ------------------------------------------------------- extern "C" symbols --
using namespace ibim;

extern "C" {

void weno(double* dfp, double* dfn, double fi[], const double dx, const int order) {
    dropin_weno(dfp, dfn, fi, dx, order);
}
void WENO_DX(double*** u, const int i, const int j, const int k, const int n, const double h, double* p, double* m,
             const int order) {
    dropin_weno_axis(0, u, i, j, k, n, h, p, m, order);
}
void WENO_DY(double*** u, const int i, const int j, const int k, const int n, const double h, double* p, double* m,
             const int order) {
    dropin_weno_axis(1, u, i, j, k, n, h, p, m, order);
}
void WENO_DZ(double*** u, const int i, const int j, const int k, const int n, const double h, double* p, double* m,
             const int order) {
    dropin_weno_axis(2, u, i, j, k, n, h, p, m, order);
}

double GaussianCurv(double*** dist, int i, int j, int k, int n, double h) {
    double d[6], mc, gc;
    dropin_second(dist, i, j, k, n, h, d);
    curvatures_from(d, &mc, &gc);
    return gc;
}
double MeanCurv(double*** dist, int i, int j, int k, int n, double h) {
    double d[6], mc, gc;
    dropin_second(dist, i, j, k, n, h, d);
    curvatures_from(d, &mc, &gc);
    return mc;
}
int BothCurvs(double*** dist, int i, int j, int k, int n, double h, double* meancurv, double* gausscurv) {
    double d[6];
    dropin_second(dist, i, j, k, n, h, d);
    curvatures_from(d, meancurv, gausscurv);
    return 0;
}
int ASSIGN_GRADIENT(int n, double h, double*** dist, double*** gx, double*** gy, double*** gz, int weno_mode) {
    int rc = dropin_assign_gradient(n, h, dist, gx, gy, gz, weno_mode);
    report("ASSIGN_GRADIENT", rc);
    return rc;
}

int CompBiCG(int size, double* density, double* densityimg, double** Kernel, double** Kernelimg, double* b,
             double* bimg) {
    int rc = dropin_compbicg(size, density, densityimg, Kernel, Kernelimg, b, bimg, 1000);  // BICGSTAB.c:159
    report("CompBiCG", rc);
    return rc;
}
int omp_CompBiCG(int size, double* density, double* densityimg, double** Kernel, double** Kernelimg, double* b,
                 double* bimg, int chunk) {
    (void)chunk;
    int rc = dropin_compbicg(size, density, densityimg, Kernel, Kernelimg, b, bimg, 2000);  // BICGSTAB.c:497
    report("omp_CompBiCG", rc);
    return rc;
}

int new_HelmholtzKernel(int size, double delta, double epsilon, double* zxlist, double* zylist, double* zzlist,
                        double* zxstarlist, double* zystarlist, double* zzstarlist, double* dist, double* gradx,
                        double* grady, double* gradz, double* curv, double* Jacobian, double* PolyW, double* SineW,
                        double* PolyResult, double* SineResult, double** PolyKernelH, double** PolyKernelHimg,
                        double** SineKernelH, double** SineKernelHimg) {
    (void)PolyW;
    (void)PolyResult;
    (void)PolyKernelH;
    (void)PolyKernelHimg;  // POLY_TEST is 0 (PDE.h:18)
    int rc = dropin_new_kernel(size, delta, epsilon, zxlist, zylist, zzlist, zxstarlist, zystarlist, zzstarlist, dist,
                               gradx, grady, gradz, curv, Jacobian, SineW, SineResult, SineKernelH, SineKernelHimg);
    report("new_HelmholtzKernel", rc);
    return rc;
}

int new_HelmholtzKernelCombo(int size, double delta, double epsilon, double*** level, double*** leveldx,
                             double*** leveldy, double*** leveldz, double*** levelmcurv, double*** levelgcurv,
                             double* zxlist, double* zylist, double* zzlist, double* zxstarlist, double* zystarlist,
                             double* zzstarlist, double* dist, double* gradx, double* grady, double* gradz,
                             double* Jacobian, double* regfactor, double* regfactorimg, double* PolyW, double* SineW,
                             double* PolyResult, double* SineResult, double** PolyKernelH, double** PolyKernelHimg,
                             double** SineKernelH, double** SineKernelHimg, double wavenum, double eta) {
    (void)PolyW;
    (void)PolyResult;
    (void)PolyKernelH;
    (void)PolyKernelHimg;
    int rc = dropin_combo(size, delta, epsilon, level, leveldx, leveldy, leveldz, levelmcurv, levelgcurv, zxlist,
                          zylist, zzlist, zxstarlist, zystarlist, zzstarlist, dist, gradx, grady, gradz, Jacobian,
                          regfactor, regfactorimg, SineW, SineResult, SineKernelH, SineKernelHimg, wavenum, eta);
    report("new_HelmholtzKernelCombo", rc < 0 ? rc : 0);
    return rc;
}

int DoubleLayerKernel(int size, double epsilon, double*** level, double*** leveldx, double*** leveldy,
                      double*** leveldz, double*** levelmcurv, double*** levelgcurv, double* zxlist, double* zylist,
                      double* zzlist, double* zxstarlist, double* zystarlist, double* zzstarlist, double* dist,
                      double* gradx, double* grady, double* gradz, double* delta, double* Jacobian, double* regfactor,
                      double* regfactorimg, double* result, double** Kernel, double** Kernelimg, double wavenum) {
    (void)wavenum;  // the reference's body uses WAVE through cal_partial
    int rc = dropin_double_layer(size, epsilon, level, leveldx, leveldy, leveldz, levelmcurv, levelgcurv, zxlist,
                                 zylist, zzlist, zxstarlist, zystarlist, zzstarlist, dist, gradx, grady, gradz, delta,
                                 Jacobian, regfactor, regfactorimg, result, Kernel, Kernelimg);
    report("DoubleLayerKernel", rc);
    return rc;
}

}  // extern "C"
