// ibim_geometry.hip -- level-set geometry and tube extraction on gfx950.
//
// Grid pass (HBM-bound stencils, FP64): one thread per node, blocks walk the
// node range in raster order so a wavefront reads 64 consecutive z-nodes of one
// (y, x) line and the +-3 x / y neighbours come from lines the neighbouring
// blocks of the same workgroup batch are reading (L2 hits).
//
// Tube: flags -> per-block counts -> one-block exclusive scan -> ordered
// scatter (raster order, as the reference's triple loops produce), then a
// per-point kernel for projection, normals, weights and the near-field value.
#include "ibim_internal.hpp"
#include "ibim_math.hpp"

namespace ibim {

namespace {

constexpr int kBlock = 256;
constexpr int kItems = 8;  // nodes per thread in the tube count / scatter kernels

// ASSIGN_GRADIENT with WENO (distance_application3D.c:56-141).
__global__ void __launch_bounds__(kBlock) k_gradient(const double* __restrict__ u, int n, double h, int order,
                                                     double* __restrict__ gx, double* __restrict__ gy,
                                                     double* __restrict__ gz) {
    const int S = n + 1;
    const size_t total = (size_t)S * S * S;
    size_t id = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (id >= total) return;
    int k = (int)(id % S);
    size_t t = id / S;
    int j = (int)(t % S), i = (int)(t / S);
    double xp, xn, yp, yn, zp, zn;
    weno_line(FlatLine{u, node_at(S, i, 0, k), (size_t)S}, j, n, h, order, &xp, &xn);
    weno_line(FlatLine{u, node_at(S, 0, j, k), (size_t)S * S}, i, n, h, order, &yp, &yn);
    weno_line(FlatLine{u, node_at(S, i, j, 0), 1}, k, n, h, order, &zp, &zn);
    double d = u[id];
    gx[id] = godunov_pick(xp, xn, d);
    gy[id] = godunov_pick(yp, yn, d);
    gz[id] = godunov_pick(zp, zn, d);
}

// BothCurvs on every node (distance_application3D.c:35-54, large3Dexample2.c:446-455).
__global__ void __launch_bounds__(kBlock) k_curvatures(const double* __restrict__ u, int n, double h,
                                                       double* __restrict__ mc, double* __restrict__ gc) {
    const int S = n + 1;
    const size_t total = (size_t)S * S * S;
    size_t id = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (id >= total) return;
    int k = (int)(id % S);
    size_t t = id / S;
    int j = (int)(t % S), i = (int)(t / S);
    double d6[6];
    second_derivs(u, n, h, i, j, k, d6);
    curvatures_from(d6, &mc[id], &gc[id]);
}

struct TubeRule {
    int kind, one_sided, interior;
    double eps, delta_in;
};

__device__ __forceinline__ bool rule_in(const TubeRule& r, double d) {
    return in_tube(d, r.kind, r.eps, r.delta_in, r.one_sided, r.interior);
}

// Per-block count of tube nodes; a block covers kBlock * kItems consecutive nodes.
__global__ void __launch_bounds__(kBlock) k_tube_count(const double* __restrict__ u, size_t total, TubeRule rule,
                                                       int* __restrict__ counts) {
    __shared__ int wsum[kBlock / 64];
    size_t base = (size_t)blockIdx.x * kBlock * kItems;
    int c = 0;
#pragma unroll
    for (int q = 0; q < kItems; ++q) {
        size_t id = base + (size_t)q * kBlock + threadIdx.x;
        if (id < total && rule_in(rule, u[id])) ++c;
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int w = 0; w < kBlock / 64; ++w) s += wsum[w];
        counts[blockIdx.x] = s;
    }
}

// Exclusive scan of the per-block counts in one workgroup; counts[nb] = total.
__global__ void __launch_bounds__(1024) k_scan_counts(int* __restrict__ counts, int nb) {
    __shared__ int sh[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 1024) {
        int i = base + (int)threadIdx.x;
        int v = i < nb ? counts[i] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {  // Hillis-Steele inclusive scan
            int add = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0;
            __syncthreads();
            sh[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < nb) counts[i] = carry + sh[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += sh[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[nb] = carry;
}

// Ordered scatter of tube node indices: node q*kBlock + t of the block lands at
// offset + (rank among earlier flagged nodes of the block), keeping raster order.
__global__ void __launch_bounds__(kBlock) k_tube_scatter(const double* __restrict__ u, size_t total, TubeRule rule,
                                                         const int* __restrict__ offsets, int* __restrict__ node) {
    __shared__ int wcnt[kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    size_t base = (size_t)blockIdx.x * kBlock * kItems;
    int running = offsets[blockIdx.x];
    for (int q = 0; q < kItems; ++q) {
        size_t id = base + (size_t)q * kBlock + threadIdx.x;
        bool f = id < total && rule_in(rule, u[id]);
        unsigned long long m = __ballot(f);
        int before_lane = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wcnt[wave] = __popcll(m);
        __syncthreads();
        int before_wave = 0, all = 0;
        for (int w = 0; w < kBlock / 64; ++w) {
            if (w < wave) before_wave += wcnt[w];
            all += wcnt[w];
        }
        if (f) node[running + before_wave + before_lane] = (int)id;
        running += all;
        __syncthreads();
    }
}

// Per-point tube data: omp_get_Ccomp's first-call block (large3Dexample2.c:4365-4447)
// for the traditional tube, new_HelmholtzKernelCombo's (HelmholtzKernel.c:278-383)
// for the combo tube. Fields are written field-major with stride `count`.
__global__ void __launch_bounds__(kBlock) k_tube_points(const double* __restrict__ u, const double* __restrict__ gx,
                                                        const double* __restrict__ gy, const double* __restrict__ gz,
                                                        const double* __restrict__ mc, const double* __restrict__ gc,
                                                        int n, double lo, double h, ibim_params prm, int kind,
                                                        const int* __restrict__ node, int count,
                                                        double* __restrict__ out, int* __restrict__ err) {
    int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= count) return;
    const int S = n + 1;
    size_t id = (size_t)node[p];
    int k = (int)(id % S);
    size_t t = id / S;
    int j = (int)(t % S), i = (int)(t / S);
    double zy = lo + h * (double)i, zx = lo + h * (double)j, zz = lo + h * (double)k;
    double d = u[id];
    double x[3] = {zx - d * gx[id], zy - d * gy[id], zz - d * gz[id]};
    double nv[3], cv[2];
    if (interp_grad(n, h, lo, gx, gy, gz, x, nv) < 0 || interp_curv(n, h, lo, mc, gc, x, cv) < 0) {
        atomicOr(err, 1);
        nv[0] = nv[1] = nv[2] = 0.0;
        cv[0] = cv[1] = 0.0;
    } else {
        // both reference callers normalise the interpolated gradient once more
        // (large3Dexample2.c:4407-4410, HelmholtzKernel.c:291-294)
        double nn = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
        nv[0] = nv[0] / nn;
        nv[1] = nv[1] / nn;
        nv[2] = nv[2] / nn;
    }
    double jac = 1 + 2. * d * mc[id] + d * d * gc[id];
    const double hcubed = h * h * h;
    double wt, wq;
    if (kind == IBIM_TUBE_TRADITIONAL) {
        wt = prm.one_sided ? 2. * delta_cos(prm.eps, d) : delta_cos(prm.eps, d);
        wq = hcubed * wt * jac;
    } else {
        double tt = fabs(d / prm.eps), r = prm.delta_in / prm.eps;
        wt = sine_weight(tt, r, prm.sine_root) / (prm.one_sided ? prm.eps : 2.0 * prm.eps);
        wq = hcubed * jac * wt;
    }
    const size_t c = (size_t)count;
    out[IBIM_F_ZX * c + p] = zx;
    out[IBIM_F_ZY * c + p] = zy;
    out[IBIM_F_ZZ * c + p] = zz;
    out[IBIM_F_SX * c + p] = x[0];
    out[IBIM_F_SY * c + p] = x[1];
    out[IBIM_F_SZ * c + p] = x[2];
    out[IBIM_F_DIST * c + p] = d;
    out[IBIM_F_NX * c + p] = nv[0];
    out[IBIM_F_NY * c + p] = nv[1];
    out[IBIM_F_NZ * c + p] = nv[2];
    out[IBIM_F_DELTA * c + p] = wt;
    out[IBIM_F_JAC * c + p] = jac;
    out[IBIM_F_W * c + p] = wq;
    out[IBIM_F_REG * c + p] = reg_value(cv[0], cv[1], prm.tau, prm.k, prm.interior);
    out[IBIM_F_HM * c + p] = cv[0];
    out[IBIM_F_GK * c + p] = cv[1];
}

unsigned blocks_for(size_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

int launch_gradient(Context* ctx, int order) {
    const size_t total = (size_t)(ctx->n + 1) * (ctx->n + 1) * (ctx->n + 1);
    TimedScope ts(ctx, "gradient");
    hipLaunchKernelGGL(k_gradient, dim3(blocks_for(total, kBlock)), dim3(kBlock), 0, ctx->stream,
                       ctx->level.as<double>(), ctx->n, ctx->h, order, ctx->gx.as<double>(), ctx->gy.as<double>(),
                       ctx->gz.as<double>());
    IBIM_HIP(hipGetLastError());
    return IBIM_OK;
}

int launch_curvatures(Context* ctx) {
    const size_t total = (size_t)(ctx->n + 1) * (ctx->n + 1) * (ctx->n + 1);
    TimedScope ts(ctx, "curvature");
    hipLaunchKernelGGL(k_curvatures, dim3(blocks_for(total, kBlock)), dim3(kBlock), 0, ctx->stream,
                       ctx->level.as<double>(), ctx->n, ctx->h, ctx->mc.as<double>(), ctx->gc.as<double>());
    IBIM_HIP(hipGetLastError());
    return IBIM_OK;
}

int launch_tube(Context* ctx, int* count_out) {
    const size_t total = (size_t)(ctx->n + 1) * (ctx->n + 1) * (ctx->n + 1);
    const unsigned nb = blocks_for(total, (unsigned)(kBlock * kItems));
    if (ctx->blockcnt.ensure(sizeof(int) * (nb + 1)) || ctx->err_flag.ensure(sizeof(int))) return IBIM_ERR_NOMEM;
    TubeRule rule{ctx->kind, ctx->prm.one_sided, ctx->prm.interior, ctx->prm.eps, ctx->prm.delta_in};
    int* counts = ctx->blockcnt.as<int>();
    {
        TimedScope ts(ctx, "tube_scan");
        hipLaunchKernelGGL(k_tube_count, dim3(nb), dim3(kBlock), 0, ctx->stream, ctx->level.as<double>(), total, rule,
                           counts);
        hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, ctx->stream, counts, (int)nb);
    }
    IBIM_HIP(hipGetLastError());
    int count = 0;
    IBIM_HIP(hipMemcpyAsync(&count, counts + nb, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    if (count <= 0) {
        set_error("the tube is empty (eps = %g, delta = %g)", ctx->prm.eps, ctx->prm.delta_in);
        return IBIM_ERR_SIZE;
    }
    if (ctx->node.ensure(sizeof(int) * (size_t)count) || ctx->tube.ensure(sizeof(double) * IBIM_NF * (size_t)count))
        return IBIM_ERR_NOMEM;
    {
        TimedScope ts(ctx, "tube_points");
        hipLaunchKernelGGL(k_tube_scatter, dim3(nb), dim3(kBlock), 0, ctx->stream, ctx->level.as<double>(), total,
                           rule, counts, ctx->node.as<int>());
        IBIM_HIP(hipMemsetAsync(ctx->err_flag.ptr, 0, sizeof(int), ctx->stream));
        hipLaunchKernelGGL(k_tube_points, dim3(blocks_for((size_t)count, kBlock)), dim3(kBlock), 0, ctx->stream,
                           ctx->level.as<double>(), ctx->gx.as<double>(), ctx->gy.as<double>(), ctx->gz.as<double>(),
                           ctx->mc.as<double>(), ctx->gc.as<double>(), ctx->n, ctx->lo, ctx->h, ctx->prm, ctx->kind,
                           ctx->node.as<int>(), count, ctx->tube.as<double>(), ctx->err_flag.as<int>());
    }
    IBIM_HIP(hipGetLastError());
    int err = 0;
    IBIM_HIP(hipMemcpyAsync(&err, ctx->err_flag.ptr, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    if (err) {
        set_error("a closest point fell outside the grid or its interpolation was NaN");
        return IBIM_ERR_GEOMETRY;
    }
    *count_out = count;
    return IBIM_OK;
}

}  // namespace ibim
