// ibim_gmres.hip -- IBIM_SOLVE_GMRES: FP64 iterative refinement around GMRES
// on the stored operator.
//
// GMRES (Saad and Schultz, SIAM J. Sci. Stat. Comput. 7, 1986) builds an
// orthonormal Krylov basis -- here by classical Gram-Schmidt applied twice
// (Giraud, Langou and Rozloznik, 2005), so each orthogonalization is two
// skinny products with the basis instead of j sequential ones -- and
// minimizes the residual over it through Givens rotations of the Hessenberg
// matrix, which the host keeps (one small copy per iteration). An iteration
// applies the operator once, to one vector: where BiCG needs A p and A^H p~
// per iteration (one fused pair sweep), GMRES needs A p only (the Y-only
// sweep). Every reduction runs in a fixed order, so solves are repeatable.
//
// Refinement: x = 0, r = b; repeat { GMRES on the stored operator for A d = r
// to tol ||b||; x += d; r = b - A64 x with the FP64 operator } until
// ||r|| / ||b|| < tol. On complex64 storage (IBIM_C64_PAIR, IBIM_MIXED_PAIR,
// IBIM_C64) that is the mixed-precision solver (entries rounded to complex64,
// products and vectors FP64, residuals FP64); on complex128 storage the FP64
// operator is the stored one and a single GMRES run is the FP64 solver.
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "ibim_internal.hpp"

namespace ibim {

namespace {

constexpr int kGT = 256;          // threads per block of the vector kernels
constexpr int kMaxBasis = 256;    // Krylov vectors per GMRES cycle (then restart)

// part[i][c] = sum over chunk c of conj(V_i[e]) w[e]: block (c, i) takes
// every kDotChunks-th group of kGT elements; a fixed-order tree over the block.
constexpr int kDotChunks = 16;
__global__ void __launch_bounds__(kGT) k_gm_dots(const double2* __restrict__ V, size_t ld, int count,
                                                 const double2* __restrict__ w, double2* __restrict__ part) {
    __shared__ double sr[kGT], si[kGT];
    const double2* v = V + (size_t)blockIdx.y * ld;
    double ar = 0.0, ai = 0.0;
    for (int e = blockIdx.x * kGT + threadIdx.x; e < count; e += kDotChunks * kGT) {
        const double2 a = v[e], b = w[e];
        ar += a.x * b.x + a.y * b.y;
        ai += a.x * b.y - a.y * b.x;
    }
    sr[threadIdx.x] = ar;
    si[threadIdx.x] = ai;
    __syncthreads();
    for (int s = kGT / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sr[threadIdx.x] += sr[threadIdx.x + s];
            si[threadIdx.x] += si[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blockIdx.y * kDotChunks + blockIdx.x] = make_double2(sr[0], si[0]);
}

// out[i] = sum_c part[i][c] in order (one thread per i)
__global__ void k_gm_dots_sum(const double2* __restrict__ part, int n, double2* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sr = 0.0, si = 0.0;
    for (int c = 0; c < kDotChunks; ++c) {
        sr += part[(size_t)i * kDotChunks + c].x;
        si += part[(size_t)i * kDotChunks + c].y;
    }
    out[i] = make_double2(sr, si);
}

// w[e] += sign sum_{i < n} V_i[e] h[i]  (n <= kMaxBasis + 1)
__global__ void __launch_bounds__(kGT) k_gm_combine(const double2* __restrict__ V, size_t ld, int count,
                                                    const double2* __restrict__ h, int n, double sign,
                                                    double2* __restrict__ w) {
    __shared__ double2 hs[kMaxBasis + 1];
    for (int i = threadIdx.x; i < n; i += kGT) hs[i] = h[i];
    __syncthreads();
    const int e = blockIdx.x * kGT + threadIdx.x;
    if (e >= count) return;
    double cr = 0.0, ci = 0.0;
    for (int i = 0; i < n; ++i) {
        const double2 v = V[(size_t)i * ld + e], c = hs[i];
        cr += v.x * c.x - v.y * c.y;
        ci += v.x * c.y + v.y * c.x;
    }
    const double2 o = w[e];
    w[e] = make_double2(o.x + sign * cr, o.y + sign * ci);
}

// v = s w (s real); out = a - b
__global__ void __launch_bounds__(kGT) k_gm_scale(const double2* __restrict__ w, double s, int count,
                                                  double2* __restrict__ v) {
    const int e = blockIdx.x * kGT + threadIdx.x;
    if (e < count) v[e] = make_double2(s * w[e].x, s * w[e].y);
}

__global__ void __launch_bounds__(kGT) k_gm_sub(const double2* __restrict__ a, const double2* __restrict__ b,
                                                int count, double2* __restrict__ out) {
    const int e = blockIdx.x * kGT + threadIdx.x;
    if (e < count) out[e] = make_double2(a[e].x - b[e].x, a[e].y - b[e].y);
}

typedef std::complex<double> cplx;

struct GmresSpace {
    double2 *V, *w, *z, *d, *r, *y, *coef;  // basis (kMaxBasis + 1 vectors), work vectors, coefficients
    double2* part;                          // dot-product partials, (kMaxBasis + 1) x kDotChunks
};

// out[i] = V_i^H w for i < n (device)
void dots(Context* ctx, const GmresSpace& g, const double2* V, int n, const double2* w, double2* out) {
    hipLaunchKernelGGL(k_gm_dots, dim3(kDotChunks, (unsigned)n), dim3(kGT), 0, ctx->stream, V, ctx->ld, ctx->count,
                       w, g.part);
    hipLaunchKernelGGL(k_gm_dots_sum, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, ctx->stream, g.part, n, out);
}

// ||v|| by the dot kernels (host value; synchronizes)
int norm_of(Context* ctx, const GmresSpace& g, const double2* v, double* out) {
    dots(ctx, g, v, 1, v, g.coef);
    double2 h;
    IBIM_HIP(hipMemcpyAsync(&h, g.coef, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    *out = sqrt(h.x);
    return IBIM_OK;
}

// GMRES(m), m <= kMaxBasis, for A d = rhs with operator `op`, from d = 0, until
// the residual estimate is below target (absolute) or `budget` products are used.
// Returns the products applied and the last estimate.
int gmres(Context* ctx, const GmresSpace& g, int op, const double2* rhs, double target, int budget, int m,
          int* products, double* estimate, double* hist, int* hist_at, int hist_len, double normb) {
    const int count = ctx->count;
    const size_t ld = ctx->ld;
    const unsigned nb = (unsigned)((count + kGT - 1) / kGT);
    IBIM_HIP(hipMemsetAsync(g.d, 0, sizeof(double2) * ld, ctx->stream));
    *products = 0;
    std::vector<cplx> H((size_t)(kMaxBasis + 1) * kMaxBasis), cs(kMaxBasis), sn(kMaxBasis), gv(kMaxBasis + 1);
    std::vector<double2> hh(2 * (kMaxBasis + 1) + 1);
    auto Hc = [&](int i, int j) -> cplx& { return H[(size_t)j * (kMaxBasis + 1) + i]; };
    bool first = true;
    for (;;) {
        // r0 = rhs - A d (the first cycle: rhs)
        if (first) {
            IBIM_HIP(hipMemcpyAsync(g.r, rhs, sizeof(double2) * count, hipMemcpyDeviceToDevice, ctx->stream));
        } else {
            if (*products >= budget) return IBIM_OK;
            if (int rc = matvec_op_y(ctx, op, g.d, g.w)) return rc;
            ++*products;
            hipLaunchKernelGGL(k_gm_sub, dim3(nb), dim3(kGT), 0, ctx->stream, rhs, g.w, count, g.r);
        }
        first = false;
        double beta;
        if (int rc = norm_of(ctx, g, g.r, &beta)) return rc;
        *estimate = beta;
        if (!(beta > target)) return IBIM_OK;
        hipLaunchKernelGGL(k_gm_scale, dim3(nb), dim3(kGT), 0, ctx->stream, g.r, 1.0 / beta, count, g.V);
        std::fill(gv.begin(), gv.end(), cplx(0.0));
        gv[0] = beta;
        int j = 0;
        bool stop = false;
        for (; j < m && *products < budget; ++j) {
            double2* vj = g.V + (size_t)j * ld;
            if (int rc = matvec_op_y(ctx, op, vj, g.w)) return rc;
            ++*products;
            // CGS2: h = V^H w, w -= V h, twice
            for (int pass = 0; pass < 2; ++pass) {
                double2* coef = g.coef + (size_t)pass * (kMaxBasis + 1);
                dots(ctx, g, g.V, j + 1, g.w, coef);
                hipLaunchKernelGGL(k_gm_combine, dim3(nb), dim3(kGT), 0, ctx->stream, g.V, ld, count, coef, j + 1,
                                   -1.0, g.w);
            }
            double2* nrm = g.coef + 2 * (size_t)(kMaxBasis + 1);
            dots(ctx, g, g.w, 1, g.w, nrm);
            IBIM_HIP(hipGetLastError());
            IBIM_HIP(hipMemcpyAsync(hh.data(), g.coef, sizeof(double2) * (2 * (kMaxBasis + 1) + 1),
                                    hipMemcpyDeviceToHost, ctx->stream));
            IBIM_HIP(hipStreamSynchronize(ctx->stream));
            for (int i = 0; i <= j; ++i)
                Hc(i, j) = cplx(hh[i].x + hh[kMaxBasis + 1 + i].x, hh[i].y + hh[kMaxBasis + 1 + i].y);
            const double hn = sqrt(hh[2 * (kMaxBasis + 1)].x);
            Hc(j + 1, j) = hn;
            // previous rotations, then the one that zeroes H(j + 1, j)
            for (int i = 0; i < j; ++i) {
                const cplx t = cs[i] * Hc(i, j) + sn[i] * Hc(i + 1, j);
                Hc(i + 1, j) = -std::conj(sn[i]) * Hc(i, j) + std::conj(cs[i]) * Hc(i + 1, j);
                Hc(i, j) = t;
            }
            const cplx a = Hc(j, j), bb = Hc(j + 1, j);
            const double rr = sqrt(std::norm(a) + std::norm(bb));
            cs[j] = rr > 0.0 ? std::conj(a) / rr : cplx(1.0);
            sn[j] = rr > 0.0 ? std::conj(bb) / rr : cplx(0.0);
            Hc(j, j) = rr;
            Hc(j + 1, j) = 0.0;
            gv[j + 1] = -std::conj(sn[j]) * gv[j];
            gv[j] = cs[j] * gv[j];
            *estimate = std::abs(gv[j + 1]);
            if (hist && *hist_at < hist_len) hist[(*hist_at)++] = *estimate / normb;
            if (!(*estimate > target) || !(hn > 0.0) || *estimate != *estimate) {
                stop = true;
                ++j;
                break;
            }
            if (j + 1 < m)
                hipLaunchKernelGGL(k_gm_scale, dim3(nb), dim3(kGT), 0, ctx->stream, g.w, 1.0 / hn, count,
                                   g.V + (size_t)(j + 1) * ld);
        }
        // y = R^{-1} g over the j columns, d += V y
        const int n = j;
        std::vector<cplx> yv(n);
        for (int i = n - 1; i >= 0; --i) {
            cplx s = gv[i];
            for (int c = i + 1; c < n; ++c) s -= Hc(i, c) * yv[c];
            yv[i] = Hc(i, i) != cplx(0.0) ? s / Hc(i, i) : cplx(0.0);
        }
        std::vector<double2> yd(n);
        for (int i = 0; i < n; ++i) yd[i] = make_double2(yv[i].real(), yv[i].imag());
        if (n > 0) {
            IBIM_HIP(hipMemcpyAsync(g.y, yd.data(), sizeof(double2) * n, hipMemcpyHostToDevice, ctx->stream));
            hipLaunchKernelGGL(k_gm_combine, dim3(nb), dim3(kGT), 0, ctx->stream, g.V, ld, count, g.y, n, 1.0, g.d);
            IBIM_HIP(hipGetLastError());
            IBIM_HIP(hipStreamSynchronize(ctx->stream));  // yd goes out of scope
        }
        if (stop || *products >= budget) return IBIM_OK;
    }
}

__global__ void __launch_bounds__(kGT) k_gm_add(double2* __restrict__ x, const double2* __restrict__ d, int count) {
    const int e = blockIdx.x * kGT + threadIdx.x;
    if (e < count) x[e] = make_double2(x[e].x + d[e].x, x[e].y + d[e].y);
}

}  // namespace

int run_gmres(Context* ctx, const ibim_solve_opts* opts, ibim_solve_info* info, double* hist) {
    if (ctx->mat_precision < 0) {
        set_error("IBIM_SOLVE_GMRES runs on a stored operator (ibim_assemble)");
        return IBIM_ERR_ARG;
    }
    const int count = ctx->count;
    const size_t ld = ctx->ld;
    const unsigned nb = (unsigned)((count + kGT - 1) / kGT);
    const int inner = OP_DENSE;
    const int outer_op = ctx->mat_precision == IBIM_C128 ? OP_DENSE : fp64_operator(ctx);
    const size_t nvec = (size_t)(kMaxBasis + 1) + 5;
    if (ctx->refine.ensure(sizeof(double2) * ld * nvec + sizeof(double2) * (2 * (kMaxBasis + 1) + 1) +
                           sizeof(double2) * kMaxBasis + sizeof(double2) * (kMaxBasis + 1) * kDotChunks + 512))
        return IBIM_ERR_NOMEM;
    GmresSpace g;
    g.V = ctx->refine.as<double2>();
    g.w = g.V + (size_t)(kMaxBasis + 1) * ld;
    g.z = g.w + ld;
    g.d = g.z + ld;
    g.r = g.d + ld;
    double2* res = g.r + ld;  // b - A64 x
    g.coef = res + ld;
    g.y = g.coef + 2 * (kMaxBasis + 1) + 1;
    g.part = g.y + kMaxBasis;
    // padding rows of the basis read as zero in the sweeps
    IBIM_HIP(hipMemsetAsync(g.V, 0, sizeof(double2) * ld * nvec, ctx->stream));
    const int max_outer = opts->max_outer > 0 ? opts->max_outer : 20;
    // the basis length: opts->chunk when set (a restart every chunk products), else kMaxBasis
    const int m = opts->chunk > 0 ? std::min(opts->chunk, kMaxBasis) : kMaxBasis;
    int total = 0, outer_max = 0, hist_at = 0, status = 4;  // 4 until a column has data
    double worst = 0.0;
    for (int c = 0; c < ctx->nrhs; ++c) {
        const double2* b = ctx->rhs.as<double2>() + (size_t)c * count;
        double2* x = ctx->density.as<double2>() + (size_t)c * count;
        IBIM_HIP(hipMemsetAsync(x, 0, sizeof(double2) * count, ctx->stream));
        IBIM_HIP(hipMemcpyAsync(res, b, sizeof(double2) * count, hipMemcpyDeviceToDevice, ctx->stream));
        double normb;
        if (int rc = norm_of(ctx, g, res, &normb)) return rc;
        if (normb == 0.0) continue;  // zero data: zero density, no iterations
        int col_status = 1, used = 0, outer = 0;
        double relres = 1.0;
        for (outer = 1; outer <= max_outer && used < opts->maxit; ++outer) {
            int products = 0;
            double est = 0.0;
            // the correction only has to bring the residual to tol ||b||
            if (int rc = gmres(ctx, g, inner, res, opts->tol * normb, opts->maxit - used, m, &products, &est,
                               c == 0 ? hist : nullptr, &hist_at, opts->maxit, normb))
                return rc;
            used += products;
            hipLaunchKernelGGL(k_gm_add, dim3(nb), dim3(kGT), 0, ctx->stream, x, g.d, count);
            if (int rc = matvec_op_y(ctx, outer_op, x, g.w)) return rc;
            hipLaunchKernelGGL(k_gm_sub, dim3(nb), dim3(kGT), 0, ctx->stream, b, g.w, count, res);
            IBIM_HIP(hipGetLastError());
            double rn;
            if (int rc = norm_of(ctx, g, res, &rn)) return rc;
            relres = rn / normb;
            if (relres != relres) {
                col_status = 2;
                break;
            }
            if (relres < opts->tol) {
                col_status = 0;
                break;
            }
        }
        total += used;
        outer_max = std::max(outer_max, std::min(outer, max_outer));
        worst = std::max(worst, relres);
        if (status == 4 || col_status != 0) status = col_status;
        report_progress(ctx, "GMRES", used, relres);
    }
    info->iterations = total;
    info->outer = outer_max;
    info->rel_residual = worst;
    info->status = status;
    ctx->last_iters = total;
    return IBIM_OK;
}

}  // namespace ibim
