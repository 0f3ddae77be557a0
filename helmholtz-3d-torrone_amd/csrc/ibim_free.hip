// ibim_free.hip -- the matrix-free sweep of the combined-field tube and FP64
// iterative refinement.
//
// When the count x count operator does not fit (or should not be stored), each
// BiCG iteration recomputes the entries of a kSweepRows x kSweepCols tile in
// registers and applies them to p (row sums) and p~ (column sums) from that one
// evaluation -- the same tile / partial-plane structure as the dense sweep, so
// the BiCG vector kernels are shared. That is the combined-field tube's
// operator (the reference's FP64 entries, ibim_entries.hpp), which has no pair
// factorisation; the traditional tube's matrix-free operators (OP_FREE64 and
// the mixed-precision OP_FREE32) sweep as pair storage with computed factors,
// one evaluation per pair (ibim_pair.hip).
//
// Refinement: x = 0, r = b; repeat { solve A32 d = r by BiCG to inner_tol;
// x += d; r = b - A64 x with the FP64 operator } until ||r|| / ||b|| < tol.
#include <cmath>
#include <vector>

#include "ibim_entries.hpp"
#include "ibim_internal.hpp"
#include "ibim_wave.hpp"

namespace ibim {

namespace {

constexpr int kFT = 256;
constexpr int kFCpt = kSweepCols / kFT;
constexpr int kFRows = kSweepRows;
constexpr int kStage = 64;

struct RowStage {
    double sx[kStage], sy[kStage], sz[kStage], nx[kStage], ny[kStage], nz[kStage], w[kStage], reg[kStage],
        dist[kStage];
    double qre[kStage], qim[kStage];
};

template <class T>
__device__ __forceinline__ T wsum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// The FP64 operator (the reference's entries) recomputed per sweep: column points in registers, row points staged in LDS, one wavefront
// reduction per row (the FP64 entry evaluation dominates this kernel).
template <int Kind>
__global__ void __launch_bounds__(kFT) k_sweep_free64(const double* __restrict__ tube, int count, ibim_params prm,
                                                    const double2* __restrict__ p, const double2* __restrict__ q,
                                                    double2* __restrict__ ypart, double2* __restrict__ zpart,
                                                    size_t pld, const int* __restrict__ done) {
    if (done != nullptr && *done) return;
    __shared__ RowStage st;
    __shared__ double ysh[2][kFT / 64][kFRows];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int c0 = blockIdx.x * kSweepCols;
    const int r0 = blockIdx.y * kFRows;
    const int nrow = min(kFRows, count - r0);
    const size_t cnt = (size_t)count;

    PointData col[kFCpt];
    double pre[kFCpt], pim[kFCpt], zre[kFCpt], zim[kFCpt];
    int cj[kFCpt];
#pragma unroll
    for (int m = 0; m < kFCpt; ++m) {
        cj[m] = c0 + t + kFT * m;
        if (cj[m] < count) {
            col[m] = load_point(tube, cnt, cj[m]);
            double2 pv = p[cj[m]];
            pre[m] = pv.x;
            pim[m] = pv.y;
        } else {
            col[m] = PointData{};
            pre[m] = pim[m] = 0.0;
        }
        zre[m] = zim[m] = 0.0;
    }
    for (int s0 = 0; s0 < nrow; s0 += kStage) {
        const int ns = min(kStage, nrow - s0);
        __syncthreads();
        if (t < ns) {
            const int i = r0 + s0 + t;
            st.sx[t] = tube[IBIM_F_SX * cnt + i];
            st.sy[t] = tube[IBIM_F_SY * cnt + i];
            st.sz[t] = tube[IBIM_F_SZ * cnt + i];
            st.nx[t] = tube[IBIM_F_NX * cnt + i];
            st.ny[t] = tube[IBIM_F_NY * cnt + i];
            st.nz[t] = tube[IBIM_F_NZ * cnt + i];
            st.w[t] = tube[IBIM_F_W * cnt + i];
            st.reg[t] = tube[IBIM_F_REG * cnt + i];
            st.dist[t] = tube[IBIM_F_DIST * cnt + i];
            double2 qv = q[i];
            st.qre[t] = qv.x;
            st.qim[t] = qv.y;
        }
        __syncthreads();
        for (int s = 0; s < ns; ++s) {
            const int i = r0 + s0 + s;
            PointData a;
            a.sx = st.sx[s];
            a.sy = st.sy[s];
            a.sz = st.sz[s];
            a.nx = st.nx[s];
            a.ny = st.ny[s];
            a.nz = st.nz[s];
            a.w = st.w[s];
            a.reg = st.reg[s];
            a.dist = st.dist[s];
            const double qre = st.qre[s], qim = st.qim[s];
            double sre = 0.0, sim = 0.0;
#pragma unroll
            for (int m = 0; m < kFCpt; ++m) {
                double er = 0.0, ei = 0.0;
                if (cj[m] < count) {
                    if (Kind == IBIM_TUBE_TRADITIONAL)
                        trad_entry(a, col[m], i == cj[m], prm.k, prm.tau, prm.dirichlet, prm.interior, prm.combined,
                                   prm.eta, &er, &ei);
                    else
                        combo_entry(a, col[m], i == cj[m], prm.k, prm.tau, prm.eta, &er, &ei);
                }
                sre += er * pre[m] - ei * pim[m];
                sim += er * pim[m] + ei * pre[m];
                zre[m] += er * qre + ei * qim;
                zim[m] += er * qim - ei * qre;
            }
            sre = wsum(sre);
            sim = wsum(sim);
            if (lane == 0) {
                ysh[0][wave][s0 + s] = sre;
                ysh[1][wave][s0 + s] = sim;
            }
        }
    }
    __syncthreads();
    for (int row = t; row < nrow; row += kFT) {
        double a = 0.0, b = 0.0;
#pragma unroll
        for (int w = 0; w < kFT / 64; ++w) {
            a += (double)ysh[0][w][row];
            b += (double)ysh[1][w][row];
        }
        ypart[(size_t)blockIdx.x * pld + r0 + row] = make_double2(a, b);
    }
#pragma unroll
    for (int m = 0; m < kFCpt; ++m)
        if (cj[m] < count) zpart[(size_t)blockIdx.y * pld + cj[m]] = make_double2((double)zre[m], (double)zim[m]);
}

// r = b - y, partial sums of |r|^2 per block.
__global__ void __launch_bounds__(256) k_residual(const double2* __restrict__ b, const double2* __restrict__ y,
                                                  double2* __restrict__ r, int count, double* __restrict__ part) {
    __shared__ double sh[4];
    double s = 0.0;
    for (int j = blockIdx.x * 256 + threadIdx.x; j < count; j += gridDim.x * 256) {
        double2 bb = b[j], yy = y[j];
        double2 rr = make_double2(bb.x - yy.x, bb.y - yy.y);
        r[j] = rr;
        s += rr.x * rr.x + rr.y * rr.y;
    }
    s = wsum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

__global__ void __launch_bounds__(256) k_add(double2* __restrict__ x, const double2* __restrict__ d, int count) {
    for (int j = blockIdx.x * 256 + threadIdx.x; j < count; j += gridDim.x * 256) {
        double2 a = x[j], e = d[j];
        x[j] = make_double2(a.x + e.x, a.y + e.y);
    }
}

unsigned nblocks(int count) {
    unsigned nb = (unsigned)((count + 255) / 256);
    return nb > 1024u ? 1024u : nb;
}

}  // namespace

// The combined-field tube's matrix-free sweep: FP64 entries for either op (the
// mixed-precision entries are the traditional kernel's).
int sweep_free_launch(Context* ctx, int op, const void* p, const void* q, void* ypart, void* zpart, const int* done) {
    if (!ctx->have_tube) {
        set_error("matrix-free sweep: no tube for the current operator (set from the host)");
        return IBIM_ERR_ARG;
    }
    if (ctx->kind != IBIM_TUBE_COMBO || (op != OP_FREE64 && op != OP_FREE32)) {
        set_error("matrix-free sweep: the traditional tube sweeps as pairs");
        return IBIM_ERR_ARG;
    }
    const int count = ctx->count;
    dim3 grid((unsigned)((count + kSweepCols - 1) / kSweepCols), (unsigned)((count + kFRows - 1) / kFRows));
    TimedScope ts(ctx, "sweep_free64");
    hipLaunchKernelGGL((k_sweep_free64<IBIM_TUBE_COMBO>), grid, dim3(kFT), 0, ctx->stream, ctx->tube.as<double>(),
                       count, ctx->prm, (const double2*)p, (const double2*)q, (double2*)ypart, (double2*)zpart,
                       ctx->ld, done);
    IBIM_HIP(hipGetLastError());
    return IBIM_OK;
}

int run_refined(Context* ctx, const ibim_solve_opts* opts, ibim_solve_info* info, double* hist) {
    const int count = ctx->count;
    const size_t ld = ctx->ld;
    const int inner = ctx->mat_precision == IBIM_C64 ? OP_DENSE : OP_FREE32;
    const int outer_op = fp64_operator(ctx);
    if (inner == OP_DENSE && outer_op == OP_DENSE) return IBIM_ERR_ARG;  // unreachable: one precision stored
    if (ctx->refine.ensure(sizeof(double2) * ld * 4 + sizeof(double) * 1024)) return IBIM_ERR_NOMEM;
    double2* r = ctx->refine.as<double2>();
    double2* d = r + ld;
    double2* y = d + ld;
    double2* z = y + ld;
    double* part = (double*)(z + ld);
    const unsigned nb = nblocks(count);
    std::vector<double> hpart(nb);
    const double inner_tol = opts->inner_tol > 0 ? opts->inner_tol : 1e-4;
    const int max_outer = opts->max_outer > 0 ? opts->max_outer : 20;
    int total_its = 0, outer = 0, hist_at = 0;
    double res = 0.0;
    int status = 1;
    for (int c = 0; c < ctx->nrhs; ++c) {
        const double2* b = ctx->rhs.as<double2>() + (size_t)c * count;
        double2* x = ctx->density.as<double2>() + (size_t)c * count;
        IBIM_HIP(hipMemsetAsync(x, 0, sizeof(double2) * count, ctx->stream));
        // ||b||
        hipLaunchKernelGGL(k_residual, dim3(nb), dim3(256), 0, ctx->stream, b, x, r, count, part);
        IBIM_HIP(hipMemcpyAsync(hpart.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, ctx->stream));
        IBIM_HIP(hipStreamSynchronize(ctx->stream));
        double nb2 = 0.0;
        for (unsigned i = 0; i < nb; ++i) nb2 += hpart[i];
        const double normb = sqrt(nb2);
        if (normb == 0.0) {
            status = 4;
            continue;
        }
        status = 1;
        for (outer = 1; outer <= max_outer; ++outer) {
            // later corrections only need to bring the residual from `res` to tol
            const double itol = outer == 1 ? inner_tol : fmax(inner_tol, fmin(0.5, 0.3 * opts->tol / res));
            ibim_solve_info in{};
            int rc = bicg_solve_op(ctx, inner, r, d, opts->maxit, itol, opts->chunk, &in, nullptr, false, false);
            if (rc) return rc;
            total_its += in.iterations;
            hipLaunchKernelGGL(k_add, dim3(nb), dim3(256), 0, ctx->stream, x, d, count);
            rc = matvec_op(ctx, outer_op, x, x, y, z);
            if (rc) return rc;
            hipLaunchKernelGGL(k_residual, dim3(nb), dim3(256), 0, ctx->stream, b, y, r, count, part);
            IBIM_HIP(hipGetLastError());
            IBIM_HIP(hipMemcpyAsync(hpart.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, ctx->stream));
            IBIM_HIP(hipStreamSynchronize(ctx->stream));
            double s = 0.0;
            for (unsigned i = 0; i < nb; ++i) s += hpart[i];
            res = sqrt(s) / normb;
            if (hist && c == 0 && hist_at < opts->maxit) hist[hist_at++] = res;
            if (res < opts->tol) {
                status = 0;
                break;
            }
            if (res != res) {
                status = 2;
                break;
            }
        }
        if (outer > max_outer) outer = max_outer;
    }
    info->iterations = total_its;
    info->outer = outer;
    info->rel_residual = res;
    info->status = status;
    ctx->last_iters = total_its;
    return IBIM_OK;
}

}  // namespace ibim
