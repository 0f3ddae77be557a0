// ibim_batch.hip -- BiCG on up to 16 right-hand sides per pass over the operator.
//
// One pass over the operator applies it to 16 vectors at once, as a complex
// GEMM on the FP64 matrix cores (v_mfma_f64_16x16x4f64, products and sums in
// FP64, so every column sees the same fixed operator as the single-vector
// path). The matrix-vector product Y = A P and the adjoint product Z = A^H Q
// are two passes; at 16 right-hand sides both are matrix-core bound, so
// fusing them would save no time and would need partial planes 16x larger.
//
//   row pass  (Y = A P):   a wave owns 16 rows and runs over all columns; the
//                          MFMA A operand is 16 rows x 4 adjacent columns,
//                          loaded directly.
//   column pass (Z = A^H Q): a wave owns 16 columns and runs over all rows; the
//                          MFMA A operand is 4 rows x 16 consecutive columns,
//                          loaded directly.
// The vector operand (a step x 16 block of P or Q) is staged in LDS once per
// workgroup and shared by its four waves.
// Pair storage and the matrix-free operators (traditional tube) run instead
// the fused 16-column pair pass of ibim_pair.hip: one pass for both products,
// the matrix-free ones evaluating each factor where the pairs would be read.
//
// Vectors are RHS-interleaved: element (point j, right-hand side r) at
// [j * 16 + r], with ld rows so padding rows read as zero. Each column keeps
// its own BiCG scalars and stop state (the reference's recurrences and stop
// rules per column, BICGSTAB.c:104-306); finished columns freeze.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "ibim_entries.hpp"
#include "ibim_internal.hpp"
#include "ibim_mfma.hpp"

namespace ibim {

namespace {

constexpr int kR = kBatchRHS;  // right-hand sides per batch: the MFMA N dimension
constexpr int kBT = 256;   // threads per workgroup (4 waves)
constexpr int kKT = 64;    // columns (row pass) or rows (column pass) staged per step
constexpr int kMaxVB = 1024;

struct BState {
    double rho[kR][2];
    double alpha[kR][2];
    double err[kR];
    double normb[kR];
    double sumb[kR];
    int done[kR];  // 0 running; else the iteration it stopped at (-1: zero data)
    int stop[kR];  // stop rule that fired (1 tol, 2 NaN, 3 stagnation, 4 zero data)
    int all_done;
    int iter;  // the iteration the vector kernels work on (kb_dot advances it)
};

// ------------------------------------------------------------ stored passes --
// Row pass: a workgroup owns 64 rows (16 per wave) and steps over the columns
// kKT at a time. The MFMA A operand (16 rows x 4 consecutive columns) is loaded
// directly: four lanes read four adjacent entries of one row (64 B for
// complex128), and the next three k-steps read the rest of those cache lines.
// The kKT x 16 block of P is staged once per workgroup.
template <class MT>
__global__ void __launch_bounds__(kBT) k_rows_stored(const MT* __restrict__ A, size_t ld, int count,
                                                     const double2* __restrict__ P, double2* __restrict__ Y,
                                                     const BState* __restrict__ st) {
    if (st != nullptr && st->all_done) return;
    __shared__ double2 ps[kKT][kR];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int row0 = (blockIdx.x * 4 + w) * 16;
    const int row = row0 + li;
    const bool rok = row < count;
    const MT* Ar = A + (size_t)(rok ? row : 0) * ld + lk;
    Acc3 y;
    for (int k0 = 0; k0 < count; k0 += kKT) {
        // k0 + kKT <= ld: both are multiples of 64 and k0 < count <= ld
        __syncthreads();
        for (int e = threadIdx.x; e < kKT * kR; e += kBT) ps[e / kR][e % kR] = P[(size_t)k0 * kR + e];
        __syncthreads();
#pragma unroll 4
        for (int s = 0; s < kKT / 4; ++s) {
            double ar = 0.0, ai = 0.0;
            if (rok) {
                const MT a = Ar[k0 + 4 * s];
                ar = (double)a.x;
                ai = (double)a.y;
            }
            const double2 p = ps[4 * s + lk][li];
            y.add(ar, ai, p.x, p.y);
        }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int r = row0 + 4 * v + lk;
        if (r < count) Y[(size_t)r * kR + li] = make_double2(y.re(v), y.im(v));
    }
}

// Complex64 row pass: lane (li, lk) reads the 16 consecutive entries of row li
// at columns k0 + 16 lk .. + 15 (128 B, eight 16-B loads), and k-step s of
// the chunk sums column k0 + 16 lk + s in the MFMA k slot lk: every column
// once per chunk, in a different order from the column index order. 1.2x the
// one-entry-per-k-step pass (11.7 vs 13.7 ms on 64,676 points, 16 right-hand
// sides); for complex128 the same runs were slower (18.9 vs 16.9 ms).
__global__ void __launch_bounds__(kBT) k_rows_stored_c64(const float2* __restrict__ A, size_t ld, int count,
                                                         const double2* __restrict__ P, double2* __restrict__ Y,
                                                         const BState* __restrict__ st) {
    if (st != nullptr && st->all_done) return;
    __shared__ double2 ps[kKT][kR];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int row0 = (blockIdx.x * 4 + w) * 16;
    const int row = row0 + li;
    const bool rok = row < count;
    // 16 B aligned: an even column of a row of ld = 64 m entries
    const float4* Ar = (const float4*)(A + (size_t)(rok ? row : 0) * ld + 16 * lk);
    Acc3 y;
    for (int k0 = 0; k0 < count; k0 += kKT) {
        // k0 + kKT <= ld: both are multiples of 64 and k0 < count <= ld
        float4 a[8];
#pragma unroll
        for (int h = 0; h < 8; ++h) a[h] = Ar[(size_t)k0 / 2 + h];
        __syncthreads();
        for (int e = threadIdx.x; e < kKT * kR; e += kBT) ps[e / kR][e % kR] = P[(size_t)k0 * kR + e];
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            const float4 v = a[s >> 1];
            const double ar = rok ? (double)((s & 1) ? v.z : v.x) : 0.0;
            const double ai = rok ? (double)((s & 1) ? v.w : v.y) : 0.0;
            const double2 p = ps[16 * lk + s][li];
            y.add(ar, ai, p.x, p.y);
        }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int r = row0 + 4 * v + lk;
        if (r < count) Y[(size_t)r * kR + li] = make_double2(y.re(v), y.im(v));
    }
}

// Column pass: a wave owns 16 columns and steps over the rows; the kKT x 16
// block of Q is staged once per workgroup.
template <class MT>
__global__ void __launch_bounds__(kBT) k_cols_stored(const MT* __restrict__ A, size_t ld, int count,
                                                     const double2* __restrict__ Q, double2* __restrict__ Z,
                                                     const BState* __restrict__ st) {
    if (st != nullptr && st->all_done) return;
    __shared__ double2 qs[kKT][kR];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int c0 = (blockIdx.x * 4 + w) * 16;
    const bool act = c0 < count;  // then c0 + 16 <= ld (multiples of 16 and 64)
    const int col = c0 + li;
    Acc3 z;
    for (int i0 = 0; i0 < count; i0 += kKT) {
        __syncthreads();
        for (int e = threadIdx.x; e < kKT * kR; e += kBT) {
            const int i = i0 + e / kR;
            qs[e / kR][e % kR] = i < count ? Q[(size_t)i0 * kR + e] : make_double2(0.0, 0.0);
        }
        __syncthreads();
#pragma unroll 4
        for (int s = 0; s < kKT / 4; ++s) {
            const int row = i0 + 4 * s + lk;
            const double2 q = qs[4 * s + lk][li];
            double ar = 0.0, ai = 0.0;
            if (act && row < count) {
                const MT a = A[(size_t)row * ld + col];
                ar = (double)a.x;
                ai = (double)a.y;
            }
            z.add(ar, -ai, q.x, q.y);  // conj(a) q
        }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int c = c0 + 4 * v + lk;
        if (c < count) Z[(size_t)c * kR + li] = make_double2(z.re(v), z.im(v));
    }
}

// ------------------------------------------------------------ vector kernels --
struct BArgs {
    double2 *x, *r, *rt, *p, *pt, *v, *w;
    const double2* b;
    double* part;     // per-block partials [block][kR][8]
    BState* st;
    double* hist;     // [maxit + 1][kR] relative residuals
    int count;
    int ref_guards;   // 1: the reference's small-data rules; 0: only exact zeros (correction solves)
};

// Thread t of a block works on right-hand side t % 16 for points
// blockIdx.x * 16 + t / 16 (+ a grid stride of gridDim.x * 16); per-column sums
// are reduced over the 4 lanes of a wave and the 4 waves in a fixed order.
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* out /* [kR][8] */) {
    __shared__ double sh[4][kR][NV];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r = threadIdx.x & 15;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        v[q] += __shfl_xor(v[q], 16);
        v[q] += __shfl_xor(v[q], 32);
    }
    if (lane < 16)
#pragma unroll
        for (int q = 0; q < NV; ++q) sh[wave][r][q] = v[q];
    __syncthreads();
    if (threadIdx.x < kR * NV) {
        const int rr = threadIdx.x / NV, q = threadIdx.x % NV;
        out[rr * 8 + q] = (sh[0][rr][q] + sh[1][rr][q]) + (sh[2][rr][q] + sh[3][rr][q]);
    }
}

// Reduce partials [nb][kR][8] to tot[kR][8] in LDS (one workgroup, fixed order).
__device__ __forceinline__ void reduce_all(const double* __restrict__ part, int nb, int nv, double (*tot)[8]) {
    __shared__ double half[2][kR * 8];
    const int t = threadIdx.x, idx = t & 127, h = t >> 7;
    double s = 0.0;
    if ((idx & 7) < nv)
        for (int b = h; b < nb; b += 2) s += part[(size_t)b * kR * 8 + idx];
    half[h][idx] = s;
    __syncthreads();
    if (t < kR * 8) tot[t >> 3][t & 7] = half[0][t] + half[1][t];
    __syncthreads();
}

__global__ void __launch_bounds__(kBT) kb_init(BArgs a) {
    const int r = threadIdx.x & 15;
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int j = blockIdx.x * 16 + (threadIdx.x >> 4); j < a.count; j += gridDim.x * 16) {
        const size_t e = (size_t)j * kR + r;
        const double2 b = a.b[e];
        const double2 c = make_double2(b.x, -b.y);
        a.x[e] = make_double2(0.0, 0.0);
        a.r[e] = b;
        a.p[e] = b;
        a.rt[e] = c;
        a.pt[e] = c;
        v[0] += b.x * c.x;
        v[1] += b.y * c.y;
        v[2] += b.y * c.x;
        v[3] += b.x * c.y;
        v[4] += b.x * b.x;
        v[5] += b.y * b.y;
        v[6] += fabs(b.x) + fabs(b.y);
    }
    block_partials<7>(v, a.part + (size_t)blockIdx.x * kR * 8);
}

__global__ void __launch_bounds__(kBT) kb_init_finish(BArgs a, int nb) {
    __shared__ double tot[kR][8];
    reduce_all(a.part, nb, 7, tot);
    if (threadIdx.x < kR) {
        const int r = threadIdx.x;
        BState* s = a.st;
        s->rho[r][0] = tot[r][0] + tot[r][1];
        s->rho[r][1] = tot[r][2] - tot[r][3];
        s->normb[r] = sqrt(tot[r][4] + tot[r][5]);
        s->sumb[r] = tot[r][6];
        s->err[r] = INFINITY;
        // BICGSTAB.c:147-156: zero density; BICGSTAB.c:138-145: perturbation of
        // tiny data (skipped for refinement corrections, whose data are small)
        const bool zero = a.ref_guards ? tot[r][6] < 1e-15 : tot[r][6] == 0.0;
        s->done[r] = zero ? -1 : 0;
        s->stop[r] = zero ? 4 : 0;
        if (a.ref_guards && !zero && tot[r][6] < 1e-13) {
            a.r[r] = make_double2(0.000000001, 0.000000001);
            a.rt[r] = make_double2(0.000000001, 0.000000001);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int all = 1;
        for (int r = 0; r < kR; ++r) all &= a.st->done[r] != 0;
        a.st->all_done = all;
        a.st->iter = 0;
    }
}

// partials of <r, r~> and <v, p~> per column
__global__ void __launch_bounds__(kBT) kb_dot(BArgs a) {
    if (a.st->all_done) {
        // the counter moves on with every launched iteration, as the columns'
        // stop iterations are compared with it
        if (blockIdx.x == 0 && threadIdx.x == 0) a.st->iter += 1;
        return;
    }
    const int r = threadIdx.x & 15;
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = blockIdx.x * 16 + (threadIdx.x >> 4); j < a.count; j += gridDim.x * 16) {
        const size_t e = (size_t)j * kR + r;
        const double2 rr = a.r[e], rt = a.rt[e], y = a.v[e], pt = a.pt[e];
        v[0] += rr.x * rt.x;
        v[1] += rr.y * rt.y;
        v[2] += rr.y * rt.x;
        v[3] += rr.x * rt.y;
        v[4] += y.x * pt.x;
        v[5] += y.y * pt.y;
        v[6] += y.y * pt.x;
        v[7] += y.x * pt.y;
    }
    block_partials<8>(v, a.part + (size_t)blockIdx.x * kR * 8);
    // the kernels after this one work on the next iteration
    if (blockIdx.x == 0 && threadIdx.x == 0) a.st->iter += 1;
}

// alpha = <r, r~> / <v, p~> per running column (BICGSTAB.c:217-229)
__global__ void __launch_bounds__(kBT) kb_alpha(BArgs a, int nb) {
    if (a.st->all_done) return;
    __shared__ double tot[kR][8];
    reduce_all(a.part, nb, 8, tot);
    if (threadIdx.x < kR) {
        const int r = threadIdx.x;
        const double ca = tot[r][0] + tot[r][1], cb = tot[r][2] - tot[r][3];
        const double cc = tot[r][4] + tot[r][5], cd = tot[r][6] - tot[r][7];
        const double den = cc * cc + cd * cd;
        a.st->alpha[r][0] = (ca * cc + cb * cd) / den;
        a.st->alpha[r][1] = (cb * cc - ca * cd) / den;
    }
}

// x += alpha p, r -= alpha v, r~ -= conj(alpha) w (BICGSTAB.c:231-242); partials
// of the new <r, r~> and ||r||^2.
__global__ void __launch_bounds__(kBT) kb_update_xr(BArgs a) {
    if (a.st->all_done) return;
    const int r = threadIdx.x & 15;
    const bool run = a.st->done[r] == 0;
    const double al = a.st->alpha[r][0], ali = a.st->alpha[r][1];
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (int j = blockIdx.x * 16 + (threadIdx.x >> 4); j < a.count; j += gridDim.x * 16) {
        const size_t e = (size_t)j * kR + r;
        double2 x = a.x[e], p = a.p[e], rr = a.r[e], vv = a.v[e], rt = a.rt[e], ww = a.w[e];
        if (run) {
            x.x = x.x + al * p.x - ali * p.y;
            x.y = x.y + ali * p.x + al * p.y;
            rr.x = rr.x - al * vv.x + ali * vv.y;
            rr.y = rr.y - ali * vv.x - al * vv.y;
            rt.x = rt.x - al * ww.x - ali * ww.y;
            rt.y = rt.y + ali * ww.x - al * ww.y;
            a.x[e] = x;
            a.r[e] = rr;
            a.rt[e] = rt;
        }
        v[0] += rr.x * rt.x;
        v[1] += rr.y * rt.y;
        v[2] += rr.y * rt.x;
        v[3] += rr.x * rt.y;
        v[4] += rr.x * rr.x;
        v[5] += rr.y * rr.y;
    }
    block_partials<6>(v, a.part + (size_t)blockIdx.x * kR * 8);
}

// beta = <r, r~>_new / <r, r~>_old (BICGSTAB.c:246-259), errors and stop rules
// (289-306) per running column; beta goes to alpha's slot for kb_update_p.
__global__ void __launch_bounds__(kBT) kb_beta(BArgs a, int nb, double tol) {
    const int it = a.st->iter;
    if (a.st->all_done) return;
    __shared__ double tot[kR][8];
    reduce_all(a.part, nb, 6, tot);
    if (threadIdx.x < kR) {
        const int r = threadIdx.x;
        BState* s = a.st;
        if (s->done[r] == 0) {
            const double ca = tot[r][0] + tot[r][1], cb = tot[r][2] - tot[r][3];
            const double o0 = s->rho[r][0], o1 = s->rho[r][1];
            const double den = o0 * o0 + o1 * o1;
            s->alpha[r][0] = (ca * o0 + cb * o1) / den;
            s->alpha[r][1] = (cb * o0 - ca * o1) / den;
            s->rho[r][0] = ca;
            s->rho[r][1] = cb;
            const double err = sqrt(tot[r][4] + tot[r][5]) / s->normb[r];
            int stop = 0;
            if (err < tol) stop = 1;
            else if (err != err) stop = 2;
            else if (fabs(s->err[r] - err) < 1e-3 * tol) stop = 3;
            s->err[r] = err;
            if (a.hist) a.hist[(size_t)it * kR + r] = err;
            if (stop) {
                s->done[r] = it;
                s->stop[r] = stop;
            }
        } else {
            s->alpha[r][0] = s->alpha[r][1] = 0.0;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int all = 1;
        for (int r = 0; r < kR; ++r) all &= a.st->done[r] != 0;
        a.st->all_done = all;
    }
}

// p = r + beta p, p~ = r~ + conj(beta) p~ (BICGSTAB.c:271-278), also for a
// column that stopped in this iteration `it`: its state is then complete, so
// residual replacement can send it back to iterate.
__global__ void __launch_bounds__(kBT) kb_update_p(BArgs a) {
    const int it = a.st->iter;
    const int r = threadIdx.x & 15;
    const int d = a.st->done[r];
    if (d != 0 && d != it) return;  // per-thread: the column is finished (no barriers below)
    const double be = a.st->alpha[r][0], bei = a.st->alpha[r][1];
    for (int j = blockIdx.x * 16 + (threadIdx.x >> 4); j < a.count; j += gridDim.x * 16) {
        const size_t e = (size_t)j * kR + r;
        const double2 rr = a.r[e], p = a.p[e], rt = a.rt[e], pt = a.pt[e];
        a.p[e] = make_double2(rr.x + be * p.x - bei * p.y, rr.y + bei * p.x + be * p.y);
        a.pt[e] = make_double2(rt.x + be * pt.x + bei * pt.y, rt.y - bei * pt.x + be * pt.y);
    }
}

// column-major (count per column) <-> RHS-interleaved, zero-filling unused columns
__global__ void kb_pack(const double2* __restrict__ src, int count, int c0, int ncol, double2* __restrict__ dst,
                        size_t rows) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * kR) return;
    const size_t j = e / kR;
    const int r = (int)(e % kR);
    dst[e] = (j < (size_t)count && r < ncol) ? src[(size_t)(c0 + r) * count + j] : make_double2(0.0, 0.0);
}

__global__ void kb_unpack(const double2* __restrict__ src, int count, int c0, int ncol, double2* __restrict__ dst) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)count * kR) return;
    const size_t j = e / kR;
    const int r = (int)(e % kR);
    if (r < ncol) dst[(size_t)(c0 + r) * count + j] = src[e];
}

}  // namespace

// Y = op(A) P for 16 RHS-interleaved vectors (rows = ld, padding rows zero).
int batch_apply(Context* ctx, int op, bool adjoint, const double2* P, double2* Y, const void* state) {
    const int count = ctx->count;
    const unsigned nb = (unsigned)((count + 63) / 64);  // 64 rows / columns per workgroup
    const BState* st = (const BState*)state;
    if (op == OP_PAIR64 || (op == OP_DENSE && ctx->mat_pair))
        return adjoint ? pair_batch_apply(ctx, op == OP_PAIR64, nullptr, P, nullptr, Y, st ? &st->all_done : nullptr)
                       : pair_batch_apply(ctx, op == OP_PAIR64, P, nullptr, Y, nullptr, st ? &st->all_done : nullptr);
    if (op == OP_FREE64 || op == OP_FREE32) {
        if (ctx->kind != IBIM_TUBE_TRADITIONAL) {
            set_error("batched matrix-free products need the traditional tube");
            return IBIM_ERR_ARG;
        }
        const bool fast = op == OP_FREE32;
        return adjoint ? pair_free_batch_apply(ctx, fast, nullptr, P, nullptr, Y, st ? &st->all_done : nullptr)
                       : pair_free_batch_apply(ctx, fast, P, nullptr, Y, nullptr, st ? &st->all_done : nullptr);
    }
    if (op != OP_DENSE) {
        set_error("batched products: unknown operator %d", op);
        return IBIM_ERR_ARG;
    }
    TimedScope ts(ctx, adjoint ? "batch_cols" : "batch_rows");
    if (ctx->mat_precision == IBIM_C128) {
        if (adjoint)
            hipLaunchKernelGGL(k_cols_stored<double2>, dim3(nb), dim3(kBT), 0, ctx->stream, ctx->mat.as<double2>(),
                               ctx->ld, count, P, Y, st);
        else
            hipLaunchKernelGGL(k_rows_stored<double2>, dim3(nb), dim3(kBT), 0, ctx->stream, ctx->mat.as<double2>(),
                               ctx->ld, count, P, Y, st);
    } else {
        if (adjoint)
            hipLaunchKernelGGL(k_cols_stored<float2>, dim3(nb), dim3(kBT), 0, ctx->stream, ctx->mat.as<float2>(),
                               ctx->ld, count, P, Y, st);
        else
            hipLaunchKernelGGL(k_rows_stored_c64, dim3(nb), dim3(kBT), 0, ctx->stream, ctx->mat.as<float2>(),
                               ctx->ld, count, P, Y, st);
    }
    IBIM_HIP(hipGetLastError());
    return IBIM_OK;
}

// V = op(A) P and W = op(A)^H Q: one pass over pair storage, else two passes.
int batch_apply2(Context* ctx, int op, const double2* P, const double2* Q, double2* V, double2* W,
                 const void* state) {
    const BState* st = (const BState*)state;
    if (op == OP_PAIR64 || (op == OP_DENSE && ctx->mat_pair))
        return pair_batch_apply(ctx, op == OP_PAIR64, P, Q, V, W, st ? &st->all_done : nullptr);
    if ((op == OP_FREE64 || op == OP_FREE32) && ctx->kind == IBIM_TUBE_TRADITIONAL)
        return pair_free_batch_apply(ctx, op == OP_FREE32, P, Q, V, W, st ? &st->all_done : nullptr);
    int rc = batch_apply(ctx, op, false, P, V, state);
    if (!rc) rc = batch_apply(ctx, op, true, Q, W, state);
    return rc;
}

// Host entry of the batched products: nvec <= 16 column-major host vectors.
int batch_apply_host(Context* ctx, int op, bool adjoint, int nvec, const double* pre, const double* pim, double* yre,
                     double* yim) {
    const int count = ctx->count;
    const size_t ld = ctx->ld, vec = ld * kR;
    if (ctx->xfer.ensure(sizeof(double2) * vec * 2)) return IBIM_ERR_NOMEM;
    double2* P = ctx->xfer.as<double2>();
    double2* Y = P + vec;
    std::vector<double2> h(vec, make_double2(0.0, 0.0));
    for (int r = 0; r < nvec; ++r)
        for (int j = 0; j < count; ++j)
            h[(size_t)j * kR + r] = make_double2(pre[(size_t)r * count + j], pim ? pim[(size_t)r * count + j] : 0.0);
    IBIM_HIP(hipMemcpyAsync(P, h.data(), sizeof(double2) * vec, hipMemcpyHostToDevice, ctx->stream));
    IBIM_HIP(hipMemsetAsync(Y, 0, sizeof(double2) * vec, ctx->stream));
    int rc = batch_apply(ctx, op, adjoint, P, Y, nullptr);
    if (rc) return rc;
    IBIM_HIP(hipMemcpyAsync(h.data(), Y, sizeof(double2) * vec, hipMemcpyDeviceToHost, ctx->stream));
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    for (int r = 0; r < nvec; ++r)
        for (int j = 0; j < count; ++j) {
            yre[(size_t)r * count + j] = h[(size_t)j * kR + r].x;
            yim[(size_t)r * count + j] = h[(size_t)j * kR + r].y;
        }
    return IBIM_OK;
}

// Blocks of the batched vector kernels (16 points each, grid-strided beyond).
static unsigned batch_blocks(const Context* ctx) {
    return (unsigned)std::min<size_t>(((size_t)ctx->count + 15) / 16, kMaxVB);
}

// Workspace of one batch: x, r, r~, p, p~, v, w, b and `extra` more
// interleaved vectors, per-block partials, the state and the residual history.
static size_t batch_workspace_bytes(const Context* ctx, int maxit, int extra) {
    const size_t vec = ctx->ld * kR;
    return sizeof(double2) * vec * (8 + extra) + sizeof(double) * (size_t)batch_blocks(ctx) * kR * 8 +
           sizeof(BState) + 256 + sizeof(double) * (size_t)(maxit + 1) * kR;
}

// The same layout at `base` (batch_workspace_bytes(ctx, maxit, extra) bytes, 256-B aligned).
static int batch_workspace_at(Context* ctx, void* base, int maxit, int extra, BArgs* a, unsigned* nvb_out,
                              double2** extra_out) {
    const int count = ctx->count;
    const size_t vec = ctx->ld * kR;
    const unsigned nvb = batch_blocks(ctx);
    const int nv = 8 + extra;
    double2* v0 = static_cast<double2*>(base);
    *a = BArgs{};
    a->x = v0;
    a->r = v0 + vec;
    a->rt = v0 + 2 * vec;
    a->p = v0 + 3 * vec;
    a->pt = v0 + 4 * vec;
    a->v = v0 + 5 * vec;
    a->w = v0 + 6 * vec;
    a->b = v0 + 7 * vec;
    if (extra_out) *extra_out = v0 + 8 * vec;
    a->part = (double*)(v0 + (size_t)nv * vec);
    a->st = (BState*)(((uintptr_t)(a->part + (size_t)nvb * kR * 8) + 255) & ~(uintptr_t)255);
    a->hist = (double*)(a->st + 1);
    a->count = count;
    a->ref_guards = 1;
    *nvb_out = nvb;
    // padding rows (count..ld) of every vector read as zero in the passes
    IBIM_HIP(hipMemsetAsync(v0, 0, sizeof(double2) * vec * nv, ctx->stream));
    return IBIM_OK;
}

static int batch_workspace(Context* ctx, int maxit, int extra, BArgs* a, unsigned* nvb_out, double2** extra_out) {
    if (ctx->batch.ensure(batch_workspace_bytes(ctx, maxit, extra))) return IBIM_ERR_NOMEM;
    return batch_workspace_at(ctx, ctx->batch.ptr, maxit, extra, a, nvb_out, extra_out);
}

struct ConvMask {
    int c[kR];
};

// R = B - Y in the columns flagged in `mask` (the others keep their R), with
// per-block partials of |B - Y|^2 (2 per column) for every column.
__global__ void __launch_bounds__(kBT) kb_replace(const double2* __restrict__ B, const double2* __restrict__ Y,
                                                  double2* __restrict__ R, int count, ConvMask mask,
                                                  double* __restrict__ part) {
    const int r = threadIdx.x & 15;
    double v[2] = {0.0, 0.0};
    for (int j = blockIdx.x * 16 + (threadIdx.x >> 4); j < count; j += gridDim.x * 16) {
        const size_t e = (size_t)j * kR + r;
        const double2 b = B[e], y = Y[e];
        const double2 rr = make_double2(b.x - y.x, b.y - y.y);
        if (mask.c[r]) R[e] = rr;
        v[0] += rr.x * rr.x;
        v[1] += rr.y * rr.y;
    }
    block_partials<2>(v, part + (size_t)blockIdx.x * kR * 8);
}

// Restart of the flagged columns from their (replaced) residual: r~ = conj(r),
// p = r, p~ = r~ (BICGSTAB.c:104-117), partials of the new <r, r~>.
__global__ void __launch_bounds__(kBT) kb_restart(BArgs a, ConvMask mask) {
    const int r = threadIdx.x & 15;
    double v[4] = {0, 0, 0, 0};
    for (int j = blockIdx.x * 16 + (threadIdx.x >> 4); j < a.count; j += gridDim.x * 16) {
        if (!mask.c[r]) continue;
        const size_t e = (size_t)j * kR + r;
        const double2 rr = a.r[e];
        const double2 c = make_double2(rr.x, -rr.y);
        a.p[e] = rr;
        a.rt[e] = c;
        a.pt[e] = c;
        v[0] += rr.x * c.x;
        v[1] += rr.y * c.y;
        v[2] += rr.y * c.x;
        v[3] += rr.x * c.y;
    }
    block_partials<4>(v, a.part + (size_t)blockIdx.x * kR * 8);
}

// rho of the restarted columns, which run again (no previous error).
__global__ void __launch_bounds__(kBT) kb_restart_finish(BArgs a, int nb, ConvMask mask) {
    __shared__ double tot[kR][8];
    reduce_all(a.part, nb, 4, tot);
    if (threadIdx.x < kR && mask.c[threadIdx.x]) {
        const int r = threadIdx.x;
        BState* s = a.st;
        s->rho[r][0] = tot[r][0] + tot[r][1];
        s->rho[r][1] = tot[r][2] - tot[r][3];
        s->err[r] = INFINITY;
        s->done[r] = 0;
        s->stop[r] = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) a.st->all_done = 0;
}

// Per-column sums of the first nv partial values (host, fixed order).
static int column_sums(Context* ctx, const double* part, unsigned nvb, int nv, std::vector<double>& out) {
    std::vector<double> h((size_t)nvb * kR * 8);
    IBIM_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    out.assign((size_t)kR * nv, 0.0);
    for (unsigned b = 0; b < nvb; ++b)
        for (int r = 0; r < kR; ++r)
            for (int q = 0; q < nv; ++q) out[(size_t)r * nv + q] += h[((size_t)b * kR + r) * 8 + q];
    return IBIM_OK;
}

// FP64 residual replacement for the batched BiCG (van der Vorst and Ye, as
// in the single-vector solver): at a host check, when a running column's
// recursive residual has dropped by `factor` since its last replacement or a
// column claims convergence, one batched pass of the FP64 operator over the
// current solutions gives r = b - A64 x for those columns. A claim whose true
// residual misses tol restarts that column's BiCG from the true residual: by
// then its shadow residual and search directions have converged with the
// recursive residual and cannot reduce the difference (on a 3,700-point
// sphere the recursive residual reached 3.5e-11 while the true one was 8e-8,
// where a resumed BiCG stayed for 100 iterations).
struct BReplace {
    int true_op;
    double factor;
    int replacements = 0;
    double next[kR];       // replace when the recursive residual falls below
    double true_err[kR];   // last true residual (-1 before the first)
    bool verified[kR];     // a convergence claim confirmed by the true residual
    int stalls[kR];        // stagnation stops checked and restarted so far
};
// Stagnation stops of a column that are checked against the true residual
// and restarted when it misses tol (the recursive residual of the FP32
// operator can stall above the FP64 one's level).
constexpr int kMaxStallRestarts = 3;

// Residual replacement at a host check (reliable batched BiCG): columns whose
// recursive residual fell by rep->factor since their last replacement, that
// claim convergence or stopped on stagnation get r = b - A64 x (one batched
// FP64 pass for all of them); refuted claims restart their column.
static int replacement_step(Context* ctx, const BArgs& a, unsigned nvb, BState* hs, BReplace* rep, double tol,
                            std::vector<double>& sums) {
    int rc = IBIM_OK;
    ConvMask m{};
    bool any = false;
    for (int r = 0; r < kR; ++r) {
        const bool claims = hs->done[r] > 0 && hs->stop[r] == 1 && !rep->verified[r];
        const bool stalled = hs->done[r] > 0 && hs->stop[r] == 3 && !rep->verified[r] &&
                             rep->stalls[r] < kMaxStallRestarts;
        const bool drop = hs->done[r] == 0 && hs->err[r] < rep->next[r];
        if (stalled) rep->stalls[r]++;
        m.c[r] = claims || stalled || drop;
        any = any || m.c[r];
    }
    // the pass computes every column: running columns that have come
    // at least half a replacement interval (sqrt(factor)) down since
    // their last replacement are replaced with the others, so the
    // columns stay in step and share the passes
    if (any)
        for (int r = 0; r < kR; ++r) {
            const double base = rep->true_err[r] >= 0.0 ? rep->true_err[r] : 1.0;
            if (hs->done[r] == 0 && hs->err[r] < base * sqrt(rep->factor)) m.c[r] = 1;
        }
    if (any) {
        // a.v is free between iterations: the true products of all columns
        rc = batch_apply(ctx, rep->true_op, false, a.x, a.v, nullptr);
        if (rc) return rc;
        hipLaunchKernelGGL(kb_replace, dim3(nvb), dim3(kBT), 0, ctx->stream, a.b, a.v, a.r, a.count, m,
                           a.part);
        IBIM_HIP(hipGetLastError());
        rc = column_sums(ctx, a.part, nvb, 2, sums);
        if (rc) return rc;
        rep->replacements++;
        ConvMask restart{};
        bool any_restart = false;
        for (int r = 0; r < kR; ++r) {
            if (!m.c[r]) continue;
            const double terr = sqrt(sums[(size_t)r * 2] + sums[(size_t)r * 2 + 1]) / hs->normb[r];
            rep->true_err[r] = terr;
            rep->next[r] = terr * rep->factor;
            if (hs->done[r] > 0) {  // a convergence claim or a checked stagnation stop
                if (terr < tol) {
                    rep->verified[r] = true;
                } else {
                    restart.c[r] = 1;
                    any_restart = true;
                }
            }
        }
        if (any_restart) {
            hipLaunchKernelGGL(kb_restart, dim3(nvb), dim3(kBT), 0, ctx->stream, a, restart);
            hipLaunchKernelGGL(kb_restart_finish, dim3(1), dim3(kBT), 0, ctx->stream, a, (int)nvb, restart);
            IBIM_HIP(hipGetLastError());
            IBIM_HIP(hipMemcpyAsync(hs, a.st, sizeof *hs, hipMemcpyDeviceToHost, ctx->stream));
            IBIM_HIP(hipStreamSynchronize(ctx->stream));
        }
    }
    return IBIM_OK;
}

// BiCG for the 16 interleaved right-hand sides in a.b (padding rows zero);
// the solutions are left in a.x and the final state in *hs. With `rep`, FP64
// residual replacement by rep->true_op.
static int bicg_batch_core(Context* ctx, int op, const BArgs& a, unsigned nvb, int maxit, double tol, int chunk,
                           BState* hs, BReplace* rep = nullptr) {
    hipLaunchKernelGGL(kb_init, dim3(nvb), dim3(kBT), 0, ctx->stream, a);
    hipLaunchKernelGGL(kb_init_finish, dim3(1), dim3(kBT), 0, ctx->stream, a, (int)nvb);
    IBIM_HIP(hipGetLastError());
    if (chunk <= 0) chunk = ctx->count > 20000 ? 2 : 16;
    int it = 1;
    IBIM_HIP(hipMemcpyAsync(hs, a.st, sizeof *hs, hipMemcpyDeviceToHost, ctx->stream));
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    if (rep)
        for (int r = 0; r < kR; ++r) {
            rep->next[r] = rep->factor;
            rep->true_err[r] = -1.0;
            rep->verified[r] = false;
            rep->stalls[r] = 0;
        }
    std::vector<double> sums;
    // one iteration: the same launches every time (the kernels read the
    // iteration from the state), so a chunk can be captured once and replayed
    auto launch_iteration = [&]() -> int {
        int rc = batch_apply2(ctx, op, a.p, a.pt, a.v, a.w, a.st);
        if (rc) return rc;
        hipLaunchKernelGGL(kb_dot, dim3(nvb), dim3(kBT), 0, ctx->stream, a);
        hipLaunchKernelGGL(kb_alpha, dim3(1), dim3(kBT), 0, ctx->stream, a, (int)nvb);
        hipLaunchKernelGGL(kb_update_xr, dim3(nvb), dim3(kBT), 0, ctx->stream, a);
        hipLaunchKernelGGL(kb_beta, dim3(1), dim3(kBT), 0, ctx->stream, a, (int)nvb, tol);
        hipLaunchKernelGGL(kb_update_p, dim3(nvb), dim3(kBT), 0, ctx->stream, a);
        return IBIM_OK;
    };
    // HIP graph of `chunk` iterations (kernels only), captured after the first
    // chunk ran eagerly (the pair passes time their workgroup shapes there)
    struct GraphExec {
        hipGraphExec_t g = nullptr;
        ~GraphExec() {
            if (g) (void)hipGraphExecDestroy(g);
        }
    } chunk_graph;
    const bool graphs = ctx->use_graphs && !ctx->profile;
    while (it <= maxit) {
        if (!hs->all_done) {
            const int last = std::min(it + chunk - 1, maxit);
            if (graphs && it > 1 && last - it + 1 == chunk) {
                if (!chunk_graph.g) {
                    hipGraph_t g = nullptr;
                    IBIM_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                    int rc = IBIM_OK;
                    for (int k = 0; k < chunk && !rc; ++k) rc = launch_iteration();
                    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
                    if (rc) {
                        if (g) (void)hipGraphDestroy(g);
                        return rc;
                    }
                    IBIM_HIP(e);
                    const hipError_t ei = hipGraphInstantiate(&chunk_graph.g, g, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(g);
                    IBIM_HIP(ei);
                }
                IBIM_HIP(hipGraphLaunch(chunk_graph.g, ctx->stream));
                it = last + 1;
            } else {
                for (; it <= last; ++it) {
                    int rc = launch_iteration();
                    if (rc) return rc;
                }
            }
            IBIM_HIP(hipGetLastError());
            IBIM_HIP(hipMemcpyAsync(hs, a.st, sizeof *hs, hipMemcpyDeviceToHost, ctx->stream));
            IBIM_HIP(hipStreamSynchronize(ctx->stream));
            double worst = 0.0;
            for (int r = 0; r < kR; ++r)
                if (hs->done[r] == 0) worst = std::max(worst, hs->err[r]);
            if (!hs->all_done) report_progress(ctx, "batched BiCG", it - 1, worst);
        }
        if (rep) {
            int rc = replacement_step(ctx, a, nvb, hs, rep, tol, sums);
            if (rc) return rc;
        }
        if (hs->all_done) break;
    }
    return IBIM_OK;
}

// After a batch's iterations: columns that stopped without a verified
// convergence claim report the true FP64 residual of their final iterate (the
// reliable solver), the solutions go to density columns c0 .. c0 + ncol - 1,
// and each column's iterations, residual and stop reason to the out arrays.
static int batch_finish(Context* ctx, const BArgs& a, unsigned nvb, const BState& hs, BReplace* rep, int c0,
                        int ncol, int maxit, int* iters, double* errs, int* stops, double* hist) {
    const int count = ctx->count;
    int rc = IBIM_OK;
    if (rep) {
        // columns that stopped without a verified convergence claim
        // (stagnation, NaN, the iteration cap) report the true FP64 residual
        // of their final iterate, not that of their last replacement
        ConvMask m{};
        bool any = false;
        for (int r = 0; r < ncol; ++r) {
            m.c[r] = hs.done[r] >= 0 && !rep->verified[r];
            any = any || m.c[r];
        }
        if (any) {
            rc = batch_apply(ctx, rep->true_op, false, a.x, a.v, nullptr);
            if (rc) return rc;
            hipLaunchKernelGGL(kb_replace, dim3(nvb), dim3(kBT), 0, ctx->stream, a.b, a.v, a.r, a.count, m, a.part);
            IBIM_HIP(hipGetLastError());
            std::vector<double> sums;
            rc = column_sums(ctx, a.part, nvb, 2, sums);
            if (rc) return rc;
            rep->replacements++;
            for (int r = 0; r < ncol; ++r)
                if (m.c[r]) rep->true_err[r] = sqrt(sums[(size_t)r * 2] + sums[(size_t)r * 2 + 1]) / hs.normb[r];
        }
    }
    hipLaunchKernelGGL(kb_unpack, dim3((unsigned)(((size_t)count * kR + 255) / 256)), dim3(256), 0, ctx->stream, a.x,
                       count, c0, ncol, ctx->density.as<double2>());
    IBIM_HIP(hipGetLastError());
    std::vector<double> h;
    if (hist) {
        h.resize((size_t)(maxit + 1) * kR);
        IBIM_HIP(hipMemcpyAsync(h.data(), a.hist, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    }
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    for (int r = 0; r < ncol; ++r) {
        const int d = hs.done[r];
        iters[c0 + r] = d > 0 ? d : (d < 0 ? 0 : maxit);
        errs[c0 + r] = d < 0 ? 0.0 : (rep && rep->true_err[r] >= 0.0 ? rep->true_err[r] : hs.err[r]);
        stops[c0 + r] = rep && rep->verified[r] ? 0 : d == 0 ? 1 : (hs.stop[r] == 1 ? 0 : hs.stop[r]);
        if (hist && c0 + r == 0)
            for (int i = 1; i <= iters[0]; ++i) hist[i - 1] = h[(size_t)i * kR];
    }
    return IBIM_OK;
}

// BiCG on columns [c0, c0 + ncol) of ctx->rhs (ncol <= 16) with operator `op`;
// densities into the same columns of ctx->density. Per-column iteration counts,
// residuals and stop reasons go to the out arrays.
static int bicg_batch(Context* ctx, int op, int c0, int ncol, int maxit, double tol, int chunk, int* iters,
                      double* errs, int* stops, double* hist, BReplace* rep = nullptr) {
    const int count = ctx->count;
    BArgs a;
    unsigned nvb;
    int rc = batch_workspace(ctx, maxit, 0, &a, &nvb, nullptr);
    if (rc) return rc;
    const unsigned pb = (unsigned)((ctx->ld * kR + 255) / 256);
    hipLaunchKernelGGL(kb_pack, dim3(pb), dim3(256), 0, ctx->stream, ctx->rhs.as<double2>(), count, c0, ncol,
                       (double2*)a.b, ctx->ld);
    BState hs{};
    rc = bicg_batch_core(ctx, op, a, nvb, maxit, tol, chunk, &hs, rep);
    if (rc) return rc;
    return batch_finish(ctx, a, nvb, hs, rep, c0, ncol, maxit, iters, errs, stops, hist);
}

// X += D; and R = B - Y with per-block partials of |R|^2 (2 per column), the
// columns flagged in `conv` getting R = 0 (their correction is then zero).
__global__ void __launch_bounds__(kBT) kb_accumulate(double2* __restrict__ X, const double2* __restrict__ D, int count) {
    const int r = threadIdx.x & 15;
    for (int j = blockIdx.x * 16 + (threadIdx.x >> 4); j < count; j += gridDim.x * 16) {
        const size_t e = (size_t)j * kR + r;
        const double2 x = X[e], d = D[e];
        X[e] = make_double2(x.x + d.x, x.y + d.y);
    }
}

__global__ void __launch_bounds__(kBT) kb_residual(const double2* __restrict__ B, const double2* __restrict__ Y,
                                                   double2* __restrict__ R, int count, ConvMask conv,
                                                   double* __restrict__ part) {
    const int r = threadIdx.x & 15;
    double v[2] = {0.0, 0.0};
    for (int j = blockIdx.x * 16 + (threadIdx.x >> 4); j < count; j += gridDim.x * 16) {
        const size_t e = (size_t)j * kR + r;
        const double2 b = B[e], y = Y[e];
        const double2 rr = make_double2(b.x - y.x, b.y - y.y);
        R[e] = conv.c[r] ? make_double2(0.0, 0.0) : rr;
        v[0] += rr.x * rr.x;
        v[1] += rr.y * rr.y;
    }
    block_partials<2>(v, part + (size_t)blockIdx.x * kR * 8);
}

// FP64 iterative refinement around batched BiCG on the FP32 operator, 16
// columns at a time: x = 0, r = b; repeat { A32 d = r by batched BiCG; x += d;
// r = b - A64 x by a batched FP64 pass } until every column's ||r|| / ||b|| < tol.
static int refined_batch(Context* ctx, int inner, int outer_op, int c0, int ncol, const ibim_solve_opts* opts,
                         int* iters, double* errs, int* stops, int* outers) {
    const int count = ctx->count;
    BArgs a;
    unsigned nvb;
    double2* ex;
    int rc = batch_workspace(ctx, opts->maxit, 2, &a, &nvb, &ex);
    if (rc) return rc;
    a.ref_guards = 0;
    const size_t vec = ctx->ld * kR;
    double2* X = ex;        // accumulated solution
    double2* B = ex + vec;  // the right-hand sides
    double2* Rb = (double2*)a.b;
    const unsigned pb = (unsigned)((vec + 255) / 256);
    hipLaunchKernelGGL(kb_pack, dim3(pb), dim3(256), 0, ctx->stream, ctx->rhs.as<double2>(), count, c0, ncol, B,
                       ctx->ld);
    IBIM_HIP(hipMemcpyAsync(Rb, B, sizeof(double2) * vec, hipMemcpyDeviceToDevice, ctx->stream));
    // ||b|| per column
    ConvMask none{};
    hipLaunchKernelGGL(kb_residual, dim3(nvb), dim3(kBT), 0, ctx->stream, B, X, a.v, count, none, a.part);
    IBIM_HIP(hipGetLastError());
    std::vector<double> sums;
    rc = column_sums(ctx, a.part, nvb, 2, sums);
    if (rc) return rc;
    double normb[kR], res[kR];
    ConvMask conv{};
    for (int r = 0; r < kR; ++r) {
        normb[r] = sqrt(sums[(size_t)r * 2] + sums[(size_t)r * 2 + 1]);
        res[r] = normb[r] > 0.0 ? 1.0 : 0.0;
        conv.c[r] = normb[r] == 0.0 || r >= ncol;
    }
    const double inner_tol = opts->inner_tol > 0 ? opts->inner_tol : 1e-4;
    const int max_outer = opts->max_outer > 0 ? opts->max_outer : 20;
    int outer = 0;
    int col_its[kR] = {};  // each column's own inner iterations, summed over the steps
    for (outer = 1; outer <= max_outer; ++outer) {
        double worst = 0.0;
        for (int r = 0; r < ncol; ++r)
            if (!conv.c[r]) worst = std::max(worst, res[r]);
        if (worst == 0.0) break;
        const double itol = outer == 1 ? inner_tol : std::max(inner_tol, std::min(0.5, 0.3 * opts->tol / worst));
        BState hs{};
        rc = bicg_batch_core(ctx, inner, a, nvb, opts->maxit, itol, opts->chunk, &hs);
        if (rc) return rc;
        // a converged column has a zero right-hand side: zero data, no iterations
        for (int r = 0; r < kR; ++r) col_its[r] += hs.done[r] > 0 ? hs.done[r] : (hs.done[r] < 0 ? 0 : opts->maxit);
        hipLaunchKernelGGL(kb_accumulate, dim3(nvb), dim3(kBT), 0, ctx->stream, X, a.x, count);
        rc = batch_apply(ctx, outer_op, false, X, a.v, nullptr);
        if (rc) return rc;
        hipLaunchKernelGGL(kb_residual, dim3(nvb), dim3(kBT), 0, ctx->stream, B, a.v, Rb, count, conv, a.part);
        IBIM_HIP(hipGetLastError());
        rc = column_sums(ctx, a.part, nvb, 2, sums);
        if (rc) return rc;
        bool all = true;
        for (int r = 0; r < ncol; ++r) {
            if (conv.c[r]) continue;
            res[r] = sqrt(sums[(size_t)r * 2] + sums[(size_t)r * 2 + 1]) / normb[r];
            if (res[r] < opts->tol) conv.c[r] = 1;
            all = all && conv.c[r];
        }
        // columns that converged now get a zero right-hand side in the next refinement step
        hipLaunchKernelGGL(kb_residual, dim3(nvb), dim3(kBT), 0, ctx->stream, B, a.v, Rb, count, conv, a.part);
        IBIM_HIP(hipGetLastError());
        if (all) break;
    }
    if (outer > max_outer) outer = max_outer;
    hipLaunchKernelGGL(kb_unpack, dim3((unsigned)(((size_t)count * kR + 255) / 256)), dim3(256), 0, ctx->stream, X,
                       count, c0, ncol, ctx->density.as<double2>());
    IBIM_HIP(hipGetLastError());
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    for (int r = 0; r < ncol; ++r) {
        iters[c0 + r] = col_its[r];
        errs[c0 + r] = res[r];
        stops[c0 + r] = normb[r] == 0.0 ? 4 : (res[r] < opts->tol ? 0 : 1);
        outers[c0 + r] = outer;
    }
    return IBIM_OK;
}

// Refinement for several right-hand sides, 16 per batch (see refined_batch).
int run_refined_batched(Context* ctx, const ibim_solve_opts* opts, ibim_solve_info* info) {
    if (ctx->kind != IBIM_TUBE_TRADITIONAL && ctx->mat_precision != IBIM_C64) return IBIM_ERR_ARG;
    const int inner = ctx->mat_precision == IBIM_C64 ? OP_DENSE : OP_FREE32;
    const int outer_op = fp64_operator(ctx);
    const int n = ctx->nrhs;
    std::vector<int> iters(n), stops(n), outers(n);
    std::vector<double> errs(n);
    for (int c0 = 0; c0 < n; c0 += kR) {
        int rc = refined_batch(ctx, inner, outer_op, c0, std::min(kR, n - c0), opts, iters.data(), errs.data(),
                               stops.data(), outers.data());
        if (rc) return rc;
    }
    int total = 0, status = 0, outer = 0;
    double worst = 0.0;
    for (int c = 0; c < n; ++c) total += iters[c];  // summed over the columns, as the batched BiCG counts
    for (int c = 0; c < n; ++c) {
        worst = std::max(worst, errs[c]);
        outer = std::max(outer, outers[c]);
        if (stops[c] && stops[c] != 4) status = stops[c];
    }
    info->iterations = total;
    info->outer = outer;
    info->rel_residual = worst;
    info->status = status;
    ctx->last_iters = total;
    return IBIM_OK;
}

// BiCG on the FP32 operator with FP64 residual replacement (as run_reliable),
// 16 right-hand sides per batch.
int run_reliable_batched(Context* ctx, const ibim_solve_opts* opts, ibim_solve_info* info, double* hist) {
    const int op = ctx->mat_precision == IBIM_C64 ? OP_DENSE : OP_FREE32;
    if (op == OP_FREE32 && ctx->kind != IBIM_TUBE_TRADITIONAL) return run_reliable(ctx, opts, info, hist);
    const int n = ctx->nrhs;
    std::vector<int> iters(n), stops(n);
    std::vector<double> errs(n);
    int outer = 0;
    for (int c0 = 0; c0 < n; c0 += kR) {
        BReplace rep{};
        rep.true_op = fp64_operator(ctx);
        rep.factor = opts->inner_tol > 0 && opts->inner_tol < 1 ? opts->inner_tol : 1e-2;
        // replacements happen at host checks: every 2 iterations unless asked
        const int chunk = opts->chunk > 0 ? opts->chunk : 2;
        int rc = bicg_batch(ctx, op, c0, std::min(kR, n - c0), opts->maxit, opts->tol, chunk, iters.data(),
                            errs.data(), stops.data(), c0 == 0 ? hist : nullptr, &rep);
        if (rc) return rc;
        outer += rep.replacements;
    }
    int total = 0, status = 0;
    double worst = 0.0;
    for (int c = 0; c < n; ++c) {
        total += iters[c];
        worst = std::max(worst, errs[c]);
        if (stops[c] && stops[c] != 4) status = stops[c];
    }
    info->iterations = total;
    info->outer = outer;
    info->rel_residual = worst;
    info->status = status;
    ctx->last_iters = total;
    return IBIM_OK;
}

// All right-hand sides, 16 per batch. The operator is the stored one when
// assembled, else the FP64 entries recomputed per pass (as run_bicg).
int run_bicg_batched(Context* ctx, const ibim_solve_opts* opts, ibim_solve_info* info, double* hist) {
    const int op = ctx->mat_precision >= 0 ? OP_DENSE : OP_FREE64;
    if (op != OP_DENSE && ctx->kind != IBIM_TUBE_TRADITIONAL) return run_bicg(ctx, opts, info, hist);
    const int n = ctx->nrhs;
    std::vector<int> iters(n), stops(n);
    std::vector<double> errs(n);
    for (int c0 = 0; c0 < n; c0 += kR) {
        int rc = bicg_batch(ctx, op, c0, std::min(kR, n - c0), opts->maxit, opts->tol, opts->chunk, iters.data(),
                            errs.data(), stops.data(), c0 == 0 ? hist : nullptr);
        if (rc) return rc;
    }
    int total = 0, status = 0;
    double worst = 0.0;
    for (int c = 0; c < n; ++c) {
        total += iters[c];
        worst = std::max(worst, errs[c]);
        if (stops[c] && stops[c] != 4) status = stops[c];
    }
    info->iterations = total;
    info->outer = 1;
    info->rel_residual = worst;
    info->status = status;
    ctx->last_iters = total;
    return IBIM_OK;
}

// ------------------------------------------------------ several problems --
// Batched BiCG (IBIM_SOLVE_BICG) or reliable batched BiCG (IBIM_SOLVE_RELIABLE)
// of every right-hand side of several pair-stored problems at once. A member
// is one problem with one batch of up to 16 of its columns; all members
// iterate in step, and each iteration applies every running member's operator
// in ONE persistent launch (pair_group_pass: a queue of 64-row items over all
// members, largest problems first), then runs each member's near field and
// BiCG vector kernels. Host checks (and residual replacements) come every
// `chunk` iterations as in the one-problem solvers, so each member takes the
// same steps, and gives the same bits, as its own batched solve with that chunk.
// All work goes to the first problem's stream for the duration of the call.
namespace {

struct GroupRun {
    Context* ctx = nullptr;
    int c0 = 0, ncol = 0;
    BArgs a{};
    unsigned nvb = 0;
    BState hs{};
    BReplace rep{};
};

}  // namespace

int run_group(Context* const* ctxs, int n, int method, const ibim_solve_opts* opts, ibim_solve_info* infos) {
    if (method == IBIM_SOLVE_GMRES || method == IBIM_SOLVE_BLOCK_GMRES) return run_group_gmres(ctxs, n, opts, infos);
    if (method != IBIM_SOLVE_BICG && method != IBIM_SOLVE_RELIABLE) {
        set_error("ibim_solve_group: BiCG, reliable BiCG, GMRES or block GMRES only");
        return IBIM_ERR_ARG;
    }
    const bool mixed = method == IBIM_SOLVE_RELIABLE;
    int source = 0;
    if (int rc = pair_group_source(ctxs, n, mixed, &source)) return rc;
    Context* lead = ctxs[0];
    const int maxit = opts->maxit;
    SharedStream shared;
    if (int rc = shared.share(ctxs, n)) return rc;
    // members, the largest problems first (their items start first)
    std::vector<GroupRun> runs;
    std::vector<int> order(n);
    for (int c = 0; c < n; ++c) order[c] = c;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return ctxs[x]->count > ctxs[y]->count; });
    for (int c : order)
        for (int c0 = 0; c0 < ctxs[c]->nrhs; c0 += kR) {
            GroupRun g;
            g.ctx = ctxs[c];
            g.c0 = c0;
            g.ncol = std::min(kR, ctxs[c]->nrhs - c0);
            runs.push_back(g);
        }
    const int nm = (int)runs.size();
    // workspaces, then the queue counter and the member table
    std::vector<size_t> off(nm + 1, 0);
    for (int m = 0; m < nm; ++m)
        off[m + 1] = off[m] + ((batch_workspace_bytes(runs[m].ctx, maxit, 0) + 255) & ~(size_t)255);
    const size_t table = off[nm] + 256;
    if (lead->group.ensure(table + sizeof(GroupMember) * nm)) return IBIM_ERR_NOMEM;
    char* base = lead->group.as<char>();
    int* queue = reinterpret_cast<int*>(base + off[nm]);
    GroupMember* dev_members = reinterpret_cast<GroupMember*>(base + table);
    std::vector<GroupMember> members(nm);
    int items = 0;
    for (int m = 0; m < nm; ++m) {
        GroupRun& g = runs[m];
        Context* ctx = g.ctx;
        if (int rc = batch_workspace_at(ctx, base + off[m], maxit, 0, &g.a, &g.nvb, nullptr)) return rc;
        const unsigned pb = (unsigned)((ctx->ld * kR + 255) / 256);
        hipLaunchKernelGGL(kb_pack, dim3(pb), dim3(256), 0, lead->stream, ctx->rhs.as<double2>(), ctx->count, g.c0,
                           g.ncol, (double2*)g.a.b, ctx->ld);
        hipLaunchKernelGGL(kb_init, dim3(g.nvb), dim3(kBT), 0, lead->stream, g.a);
        hipLaunchKernelGGL(kb_init_finish, dim3(1), dim3(kBT), 0, lead->stream, g.a, (int)g.nvb);
        IBIM_HIP(hipGetLastError());
        GroupMember& d = members[m];
        d.F = ctx->mat.ptr;
        d.blk = ctx->pair_blk.as<int>();
        d.pts = ctx->pair_pts.ptr;
        d.P = g.a.p;
        d.Q = g.a.pt;
        d.Y = g.a.v;
        d.Z = g.a.w;
        d.stop = &g.a.st->all_done;
        d.prm = ctx->prm;
        d.count = ctx->count;
        d.item0 = items;
        items += (ctx->count + 63) / 64;
        if (mixed) {
            g.rep.true_op = fp64_operator(ctx);
            g.rep.factor = opts->inner_tol > 0 && opts->inner_tol < 1 ? opts->inner_tol : 1e-2;
            for (int r = 0; r < kR; ++r) {
                g.rep.next[r] = g.rep.factor;
                g.rep.true_err[r] = -1.0;
                g.rep.verified[r] = false;
                g.rep.stalls[r] = 0;
            }
        }
    }
    IBIM_HIP(hipMemcpyAsync(dev_members, members.data(), sizeof(GroupMember) * nm, hipMemcpyHostToDevice,
                            lead->stream));
    for (GroupRun& g : runs) IBIM_HIP(hipMemcpyAsync(&g.hs, g.a.st, sizeof g.hs, hipMemcpyDeviceToHost, lead->stream));
    IBIM_HIP(hipStreamSynchronize(lead->stream));
    // host checks as in the one-problem solvers: every 2 iterations for the
    // reliable solver unless asked; BiCG: 2 when the largest problem has more
    // than 20,000 points, else 16
    int chunk = opts->chunk;
    if (chunk <= 0) chunk = mixed ? 2 : (runs[0].ctx->count > 20000 ? 2 : 16);
    std::vector<double> sums;
    // one iteration of every batch still running at the last host check: the
    // same launches every time (the kernels read the iteration and stop flags
    // from the device states), so a chunk of them can be captured and replayed
    auto launch_iteration = [&]() -> int {
        if (int rc = pair_group_pass(lead, source, dev_members, nm, items, queue)) return rc;
        for (GroupRun& g : runs) {
            if (g.hs.all_done) continue;  // stays done until a host check restarts it
            const BArgs& a = g.a;
            if (int rc = pair_batch_near(g.ctx, a.p, a.pt, a.v, a.w, &a.st->all_done)) return rc;
            hipLaunchKernelGGL(kb_dot, dim3(g.nvb), dim3(kBT), 0, lead->stream, a);
            hipLaunchKernelGGL(kb_alpha, dim3(1), dim3(kBT), 0, lead->stream, a, (int)g.nvb);
            hipLaunchKernelGGL(kb_update_xr, dim3(g.nvb), dim3(kBT), 0, lead->stream, a);
            hipLaunchKernelGGL(kb_beta, dim3(1), dim3(kBT), 0, lead->stream, a, (int)g.nvb, opts->tol);
            hipLaunchKernelGGL(kb_update_p, dim3(g.nvb), dim3(kBT), 0, lead->stream, a);
        }
        return IBIM_OK;
    };
    // HIP graph replay (ibim_set_graphs on the first problem, not while
    // profiling): full chunks after the first are captured once per set of
    // running batches and replayed; a host check that changes the set (a batch
    // finished or was restarted) makes the next chunk capture again
    struct GraphExec {
        hipGraphExec_t g = nullptr;
        ~GraphExec() {
            if (g) (void)hipGraphExecDestroy(g);
        }
    } chunk_graph;
    std::vector<char> graph_live;  // the running batches the captured chunk launches
    const bool graphs = lead->use_graphs && !lead->profile;
    int it = 1;
    while (it <= maxit) {
        bool live = false;
        std::vector<char> live_set(runs.size());
        for (size_t m = 0; m < runs.size(); ++m) {
            live_set[m] = !runs[m].hs.all_done;
            live = live || live_set[m];
        }
        if (live) {
            const int last = std::min(it + chunk - 1, maxit);
            if (graphs && it > 1 && last - it + 1 == chunk) {
                if (chunk_graph.g && graph_live != live_set) {
                    IBIM_HIP(hipGraphExecDestroy(chunk_graph.g));
                    chunk_graph.g = nullptr;
                }
                if (!chunk_graph.g) {
                    hipGraph_t g = nullptr;
                    IBIM_HIP(hipStreamBeginCapture(lead->stream, hipStreamCaptureModeThreadLocal));
                    int rc = IBIM_OK;
                    for (int c = 0; c < chunk && !rc; ++c) rc = launch_iteration();
                    const hipError_t e = hipStreamEndCapture(lead->stream, &g);
                    if (rc) {
                        if (g) (void)hipGraphDestroy(g);
                        return rc;
                    }
                    IBIM_HIP(e);
                    const hipError_t ei = hipGraphInstantiate(&chunk_graph.g, g, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(g);
                    IBIM_HIP(ei);
                    graph_live = live_set;
                }
                IBIM_HIP(hipGraphLaunch(chunk_graph.g, lead->stream));
                it = last + 1;
            } else {
                for (; it <= last; ++it)
                    if (int rc = launch_iteration()) return rc;
            }
            IBIM_HIP(hipGetLastError());
            double worst = 0.0;
            for (GroupRun& g : runs) {
                if (g.hs.all_done) continue;
                IBIM_HIP(hipMemcpyAsync(&g.hs, g.a.st, sizeof g.hs, hipMemcpyDeviceToHost, lead->stream));
            }
            IBIM_HIP(hipStreamSynchronize(lead->stream));
            for (const GroupRun& g : runs)
                for (int r = 0; r < kR; ++r)
                    if (g.hs.done[r] == 0) worst = std::max(worst, g.hs.err[r]);
            report_progress(lead, "grouped batched BiCG", it - 1, worst);
        }
        bool all = true;
        for (GroupRun& g : runs) {
            if (mixed)
                if (int rc = replacement_step(g.ctx, g.a, g.nvb, &g.hs, &g.rep, opts->tol, sums)) return rc;
            all = all && g.hs.all_done;
        }
        if (all) break;
    }
    // per problem: its columns' results, summed as the one-problem solvers do
    for (int c = 0; c < n; ++c) {
        Context* ctx = ctxs[c];
        const int nr = ctx->nrhs;
        std::vector<int> iters(nr), stops(nr);
        std::vector<double> errs(nr);
        int replacements = 0;
        for (GroupRun& g : runs) {
            if (g.ctx != ctx) continue;
            if (int rc = batch_finish(ctx, g.a, g.nvb, g.hs, mixed ? &g.rep : nullptr, g.c0, g.ncol, maxit,
                                      iters.data(), errs.data(), stops.data(), nullptr))
                return rc;
            replacements += g.rep.replacements;
        }
        ibim_solve_info& info = infos[c];
        info = ibim_solve_info{};
        double worst = 0.0;
        for (int col = 0; col < nr; ++col) {
            info.iterations += iters[col];
            worst = std::max(worst, errs[col]);
            if (stops[col] && stops[col] != 4) info.status = stops[col];
        }
        info.outer = mixed ? replacements : 1;
        info.rel_residual = worst;
        ctx->last_iters = info.iterations;
    }
    return IBIM_OK;
}

}  // namespace ibim
