// ibim_gmres.hip -- IBIM_SOLVE_GMRES: FP64 iterative refinement around GMRES
// on the stored operator, for one right-hand side at a time or for 16 at once.
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
// residuals FP64); on complex128 storage the FP64 operator is the stored one
// and a single GMRES run is the FP64 solver. With nothing stored, GMRES runs on
// the mixed matrix-free operator (OP_FREE32) and the residuals on the FP64
// matrix-free one.
//
// Several right-hand sides (ibim_set_nrhs, opts->batch == 0) run 16 columns
// per batch: the vectors are right-hand-side interleaved (point j, column r at
// j * 16 + r), one 16-column pass of ibim_batch.hip / ibim_pair.hip applies
// the operator to all of them (Y = A P only; on complex64 pairs with
// ibim_set_fp32_products, on the FP32 matrix cores), and each column keeps its
// own Hessenberg matrix, rotations and stop rule. A column that has reached
// its target gets zero basis vectors from then on (its products are exact
// zeros), so it neither moves nor disturbs the others. The Gram-Schmidt
// kernels are the single-vector ones with a column index: with one column
// they do the same operations in the same order.
#include <algorithm>
#include <cmath>
#include <complex>
#include <type_traits>
#include <vector>

#include "ibim_internal.hpp"
#include "ibim_mfma.hpp"

namespace ibim {

namespace {

constexpr int kGT = 256;          // threads per block of the vector kernels
constexpr int kMaxBasis = 256;    // Krylov vectors per GMRES cycle (then restart), one column
constexpr int kMaxBasisB = 128;   // the same for 16-column batches (vectors 16x as long)

template <int C>
constexpr int max_basis() {
    return C == 1 ? kMaxBasis : kMaxBasisB;
}
// element chunks per basis vector in the dot products (one wave each)
template <int C>
constexpr int dot_chunks() {
    return C == 1 ? 32 : 256;
}

// part[(i D + c) C + col] = sum over chunk c of conj(V_i[e]) w[e] over the
// elements e of column col (e % C == col), D = dot_chunks<C>(): one wave per
// (chunk, basis vector) -- four basis vectors per block -- whose lanes stride
// over the chunk (chunk lengths are multiples of 64, so lane l only ever sees
// column l % C), then a fixed butterfly over the lanes of each column.
template <int C>
__global__ void __launch_bounds__(kGT) k_gm_dots(const double2* __restrict__ V, size_t vs, int n, int nv,
                                                 const double2* __restrict__ w, double2* __restrict__ part) {
    constexpr int D = dot_chunks<C>();
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.y * (kGT / 64) + (threadIdx.x >> 6);
    if (i >= nv) return;  // a whole wave; no barrier below
    const int c = blockIdx.x;
    const int len = ((n + D - 1) / D + 63) / 64 * 64;
    const int e1 = min(n, (c + 1) * len);
    const double2* v = V + (size_t)i * vs;
    double ar = 0.0, ai = 0.0;
#pragma unroll 4
    for (int e = c * len + lane; e < e1; e += 64) {
        const double2 a = v[e], b = w[e];
        ar += a.x * b.x + a.y * b.y;
        ai += a.x * b.y - a.y * b.x;
    }
#pragma unroll
    for (int off = 32; off >= C; off >>= 1) {
        ar += __shfl_xor(ar, off);
        ai += __shfl_xor(ai, off);
    }
    if (lane < C) part[((size_t)i * D + c) * C + lane] = make_double2(ar, ai);
}

// out[i C + col] = sum_c part[(i D + c) C + col] in order (one thread per entry)
template <int C>
__global__ void k_gm_dots_sum(const double2* __restrict__ part, int n, double2* __restrict__ out) {
    constexpr int D = dot_chunks<C>();
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * C) return;
    const int i = t / C, col = t % C;
    double sr = 0.0, si = 0.0;
    for (int c = 0; c < D; ++c) {
        sr += part[((size_t)i * D + c) * C + col].x;
        si += part[((size_t)i * D + c) * C + col].y;
    }
    out[t] = make_double2(sr, si);
}

// w[e] += sign sum_{i < nv} V_i[e] h[i C + e % C]  (nv <= max_basis<C>() + 1)
template <int C>
__global__ void __launch_bounds__(kGT) k_gm_combine(const double2* __restrict__ V, size_t vs, int n,
                                                    const double2* __restrict__ h, int nv, double sign,
                                                    double2* __restrict__ w) {
    __shared__ double2 hs[(max_basis<C>() + 1) * C];
    for (int i = threadIdx.x; i < nv * C; i += kGT) hs[i] = h[i];
    __syncthreads();
    const int e = blockIdx.x * kGT + threadIdx.x;
    if (e >= n) return;
    const int col = e % C;
    double cr = 0.0, ci = 0.0;
    for (int i = 0; i < nv; ++i) {
        const double2 v = V[(size_t)i * vs + e], c = hs[i * C + col];
        cr += v.x * c.x - v.y * c.y;
        ci += v.x * c.y + v.y * c.x;
    }
    const double2 o = w[e];
    w[e] = make_double2(o.x + sign * cr, o.y + sign * ci);
}

// One real factor per column.
template <int C>
struct Scales {
    double s[C];
};

// v[e] = s[e % C] w[e]
template <int C>
__global__ void __launch_bounds__(kGT) k_gm_scale(const double2* __restrict__ w, Scales<C> s, int n,
                                                  double2* __restrict__ v) {
    const int e = blockIdx.x * kGT + threadIdx.x;
    if (e < n) {
        const double f = s.s[e % C];
        v[e] = make_double2(f * w[e].x, f * w[e].y);
    }
}

// out = a - b; x += d
__global__ void __launch_bounds__(kGT) k_gm_sub(const double2* __restrict__ a, const double2* __restrict__ b,
                                                int n, double2* __restrict__ out) {
    const int e = blockIdx.x * kGT + threadIdx.x;
    if (e < n) out[e] = make_double2(a[e].x - b[e].x, a[e].y - b[e].y);
}

__global__ void __launch_bounds__(kGT) k_gm_add(double2* __restrict__ x, const double2* __restrict__ d, int n) {
    const int e = blockIdx.x * kGT + threadIdx.x;
    if (e < n) x[e] = make_double2(x[e].x + d[e].x, x[e].y + d[e].y);
}

// Columns c0 .. c0 + ncol - 1 of the column-major right-hand sides into a
// 16-column interleaved block (other columns zero), and back.
__global__ void k_gm_pack(const double2* __restrict__ rhs, int count, int c0, int ncol, double2* __restrict__ B) {
    const size_t e = (size_t)blockIdx.x * kGT + threadIdx.x;
    if (e >= (size_t)count * kBatchRHS) return;
    const int j = (int)(e / kBatchRHS), r = (int)(e % kBatchRHS);
    B[e] = r < ncol ? rhs[(size_t)(c0 + r) * count + j] : make_double2(0.0, 0.0);
}

__global__ void k_gm_unpack(const double2* __restrict__ X, int count, int c0, int ncol, double2* __restrict__ out) {
    const size_t e = (size_t)blockIdx.x * kGT + threadIdx.x;
    if (e >= (size_t)count * kBatchRHS) return;
    const int j = (int)(e / kBatchRHS), r = (int)(e % kBatchRHS);
    if (r < ncol) out[(size_t)(c0 + r) * count + j] = X[e];
}

// ---------------------------------------------------------- block GMRES --
// Block products of 16-column blocks (point j, column a at j * 16 + a) on the
// FP64 matrix cores: one v_mfma_f64_16x16x4f64 k-step takes 4 points, and
// lane l loads element 64 s + l of both blocks -- A[a][k] = conj(V[4 s + k][a])
// and B[k][b] = W[4 s + k][b] sit at the same index -- so the D tile holds
// V^H W (Acc3: three real products per complex one).
constexpr int kBlockChunks = 256;  // point chunks per basis block in the block products
constexpr int kMaxBlocks = 64;    // blocks per block-GMRES cycle (then restart)

// part[(i kBlockChunks + c) 256 + 16 a + b] = sum over chunk c of conj(V_i[j][a]) W[j][b]:
// a wave takes chunk c of four basis blocks, so each element of W is loaded
// once for four products
constexpr int kBlocksPerWave = 4;
__global__ void __launch_bounds__(kGT) k_bg_dots(const double2* __restrict__ V, size_t vs, int count, int nv,
                                                 const double2* __restrict__ W, double2* __restrict__ part) {
    const int lane = threadIdx.x & 63;
    const int i0 = (blockIdx.y * (kGT / 64) + (threadIdx.x >> 6)) * kBlocksPerWave;
    if (i0 >= nv) return;  // a whole wave; no barrier below
    const int c = blockIdx.x;
    const int steps = (count + 3) / 4;
    const int per = (steps + kBlockChunks - 1) / kBlockChunks;
    const int s1 = min(steps, (c + 1) * per);
    const size_t n = (size_t)count * kBatchRHS;
    Acc3 acc[kBlocksPerWave];
    for (int s = c * per; s < s1; ++s) {
        const size_t e = (size_t)s * 64 + lane;
        const double2 b = e < n ? W[e] : make_double2(0.0, 0.0);
#pragma unroll
        for (int q = 0; q < kBlocksPerWave; ++q) {
            if (i0 + q >= nv) break;  // wave-uniform
            const double2 a = e < n ? V[(size_t)(i0 + q) * vs + e] : make_double2(0.0, 0.0);
            acc[q].add(a.x, -a.y, b.x, b.y);
        }
    }
#pragma unroll
    for (int q = 0; q < kBlocksPerWave; ++q) {
        if (i0 + q >= nv) break;
        double2* out = part + ((size_t)(i0 + q) * kBlockChunks + c) * 256;
#pragma unroll
        for (int v = 0; v < 4; ++v)
            out[(4 * v + lane / 16) * 16 + lane % 16] = make_double2(acc[q].re(v), acc[q].im(v));
    }
}

// out[i 256 + t] = sum_c part[(i kBlockChunks + c) 256 + t] in order
__global__ void k_bg_dots_sum(const double2* __restrict__ part, int nv, double2* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv * 256) return;
    const int i = t / 256, r = t % 256;
    double sr = 0.0, si = 0.0;
    for (int c = 0; c < kBlockChunks; ++c) {
        sr += part[((size_t)i * kBlockChunks + c) * 256 + r].x;
        si += part[((size_t)i * kBlockChunks + c) * 256 + r].y;
    }
    out[t] = make_double2(sr, si);
}

// W[j][b] = (accumulate ? W[j][b] : 0) + sign sum_{i < nv} sum_a V_i[j][a] M_i[a][b]
// (M_i: 16 x 16 blocks, row-major, M + 256 i) on the FP64 matrix cores: a wave
// owns 16 points, A[p][k] = V_i[p][a0 + k], B[k][b] = M_i[a0 + k][b], four
// k-steps per basis block
__global__ void __launch_bounds__(kGT) k_bg_gemm(const double2* __restrict__ V, size_t vs, int count,
                                                 const double2* __restrict__ M, int nv, double sign, bool accumulate,
                                                 double2* __restrict__ W) {
    const int lane = threadIdx.x & 63;
    const int e0 = (blockIdx.x * (kGT / 64) + (threadIdx.x >> 6)) * 16;
    if (e0 >= count) return;  // a whole wave; no barrier below
    const int pt = e0 + lane % 16;
    Acc3 acc;
    for (int i = 0; i < nv; ++i) {
        const double2* v = V + (size_t)i * vs;
        const double2* m = M + (size_t)i * 256;
#pragma unroll
        for (int a0 = 0; a0 < kBatchRHS; a0 += 4) {
            const double2 x = pt < count ? v[(size_t)pt * kBatchRHS + a0 + lane / 16] : make_double2(0.0, 0.0);
            const double2 y = m[(a0 + lane / 16) * 16 + lane % 16];
            acc.add(x.x, x.y, y.x, y.y);
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = e0 + 4 * q + lane / 16;
        if (r < count) {
            const size_t e = (size_t)r * kBatchRHS + lane % 16;
            const double2 o = accumulate ? W[e] : make_double2(0.0, 0.0);
            W[e] = make_double2(o.x + sign * acc.re(q), o.y + sign * acc.im(q));
        }
    }
}

// B[j][c] = v[j] for the points j of piece c -- [c count / 16, (c + 1) count / 16):
// tube points come in grid order, so a piece is one part of the surface --
// and 0 elsewhere: the columns sum to v.
__global__ void k_gm_split(const double2* __restrict__ v, int count, double2* __restrict__ B) {
    const size_t e = (size_t)blockIdx.x * kGT + threadIdx.x;
    if (e >= (size_t)count * kBatchRHS) return;
    const int j = (int)(e / kBatchRHS), c = (int)(e % kBatchRHS);
    const long long lo = (long long)c * count / kBatchRHS, hi = (long long)(c + 1) * count / kBatchRHS;
    B[e] = j >= lo && j < hi ? v[j] : make_double2(0.0, 0.0);
}

// x[j] += D[j][0]
__global__ void k_gm_add_col0(double2* __restrict__ x, const double2* __restrict__ D, int count) {
    const int j = blockIdx.x * kGT + threadIdx.x;
    if (j < count) x[j] = make_double2(x[j].x + D[(size_t)j * kBatchRHS].x, x[j].y + D[(size_t)j * kBatchRHS].y);
}

typedef std::complex<double> cplx;

// Device workspace of one solve. Vectors hold C interleaved columns of ld rows
// (vs = ld C elements; padding rows stay zero).
struct GmresSpace {
    double2 *V, *w, *d, *r;  // basis (max_basis + 1 vectors), A v, the correction, GMRES's residual
    double2 *res, *b, *x;    // refinement: residual, right-hand side (batched: packed), solution
    double2 *y, *coef;       // combination coefficients, Gram-Schmidt coefficients and norms
    double2* part;           // dot-product partials
    size_t vs;               // elements per vector
    int n;                   // elements in use (count C)
};

template <int C>
size_t space_bytes(const Context* ctx) {
    const int mb = max_basis<C>();
    const size_t vs = ctx->ld * C;
    return sizeof(double2) * (vs * ((size_t)mb + 1 + 6) + (size_t)C * (2 * (mb + 1) + 1) + (size_t)C * mb +
                              (size_t)C * (mb + 1) * dot_chunks<C>()) +
           512;
}

// The workspace at `base` (space_bytes<C>(ctx) bytes, 256-B aligned).
template <int C>
int space_at(Context* ctx, void* base, GmresSpace* g) {
    const int mb = max_basis<C>();
    g->vs = ctx->ld * C;
    g->n = ctx->count * C;
    g->V = static_cast<double2*>(base);
    g->w = g->V + ((size_t)mb + 1) * g->vs;
    g->d = g->w + g->vs;
    g->r = g->d + g->vs;
    g->res = g->r + g->vs;
    g->b = g->res + g->vs;
    g->x = g->b + g->vs;
    g->coef = g->x + g->vs;
    g->y = g->coef + (size_t)C * (2 * (mb + 1) + 1);
    g->part = g->y + (size_t)C * mb;
    // padding rows of every vector read as zero in the operator passes
    IBIM_HIP(hipMemsetAsync(g->V, 0, sizeof(double2) * g->vs * ((size_t)mb + 1 + 6), ctx->stream));
    return IBIM_OK;
}

template <int C>
int make_space(Context* ctx, GmresSpace* g) {
    if (ctx->refine.ensure(space_bytes<C>(ctx))) return IBIM_ERR_NOMEM;
    return space_at<C>(ctx, ctx->refine.ptr, g);
}

// out[i C + col] = V_i^H w per column, i < nv (device)
template <int C>
void dots(Context* ctx, const GmresSpace& g, const double2* V, int nv, const double2* w, double2* out) {
    hipLaunchKernelGGL(k_gm_dots<C>, dim3(dot_chunks<C>(), (unsigned)((nv + 3) / 4)), dim3(kGT), 0, ctx->stream, V,
                       g.vs, g.n, nv, w, g.part);
    hipLaunchKernelGGL(k_gm_dots_sum<C>, dim3((unsigned)((nv * C + 63) / 64)), dim3(64), 0, ctx->stream, g.part, nv,
                       out);
}

// ||v|| per column (host values; synchronizes)
template <int C>
int norms_of(Context* ctx, const GmresSpace& g, const double2* v, double* out) {
    dots<C>(ctx, g, v, 1, v, g.coef);
    double2 h[C];
    IBIM_HIP(hipMemcpyAsync(h, g.coef, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    for (int c = 0; c < C; ++c) out[c] = sqrt(h[c].x);
    return IBIM_OK;
}

// The operator on C-column vectors: the single-vector Y-only sweep, or the
// 16-column pass (Y = A P alone).
template <int C>
int apply_op(Context* ctx, int op, const double2* p, double2* y) {
    if constexpr (C == 1)
        return matvec_op_y(ctx, op, p, y);
    else
        return batch_apply(ctx, op, false, p, y, nullptr);
}

// GMRES(m), m <= max_basis<C>(), for A d = rhs per column, from d = 0, until
// each column's residual estimate is below its target (absolute; +inf leaves a
// column out) or `budget` products are used. Resumable, so that several runs
// can share one operator pass: the caller applies the operator to in() into
// out(), then calls launch_post() (device work and the copy of its results
// to the host), synchronizes the stream, and calls host_post(); until done().
template <int C>
struct GmresRun {
    enum Phase { kIterate, kRestart, kDone };
    static constexpr int MB = max_basis<C>();
    Context* ctx = nullptr;
    GmresSpace g{};
    const double2* rhs = nullptr;
    double target[C] = {};
    int budget = 0, m = 0;
    int passes = 0;               // operator applications
    int* col_products = nullptr;  // += the products each column took part in
    double estimate[C] = {};
    double* hist = nullptr;  // column 0's estimate / normb per product
    int* hist_at = nullptr;
    int hist_len = 0;
    double normb = 1.0;
    Phase phase = kDone;
    int j = 0;
    bool active[C] = {};
    int ncols[C] = {};
    // per column: Hessenberg (column-major, MB + 1 rows), rotations, rotated residual
    std::vector<cplx> H, cs, sn, gv;
    std::vector<double2> hh, yd;  // host copies of the coefficients; the combination

    static size_t space_bytes(const Context* ctx) { return ibim::space_bytes<C>(ctx); }
    static int space_at(Context* ctx, void* base, GmresSpace* s) { return ibim::space_at<C>(ctx, base, s); }

    cplx& Hc(int c, int i, int k) { return H[((size_t)c * MB + k) * (MB + 1) + i]; }
    bool done() const { return phase == kDone; }
    const double2* in() const { return phase == kIterate ? g.V + (size_t)j * g.vs : g.d; }
    double2* out() const { return g.w; }
    unsigned blocks() const { return (unsigned)((g.n + kGT - 1) / kGT); }

    // d = 0, r = rhs, the first cycle's norms (synchronizes)
    int start() {
        H.assign((size_t)C * (MB + 1) * MB, cplx(0.0));
        cs.assign((size_t)C * MB, cplx(0.0));
        sn.assign((size_t)C * MB, cplx(0.0));
        gv.assign((size_t)C * (MB + 1), cplx(0.0));
        hh.assign((size_t)C * (2 * (MB + 1) + 1), make_double2(0.0, 0.0));
        passes = 0;
        for (int c = 0; c < C; ++c) active[c] = false;
        IBIM_HIP(hipMemsetAsync(g.d, 0, sizeof(double2) * g.vs, ctx->stream));
        IBIM_HIP(hipMemcpyAsync(g.r, rhs, sizeof(double2) * g.n, hipMemcpyDeviceToDevice, ctx->stream));
        dots<C>(ctx, g, g.r, 1, g.r, g.coef);
        IBIM_HIP(hipMemcpyAsync(hh.data(), g.coef, sizeof(double2) * C, hipMemcpyDeviceToHost, ctx->stream));
        IBIM_HIP(hipStreamSynchronize(ctx->stream));
        return cycle_begin();
    }

    // after out() = A in(): the device side of the step and its host copy
    int launch_post() {
        ++passes;
        for (int c = 0; c < C; ++c) col_products[c] += active[c];
        if (phase == kRestart) {
            // r0 = rhs - A d
            hipLaunchKernelGGL(k_gm_sub, dim3(blocks()), dim3(kGT), 0, ctx->stream, rhs, g.w, g.n, g.r);
            dots<C>(ctx, g, g.r, 1, g.r, g.coef);
            IBIM_HIP(hipGetLastError());
            IBIM_HIP(hipMemcpyAsync(hh.data(), g.coef, sizeof(double2) * C, hipMemcpyDeviceToHost, ctx->stream));
            return IBIM_OK;
        }
        // CGS2: h = V^H w, w -= V h, twice; then ||w||
        for (int pass = 0; pass < 2; ++pass) {
            double2* coef = g.coef + (size_t)pass * (MB + 1) * C;
            dots<C>(ctx, g, g.V, j + 1, g.w, coef);
            hipLaunchKernelGGL(k_gm_combine<C>, dim3(blocks()), dim3(kGT), 0, ctx->stream, g.V, g.vs, g.n, coef,
                               j + 1, -1.0, g.w);
        }
        dots<C>(ctx, g, g.w, 1, g.w, g.coef + 2 * (size_t)(MB + 1) * C);
        IBIM_HIP(hipGetLastError());
        IBIM_HIP(hipMemcpyAsync(hh.data(), g.coef, sizeof(double2) * hh.size(), hipMemcpyDeviceToHost, ctx->stream));
        return IBIM_OK;
    }

    // after the stream has caught up with launch_post
    int host_post() {
        if (phase == kRestart) return cycle_begin();
        Scales<C> sc;
        bool any = false;
        for (int c = 0; c < C; ++c) {
            if (!active[c]) {
                sc.s[c] = 0.0;
                continue;
            }
            cplx* gc = gv.data() + (size_t)c * (MB + 1);
            cplx* csc = cs.data() + (size_t)c * MB;
            cplx* snc = sn.data() + (size_t)c * MB;
            for (int i = 0; i <= j; ++i) {
                const double2 a = hh[(size_t)i * C + c], b = hh[((size_t)(MB + 1) + i) * C + c];
                Hc(c, i, j) = cplx(a.x + b.x, a.y + b.y);
            }
            const double hn = sqrt(hh[2 * (size_t)(MB + 1) * C + c].x);
            Hc(c, j + 1, j) = hn;
            // previous rotations, then the one that zeroes H(j + 1, j)
            for (int i = 0; i < j; ++i) {
                const cplx t = csc[i] * Hc(c, i, j) + snc[i] * Hc(c, i + 1, j);
                Hc(c, i + 1, j) = -std::conj(snc[i]) * Hc(c, i, j) + std::conj(csc[i]) * Hc(c, i + 1, j);
                Hc(c, i, j) = t;
            }
            const cplx a = Hc(c, j, j), bb = Hc(c, j + 1, j);
            const double rr = sqrt(std::norm(a) + std::norm(bb));
            csc[j] = rr > 0.0 ? std::conj(a) / rr : cplx(1.0);
            snc[j] = rr > 0.0 ? std::conj(bb) / rr : cplx(0.0);
            Hc(c, j, j) = rr;
            Hc(c, j + 1, j) = 0.0;
            gc[j + 1] = -std::conj(snc[j]) * gc[j];
            gc[j] = csc[j] * gc[j];
            estimate[c] = std::abs(gc[j + 1]);
            if (c == 0 && hist && *hist_at < hist_len) hist[(*hist_at)++] = estimate[c] / normb;
            if (!(estimate[c] > target[c]) || !(hn > 0.0) || estimate[c] != estimate[c]) {
                active[c] = false;  // done with j + 1 basis vectors
                ncols[c] = j + 1;
                sc.s[c] = 0.0;
            } else {
                sc.s[c] = 1.0 / hn;
                any = true;
            }
        }
        ++j;
        if (!any || j >= m || passes >= budget) return cycle_end();
        hipLaunchKernelGGL(k_gm_scale<C>, dim3(blocks()), dim3(kGT), 0, ctx->stream, g.w, sc, g.n,
                           g.V + (size_t)j * g.vs);
        IBIM_HIP(hipGetLastError());
        return IBIM_OK;
    }

    // a cycle from the residual r (its norms in hh): v0 = r / ||r|| per column
    int cycle_begin() {
        Scales<C> sc;
        bool any = false;
        for (int c = 0; c < C; ++c) {
            const double beta = sqrt(hh[c].x);
            estimate[c] = beta;
            active[c] = beta > target[c];
            sc.s[c] = active[c] ? 1.0 / beta : 0.0;
            any = any || active[c];
            ncols[c] = 0;
            std::fill(gv.begin() + (size_t)c * (MB + 1), gv.begin() + (size_t)(c + 1) * (MB + 1), cplx(0.0));
            gv[(size_t)c * (MB + 1)] = beta;
        }
        j = 0;
        if (!any) {
            phase = kDone;
            return IBIM_OK;
        }
        hipLaunchKernelGGL(k_gm_scale<C>, dim3(blocks()), dim3(kGT), 0, ctx->stream, g.r, sc, g.n, g.V);
        IBIM_HIP(hipGetLastError());
        phase = kIterate;
        return IBIM_OK;
    }

    // y = R^{-1} g over each column's basis vectors, d += V y; then a restart
    // for the columns that ran out of basis vectors, budget permitting
    int cycle_end() {
        bool again = false;
        int nmax = 0;
        for (int c = 0; c < C; ++c) {
            if (active[c]) {
                ncols[c] = j;
                again = true;
            }
            nmax = std::max(nmax, ncols[c]);
        }
        yd.assign((size_t)nmax * C, make_double2(0.0, 0.0));
        for (int c = 0; c < C; ++c) {
            const int nc = ncols[c];
            const cplx* gc = gv.data() + (size_t)c * (MB + 1);
            std::vector<cplx> yv(nc);
            for (int i = nc - 1; i >= 0; --i) {
                cplx s = gc[i];
                for (int k = i + 1; k < nc; ++k) s -= Hc(c, i, k) * yv[k];
                yv[i] = Hc(c, i, i) != cplx(0.0) ? s / Hc(c, i, i) : cplx(0.0);
            }
            for (int i = 0; i < nc; ++i) yd[(size_t)i * C + c] = make_double2(yv[i].real(), yv[i].imag());
        }
        if (nmax > 0) {
            IBIM_HIP(hipMemcpyAsync(g.y, yd.data(), sizeof(double2) * yd.size(), hipMemcpyHostToDevice, ctx->stream));
            hipLaunchKernelGGL(k_gm_combine<C>, dim3(blocks()), dim3(kGT), 0, ctx->stream, g.V, g.vs, g.n, g.y, nmax,
                               1.0, g.d);
            IBIM_HIP(hipGetLastError());
        }
        phase = again && passes < budget ? kRestart : kDone;
        if (phase == kDone) IBIM_HIP(hipStreamSynchronize(ctx->stream));  // yd stays untouched until d is formed
        return IBIM_OK;
    }
};

// Block GMRES (Saad, Iterative Methods for Sparse Linear Systems, 2003, 6.12)
// over one batch of 16 columns: ONE Krylov space for all of them, built from
// the residual block R0 = V0 B0. Iteration j: W = A V_j (one 16-column pass),
// block classical Gram-Schmidt applied twice against V_0 .. V_j (16 x 16
// coefficient blocks on the matrix cores), then W = V_{j+1} R by Cholesky QR
// applied twice (Fukaya et al., 2014; the 16 x 16 Cholesky on the host); the
// block Hessenberg matrix is reduced by Givens rotations as it grows, and the
// rotated B0 gives every column's residual. Each column of a batch searches
// the whole block space (16 (j + 1) dimensions after j + 1 passes) where
// per-column GMRES searches its own (j + 1): on problems whose columns share
// slow modes -- near an interior resonance, where the double layer's
// per-column GMRES needs ~140 products -- the block space finds them in a
// fraction of the passes. Columns at their target before the cycle starts
// stay out of it (zero basis columns); a Cholesky pivot below 1e-24 of the
// block's largest (a basis block that has lost a direction) ends the cycle
// and the next one restarts from the true residuals of the columns left.
// The same resumable interface as GmresRun.
struct BlockRun {
    enum Phase { kIterate, kRestart, kDone };
    static constexpr int P = kBatchRHS;
    static constexpr int MB = kMaxBlocks;
    Context* ctx = nullptr;
    GmresSpace g{};
    const double2* rhs = nullptr;
    double target[P] = {};
    int budget = 0, m = 0;
    int passes = 0;
    int* col_products = nullptr;
    double estimate[P] = {};
    Phase phase = kDone;
    // enlarged mode (one right-hand side split over the columns): the block
    // space is built as above, but ONE residual -- the sum of the columns --
    // is minimized over all of it; column 0 carries its estimate, target and
    // correction, and a cycle ends the run (the refinement restarts it)
    bool single = false;
    bool adjoint = false;  // the passes apply A^H (the adjoint solve of a gradient)
    int j = 0;            // basis blocks of the cycle with their product applied
    int p = 0;            // columns in the cycle
    int idx[P] = {};      // their slots
    bool active[P] = {};  // slot in the cycle
    // block Hessenberg ((MB + 1) p x MB p, column-major), rotated right-hand side
    // ((MB + 1) p x p), Givens rotations of rows (r, r + 1) in order
    std::vector<cplx> H, E;
    struct Rot {
        int r;
        cplx c, s;
    };
    std::vector<Rot> rots;
    std::vector<double2> hh;      // device results on the host
    std::vector<double2> upm, upy;  // uploads: a 16 x 16 inverse, the combination blocks (kept until synchronized)
    double2* mat = nullptr;       // device 16 x 16 upload slot
    double2* coefs = nullptr;     // device: 2 (MB + 1) coefficient blocks, then the Gram block
    int hrows = 0;

    static size_t space_bytes(const Context* ctx) {
        const size_t vs = ctx->ld * P;
        const size_t part = std::max((size_t)(MB + 1) * kBlockChunks * 256, (size_t)dot_chunks<P>() * P);
        return sizeof(double2) * (vs * ((size_t)MB + 1 + 6) + (2 * (size_t)(MB + 1) + 1) * 256 +
                                  ((size_t)MB + 1) * 256 + part) +
               512;
    }
    static int space_at(Context* ctx, void* base, GmresSpace* s) {
        s->vs = ctx->ld * P;
        s->n = ctx->count * P;
        s->V = static_cast<double2*>(base);
        s->w = s->V + ((size_t)MB + 1) * s->vs;
        s->d = s->w + s->vs;
        s->r = s->d + s->vs;
        s->res = s->r + s->vs;
        s->b = s->res + s->vs;
        s->x = s->b + s->vs;
        s->coef = s->x + s->vs;
        s->y = s->coef + (2 * (size_t)(MB + 1) + 1) * 256;
        s->part = s->y + ((size_t)MB + 1) * 256;
        IBIM_HIP(hipMemsetAsync(s->V, 0, sizeof(double2) * s->vs * ((size_t)MB + 1 + 6), ctx->stream));
        return IBIM_OK;
    }

    cplx& Hc(int r, int c) { return H[(size_t)c * hrows + r]; }
    cplx& Ec(int r, int c) { return E[(size_t)c * hrows + r]; }
    bool done() const { return phase == kDone; }
    const double2* in() const { return phase == kIterate ? g.V + (size_t)j * g.vs : g.d; }
    double2* out() const { return g.w; }
    unsigned blocks() const { return (unsigned)((g.n + kGT - 1) / kGT); }

    // out[i] = V_i^H W for i < nv (16 x 16 blocks, device)
    void block_dots(const double2* V, int nv, const double2* W, double2* outp) {
        const int per_block = (kGT / 64) * kBlocksPerWave;
        hipLaunchKernelGGL(k_bg_dots, dim3(kBlockChunks, (unsigned)((nv + per_block - 1) / per_block)), dim3(kGT), 0,
                           ctx->stream, V, g.vs, ctx->count, nv, W, g.part);
        hipLaunchKernelGGL(k_bg_dots_sum, dim3((unsigned)((nv * 256 + 255) / 256)), dim3(256), 0, ctx->stream, g.part,
                           nv, outp);
    }
    // W = (accumulate ? W : 0) + sign sum_i V_i M_i
    void block_gemm(const double2* V, const double2* M, int nv, double sign, bool accumulate, double2* W) {
        const unsigned tiles = (unsigned)((ctx->count + 16 * (kGT / 64) - 1) / (16 * (kGT / 64)));
        hipLaunchKernelGGL(k_bg_gemm, dim3(tiles), dim3(kGT), 0, ctx->stream, V, g.vs, ctx->count, M, nv, sign,
                           accumulate, W);
    }
    // the Gram block of W (device) -> host (synchronizes)
    int gram(const double2* W, cplx* G) {
        double2* gd = coefs + 2 * (size_t)(MB + 1) * 256;
        block_dots(W, 1, W, gd);
        IBIM_HIP(hipGetLastError());
        hh.resize(256);
        IBIM_HIP(hipMemcpyAsync(hh.data(), gd, sizeof(double2) * 256, hipMemcpyDeviceToHost, ctx->stream));
        IBIM_HIP(hipStreamSynchronize(ctx->stream));
        for (int t = 0; t < 256; ++t) G[t] = cplx(hh[t].x, hh[t].y);
        return IBIM_OK;
    }
    // R (p x p upper, row-major) with G[idx][idx] = R^H R; false on a pivot
    // below 1e-24 of the largest diagonal entry
    bool cholesky(const cplx* G, cplx* R) const {
        double dmax = 0.0;
        for (int a = 0; a < p; ++a) dmax = std::max(dmax, G[idx[a] * 16 + idx[a]].real());
        for (int t = 0; t < p * p; ++t) R[t] = 0.0;
        for (int k = 0; k < p; ++k) {
            double d = G[idx[k] * 16 + idx[k]].real();
            for (int q = 0; q < k; ++q) d -= std::norm(R[q * p + k]);
            if (!(d > 1e-24 * dmax)) return false;
            const double rk = sqrt(d);
            R[k * p + k] = rk;
            for (int c = k + 1; c < p; ++c) {
                cplx v = G[idx[k] * 16 + idx[c]];
                for (int q = 0; q < k; ++q) v -= std::conj(R[q * p + k]) * R[q * p + c];
                R[k * p + c] = v / rk;
            }
        }
        return true;
    }
    // upload the inverse of R (p x p upper) as a 16 x 16 block over the slots
    int upload_inverse(const cplx* R) {
        std::vector<cplx> Ri((size_t)p * p, cplx(0.0));
        for (int c = 0; c < p; ++c) {  // column c of R^{-1}: R x = e_c
            for (int r = c; r >= 0; --r) {
                cplx v = r == c ? cplx(1.0) : cplx(0.0);
                for (int q = r + 1; q <= c; ++q) v -= R[r * p + q] * Ri[q * p + c];
                Ri[r * p + c] = v / R[r * p + r];
            }
        }
        // the previous upload from upm has completed: every call follows a gram() synchronization
        upm.assign(256, make_double2(0.0, 0.0));
        for (int a = 0; a < p; ++a)
            for (int b = 0; b < p; ++b)
                upm[idx[a] * 16 + idx[b]] = make_double2(Ri[a * p + b].real(), Ri[a * p + b].imag());
        IBIM_HIP(hipMemcpyAsync(mat, upm.data(), sizeof(double2) * 256, hipMemcpyHostToDevice, ctx->stream));
        return IBIM_OK;
    }
    // dst = src R^{-1} with R^H R = the Gram block of src, twice (Cholesky QR 2);
    // Rout = R2 R1. False when a pivot fails. tmp: a free vector.
    int orthonormalize(const double2* src, double2* dst, double2* tmp, cplx* Rout, bool* ok) {
        std::vector<cplx> G(256), R1((size_t)p * p), R2((size_t)p * p);
        if (int rc = gram(src, G.data())) return rc;
        *ok = cholesky(G.data(), R1.data());
        if (!*ok) return IBIM_OK;
        if (int rc = upload_inverse(R1.data())) return rc;
        block_gemm(src, mat, 1, 1.0, false, tmp);
        if (int rc = gram(tmp, G.data())) return rc;  // also: the upload is done
        *ok = cholesky(G.data(), R2.data());
        if (!*ok) return IBIM_OK;
        if (int rc = upload_inverse(R2.data())) return rc;
        block_gemm(tmp, mat, 1, 1.0, false, dst);
        IBIM_HIP(hipGetLastError());
        for (int a = 0; a < p; ++a)
            for (int b = 0; b < p; ++b) {
                cplx v = 0.0;
                for (int q = 0; q < p; ++q) v += R2[a * p + q] * R1[q * p + b];
                Rout[a * p + b] = v;
            }
        return IBIM_OK;
    }

    int start() {
        passes = 0;
        mat = g.y + (size_t)MB * 256;
        coefs = g.coef;
        IBIM_HIP(hipMemsetAsync(g.d, 0, sizeof(double2) * g.vs, ctx->stream));
        IBIM_HIP(hipMemcpyAsync(g.r, rhs, sizeof(double2) * g.n, hipMemcpyDeviceToDevice, ctx->stream));
        return cycle_begin();
    }

    // a cycle from the residual block r: the columns above their targets, V_0 R0 = r
    int cycle_begin() {
        std::vector<cplx> G(256);
        if (int rc = gram(g.r, G.data())) return rc;
        p = 0;
        double total = 0.0;
        for (int c = 0; c < P; ++c) {
            estimate[c] = sqrt(std::max(G[c * 16 + c].real(), 0.0));
            // enlarged mode: every piece of the one residual above 1e-12 of the largest
            active[c] = single ? estimate[c] > 0.0 : estimate[c] > target[c];
        }
        if (single) {
            double big = 0.0;
            for (int c = 0; c < P; ++c) big = std::max(big, estimate[c]);
            for (int c = 0; c < P; ++c) active[c] = estimate[c] > 1e-12 * big;
        }
        for (int c = 0; c < P; ++c)
            if (active[c]) idx[p++] = c;
        if (single) {  // the pieces have disjoint supports: their norms add
            for (int c = 0; c < P; ++c) total += estimate[c] * estimate[c];
            for (int c = 1; c < P; ++c) estimate[c] = 0.0;
            estimate[0] = sqrt(total);
        }
        j = 0;
        if (p == 0 || (single && !(estimate[0] > target[0]))) {
            phase = kDone;
            return IBIM_OK;
        }
        std::vector<cplx> R0((size_t)p * p);
        bool ok = false;
        // V_0 in the slot of basis block 0; g.w is free until the first product
        if (int rc = orthonormalize(g.r, g.V, g.w, R0.data(), &ok)) return rc;
        if (!ok && single) {  // disjoint pieces: only a piece below 1e-12 of the largest can fail
            phase = kDone;
            return IBIM_OK;
        }
        if (!ok) {
            // dependent residual columns: the cycle takes the largest one alone
            int best = idx[0];
            for (int a = 0; a < p; ++a)
                if (estimate[idx[a]] > estimate[best]) best = idx[a];
            for (int c = 0; c < P; ++c) active[c] = c == best;
            p = 1;
            idx[0] = best;
            if (int rc = orthonormalize(g.r, g.V, g.w, R0.data(), &ok)) return rc;
            if (!ok) {
                phase = kDone;
                return IBIM_OK;
            }
        }
        hrows = (MB + 1) * p;
        H.assign((size_t)hrows * MB * p, cplx(0.0));
        E.assign((size_t)hrows * ecols(), cplx(0.0));
        rots.clear();
        for (int a = 0; a < p; ++a)
            for (int b = a; b < p; ++b) {
                if (single)
                    Ec(a, 0) += R0[a * p + b];  // r = V_0 R0 (1, ..., 1)^T
                else
                    Ec(a, b) = R0[a * p + b];
            }
        phase = kIterate;
        return IBIM_OK;
    }

    int launch_post() {
        ++passes;
        if (single)
            ++col_products[0];
        else
            for (int c = 0; c < P; ++c) col_products[c] += active[c];
        if (phase == kRestart) {
            hipLaunchKernelGGL(k_gm_sub, dim3(blocks()), dim3(kGT), 0, ctx->stream, rhs, g.w, g.n, g.r);
            IBIM_HIP(hipGetLastError());
            return IBIM_OK;
        }
        // block CGS2: C = V^H W, W -= V C, twice
        for (int pass = 0; pass < 2; ++pass) {
            double2* c = coefs + (size_t)pass * (MB + 1) * 256;
            block_dots(g.V, j + 1, g.w, c);
            block_gemm(g.V, c, j + 1, -1.0, true, g.w);
        }
        IBIM_HIP(hipGetLastError());
        hh.resize(2 * (size_t)(MB + 1) * 256);
        for (int pass = 0; pass < 2; ++pass)
            IBIM_HIP(hipMemcpyAsync(hh.data() + (size_t)pass * (MB + 1) * 256, coefs + (size_t)pass * (MB + 1) * 256,
                                    sizeof(double2) * 256 * (j + 1), hipMemcpyDeviceToHost, ctx->stream));
        return IBIM_OK;
    }

    // rotation zeroing H(r + 1, col) against H(r, col), applied from column col on
    void rotate(int r, int col, int last) {
        const cplx a = Hc(r, col), bb = Hc(r + 1, col);
        const double rr = sqrt(std::norm(a) + std::norm(bb));
        const cplx c = rr > 0.0 ? std::conj(a) / rr : cplx(1.0), sn = rr > 0.0 ? std::conj(bb) / rr : cplx(0.0);
        rots.push_back({r, c, sn});
        for (int q = col; q <= last; ++q) {
            const cplx x = Hc(r, q), y = Hc(r + 1, q);
            Hc(r, q) = c * x + sn * y;
            Hc(r + 1, q) = -std::conj(sn) * x + std::conj(c) * y;
        }
        for (int q = 0; q < ecols(); ++q) {
            const cplx x = Ec(r, q), y = Ec(r + 1, q);
            Ec(r, q) = c * x + sn * y;
            Ec(r + 1, q) = -std::conj(sn) * x + std::conj(c) * y;
        }
    }
    int ecols() const { return single ? 1 : p; }

    int host_post() {
        if (phase == kRestart) return cycle_begin();
        // block column j of H: the coefficient blocks, then W = V_{j+1} R
        const int c0 = j * p;
        for (int i = 0; i <= j; ++i)
            for (int a = 0; a < p; ++a)
                for (int b = 0; b < p; ++b) {
                    const double2 x = hh[(size_t)i * 256 + idx[a] * 16 + idx[b]];
                    const double2 y = hh[((size_t)(MB + 1) + i) * 256 + idx[a] * 16 + idx[b]];
                    Hc(i * p + a, c0 + b) = cplx(x.x + y.x, x.y + y.y);
                }
        std::vector<cplx> R((size_t)p * p, cplx(0.0));
        bool ok = false;
        // V_{j+1} in basis slot j + 1 (g.r is free between restarts)
        if (j + 1 <= MB)
            if (int rc = orthonormalize(g.w, g.V + (size_t)(j + 1) * g.vs, g.r, R.data(), &ok)) return rc;
        for (int a = 0; a < p; ++a)
            for (int b = a; b < p; ++b) Hc((j + 1) * p + a, c0 + b) = ok ? R[a * p + b] : cplx(0.0);
        // earlier rotations, then the new subdiagonal block eliminated column by column
        for (const Rot& t : rots)
            for (int q = c0; q < c0 + p; ++q) {
                const cplx x = Hc(t.r, q), y = Hc(t.r + 1, q);
                Hc(t.r, q) = t.c * x + t.s * y;
                Hc(t.r + 1, q) = -std::conj(t.s) * x + std::conj(t.c) * y;
            }
        for (int b = 0; b < p; ++b) {
            const int col = c0 + b;
            for (int r = (j + 1) * p + b - 1; r >= col; --r) rotate(r, col, c0 + p - 1);
        }
        ++j;
        bool any = false;
        for (int q = 0; q < ecols(); ++q) {
            const int c = single ? 0 : idx[q];
            double s2 = 0.0;
            for (int r = j * p; r < (j + 1) * p; ++r) s2 += std::norm(Ec(r, q));
            estimate[c] = sqrt(s2);
            any = any || estimate[c] > target[c] || estimate[c] != estimate[c];
        }
        if (!ok || !any || j >= m || passes >= budget) return cycle_end(any && !single);
        return IBIM_OK;
    }

    // Y = R^{-1} E over the j block columns, d += sum_i V_i Y_i; then a restart
    // for the columns still above their targets, budget permitting
    int cycle_end(bool again) {
        const int nc = j * p;
        std::vector<cplx> Y((size_t)nc * p, cplx(0.0));
        for (int q = 0; q < ecols(); ++q)
            for (int r = nc - 1; r >= 0; --r) {
                cplx v = Ec(r, q);
                for (int t = r + 1; t < nc; ++t) v -= Hc(r, t) * Y[(size_t)q * nc + t];
                Y[(size_t)q * nc + r] = Hc(r, r) != cplx(0.0) ? v / Hc(r, r) : cplx(0.0);
            }
        upy.assign((size_t)j * 256, make_double2(0.0, 0.0));
        for (int i = 0; i < j; ++i)
            for (int a = 0; a < p; ++a)
                for (int q = 0; q < ecols(); ++q) {
                    const cplx v = Y[(size_t)q * nc + i * p + a];
                    // enlarged mode: the correction is column 0 of d
                    upy[(size_t)i * 256 + idx[a] * 16 + (single ? 0 : idx[q])] = make_double2(v.real(), v.imag());
                }
        IBIM_HIP(hipMemcpyAsync(g.y, upy.data(), sizeof(double2) * upy.size(), hipMemcpyHostToDevice, ctx->stream));
        block_gemm(g.V, g.y, j, 1.0, true, g.d);
        IBIM_HIP(hipGetLastError());
        IBIM_HIP(hipStreamSynchronize(ctx->stream));  // upy stays untouched until d is formed
        phase = again && passes < budget ? kRestart : kDone;
        return IBIM_OK;
    }
};

// One GMRES solve on one context (see GmresRun). Adds each column's products
// to col_products; returns the passes applied and the last estimates.
template <int C>
int gmres(Context* ctx, const GmresSpace& g, int op, const double2* rhs, const double* target, int budget, int m,
          int* passes, int* col_products, double* estimate, double* hist, int* hist_at, int hist_len, double normb) {
    GmresRun<C> run;
    run.ctx = ctx;
    run.g = g;
    run.rhs = rhs;
    for (int c = 0; c < C; ++c) run.target[c] = target[c];
    run.budget = budget;
    run.m = m;
    run.col_products = col_products;
    run.hist = hist;
    run.hist_at = hist_at;
    run.hist_len = hist_len;
    run.normb = normb;
    int rc = run.start();
    while (!rc && !run.done()) {
        rc = apply_op<C>(ctx, op, run.in(), run.out());
        if (!rc) rc = run.launch_post();
        if (!rc) IBIM_HIP(hipStreamSynchronize(ctx->stream));
        if (!rc) rc = run.host_post();
    }
    *passes = run.passes;
    for (int c = 0; c < C; ++c) estimate[c] = run.estimate[c];
    return rc;
}

// The FP64 refinement residual of a batch: one 16-column pass per refinement
// step, so its workgroup shape is not timed.
int residual_pass(Context* ctx, int op, const double2* x, double2* y) {
    const bool old = ctx->batch_untimed;
    ctx->batch_untimed = true;
    const int rc = batch_apply(ctx, op, false, x, y, nullptr);
    ctx->batch_untimed = old;
    return rc;
}

// Runs a resumable solve (GmresRun, BlockRun) to its end on one context.
template <class Run>
int solve_run(Run& run, int op) {
    int rc = run.start();
    while (!rc && !run.done()) {
        if constexpr (std::is_same<Run, BlockRun>::value)
            rc = batch_apply(run.ctx, op, run.adjoint, run.in(), run.out(), nullptr);
        else
            rc = apply_op<kBatchRHS>(run.ctx, op, run.in(), run.out());
        if (!rc) rc = run.launch_post();
        if (!rc) IBIM_HIP(hipStreamSynchronize(run.ctx->stream));
        if (!rc) rc = run.host_post();
    }
    return rc;
}

// Basis length of a batched run: opts->chunk when set, else the largest.
template <class Run>
int basis_of(const ibim_solve_opts* opts);
template <>
int basis_of<GmresRun<kBatchRHS>>(const ibim_solve_opts* opts) {
    return opts->chunk > 0 ? std::min(opts->chunk, kMaxBasisB) : kMaxBasisB;
}
template <>
int basis_of<BlockRun>(const ibim_solve_opts* opts) {
    return opts->chunk > 0 ? std::min(opts->chunk, kMaxBlocks) : kMaxBlocks;
}

// The operators of a refinement: GMRES on the stored one, residuals with the
// FP64 one (complex128 storage: itself); with nothing stored, GMRES on the
// mixed matrix-free operator and residuals with the FP64 matrix-free one.
void refinement_ops(const Context* ctx, int* inner, int* outer) {
    if (ctx->mat_precision < 0) {
        *inner = OP_FREE32;
        *outer = OP_FREE64;
        return;
    }
    *inner = OP_DENSE;
    *outer = ctx->mat_precision == IBIM_C128 ? OP_DENSE : fp64_operator(ctx);
}

// Each GMRES run of a refinement aims at kInnerMargin tol ||b|| when GMRES runs
// on a stored operator and the FP64 residual has to be computed (the FP64
// matrix-free operator): the FP64 residual of a correction computed on the
// lower-precision operator lands somewhat above GMRES's own estimate, and one
// FP64 pass then costs several GMRES products (star surface on 256^3, 16 plane
// waves, FP32 products: 357 against 127 ms; true residual 1.09e-7 after a run
// to tol = 1e-7, so a second run and a second FP64 pass, 5.4 s in all; 6.6e-8
// after a run to 0.5 tol, 5.2 s). Where the FP64 pass is cheap (stored
// complex128 copy, or both operators computed) the second step costs less than
// the margin's extra products (torus 128^3, matrix-free: 0.785 against
// 0.754 s), so GMRES aims at tol itself.
constexpr double kInnerMargin = 0.5;
double inner_margin(int inner, int outer) { return inner == OP_DENSE && outer == OP_FREE64 ? kInnerMargin : 1.0; }

// One right-hand side at a time.
int gmres_columns(Context* ctx, const ibim_solve_opts* opts, ibim_solve_info* info, double* hist) {
    int inner, outer_op;
    refinement_ops(ctx, &inner, &outer_op);
    const int count = ctx->count;
    const unsigned nb = (unsigned)((count + kGT - 1) / kGT);
    GmresSpace g;
    if (int rc = make_space<1>(ctx, &g)) return rc;
    const int max_outer = opts->max_outer > 0 ? opts->max_outer : 20;
    // the basis length: opts->chunk when set (a restart every chunk products), else kMaxBasis
    const int m = opts->chunk > 0 ? std::min(opts->chunk, kMaxBasis) : kMaxBasis;
    int total = 0, outer_max = 0, hist_at = 0, status = 4;  // 4 until a column has data
    double worst = 0.0;
    for (int c = 0; c < ctx->nrhs; ++c) {
        const double2* b = ctx->rhs.as<double2>() + (size_t)c * count;
        double2* x = ctx->density.as<double2>() + (size_t)c * count;
        IBIM_HIP(hipMemsetAsync(x, 0, sizeof(double2) * count, ctx->stream));
        IBIM_HIP(hipMemcpyAsync(g.res, b, sizeof(double2) * count, hipMemcpyDeviceToDevice, ctx->stream));
        double normb;
        if (int rc = norms_of<1>(ctx, g, g.res, &normb)) return rc;
        if (normb == 0.0) continue;  // zero data: zero density, no iterations
        int col_status = 1, used = 0, outer = 0;
        double relres = 1.0;
        for (outer = 1; outer <= max_outer && used < opts->maxit; ++outer) {
            int products = 0, passes = 0;
            double est = 0.0;
            // the correction only has to bring the residual to tol ||b||
            const double target = inner_margin(inner, outer_op) * opts->tol * normb;
            if (int rc = gmres<1>(ctx, g, inner, g.res, &target, opts->maxit - used, m, &passes, &products, &est,
                                  c == 0 ? hist : nullptr, &hist_at, opts->maxit, normb))
                return rc;
            used += passes;
            hipLaunchKernelGGL(k_gm_add, dim3(nb), dim3(kGT), 0, ctx->stream, x, g.d, count);
            if (int rc = matvec_op_y(ctx, outer_op, x, g.w)) return rc;
            hipLaunchKernelGGL(k_gm_sub, dim3(nb), dim3(kGT), 0, ctx->stream, b, g.w, count, g.res);
            IBIM_HIP(hipGetLastError());
            double rn;
            if (int rc = norms_of<1>(ctx, g, g.res, &rn)) return rc;
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

// 16 right-hand sides per batch: the same refinement per column, every
// product of a batch one 16-column pass (Run: GmresRun<16>, a Krylov space per
// column; BlockRun, one block Krylov space per batch). Iterations are summed
// over the columns (each column counts the passes it took part in), as the
// batched BiCG counts them.
template <class Run>
int gmres_batched(Context* ctx, const ibim_solve_opts* opts, ibim_solve_info* info, double* hist) {
    constexpr int C = kBatchRHS;
    int inner, outer_op;
    refinement_ops(ctx, &inner, &outer_op);
    const int count = ctx->count;
    GmresSpace g;
    if (ctx->refine.ensure(Run::space_bytes(ctx))) return IBIM_ERR_NOMEM;
    if (int rc = Run::space_at(ctx, ctx->refine.ptr, &g)) return rc;
    const unsigned nb = (unsigned)((g.n + kGT - 1) / kGT);
    const int max_outer = opts->max_outer > 0 ? opts->max_outer : 20;
    const int m = basis_of<Run>(opts);
    int total = 0, outer_max = 0, hist_at = 0, status = 4;
    double worst = 0.0;
    for (int c0 = 0; c0 < ctx->nrhs; c0 += C) {
        const int ncol = std::min(C, ctx->nrhs - c0);
        hipLaunchKernelGGL(k_gm_pack, dim3(nb), dim3(kGT), 0, ctx->stream, ctx->rhs.as<double2>(), count, c0, ncol,
                           g.b);
        IBIM_HIP(hipGetLastError());
        IBIM_HIP(hipMemsetAsync(g.x, 0, sizeof(double2) * g.vs, ctx->stream));
        IBIM_HIP(hipMemcpyAsync(g.res, g.b, sizeof(double2) * g.n, hipMemcpyDeviceToDevice, ctx->stream));
        double normb[C], relres[C], target[C];
        int col_status[C], col_used[C] = {};
        if (int rc = norms_of<C>(ctx, g, g.res, normb)) return rc;
        bool any = false;
        for (int c = 0; c < C; ++c) {
            const bool live = c < ncol && normb[c] > 0.0;
            relres[c] = live ? 1.0 : 0.0;
            col_status[c] = live ? 1 : 4;  // zero data: zero density, no iterations
            target[c] = live ? inner_margin(inner, outer_op) * opts->tol * normb[c] : INFINITY;
            any = any || live;
        }
        int used = 0, steps = 0;
        while (any && steps < max_outer && used < opts->maxit) {
            ++steps;
            Run run;
            run.ctx = ctx;
            run.g = g;
            run.rhs = g.res;
            for (int c = 0; c < C; ++c) run.target[c] = target[c];
            run.budget = opts->maxit - used;
            run.m = m;
            run.col_products = col_used;
            if constexpr (std::is_same<Run, GmresRun<C>>::value) {
                run.hist = c0 == 0 ? hist : nullptr;
                run.hist_at = &hist_at;
                run.hist_len = opts->maxit;
                run.normb = normb[0];
            }
            if (int rc = solve_run(run, inner)) return rc;
            used += run.passes;
            hipLaunchKernelGGL(k_gm_add, dim3(nb), dim3(kGT), 0, ctx->stream, g.x, g.d, g.n);
            if (int rc = residual_pass(ctx, outer_op, g.x, g.w)) return rc;
            hipLaunchKernelGGL(k_gm_sub, dim3(nb), dim3(kGT), 0, ctx->stream, g.b, g.w, g.n, g.res);
            IBIM_HIP(hipGetLastError());
            double rn[C];
            if (int rc = norms_of<C>(ctx, g, g.res, rn)) return rc;
            any = false;
            for (int c = 0; c < C; ++c) {
                if (col_status[c] != 1) continue;
                relres[c] = rn[c] / normb[c];
                if (relres[c] != relres[c])
                    col_status[c] = 2;
                else if (relres[c] < opts->tol)
                    col_status[c] = 0;
                if (col_status[c] != 1) target[c] = INFINITY;  // out of the next corrections
                any = any || col_status[c] == 1;
            }
        }
        hipLaunchKernelGGL(k_gm_unpack, dim3(nb), dim3(kGT), 0, ctx->stream, g.x, count, c0, ncol,
                           ctx->density.as<double2>());
        IBIM_HIP(hipGetLastError());
        IBIM_HIP(hipStreamSynchronize(ctx->stream));
        outer_max = std::max(outer_max, steps);
        for (int c = 0; c < ncol; ++c) {
            total += col_used[c];
            if (col_status[c] == 4) continue;
            worst = std::max(worst, relres[c]);
            if (status == 4 || col_status[c] != 0) status = col_status[c];
        }
        report_progress(ctx, std::is_same<Run, BlockRun>::value ? "block GMRES" : "GMRES (16 columns)", used, worst);
    }
    info->iterations = total;
    info->outer = outer_max;
    info->rel_residual = worst;
    info->status = status;
    ctx->last_iters = total;
    return IBIM_OK;
}

// Enlarged GMRES (Grigori, Moufawad and Nataf, SIAM J. Sci. Comput. 38, 2016)
// for one right-hand side at a time, inside the same FP64 refinement: each
// refinement step splits its residual over 16 pieces of the surface
// (k_gm_split), builds one block Krylov space from them with the 16-column
// passes (BlockRun, enlarged mode) and minimizes the one residual over all of
// it. That space contains GMRES's own, so a pass never does worse than a GMRES
// product, and it holds 16 directions per pass: on the double layer, whose
// slow modes GMRES finds one product at a time, a fraction of the passes
// (numpy model on the oracle's torus operator, 32^3 and 48^3, k = 10: 136
// GMRES products, 18 passes).
struct EnlargedSpace {
    GmresSpace g, g1;  // the block space; its single-vector view for the norms
    double2 *res, *ax, *ay;
};

int enlarged_space(Context* ctx, EnlargedSpace* e) {
    const size_t ld = ctx->ld;
    const size_t bbytes = (BlockRun::space_bytes(ctx) + 255) & ~(size_t)255;
    if (ctx->refine.ensure(bbytes + sizeof(double2) * ld * 3)) return IBIM_ERR_NOMEM;
    if (int rc = BlockRun::space_at(ctx, ctx->refine.ptr, &e->g)) return rc;
    e->res = reinterpret_cast<double2*>(ctx->refine.as<char>() + bbytes);  // b - op(A64) x
    e->ax = e->res + ld;
    e->ay = e->ax + ld;
    e->g1 = e->g;
    e->g1.vs = ld;
    e->g1.n = ctx->count;
    return IBIM_OK;
}

// One right-hand side b: x = 0, then refinement steps of enlarged GMRES with
// `inner` (A^H when adjoint) and FP64 residuals with `outer`, until
// ||b - op(A) x|| < tol ||b||. Reports the passes, steps, residual and status
// (0 converged, 1 budget, 2 NaN, 4 zero data).
int enlarged_core(Context* ctx, const EnlargedSpace& e, int inner, int outer_op, bool adjoint, const double2* b,
                  double2* x, double tol, int maxit, int max_outer, int m, int* used_out, int* steps_out,
                  double* relres_out, int* status_out) {
    const int count = ctx->count;
    const unsigned nb1 = (unsigned)((count + kGT - 1) / kGT), nbb = (unsigned)((e.g.n + kGT - 1) / kGT);
    IBIM_HIP(hipMemsetAsync(x, 0, sizeof(double2) * count, ctx->stream));
    IBIM_HIP(hipMemcpyAsync(e.res, b, sizeof(double2) * count, hipMemcpyDeviceToDevice, ctx->stream));
    double normb;
    if (int rc = norms_of<1>(ctx, e.g1, e.res, &normb)) return rc;
    *used_out = *steps_out = 0;
    *relres_out = 0.0;
    *status_out = 4;
    if (normb == 0.0) return IBIM_OK;  // zero data: zero solution, no iterations
    int col_status = 1, used = 0, steps = 0;
    double relres = 1.0;
    while (steps < max_outer && used < maxit) {
        ++steps;
        hipLaunchKernelGGL(k_gm_split, dim3(nbb), dim3(kGT), 0, ctx->stream, e.res, count, e.g.b);
        IBIM_HIP(hipGetLastError());
        BlockRun run;
        run.single = true;
        run.adjoint = adjoint;
        run.ctx = ctx;
        run.g = e.g;
        run.rhs = e.g.b;
        for (int q = 0; q < kBatchRHS; ++q) run.target[q] = INFINITY;
        run.target[0] = inner_margin(inner, outer_op) * tol * normb;
        run.budget = maxit - used;
        run.m = m;
        int prods[kBatchRHS] = {};
        run.col_products = prods;
        // a few tens of passes per solve: the 64-row shape, untimed
        const bool untimed = ctx->batch_untimed;
        ctx->batch_untimed = true;
        const int rc = solve_run(run, inner);
        ctx->batch_untimed = untimed;
        if (rc) return rc;
        used += run.passes;
        hipLaunchKernelGGL(k_gm_add_col0, dim3(nb1), dim3(kGT), 0, ctx->stream, x, e.g.d, count);
        if (adjoint) {  // A^H x: the fused sweep's second product
            if (int rc2 = matvec_op(ctx, outer_op, x, x, e.ay, e.ax)) return rc2;
        } else if (int rc2 = matvec_op_y(ctx, outer_op, x, e.ax)) {
            return rc2;
        }
        hipLaunchKernelGGL(k_gm_sub, dim3(nb1), dim3(kGT), 0, ctx->stream, b, e.ax, count, e.res);
        IBIM_HIP(hipGetLastError());
        double rn;
        if (int rc2 = norms_of<1>(ctx, e.g1, e.res, &rn)) return rc2;
        relres = rn / normb;
        if (relres != relres) {
            col_status = 2;
            break;
        }
        if (relres < tol) {
            col_status = 0;
            break;
        }
    }
    *used_out = used;
    *steps_out = steps;
    *relres_out = relres;
    *status_out = col_status;
    return IBIM_OK;
}

int gmres_enlarged(Context* ctx, const ibim_solve_opts* opts, ibim_solve_info* info) {
    int inner, outer_op;
    refinement_ops(ctx, &inner, &outer_op);
    const int count = ctx->count;
    EnlargedSpace e;
    if (int rc = enlarged_space(ctx, &e)) return rc;
    const int max_outer = opts->max_outer > 0 ? opts->max_outer : 20;
    const int m = basis_of<BlockRun>(opts);
    int total = 0, outer_max = 0, status = 4;
    double worst = 0.0;
    for (int c = 0; c < ctx->nrhs; ++c) {
        const double2* b = ctx->rhs.as<double2>() + (size_t)c * count;
        double2* x = ctx->density.as<double2>() + (size_t)c * count;
        int used, steps, col_status;
        double relres;
        if (int rc = enlarged_core(ctx, e, inner, outer_op, false, b, x, opts->tol, opts->maxit, max_outer, m, &used,
                                   &steps, &relres, &col_status))
            return rc;
        if (col_status == 4) continue;
        total += used;
        outer_max = std::max(outer_max, steps);
        worst = std::max(worst, relres);
        if (status == 4 || col_status != 0) status = col_status;
        report_progress(ctx, "enlarged GMRES", used, relres);
    }
    info->iterations = total;
    info->outer = outer_max;
    info->rel_residual = worst;
    info->status = status;
    ctx->last_iters = total;
    return IBIM_OK;
}

// One batch of up to 16 columns of one problem in a grouped solve.
template <class Run>
struct GroupBatch {
    Context* ctx = nullptr;
    int c0 = 0, ncol = 0;
    int inner = OP_DENSE, outer = OP_DENSE;
    GmresSpace g{};
    Run run;
    double normb[kBatchRHS] = {}, relres[kBatchRHS] = {}, target[kBatchRHS] = {};
    int status[kBatchRHS] = {}, used[kBatchRHS] = {};
    int passes = 0, outer_steps = 0;
    bool running = false;
};

// Several pair-stored problems at once (ibim_solve_group): every batch of 16
// columns of every problem runs the refinement of gmres_batched, and the
// GMRES products of all batches still iterating are ONE persistent launch per
// iteration (pair_group_pass, Y = A P alone; largest problems first), followed
// by each batch's near field and Gram-Schmidt work and one host
// synchronization for all. The refinement steps are taken together: a batch
// whose GMRES has finished waits for the others before its FP64 residual.
// Each batch takes the same steps, and gives the same bits, as its own
// batched solve.
template <class Run>
int group_gmres(Context* const* ctxs, int n, const ibim_solve_opts* opts, ibim_solve_info* infos) {
    constexpr int C = kBatchRHS;
    int source = 0;
    if (int rc = pair_group_source(ctxs, n, false, &source)) return rc;
    Context* lead = ctxs[0];
    SharedStream shared;
    if (int rc = shared.share(ctxs, n)) return rc;
    std::vector<int> order(n);
    for (int c = 0; c < n; ++c) order[c] = c;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return ctxs[x]->count > ctxs[y]->count; });
    std::vector<GroupBatch<Run>> bs;
    for (int c : order)
        for (int c0 = 0; c0 < ctxs[c]->nrhs; c0 += C) {
            GroupBatch<Run> b;
            b.ctx = ctxs[c];
            b.c0 = c0;
            b.ncol = std::min(C, ctxs[c]->nrhs - c0);
            bs.push_back(std::move(b));
        }
    const int nb = (int)bs.size();
    // workspaces, then the queue counter and the member table
    std::vector<size_t> off(nb + 1, 0);
    for (int m = 0; m < nb; ++m) off[m + 1] = off[m] + ((Run::space_bytes(bs[m].ctx) + 255) & ~(size_t)255);
    const size_t table = off[nb] + 256;
    if (lead->group.ensure(table + sizeof(GroupMember) * nb)) return IBIM_ERR_NOMEM;
    char* base = lead->group.as<char>();
    int* queue = reinterpret_cast<int*>(base + off[nb]);
    GroupMember* dev_members = reinterpret_cast<GroupMember*>(base + table);
    const int maxit = opts->maxit;
    const int max_outer = opts->max_outer > 0 ? opts->max_outer : 20;
    const int mlen = basis_of<Run>(opts);
    for (int m = 0; m < nb; ++m) {
        GroupBatch<Run>& b = bs[m];
        Context* ctx = b.ctx;
        refinement_ops(ctx, &b.inner, &b.outer);
        if (int rc = Run::space_at(ctx, base + off[m], &b.g)) return rc;
        const unsigned blocks = (unsigned)((b.g.n + kGT - 1) / kGT);
        hipLaunchKernelGGL(k_gm_pack, dim3(blocks), dim3(kGT), 0, ctx->stream, ctx->rhs.as<double2>(), ctx->count,
                           b.c0, b.ncol, b.g.b);
        IBIM_HIP(hipGetLastError());
        IBIM_HIP(hipMemcpyAsync(b.g.res, b.g.b, sizeof(double2) * b.g.n, hipMemcpyDeviceToDevice, ctx->stream));
        if (int rc = norms_of<C>(ctx, b.g, b.g.res, b.normb)) return rc;
        for (int c = 0; c < C; ++c) {
            const bool live = c < b.ncol && b.normb[c] > 0.0;
            b.relres[c] = live ? 1.0 : 0.0;
            b.status[c] = live ? 1 : 4;
            b.target[c] = live ? inner_margin(b.inner, b.outer) * opts->tol * b.normb[c] : INFINITY;
            b.running = b.running || live;
        }
    }
    std::vector<GroupMember> tab;
    for (int outer = 1; outer <= max_outer; ++outer) {
        // a GMRES solve for every batch with running columns and budget left
        std::vector<int> live, stepped;
        for (int m = 0; m < nb; ++m) {
            GroupBatch<Run>& b = bs[m];
            if (!b.running || b.passes >= maxit) continue;
            Run& r = b.run;
            r.ctx = b.ctx;
            r.g = b.g;
            r.rhs = b.g.res;
            for (int c = 0; c < C; ++c) r.target[c] = b.target[c];
            r.budget = maxit - b.passes;
            r.m = mlen;
            r.col_products = b.used;
            if (int rc = r.start()) return rc;
            stepped.push_back(m);
            if (!r.done()) live.push_back(m);
        }
        if (stepped.empty()) break;
        while (!live.empty()) {
            tab.clear();
            int items = 0;
            for (int m : live) {
                const GroupBatch<Run>& b = bs[m];
                const Context* ctx = b.ctx;
                GroupMember d{};
                d.F = ctx->mat.ptr;
                d.blk = ctx->pair_blk.as<int>();
                d.pts = ctx->pair_pts.ptr;
                d.P = b.run.in();
                d.Y = b.run.out();
                d.prm = ctx->prm;
                d.count = ctx->count;
                d.item0 = items;
                items += (ctx->count + 63) / 64;
                tab.push_back(d);
            }
            IBIM_HIP(hipMemcpyAsync(dev_members, tab.data(), sizeof(GroupMember) * tab.size(), hipMemcpyHostToDevice,
                                    lead->stream));
            if (int rc = pair_group_pass(lead, source, dev_members, (int)tab.size(), items, queue, false)) return rc;
            for (int m : live) {
                GroupBatch<Run>& b = bs[m];
                if (int rc = pair_batch_near(b.ctx, b.run.in(), nullptr, b.run.out(), nullptr, nullptr)) return rc;
                if (int rc = b.run.launch_post()) return rc;
            }
            IBIM_HIP(hipStreamSynchronize(lead->stream));  // also: tab may change
            std::vector<int> still;
            for (int m : live) {
                if (int rc = bs[m].run.host_post()) return rc;
                if (!bs[m].run.done()) still.push_back(m);
            }
            live.swap(still);
        }
        // the FP64 residual of every batch that took a step
        double worst = 0.0;
        for (int m : stepped) {
            GroupBatch<Run>& b = bs[m];
            Context* ctx = b.ctx;
            const unsigned blocks = (unsigned)((b.g.n + kGT - 1) / kGT);
            b.passes += b.run.passes;
            b.outer_steps = outer;
            hipLaunchKernelGGL(k_gm_add, dim3(blocks), dim3(kGT), 0, ctx->stream, b.g.x, b.g.d, b.g.n);
            if (int rc = residual_pass(ctx, b.outer, b.g.x, b.g.w)) return rc;
            hipLaunchKernelGGL(k_gm_sub, dim3(blocks), dim3(kGT), 0, ctx->stream, b.g.b, b.g.w, b.g.n, b.g.res);
            IBIM_HIP(hipGetLastError());
            double rn[C];
            if (int rc = norms_of<C>(ctx, b.g, b.g.res, rn)) return rc;
            b.running = false;
            for (int c = 0; c < C; ++c) {
                if (b.status[c] != 1) continue;
                b.relres[c] = rn[c] / b.normb[c];
                if (b.relres[c] != b.relres[c])
                    b.status[c] = 2;
                else if (b.relres[c] < opts->tol)
                    b.status[c] = 0;
                if (b.status[c] != 1) b.target[c] = INFINITY;
                b.running = b.running || b.status[c] == 1;
                if (b.status[c] == 1) worst = std::max(worst, b.relres[c]);
            }
        }
        report_progress(lead, std::is_same<Run, BlockRun>::value ? "grouped block GMRES, refinement"
                                                                 : "grouped GMRES, refinement",
                        outer, worst);
    }
    // per problem: its columns' results, summed as the one-problem solver does
    for (int c = 0; c < n; ++c) {
        Context* ctx = ctxs[c];
        ibim_solve_info& info = infos[c];
        info = ibim_solve_info{};
        int status = 4;
        for (GroupBatch<Run>& b : bs) {
            if (b.ctx != ctx) continue;
            const unsigned blocks = (unsigned)((b.g.n + kGT - 1) / kGT);
            hipLaunchKernelGGL(k_gm_unpack, dim3(blocks), dim3(kGT), 0, ctx->stream, b.g.x, ctx->count, b.c0, b.ncol,
                               ctx->density.as<double2>());
            IBIM_HIP(hipGetLastError());
            info.outer = std::max(info.outer, b.outer_steps);
            for (int col = 0; col < b.ncol; ++col) {
                info.iterations += b.used[col];
                if (b.status[col] == 4) continue;
                info.rel_residual = std::max(info.rel_residual, b.relres[col]);
                if (status == 4 || b.status[col] != 0) status = b.status[col];
            }
        }
        info.status = status;
        ctx->last_iters = info.iterations;
    }
    IBIM_HIP(hipStreamSynchronize(lead->stream));
    return IBIM_OK;
}

}  // namespace

int enlarged_solve_op(Context* ctx, int op, bool adjoint, const void* b_dev, void* x_dev, int maxit, double tol,
                      ibim_solve_info* info) {
    EnlargedSpace e;
    if (int rc = enlarged_space(ctx, &e)) return rc;
    int used, steps, status;
    double relres;
    if (int rc = enlarged_core(ctx, e, op, op, adjoint, static_cast<const double2*>(b_dev), static_cast<double2*>(x_dev),
                               tol, maxit, 20, kMaxBlocks, &used, &steps, &relres, &status))
        return rc;
    info->iterations = used;
    info->outer = steps;
    info->rel_residual = relres;
    info->status = status;
    return IBIM_OK;
}

int run_group_gmres(Context* const* ctxs, int n, const ibim_solve_opts* opts, ibim_solve_info* infos) {
    return opts->method == IBIM_SOLVE_BLOCK_GMRES ? group_gmres<BlockRun>(ctxs, n, opts, infos)
                                                  : group_gmres<GmresRun<kBatchRHS>>(ctxs, n, opts, infos);
}

int run_gmres(Context* ctx, const ibim_solve_opts* opts, ibim_solve_info* info, double* hist) {
    // 16-column passes where the batched products exist: every operator of
    // the traditional tube, or dense complex128 storage refined against itself
    const bool batched = ctx->nrhs > 1 && opts->batch == 0 &&
                         (ctx->kind == IBIM_TUBE_TRADITIONAL || ctx->mat_precision == IBIM_C128);
    if (!batched) {
        // one column at a time: enlarged GMRES where the 16-column passes exist
        const bool passes = ctx->kind == IBIM_TUBE_TRADITIONAL || ctx->mat_precision == IBIM_C128;
        return opts->method == IBIM_SOLVE_BLOCK_GMRES && passes ? gmres_enlarged(ctx, opts, info)
                                                                : gmres_columns(ctx, opts, info, hist);
    }
    return opts->method == IBIM_SOLVE_BLOCK_GMRES ? gmres_batched<BlockRun>(ctx, opts, info, hist)
                                                  : gmres_batched<GmresRun<kBatchRHS>>(ctx, opts, info, hist);
}

}  // namespace ibim
