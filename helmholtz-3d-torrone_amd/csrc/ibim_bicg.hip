// ibim_bicg.hip -- the reference's complex BiCG (BICGSTAB.c:5-339/344-765) on a
// dense operator in HBM, with the matrix streamed once per iteration.
//
// Fused sweep. One pass over A gives both v = A p and w = A^H p~. A workgroup
// owns a kMvRows x kMvCols tile; thread t owns columns c0 + t + 256 m
// (m < kMvCpt), so every row of the tile is read as kMvCpt coalesced 1 KiB
// (c128) wavefront loads. Per row, the thread's A p contribution is reduced
// across the wavefront with an xor butterfly and parked in LDS; the A^H p~
// contribution stays in registers for the whole tile (p~ of the row is a
// wave-uniform load). Tile results go to two partial planes
// (ypart[column tile][row], zpart[row tile][column]) which the next kernel sums
// in a fixed order: no atomics, so repeated solves are bitwise identical.
//
// BiCG scalars never leave the device: each iteration is four launches
// (sweep, dot, update x/r/r~, update p/p~); every workgroup of the update
// kernels re-reduces the previous kernel's per-block partials in the same fixed
// order, so all of them hold identical alpha / beta. A per-iteration record
// (BicgScalars) carries rho, alpha, beta, the residual and the stop flag; once a
// stop rule fires, later launches of the chunk return at entry.
#include <cmath>
#include <vector>

#include "ibim_internal.hpp"
#include "ibim_wave.hpp"

namespace ibim {

namespace {

constexpr int kMvThreads = 256;
constexpr int kMvCpt = 4;
constexpr int kMvCols = kMvThreads * kMvCpt;
constexpr int kMvRows = kSweepRows;
static_assert(kMvCols == kSweepCols, "sweep tiles must match the partial planes");
constexpr int kMvUnroll = 4;
constexpr int kVecThreads = 256;
constexpr int kMaxVecBlocks = 1024;

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Entries are applied in FP64 whatever their storage: a stored complex64
// operator is then one fixed matrix for BiCG (FP32 products or sums make it
// vector-dependent at the 1e-5 level, which doubled the iteration count on a
// 65k-point torus). With the transposed row reduction (wave_sum8_transposed)
// the FP64 sweep over complex64 storage still streams at the HBM read rate.

static_assert(kMvUnroll == 4, "the transposed row reduction packs 4 rows x (re, im)");

// The row loop of one tile. Guard = false for interior tiles (all kMvRows rows
// and kMvCols columns in range): unpredicated loads keep the complex64 sweep at
// the HBM rate (per-element guards cost it 40%).
template <bool Guard, class MT>
__device__ __forceinline__ void sweep_rows(const MT* __restrict__ Ab, size_t ld, int nrow, const bool (&cok)[kMvCpt],
                                           const double (&pre)[kMvCpt], const double (&pim)[kMvCpt],
                                           double (&zre)[kMvCpt], double (&zim)[kMvCpt],
                                           const double2* __restrict__ qt, double (&ysh)[2][kMvThreads / 64][kMvRows],
                                           int lane, int wave, int my_val) {
    for (int rr = 0; rr < nrow; rr += kMvUnroll) {
        MT a[kMvUnroll][kMvCpt];
#pragma unroll
        for (int u = 0; u < kMvUnroll; ++u)
#pragma unroll
            for (int m = 0; m < kMvCpt; ++m) {
                if (!Guard || (rr + u < nrow && cok[m])) {
                    a[u][m] = load_stream(Ab + (size_t)(rr + u) * ld + kMvThreads * m);
                } else {
                    a[u][m].x = 0;
                    a[u][m].y = 0;
                }
            }
        double rows[2 * kMvUnroll];
        double zr[kMvCpt], zi[kMvCpt];
#pragma unroll
        for (int m = 0; m < kMvCpt; ++m) zr[m] = zi[m] = 0.0;
#pragma unroll
        for (int u = 0; u < kMvUnroll; ++u) {
            const double2 qv = (!Guard || rr + u < nrow) ? qt[rr + u] : make_double2(0.0, 0.0);
            double sre = 0.0, sim = 0.0;
#pragma unroll
            for (int m = 0; m < kMvCpt; ++m) {
                // explicit FMAs: the file is built without contraction for the
                // BiCG vector updates, and separate multiplies and adds would
                // drop the complex64 sweep off the HBM rate
                const double ar = (double)a[u][m].x, ai = (double)a[u][m].y;
                sre = fma(ar, pre[m], fma(-ai, pim[m], sre));
                sim = fma(ar, pim[m], fma(ai, pre[m], sim));
                zr[m] = fma(ar, qv.x, fma(ai, qv.y, zr[m]));  // conj(a) q
                zi[m] = fma(ar, qv.y, fma(-ai, qv.x, zi[m]));
            }
            rows[2 * u] = sre;
            rows[2 * u + 1] = sim;
        }
#pragma unroll
        for (int m = 0; m < kMvCpt; ++m) {
            zre[m] += zr[m];
            zim[m] += zi[m];
        }
        const double tot = wave_sum8_transposed(rows, lane);
        const int row = rr + (my_val >> 1);
        if ((lane & 7) == 0 && (!Guard || row < nrow)) ysh[my_val & 1][wave][row] = tot;
    }
}

template <class MT>
__global__ void __launch_bounds__(kMvThreads) k_sweep(const MT* __restrict__ A, size_t ld, int count,
                                                      const double2* __restrict__ p, const double2* __restrict__ q,
                                                      double2* __restrict__ ypart, double2* __restrict__ zpart,
                                                      const int* __restrict__ done) {
    if (done != nullptr && *done) return;
    __shared__ double ysh[2][kMvThreads / 64][kMvRows];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int c0 = blockIdx.x * kMvCols;
    const int r0 = blockIdx.y * kMvRows;
    const int nrow = min(kMvRows, count - r0);
    const int my_val = wave_sum8_index(lane);

    double pre[kMvCpt], pim[kMvCpt], zre[kMvCpt], zim[kMvCpt];
    bool cok[kMvCpt];
#pragma unroll
    for (int m = 0; m < kMvCpt; ++m) {
        int c = c0 + t + kMvThreads * m;
        cok[m] = c < count;
        double2 pv = cok[m] ? p[c] : make_double2(0.0, 0.0);
        pre[m] = pv.x;
        pim[m] = pv.y;
        zre[m] = zim[m] = 0.0;
    }
    const MT* Ab = A + (size_t)r0 * ld + c0 + t;
    if (nrow == kMvRows && c0 + kMvCols <= count)
        sweep_rows<false>(Ab, ld, nrow, cok, pre, pim, zre, zim, q + r0, ysh, lane, wave, my_val);
    else
        sweep_rows<true>(Ab, ld, nrow, cok, pre, pim, zre, zim, q + r0, ysh, lane, wave, my_val);
    __syncthreads();
    for (int row = t; row < nrow; row += kMvThreads) {
        double sre = 0.0, sim = 0.0;
#pragma unroll
        for (int w = 0; w < kMvThreads / 64; ++w) {
            sre += ysh[0][w][row];
            sim += ysh[1][w][row];
        }
        ypart[(size_t)blockIdx.x * ld + r0 + row] = make_double2(sre, sim);
    }
#pragma unroll
    for (int m = 0; m < kMvCpt; ++m) {
        int c = c0 + t + kMvThreads * m;
        if (cok[m]) zpart[(size_t)blockIdx.y * ld + c] = make_double2(zre[m], zim[m]);
    }
}

// Fixed-order block reduction of NV doubles per thread; result valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV], double* sh /* NV * 4 */) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double s = wave_sum(v[q]);
        if (lane == 0) sh[q * 4 + wave] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) v[q] = (sh[q * 4 + 0] + sh[q * 4 + 1]) + (sh[q * 4 + 2] + sh[q * 4 + 3]);
    }
}

// Every workgroup reduces the NV per-block partials of the previous kernel in
// the same order (32 threads per value, blocks strided by 32, xor tree).
template <int NV>
__device__ __forceinline__ void reduce_partials(const double* __restrict__ part, int nblocks, double* out /* LDS NV */) {
    static_assert(NV * 32 <= kVecThreads, "too many partial values");
    const int t = threadIdx.x, q = t >> 5, s = t & 31;
    double acc = 0.0;
    if (q < NV)
        for (int b = s; b < nblocks; b += 32) acc += part[(size_t)b * NV + q];
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 32);
    if (q < NV && s == 0) out[q] = acc;
    __syncthreads();
}

struct VecArgs {
    double2 *x, *r, *rt, *p, *pt, *v, *w;
    const double2* b;
    // Partial planes of the last sweep. Dense and matrix-free sweeps: v sums
    // ncol_tiles planes of ypart, w sums nrow_tiles planes of zpart. Pair
    // storage (pair = 1): ypart and zpart alike hold ncol_tiles row-side
    // planes, nrow_tiles column-side planes and a near-field plane, of which
    // element j sums those its tiles wrote (ibim_pair.hip).
    double2 *ypart, *zpart;
    int ncol_tiles, nrow_tiles;
    int pair;
    size_t ld;
    double* part;  // per-block partials
    BicgScalars* sc;
    double* norms;  // [0] = ||b||, [1] = sum |re|+|im| of b
    int* done;
    // the iteration the vector kernels work on, kept on the device so that a
    // chunk of iterations is the same launches every time (HIP graph replay);
    // k_bicg_dot advances it, null outside the BiCG loop
    int* iter;
    int count;
};

__device__ __forceinline__ double2 sum_planes(const double2* __restrict__ part, int nplanes, size_t ld, int j) {
    double sr = 0.0, si = 0.0;
#pragma unroll 8
    for (int c = 0; c < nplanes; ++c) {
        const double2 y = part[(size_t)c * ld + j];
        sr += y.x;
        si += y.y;
    }
    return make_double2(sr, si);
}

__device__ __forceinline__ double2 sum_pair_planes(const double2* __restrict__ part, int ncol, int nrow, int count,
                                                   size_t ld, int j) {
    const int blk = j / kSweepRows;
    const int nt = (count - blk * kSweepRows + kSweepCols - 1) / kSweepCols;
    double sr = 0.0, si = 0.0;
    // unrolled for loads in flight; the sums keep the plane order
#pragma unroll 8
    for (int c = 0; c < nt; ++c) {
        const double2 y = part[(size_t)c * ld + j];
        sr += y.x;
        si += y.y;
    }
#pragma unroll 8
    for (int c = 0; c <= blk; ++c) {
        const double2 y = part[(size_t)(ncol + c) * ld + j];
        sr += y.x;
        si += y.y;
    }
    const double2 y = part[(size_t)(ncol + nrow) * ld + j];
    return make_double2(sr + y.x, si + y.y);
}

// (A p)_j and (A^H q)_j of the last sweep.
__device__ __forceinline__ double2 sweep_y(const VecArgs& a, int j) {
    return a.pair ? sum_pair_planes(a.ypart, a.ncol_tiles, a.nrow_tiles, a.count, a.ld, j)
                  : sum_planes(a.ypart, a.ncol_tiles, a.ld, j);
}
__device__ __forceinline__ double2 sweep_z(const VecArgs& a, int j) {
    return a.pair ? sum_pair_planes(a.zpart, a.ncol_tiles, a.nrow_tiles, a.count, a.ld, j)
                  : sum_planes(a.zpart, a.nrow_tiles, a.ld, j);
}

// x = 0, r = p = b, r~ = p~ = conj(b) (BICGSTAB.c:104-117); partials of
// <r, r~> (4 sums), ||b||^2 (2 sums) and sum |b_re| + |b_im|.
__global__ void __launch_bounds__(kVecThreads) k_bicg_init(VecArgs a) {
    __shared__ double sh[7 * 4];
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int j = blockIdx.x * kVecThreads + threadIdx.x; j < a.count; j += gridDim.x * kVecThreads) {
        double2 b = a.b[j];
        double2 c = make_double2(b.x, (-1.0) * b.y);
        a.x[j] = make_double2(0.0, 0.0);
        a.r[j] = b;
        a.p[j] = b;
        a.rt[j] = c;
        a.pt[j] = c;
        v[0] += b.x * c.x;
        v[1] += b.y * c.y;
        v[2] += b.y * c.x;
        v[3] += b.x * c.y;
        v[4] += b.x * b.x;
        v[5] += b.y * b.y;
        v[6] += fabs(b.x) + fabs(b.y);
    }
    block_reduce<7>(v, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < 7; ++q) a.part[(size_t)blockIdx.x * 7 + q] = v[q];
}

__global__ void __launch_bounds__(kVecThreads) k_bicg_init_finish(VecArgs a, int nblocks) {
    __shared__ double tot[7];
    reduce_partials<7>(a.part, nblocks, tot);
    if (threadIdx.x == 0) {
        BicgScalars s{};
        s.rho_re = tot[0] + tot[1];  // innerrold (BICGSTAB.c:127-128)
        s.rho_im = tot[2] - tot[3];
        s.err = INFINITY;            // no previous error before iteration 1
        a.sc[0] = s;
        a.norms[0] = sqrt(tot[4] + tot[5]);
        a.norms[1] = tot[6];
        *a.done = 0;
        if (a.iter) *a.iter = 0;
    }
}

// The zero-data perturbation of BICGSTAB.c:138-145 (after rho was taken).
__global__ void k_bicg_perturb(VecArgs a) {
    a.r[0] = make_double2(0.000000001, 0.000000001);
    a.rt[0] = make_double2(0.000000001, 0.000000001);
}

// v, w from the partial planes; partials of <r, r~> and <v, p~>.
__global__ void __launch_bounds__(kVecThreads) k_bicg_dot(VecArgs a) {
    if (*a.done) return;
    __shared__ double sh[8 * 4];
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = blockIdx.x * kVecThreads + threadIdx.x; j < a.count; j += gridDim.x * kVecThreads) {
        const double2 y = sweep_y(a, j), z = sweep_z(a, j);
        const double yr = y.x, yi = y.y;
        a.v[j] = y;
        a.w[j] = z;
        double2 r = a.r[j], rt = a.rt[j], pt = a.pt[j];
        v[0] += r.x * rt.x;
        v[1] += r.y * rt.y;
        v[2] += r.y * rt.x;
        v[3] += r.x * rt.y;
        v[4] += yr * pt.x;
        v[5] += yi * pt.y;
        v[6] += yi * pt.x;
        v[7] += yr * pt.y;
    }
    block_reduce<8>(v, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < 8; ++q) a.part[(size_t)blockIdx.x * 8 + q] = v[q];
    // the update kernels after this one work on the next iteration
    if (a.iter && blockIdx.x == 0 && threadIdx.x == 0) *a.iter += 1;
}

// alpha = <r, r~> / <v, p~> (BICGSTAB.c:217-229); x += alpha p, r -= alpha v,
// r~ -= conj(alpha) w (231-242); partials of the new <r, r~> and ||r||^2.
__global__ void __launch_bounds__(kVecThreads) k_bicg_update_xr(VecArgs a, int nblocks, double* part_out) {
    if (*a.done) return;
    const int it = *a.iter;
    __shared__ double tot[8];
    __shared__ double sh[6 * 4];
    reduce_partials<8>(a.part, nblocks, tot);
    const double ca = tot[0] + tot[1], cb = tot[2] - tot[3];
    const double cc = tot[4] + tot[5], cd = tot[6] - tot[7];
    const double num = ca * cc + cb * cd, numi = cb * cc - ca * cd, den = cc * cc + cd * cd;
    const double al = num / den, ali = numi / den;
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (int j = blockIdx.x * kVecThreads + threadIdx.x; j < a.count; j += gridDim.x * kVecThreads) {
        double2 x = a.x[j], p = a.p[j], r = a.r[j], vv = a.v[j], rt = a.rt[j], ww = a.w[j];
        x.x = x.x + al * p.x - ali * p.y;
        x.y = x.y + ali * p.x + al * p.y;
        r.x = r.x - al * vv.x + ali * vv.y;
        r.y = r.y - ali * vv.x - al * vv.y;
        rt.x = rt.x - al * ww.x - ali * ww.y;
        rt.y = rt.y + ali * ww.x - al * ww.y;
        a.x[j] = x;
        a.r[j] = r;
        a.rt[j] = rt;
        v[0] += r.x * rt.x;
        v[1] += r.y * rt.y;
        v[2] += r.y * rt.x;
        v[3] += r.x * rt.y;
        v[4] += r.x * r.x;
        v[5] += r.y * r.y;
    }
    block_reduce<6>(v, sh);
    if (threadIdx.x == 0) {
        for (int q = 0; q < 6; ++q) part_out[(size_t)blockIdx.x * 6 + q] = v[q];
        if (blockIdx.x == 0) {
            a.sc[it].alpha_re = al;
            a.sc[it].alpha_im = ali;
        }
    }
}

// beta = <r, r~>_new / <r, r~>_old (BICGSTAB.c:246-259); p = r + beta p,
// p~ = r~ + conj(beta) p~ (271-278); error and stop rules (289-306).
__global__ void __launch_bounds__(kVecThreads) k_bicg_update_p(VecArgs a, const double* __restrict__ part_in,
                                                               int nblocks, double tol) {
    if (*a.done) return;
    const int it = *a.iter;
    __shared__ double tot[6];
    reduce_partials<6>(part_in, nblocks, tot);
    const BicgScalars old = a.sc[it - 1];
    const double ca = tot[0] + tot[1], cb = tot[2] - tot[3];
    const double num = ca * old.rho_re + cb * old.rho_im, numi = cb * old.rho_re - ca * old.rho_im;
    const double den = old.rho_re * old.rho_re + old.rho_im * old.rho_im;
    const double be = num / den, bei = numi / den;
    for (int j = blockIdx.x * kVecThreads + threadIdx.x; j < a.count; j += gridDim.x * kVecThreads) {
        double2 r = a.r[j], p = a.p[j], rt = a.rt[j], pt = a.pt[j];
        double2 np, npt;
        np.x = r.x + be * p.x - bei * p.y;
        np.y = r.y + bei * p.x + be * p.y;
        npt.x = rt.x + be * pt.x + bei * pt.y;
        npt.y = rt.y - bei * pt.x + be * pt.y;
        a.p[j] = np;
        a.pt[j] = npt;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const double err = sqrt(tot[4] + tot[5]) / a.norms[0];
        BicgScalars s = a.sc[it];
        s.rho_re = ca;
        s.rho_im = cb;
        s.beta_re = be;
        s.beta_im = bei;
        s.err = err;
        int stop = 0;
        if (err < tol) stop = 1;
        else if (err != err) stop = 2;
        else if (fabs(old.err - err) < 1e-3 * tol) stop = 3;
        s.stop = stop;
        a.sc[it] = s;
        if (stop) *a.done = it;
    }
}

// r = b - A x from the partial planes of a sweep over x (FP64 residual
// replacement), with per-block partials of |r|^2.
__global__ void __launch_bounds__(kVecThreads) k_bicg_replace(VecArgs a) {
    __shared__ double sh[2 * 4];
    double v[2] = {0.0, 0.0};
    for (int j = blockIdx.x * kVecThreads + threadIdx.x; j < a.count; j += gridDim.x * kVecThreads) {
        const double2 y = sweep_y(a, j);
        const double2 bb = a.b[j];
        const double2 rr = make_double2(bb.x - y.x, bb.y - y.y);
        a.r[j] = rr;
        v[0] += rr.x * rr.x;
        v[1] += rr.y * rr.y;
    }
    block_reduce<2>(v, sh);
    if (threadIdx.x == 0) {
        a.part[(size_t)blockIdx.x * 2] = v[0];
        a.part[(size_t)blockIdx.x * 2 + 1] = v[1];
    }
}

// Restart from the (replaced) residual: r~ = conj(r), p = r, p~ = r~
// (BICGSTAB.c:104-117), with per-block partials of the new <r, r~>.
__global__ void __launch_bounds__(kVecThreads) k_bicg_restart(VecArgs a) {
    __shared__ double sh[4 * 4];
    double v[4] = {0, 0, 0, 0};
    for (int j = blockIdx.x * kVecThreads + threadIdx.x; j < a.count; j += gridDim.x * kVecThreads) {
        const double2 r = a.r[j];
        const double2 c = make_double2(r.x, (-1.0) * r.y);
        a.p[j] = r;
        a.rt[j] = c;
        a.pt[j] = c;
        v[0] += r.x * c.x;
        v[1] += r.y * c.y;
        v[2] += r.y * c.x;
        v[3] += r.x * c.y;
    }
    block_reduce<4>(v, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; ++q) a.part[(size_t)blockIdx.x * 4 + q] = v[q];
}

// The restarted state goes into record `at` (read as the previous iteration's
// by the next one): rho = <r, r~>, no previous error, running again.
__global__ void __launch_bounds__(kVecThreads) k_bicg_restart_finish(VecArgs a, int nblocks, int at) {
    __shared__ double tot[4];
    reduce_partials<4>(a.part, nblocks, tot);
    if (threadIdx.x == 0) {
        BicgScalars s = a.sc[at];
        s.rho_re = tot[0] + tot[1];
        s.rho_im = tot[2] - tot[3];
        s.err = INFINITY;
        s.stop = 0;
        a.sc[at] = s;
        *a.done = 0;
        *a.iter = at;  // the next k_bicg_dot starts iteration at + 1
    }
}

unsigned vec_blocks(int count) {
    unsigned nb = (unsigned)((count + kVecThreads - 1) / kVecThreads);
    return nb > (unsigned)kMaxVecBlocks ? (unsigned)kMaxVecBlocks : nb;
}

}  // namespace

// One sweep y = A p, z = A^H q of operator `op` into the partial planes.
static int sweep_any(Context* ctx, int op, const double2* p, const double2* q, double2* ypart, double2* zpart,
                     const int* done, bool with_z = true) {
    if (op == OP_PAIR64) return sweep_pair_launch(ctx, true, p, q, ypart, zpart, done, with_z);
    if (free_pair_op(ctx, op)) return sweep_pair_free_launch(ctx, op == OP_FREE32, p, q, ypart, zpart, done, with_z);
    if (op != OP_DENSE) return sweep_free_launch(ctx, op, p, q, ypart, zpart, done);
    if (ctx->mat_pair) return sweep_pair_launch(ctx, false, p, q, ypart, zpart, done, with_z);
    const int count = ctx->count;
    dim3 grid((unsigned)((count + kMvCols - 1) / kMvCols), (unsigned)((count + kMvRows - 1) / kMvRows));
    TimedScope ts(ctx, "matvec");
    if (ctx->mat_precision == IBIM_C128)
        hipLaunchKernelGGL(k_sweep<double2>, grid, dim3(kMvThreads), 0, ctx->stream, ctx->mat.as<double2>(), ctx->ld,
                           count, p, q, ypart, zpart, done);
    else
        hipLaunchKernelGGL(k_sweep<float2>, grid, dim3(kMvThreads), 0, ctx->stream, ctx->mat.as<float2>(), ctx->ld,
                           count, p, q, ypart, zpart, done);
    IBIM_HIP(hipGetLastError());
    return IBIM_OK;
}

namespace {

struct Workspace {
    VecArgs a{};
    unsigned nvb = 0;
    double* part2 = nullptr;
};

int prepare(Context* ctx, int maxit, Workspace* ws) {
    const int count = ctx->count;
    const size_t ld = ctx->ld;
    const int nct = (count + kMvCols - 1) / kMvCols, nrt = (count + kMvRows - 1) / kMvRows;
    const unsigned nvb = vec_blocks(count);
    const size_t vbytes = sizeof(double2) * ld;
    // pair storage and the FP64 matrix-free operator of a traditional tube:
    // row-side, column-side and near planes for v and w alike
    const bool pairs = ctx->mat_pair || ctx->kind == IBIM_TUBE_TRADITIONAL;
    const size_t nplanes = pairs ? 2 * (size_t)(nct + nrt + 1) : (size_t)(nct + nrt);
    if (ctx->work.ensure(vbytes * 7) || ctx->partial.ensure(sizeof(double2) * ld * nplanes) ||
        ctx->scal.ensure(sizeof(BicgScalars) * (size_t)(maxit + 1) + sizeof(double) * (size_t)nvb * 16 + 64) ||
        ctx->done.ensure(sizeof(int) * 4))
        return IBIM_ERR_NOMEM;
    char* w = ctx->work.as<char>();
    VecArgs& a = ws->a;
    a.x = (double2*)(w + 0 * vbytes);
    a.r = (double2*)(w + 1 * vbytes);
    a.rt = (double2*)(w + 2 * vbytes);
    a.p = (double2*)(w + 3 * vbytes);
    a.pt = (double2*)(w + 4 * vbytes);
    a.v = (double2*)(w + 5 * vbytes);
    a.w = (double2*)(w + 6 * vbytes);
    a.ypart = ctx->partial.as<double2>();
    a.zpart = a.ypart + ld * (size_t)nct;
    a.ncol_tiles = nct;
    a.nrow_tiles = nrt;
    a.pair = 0;
    a.ld = ld;
    a.sc = ctx->scal.as<BicgScalars>();
    double* tail = (double*)(a.sc + (maxit + 1));
    a.part = tail;
    ws->part2 = tail + (size_t)nvb * 8;
    a.norms = ws->part2 + (size_t)nvb * 6;
    a.done = ctx->done.as<int>();
    a.iter = nullptr;  // set by the BiCG loop only
    a.count = count;
    ws->nvb = nvb;
    return IBIM_OK;
}

// The workspace's vector arguments with the partial-plane layout of operator op.
VecArgs with_layout(const VecArgs& base, const Context* ctx, int op) {
    VecArgs a = base;
    a.pair = pair_layout_op(ctx, op) ? 1 : 0;
    a.zpart = a.ypart + a.ld * (size_t)(a.pair ? a.ncol_tiles + a.nrow_tiles + 1 : a.ncol_tiles);
    return a;
}

// One BiCG solve of A x = b for b, x on the device (length count). Returns
// iterations and status in info; hist (host, may be NULL) gets the residuals.
// FP64 residual replacement (van der Vorst and Ye, SIAM J. Sci. Comput. 22,
// 2000): BiCG runs on a lower-precision operator; whenever its recursive
// residual has dropped by `factor` since the last replacement, and when it
// claims convergence, r is replaced by the true residual b - A x of the FP64
// operator `true_op`. The Krylov recurrences carry on (no restart).
struct Replace {
    int true_op;
    double factor;
    int replacements;
    double true_err;
};
constexpr int kMaxStallRestarts = 3;

// r = b - A_true x into the BiCG workspace; returns ||r|| / ||b||.
int true_residual(Context* ctx, const VecArgs& base, unsigned nvb, Replace* rep, double normb, double* err) {
    const VecArgs a = with_layout(base, ctx, rep->true_op);
    int rc = sweep_any(ctx, rep->true_op, a.x, a.x, a.ypart, a.zpart, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bicg_replace, dim3(nvb), dim3(kVecThreads), 0, ctx->stream, a);
    IBIM_HIP(hipGetLastError());
    std::vector<double> h(2 * (size_t)nvb);
    IBIM_HIP(hipMemcpyAsync(h.data(), a.part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    double s2 = 0.0;
    for (double v : h) s2 += v;
    *err = sqrt(s2) / normb;
    rep->replacements++;
    rep->true_err = *err;
    return IBIM_OK;
}

// ref_guards: the reference's rules for tiny data (BICGSTAB.c:133-156: zero
// density below 1e-15, a 1e-9 perturbation below 1e-13, in sum |b|); off for
// correction and adjoint solves, whose data are small by construction, where
// only an exactly zero right-hand side gives a zero solution.
int bicg_device(Context* ctx, int op, Workspace& ws, const double2* b, double2* x_out, int maxit, double tol,
                int chunk, ibim_solve_info* info, double* hist, bool adjoint = false, Replace* rep = nullptr,
                bool ref_guards = true) {
    // the matrix-free pair tables are built here, before any graph capture
    if (free_pair_op(ctx, op) || (rep && free_pair_op(ctx, rep->true_op)))
        if (int rc = prepare_pair_free(ctx)) return rc;
    VecArgs a = with_layout(ws.a, ctx, op);
    a.b = b;
    a.iter = ctx->done.as<int>() + 1;
    // The operator's adjoint system A^H x = b runs the same recurrences: one
    // sweep with (p~, p) in place of (p, p~) gives A p~ and A^H p, so the dot
    // kernel reads the two partial planes the other way round.
    VecArgs ad = a;
    if (adjoint) {
        ad.ypart = a.zpart;
        ad.zpart = a.ypart;
        if (!a.pair) {  // the pair layout is the same for v and w
            ad.ncol_tiles = a.nrow_tiles;
            ad.nrow_tiles = a.ncol_tiles;
        }
    }
    const unsigned nvb = ws.nvb;
    hipLaunchKernelGGL(k_bicg_init, dim3(nvb), dim3(kVecThreads), 0, ctx->stream, a);
    hipLaunchKernelGGL(k_bicg_init_finish, dim3(1), dim3(kVecThreads), 0, ctx->stream, a, (int)nvb);
    IBIM_HIP(hipGetLastError());
    double norms[2];
    IBIM_HIP(hipMemcpyAsync(norms, a.norms, sizeof norms, hipMemcpyDeviceToHost, ctx->stream));
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    info->iterations = 0;
    info->rel_residual = 0.0;
    if (ref_guards && norms[1] < 1e-13 && norms[1] >= 1e-15)
        hipLaunchKernelGGL(k_bicg_perturb, dim3(1), dim3(1), 0, ctx->stream, a);
    if (ref_guards ? norms[1] < 1e-15 : norms[1] == 0.0) {  // zero data: zero density (BICGSTAB.c:147-156)
        IBIM_HIP(hipMemsetAsync(x_out, 0, sizeof(double2) * ctx->count, ctx->stream));
        info->status = 4;
        return IBIM_OK;
    }
    int it = 1, done = 0;
    bool verified = false;  // the last convergence claim was confirmed by the true residual
    int stalls = 0;         // stagnation stops checked against the true residual
    if (chunk <= 0) chunk = ctx->count > 20000 ? 2 : 16;
    double next_replace = rep ? rep->factor : 0.0;
    // one iteration: the same launches every time (the kernels read the
    // iteration from the device), so a chunk can be captured once and replayed
    auto launch_iteration = [&]() -> int {
        int rc = adjoint ? sweep_any(ctx, op, a.pt, a.p, a.ypart, a.zpart, a.done)
                         : sweep_any(ctx, op, a.p, a.pt, a.ypart, a.zpart, a.done);
        if (rc) return rc;
        hipLaunchKernelGGL(k_bicg_dot, dim3(nvb), dim3(kVecThreads), 0, ctx->stream, adjoint ? ad : a);
        hipLaunchKernelGGL(k_bicg_update_xr, dim3(nvb), dim3(kVecThreads), 0, ctx->stream, a, (int)nvb, ws.part2);
        hipLaunchKernelGGL(k_bicg_update_p, dim3(nvb), dim3(kVecThreads), 0, ctx->stream, a, ws.part2, (int)nvb,
                           tol);
        return IBIM_OK;
    };
    // HIP graph of `chunk` iterations (no memset or copy nodes: kernels only);
    // not with profiling, whose timing events wrap the individual launches
    struct GraphExec {
        hipGraphExec_t g = nullptr;
        ~GraphExec() {
            if (g) (void)hipGraphExecDestroy(g);
        }
    } chunk_graph;
    const bool graphs = ctx->use_graphs && !ctx->profile;
    while (it <= maxit && !done) {
        int last = it + chunk - 1 < maxit ? it + chunk - 1 : maxit;
        if (graphs && last - it + 1 == chunk) {
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
        IBIM_HIP(hipMemcpyAsync(&done, a.done, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        BicgScalars at_last{};
        if (rep || ctx->progress)
            IBIM_HIP(hipMemcpyAsync(&at_last, a.sc + last, sizeof at_last, hipMemcpyDeviceToHost, ctx->stream));
        IBIM_HIP(hipStreamSynchronize(ctx->stream));
        if (!done) report_progress(ctx, adjoint ? "adjoint BiCG" : "BiCG", last, at_last.err);
        if (!rep) continue;
        BicgScalars cur = at_last;
        if (done) {
            IBIM_HIP(hipMemcpyAsync(&cur, a.sc + done, sizeof cur, hipMemcpyDeviceToHost, ctx->stream));
            IBIM_HIP(hipStreamSynchronize(ctx->stream));
        }
        // a stagnation stop is checked like a claim (the FP32 operator's
        // recursive residual can stall above the FP64 level) up to
        // kMaxStallRestarts times; NaN ends the solve
        const bool stalled = done && cur.stop == 3 && stalls < kMaxStallRestarts;
        if (stalled) ++stalls;
        const bool claims = done && (cur.stop == 1 || stalled);
        if (done && !claims) break;
        if (!claims && !(cur.err < next_replace)) continue;
        double terr = 0.0;
        int rc = true_residual(ctx, a, nvb, rep, norms[0], &terr);
        if (rc) return rc;
        if (claims && terr < tol) {
            verified = true;
            break;
        }
        // A convergence claim the true residual refutes: restart from the
        // last real iteration with the true residual. By then the shadow
        // residual and the search directions have converged with the
        // recursive residual and cannot reduce the difference, so BiCG
        // restarts (r~ = conj(r), p = r) instead of resuming.
        if (done) {
            it = done + 1;
            hipLaunchKernelGGL(k_bicg_restart, dim3(nvb), dim3(kVecThreads), 0, ctx->stream, a);
            hipLaunchKernelGGL(k_bicg_restart_finish, dim3(1), dim3(kVecThreads), 0, ctx->stream, a, (int)nvb,
                               done);
            IBIM_HIP(hipGetLastError());
            done = 0;
        }
        next_replace = terr * rep->factor;
    }
    // a solve that stopped without a confirmed claim (stagnation, NaN, the
    // iteration cap) reports the true residual of its final iterate
    if (rep && !verified) {
        double terr = 0.0;
        if (int rc = true_residual(ctx, a, nvb, rep, norms[0], &terr)) return rc;
    }
    const int iters = done ? done : maxit;
    std::vector<BicgScalars> sc((size_t)iters + 1);
    IBIM_HIP(hipMemcpyAsync(sc.data(), a.sc, sizeof(BicgScalars) * sc.size(), hipMemcpyDeviceToHost, ctx->stream));
    if (x_out != a.x)
        IBIM_HIP(hipMemcpyAsync(x_out, a.x, sizeof(double2) * ctx->count, hipMemcpyDeviceToDevice, ctx->stream));
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    if (hist)
        for (int i = 1; i <= iters; ++i) hist[i - 1] = sc[i].err;
    info->iterations = iters;
    info->rel_residual = rep && rep->replacements ? rep->true_err : sc[iters].err;
    info->status = done ? (sc[iters].stop == 1 ? 0 : sc[iters].stop) : 1;
    if (verified) info->status = 0;  // the true residual met tol (also after a checked stagnation stop)
    return IBIM_OK;
}

}  // namespace

int matvec_op(Context* ctx, int op, const void* p_dev, const void* q_dev, void* y_dev, void* z_dev) {
    Workspace ws;
    int rc = prepare(ctx, 1, &ws);
    if (rc) return rc;
    VecArgs a = with_layout(ws.a, ctx, op);
    rc = sweep_any(ctx, op, (const double2*)p_dev, (const double2*)q_dev, a.ypart, a.zpart, nullptr);
    if (rc) return rc;
    // sum the partial planes with the dot kernel (its dot products are unused)
    a.v = (double2*)y_dev;
    a.w = (double2*)z_dev;
    a.r = (double2*)p_dev;
    a.rt = (double2*)q_dev;
    a.pt = (double2*)q_dev;
    IBIM_HIP(hipMemsetAsync(a.done, 0, sizeof(int), ctx->stream));
    hipLaunchKernelGGL(k_bicg_dot, dim3(ws.nvb), dim3(kVecThreads), 0, ctx->stream, a);
    IBIM_HIP(hipGetLastError());
    return IBIM_OK;
}

namespace {

__global__ void __launch_bounds__(kVecThreads) k_sum_y(VecArgs a) {
    for (int j = blockIdx.x * kVecThreads + threadIdx.x; j < a.count; j += gridDim.x * kVecThreads)
        a.v[j] = sweep_y(a, j);
}

}  // namespace

int matvec_op_y(Context* ctx, int op, const void* p_dev, void* y_dev) {
    Workspace ws;
    int rc = prepare(ctx, 1, &ws);
    if (rc) return rc;
    VecArgs a = with_layout(ws.a, ctx, op);
    rc = sweep_any(ctx, op, (const double2*)p_dev, (const double2*)p_dev, a.ypart, a.zpart, nullptr, false);
    if (rc) return rc;
    a.v = (double2*)y_dev;
    hipLaunchKernelGGL(k_sum_y, dim3(ws.nvb), dim3(kVecThreads), 0, ctx->stream, a);
    IBIM_HIP(hipGetLastError());
    return IBIM_OK;
}

namespace {

__global__ void k_clear_flag(int* flag) { *flag = 0; }

// The layout of a single product's workspace carved from one buffer: partial
// planes, per-block partials of the plane-summing kernel, its stop flag.
struct ScratchLayout {
    size_t planes, part, flag, total;
};

ScratchLayout scratch_layout(const Context* ctx, int op) {
    const int count = ctx->count;
    const size_t ld = ctx->ld;
    const size_t nct = (size_t)(count + kMvCols - 1) / kMvCols, nrt = (size_t)(count + kMvRows - 1) / kMvRows;
    const size_t nplanes = pair_layout_op(ctx, op) ? 2 * (nct + nrt + 1) : nct + nrt;
    ScratchLayout L;
    L.planes = 0;
    L.part = (sizeof(double2) * ld * nplanes + 255) / 256 * 256;
    L.flag = L.part + (sizeof(double) * (size_t)vec_blocks(count) * 8 + 255) / 256 * 256;
    L.total = L.flag + 256;
    return L;
}

}  // namespace

size_t matvec_scratch_bytes(const Context* ctx, int op) { return scratch_layout(ctx, op).total; }

// matvec_op with its workspace in `scratch` (matvec_scratch_bytes(ctx, op)
// bytes of device memory): no allocation, no memset (the flag is cleared by a
// kernel), no timing events.
int matvec_op_scratch(Context* ctx, int op, const void* p_dev, const void* q_dev, void* y_dev, void* z_dev,
                      void* scratch) {
    const ScratchLayout L = scratch_layout(ctx, op);
    char* base = (char*)scratch;
    const int count = ctx->count;
    VecArgs a{};
    a.ld = ctx->ld;
    a.ncol_tiles = (count + kMvCols - 1) / kMvCols;
    a.nrow_tiles = (count + kMvRows - 1) / kMvRows;
    a.ypart = (double2*)(base + L.planes);
    a.part = (double*)(base + L.part);
    a.done = (int*)(base + L.flag);
    a.iter = nullptr;
    a.count = count;
    a = with_layout(a, ctx, op);
    int rc = sweep_any(ctx, op, (const double2*)p_dev, (const double2*)q_dev, a.ypart, a.zpart, nullptr);
    if (rc) return rc;
    a.v = (double2*)y_dev;
    a.w = (double2*)z_dev;
    a.r = (double2*)p_dev;
    a.rt = (double2*)q_dev;
    a.pt = (double2*)q_dev;
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, ctx->stream, a.done);
    hipLaunchKernelGGL(k_bicg_dot, dim3(vec_blocks(count)), dim3(kVecThreads), 0, ctx->stream, a);
    IBIM_HIP(hipGetLastError());
    return IBIM_OK;
}

int matvec_dense(Context* ctx, const void* p_dev, const void* q_dev, void* y_dev, void* z_dev) {
    return matvec_op(ctx, OP_DENSE, p_dev, q_dev, y_dev, z_dev);
}

int bicg_solve_op(Context* ctx, int op, const void* b_dev, void* x_dev, int maxit, double tol, int chunk,
                  ibim_solve_info* info, double* hist, bool adjoint, bool ref_guards) {
    Workspace ws;
    int rc = prepare(ctx, maxit, &ws);
    if (rc) return rc;
    return bicg_device(ctx, op, ws, (const double2*)b_dev, (double2*)x_dev, maxit, tol, chunk, info, hist, adjoint,
                       nullptr, ref_guards);
}

// BiCG on the FP32 operator (the stored complex-FP32 matrix when there is one,
// else FP32 entries recomputed per sweep) with FP64 residual replacement by the
// FP64 operator (stored complex-FP64, else recomputed FP64 entries).
int run_reliable(Context* ctx, const ibim_solve_opts* opts, ibim_solve_info* info, double* hist) {
    const int op = ctx->mat_precision == IBIM_C64 ? OP_DENSE : OP_FREE32;
    const int true_op = fp64_operator(ctx);
    Workspace ws;
    int rc = prepare(ctx, opts->maxit, &ws);
    if (rc) return rc;
    int total = 0, outer = 0;
    ibim_solve_info last{};
    for (int c = 0; c < ctx->nrhs; ++c) {
        const double2* b = ctx->rhs.as<double2>() + (size_t)c * ctx->count;
        double2* x = ctx->density.as<double2>() + (size_t)c * ctx->count;
        Replace rep{true_op, opts->inner_tol > 0 && opts->inner_tol < 1 ? opts->inner_tol : 1e-2, 0, 0.0};
        // replacements happen at host checks: every 2 iterations unless asked
        const int chunk = opts->chunk > 0 ? opts->chunk : 2;
        rc = bicg_device(ctx, op, ws, b, x, opts->maxit, opts->tol, chunk, &last, c == 0 ? hist : nullptr, false,
                         &rep);
        if (rc) return rc;
        total += last.iterations;
        outer += rep.replacements;
    }
    *info = last;
    info->iterations = total;
    info->outer = outer;
    ctx->last_iters = total;
    return IBIM_OK;
}

int run_bicg(Context* ctx, const ibim_solve_opts* opts, ibim_solve_info* info, double* hist) {
    // the stored operator when there is one, else the FP64 entries recomputed per sweep
    const int op = ctx->mat_precision >= 0 ? OP_DENSE : OP_FREE64;
    Workspace ws;
    int rc = prepare(ctx, opts->maxit, &ws);
    if (rc) return rc;
    int total = 0;
    ibim_solve_info last{};
    for (int c = 0; c < ctx->nrhs; ++c) {
        const double2* b = ctx->rhs.as<double2>() + (size_t)c * ctx->count;
        double2* x = ctx->density.as<double2>() + (size_t)c * ctx->count;
        rc = bicg_device(ctx, op, ws, b, x, opts->maxit, opts->tol, opts->chunk, &last, c == 0 ? hist : nullptr);
        if (rc) return rc;
        total += last.iterations;
    }
    *info = last;
    info->iterations = total;
    info->outer = 1;
    ctx->last_iters = total;
    return IBIM_OK;
}

}  // namespace ibim
