// ibim_pair.hip -- the traditional operator stored once per unordered pair
// (IBIM_C128_PAIR / IBIM_C64_PAIR), and its matrix-free forms (OP_FREE64,
// OP_FREE32), which run the same sweeps and 16-column passes with each F
// computed where it would be read instead of loaded.
//
// For rho_ij > tau the traditional entry (large3Dexample2.c:4463-4564) factors
// into a part symmetric in (i, j) and two real point factors:
//   A_ij = F_ij alpha_ij,   A_ji = F_ij beta_ij,   F_ij = sigma (c1 + i c2)(rho_ij)
// with grad_x Phi = (c1 + i c2) r and sigma = -1 for the interior problem, +1
// otherwise. With r_ij = x*_i - x*_j:
//   Dirichlet (normal of the column point): alpha_ij = w_j r_ij.n_j = x_i.m_j - c_j,
//     beta_ij = w_i r_ji.n_i = x_j.m_i - c_i, per point m = w n, c = x.m;
//   Neumann (normal of the row point): alpha_ij = w_j r_ij.n_i = w_j (c_i - x_j.m_i),
//     beta_ij = w_i r_ji.n_j = w_i (c_j - x_i.m_j), per point m = n, c = x.n.
// Only F of the pairs i < j is stored, so a sweep reads half the bytes of the
// dense operator and rebuilds alpha, beta from the points (FP64; the points
// are wave-uniform for the row and in registers for the column). One pass over
// F gives v = A p and w = A^H q:
//   row side    v_i += A_ij p_j,        w_i += conj(A_ji) q_j
//   column side v_j += A_ji p_i,        w_j += conj(A_ij) q_i
// Near-field pairs (rho <= tau, the diagonal included) store F = 0; their
// entries w_j reg_q (+1/2 on the diagonal) come from a per-point list of
// near neighbours (sorted, so the sums have a fixed order) in a small kernel
// that writes one more partial plane.
//
// The combined-field Dirichlet operator (ibim_params.combined, kernel
// kCombined) adds -i eta w_j Phi_ij with Phi = e^{ik rho} / (4 pi rho), also
// symmetric in (i, j). Its pairs store G_ij = Phi(rho_ij) instead of F: with
// u = 1 / rho, F = G (i k u - u^2), so
//   A_ij = G [alpha_ij (i k u - u^2) - i eta w_j],  A_ji = G [beta_ij (i k u - u^2) - i eta w_i],
// one stored value per pair as before, u rebuilt from the points in the sweep.
// Its near entries are w_j (reg_j - i eta S_near) with S_near the disc mean of
// Phi (single_layer_near).
//
// Storage: row blocks of kSweepRows rows; row block I holds tiles t = 0 ..
// T_I - 1 of kSweepRows x kSweepCols values covering columns
// I kSweepRows + t kSweepCols .. + kSweepCols - 1 (the first kSweepRows of them
// form the diagonal block, where F = 0 for j <= i). Tiles are stored one after
// another, each row-major; values outside the operator are zero.
//
// Partial planes (complex FP64, leading dimension ld), for v in ypart and for
// w in zpart alike: ncol = ceil(count / kSweepCols) row-side planes (tile t of
// any row block writes plane t), nrow = ceil(count / kSweepRows) column-side
// planes (tile (I, t) writes plane ncol + I), and the near-field plane
// ncol + nrow. Element j of block J = j / kSweepRows sums row-side planes
// t < T_J, column-side planes I <= J and the near plane, in that order.
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "ibim_entries.hpp"
#include "ibim_internal.hpp"
#include "ibim_mfma.hpp"
#include "ibim_wave.hpp"

namespace ibim {

namespace {

constexpr int kPT = 256;                 // threads per tile
constexpr int kPCpt = kSweepCols / kPT;  // columns per thread
constexpr int kPR = kSweepRows;
constexpr int kPC = kSweepCols;
constexpr int kPairAsmRows = 64;  // rows per assembly workgroup
static_assert(kPCpt == 4, "the column loops below assume 4 columns per thread");
static_assert(kPR % kPairAsmRows == 0, "assembly row chunks must tile a row block");

// Point data of the pair sweep: position, m and c as defined above, and w.
struct alignas(64) PairPoint {
    double x, y, z, mx, my, mz, c, w;
};

template <class Store>
__device__ __forceinline__ void store_pair(Store* dst, double re, double im);

template <>
__device__ __forceinline__ void store_pair<double2>(double2* dst, double re, double im) {
    *dst = make_double2(re, im);
}

template <>
__device__ __forceinline__ void store_pair<float2>(float2* dst, double re, double im) {
    *dst = make_float2((float)re, (float)im);
}

// The near-field decision of trad_entry (the same FP64 expression, so the
// pair operator and the dense one agree on which pairs are near).
__device__ __forceinline__ double pair_rho(const PointData& a, const PointData& b) {
    double d1 = a.sx - b.sx, d2 = a.sy - b.sy, d3 = a.sz - b.sz;
    return sqrt(d1 * d1 + d2 * d2 + d3 * d3);
}

// alpha_ij (factor of A_ij) and beta_ij (factor of A_ji) for row point R = i
// and column point C = j. Precise: from the separation r_ij = x_i - x_j (3 more
// FP64 operations per pair), as the dense entries form r.n; else from the
// per-point c, which cancels for nearly tangential pairs (|x| |m| / |r.m| ulps,
// below the rounding of complex64 factors; 10x the dense operator's rounding
// noise with complex128 factors).
template <bool Dirichlet, bool Precise>
__device__ __forceinline__ void pair_factors(const PairPoint& R, double cx, double cy, double cz, double cmx,
                                             double cmy, double cmz, double cc, double cw, double* alpha,
                                             double* beta) {
    if (Precise) {
        const double d1 = R.x - cx, d2 = R.y - cy, d3 = R.z - cz;
        if (Dirichlet) {
            *alpha = fma(d1, cmx, fma(d2, cmy, d3 * cmz));
            *beta = -fma(d1, R.mx, fma(d2, R.my, d3 * R.mz));
        } else {
            *alpha = cw * fma(d1, R.mx, fma(d2, R.my, d3 * R.mz));
            *beta = R.w * -fma(d1, cmx, fma(d2, cmy, d3 * cmz));
        }
        return;
    }
    if (Dirichlet) {
        *alpha = fma(R.x, cmx, fma(R.y, cmy, fma(R.z, cmz, -cc)));
        *beta = fma(cx, R.mx, fma(cy, R.my, fma(cz, R.mz, -R.c)));
    } else {
        *alpha = cw * fma(-cx, R.mx, fma(-cy, R.my, fma(-cz, R.mz, R.c)));
        *beta = R.w * fma(-R.x, cmx, fma(-R.y, cmy, fma(-R.z, cmz, cc)));
    }
}

__global__ void __launch_bounds__(256) k_pair_points(const double* __restrict__ tube, int count, int dirichlet,
                                                     PairPoint* __restrict__ pts, int npad) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= npad) return;
    PairPoint o{};
    if (p < count) {
        const PointData d = load_point(tube, (size_t)count, p);
        o.x = d.sx;
        o.y = d.sy;
        o.z = d.sz;
        o.w = d.w;
        if (dirichlet) {
            o.mx = d.w * d.nx;
            o.my = d.w * d.ny;
            o.mz = d.w * d.nz;
        } else {
            o.mx = d.nx;
            o.my = d.ny;
            o.mz = d.nz;
        }
        o.c = d.sx * o.mx + d.sy * o.my + d.sz * o.mz;
    }
    pts[p] = o;
}

// F (G for the combined operator) of one tile, kPairAsmRows rows per
// workgroup (and, for the mixed storage, the complex128 copy in F64); counts
// near pairs per point.
template <class Store>
__global__ void __launch_bounds__(kPT) k_pair_assemble(const double* __restrict__ tube, int count,
                                                       const int2* __restrict__ tiles, ibim_params prm,
                                                       Store* __restrict__ F, double2* __restrict__ F64,
                                                       int* __restrict__ ncnt) {
    const int2 tl = tiles[blockIdx.x];
    const int r0 = tl.x * kPR, c0 = r0 + tl.y * kPC;
    const size_t cnt = (size_t)count;
    const double sigma = prm.interior ? -1.0 : 1.0;
    PointData b[kPCpt];
#pragma unroll
    for (int m = 0; m < kPCpt; ++m) {
        const int j = c0 + threadIdx.x + kPT * m;
        b[m] = j < count ? load_point(tube, cnt, j) : PointData{};
    }
    const size_t tb = (size_t)blockIdx.x * kPR * kPC + threadIdx.x;
    Store* Fb = F + tb;
    const int rb = blockIdx.y * kPairAsmRows;
    for (int r = rb; r < rb + kPairAsmRows; ++r) {
        const int i = r0 + r;
        const PointData a = i < count ? load_point(tube, cnt, i) : PointData{};
#pragma unroll
        for (int m = 0; m < kPCpt; ++m) {
            const int j = c0 + threadIdx.x + kPT * m;
            double fr = 0.0, fi = 0.0;
            if (i < j && j < count) {
                const double rho = pair_rho(a, b[m]);
                if (rho > prm.tau) {
                    if (prm.combined) {
                        phi_value(prm.k, rho, &fr, &fi);  // G_ij (exterior only)
                    } else {
                        double c1, c2;
                        grad_phi_coeffs(prm.k, rho, &c1, &c2);
                        fr = sigma * c1;
                        fi = sigma * c2;
                    }
                } else {
                    atomicAdd(&ncnt[i], 1);
                    atomicAdd(&ncnt[j], 1);
                }
            }
            store_pair<Store>(Fb + (size_t)r * kPC + kPT * m, fr, fi);
            if (F64 != nullptr) F64[tb + (size_t)r * kPC + kPT * m] = make_double2(fr, fi);
        }
    }
}

// The near-neighbour lists: j enters i's list and i enters j's for every near
// pair i < j (order fixed afterwards by k_near_sort).
__global__ void __launch_bounds__(kPT) k_near_fill(const double* __restrict__ tube, int count,
                                                   const int2* __restrict__ tiles, double tau,
                                                   const int* __restrict__ off, int* __restrict__ cur,
                                                   int* __restrict__ idx) {
    const int2 tl = tiles[blockIdx.x];
    const int r0 = tl.x * kPR, c0 = r0 + tl.y * kPC;
    const size_t cnt = (size_t)count;
    PointData b[kPCpt];
#pragma unroll
    for (int m = 0; m < kPCpt; ++m) {
        const int j = c0 + threadIdx.x + kPT * m;
        b[m] = j < count ? load_point(tube, cnt, j) : PointData{};
    }
    const int rb = blockIdx.y * kPairAsmRows;
    for (int r = rb; r < rb + kPairAsmRows; ++r) {
        const int i = r0 + r;
        if (i >= count) break;
        const PointData a = load_point(tube, cnt, i);
#pragma unroll
        for (int m = 0; m < kPCpt; ++m) {
            const int j = c0 + threadIdx.x + kPT * m;
            if (i < j && j < count && !(pair_rho(a, b[m]) > tau)) {
                // the counting pass made the same decisions; the bound only
                // keeps the lists in their slots
                const int si = atomicAdd(&cur[i], 1), sj = atomicAdd(&cur[j], 1);
                if (si < off[i + 1] - off[i]) idx[off[i] + si] = j;
                if (sj < off[j + 1] - off[j]) idx[off[j] + sj] = i;
            }
        }
    }
}

// Near pairs per point without an assembly (the matrix-free pair operator):
// k_pair_assemble's count, the same decision on the same FP64 distance.
__global__ void __launch_bounds__(kPT) k_near_count(const double* __restrict__ tube, int count,
                                                    const int2* __restrict__ tiles, double tau,
                                                    int* __restrict__ ncnt) {
    const int2 tl = tiles[blockIdx.x];
    const int r0 = tl.x * kPR, c0 = r0 + tl.y * kPC;
    const size_t cnt = (size_t)count;
    PointData b[kPCpt];
#pragma unroll
    for (int m = 0; m < kPCpt; ++m) {
        const int j = c0 + threadIdx.x + kPT * m;
        b[m] = j < count ? load_point(tube, cnt, j) : PointData{};
    }
    const int rb = blockIdx.y * kPairAsmRows;
    for (int r = rb; r < rb + kPairAsmRows; ++r) {
        const int i = r0 + r;
        if (i >= count) break;
        const PointData a = load_point(tube, cnt, i);
#pragma unroll
        for (int m = 0; m < kPCpt; ++m) {
            const int j = c0 + threadIdx.x + kPT * m;
            if (i < j && j < count && !(pair_rho(a, b[m]) > tau)) {
                atomicAdd(&ncnt[i], 1);
                atomicAdd(&ncnt[j], 1);
            }
        }
    }
}

__global__ void __launch_bounds__(256) k_near_sort(const int* __restrict__ off, int* __restrict__ idx, int count) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    const int b = off[i], e = off[i + 1];
    for (int s = b + 1; s < e; ++s) {
        const int v = idx[s];
        int t = s - 1;
        while (t >= b && idx[t] > v) {
            idx[t + 1] = idx[t];
            --t;
        }
        idx[t + 1] = v;
    }
}

// c1, c2 of grad_phi_coeffs from rho^2 for the mixed-precision operator: 1/rho
// from the FP32 reciprocal square root refined by one FP64 Newton step
// (relative error ~3e-14), rho = rho^2 / rho, the phase kappa rho reduced in
// FP64 (revolutions) and its sine and cosine from the FP32 polynomial (about
// one FP32 ulp); c1, c2 in FP64.
__device__ __forceinline__ void grad_phi_coeffs_fast(double k, double kappa, double rho2, double* c1, double* c2) {
    const double i0 = (double)__builtin_amdgcn_rsqf((float)rho2);
    const double ir = i0 * fma(-0.5 * rho2, i0 * i0, 1.5);
    const double rho = rho2 * ir;
    const double t = rho * kappa;
    float s32, c32;
    sincos_rev_f32(t - rint(t), &s32, &c32);
    const double sv = (double)s32, cv = (double)c32;
    const double den = (1.0 / (4.0 * kPi)) * ir * ir;
    *c1 = -1.0 * (k * sv + cv * ir) * den;
    *c2 = (k * cv - sv * ir) * den;
}

// sqrt(rho2) > tau, the near lists' far-field rule, without the FP64 square
// root unless rho2 lies within a relative 1e-9 of tau^2 (tau2_lo, tau2_hi),
// where the rounding of the square root could decide.
__device__ __forceinline__ bool far_pair(double rho2, double tau, double tau2_lo, double tau2_hi) {
    if (!(tau > 0.0)) return sqrt(rho2) > tau;  // tau <= 0 (or NaN): the rule itself
    if (rho2 > tau2_hi) return true;
    if (rho2 < tau2_lo) return false;
    return sqrt(rho2) > tau;
}

// Which operator the pair kernels build: the traditional kernel with the
// normal of the row point (Neumann) or of the column point (Dirichlet), or the
// combined-field Dirichlet operator.
enum PairKernel { kNeumann = 0, kDirichlet = 1, kCombined = 2 };

int pair_kernel(const Context* ctx) {
    if (ctx->prm.dirichlet == 0) return kNeumann;
    return ctx->prm.combined ? kCombined : kDirichlet;
}

// 1 / rho from rho^2: the FP32 reciprocal square root refined by Steps FP64
// Newton steps (two: within a few ulp, the FP64 operators; one: ~3e-14, the
// mixed ones). rho^2 is floored at 1e-30 so that the zero factors of
// coincident or padding points stay zero in the products.
template <int Steps>
__device__ __forceinline__ double inv_rho(double rho2) {
    const double r2 = fmax(rho2, 1e-30);
    double ir = (double)__builtin_amdgcn_rsqf((float)r2);
#pragma unroll
    for (int s = 0; s < Steps; ++s) ir = ir * fma(-0.5 * r2, ir * ir, 1.5);
    return ir;
}

// G = Phi(rho) for the mixed-precision combined operator (the phase reduced in
// FP64, FP32 sine and cosine, as grad_phi_coeffs_fast), and u = 1 / rho.
__device__ __forceinline__ void phi_fast(double kappa, double rho2, double* gr, double* gi, double* u) {
    const double ir = inv_rho<1>(rho2);
    const double t = (rho2 * ir) * kappa;
    float s32, c32;
    sincos_rev_f32(t - rint(t), &s32, &c32);
    const double sc = ir * (1.0 / (4.0 * kPi));
    *gr = (double)c32 * sc;
    *gi = (double)s32 * sc;
    *u = ir;
}

// The combined-field entries of a pair from G = Phi(rho), u = 1 / rho and the
// point factors: A_ij = G [alpha (i k u - u^2) - i eta w_j] and A_ji likewise
// with beta and w_i.
__device__ __forceinline__ void combined_entries(double gr, double gi, double u, double k, double eta, double al,
                                                 double be, double wi, double wj, double* ar, double* ai, double* br,
                                                 double* bi) {
    const double u2 = u * u, ku = k * u;
    const double xr = -u2 * al, xi = fma(ku, al, -eta * wj);
    const double yr = -u2 * be, yi = fma(ku, be, -eta * wi);
    *ar = fma(gr, xr, -gi * xi);
    *ai = fma(gr, xi, gi * xr);
    *br = fma(gr, yr, -gi * yi);
    *bi = fma(gr, yi, gi * yr);
}

// Where a pair sweep takes its factors F: the stored tiles, or evaluated where
// they would be read -- with k_pair_assemble's operations (the complex128 pair
// storage's bits: the FP64 matrix-free operator) or with
// grad_phi_coeffs_fast (the mixed-precision matrix-free operator); one
// evaluation per pair either way.
// kCachedFactors: the FP64 matrix-free operator with the complex128 pairs of
// its first row blocks kept in HBM (read where cached, computed elsewhere).
enum PairFactors { kStoredFactors = 0, kExactFactors = 1, kFastFactors = 2, kCachedFactors = 3 };

// One tile of the pair sweep (see the file comment for the planes written).
// Matrix-free sweeps are held to 2 waves per SIMD: the FP64 factor's registers
// otherwise leave the Neumann variant at one (14.8 vs 10.1 ms on 64,676 points).
// WithZ = false: v = A p alone (q, the w planes and the A^H products skipped).
template <class MT, int Kern, int Factors = kStoredFactors, bool WithZ = true>
__global__ void __launch_bounds__(kPT, Factors == kStoredFactors ? 1 : 2)
    k_pair_sweep(const MT* __restrict__ F, const int2* __restrict__ tiles,
                                                    const PairPoint* __restrict__ pts, int count, size_t ld,
                                                    int ncol, const double2* __restrict__ p,
                                                    const double2* __restrict__ q, double2* __restrict__ ypart,
                                                    double2* __restrict__ zpart, const int* __restrict__ done,
                                                    ibim_params prm) {
    constexpr bool Computed = Factors != kStoredFactors;
    constexpr bool Dirichlet = Kern != kNeumann, Combined = Kern == kCombined;
    constexpr bool Precise = sizeof(MT) == sizeof(double2);
    static_assert(!Computed || Precise, "computed factors are complex128");
    if (done != nullptr && *done) return;
    __shared__ double sh[4][kPT / 64][kPR];  // (v re, v im, w re, w im) x wave x row
    const int2 tl = tiles[blockIdx.x];
    const int r0 = tl.x * kPR, c0 = r0 + tl.y * kPC;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int my_val = wave_sum8_index(lane);

    double cx[kPCpt], cy[kPCpt], cz[kPCpt], cmx[kPCpt], cmy[kPCpt], cmz[kPCpt], cc[kPCpt], cw[kPCpt];
    double pre[kPCpt], pim[kPCpt], qre[kPCpt], qim[kPCpt];
    double vre[kPCpt], vim[kPCpt], wre[kPCpt], wim[kPCpt];
#pragma unroll
    for (int m = 0; m < kPCpt; ++m) {
        const int j = c0 + t + kPT * m;  // < count + kPC: inside the padded point table
        const PairPoint P = pts[j];
        cx[m] = P.x;
        cy[m] = P.y;
        cz[m] = P.z;
        cmx[m] = P.mx;
        cmy[m] = P.my;
        cmz[m] = P.mz;
        cc[m] = P.c;
        cw[m] = P.w;
        // F = 0 outside the operator, so clamped loads contribute nothing
        const int jc = j < count ? j : count - 1;
        const double2 pv = p[jc], qv = WithZ ? q[jc] : make_double2(0.0, 0.0);
        pre[m] = pv.x;
        pim[m] = pv.y;
        qre[m] = qv.x;
        qim[m] = qv.y;
        vre[m] = vim[m] = wre[m] = wim[m] = 0.0;
    }
    const MT* Fb = Computed ? nullptr : F + (size_t)blockIdx.x * kPR * kPC + t;
    // the next two rows' values are loaded while the current two are applied
    MT nxt[2][kPCpt];
    if (!Computed) {
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int m = 0; m < kPCpt; ++m) nxt[u][m] = load_stream(Fb + (size_t)u * kPC + kPT * m);
    }
    const double sigma = prm.interior ? -1.0 : 1.0;
    const double kappa = prm.k / (2.0 * kPi);
    const double tau2_lo = prm.tau * prm.tau * (1.0 - 1e-9), tau2_hi = prm.tau * prm.tau * (1.0 + 1e-9);
    for (int rr = 0; rr < kPR; rr += 2) {
        MT a[2][kPCpt];
        if (!Computed) {
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int m = 0; m < kPCpt; ++m) a[u][m] = nxt[u][m];
        }
        if (!Computed && rr + 2 < kPR) {
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int m = 0; m < kPCpt; ++m) nxt[u][m] = load_stream(Fb + (size_t)(rr + 2 + u) * kPC + kPT * m);
        }
        double rows[8];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int i = r0 + rr + u;  // < count + kPR: inside the padded point table
            const PairPoint R = pts[i];
            const int ic = i < count ? i : count - 1;
            const double2 pi = p[ic], qi = WithZ ? q[ic] : make_double2(0.0, 0.0);
            double svr = 0.0, svi = 0.0, swr = 0.0, swi = 0.0;
#pragma unroll
            for (int m = 0; m < kPCpt; ++m) {
                double rho2 = 0.0, uinv = 0.0;
                if (Computed || Combined) {
                    const double d1 = R.x - cx[m], d2 = R.y - cy[m], d3 = R.z - cz[m];
                    rho2 = d1 * d1 + d2 * d2 + d3 * d3;
                }
                if (Computed) {
                    // F (G) of pair (i, j), 0 where the storage holds 0 (j <= i,
                    // past count, the near field)
                    const int j = c0 + t + kPT * m;
                    a[u][m].x = 0.0;
                    a[u][m].y = 0.0;
                    if (i < j && j < count) {
                        double c1, c2;
                        if (Factors == kExactFactors) {
                            const double rho = sqrt(rho2);
                            if (rho > prm.tau) {
                                if (Combined) {
                                    phi_value(prm.k, rho, &c1, &c2);
                                    a[u][m].x = c1;
                                    a[u][m].y = c2;
                                } else {
                                    grad_phi_coeffs(prm.k, rho, &c1, &c2);
                                    a[u][m].x = sigma * c1;
                                    a[u][m].y = sigma * c2;
                                }
                            }
                        } else if (far_pair(rho2, prm.tau, tau2_lo, tau2_hi)) {
                            if (Combined) {
                                phi_fast(kappa, rho2, &c1, &c2, &uinv);
                                a[u][m].x = c1;
                                a[u][m].y = c2;
                            } else {
                                grad_phi_coeffs_fast(prm.k, kappa, rho2, &c1, &c2);
                                a[u][m].x = sigma * c1;
                                a[u][m].y = sigma * c2;
                            }
                        }
                    }
                }
                double al, be;
                pair_factors<Dirichlet, Precise>(R, cx[m], cy[m], cz[m], cmx[m], cmy[m], cmz[m], cc[m], cw[m], &al,
                                                 &be);
                const double fr = (double)a[u][m].x, fi = (double)a[u][m].y;
                double ar, ai, br, bi;  // A_ij, A_ji
                if (Combined) {
                    if (Factors != kFastFactors) uinv = inv_rho<Precise ? 2 : 1>(rho2);
                    combined_entries(fr, fi, uinv, prm.k, prm.eta, al, be, R.w, cw[m], &ar, &ai, &br, &bi);
                } else {
                    ar = fr * al;
                    ai = fi * al;
                    br = fr * be;
                    bi = fi * be;
                }
                svr = fma(ar, pre[m], fma(-ai, pim[m], svr));  // A_ij p_j
                svi = fma(ar, pim[m], fma(ai, pre[m], svi));
                vre[m] = fma(br, pi.x, fma(-bi, pi.y, vre[m]));  // A_ji p_i
                vim[m] = fma(br, pi.y, fma(bi, pi.x, vim[m]));
                if (WithZ) {
                    swr = fma(br, qre[m], fma(bi, qim[m], swr));  // conj(A_ji) q_j
                    swi = fma(br, qim[m], fma(-bi, qre[m], swi));
                    wre[m] = fma(ar, qi.x, fma(ai, qi.y, wre[m]));  // conj(A_ij) q_i
                    wim[m] = fma(ar, qi.y, fma(-ai, qi.x, wim[m]));
                }
            }
            rows[4 * u] = svr;
            rows[4 * u + 1] = svi;
            rows[4 * u + 2] = swr;
            rows[4 * u + 3] = swi;
        }
        if constexpr (WithZ) {
            const double tot = wave_sum8_transposed_x(rows, lane);
            if ((lane & 7) == 0) sh[my_val & 3][wave][rr + (my_val >> 2)] = tot;
        } else {
            const double r4[4] = {rows[0], rows[1], rows[4], rows[5]};
            const double tot = wave_sum4_transposed_x(r4);
            if ((lane & 15) == 0) sh[(lane >> 4) & 1][wave][rr + (lane >> 5)] = tot;
        }
    }
    __syncthreads();
    for (int row = t; row < kPR; row += kPT) {
        const int i = r0 + row;
        if (i >= count) break;
        double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < (WithZ ? 4 : 2); ++k)
            s[k] = (sh[k][0][row] + sh[k][1][row]) + (sh[k][2][row] + sh[k][3][row]);
        ypart[(size_t)tl.y * ld + i] = make_double2(s[0], s[1]);
        if (WithZ) zpart[(size_t)tl.y * ld + i] = make_double2(s[2], s[3]);
    }
    const size_t cp = (size_t)(ncol + tl.x) * ld;
#pragma unroll
    for (int m = 0; m < kPCpt; ++m) {
        const int j = c0 + t + kPT * m;
        if (j < count) {
            ypart[cp + j] = make_double2(vre[m], vim[m]);
            if (WithZ) zpart[cp + j] = make_double2(wre[m], wim[m]);
        }
    }
}

// Near-field entries (w_j reg_q, +1/2 on the diagonal; q = j for Dirichlet,
// i for Neumann) applied to p and, conjugate-transposed, to q. The combined
// operator adds w_j (nr + i ni), its single layer's disc mean times -i eta
// (combined_near), to every near entry.
template <int Kern>
__global__ void __launch_bounds__(256) k_pair_near(const double* __restrict__ tube, int count,
                                                   const int* __restrict__ off, const int* __restrict__ idx,
                                                   const double2* __restrict__ p, const double2* __restrict__ q,
                                                   double2* __restrict__ ynear, double2* __restrict__ znear,
                                                   const int* __restrict__ done, double nr, double ni) {
    constexpr bool Dirichlet = Kern != kNeumann, Combined = Kern == kCombined;
    if (done != nullptr && *done) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    const double* W = tube + (size_t)IBIM_F_W * count;
    const double* G = tube + (size_t)IBIM_F_REG * count;
    const double wi = W[i], gi = G[i];
    const double2 pi = p[i], qi = q[i];
    double yr, yi, zr, zi;
    if (Combined) {
        const double ar = wi * (gi + nr) + 0.5, ai = wi * ni;
        yr = ar * pi.x - ai * pi.y;
        yi = ar * pi.y + ai * pi.x;
        zr = ar * qi.x + ai * qi.y;
        zi = ar * qi.y - ai * qi.x;
    } else {
        const double aii = wi * gi + 0.5;
        yr = aii * pi.x;
        yi = aii * pi.y;
        zr = aii * qi.x;
        zi = aii * qi.y;
    }
    const int e = off[i + 1];
#pragma unroll 4
    for (int s = off[i]; s < e; ++s) {
        const int j = idx[s];
        const double wj = W[j], gj = G[j];
        const double2 pj = p[j], qj = q[j];
        if (Combined) {
            const double ar = wj * (gj + nr), ai = wj * ni;  // A_ij
            const double br = wi * (gi + nr), bi = wi * ni;  // A_ji
            yr += ar * pj.x - ai * pj.y;
            yi += ar * pj.y + ai * pj.x;
            zr += br * qj.x + bi * qj.y;
            zi += br * qj.y - bi * qj.x;
        } else {
            const double aij = Dirichlet ? wj * gj : wj * gi;
            const double aji = Dirichlet ? wi * gi : wi * gj;
            yr += aij * pj.x;
            yi += aij * pj.y;
            zr += aji * qj.x;
            zi += aji * qj.y;
        }
    }
    ynear[i] = make_double2(yr, yi);
    znear[i] = make_double2(zr, zi);
}

// Dense rows row0 .. row0 + nrows - 1 of the pair operator (the sweep's
// arithmetic), for inspection and tests.
template <class MT, int Kern>
__global__ void __launch_bounds__(256) k_pair_rows(const MT* __restrict__ F, const int* __restrict__ blk_start,
                                                   const PairPoint* __restrict__ pts, const double* __restrict__ tube,
                                                   int count, ibim_params prm, int row0, int nrows,
                                                   double2* __restrict__ out) {
    constexpr bool Dirichlet = Kern != kNeumann, Combined = Kern == kCombined;
    constexpr bool Precise = sizeof(MT) == sizeof(double2);
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int r = blockIdx.y;
    if (j >= count || r >= nrows) return;
    const int i = row0 + r;
    const size_t cnt = (size_t)count;
    double re, im;
    const PointData a = load_point(tube, cnt, i), b = load_point(tube, cnt, j);
    if (!(pair_rho(a, b) > prm.tau)) {
        if (Combined) {
            double nr, ni;
            combined_near(prm.k, prm.tau, prm.eta, &nr, &ni);
            re = b.w * (b.reg + nr);
            im = b.w * ni;
        } else {
            re = b.w * (Dirichlet ? b.reg : a.reg);
            im = 0.0;
        }
        if (i == j) re += 0.5;
    } else {
        const int lo = i < j ? i : j, hi = i < j ? j : i;
        const int I = lo / kPR;
        const int t = (hi - I * kPR) / kPC;
        const size_t tile = (size_t)blk_start[I] + t;
        const MT f = F[tile * kPR * kPC + (size_t)(lo - I * kPR) * kPC + (hi - I * kPR - t * kPC)];
        const PairPoint L = pts[lo], H = pts[hi];
        double al, be;
        pair_factors<Dirichlet, Precise>(L, H.x, H.y, H.z, H.mx, H.my, H.mz, H.c, H.w, &al, &be);
        if (Combined) {
            const double d1 = L.x - H.x, d2 = L.y - H.y, d3 = L.z - H.z;
            const double u = inv_rho<Precise ? 2 : 1>(d1 * d1 + d2 * d2 + d3 * d3);
            double ar, ai, br, bi;
            combined_entries((double)f.x, (double)f.y, u, prm.k, prm.eta, al, be, L.w, H.w, &ar, &ai, &br, &bi);
            re = i < j ? ar : br;
            im = i < j ? ai : bi;
        } else {
            const double fac = i < j ? al : be;
            re = (double)f.x * fac;
            im = (double)f.y * fac;
        }
    }
    out[(size_t)r * count + j] = make_double2(re, im);
}

// ------------------------------------------------ 16 right-hand sides a pass --
// Y = A P and Z = A^H Q for 16 RHS-interleaved vectors (element (j, r) at
// j * 16 + r) in one pass over the pair storage, on the FP64 matrix cores. A
// workgroup owns 64 rows of one row block (16 per wave) and walks every column
// once: j < its rows from the column strips of the tiles (J, t_J), J <= I, that
// hold the pairs (j, i), then j >= its rows from the row block's own tiles.
// Each F feeds A_ij = F a(i, j) into Y and conj(A_ji) = conj(F) a(j, i) into
// Z, with a(i, j), a(j, i) = pair_factors(i, j) -- the same FP64 operations
// the single-vector sweep performs for that pair, so the entries are the same
// bits. The rows are reduced inside the wave, so no partial planes are
// written; both products read the pairs twice per pass (N^2 values), the bytes
// of one dense pass for the matrix-core work of two.
// Waves rows-of-16 per workgroup (1, 2 or 4) is the tuned shape: more, smaller
// workgroups fill the GPU on small operators; larger ones share each staged
// chunk among more waves. Every wave owns its 16 rows and walks the columns in
// the same order, so all shapes give the same bits.
constexpr int kBatchK = 64;  // columns staged per step
static_assert(kPR % 64 == 0 && kPC % 64 == 0, "a workgroup's rows share their tiles");

// Factors other than kStoredFactors: F unused; every visit evaluates F of the
// pair it would have read (kExactFactors: the complex128 pair storage's bits,
// the FP64 matrix-free operator; kFastFactors: the mixed-precision one).
// The rows rowW .. rowW + 16 Waves - 1 of the pass (one workgroup's share; the
// body of k_pair_batch and of the grouped persistent pass).
template <class MT, int Kern, bool WithY, bool WithZ, int Waves, int Factors>
__device__ __forceinline__ void pair_batch_rows(const MT* __restrict__ F, const int* __restrict__ blk_start,
                                                const PairPoint* __restrict__ pts, int count, const ibim_params& prm,
                                                const double2* __restrict__ P, const double2* __restrict__ Q,
                                                double2* __restrict__ Y, double2* __restrict__ Z, int rowW,
                                                int cached = 0) {
    constexpr bool Computed = Factors != kStoredFactors;
    constexpr bool Exact = Factors == kExactFactors || Factors == kCachedFactors;
    constexpr bool Dirichlet = Kern != kNeumann, Combined = Kern == kCombined;
    static_assert(!Computed || sizeof(MT) == sizeof(double2), "computed factors are complex128");
    constexpr bool Precise = sizeof(MT) == sizeof(double2);
    constexpr int R = kBatchRHS;
    constexpr int kThreads = 64 * Waves;
    __shared__ double2 ps[WithY ? kBatchK : 1][R];
    __shared__ double2 qs[WithZ ? kBatchK : 1][R];
    __shared__ PairPoint sp[kBatchK];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    // rowW < count
    const int I = rowW / kPR;
    const int i0 = rowW + 16 * w;
    const int i = i0 + li;  // < count + kPC: inside the padded point table
    const PairPoint Ri = pts[i];
    Acc3 y, z;

    // P, Q rows j0 .. j0 + 63 (zero from count on) and their points
    auto stage = [&](int j0) {
        __syncthreads();
        for (int e = threadIdx.x; e < kBatchK * R; e += kThreads) {
            const int j = j0 + e / R;
            if (WithY) ps[e / R][e % R] = j < count ? P[(size_t)j0 * R + e] : make_double2(0.0, 0.0);
            if (WithZ) qs[e / R][e % R] = j < count ? Q[(size_t)j0 * R + e] : make_double2(0.0, 0.0);
        }
        if (threadIdx.x < kBatchK) sp[threadIdx.x] = pts[j0 + threadIdx.x];  // j0 + 63 < count + kPC
        __syncthreads();
    };
    // F of pair (i, j) as k_pair_assemble stores it (0 when not `keep`: the
    // visit's zero pattern, the near field and everything past count)
    const double sigma = prm.interior ? -1.0 : 1.0;
    const double kappa = prm.k / (2.0 * kPi);
    const double tau2_lo = prm.tau * prm.tau * (1.0 - 1e-9), tau2_hi = prm.tau * prm.tau * (1.0 + 1e-9);
    // u: 1 / rho of the pair where the mixed combined factor computed it
    auto factor = [&](int jj, bool keep, double* u) -> MT {
        MT f;
        f.x = 0.0;
        f.y = 0.0;
        *u = 0.0;
        if (keep) {
            const PairPoint& C = sp[jj];
            const double d1 = Ri.x - C.x, d2 = Ri.y - C.y, d3 = Ri.z - C.z;
            const double rho2 = d1 * d1 + d2 * d2 + d3 * d3;
            double c1, c2;
            if (Exact) {
                const double rho = sqrt(rho2);
                if (rho > prm.tau) {
                    if (Combined) {
                        phi_value(prm.k, rho, &c1, &c2);
                        f.x = c1;
                        f.y = c2;
                    } else {
                        grad_phi_coeffs(prm.k, rho, &c1, &c2);
                        f.x = sigma * c1;
                        f.y = sigma * c2;
                    }
                }
            } else if (far_pair(rho2, prm.tau, tau2_lo, tau2_hi)) {
                if (Combined) {
                    phi_fast(kappa, rho2, &c1, &c2, u);
                    f.x = c1;
                    f.y = c2;
                } else {
                    grad_phi_coeffs_fast(prm.k, kappa, rho2, &c1, &c2);
                    f.x = sigma * c1;
                    f.y = sigma * c2;
                }
            }
        }
        return f;
    };
    // one k-step: this lane's entry F (row i, column j = staged index jj)
    auto apply = [&](const MT f, double u, int jj) {
        const PairPoint& C = sp[jj];
        double al, be;
        pair_factors<Dirichlet, Precise>(Ri, C.x, C.y, C.z, C.mx, C.my, C.mz, C.c, C.w, &al, &be);
        const double fr = (double)f.x, fi = (double)f.y;
        double ar, ai, br, bi;  // A_ij, A_ji
        if (Combined) {
            if (Factors != kFastFactors) {
                const double d1 = Ri.x - C.x, d2 = Ri.y - C.y, d3 = Ri.z - C.z;
                u = inv_rho<Precise ? 2 : 1>(d1 * d1 + d2 * d2 + d3 * d3);
            }
            combined_entries(fr, fi, u, prm.k, prm.eta, al, be, Ri.w, C.w, &ar, &ai, &br, &bi);
        } else {
            ar = fr * al;
            ai = fi * al;
            br = fr * be;
            bi = fi * be;
        }
        if (WithY) {
            const double2 p = ps[jj][li];
            y.add(ar, ai, p.x, p.y);  // A_ij p_j
        }
        if (WithZ) {
            const double2 q = qs[jj][li];
            z.add(br, -bi, q.x, q.y);  // conj(A_ji) q_j
        }
    };

    // j < rows: pairs (j, i) in the column strip of tile (J, t_J); for J = I
    // the strip's rows j >= i hold zeros (the diagonal block's lower half)
    for (int J = 0; J <= I; ++J) {
        const int cJ = rowW - J * kPR;  // column of rowW relative to row block J
        const int t = cJ / kPC;
        const bool load = !Computed || (Factors == kCachedFactors && J < cached);  // wave-uniform
        const MT* T = !load ? nullptr
                            : F + ((size_t)blk_start[J] + t) * kPR * kPC + (size_t)(i0 - J * kPR - t * kPC) + li;
        for (int rj0 = 0; rj0 < kPR; rj0 += kBatchK) {
            stage(J * kPR + rj0);
#pragma unroll 8
            for (int s = 0; s < kBatchK / 4; ++s) {
                const int jj = 4 * s + lk;
                if (!load) {
                    double u;
                    const MT f = factor(jj, J * kPR + rj0 + jj < i && i < count, &u);
                    apply(f, u, jj);
                } else {
                    apply(T[(size_t)(rj0 + jj) * kPC], 0.0, jj);
                }
            }
        }
    }
    // j >= rows: row i of the row block's tiles
    const int nt = blk_start[I + 1] - blk_start[I];
    const bool load_own = !Computed || (Factors == kCachedFactors && I < cached);
    const MT* Trow = !load_own ? nullptr : F + (size_t)blk_start[I] * kPR * kPC + (size_t)(i - I * kPR) * kPC;
    for (int t = 0; t < nt; ++t) {
        const int c0 = I * kPR + t * kPC;
        const MT* Tr = !load_own ? nullptr : Trow + (size_t)t * kPR * kPC;
        for (int cb = 0; cb < kPC; cb += kBatchK) {
            stage(c0 + cb);
#pragma unroll 8
            for (int s = 0; s < kBatchK / 4; ++s) {
                const int jj = 4 * s + lk;
                if (!load_own) {
                    const int j = c0 + cb + jj;
                    double u;
                    const MT f = factor(jj, i < j && j < count, &u);
                    apply(f, u, jj);
                } else {
                    apply(Tr[cb + jj], 0.0, jj);
                }
            }
        }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int r = i0 + 4 * v + lk;
        if (r < count) {
            if (WithY) Y[(size_t)r * R + li] = make_double2(y.re(v), y.im(v));
            if (WithZ) Z[(size_t)r * R + li] = make_double2(z.re(v), z.im(v));
        }
    }
}

template <class MT, int Kern, bool WithY, bool WithZ, int Waves, int Factors>
__global__ void __launch_bounds__(64 * Waves) k_pair_batch(const MT* __restrict__ F, const int* __restrict__ blk_start,
                                                           const PairPoint* __restrict__ pts, int count,
                                                           ibim_params prm, const double2* __restrict__ P,
                                                           const double2* __restrict__ Q, double2* __restrict__ Y,
                                                           double2* __restrict__ Z, const int* __restrict__ stop,
                                                           int cached) {
    if (stop != nullptr && *stop) return;
    pair_batch_rows<MT, Kern, WithY, WithZ, Waves, Factors>(F, blk_start, pts, count, prm, P, Q, Y, Z,
                                                            blockIdx.x * 16 * Waves, cached);
}

// Near-field entries of the batched products (one thread per (row, RHS)),
// added after the far part in list order; the combined operator's single-layer
// term (nr, ni) as in k_pair_near.
template <int Kern, bool WithY, bool WithZ>
__global__ void __launch_bounds__(256) k_pair_near_batch(const double* __restrict__ tube, int count,
                                                         const int* __restrict__ off, const int* __restrict__ idx,
                                                         const double2* __restrict__ P, const double2* __restrict__ Q,
                                                         double2* __restrict__ Y, double2* __restrict__ Z,
                                                         const int* __restrict__ stop, double nr, double ni) {
    constexpr bool Dirichlet = Kern != kNeumann, Combined = Kern == kCombined;
    constexpr int R = kBatchRHS;
    if (stop != nullptr && *stop) return;
    const int g = blockIdx.x * 256 + threadIdx.x;
    const int i = g / R, r = g % R;
    if (i >= count) return;
    const double* W = tube + (size_t)IBIM_F_W * count;
    const double* G = tube + (size_t)IBIM_F_REG * count;
    const double wi = W[i], gi = G[i];
    // the diagonal entry (re, im)
    const double aii = Combined ? wi * (gi + nr) + 0.5 : wi * gi + 0.5;
    const double aiii = Combined ? wi * ni : 0.0;
    double2 yv = make_double2(0.0, 0.0), zv = make_double2(0.0, 0.0);
    if (WithY) {
        const double2 p = P[(size_t)i * R + r];
        yv = Combined ? make_double2(aii * p.x - aiii * p.y, aii * p.y + aiii * p.x)
                      : make_double2(aii * p.x, aii * p.y);
    }
    if (WithZ) {
        const double2 q = Q[(size_t)i * R + r];
        zv = Combined ? make_double2(aii * q.x + aiii * q.y, aii * q.y - aiii * q.x)
                      : make_double2(aii * q.x, aii * q.y);
    }
    const int e = off[i + 1];
    for (int s = off[i]; s < e; ++s) {
        const int j = idx[s];
        const double wj = W[j], gj = G[j];
        if (WithY) {
            const double2 p = P[(size_t)j * R + r];
            if (Combined) {
                const double ar = wj * (gj + nr), ai = wj * ni;
                yv.x += ar * p.x - ai * p.y;
                yv.y += ar * p.y + ai * p.x;
            } else {
                const double aij = Dirichlet ? wj * gj : wj * gi;
                yv.x += aij * p.x;
                yv.y += aij * p.y;
            }
        }
        if (WithZ) {
            const double2 q = Q[(size_t)j * R + r];
            if (Combined) {
                const double br = wi * (gi + nr), bi = wi * ni;
                zv.x += br * q.x + bi * q.y;
                zv.y += br * q.y - bi * q.x;
            } else {
                const double aji = Dirichlet ? wi * gi : wi * gj;
                zv.x += aji * q.x;
                zv.y += aji * q.y;
            }
        }
    }
    if (WithY) {
        double2& o = Y[(size_t)i * R + r];
        o = make_double2(o.x + yv.x, o.y + yv.y);
    }
    if (WithZ) {
        double2& o = Z[(size_t)i * R + r];
        o = make_double2(o.x + zv.x, o.y + zv.y);
    }
}

template <class MT, int Kern, bool WithY, bool WithZ, int Factors>
void launch_pair_batch_far(Context* ctx, int waves, const MT* F, const double2* P, const double2* Q, double2* Y,
                           double2* Z, const int* stop) {
    const int count = ctx->count;
    const int rows = 16 * waves;
    const dim3 grid((unsigned)((count + rows - 1) / rows)), block((unsigned)(64 * waves));
    const int* blk = ctx->pair_blk.as<int>();
    const PairPoint* pts = ctx->pair_pts.as<PairPoint>();
    const int cached = Factors == kCachedFactors ? ctx->free_cached_blocks : 0;
    if (waves == 1)
        hipLaunchKernelGGL((k_pair_batch<MT, Kern, WithY, WithZ, 1, Factors>), grid, block, 0, ctx->stream, F,
                           blk, pts, count, ctx->prm, P, Q, Y, Z, stop, cached);
    else if (waves == 2)
        hipLaunchKernelGGL((k_pair_batch<MT, Kern, WithY, WithZ, 2, Factors>), grid, block, 0, ctx->stream, F,
                           blk, pts, count, ctx->prm, P, Q, Y, Z, stop, cached);
    else
        hipLaunchKernelGGL((k_pair_batch<MT, Kern, WithY, WithZ, 4, Factors>), grid, block, 0, ctx->stream, F,
                           blk, pts, count, ctx->prm, P, Q, Y, Z, stop, cached);
}

// The workgroup shape of the far pass: IBIM_BATCH_WAVES (1, 2, 4) if set, else
// the fastest of the three timed on this call's own operands (once per
// context, storage and product kind; the outputs are the same bits).
template <class MT, int Kern, bool WithY, bool WithZ, int Factors>
int tuned_waves(Context* ctx, int slot, const MT* F, const double2* P, const double2* Q, double2* Y, double2* Z,
                const int* stop) {
    if (const char* env = getenv("IBIM_BATCH_WAVES")) {
        const int w = atoi(env);
        if (w == 1 || w == 2 || w == 4) return w;
    }
    if (ctx->batch_waves[slot]) return ctx->batch_waves[slot];
    if (ctx->batch_untimed) return 4;
    // Operators of 2^17 points and more (the 256^3 surfaces: 0.3-0.5 s per
    // pass) take 64 rows per workgroup without timing: that shape won at
    // every size measured alone, and each timed shape costs a whole pass of
    // the caller's solve.
    if (ctx->count >= (1 << 17)) return ctx->batch_waves[slot] = 4;
    // no timing inside a stream capture (events would be captured): any shape
    // gives the same bits
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(ctx->stream, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) return 4;
    // two interleaved rounds, each shape's faster time: other streams sharing
    // the GPU disturb single timings
    const int cand[3] = {4, 2, 1};
    float best_of[3] = {-1.0f, -1.0f, -1.0f};
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return 4;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return 4;
    }
    // Passes of more than kLongPass ms (256^3 operators: 0.3-0.9 s each) are
    // timed once, and the 16-row shape only if 32 rows beat 64: narrower
    // shapes were slower at every size measured alone, and the timing is part
    // of the caller's solve.
    constexpr float kLongPass = 50.0f;
    bool long_pass = false;
    for (int round = 0; round < 2 && !long_pass; ++round)
        for (int c = 0; c < 3; ++c) {
            if (long_pass && c == 2 && best_of[1] >= best_of[0]) break;
            (void)hipEventRecord(e0, ctx->stream);
            launch_pair_batch_far<MT, Kern, WithY, WithZ, Factors>(ctx, cand[c], F, P, Q, Y, Z, stop);
            (void)hipEventRecord(e1, ctx->stream);
            float ms = 0.0f;
            if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess &&
                (best_of[c] < 0.0f || ms < best_of[c]))
                best_of[c] = ms;
            if (c == 0 && best_of[0] > kLongPass) long_pass = true;
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int pick = 0;
    for (int c = 0; c < 3; ++c)
        if (best_of[c] >= 0.0f && (pick == 0 || best_of[c] < best_of[pick - 1])) pick = c + 1;
    if (!pick) return 4;
    // a pass stopped by the flag times nothing: keep it untuned
    if (best_of[pick - 1] > 0.02f) ctx->batch_waves[slot] = cand[pick - 1];
    return cand[pick - 1];
}

template <class MT, int Kern, bool WithY, bool WithZ, int Factors>
void launch_pair_batch(Context* ctx, int slot, const MT* F, const double2* P, const double2* Q, double2* Y,
                       double2* Z, const int* stop) {
    const int count = ctx->count;
    const int waves = tuned_waves<MT, Kern, WithY, WithZ, Factors>(ctx, slot, F, P, Q, Y, Z, stop);
    TimedScope ts(ctx, Factors == kExactFactors || Factors == kCachedFactors ? "batch_free64"
                       : Factors == kFastFactors ? "batch_free32"
                       : slot >= 6               ? "batch_pair64"
                                                 : "batch_pair");
    launch_pair_batch_far<MT, Kern, WithY, WithZ, Factors>(ctx, waves, F, P, Q, Y, Z, stop);
    double nr = 0.0, ni = 0.0;
    if (Kern == kCombined) combined_near(ctx->prm.k, ctx->prm.tau, ctx->prm.eta, &nr, &ni);
    hipLaunchKernelGGL((k_pair_near_batch<Kern, WithY, WithZ>),
                       dim3((unsigned)(((size_t)count * kBatchRHS + 255) / 256)), dim3(256), 0, ctx->stream,
                       ctx->tube.as<double>(), count, ctx->near_off.as<int>(), ctx->near_idx.as<int>(), P, Q, Y, Z,
                       stop, nr, ni);
}

// slot: factors (0 complex64, 3 complex128, 6 complex128 copy, 9 exact and 12
// fast computed) + product kind
template <class MT, int Kern, int Factors = kStoredFactors>
void launch_pair_batch_any(Context* ctx, int slot0, const MT* F, const double2* P, const double2* Q, double2* Y,
                           double2* Z, const int* stop) {
    if (Y && Z)
        launch_pair_batch<MT, Kern, true, true, Factors>(ctx, slot0, F, P, Q, Y, Z, stop);
    else if (Y)
        launch_pair_batch<MT, Kern, true, false, Factors>(ctx, slot0 + 1, F, P, Q, Y, Z, stop);
    else
        launch_pair_batch<MT, Kern, false, true, Factors>(ctx, slot0 + 2, F, P, Q, Y, Z, stop);
}

// ------------------------------------------- the same pass on FP32 matrix cores --
// Complex64 pair storage only (the mixed solvers' inner operator): the walk of
// k_pair_batch, with the entries formed in FP32 -- the separation r = x_i - x_j
// in FP64, rounded once; alpha, beta, rho^2, 1 / rho (the FP32 reciprocal square
// root) and the combined-field entries in FP32 -- and multiplied with FP32 copies
// of P and Q on v_mfma_f32_16x16x4f32 (twice the FP64 rate), the products summed
// in FP32 over each staged chunk of kBatchK columns and the chunk sums in FP64.
// Entry errors stay at the level of the complex64 factors (relative to the
// largest entries); the products add FP32 rounding that depends on the vector,
// which the mixed solvers' FP64 residuals absorb.
typedef float f4 __attribute__((ext_vector_type(4)));
typedef float fv2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ f4 mfma32(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// D = A B + C for a 16x4 A, 4x16 B: lane l holds A[l % 16][l / 16],
// B[l / 16][l % 16] and D[4 (l / 16) + v][l % 16] in component v. Complex
// products take Gauss's three real products, as Acc3 does in FP64.
// Raw buffer loads (vector offset + scalar offset + immediate) from a
// wave-uniform base: one address register per lane for a whole chunk, and
// reads at or past `bytes` return zero.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_of(const void* base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
// Aux is the cache policy: 0 for data other waves read again (P, Q, the
// points), kStream (nt) for the pair values, read once per visit like the
// single-vector sweeps' streaming loads.
constexpr int kStream = 2;
template <int Aux = 0>
__device__ __forceinline__ fv2 ld8(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(fv2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, Aux));
}
template <int Aux = 0>
__device__ __forceinline__ f4 ld16(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, Aux));
}

// The walk of k_pair_batch in chunks of kBatchK columns, software-pipelined:
// while a chunk's 16 k-steps run on the matrix cores, the next chunk's pair
// values are in flight into registers and its P, Q and column points into
// registers for the other half of a double-buffered LDS stage (one barrier per
// chunk); inside a chunk each k-step's LDS operands are read one k-step ahead.
// Lane (l % 16, l / 16) takes chunk column 8 (s / 2) + 2 (l / 16) + s % 2 at
// k-step s: along a column strip 16 lanes read 128 contiguous bytes, and
// along a row block's own tiles each 16-byte load reads 64 contiguous bytes
// of each of 16 rows (the four lanes of a row side by side; two consecutive
// loads fill the same 128-byte lines). P and Q
// are staged as FP32 Gauss triples (pr, pi - pr, pr + pi) and (qr, qi - qr,
// -(qr + qi)), so a k-step spends no vector instruction on the vector side of
// the three real products.
// Rows rowW .. rowW + 63 (one 256-thread workgroup's share; the body of
// k_pair_batch32 and of the grouped persistent pass).
template <int Kern, bool WithY, bool WithZ>
__device__ __forceinline__ void pair_batch32_rows(const float2* __restrict__ F, const int* __restrict__ blk_start,
                                                  const PairPoint* __restrict__ pts, int count,
                                                  const ibim_params& prm, const double2* __restrict__ P,
                                                  const double2* __restrict__ Q, double2* __restrict__ Y,
                                                  double2* __restrict__ Z, int rowW) {
    constexpr bool Dirichlet = Kern != kNeumann, Combined = Kern == kCombined;
    constexpr int R = kBatchRHS, KC = kBatchK, KS = kBatchK / 4;
    constexpr int kStrip = kPR / KC, kOwn = kPC / KC;  // chunks per strip of a block, per own tile
    static_assert(R == 16 && KS == 16 && kStrip == 8 && kOwn == 16, "chunk indexing below");
    static_assert(KC * R == 4 * 256 && KC * 8 == 2 * 256, "staging: 4 vector and 2 point values per thread");
    __shared__ f4 ps[WithY ? 2 : 1][WithY ? KC : 1][R];
    __shared__ f4 qs[WithZ ? 2 : 1][WithZ ? KC : 1][R];
    __shared__ double spd[2][KC][8];  // column points: x, y, z as staged (the rest unused)
    __shared__ f4 spf[2][KC][2];      // (mx, my, mz, s) of each column, then unused slots
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int li = lane & 15, lk = lane >> 4;
    // rowW < count
    const int I = rowW / kPR;
    const int i0 = rowW + 16 * w;
    const int i = i0 + li;  // < count + kPC: inside the padded point table
    const PairPoint Ri = pts[i];
    const float rmx = (float)Ri.mx, rmy = (float)Ri.my, rmz = (float)Ri.mz;
    const float rs = Combined ? (float)(prm.eta * Ri.w) : (float)Ri.w;
    const float k = (float)prm.k;
    const int n1 = (I + 1) * kStrip;  // column-strip chunks (j < rows), then own-tile chunks
    const int nq = n1 + (blk_start[I + 1] - blk_start[I]) * kOwn;
    // lane offsets (bytes) of the two chunk kinds inside a chunk's F region
    const int voff_strip = (2 * lk * kPC + li) * 8, voff_own = (li * kPC + 2 * lk) * 8;
    // the chunk column of this lane's k-step s
    auto col_of = [&](int s) { return 8 * (s >> 1) + 2 * lk + (s & 1); };

    // chunk q: its first column and this wave's offset into F; true for a
    // column strip (F(j, i) down a tile column), false for an own tile (F(i, j))
    auto where = [&](int q, size_t* off, int* j0) -> bool {
        if (q < n1) {
            const int J = q / kStrip, rj0 = (q % kStrip) * KC;
            const int t = (rowW - J * kPR) / kPC;
            *off = ((size_t)blk_start[J] + t) * kPR * kPC + (size_t)rj0 * kPC + (size_t)(i0 - J * kPR - t * kPC);
            *j0 = J * kPR + rj0;
            return true;
        }
        const int qq = q - n1, t = qq / kOwn, cb = (qq % kOwn) * KC;
        *off = ((size_t)blk_start[I] + t) * kPR * kPC + (size_t)(i0 - I * kPR) * kPC + cb;
        *j0 = I * kPR + t * kPC + cb;
        return false;
    };
    // this lane's 16 pair values of chunk q (the tile positions k_pair_batch
    // reads, permuted inside the chunk); a chunk's region spans < 1 MB of F
    auto load = [&](int q, fv2 (&f)[KS]) {
        size_t off;
        int j0;
        const bool strip = where(q, &off, &j0);
        const __amdgpu_buffer_rsrc_t r = buffer_of(F + off, 1u << 20);
        if (strip) {
#pragma unroll
            for (int s = 0; s < KS; ++s) f[s] = ld8<kStream>(r, voff_strip, (col_of(s) - 2 * lk) * kPC * 8);
        } else {
#pragma unroll
            for (int s = 0; s < KS / 2; ++s) {
                const f4 v = ld16<kStream>(r, voff_own + 64 * s, 0);
                f[2 * s] = fv2{v.x, v.y};
                f[2 * s + 1] = fv2{v.z, v.w};
            }
        }
    };
    // staging of chunk q through registers: P, Q rows j0 .. j0 + 63 (zero from
    // count on, by the buffer's size) and the 8 doubles of each column's PairPoint
    f4 pv[4], qv[4];
    fv2 pt[2];
    auto fetch = [&](int q) {
        size_t off;
        int j0;
        (void)where(q, &off, &j0);
        const unsigned bytes = j0 < count ? (unsigned)(count - j0) * R * 16u : 0u;
        const __amdgpu_buffer_rsrc_t rp = buffer_of(P + (size_t)j0 * R, WithY ? bytes : 0u);
        const __amdgpu_buffer_rsrc_t rq = buffer_of(Q + (size_t)j0 * R, WithZ ? bytes : 0u);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (WithY) pv[u] = ld16(rp, tid * 16, u * 4096);
            if (WithZ) qv[u] = ld16(rq, tid * 16, u * 4096);
        }
        const __amdgpu_buffer_rsrc_t rt = buffer_of(pts + j0, KC * sizeof(PairPoint));  // j0 + 63 < count + kPC
#pragma unroll
        for (int u = 0; u < 2; ++u) pt[u] = ld8(rt, tid * 8, u * 2048);
    };
    auto put = [&](int b) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u;
            if (WithY) {
                const f4 v = pv[u];
                const float pr = (float)__builtin_bit_cast(double, fv2{v.x, v.y});
                const float pi = (float)__builtin_bit_cast(double, fv2{v.z, v.w});
                ps[b][e / R][e % R] = f4{pr, pi - pr, pr + pi, 0.0f};
            }
            if (WithZ) {
                const f4 v = qv[u];
                const float qr = (float)__builtin_bit_cast(double, fv2{v.x, v.y});
                const float qi = (float)__builtin_bit_cast(double, fv2{v.z, v.w});
                qs[b][e / R][e % R] = f4{qr, qi - qr, -(qr + qi), 0.0f};
            }
        }
        // PairPoint field order x, y, z, mx, my, mz, c, w: every field as a
        // double, and as a float in slot (mx, my, mz, s, c, x, y, z)
        const int fld = tid & 7;
        const int slot = fld == 7 ? 3 : fld == 6 ? 4 : fld >= 3 ? fld - 3 : fld + 5;
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int c = (tid + 256 * u) >> 3;
            const double v = __builtin_bit_cast(double, pt[u]);
            spd[b][c][fld] = v;
            reinterpret_cast<float*>(spf[b][c])[slot] = (float)(Combined && fld == 7 ? prm.eta * v : v);
        }
    };

    double yre[4] = {0, 0, 0, 0}, yim[4] = {0, 0, 0, 0}, zre[4] = {0, 0, 0, 0}, zim[4] = {0, 0, 0, 0};
    // the LDS operands of one k-step: the column's position and factors, its P
    // and Q triples
    struct Col {
        double x, y, z;
        f4 cf, p, q;
    };
    auto read_col = [&](int b, int jj) {
        Col c;
        c.x = spd[b][jj][0];
        c.y = spd[b][jj][1];
        c.z = spd[b][jj][2];
        c.cf = spf[b][jj][0];
        if (WithY) c.p = ps[b][jj][li];
        if (WithZ) c.q = qs[b][jj][li];
        return c;
    };
    // the 16 k-steps of a staged chunk, FP32 sums, then into the FP64 totals
    auto compute = [&](const fv2 (&f)[KS], int b) {
        f4 y1 = {0, 0, 0, 0}, y2 = y1, y3 = y1, z1 = y1, z2 = y1, z3 = y1;
        Col nx = read_col(b, col_of(0));
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const Col c = nx;
            if (s + 1 < KS) nx = read_col(b, col_of(s + 1));
            // the next k-step's LDS reads stay ahead of this one's work (the
            // scheduler would otherwise sink them next to their use and wait
            // out the LDS latency every k-step)
            __builtin_amdgcn_sched_barrier(0);
            const float d1 = (float)(Ri.x - c.x), d2 = (float)(Ri.y - c.y), d3 = (float)(Ri.z - c.z);
            float al, be;  // alpha_ij, beta_ij from the separation
            if (Dirichlet) {
                al = fmaf(d1, c.cf.x, fmaf(d2, c.cf.y, d3 * c.cf.z));
                be = -fmaf(d1, rmx, fmaf(d2, rmy, d3 * rmz));
            } else {
                al = c.cf.w * fmaf(d1, rmx, fmaf(d2, rmy, d3 * rmz));
                be = rs * -fmaf(d1, c.cf.x, fmaf(d2, c.cf.y, d3 * c.cf.z));
            }
            const float fx = f[s].x, fy = f[s].y;
            float ar, ai, br, bi;  // A_ij, A_ji
            if (Combined) {
                const float r2 = fmaxf(fmaf(d1, d1, fmaf(d2, d2, d3 * d3)), 1e-30f);
                const float u = __builtin_amdgcn_rsqf(r2);
                const float nu2 = -u * u, ku = k * u;
                const float xr = nu2 * al, xi = fmaf(ku, al, -c.cf.w);
                const float yr = nu2 * be, yi = fmaf(ku, be, -rs);
                ar = fmaf(fx, xr, -fy * xi);
                ai = fmaf(fx, xi, fy * xr);
                br = fmaf(fx, yr, -fy * yi);
                bi = fmaf(fx, yi, fy * yr);
            } else {
                ar = fx * al;
                ai = fy * al;
                br = fx * be;
                bi = fy * be;
            }
            if (WithY) {  // A_ij p_j
                y1 = mfma32(ar + ai, c.p.x, y1);
                y2 = mfma32(ar, c.p.y, y2);
                y3 = mfma32(ai, c.p.z, y3);
            }
            if (WithZ) {  // conj(A_ji) q_j = (br - i bi) q_j
                z1 = mfma32(br - bi, c.q.x, z1);
                z2 = mfma32(br, c.q.y, z2);
                z3 = mfma32(bi, c.q.z, z3);
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            if (WithY) {
                yre[v] += (double)y1[v] - (double)y3[v];
                yim[v] += (double)y1[v] + (double)y2[v];
            }
            if (WithZ) {
                zre[v] += (double)z1[v] - (double)z3[v];
                zim[v] += (double)z1[v] + (double)z2[v];
            }
        }
    };
    // one chunk: next chunk's loads issued first (staging before the pair
    // values, so put() waits for the staging only), this chunk's k-steps, the
    // next stage written, one barrier
    auto step = [&](int q, const fv2 (&cur)[KS], fv2 (&nxt)[KS]) {
        const bool more = q + 1 < nq;
        if (more) {
            fetch(q + 1);
            load(q + 1, nxt);
        }
        compute(cur, q & 1);
        if (more) put((q + 1) & 1);
        __syncthreads();
    };

    fv2 fa[KS], fb[KS];
    fetch(0);
    load(0, fa);
    put(0);
    __syncthreads();
    for (int q = 0; q < nq; q += 2) {
        step(q, fa, fb);
        if (q + 1 < nq) step(q + 1, fb, fa);
    }
    // the FP32 instruction's D layout: lane (li, lk) holds rows 4 lk + v of
    // column li (the FP64 one holds rows 4 v + lk)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int r = i0 + 4 * lk + v;
        if (r < count) {
            if (WithY) Y[(size_t)r * R + li] = make_double2(yre[v], yim[v]);
            if (WithZ) Z[(size_t)r * R + li] = make_double2(zre[v], zim[v]);
        }
    }
}

template <int Kern, bool WithY, bool WithZ>
__global__ void __launch_bounds__(256, 2) k_pair_batch32(const float2* __restrict__ F,
                                                         const int* __restrict__ blk_start,
                                                         const PairPoint* __restrict__ pts, int count,
                                                         ibim_params prm, const double2* __restrict__ P,
                                                         const double2* __restrict__ Q, double2* __restrict__ Y,
                                                         double2* __restrict__ Z, const int* __restrict__ stop) {
    if (stop != nullptr && *stop) return;
    pair_batch32_rows<Kern, WithY, WithZ>(F, blk_start, pts, count, prm, P, Q, Y, Z, blockIdx.x * 64);
}

// Calls fn(std::integral_constant<int, K>) with the context's pair kernel K.
template <class Fn>
void with_kernel(const Context* ctx, Fn&& fn) {
    switch (pair_kernel(ctx)) {
        case kNeumann:
            fn(std::integral_constant<int, kNeumann>{});
            break;
        case kDirichlet:
            fn(std::integral_constant<int, kDirichlet>{});
            break;
        default:
            fn(std::integral_constant<int, kCombined>{});
            break;
    }
}

// The combined operator's near-field single-layer term (0 otherwise).
void near_term(const Context* ctx, double* nr, double* ni) {
    *nr = *ni = 0.0;
    if (pair_kernel(ctx) == kCombined) combined_near(ctx->prm.k, ctx->prm.tau, ctx->prm.eta, nr, ni);
}

// The FP32 matrix-core pass over complex64 pairs (64 rows per workgroup) and
// the FP64 near field after it.
template <int Kern, bool WithY, bool WithZ>
void launch_pair_batch32(Context* ctx, const float2* F, const double2* P, const double2* Q, double2* Y, double2* Z,
                         const int* stop) {
    const int count = ctx->count;
    TimedScope ts(ctx, "batch_pair32");
    hipLaunchKernelGGL((k_pair_batch32<Kern, WithY, WithZ>), dim3((unsigned)((count + 63) / 64)), dim3(256), 0,
                       ctx->stream, F, ctx->pair_blk.as<int>(), ctx->pair_pts.as<PairPoint>(), count, ctx->prm, P, Q,
                       Y, Z, stop);
    double nr = 0.0, ni = 0.0;
    if (Kern == kCombined) combined_near(ctx->prm.k, ctx->prm.tau, ctx->prm.eta, &nr, &ni);
    hipLaunchKernelGGL((k_pair_near_batch<Kern, WithY, WithZ>),
                       dim3((unsigned)(((size_t)count * kBatchRHS + 255) / 256)), dim3(256), 0, ctx->stream,
                       ctx->tube.as<double>(), count, ctx->near_off.as<int>(), ctx->near_idx.as<int>(), P, Q, Y, Z,
                       stop, nr, ni);
}

template <int Kern>
void launch_pair_batch32_any(Context* ctx, const float2* F, const double2* P, const double2* Q, double2* Y,
                             double2* Z, const int* stop) {
    if (Y && Z)
        launch_pair_batch32<Kern, true, true>(ctx, F, P, Q, Y, Z, stop);
    else if (Y)
        launch_pair_batch32<Kern, true, false>(ctx, F, P, Q, Y, Z, stop);
    else
        launch_pair_batch32<Kern, false, true>(ctx, F, P, Q, Y, Z, stop);
}

// ------------------------------------------------ several operators, one launch --
// The far-field 16-column passes (A P and A^H Q) of several pair-stored
// operators in one persistent launch. A member is one operator with one batch
// of its right-hand sides; its work items are its 64-row blocks, numbered member
// after member (the caller puts the largest operators first). Each workgroup
// takes the next item from the queue counter until none is left, so the
// uneven members fill the GPU together and the longest items start first. An
// item runs the rows of k_pair_batch (64 rows per workgroup) or k_pair_batch32,
// so every member gets the bits of its own pass; items of a member whose stop
// flag is set are skipped.
__global__ void k_group_reset(int* queue) {
    if (threadIdx.x == 0) *queue = 0;
}

template <class MT, int Kern, bool FP32, bool WithZ>
__global__ void __launch_bounds__(256, FP32 ? 2 : 1) k_pair_batch_group(const GroupMember* __restrict__ mem,
                                                                        int nmem, int nitems,
                                                                        int* __restrict__ queue) {
    __shared__ int next;
    for (;;) {
        __syncthreads();  // the previous item's LDS reads and `next` are done
        if (threadIdx.x == 0) next = atomicAdd(queue, 1);
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(next);
        if (item >= nitems) return;
        int m = 0;
        while (m + 1 < nmem && mem[m + 1].item0 <= item) ++m;
        const GroupMember& g = mem[m];
        if (g.stop != nullptr && *g.stop) continue;
        const ibim_params prm = g.prm;
        const int rowW = (item - g.item0) * 64;
        if constexpr (FP32)
            pair_batch32_rows<Kern, true, WithZ>(static_cast<const float2*>(g.F), g.blk,
                                                 static_cast<const PairPoint*>(g.pts), g.count, prm, g.P, g.Q, g.Y,
                                                 g.Z, rowW);
        else
            pair_batch_rows<MT, Kern, true, WithZ, 4, kStoredFactors>(static_cast<const MT*>(g.F), g.blk,
                                                                      static_cast<const PairPoint*>(g.pts), g.count,
                                                                      prm, g.P, g.Q, g.Y, g.Z, rowW);
    }
}

template <class MT, int Kern, bool FP32, bool WithZ>
int launch_group(Context* lead, const GroupMember* mem, int nmem, int nitems, int* queue) {
    auto kern = k_pair_batch_group<MT, Kern, FP32, WithZ>;
    int per_cu = 0, cus = 0;
    IBIM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, 0));
    IBIM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, lead->device));
    const int grid = std::max(1, std::min(nitems, std::max(per_cu, 1) * cus));
    hipLaunchKernelGGL(k_group_reset, dim3(1), dim3(64), 0, lead->stream, queue);
    TimedScope ts(lead, "batch_group");
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(256), 0, lead->stream, mem, nmem, nitems, queue);
    return IBIM_OK;
}

int near_lists(Context* ctx, dim3 grid);

template <class MT>
int assemble_pair_as(Context* ctx, const std::vector<int2>& tiles, bool with64) {
    const int count = ctx->count;
    const unsigned nt = (unsigned)tiles.size();
    const double* tube = ctx->tube.as<double>();
    int* cnt = ctx->near_cur.as<int>();
    IBIM_HIP(hipMemsetAsync(cnt, 0, sizeof(int) * count, ctx->stream));
    const int npad = count + kPC;
    hipLaunchKernelGGL(k_pair_points, dim3((unsigned)((npad + 255) / 256)), dim3(256), 0, ctx->stream, tube, count,
                       ctx->prm.dirichlet != 0, ctx->pair_pts.as<PairPoint>(), npad);
    const dim3 grid(nt, kPR / kPairAsmRows);
    {
        TimedScope ts(ctx, "assemble");
        hipLaunchKernelGGL(k_pair_assemble<MT>, grid, dim3(kPT), 0, ctx->stream, tube, count,
                           ctx->pair_tiles.as<int2>(), ctx->prm, ctx->mat.as<MT>(),
                           with64 ? ctx->mat64.as<double2>() : nullptr, cnt);
    }
    IBIM_HIP(hipGetLastError());
    return near_lists(ctx, grid);
}

// The near-neighbour lists from the per-point counts in near_cur.
int near_lists(Context* ctx, dim3 grid) {
    const int count = ctx->count;
    const double* tube = ctx->tube.as<double>();
    int* cnt = ctx->near_cur.as<int>();
    // offsets: exclusive scan on the host (count ints)
    std::vector<int> h((size_t)count + 1, 0);
    IBIM_HIP(hipMemcpyAsync(h.data(), cnt, sizeof(int) * count, hipMemcpyDeviceToHost, ctx->stream));
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    long long run = 0;
    for (int i = 0; i < count; ++i) {
        const int c = h[i];
        h[i] = (int)run;
        run += c;
    }
    if (run > 0x7fffffffLL) {
        set_error("ibim_assemble: %lld near-field pairs overflow the 32-bit lists", run);
        return IBIM_ERR_SIZE;
    }
    h[count] = (int)run;
    if (ctx->near_idx.ensure(sizeof(int) * (size_t)(run > 0 ? run : 1))) return IBIM_ERR_NOMEM;
    IBIM_HIP(hipMemcpyAsync(ctx->near_off.ptr, h.data(), sizeof(int) * ((size_t)count + 1), hipMemcpyHostToDevice,
                            ctx->stream));
    IBIM_HIP(hipMemsetAsync(cnt, 0, sizeof(int) * count, ctx->stream));
    hipLaunchKernelGGL(k_near_fill, grid, dim3(kPT), 0, ctx->stream, tube, count, ctx->pair_tiles.as<int2>(),
                       ctx->prm.tau, ctx->near_off.as<int>(), cnt, ctx->near_idx.as<int>());
    hipLaunchKernelGGL(k_near_sort, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, ctx->stream,
                       ctx->near_off.as<int>(), ctx->near_idx.as<int>(), count);
    IBIM_HIP(hipGetLastError());
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    return IBIM_OK;
}

// Tiles (I, t) of every row block, each block's first tile, and the buffers of
// the point table and near lists; uploads the tile lists.
int pair_layout(Context* ctx, std::vector<int2>& tiles) {
    const int count = ctx->count;
    std::vector<int> blk;
    tiles.clear();
    for (int I = 0; I * kPR < count; ++I) {
        blk.push_back((int)tiles.size());
        const int T = (count - I * kPR + kPC - 1) / kPC;
        for (int t = 0; t < T; ++t) tiles.push_back(make_int2(I, t));
    }
    blk.push_back((int)tiles.size());
    if (ctx->pair_tiles.ensure(sizeof(int2) * tiles.size()) || ctx->pair_blk.ensure(sizeof(int) * blk.size()) ||
        ctx->pair_pts.ensure(sizeof(PairPoint) * ((size_t)count + kPC)) ||
        ctx->near_off.ensure(sizeof(int) * ((size_t)count + 1)) || ctx->near_cur.ensure(sizeof(int) * count))
        return IBIM_ERR_NOMEM;
    IBIM_HIP(hipMemcpyAsync(ctx->pair_tiles.ptr, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice,
                            ctx->stream));
    IBIM_HIP(hipMemcpyAsync(ctx->pair_blk.ptr, blk.data(), sizeof(int) * blk.size(), hipMemcpyHostToDevice,
                            ctx->stream));
    IBIM_HIP(hipStreamSynchronize(ctx->stream));  // the host vectors go out of scope
    return IBIM_OK;
}

}  // namespace

int launch_assemble_pair(Context* ctx, int precision, bool with64) {
    if (ctx->kind != IBIM_TUBE_TRADITIONAL) {
        set_error("ibim_assemble: pair storage needs the traditional tube");
        return IBIM_ERR_ARG;
    }
    const int count = ctx->count;
    const size_t esz = precision == IBIM_C128 ? sizeof(double2) : sizeof(float2);
    ctx->mat_precision = -1;
    ctx->mat_pair = ctx->mat_pair64 = false;
    ctx->pair_aux = ctx->free_cache_valid = false;
    ctx->ld = ((size_t)count + 63) / 64 * 64;
    std::vector<int2> tiles;
    if (int rc = pair_layout(ctx, tiles)) return rc;
    if (ctx->mat.ensure(esz * tiles.size() * kPR * kPC) ||
        (with64 && ctx->mat64.ensure(sizeof(double2) * tiles.size() * kPR * kPC)))
        return IBIM_ERR_NOMEM;
    const int rc = precision == IBIM_C128 ? assemble_pair_as<double2>(ctx, tiles, false)
                                          : assemble_pair_as<float2>(ctx, tiles, with64);
    if (rc) return rc;
    ctx->pair_ntiles = (int)tiles.size();
    for (int& w : ctx->batch_waves) w = 0;  // a new operator is tuned afresh
    ctx->pair_aux = true;
    ctx->mat_precision = precision;
    ctx->mat_pair = true;
    ctx->mat_pair64 = with64 && precision != IBIM_C128;
    ctx->regcount = 0;
    return IBIM_OK;
}

// The pair layout, point table and near lists of the current tube without the
// factors (kept until the tube changes; a pair assembly builds the same).
int prepare_pair_free(Context* ctx) {
    if (ctx->pair_aux) return IBIM_OK;
    if (ctx->kind != IBIM_TUBE_TRADITIONAL || ctx->count <= 0 || !ctx->have_tube) {
        set_error("matrix-free pair products need the traditional tube of the current operator");
        return IBIM_ERR_ARG;
    }
    const int count = ctx->count;
    std::vector<int2> tiles;
    if (int rc = pair_layout(ctx, tiles)) return rc;
    const double* tube = ctx->tube.as<double>();
    IBIM_HIP(hipMemsetAsync(ctx->near_cur.ptr, 0, sizeof(int) * count, ctx->stream));
    const int npad = count + kPC;
    hipLaunchKernelGGL(k_pair_points, dim3((unsigned)((npad + 255) / 256)), dim3(256), 0, ctx->stream, tube, count,
                       ctx->prm.dirichlet != 0, ctx->pair_pts.as<PairPoint>(), npad);
    const dim3 grid((unsigned)tiles.size(), kPR / kPairAsmRows);
    hipLaunchKernelGGL(k_near_count, grid, dim3(kPT), 0, ctx->stream, tube, count, ctx->pair_tiles.as<int2>(),
                       ctx->prm.tau, ctx->near_cur.as<int>());
    IBIM_HIP(hipGetLastError());
    if (int rc = near_lists(ctx, grid)) return rc;
    ctx->pair_ntiles = (int)tiles.size();
    for (int w = 9; w < 15; ++w) ctx->batch_waves[w] = 0;
    ctx->pair_aux = true;
    return IBIM_OK;
}

int sweep_pair_launch(Context* ctx, bool fp64_copy, const void* p, const void* q, void* ypart, void* zpart,
                      const int* done, bool with_z) {
    if (!ctx->mat_pair || (fp64_copy && !ctx->mat_pair64)) {
        set_error("pair sweep without the pair storage it needs");
        return IBIM_ERR_ARG;
    }
    const int count = ctx->count;
    const size_t ld = ctx->ld;
    const int ncol = (count + kPC - 1) / kPC, nrow = (count + kPR - 1) / kPR;
    auto P = (const double2*)p;
    auto Q = (const double2*)q;
    auto Y = (double2*)ypart;
    auto Z = (double2*)zpart;
    const int2* tiles = ctx->pair_tiles.as<int2>();
    const PairPoint* pts = ctx->pair_pts.as<PairPoint>();
    const bool wide = fp64_copy || ctx->mat_precision == IBIM_C128;
    double nr, ni;
    near_term(ctx, &nr, &ni);
    with_kernel(ctx, [&](auto K) {
        constexpr int Kern = decltype(K)::value;
        {
            TimedScope ts(ctx, fp64_copy ? "matvec64" : "matvec");
            const dim3 grid((unsigned)ctx->pair_ntiles);
            if (wide) {
                const double2* F = (fp64_copy ? ctx->mat64 : ctx->mat).as<double2>();
                if (with_z)
                    hipLaunchKernelGGL((k_pair_sweep<double2, Kern>), grid, dim3(kPT), 0, ctx->stream, F, tiles, pts,
                                       count, ld, ncol, P, Q, Y, Z, done, ctx->prm);
                else
                    hipLaunchKernelGGL((k_pair_sweep<double2, Kern, kStoredFactors, false>), grid, dim3(kPT), 0,
                                       ctx->stream, F, tiles, pts, count, ld, ncol, P, Q, Y, Z, done, ctx->prm);
            } else {
                const float2* F = ctx->mat.as<float2>();
                if (with_z)
                    hipLaunchKernelGGL((k_pair_sweep<float2, Kern>), grid, dim3(kPT), 0, ctx->stream, F, tiles, pts,
                                       count, ld, ncol, P, Q, Y, Z, done, ctx->prm);
                else
                    hipLaunchKernelGGL((k_pair_sweep<float2, Kern, kStoredFactors, false>), grid, dim3(kPT), 0,
                                       ctx->stream, F, tiles, pts, count, ld, ncol, P, Q, Y, Z, done, ctx->prm);
            }
        }
        TimedScope ts(ctx, "near");
        const size_t np = (size_t)(ncol + nrow) * ld;
        hipLaunchKernelGGL((k_pair_near<Kern>), dim3((unsigned)((count + 255) / 256)), dim3(256), 0, ctx->stream,
                           ctx->tube.as<double>(), count, ctx->near_off.as<int>(), ctx->near_idx.as<int>(), P, Q,
                           Y + np, Z + np, done, nr, ni);
    });
    IBIM_HIP(hipGetLastError());
    return IBIM_OK;
}

// The FP64 matrix-free operator's cache (ibim_set_free_cache): the complex128
// pairs of the first row blocks whose tiles fit in free_cache_max bytes, written
// by the pair assembly's kernel (so they are the bits the sweeps would compute),
// as the reference keeps the operator rows that fit and recomputes the rest
// (FASTSIZEBOUND, large3Dexample2.c:27,3905-3926). Built once per tube.
int ensure_free_cache(Context* ctx) {
    if (ctx->free_cache_valid) return IBIM_OK;
    const int count = ctx->count;
    const size_t tile = sizeof(double2) * kPR * kPC;
    int blocks = 0, tiles = 0;
    for (int I = 0; I * kPR < count; ++I) {
        const int T = (count - I * kPR + kPC - 1) / kPC;
        if ((size_t)(tiles + T) * tile > ctx->free_cache_max) break;
        tiles += T;
        blocks = I + 1;
    }
    ctx->free_cached_blocks = blocks;
    ctx->free_cached_tiles = tiles;
    if (tiles > 0) {
        if (ctx->free_cache.ensure((size_t)tiles * tile)) return IBIM_ERR_NOMEM;
        // near counts go to the build cursors, scratch once the lists exist
        TimedScope ts(ctx, "assemble");
        hipLaunchKernelGGL(k_pair_assemble<double2>, dim3((unsigned)tiles, kPR / kPairAsmRows), dim3(kPT), 0,
                           ctx->stream, ctx->tube.as<double>(), count, ctx->pair_tiles.as<int2>(), ctx->prm,
                           ctx->free_cache.as<double2>(), nullptr, ctx->near_cur.as<int>());
        IBIM_HIP(hipGetLastError());
    }
    ctx->free_cache_valid = true;
    return IBIM_OK;
}

int sweep_pair_free_launch(Context* ctx, bool fast, const void* p, const void* q, void* ypart, void* zpart,
                           const int* done, bool with_z) {
    if (int rc = prepare_pair_free(ctx)) return rc;
    if (!fast && ctx->free_cache_max > 0)
        if (int rc = ensure_free_cache(ctx)) return rc;
    const int count = ctx->count;
    const size_t ld = ctx->ld;
    const int ncol = (count + kPC - 1) / kPC, nrow = (count + kPR - 1) / kPR;
    auto P = (const double2*)p;
    auto Q = (const double2*)q;
    auto Y = (double2*)ypart;
    auto Z = (double2*)zpart;
    double nr, ni;
    near_term(ctx, &nr, &ni);
    with_kernel(ctx, [&](auto K) {
        constexpr int Kern = decltype(K)::value;
        // both products, or A p alone (with_z false: the same A p bits)
        auto sweeps = [&](auto WZ) {
            constexpr bool WithZ = decltype(WZ)::value;
            TimedScope ts(ctx, fast ? "sweep_free32" : "sweep_free64");
            if (fast)
                hipLaunchKernelGGL((k_pair_sweep<double2, Kern, kFastFactors, WithZ>),
                                   dim3((unsigned)ctx->pair_ntiles), dim3(kPT), 0, ctx->stream, nullptr,
                                   ctx->pair_tiles.as<int2>(), ctx->pair_pts.as<PairPoint>(), count, ld, ncol, P, Q,
                                   Y, Z, done, ctx->prm);
            else {
                // cached tiles (the first ones) read their factors, the others compute them
                const int nc = ctx->free_cache_max > 0 ? ctx->free_cached_tiles : 0;
                if (nc > 0)
                    hipLaunchKernelGGL((k_pair_sweep<double2, Kern, kStoredFactors, WithZ>), dim3((unsigned)nc),
                                       dim3(kPT), 0, ctx->stream, ctx->free_cache.as<double2>(),
                                       ctx->pair_tiles.as<int2>(), ctx->pair_pts.as<PairPoint>(), count, ld, ncol, P,
                                       Q, Y, Z, done, ctx->prm);
                if (ctx->pair_ntiles > nc)
                    hipLaunchKernelGGL((k_pair_sweep<double2, Kern, kExactFactors, WithZ>),
                                       dim3((unsigned)(ctx->pair_ntiles - nc)), dim3(kPT), 0, ctx->stream, nullptr,
                                       ctx->pair_tiles.as<int2>() + nc, ctx->pair_pts.as<PairPoint>(), count, ld,
                                       ncol, P, Q, Y, Z, done, ctx->prm);
            }
        };
        if (with_z)
            sweeps(std::true_type{});
        else
            sweeps(std::false_type{});
        TimedScope ts(ctx, "near");
        const size_t np = (size_t)(ncol + nrow) * ld;
        hipLaunchKernelGGL((k_pair_near<Kern>), dim3((unsigned)((count + 255) / 256)), dim3(256), 0, ctx->stream,
                           ctx->tube.as<double>(), count, ctx->near_off.as<int>(), ctx->near_idx.as<int>(), P, Q,
                           Y + np, Z + np, done, nr, ni);
    });
    IBIM_HIP(hipGetLastError());
    return IBIM_OK;
}

int pair_batch_apply(Context* ctx, bool fp64_copy, const double2* P, const double2* Q, double2* Y, double2* Z,
                     const int* stop) {
    if (!ctx->mat_pair || (fp64_copy && !ctx->mat_pair64) || (!Y && !Z) || (Y && !P) || (Z && !Q)) {
        set_error("batched pair products without the pair storage or vectors they need");
        return IBIM_ERR_ARG;
    }
    with_kernel(ctx, [&](auto K) {
        constexpr int Kern = decltype(K)::value;
        if (fp64_copy || ctx->mat_precision == IBIM_C128)
            launch_pair_batch_any<double2, Kern>(ctx, fp64_copy ? 6 : 3, (fp64_copy ? ctx->mat64 : ctx->mat).as<double2>(),
                                                 P, Q, Y, Z, stop);
        else if (ctx->fp32_products)
            launch_pair_batch32_any<Kern>(ctx, ctx->mat.as<float2>(), P, Q, Y, Z, stop);
        else
            launch_pair_batch_any<float2, Kern>(ctx, 0, ctx->mat.as<float2>(), P, Q, Y, Z, stop);
    });
    IBIM_HIP(hipGetLastError());
    return IBIM_OK;
}

int pair_batch_near(Context* ctx, const double2* P, const double2* Q, double2* Y, double2* Z, const int* stop) {
    // both products, or A P alone (Q and Z null)
    if (!ctx->pair_aux || !Y || !P || (!Q != !Z)) {
        set_error("batched near field without the near lists or vectors it needs");
        return IBIM_ERR_ARG;
    }
    double nr, ni;
    near_term(ctx, &nr, &ni);
    const dim3 grid((unsigned)(((size_t)ctx->count * kBatchRHS + 255) / 256));
    with_kernel(ctx, [&](auto K) {
        constexpr int Kern = decltype(K)::value;
        if (Z)
            hipLaunchKernelGGL((k_pair_near_batch<Kern, true, true>), grid, dim3(256), 0, ctx->stream,
                               ctx->tube.as<double>(), ctx->count, ctx->near_off.as<int>(), ctx->near_idx.as<int>(),
                               P, Q, Y, Z, stop, nr, ni);
        else
            hipLaunchKernelGGL((k_pair_near_batch<Kern, true, false>), grid, dim3(256), 0, ctx->stream,
                               ctx->tube.as<double>(), ctx->count, ctx->near_off.as<int>(), ctx->near_idx.as<int>(),
                               P, Q, Y, Z, stop, nr, ni);
    });
    IBIM_HIP(hipGetLastError());
    return IBIM_OK;
}

int pair_group_source(Context* const* ctxs, int n, bool mixed, int* source) {
    const Context* a = ctxs[0];
    for (int c = 0; c < n; ++c) {
        const Context* b = ctxs[c];
        if (!b->have_tube || !b->mat_pair || !b->pair_aux || b->count <= 0 || b->nrhs <= 0) {
            set_error("ibim_solve_group: problem %d has no pair storage of its tube or no right-hand sides", c);
            return IBIM_ERR_ARG;
        }
        if (b->device != a->device || pair_kernel(b) != pair_kernel(a) || b->mat_precision != a->mat_precision ||
            b->fp32_products != a->fp32_products) {
            set_error("ibim_solve_group: problem %d differs from problem 0 in device, kernel (Dirichlet / Neumann / "
                      "combined), pair precision or FP32 products", c);
            return IBIM_ERR_ARG;
        }
        for (int d = 0; d < c; ++d)
            if (ctxs[d] == b) {
                set_error("ibim_solve_group: problem %d appears twice", c);
                return IBIM_ERR_ARG;
            }
    }
    if (mixed && (a->mat_precision != IBIM_C64 || !a->have_tube)) {
        set_error("ibim_solve_group: the reliable solver runs on complex64 pairs (IBIM_C64_PAIR, IBIM_MIXED_PAIR)");
        return IBIM_ERR_ARG;
    }
    *source = a->mat_precision == IBIM_C128 ? GROUP_C128 : a->fp32_products ? GROUP_C64_FP32 : GROUP_C64;
    return IBIM_OK;
}

int pair_group_pass(Context* lead, int source, const GroupMember* dev_members, int nmembers, int nitems,
                    int* queue, bool with_z) {
    int rc = IBIM_OK;
    with_kernel(lead, [&](auto K) {
        constexpr int Kern = decltype(K)::value;
        auto go = [&](auto Z) {
            constexpr bool WithZ = decltype(Z)::value;
            if (source == GROUP_C128)
                rc = launch_group<double2, Kern, false, WithZ>(lead, dev_members, nmembers, nitems, queue);
            else if (source == GROUP_C64_FP32)
                rc = launch_group<float2, Kern, true, WithZ>(lead, dev_members, nmembers, nitems, queue);
            else
                rc = launch_group<float2, Kern, false, WithZ>(lead, dev_members, nmembers, nitems, queue);
        };
        if (with_z)
            go(std::true_type{});
        else
            go(std::false_type{});
    });
    if (rc) return rc;
    IBIM_HIP(hipGetLastError());
    return IBIM_OK;
}

// The matrix-free batched products: the fused pair pass evaluating each factor
// where the complex128 pair storage would read it (FP64: the same bits; fast:
// the mixed-precision factors).
int pair_free_batch_apply(Context* ctx, bool fast, const double2* P, const double2* Q, double2* Y, double2* Z,
                          const int* stop) {
    if ((!Y && !Z) || (Y && !P) || (Z && !Q)) {
        set_error("batched matrix-free pair products without the vectors they need");
        return IBIM_ERR_ARG;
    }
    if (int rc = prepare_pair_free(ctx)) return rc;
    const bool cache = !fast && ctx->free_cache_max > 0;
    if (cache)
        if (int rc = ensure_free_cache(ctx)) return rc;
    with_kernel(ctx, [&](auto K) {
        constexpr int Kern = decltype(K)::value;
        if (fast)
            launch_pair_batch_any<double2, Kern, kFastFactors>(ctx, 12, nullptr, P, Q, Y, Z, stop);
        else if (cache && ctx->free_cached_blocks > 0)
            launch_pair_batch_any<double2, Kern, kCachedFactors>(ctx, 9, ctx->free_cache.as<double2>(), P, Q, Y, Z,
                                                                 stop);
        else
            launch_pair_batch_any<double2, Kern, kExactFactors>(ctx, 9, nullptr, P, Q, Y, Z, stop);
    });
    IBIM_HIP(hipGetLastError());
    return IBIM_OK;
}

int pair_rows(Context* ctx, int row0, int nrows, double2* out_dev) {
    const dim3 grid((unsigned)((ctx->count + 255) / 256), (unsigned)nrows);
    const PairPoint* pts = ctx->pair_pts.as<PairPoint>();
    const double* tube = ctx->tube.as<double>();
    const int* blk = ctx->pair_blk.as<int>();
    with_kernel(ctx, [&](auto K) {
        constexpr int Kern = decltype(K)::value;
        if (ctx->mat_precision == IBIM_C128)
            hipLaunchKernelGGL((k_pair_rows<double2, Kern>), grid, dim3(256), 0, ctx->stream, ctx->mat.as<double2>(),
                               blk, pts, tube, ctx->count, ctx->prm, row0, nrows, out_dev);
        else
            hipLaunchKernelGGL((k_pair_rows<float2, Kern>), grid, dim3(256), 0, ctx->stream, ctx->mat.as<float2>(),
                               blk, pts, tube, ctx->count, ctx->prm, row0, nrows, out_dev);
    });
    IBIM_HIP(hipGetLastError());
    return IBIM_OK;
}

}  // namespace ibim
