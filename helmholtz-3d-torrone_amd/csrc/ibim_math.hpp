// ibim_math.hpp -- scalar math of the IBIM path, shared by the gfx950 kernels
// and by the host-side scalar entry points (weno, BothCurvs, ...).
//
// Grid layout: node (i, j, k) = (y, x, z) at ((i*S)+j)*S+k with S = n+1, k
// fastest, the reference's double*** u[i][j][k] flattened
// (large3Dexample2.c:413-421). Wherever the reference's operation order decides
// a discrete branch (limiter choices, min-abs ties) the same order is used, and
// these translation units are built with -ffp-contract=off, so the GPU geometry
// rounds like the reference's x86-64 build.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstddef>
#include <cstdint>

#define IBIM_HD __host__ __device__ __forceinline__

namespace ibim {

constexpr double kPi = 3.14159265358979323846264338328;  // constants.h:6

IBIM_HD size_t node_at(int S, int i, int j, int k) { return ((size_t)i * S + j) * S + k; }

// Magnitude used by every branch decision; for dual numbers the branch is taken
// on the value, so derivatives are those of the branch the value selects.
IBIM_HD double absv(double a) { return fabs(a); }

// Forward-mode dual number with ND derivative directions (exact vector-Jacobian
// products of the piecewise-smooth stencils, used by the adjoint).
template <int ND>
struct Dual {
    double v;
    double d[ND];
    IBIM_HD Dual() : v(0.0) {
        for (int q = 0; q < ND; ++q) d[q] = 0.0;
    }
    IBIM_HD Dual(double x) : v(x) {
        for (int q = 0; q < ND; ++q) d[q] = 0.0;
    }
    IBIM_HD static Dual seed(double x, int q) {
        Dual r(x);
        r.d[q] = 1.0;
        return r;
    }
};
template <int ND>
IBIM_HD double absv(const Dual<ND>& a) { return fabs(a.v); }
template <int ND>
IBIM_HD Dual<ND> operator+(const Dual<ND>& a, const Dual<ND>& b) {
    Dual<ND> r(a.v + b.v);
    for (int q = 0; q < ND; ++q) r.d[q] = a.d[q] + b.d[q];
    return r;
}
template <int ND>
IBIM_HD Dual<ND> operator-(const Dual<ND>& a, const Dual<ND>& b) {
    Dual<ND> r(a.v - b.v);
    for (int q = 0; q < ND; ++q) r.d[q] = a.d[q] - b.d[q];
    return r;
}
template <int ND>
IBIM_HD Dual<ND> operator-(const Dual<ND>& a) {
    Dual<ND> r(-a.v);
    for (int q = 0; q < ND; ++q) r.d[q] = -a.d[q];
    return r;
}
template <int ND>
IBIM_HD Dual<ND> operator*(const Dual<ND>& a, const Dual<ND>& b) {
    Dual<ND> r(a.v * b.v);
    for (int q = 0; q < ND; ++q) r.d[q] = a.d[q] * b.v + a.v * b.d[q];
    return r;
}
template <int ND>
IBIM_HD Dual<ND> operator*(double a, const Dual<ND>& b) {
    Dual<ND> r(a * b.v);
    for (int q = 0; q < ND; ++q) r.d[q] = a * b.d[q];
    return r;
}
template <int ND>
IBIM_HD Dual<ND> operator*(const Dual<ND>& a, double b) {
    return b * a;
}
template <int ND>
IBIM_HD Dual<ND> operator/(const Dual<ND>& a, const Dual<ND>& b) {
    Dual<ND> r(a.v / b.v);
    const double ib = 1.0 / b.v;
    for (int q = 0; q < ND; ++q) r.d[q] = (a.d[q] - r.v * b.d[q]) * ib;
    return r;
}
template <int ND>
IBIM_HD Dual<ND> operator/(const Dual<ND>& a, double b) {
    Dual<ND> r(a.v / b);
    for (int q = 0; q < ND; ++q) r.d[q] = a.d[q] / b;
    return r;
}
template <int ND>
IBIM_HD Dual<ND> operator/(double a, const Dual<ND>& b) {
    return Dual<ND>(a) / b;
}
template <int ND>
IBIM_HD Dual<ND> operator+(double a, const Dual<ND>& b) {
    Dual<ND> r = b;
    r.v = a + b.v;
    return r;
}
template <int ND>
IBIM_HD Dual<ND> operator+(const Dual<ND>& a, double b) {
    return b + a;
}
template <int ND>
IBIM_HD Dual<ND> operator-(const Dual<ND>& a, double b) {
    Dual<ND> r = a;
    r.v = a.v - b;
    return r;
}
template <int ND>
IBIM_HD Dual<ND> operator-(double a, const Dual<ND>& b) {
    return Dual<ND>(a) - b;
}
template <int ND>
IBIM_HD Dual<ND>& operator+=(Dual<ND>& a, const Dual<ND>& b) {
    a = a + b;
    return a;
}

// ---------------------------------------------------------------------------
// WENO one-sided derivatives (weno.c:259-345): Jiang-Peng HJ-WENO3 / WENO5,
// eps = 1e-12, ideal weights 1:6:3. f holds 5 (order 3) or 7 (order 5) samples.
// ---------------------------------------------------------------------------
template <class T>
IBIM_HD void weno_1d(const T* f, double dx, int order, T* dplus, T* dminus) {
    const double eps = 1e-12;
    const double ih = 1.0 / dx;
    if (order == 3) {
        T d0 = f[1] - f[0], d1 = f[2] - f[1], d2 = f[3] - f[2], d3 = f[4] - f[3];
        T central = d1 + d2;
        T s0 = d1 - d0, s1 = d2 - d1, s2 = d3 - d2;
        T r = (eps + s2 * s2) / (eps + s1 * s1);
        T w = 1. / (1. + 2. * r * r);
        *dplus = .5 * ih * (central - w * (d3 - 2. * d2 + d1));
        r = (eps + s0 * s0) / (eps + s1 * s1);
        w = 1 / (1 + 2. * r * r);
        *dminus = .5 * ih * (central - w * (d0 - 2. * d1 + d2));
        return;
    }
    T q[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) q[m] = (f[m + 1] - f[m]) * ih;
    T central = (-q[1] + 7. * (q[2] + q[3]) - q[4]) / 12.;
    T s[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) s[m] = q[m + 1] - q[m];
    T b0 = 13. * (s[4] - s[3]) * (s[4] - s[3]) + 3. * (s[4] - 3. * s[3]) * (s[4] - 3. * s[3]);
    T b1 = 13. * (s[3] - s[2]) * (s[3] - s[2]) + 3. * (s[3] + s[2]) * (s[3] + s[2]);
    T b2 = 13. * (s[2] - s[1]) * (s[2] - s[1]) + 3. * (3. * s[2] - s[1]) * (3. * s[2] - s[1]);
    T a0 = 1. / ((eps + b0) * (eps + b0));
    T a1 = 6. / ((eps + b1) * (eps + b1));
    T a2 = 3. / ((eps + b2) * (eps + b2));
    T inv = 1. / (a0 + a1 + a2);
    T w0 = a0 * inv, w2 = a2 * inv;
    *dplus = central + w0 * (s[4] - 2. * s[3] + s[2]) / 3. + (w2 - .5) * (s[3] - 2. * s[2] + s[1]) / 6.;
    b0 = 13. * (s[0] - s[1]) * (s[0] - s[1]) + 3. * (s[0] - 3. * s[1]) * (s[0] - 3. * s[1]);
    b1 = 13. * (s[1] - s[2]) * (s[1] - s[2]) + 3. * (s[1] + s[2]) * (s[1] + s[2]);
    b2 = 13. * (s[2] - s[3]) * (s[2] - s[3]) + 3. * (3. * s[2] - s[3]) * (3. * s[2] - s[3]);
    a0 = 1. / ((eps + b0) * (eps + b0));
    a1 = 6. / ((eps + b1) * (eps + b1));
    a2 = 3. / ((eps + b2) * (eps + b2));
    inv = 1. / (a0 + a1 + a2);
    w0 = a0 * inv;
    w2 = a2 * inv;
    *dminus = central - w0 * (s[0] - 2. * s[1] + s[2]) / 3. - (w2 - .5) * (s[1] - 2. * s[2] + s[3]) / 6.;
}

// Mirror reflection at the grid faces (weno.c:131-139).
IBIM_HD int mirror(int m, int n) { return m < 0 ? -m : (m > n ? 2 * n - m : m); }

// Line accessor over a flat grid: value at absolute index m along one axis.
struct FlatLine {
    const double* u;
    size_t base, stride;
    IBIM_HD double operator()(int m) const { return u[base + (size_t)m * stride]; }
};

// Plane accessor over a flat grid: value at offsets (p, q) along axes (a, b)
// from the node at flat offset o.
struct FlatPlane {
    const double* u;
    ptrdiff_t o, sa, sb;
    IBIM_HD double operator()(int p, int q) const { return u[o + (ptrdiff_t)p * sa + (ptrdiff_t)q * sb]; }
};

// One axis of WENO_DX / WENO_DY / WENO_DZ (weno.c:117-254).
template <class Line, class T>
IBIM_HD void weno_line(const Line& V, int c, int n, double h, int order, T* dp, T* dm) {
    T f[7];
    const int half = order == 5 ? 3 : 2;
    for (int o = -half; o <= half; ++o) f[o + half] = V(mirror(c + o, n));
    weno_1d(f, h, order, dp, dm);
}

// Godunov upwind choice (distance_application3D.c:88-112).
IBIM_HD double valof(double a) { return a; }
template <int ND>
IBIM_HD double valof(const Dual<ND>& a) { return a.v; }

template <class T>
IBIM_HD T godunov_pick(T dp, T dn, double d) {
    const T zero(0.0);
    T p_hi = 0.0 >= valof(dp) ? zero : dp, p_lo = 0.0 <= valof(dp) ? zero : dp;
    T n_hi = 0.0 >= valof(dn) ? zero : dn, n_lo = 0.0 <= valof(dn) ? zero : dn;
    if (d > 0.0) return absv(p_lo) > absv(n_hi) ? p_lo : n_hi;
    return absv(p_hi) > absv(n_lo) ? p_hi : n_lo;
}

// ---------------------------------------------------------------------------
// Limited second derivatives (finite_difference3D.c:220-721).
// ---------------------------------------------------------------------------
template <class T>
IBIM_HD T pick_smaller(T a, T b) {
    return absv(a) < absv(b) ? a : b;
}
template <class T>
IBIM_HD T pick_smallest(T a, T b, T c) {
    return pick_smaller(pick_smaller(a, b), c);
}

// XDER2 / YDER2 / ZDER2 along a strided line; tri_order keeps each function's
// argument order of the three-way min-abs (XDER2: fwd, cen, bwd; others fwd, bwd, cen).
template <class Line>
IBIM_HD auto d2_line(const Line& V, int c, int n, double h, int tri_order) -> decltype(V(0)) {
    using T = decltype(V(0));
#define IBIM_V(m) V(m)
    const double hh = h * h;
    if (c < 0) {
        int a = -c;
        return (IBIM_V(a + 1) - 2.0 * IBIM_V(a) + IBIM_V(a - 1)) / hh;
    }
    if (c == 0) return 2.0 * (IBIM_V(1) - IBIM_V(0)) / hh;
    if (c == n) return 2.0 * (IBIM_V(n - 1) - IBIM_V(n)) / hh;
    if (c > n) {
        int a = 2 * n - c;
        return (IBIM_V(a + 1) - 2.0 * IBIM_V(a) + IBIM_V(a - 1)) / hh;
    }
    if (c < 3) {
        T fwd = (2.0 * IBIM_V(c) - 5.0 * IBIM_V(c + 1) + 4.0 * IBIM_V(c + 2) - IBIM_V(c + 3)) / hh;
        T cen = (IBIM_V(c + 1) - 2.0 * IBIM_V(c) + IBIM_V(c - 1)) / hh;
        return pick_smaller(fwd, cen);
    }
    if (c > n - 3) {
        T bwd = (2.0 * IBIM_V(c) - 5.0 * IBIM_V(c - 1) + 4.0 * IBIM_V(c - 2) - IBIM_V(c - 3)) / hh;
        T cen = (IBIM_V(c + 1) - 2.0 * IBIM_V(c) + IBIM_V(c - 1)) / hh;
        return pick_smaller(bwd, cen);
    }
    T v_m3 = IBIM_V(c - 3), v_m2 = IBIM_V(c - 2), v_m1 = IBIM_V(c - 1), v0 = IBIM_V(c);
    T v1 = IBIM_V(c + 1), v2 = IBIM_V(c + 2), v3 = IBIM_V(c + 3);
#undef IBIM_V
    const double jpp = valof(v2) - valof(v1), jp = valof(v1) - valof(v0);
    const double jn = valof(v0) - valof(v_m1), jnn = valof(v_m1) - valof(v_m2);
    T fwd = (2.0 * v0 - 5.0 * v1 + 4.0 * v2 - v3) / hh;
    T cen = (v1 - 2.0 * v0 + v_m1) / hh;
    T bwd = (2.0 * v0 - 5.0 * v_m1 + 4.0 * v_m2 - v_m3) / hh;
    if (jp * jnn <= 0.0 && jpp * jn <= 0.0)
        return tri_order == 0 ? pick_smallest(fwd, cen, bwd) : pick_smallest(fwd, bwd, cen);
    if (jp * jnn <= 0.0) return pick_smaller(fwd, cen);
    if (jpp * jn <= 0.0) return pick_smaller(bwd, cen);
    return cen;
}

// Mixed derivative over axes (a, b) with the node at plane offset (0, 0):
// secondpartialxy/xz/yz. Nine candidates built from central and second-order
// one-sided first differences ({1,-4,3}); the smallest in magnitude wins,
// scanned in (da, db) order starting from 1e6 (LARGENUM).
template <class Plane>
IBIM_HD auto dmixed(const Plane& Wp, int ia, int ib, int n, double h) -> decltype(Wp(0, 0)) {
    using T = decltype(Wp(0, 0));
#define IBIM_W(p, q) Wp((p), (q))
    const double c3[3] = {1.0, -4.0, 3.0};
    if (ia == 0 || ia == n || ib == 0 || ib == n) return T(0.0);
    if (ia < 2 || ia > n - 2 || ib < 2 || ib > n - 2)
        return (IBIM_W(1, 1) - IBIM_W(-1, 1) - IBIM_W(1, -1) + IBIM_W(-1, -1)) / (4.0 * h * h);
    T best(1000000.0);
    for (int da = 0; da < 3; ++da)
        for (int db = 0; db < 3; ++db) {
            T acc(0.0);
            if (da == 1 && db == 1) {
                acc = (IBIM_W(1, 1) - IBIM_W(-1, 1) - IBIM_W(1, -1) + IBIM_W(-1, -1));
            } else if (da == 1) {
                int s = db - 1;
                for (int m = 2; m >= 0; --m)
                    acc += (-1.0 * (double)s * c3[2 - m] * (IBIM_W(1, s * m) - IBIM_W(-1, s * m)));
            } else if (db == 1) {
                int s = da - 1;
                for (int m = 2; m >= 0; --m)
                    acc += (-1.0 * (double)s * c3[2 - m] * (IBIM_W(s * m, 1) - IBIM_W(s * m, -1)));
            } else {
                int sa_ = da - 1, sb_ = db - 1;
                for (int ma = 2; ma >= 0; --ma)
                    for (int mb = 2; mb >= 0; --mb)
                        acc += (c3[2 - ma] * c3[2 - mb] * (double)(sa_ * sb_) * IBIM_W(sa_ * ma, sb_ * mb));
            }
            if (absv(acc) < absv(best)) best = acc;
        }
#undef IBIM_W
    return best / (4.0 * h * h);
}

// The six second derivatives {xx, yy, zz, xy, xz, yz} at node (i, j, k).
IBIM_HD void second_derivs(const double* u, int n, double h, int i, int j, int k, double* o6) {
    const int S = n + 1;
    o6[0] = d2_line(FlatLine{u, node_at(S, i, 0, k), (size_t)S}, j, n, h, 0);
    o6[1] = d2_line(FlatLine{u, node_at(S, 0, j, k), (size_t)S * S}, i, n, h, 1);
    o6[2] = d2_line(FlatLine{u, node_at(S, i, j, 0), 1}, k, n, h, 1);
    const ptrdiff_t o = (ptrdiff_t)node_at(S, i, j, k);
    o6[3] = dmixed(FlatPlane{u, o, (ptrdiff_t)S * S, (ptrdiff_t)S}, i, j, n, h);
    o6[4] = dmixed(FlatPlane{u, o, (ptrdiff_t)S, 1}, j, k, n, h);
    o6[5] = dmixed(FlatPlane{u, o, (ptrdiff_t)S * S, 1}, i, k, n, h);
}

// Mean and Gaussian curvature from the six derivatives (distance_application3D.c:35-54).
IBIM_HD void curvatures_from(const double* d, double* mean, double* gauss) {
    double diag = d[0] * d[1] + d[1] * d[2] + d[0] * d[2];
    double cross = d[3] * d[3] + d[5] * d[5] + d[4] * d[4];
    *mean = -0.5 * (d[0] + d[1] + d[2]);
    *gauss = diag - cross;
}

// ---------------------------------------------------------------------------
// Trilinear sampling (interpolate3D.c).
// ---------------------------------------------------------------------------
IBIM_HD double trilin(const double* g, int S, int i, int j, int k, double lx, double ly, double lz, double h) {
    return (g[node_at(S, i, j, k)] * (h - lx) * (h - ly) * (h - lz) +
            g[node_at(S, i, j, k + 1)] * (h - lx) * (h - ly) * lz +
            g[node_at(S, i + 1, j, k)] * (h - lx) * ly * (h - lz) + g[node_at(S, i + 1, j, k + 1)] * (h - lx) * ly * lz +
            g[node_at(S, i, j + 1, k)] * lx * (h - ly) * (h - lz) + g[node_at(S, i, j + 1, k + 1)] * lx * (h - ly) * lz +
            g[node_at(S, i + 1, j + 1, k)] * lx * ly * (h - lz) + g[node_at(S, i + 1, j + 1, k + 1)] * lx * ly * lz) /
           (h * h * h);
}

// cal_interpgrad (interpolate3D.c:5-102): 0 ok, 1 |grad| off by > 0.5 (still
// normalised), -1 outside its cell or the grid, or NaN (the reference exits).
IBIM_HD int interp_grad(int n, double h, double lo, const double* gx, const double* gy, const double* gz,
                        const double* p, double* out) {
    const int S = n + 1;
    int i = (int)((p[1] - lo) / h), j = (int)((p[0] - lo) / h), k = (int)((p[2] - lo) / h);
    double lx = p[0] - lo - (double)j * h, ly = p[1] - lo - (double)i * h, lz = p[2] - lo - (double)k * h;
    // the reference exits when x* is outside its cell (interpolate3D.c:31-36);
    // rounding of (x - lo) / h can put it a few ulps outside, which is accepted
    const double slack = 1e-12 * h;
    if (lx > h + slack || lx < -slack || ly > h + slack || ly < -slack || lz < -slack || lz > h + slack) return -1;
    if (i < 0 || j < 0 || k < 0 || i >= n || j >= n || k >= n) return -1;
    double vx = trilin(gx, S, i, j, k, lx, ly, lz, h);
    double vy = trilin(gy, S, i, j, k, lx, ly, lz, h);
    double vz = trilin(gz, S, i, j, k, lx, ly, lz, h);
    if (vx != vx || vy != vy || vz != vz) return -1;
    double nrm = sqrt(vx * vx + vy * vy + vz * vz);
    out[0] = vx / nrm;
    out[1] = vy / nrm;
    out[2] = vz / nrm;
    return fabs(nrm - 1.0) > 0.5 ? 1 : 0;
}

// cal_BothCurvs (interpolate3D.c:569-605), with its 1e-19 nudge of the cell index.
IBIM_HD int interp_curv(int n, double h, double lo, const double* mc, const double* gc, const double* p,
                        double* out) {
    const int S = n + 1;
    int j = (int)((p[0] + 1e-19 - lo) / h), i = (int)((p[1] + 1e-19 - lo) / h), k = (int)((p[2] + 1e-19 - lo) / h);
    if (i < 0 || j < 0 || k < 0 || i >= n || j >= n || k >= n) return -1;
    double lx = p[0] - lo - (double)j * h, ly = p[1] - lo - (double)i * h, lz = p[2] - lo - (double)k * h;
    double hc = h * h * h;
    double a = mc[node_at(S, i, j, k)] * (h - lx) * (h - ly) * (h - lz) +
               mc[node_at(S, i + 1, j, k)] * (h - lx) * ly * (h - lz) +
               mc[node_at(S, i, j + 1, k)] * lx * (h - ly) * (h - lz) + mc[node_at(S, i + 1, j + 1, k)] * lx * ly * (h - lz) +
               mc[node_at(S, i, j, k + 1)] * (h - lx) * (h - ly) * lz + mc[node_at(S, i + 1, j, k + 1)] * (h - lx) * ly * lz +
               mc[node_at(S, i, j + 1, k + 1)] * lx * (h - ly) * lz + mc[node_at(S, i + 1, j + 1, k + 1)] * lx * ly * lz;
    double b = gc[node_at(S, i, j, k)] * (h - lx) * (h - ly) * (h - lz) +
               gc[node_at(S, i + 1, j, k)] * (h - lx) * ly * (h - lz) +
               gc[node_at(S, i, j + 1, k)] * lx * (h - ly) * (h - lz) + gc[node_at(S, i + 1, j + 1, k)] * lx * ly * (h - lz) +
               gc[node_at(S, i, j, k + 1)] * (h - lx) * (h - ly) * lz + gc[node_at(S, i + 1, j, k + 1)] * (h - lx) * ly * lz +
               gc[node_at(S, i, j + 1, k + 1)] * lx * (h - ly) * lz + gc[node_at(S, i + 1, j + 1, k + 1)] * lx * ly * lz;
    out[0] = a / hc;
    out[1] = b / hc;
    return 0;
}

// ---------------------------------------------------------------------------
// Quadrature weights and the near-field value (functions.c, large3Dexample2.c).
// ---------------------------------------------------------------------------
// Cosine delta (functions.c:366-374).
IBIM_HD double delta_cos(double eps, double x) {
    if (fabs(x) <= eps) return (1.0 / (2.0 * eps)) * (1.0 + cos(kPi * x / eps));
    return 0.0;
}

// SineWeight, first moment (functions.c:835-866, constants functions.h:25-26);
// `root` is the inner root of the exponent (0.1 in the reference).
IBIM_HD double sine_weight(double t, double tau, double root) {
    if (t <= tau || t >= 1.0) return 0.0;
    return exp(0.5 / ((t - root) * (t - 1.))) * (-759.2781934172483 * t + 446.2604260472818);
}

// PolyWeight, first moment with smooth ends (functions.c:793-832).
IBIM_HD double poly_weight(double t, double tau) {
    if (t <= tau || t >= 1.0) return 0.0;
    double tau2 = tau * tau;
    double lin = 7. * (1 + tau) * t + (-2. * (2. + 3. * tau + 2. * tau2));
    return 60. * (t - 1.) * (t - 1.) * (t - tau) * (t - tau) * lin / pow(tau - 1., 7.0);
}

// Curvature regularisation of the near field (large3Dexample2.c:4432-4445,
// HelmholtzKernel.c:350-362); exterior problems take the negative sign.
IBIM_HD double reg_value(double hm, double gk, double tau, double k, int interior) {
    double t1 = hm / (4. * kPi * tau);
    double t2 = -5. * tau * (2. * hm) * (4. * hm * hm - gk) / (512. * kPi);
    double t3 = -25. * tau * (2. * hm) * gk / (1536. * kPi);
    double t4 = k * k * tau * 2. * hm / (48. * kPi);
    return interior ? t1 + t2 + t3 + t4 : -1. * (t1 + t2 + t3 + t4);
}

// Traditional tube membership (large3Dexample2.c:4354-4363) and combo tube
// membership (HelmholtzKernel.c:244-261).
IBIM_HD bool in_tube(double d, int kind, double eps, double delta_in, int one_sided, int interior) {
    if (kind == 0) return one_sided ? !(d > eps || d < 0.) : !(fabs(d) > eps);
    if (one_sided) return interior ? !(d > eps || d < delta_in) : !(d < -1. * eps || d > -1. * delta_in);
    return !(fabs(d) > eps || fabs(d) < delta_in);
}

// grad_x Phi(x, y) = (c1 + i c2) (x - y) for Phi = e^{ik rho} / (4 pi rho)
// (large3Dexample2.c:4478-4487).
IBIM_HD void grad_phi_coeffs(double k, double rho, double* c1, double* c2) {
    const double rfp = 1. / (4. * kPi);
    double rr = 1. / rho, sv, cv;
    sincos(k * rho, &sv, &cv);
    double esv = sv * rr, ecv = cv * rr, den = rfp * rr * rr;
    *c1 = -1. * (k * sv + ecv) * den;
    *c2 = (k * cv - esv) * den;
}

// Phi(x, y) = e^{ik rho} / (4 pi rho), the single-layer kernel of the
// combined-field Dirichlet operator.
IBIM_HD void phi_value(double k, double rho, double* re, double* im) {
    const double rfp = 1. / (4. * kPi);
    double rr = 1. / rho, sv, cv;
    sincos(k * rho, &sv, &cv);
    *re = cv * rr * rfp;
    *im = sv * rr * rfp;
}

// Near-field value of the single layer: the mean of Phi over a flat disc of
// radius tau about the point, (e^{ik tau} - 1) / (2 pi i k tau^2)
// = [sin(k tau) + i (1 - cos(k tau))] / (2 pi k tau^2), written with sinc
// factors so that k = 0 gives the Laplace value 1 / (2 pi tau).
IBIM_HD void single_layer_near(double k, double tau, double* re, double* im) {
    const double x = k * tau, hx = 0.5 * x;
    const double sx = x != 0.0 ? sin(x) / x : 1.0;
    const double sh = hx != 0.0 ? sin(hx) / hx : 1.0;
    *re = sx / (2. * kPi * tau);
    *im = k * sh * sh / (4. * kPi);
}

// The combined-field coupling's near-field term -i eta (S_re + i S_im) added
// to the double layer's real near value: (eta S_im, -eta S_re).
IBIM_HD void combined_near(double k, double tau, double eta, double* re, double* im) {
    double sr, si;
    single_layer_near(k, tau, &sr, &si);
    *re = eta * si;
    *im = -1. * eta * sr;
}

}  // namespace ibim
