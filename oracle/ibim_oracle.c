/*
 * ibim_oracle.c -- CPU FP64 restatement of the reference IBIM hot path.
 *
 * TEST INFRASTRUCTURE ONLY: it checks the HIP library's results and is never the
 * measured or shipped path; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.
 *
 * Every function names the reference lines it restates. Where the reference's
 * operation order decides a discrete branch (limiter choices, min-abs ties) or
 * the last bit of a stored value, the same order is kept so that this file and
 * the reference build agree bitwise on the golden fixtures; it is built with
 * -ffp-contract=off for the same reason.
 *
 * Parity status per function is recorded in the test files that pin it
 * (tests/test_oracle_golden.py, tests/test_oracle_math.py).
 */
#include "ibim_oracle.h"

#include <complex.h>
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#define IBO_PI 3.14159265358979323846264338328 /* constants.h:6 */

static inline size_t at(int S, int i, int j, int k) { return ((size_t)i * S + j) * S + k; }

/* ------------------------------------------------------------------------- */
/* WENO one-sided derivatives: weno.c:259-345 (Jiang-Peng HJ-WENO3/5).        */
/* ------------------------------------------------------------------------- */
void ibo_weno(const double *f, double dx, int order, double *dplus, double *dminus)
{
    const double eps = 1e-12; /* weno.c:264 */
    const double ih = 1.0 / dx;
    if (order == 3) { /* weno.c:270-294, centre at f[2] */
        double d0 = f[1] - f[0], d1 = f[2] - f[1], d2 = f[3] - f[2], d3 = f[4] - f[3];
        double central = d1 + d2;
        double s0 = d1 - d0, s1 = d2 - d1, s2 = d3 - d2;
        double r = (eps + s2 * s2) / (eps + s1 * s1);
        double w = 1. / (1. + 2. * r * r);
        *dplus = .5 * ih * (central - w * (d3 - 2. * d2 + d1));
        r = (eps + s0 * s0) / (eps + s1 * s1);
        w = 1 / (1 + 2. * r * r);
        *dminus = .5 * ih * (central - w * (d0 - 2. * d1 + d2));
        return;
    }
    /* WENO-5, weno.c:296-343, centre at f[3]; first differences scaled by 1/dx. */
    double q[6];
    for (int m = 0; m < 6; ++m) q[m] = (f[m + 1] - f[m]) * ih;
    double central = (-q[1] + 7. * (q[2] + q[3]) - q[4]) / 12.;
    double s[5];
    for (int m = 0; m < 5; ++m) s[m] = q[m + 1] - q[m]; /* second differences */
    /* forward (+) side uses s[4], s[3], s[2], s[1] */
    double b0 = 13. * (s[4] - s[3]) * (s[4] - s[3]) + 3. * (s[4] - 3. * s[3]) * (s[4] - 3. * s[3]);
    double b1 = 13. * (s[3] - s[2]) * (s[3] - s[2]) + 3. * (s[3] + s[2]) * (s[3] + s[2]);
    double b2 = 13. * (s[2] - s[1]) * (s[2] - s[1]) + 3. * (3. * s[2] - s[1]) * (3. * s[2] - s[1]);
    double a0 = 1. / ((eps + b0) * (eps + b0));
    double a1 = 6. / ((eps + b1) * (eps + b1));
    double a2 = 3. / ((eps + b2) * (eps + b2));
    double inv = 1. / (a0 + a1 + a2);
    double w0 = a0 * inv, w2 = a2 * inv;
    *dplus = central + w0 * (s[4] - 2. * s[3] + s[2]) / 3. + (w2 - .5) * (s[3] - 2. * s[2] + s[1]) / 6.;
    /* backward (-) side uses s[0], s[1], s[2], s[3] */
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

/* Mirror reflection of an index outside [0, n]: weno.c:131-139. */
static inline int mirror(int m, int n) { return m < 0 ? -m : (m > n ? 2 * n - m : m); }

/* WENO_DX / WENO_DY / WENO_DZ: weno.c:117-254. axis 0 = x (j), 1 = y (i), 2 = z (k). */
void ibo_weno_axis(const double *u, int n, double h, int i, int j, int k, int axis, int order, double *dplus,
                   double *dminus)
{
    const int S = n + 1, half = order == 5 ? 3 : 2;
    double f[7];
    for (int o = -half; o <= half; ++o) {
        int ii = i, jj = j, kk = k;
        if (axis == 0) jj = mirror(j + o, n);
        else if (axis == 1) ii = mirror(i + o, n);
        else kk = mirror(k + o, n);
        f[o + half] = u[at(S, ii, jj, kk)];
    }
    ibo_weno(f, h, order, dplus, dminus);
}

/* Godunov upwind choice of ASSIGN_GRADIENT: distance_application3D.c:88-112. */
static inline double godunov_pick(double dp, double dn, double d)
{
    double p_hi = 0.0 >= dp ? 0.0 : dp, p_lo = 0.0 <= dp ? 0.0 : dp; /* max(0,Dp), min(0,Dp) */
    double n_hi = 0.0 >= dn ? 0.0 : dn, n_lo = 0.0 <= dn ? 0.0 : dn; /* max(0,Dn), min(0,Dn) */
    if (d > 0.0) return fabs(p_lo) > fabs(n_hi) ? p_lo : n_hi;
    return fabs(p_hi) > fabs(n_lo) ? p_hi : n_lo;
}

/* ASSIGN_GRADIENT with WENO (weno_mode 3 or 5): distance_application3D.c:56-141. */
void ibo_gradient(const double *u, int n, double h, int order, double *gx, double *gy, double *gz)
{
    const int S = n + 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < S; ++i)
        for (int j = 0; j < S; ++j)
            for (int k = 0; k < S; ++k) {
                double xp, xn, yp, yn, zp, zn;
                ibo_weno_axis(u, n, h, i, j, k, 0, order, &xp, &xn);
                ibo_weno_axis(u, n, h, i, j, k, 1, order, &yp, &yn);
                ibo_weno_axis(u, n, h, i, j, k, 2, order, &zp, &zn);
                size_t id = at(S, i, j, k);
                double d = u[id];
                gx[id] = godunov_pick(xp, xn, d);
                gy[id] = godunov_pick(yp, yn, d);
                gz[id] = godunov_pick(zp, zn, d);
            }
}

/* ------------------------------------------------------------------------- */
/* Limited second derivatives: finite_difference3D.c:220-390.                */
/* ------------------------------------------------------------------------- */
static inline double pick_smaller(double a, double b) { return fabs(a) < fabs(b) ? a : b; } /* tools.c:334-337 */
static inline double pick_smallest(double a, double b, double c) { return pick_smaller(pick_smaller(a, b), c); } /* tools.c:339-344 */

/* Pure second derivative along one axis. v(m) returns u at offset m along the
 * axis from node index c in [0, n]; `tri_order` reproduces the per-axis
 * argument order of the three-way min-abs (XDER2 uses p,c,n; YDER2/ZDER2 p,n,c). */
static double d2_line(const double *u, size_t base, size_t stride, int c, int n, double h, int tri_order)
{
#define V(m) u[base + (size_t)(m) * stride]
    const double hh = h * h;
    if (c < 0) { int a = -c; return (V(a + 1) - 2.0 * V(a) + V(a - 1)) / hh; }
    if (c == 0) return 2.0 * (V(1) - V(0)) / hh;
    if (c == n) return 2.0 * (V(n - 1) - V(n)) / hh;
    if (c > n) { int a = 2 * n - c; return (V(a + 1) - 2.0 * V(a) + V(a - 1)) / hh; }
    if (c < 3) {
        double fwd = (2.0 * V(c) - 5.0 * V(c + 1) + 4.0 * V(c + 2) - V(c + 3)) / hh;
        double cen = (V(c + 1) - 2.0 * V(c) + V(c - 1)) / hh;
        return pick_smaller(fwd, cen);
    }
    if (c > n - 3) {
        double bwd = (2.0 * V(c) - 5.0 * V(c - 1) + 4.0 * V(c - 2) - V(c - 3)) / hh;
        double cen = (V(c + 1) - 2.0 * V(c) + V(c - 1)) / hh;
        return pick_smaller(bwd, cen);
    }
    double jpp = V(c + 2) - V(c + 1), jp = V(c + 1) - V(c), jn = V(c) - V(c - 1), jnn = V(c - 1) - V(c - 2);
    double fwd = (2.0 * V(c) - 5.0 * V(c + 1) + 4.0 * V(c + 2) - V(c + 3)) / hh;
    double cen = (V(c + 1) - 2.0 * V(c) + V(c - 1)) / hh;
    double bwd = (2.0 * V(c) - 5.0 * V(c - 1) + 4.0 * V(c - 2) - V(c - 3)) / hh;
#undef V
    if (jp * jnn <= 0.0 && jpp * jn <= 0.0)
        return tri_order == 0 ? pick_smallest(fwd, cen, bwd) : pick_smallest(fwd, bwd, cen);
    if (jp * jnn <= 0.0) return pick_smaller(fwd, cen);
    if (jpp * jn <= 0.0) return pick_smaller(bwd, cen);
    return cen;
}

/* XDER2 / YDER2 / ZDER2. axis 0 = x (j), 1 = y (i), 2 = z (k). */
double ibo_d2(const double *u, int n, double h, int i, int j, int k, int axis)
{
    const int S = n + 1;
    if (axis == 0) return d2_line(u, at(S, i, 0, k), (size_t)S, j, n, h, 0);
    if (axis == 1) return d2_line(u, at(S, 0, j, k), (size_t)S * S, i, n, h, 1);
    return d2_line(u, at(S, i, j, 0), 1, k, n, h, 1);
}

/* Mixed derivative over axes (a, b): secondpartialxy/xz/yz,
 * finite_difference3D.c:393-721. The nine candidates combine central and
 * second-order one-sided first differences (weights {1,-4,3}); the one with the
 * smallest magnitude, scanned in (da, db) order, is returned. */
static double dmixed_ab(const double *u, size_t o, size_t sa, size_t sb, int ia, int ib, int n, double h)
{
#define W(p, q) u[o + (ptrdiff_t)(p) * (ptrdiff_t)sa + (ptrdiff_t)(q) * (ptrdiff_t)sb]
    static const double c3[3] = {1.0, -4.0, 3.0};
    if (ia == 0 || ia == n || ib == 0 || ib == n) return 0.0;
    if (ia < 2 || ia > n - 2 || ib < 2 || ib > n - 2)
        return (W(1, 1) - W(-1, 1) - W(1, -1) + W(-1, -1)) / (4.0 * h * h);
    double best = 1000000; /* LARGENUM, constants.h:8 */
    for (int da = 0; da < 3; ++da)
        for (int db = 0; db < 3; ++db) {
            double acc = 0.0;
            if (da == 1 && db == 1) {
                acc = (W(1, 1) - W(-1, 1) - W(1, -1) + W(-1, -1));
            } else if (da == 1) {
                int s = db - 1;
                for (int m = 2; m >= 0; --m) acc += (-1.0 * (double)s * c3[2 - m] * (W(1, s * m) - W(-1, s * m)));
            } else if (db == 1) {
                int s = da - 1;
                for (int m = 2; m >= 0; --m) acc += (-1.0 * (double)s * c3[2 - m] * (W(s * m, 1) - W(s * m, -1)));
            } else {
                int sa_ = da - 1, sb_ = db - 1;
                for (int ma = 2; ma >= 0; --ma)
                    for (int mb = 2; mb >= 0; --mb)
                        acc += (c3[2 - ma] * c3[2 - mb] * (double)(sa_ * sb_) * W(sa_ * ma, sb_ * mb));
            }
            if (fabs(acc) < fabs(best)) best = acc;
        }
#undef W
    return best / (4.0 * h * h);
}

/* pair 0 = xy (a = y index i, b = x index j), 1 = xz (a = x, b = z), 2 = yz (a = y, b = z). */
double ibo_dmixed(const double *u, int n, double h, int i, int j, int k, int pair)
{
    const int S = n + 1;
    size_t o = at(S, i, j, k);
    if (pair == 0) return dmixed_ab(u, o, (size_t)S * S, (size_t)S, i, j, n, h);
    if (pair == 1) return dmixed_ab(u, o, (size_t)S, 1, j, k, n, h);
    return dmixed_ab(u, o, (size_t)S * S, 1, i, k, n, h);
}

/* BothCurvs over every node (distance_application3D.c:35-54, driven by
 * large3Dexample2.c:446-455): H = -(dxx+dyy+dzz)/2, K = sum of 2x2 Hessian minors. */
void ibo_curvatures(const double *u, int n, double h, double *mc, double *gc)
{
    const int S = n + 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < S; ++i)
        for (int j = 0; j < S; ++j)
            for (int k = 0; k < S; ++k) {
                double xx = ibo_d2(u, n, h, i, j, k, 0), yy = ibo_d2(u, n, h, i, j, k, 1);
                double zz = ibo_d2(u, n, h, i, j, k, 2);
                double xy = ibo_dmixed(u, n, h, i, j, k, 0), xz = ibo_dmixed(u, n, h, i, j, k, 1);
                double yz = ibo_dmixed(u, n, h, i, j, k, 2);
                double diag = xx * yy + yy * zz + xx * zz;
                double cross = xy * xy + yz * yz + xz * xz;
                size_t id = at(S, i, j, k);
                mc[id] = -0.5 * (xx + yy + zz);
                gc[id] = diag - cross;
            }
}

/* ------------------------------------------------------------------------- */
/* Trilinear sampling: interpolate3D.c.                                       */
/* ------------------------------------------------------------------------- */
static inline double trilin(const double *g, int S, int i, int j, int k, double lx, double ly, double lz, double h)
{
    /* corner weights as in interpolate3D.c:49-52 (x <-> j, y <-> i, z <-> k) */
    return (g[at(S, i, j, k)] * (h - lx) * (h - ly) * (h - lz) + g[at(S, i, j, k + 1)] * (h - lx) * (h - ly) * lz +
            g[at(S, i + 1, j, k)] * (h - lx) * ly * (h - lz) + g[at(S, i + 1, j, k + 1)] * (h - lx) * ly * lz +
            g[at(S, i, j + 1, k)] * lx * (h - ly) * (h - lz) + g[at(S, i, j + 1, k + 1)] * lx * (h - ly) * lz +
            g[at(S, i + 1, j + 1, k)] * lx * ly * (h - lz) + g[at(S, i + 1, j + 1, k + 1)] * lx * ly * lz) /
           (h * h * h);
}

/* cal_interpgrad: interpolate3D.c:5-102. Returns 0, 1 (|grad| off by > 0.5,
 * still normalised) or -1 (point outside its cell or the grid, or NaN; the
 * reference exits there). */
int ibo_interp_grad(int n, double h, double lo, const double *gx, const double *gy, const double *gz, const double *p,
                    double *out)
{
    const int S = n + 1;
    int i = (int)((p[1] - lo) / h), j = (int)((p[0] - lo) / h), k = (int)((p[2] - lo) / h);
    double lx = p[0] - lo - (double)j * h, ly = p[1] - lo - (double)i * h, lz = p[2] - lo - (double)k * h;
    /* the reference exits when x* is outside its cell (interpolate3D.c:31-36);
     * rounding of (x - lo) / h can put it a few ulps outside, which is accepted */
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

/* cal_BothCurvs: interpolate3D.c:569-605 (note the 1e-19 nudge of the cell index). */
void ibo_interp_curv(int n, double h, double lo, const double *mc, const double *gc, const double *p, double *out)
{
    const int S = n + 1;
    int j = (int)((p[0] + 1e-19 - lo) / h), i = (int)((p[1] + 1e-19 - lo) / h), k = (int)((p[2] + 1e-19 - lo) / h);
    double lx = p[0] - lo - (double)j * h, ly = p[1] - lo - (double)i * h, lz = p[2] - lo - (double)k * h;
    double hc = h * h * h;
    double a = mc[at(S, i, j, k)] * (h - lx) * (h - ly) * (h - lz) + mc[at(S, i + 1, j, k)] * (h - lx) * ly * (h - lz) +
               mc[at(S, i, j + 1, k)] * lx * (h - ly) * (h - lz) + mc[at(S, i + 1, j + 1, k)] * lx * ly * (h - lz) +
               mc[at(S, i, j, k + 1)] * (h - lx) * (h - ly) * lz + mc[at(S, i + 1, j, k + 1)] * (h - lx) * ly * lz +
               mc[at(S, i, j + 1, k + 1)] * lx * (h - ly) * lz + mc[at(S, i + 1, j + 1, k + 1)] * lx * ly * lz;
    double b = gc[at(S, i, j, k)] * (h - lx) * (h - ly) * (h - lz) + gc[at(S, i + 1, j, k)] * (h - lx) * ly * (h - lz) +
               gc[at(S, i, j + 1, k)] * lx * (h - ly) * (h - lz) + gc[at(S, i + 1, j + 1, k)] * lx * ly * (h - lz) +
               gc[at(S, i, j, k + 1)] * (h - lx) * (h - ly) * lz + gc[at(S, i + 1, j, k + 1)] * (h - lx) * ly * lz +
               gc[at(S, i, j + 1, k + 1)] * lx * (h - ly) * lz + gc[at(S, i + 1, j + 1, k + 1)] * lx * ly * lz;
    out[0] = a / hc;
    out[1] = b / hc;
}

/* cal_dist: interpolate3D.c:407-566 -- trilinear d minus a correction built
 * from the smallest-magnitude pure second derivative over the cell corners. */
static double smallest_corner_d2(const double *u, int n, double h, int i, int j, int k, int axis)
{
    double c111 = ibo_d2(u, n, h, i, j, k, axis), c211 = ibo_d2(u, n, h, i + 1, j, k, axis);
    double c121 = ibo_d2(u, n, h, i, j + 1, k, axis), c221 = ibo_d2(u, n, h, i + 1, j + 1, k, axis);
    double c112 = ibo_d2(u, n, h, i, j, k + 1, axis), c212 = ibo_d2(u, n, h, i + 1, j, k + 1, axis);
    double c122 = ibo_d2(u, n, h, i, j + 1, k + 1, axis), c222 = ibo_d2(u, n, h, i + 1, j + 1, k + 1, axis);
    double m11 = pick_smaller(c111, c112), m12 = pick_smaller(c121, c122);
    double m21 = pick_smaller(c211, c212), m22 = pick_smaller(c221, c222);
    return pick_smaller(pick_smaller(m11, m12), pick_smaller(m21, m22));
}

double ibo_interp_dist(const double *u, int n, double h, double lo, const double *p)
{
    const int S = n + 1;
    int jx = (int)((p[0] - lo) / h), iy = (int)((p[1] - lo) / h), kz = (int)((p[2] - lo) / h);
    double lx = p[0] - lo - jx * h, ly = p[1] - lo - iy * h, lz = p[2] - lo - kz * h;
    double sxx = smallest_corner_d2(u, n, h, iy, jx, kz, 0);
    double syy = smallest_corner_d2(u, n, h, iy, jx, kz, 1);
    double szz = smallest_corner_d2(u, n, h, iy, jx, kz, 2);
    double lin = (u[at(S, iy, jx, kz)] * (h - lx) * (h - ly) * (h - lz) + u[at(S, iy + 1, jx, kz)] * (h - lx) * ly * (h - lz) +
                  u[at(S, iy, jx + 1, kz)] * (lx) * (h - ly) * (h - lz) + u[at(S, iy + 1, jx + 1, kz)] * lx * ly * (h - lz) +
                  u[at(S, iy, jx, kz + 1)] * (h - lx) * (h - ly) * lz + u[at(S, iy + 1, jx, kz + 1)] * (h - lx) * ly * lz +
                  u[at(S, iy, jx + 1, kz + 1)] * (lx) * (h - ly) * lz + u[at(S, iy + 1, jx + 1, kz + 1)] * lx * ly * lz) /
                 (h * h * h);
    return lin - 0.5 * (sxx * (h - lx) * lx + syy * (h - ly) * ly + szz * (h - lz) * lz);
}

/* ------------------------------------------------------------------------- */
/* L1 primitives: functions.c.                                                */
/* ------------------------------------------------------------------------- */
/* cal_delta, cosine mode: functions.c:366-374. */
double ibo_delta_cos(double eps, double x)
{
    if (fabs(x) <= eps) return (1.0 / (2.0 * eps)) * (1.0 + cos(IBO_PI * x / eps));
    return 0.0;
}

/* SineWeight, first-moment branch: functions.c:835-866 with constants
 * functions.h:25-26. The reference fixes the inner root of the exponent at 0.1;
 * `root` carries it (0.1 reproduces the reference). */
double ibo_sine_weight(double t, double tau, double root)
{
    if (t <= tau || t >= 1.0) return 0.0;
    return exp(0.5 / ((t - root) * (t - 1.))) * (-759.2781934172483 * t + 446.2604260472818);
}

/* PolyWeight, first moment with smooth ends: functions.c:793-832. */
double ibo_poly_weight(double t, double tau)
{
    if (t <= tau || t >= 1.0) return 0.0;
    double tau2 = tau * tau;
    double lin = 7. * (1 + tau) * t + (-2. * (2. + 3. * tau + 2. * tau2));
    return 60. * (t - 1.) * (t - 1.) * (t - tau) * (t - tau) * lin / pow(tau - 1., 7.0);
}

/* ------------------------------------------------------------------------- */
/* Tubes.                                                                     */
/* ------------------------------------------------------------------------- */
static int in_tube(double d, const ibo_params *prm, int kind)
{
    if (kind == 0) { /* large3Dexample2.c:4354-4363 */
        if (prm->one_sided) return !(d > prm->eps || d < 0.);
        return !(fabs(d) > prm->eps);
    }
    /* combo: HelmholtzKernel.c:244-261 (exterior one-sided side when one_sided) */
    if (prm->one_sided) {
        if (prm->interior) return !(d > prm->eps || d < prm->delta_in);
        return !(d < -1. * prm->eps || d > -1. * prm->delta_in);
    }
    return !(fabs(d) > prm->eps || fabs(d) < prm->delta_in);
}

int ibo_tube_count(const double *u, int n, const ibo_params *prm, int kind)
{
    const size_t S = (size_t)n + 1, total = S * S * S;
    int c = 0;
    for (size_t id = 0; id < total; ++id) c += in_tube(u[id], prm, kind);
    return c;
}

/* Near-field curvature regularisation with the exterior sign
 * (large3Dexample2.c:4432-4445 == HelmholtzKernel.c:350-362). */
static double reg_value(double hm, double gk, double tau, double k, int interior)
{
    double t1 = hm / (4. * IBO_PI * tau);
    double t2 = -5. * tau * (2. * hm) * (4. * hm * hm - gk) / (512. * IBO_PI);
    double t3 = -25. * tau * (2. * hm) * gk / (1536. * IBO_PI);
    double t4 = k * k * tau * 2. * hm / (48. * IBO_PI);
    return interior ? t1 + t2 + t3 + t4 : -1. * (t1 + t2 + t3 + t4);
}

/* Traditional tube: omp_get_Ccomp's first-call block (large3Dexample2.c:4340-4452).
 * Combo tube: new_HelmholtzKernelCombo (HelmholtzKernel.c:223-400).
 * Returns the number of points written, or -1 if it differs from count, or -2
 * if a closest point falls outside the grid (the reference exits in both). */
int ibo_tube_build(const double *u, const double *gx, const double *gy, const double *gz, const double *mc,
                   const double *gc, int n, double h, double lo, const ibo_params *prm, int kind, int count,
                   int *node, double *out)
{
    const int S = n + 1;
    const double hcubed = h * h * h;
    int c = 0;
    for (int i = 0; i < S; ++i) {
        double zy = lo + h * (double)i;
        for (int j = 0; j < S; ++j) {
            double zx = lo + h * (double)j;
            for (int k = 0; k < S; ++k) {
                double zz = lo + h * (double)k;
                size_t id = at(S, i, j, k);
                double d = u[id];
                if (!in_tube(d, prm, kind)) continue;
                if (c >= count) return -1;
                double p[3] = {zx - d * gx[id], zy - d * gy[id], zz - d * gz[id]};
                double nv[3], cv[2];
                if (ibo_interp_grad(n, h, lo, gx, gy, gz, p, nv) < 0) return -2;
                /* both callers normalise once more (large3Dexample2.c:4407-4410,
                 * HelmholtzKernel.c:291-294) */
                double nn = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
                nv[0] = nv[0] / nn;
                nv[1] = nv[1] / nn;
                nv[2] = nv[2] / nn;
                double jac = 1 + 2. * d * mc[id] + d * d * gc[id];
                ibo_interp_curv(n, h, lo, mc, gc, p, cv);
                double wt, wq;
                if (kind == 0) {
                    wt = prm->one_sided ? 2. * ibo_delta_cos(prm->eps, d) : ibo_delta_cos(prm->eps, d);
                    wq = hcubed * wt * jac;
                } else {
                    double t = fabs(d / prm->eps), r = prm->delta_in / prm->eps;
                    wt = ibo_sine_weight(t, r, prm->sine_root) / (prm->one_sided ? prm->eps : 2.0 * prm->eps);
                    wq = hcubed * jac * wt;
                }
                node[c] = (int)id;
                out[IBO_ZX * count + c] = zx;
                out[IBO_ZY * count + c] = zy;
                out[IBO_ZZ * count + c] = zz;
                out[IBO_SX * count + c] = p[0];
                out[IBO_SY * count + c] = p[1];
                out[IBO_SZ * count + c] = p[2];
                out[IBO_DIST * count + c] = d;
                out[IBO_NX * count + c] = nv[0];
                out[IBO_NY * count + c] = nv[1];
                out[IBO_NZ * count + c] = nv[2];
                out[IBO_DELTA * count + c] = wt;
                out[IBO_JAC * count + c] = jac;
                out[IBO_W * count + c] = wq;
                out[IBO_REG * count + c] = reg_value(cv[0], cv[1], prm->tau, prm->k, prm->interior);
                out[IBO_HM * count + c] = cv[0];
                out[IBO_GK * count + c] = cv[1];
                ++c;
            }
        }
    }
    return c == count ? c : -1;
}

/* ------------------------------------------------------------------------- */
/* Kernel entries.                                                            */
/* ------------------------------------------------------------------------- */
#define T(f, p) tube[(size_t)(f) * (size_t)count + (size_t)(p)]

/* Near-field value of the single layer of the combined-field operator: the
 * mean of e^{ik r} / (4 pi r) over a flat disc of radius tau,
 * (1 / (pi tau^2)) int_0^tau e^{ik r} / (4 pi r) 2 pi r dr
 * = (e^{ik tau} - 1) / (2 pi i k tau^2), and 1 / (2 pi tau) at k = 0. */
static double complex single_layer_disc_mean(double k, double tau)
{
    if (k == 0.0) return 1.0 / (2.0 * IBO_PI * tau);
    return (cexp(I * k * tau) - 1.0) / (2.0 * IBO_PI * I * k * tau * tau);
}

/* Traditional kernel: omp_get_Ccomp entry block (large3Dexample2.c:4463-4564).
 * A_ij = w_j grad_x Phi(x*_i, y*_j) . n_(i or j) (+1/2 on the diagonal) for
 * rho > tau, else w_j reg (reg of row i for Neumann, of column j for
 * Dirichlet -- the reference indexes this with a stale loop variable).
 * combined (no counterpart in the reference, whose combined field is the
 * Neumann combo kernel, HelmholtzKernel.c:509-521): the exterior Dirichlet
 * operator 1/2 I + K - i eta S of the potential u = (D - i eta S) phi, with the
 * single layer's near field taken as its disc mean. */
static void trad_entry(int count, const double *tube, const ibo_params *prm, int pi, int pj, double *re, double *im)
{
    const double rfp = 1. / (4. * IBO_PI);
    double d1 = T(IBO_SX, pi) - T(IBO_SX, pj), d2 = T(IBO_SY, pi) - T(IBO_SY, pj), d3 = T(IBO_SZ, pi) - T(IBO_SZ, pj);
    double rho = sqrt(d1 * d1 + d2 * d2 + d3 * d3);
    double pr, pim;
    if (rho > prm->tau) {
        double rr = 1. / rho, arg = prm->k * rho, sv = sin(arg), cv = cos(arg);
        double esv = sv * rr, ecv = cv * rr, den = rfp * rr * rr;
        double c1 = -1. * (prm->k * sv + ecv) * den, c2 = (prm->k * cv - esv) * den;
        int q = prm->dirichlet ? pj : pi;
        double ip = d1 * T(IBO_NX, q) + d2 * T(IBO_NY, q) + d3 * T(IBO_NZ, q);
        if (prm->interior) { pr = (-1.0) * c1 * ip; pim = (-1.0) * c2 * ip; }
        else { pr = c1 * ip; pim = c2 * ip; }
    } else {
        pr = T(IBO_REG, prm->dirichlet ? pj : pi);
        pim = 0.0;
    }
    if (prm->combined) {
        double complex s = rho > prm->tau ? cexp(I * prm->k * rho) / (4. * IBO_PI * rho)
                                          : single_layer_disc_mean(prm->k, prm->tau);
        double complex t = -I * prm->eta * s;
        pr += creal(t);
        pim += cimag(t);
    }
    *re = T(IBO_W, pj) * pr;
    *im = T(IBO_W, pj) * pim;
    if (pi == pj) *re += 0.5;
}

void ibo_kernel_trad_rows(int count, const double *tube, const ibo_params *prm, int row0, int nrows, double *are,
                          double *aim)
{
#pragma omp parallel for schedule(static)
    for (int r = 0; r < nrows; ++r)
        for (int j = 0; j < count; ++j)
            trad_entry(count, tube, prm, row0 + r, j, &are[(size_t)r * count + j], &aim[(size_t)r * count + j]);
}

/* Combined-field kernel: HelmholtzKernel.c:410-607 (Sine weight branch).
 * Returns the number of near-field (rho <= tau) entries in the rows. */
int ibo_kernel_combo_rows(int count, const double *tube, const ibo_params *prm, int row0, int nrows, double *are,
                          double *aim)
{
    const double rfp = 1. / (4. * IBO_PI);
    const double k = prm->k, eta = prm->eta;
    int regs = 0;
#pragma omp parallel for schedule(static) reduction(+ : regs)
    for (int r = 0; r < nrows; ++r) {
        int i = row0 + r;
        double sxi = T(IBO_SX, i), syi = T(IBO_SY, i), szi = T(IBO_SZ, i);
        double nxi = T(IBO_NX, i), nyi = T(IBO_NY, i), nzi = T(IBO_NZ, i);
        for (int j = 0; j < count; ++j) {
            double sxj = T(IBO_SX, j), syj = T(IBO_SY, j), szj = T(IBO_SZ, j), Rj = T(IBO_W, j);
            /* single layer: -i eta (dG/dn_x R_j - delta_ij / 2) */
            double e1 = sxi - sxj, e2 = syi - syj, e3 = szi - szj;
            double rho = sqrt(e1 * e1 + e2 * e2 + e3 * e3);
            double pr, pim;
            if (rho > prm->tau) {
                double rr = 1. / rho, arg = k * rho, sv = sin(arg), cv = cos(arg);
                double esv = sv * rr, ecv = cv * rr, den = rfp * rr * rr;
                double c1 = -1. * (k * sv + ecv) * den, c2 = (k * cv - esv) * den;
                double ip = -1. * (e1 * nxi + e2 * nyi + e3 * nzi);
                pr = c1 * ip;
                pim = c2 * ip;
            } else {
                regs += 1;
                pr = T(IBO_REG, i);
                pim = 0.0;
            }
            double tr = pr * Rj, ti = pim * Rj;
            if (i == j) tr -= 0.5;
            double sre = eta * ti, sim = -1. * eta * tr;
            /* double layer at x = x*_i + |d_j| n_i */
            double ad = fabs(T(IBO_DIST, j));
            double x1 = sxi + ad * nxi, x2 = syi + ad * nyi, x3 = szi + ad * nzi;
            double f1 = x1 - sxj, f2 = x2 - syj, f3 = x3 - szj;
            double dn = sqrt(f1 * f1 + f2 * f2 + f3 * f3), rd = 1. / dn, arg = k * dn;
            double nxj = T(IBO_NX, j), nyj = T(IBO_NY, j), nzj = T(IBO_NZ, j);
            double ip1 = nxi * nxj + nyi * nyj + nzi * nzj;
            double ip2 = f1 * nxi + f2 * nyi + f3 * nzi;
            double ip3 = f1 * nxj + f2 * nyj + f3 * nzj;
            double ct1 = ip1 * rd, ct2 = ip2 * rd * ip3 * rd;
            double cvv = cos(arg) * rd, svv = sin(arg) * rd;
            double t1r = -1. * (cvv * rd + k * svv) * ct1, t1i = (k * cvv - svv * rd) * ct1;
            double t2r = -1. * k * k * cvv * ct2, t2i = -1. * k * k * svv * ct2;
            double t3r = 3. * k * svv * ct2 * rd, t3i = -3. * k * cvv * ct2 * rd;
            double t4r = 3. * cvv * rd * ct2 * rd, t4i = 3. * svv * rd * ct2 * rd;
            double dre = Rj * (t1r + t2r + t3r + t4r) * rfp, dim = Rj * (t1i + t2i + t3i + t4i) * rfp;
            are[(size_t)r * count + j] = sre + dre;
            aim[(size_t)r * count + j] = sim + dim;
        }
    }
    return regs;
}

/* ------------------------------------------------------------------------- */
/* Complex BiCG: BICGSTAB.c:5-339 (serial) and 344-765 (OpenMP).              */
/* ------------------------------------------------------------------------- */
static double dot(int n, const double *a, const double *b)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

typedef void (*sweep_fn)(int n, const double *are, const double *aim, const double *pr, const double *pi,
                         const double *qr, const double *qi, double *vr, double *vi, double *wr, double *wi);

/* v = A p (rows), w = A^H q (columns): BICGSTAB.c:164-210. */
static void sweep_serial(int n, const double *are, const double *aim, const double *pr, const double *pi,
                         const double *qr, const double *qi, double *vr, double *vi, double *wr, double *wi)
{
    for (int j = 0; j < n; ++j) {
        double s = 0.0, si = 0.0;
        const double *ar = are + (size_t)j * n, *ai = aim + (size_t)j * n;
        for (int k = 0; k < n; ++k) {
            s += (ar[k] * pr[k] - ai[k] * pi[k]);
            si += (ar[k] * pi[k] + ai[k] * pr[k]);
        }
        vr[j] = s;
        vi[j] = si;
    }
    for (int j = 0; j < n; ++j) {
        double s = 0.0, si = 0.0;
        for (int k = 0; k < n; ++k) {
            double a = are[(size_t)k * n + j], b = aim[(size_t)k * n + j];
            s += (a * qr[k] + b * qi[k]);
            si += (a * qi[k] - b * qr[k]);
        }
        wr[j] = s;
        wi[j] = si;
    }
}

/* Same arithmetic with the rows / columns split over OpenMP threads
 * (BICGSTAB.c:511-552, schedule(static, chunk)); each output is still summed
 * serially, so the result does not depend on the thread count. */
static void sweep_omp(int n, const double *are, const double *aim, const double *pr, const double *pi,
                      const double *qr, const double *qi, double *vr, double *vi, double *wr, double *wi)
{
#pragma omp parallel for schedule(static, 16)
    for (int j = 0; j < n; ++j) {
        double s = 0.0, si = 0.0;
        const double *ar = are + (size_t)j * n, *ai = aim + (size_t)j * n;
        for (int k = 0; k < n; ++k) {
            s += (ar[k] * pr[k] - ai[k] * pi[k]);
            si += (ar[k] * pi[k] + ai[k] * pr[k]);
        }
        vr[j] = s;
        vi[j] = si;
    }
#pragma omp parallel for schedule(static, 16)
    for (int j = 0; j < n; ++j) {
        double s = 0.0, si = 0.0;
        for (int k = 0; k < n; ++k) {
            double a = are[(size_t)k * n + j], b = aim[(size_t)k * n + j];
            s += (a * qr[k] + b * qi[k]);
            si += (a * qi[k] - b * qr[k]);
        }
        wr[j] = s;
        wi[j] = si;
    }
}

void ibo_sweep_rows(int nrows, int ncols, const double *are, const double *aim, const double *pr, const double *pi,
                    const double *qr, const double *qi, double *vr, double *vi, double *wr, double *wi)
{
#pragma omp parallel for schedule(static, 16)
    for (int j = 0; j < nrows; ++j) {
        double s = 0.0, si = 0.0;
        const double *ar = are + (size_t)j * ncols, *ai = aim + (size_t)j * ncols;
        for (int k = 0; k < ncols; ++k) {
            s += (ar[k] * pr[k] - ai[k] * pi[k]);
            si += (ar[k] * pi[k] + ai[k] * pr[k]);
        }
        vr[j] = s;
        vi[j] = si;
    }
#pragma omp parallel for schedule(static, 16)
    for (int j = 0; j < ncols; ++j) {
        double s = 0.0, si = 0.0;
        for (int k = 0; k < nrows; ++k) {
            double a = are[(size_t)k * ncols + j], b = aim[(size_t)k * ncols + j];
            s += (a * qr[k] + b * qi[k]);
            si += (a * qi[k] - b * qr[k]);
        }
        wr[j] = s;
        wi[j] = si;
    }
}

static int bicg_core(int n, const double *are, const double *aim, const double *bre, const double *bim, double *xre,
                     double *xim, int maxit, double tol, double *hist, sweep_fn sweep)
{
    double *buf = (double *)calloc((size_t)n * 16, sizeof(double));
    double *x = buf, *xi = x + n, *r = xi + n, *ri = r + n, *rt = ri + n, *rti = rt + n, *p = rti + n, *pi = p + n;
    double *pt = pi + n, *pti = pt + n, *v = pti + n, *vi = v + n, *w = vi + n, *wi = w + n, *po = wi + n,
           *poi = po + n;
    double normb = sqrt(dot(n, bre, bre) + dot(n, bim, bim));
    for (int c = 0; c < n; ++c) {
        r[c] = bre[c];
        ri[c] = bim[c];
        p[c] = r[c];
        pi[c] = ri[c];
        rt[c] = r[c];
        rti[c] = (-1.0) * ri[c]; /* shadow residual r~ = conj(b): BICGSTAB.c:113-116 */
        pt[c] = rt[c];
        pti[c] = rti[c];
    }
    double rho = dot(n, r, rt) + dot(n, ri, rti), rhoi = dot(n, ri, rt) - dot(n, r, rti);
    double bcsum = 0.0;
    for (int c = 0; c < n; ++c) {
        bcsum += fabs(r[c]);
        bcsum += fabs(ri[c]);
    }
    if (bcsum < pow(10, -13.0)) { /* BICGSTAB.c:138-145 */
        r[0] = 0.000000001;
        ri[0] = 0.000000001;
        rt[0] = 0.000000001;
        rti[0] = 0.000000001;
    }
    int its = 0;
    if (bcsum < pow(10, -15.0)) { /* zero data gives a zero density: BICGSTAB.c:147-156 */
        memset(xre, 0, sizeof(double) * n);
        memset(xim, 0, sizeof(double) * n);
        free(buf);
        return 0;
    }
    double err = INFINITY, prev;
    for (int it = 1; it <= maxit; ++it) {
        sweep(n, are, aim, p, pi, pt, pti, v, vi, w, wi);
        /* alpha = <r, r~> / <A p, p~>, <a, b> = sum a conj(b): BICGSTAB.c:217-229 */
        double ca = dot(n, r, rt) + dot(n, ri, rti), cb = dot(n, ri, rt) - dot(n, r, rti);
        double cc = dot(n, v, pt) + dot(n, vi, pti), cd = dot(n, vi, pt) - dot(n, v, pti);
        double den = cc * cc + cd * cd;
        double al = (ca * cc + cb * cd) / den, ali = (cb * cc - ca * cd) / den;
        for (int j = 0; j < n; ++j) { /* BICGSTAB.c:231-242 */
            x[j] = x[j] + al * p[j] - ali * pi[j];
            xi[j] = xi[j] + ali * p[j] + al * pi[j];
            r[j] = r[j] - al * v[j] + ali * vi[j];
            ri[j] = ri[j] - ali * v[j] - al * vi[j];
            rt[j] = rt[j] - al * w[j] - ali * wi[j];
            rti[j] = rti[j] + ali * w[j] - al * wi[j];
        }
        /* beta = <r, r~>_new / <r, r~>_old: BICGSTAB.c:246-259 */
        ca = dot(n, r, rt) + dot(n, ri, rti);
        cb = dot(n, ri, rt) - dot(n, r, rti);
        den = rho * rho + rhoi * rhoi;
        double be = (ca * rho + cb * rhoi) / den, bei = (cb * rho - ca * rhoi) / den;
        rho = ca;
        rhoi = cb;
        memcpy(po, p, sizeof(double) * n * 2); /* p, pi are contiguous */
        double *pto = (double *)malloc(sizeof(double) * 2 * n);
        memcpy(pto, pt, sizeof(double) * n * 2);
        for (int j = 0; j < n; ++j) { /* BICGSTAB.c:271-278 */
            p[j] = r[j] + be * po[j] - bei * poi[j];
            pi[j] = ri[j] + bei * po[j] + be * poi[j];
            pt[j] = rt[j] + be * pto[j] + bei * pto[n + j];
            pti[j] = rti[j] - bei * pto[j] + be * pto[n + j];
        }
        free(pto);
        prev = err;
        err = sqrt(dot(n, r, r) + dot(n, ri, ri)) / normb;
        if (hist) hist[it - 1] = err;
        its = it;
        /* stop rules: BICGSTAB.c:301-306; the first iteration has no previous
         * error (uninitialised in the reference), taken as +inf here */
        if (err < tol) break;
        if (err != err) break;
        if (fabs(prev - err) < pow(10, -3) * tol) break;
    }
    memcpy(xre, x, sizeof(double) * n);
    memcpy(xim, xi, sizeof(double) * n);
    free(buf);
    return its;
}

int ibo_bicg(int n, const double *are, const double *aim, const double *bre, const double *bim, double *xre,
             double *xim, int maxit, double tol, double *hist)
{
    return bicg_core(n, are, aim, bre, bim, xre, xim, maxit, tol, hist, sweep_serial);
}

int ibo_bicg_omp(int n, const double *are, const double *aim, const double *bre, const double *bim, double *xre,
                 double *xim, int maxit, double tol, double *hist)
{
    return bicg_core(n, are, aim, bre, bim, xre, xim, maxit, tol, hist, sweep_omp);
}

/* ------------------------------------------------------------------------- */
/* Right-hand sides and field evaluation.                                     */
/* ------------------------------------------------------------------------- */
/* Point-source data Phi(x*, x0) = e^{ik r} / (4 pi r) (the analytic exterior
 * field of an interior source; the reference's cal_u, large3Dexample2.c:4642,
 * is the same field times -4 pi / k). */
void ibo_rhs_point_source(int count, const double *tube, double k, const double *x0, double *bre, double *bim)
{
    for (int p = 0; p < count; ++p) {
        double a = T(IBO_SX, p) - x0[0], b = T(IBO_SY, p) - x0[1], c = T(IBO_SZ, p) - x0[2];
        double r = sqrt(a * a + b * b + c * c);
        bre[p] = cos(k * r) / (4. * IBO_PI * r);
        bim[p] = sin(k * r) / (4. * IBO_PI * r);
    }
}

/* Neumann data of the incident plane wave amp e^{ik dir.x} along the
 * interpolated unit normal at x*: cal_partialu / cal_partialuimg with
 * INCIDENTALWAVE (large3Dexample2.c:4735-4746, 4792-4810). interp_n holds the
 * once-normalised trilinear gradient at x* (3 x count, field-major). */
void ibo_rhs_plane_neumann(int count, const double *tube, const double *interp_n, double k, double amp,
                           const double *dir, double *bre, double *bim)
{
    for (int p = 0; p < count; ++p) {
        double x = T(IBO_SX, p), y = T(IBO_SY, p), z = T(IBO_SZ, p);
        double dx = interp_n[p], dy = interp_n[count + p], dz = interp_n[2 * count + p];
        double ph = x * dir[0] + y * dir[1] + z * dir[2];
        double dn = -1. * (dir[0] * dx + dir[1] * dy + dir[2] * dz);
        bre[p] = amp * k * dn * sin(k * ph);
        bim[p] = -amp * k * dn * cos(k * ph);
    }
}

/* Sound-soft data of an incident plane wave: b = -amp e^{ik dir.x*}. */
void ibo_rhs_plane_dirichlet(int count, const double *tube, double k, double amp, const double *dir, double *bre,
                             double *bim)
{
    for (int p = 0; p < count; ++p) {
        double ph = k * (T(IBO_SX, p) * dir[0] + T(IBO_SY, p) * dir[1] + T(IBO_SZ, p) * dir[2]);
        bre[p] = -amp * cos(ph);
        bim[p] = -amp * sin(ph);
    }
}

/* Field at receivers. mode 0: u = sum_j (-Phi(x, y*_j)) phi_j w_j, zero for
 * |x - y*_j| <= tau (NeumannTest, large3Dexample2.c:1810-1824). mode 1: the
 * double layer matching the Dirichlet kernel, u = sum_j w_j grad_x Phi(x, y*_j)
 * . n_j phi_j. mode 2: combo sample kernel (Combo_NeumannTest,
 * large3Dexample2.c:1980-2058), with R_j the combo weight and the near field
 * taken from reg of point j. */
void ibo_eval(int count, const double *tube, const ibo_params *prm, const double *phre, const double *phim, int nrecv,
              const double *xyz, int mode, double *ure, double *uim)
{
    const double k = prm->k;
#pragma omp parallel for schedule(static)
    for (int q = 0; q < nrecv; ++q) {
        const double *x = xyz + 3 * (size_t)q;
        double sr = 0.0, si = 0.0;
        for (int j = 0; j < count; ++j) {
            double a = x[0] - T(IBO_SX, j), b = x[1] - T(IBO_SY, j), c = x[2] - T(IBO_SZ, j);
            double rho = sqrt(a * a + b * b + c * c);
            double kr, ki;
            if (mode == 0) {
                if (rho > prm->tau) {
                    kr = cos(k * rho) / (-4. * IBO_PI * rho);
                    ki = sin(k * rho) / (-4. * IBO_PI * rho);
                } else {
                    kr = 0.;
                    ki = 0.;
                }
            } else if (mode == 1 || mode == 3) {
                if (rho > prm->tau) {
                    double sv = sin(k * rho), cv = cos(k * rho), den = 1. / (4. * IBO_PI * rho * rho);
                    double c1 = -(k * sv + cv / rho) * den, c2 = (k * cv - sv / rho) * den;
                    double ip = a * T(IBO_NX, j) + b * T(IBO_NY, j) + c * T(IBO_NZ, j);
                    kr = c1 * ip;
                    ki = c2 * ip;
                    if (mode == 3) {
                        double complex t = -I * prm->eta * cexp(I * k * rho) / (4. * IBO_PI * rho);
                        kr += creal(t);
                        ki += cimag(t);
                    }
                } else {
                    kr = 0.;
                    ki = 0.;
                }
            } else {
                double sv = sin(k * rho), cv = cos(k * rho);
                double ip = a * T(IBO_NX, j) + b * T(IBO_NY, j) + c * T(IBO_NZ, j);
                double f1 = 4. * IBO_PI * rho, f2 = f1 * rho;
                if (rho > prm->tau) {
                    kr = (cv / rho + k * sv) * ip / f2;
                    ki = (sv / rho - k * cv) * ip / f2;
                } else {
                    kr = T(IBO_REG, j);
                    ki = 0.0;
                }
                kr += prm->eta * sv / f1;
                ki += -1. * prm->eta * cv / f1;
            }
            double wj = T(IBO_W, j);
            sr += (kr * phre[j] - ki * phim[j]) * wj;
            si += (kr * phim[j] + ki * phre[j]) * wj;
        }
        ure[q] = sr;
        uim[q] = si;
    }
}
#undef T
