// ibim_entries.hpp -- operator entries of the two IBIM formulations, FP64.
//
// Point data is read from the field-major tube (IBIM_F_* fields, stride count).
// The operation order follows the reference entry by entry (cited per function)
// so that the GPU operator agrees with the CPU oracle to a few ulp.
#pragma once

#include "../../include/ibim.h"
#include "ibim_math.hpp"

namespace ibim {

struct PointData {  // the per-point values an entry needs
    double sx, sy, sz, nx, ny, nz, w, reg, dist;
};

__device__ __forceinline__ PointData load_point(const double* __restrict__ tube, size_t count, int p) {
    PointData d;
    d.sx = tube[IBIM_F_SX * count + p];
    d.sy = tube[IBIM_F_SY * count + p];
    d.sz = tube[IBIM_F_SZ * count + p];
    d.nx = tube[IBIM_F_NX * count + p];
    d.ny = tube[IBIM_F_NY * count + p];
    d.nz = tube[IBIM_F_NZ * count + p];
    d.w = tube[IBIM_F_W * count + p];
    d.reg = tube[IBIM_F_REG * count + p];
    d.dist = tube[IBIM_F_DIST * count + p];
    return d;
}

// Traditional kernel entry, omp_get_Ccomp (large3Dexample2.c:4463-4564):
// A_ij = w_j grad_x Phi(x*_i, y*_j) . n_q (+1/2 on the diagonal) for rho > tau,
// q = i (Neumann) or j (Dirichlet); w_j * reg_q in the near field (the
// reference indexes reg with a stale loop variable there; q is the intended
// point). combined (exterior Dirichlet only) adds the combined-field coupling
// -i eta w_j Phi(x*_i, y*_j), in the near field -i eta w_j times the disc mean
// of Phi (single_layer_near): A = 1/2 I + K - i eta S.
__device__ __forceinline__ void trad_entry(const PointData& a, const PointData& b, bool diag, double k, double tau,
                                           int dirichlet, int interior, int combined, double eta, double* re,
                                           double* im) {
    double d1 = a.sx - b.sx, d2 = a.sy - b.sy, d3 = a.sz - b.sz;
    double rho = sqrt(d1 * d1 + d2 * d2 + d3 * d3);
    double pr, pim;
    if (rho > tau) {
        double c1, c2;
        grad_phi_coeffs(k, rho, &c1, &c2);
        double ip = dirichlet ? d1 * b.nx + d2 * b.ny + d3 * b.nz : d1 * a.nx + d2 * a.ny + d3 * a.nz;
        if (interior) {
            pr = (-1.0) * c1 * ip;
            pim = (-1.0) * c2 * ip;
        } else {
            pr = c1 * ip;
            pim = c2 * ip;
        }
        if (combined) {
            double gr, gi;
            phi_value(k, rho, &gr, &gi);
            pr = pr + eta * gi;
            pim = pim - eta * gr;
        }
    } else {
        pr = dirichlet ? b.reg : a.reg;
        pim = 0.0;
        if (combined) {
            double nr, ni;
            combined_near(k, tau, eta, &nr, &ni);
            pr = pr + nr;
            pim = ni;
        }
    }
    *re = b.w * pr;
    *im = b.w * pim;
    if (diag) *re += 0.5;
}

// Combined-field entry, new_HelmholtzKernelCombo (HelmholtzKernel.c:438-588):
// -i eta (dG/dn_x(x*_i, y*_j) R_j - delta_ij / 2) plus the double layer
// d^2 G / dn_x dn_y at x = x*_i + |d_j| n_i, times R_j / (4 pi).
// Returns 1 when the single-layer part took the near-field value.
__device__ __forceinline__ int combo_entry(const PointData& a, const PointData& b, bool diag, double k, double tau,
                                           double eta, double* re, double* im) {
    const double rfp = 1. / (4. * kPi);
    double e1 = a.sx - b.sx, e2 = a.sy - b.sy, e3 = a.sz - b.sz;
    double rho = sqrt(e1 * e1 + e2 * e2 + e3 * e3);
    double pr, pim;
    int near = 0;
    if (rho > tau) {
        double c1, c2;
        grad_phi_coeffs(k, rho, &c1, &c2);
        double ip = -1. * (e1 * a.nx + e2 * a.ny + e3 * a.nz);
        pr = c1 * ip;
        pim = c2 * ip;
    } else {
        near = 1;
        pr = a.reg;
        pim = 0.0;
    }
    double tr = pr * b.w, ti = pim * b.w;
    if (diag) tr -= 0.5;
    double sre = eta * ti, sim = -1. * eta * tr;
    double ad = fabs(b.dist);
    double x1 = a.sx + ad * a.nx, x2 = a.sy + ad * a.ny, x3 = a.sz + ad * a.nz;
    double f1 = x1 - b.sx, f2 = x2 - b.sy, f3 = x3 - b.sz;
    double dn = sqrt(f1 * f1 + f2 * f2 + f3 * f3), rd = 1. / dn;
    double ip1 = a.nx * b.nx + a.ny * b.ny + a.nz * b.nz;
    double ip2 = f1 * a.nx + f2 * a.ny + f3 * a.nz;
    double ip3 = f1 * b.nx + f2 * b.ny + f3 * b.nz;
    double ct1 = ip1 * rd, ct2 = ip2 * rd * ip3 * rd;
    double sv, cv;
    sincos(k * dn, &sv, &cv);
    double cvv = cv * rd, svv = sv * rd;
    double t1r = -1. * (cvv * rd + k * svv) * ct1, t1i = (k * cvv - svv * rd) * ct1;
    double t2r = -1. * k * k * cvv * ct2, t2i = -1. * k * k * svv * ct2;
    double t3r = 3. * k * svv * ct2 * rd, t3i = -3. * k * cvv * ct2 * rd;
    double t4r = 3. * cvv * rd * ct2 * rd, t4i = 3. * svv * rd * ct2 * rd;
    double dre = b.w * (t1r + t2r + t3r + t4r) * rfp, dim = b.w * (t1i + t2i + t3i + t4i) * rfp;
    *re = sre + dre;
    *im = sim + dim;
    return near;
}

// sin and cos of 2 pi f for a reduced phase f in [-1/2, 1/2] revolutions, in
// FP32: quadrant q = rint(4 f) and x = 2 pi (f - q / 4) in [-pi/4, pi/4] are
// formed in FP64 and x is rounded once; Taylor polynomials in x^2 (sin to x^9,
// cos to x^10: truncation below 2e-9) and a quarter-turn rotation by q. About
// one FP32 ulp, without the systematic error of the hardware sin / cos units.
__device__ __forceinline__ void sincos_rev_f32(double f, float* s, float* c) {
    const double q = rint(4.0 * f);
    const float x = (float)((2.0 * kPi) * fma(-0.25, q, f));
    const float x2 = x * x;
    float ps = 1.0f / 362880.0f;                 //  1/9!
    ps = fmaf(ps, x2, -1.0f / 5040.0f);          // -1/7!
    ps = fmaf(ps, x2, 1.0f / 120.0f);            //  1/5!
    ps = fmaf(ps, x2, -1.0f / 6.0f);             // -1/3!
    const float sx = fmaf(ps * x2, x, x);
    float pc = -1.0f / 3628800.0f;               // -1/10!
    pc = fmaf(pc, x2, 1.0f / 40320.0f);          //  1/8!
    pc = fmaf(pc, x2, -1.0f / 720.0f);           // -1/6!
    pc = fmaf(pc, x2, 1.0f / 24.0f);             //  1/4!
    pc = fmaf(pc, x2, -0.5f);
    const float cx = fmaf(pc, x2, 1.0f);
    const int quad = ((int)q) & 3;               // q in [-2, 2]
    *s = quad == 0 ? sx : quad == 1 ? cx : quad == 2 ? -sx : -cx;
    *c = quad == 0 ? cx : quad == 1 ? -sx : quad == 2 ? -cx : sx;
}

}  // namespace ibim
