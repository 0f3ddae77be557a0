// ibim_adjoint.hip -- gradient of a receiver misfit with respect to the level
// set and the wavenumber (traditional formulation).
//
//   J = sum_r |u_r - t_r|^2,  u = E phi,  A phi = b.
//   dJ = 2 Re[e^H dE phi] + 2 Re[lambda^H (db - dA phi)],  A^H lambda = E^H e.
//
// 1. e = u - t; g = E^H e and the evaluator's entry derivatives (k_adj_eval).
// 2. lambda from the same BiCG on the adjoint system (the fused sweep already
//    produces A^H products).
// 3. Operator entry derivatives recomputed on the fly in two passes with the
//    other index streamed through LDS: per row i (x*_i, and n_i / reg_i for
//    Neumann, k) and per column j (y*_j, w_j, and n_j / reg_j for Dirichlet).
//    No derivative matrix is stored -- tiles are recomputed from the tube.
// 4. The right-hand side's derivative (point source / Dirichlet plane wave).
// 5. Tube data -> grid fields: projection x* = z - d grad d, the twice
//    normalised trilinear normal, the weight h^3 delta(d) J, and the
//    regularisation from trilinear H*, K* (atomics into grid adjoints).
// 6. Grid fields -> level-set values: WENO / Godunov gradient and the limited
//    second derivatives of every node with a nonzero adjoint, by forward-mode
//    dual numbers over each stencil (branch choices held fixed).
#include <cmath>
#include <vector>

#include "ibim_entries.hpp"
#include "ibim_internal.hpp"

namespace ibim {

namespace {

constexpr int kT = 256;
constexpr int kStageA = 128;

// per-point adjoint fields, field-major with stride count
enum { A_X = 0, A_Y, A_Z, A_NX, A_NY, A_NZ, A_W, A_REG, A_K, A_NF };

// c(rho) = e^{ik rho} (ik rho - 1) / (4 pi rho^3) with grad_x Phi = c (x - y),
// its rho-derivative e^{ik rho} (3 - 3ik rho - k^2 rho^2) / (4 pi rho^4) and
// its k-derivative -k e^{ik rho} / (4 pi rho).
struct CD {
    double cr, ci, pr, pi, kr, ki;
};

__device__ __forceinline__ CD cder(double k, double rho) {
    double s, c;
    sincos(k * rho, &s, &c);
    const double kr = k * rho;
    const double inv3 = 1.0 / (4.0 * kPi * rho * rho * rho);
    const double inv4 = inv3 / rho;
    const double inv1 = 1.0 / (4.0 * kPi * rho);
    const double a = 3.0 - kr * kr, b = -3.0 * kr;
    CD o;
    o.cr = (-c - s * kr) * inv3;
    o.ci = (c * kr - s) * inv3;
    o.pr = (c * a - s * b) * inv4;
    o.pi = (s * a + c * b) * inv4;
    o.kr = -k * c * inv1;
    o.ki = -k * s * inv1;
    return o;
}

__device__ __forceinline__ double re_mul(double ar, double ai, double xr, double xi) { return ar * xr - ai * xi; }

struct Pt {
    double sx, sy, sz, nx, ny, nz, w, reg;
};

__device__ __forceinline__ Pt load_pt(const double* __restrict__ t, size_t c, int p) {
    return Pt{t[IBIM_F_SX * c + p], t[IBIM_F_SY * c + p], t[IBIM_F_SZ * c + p], t[IBIM_F_NX * c + p],
              t[IBIM_F_NY * c + p], t[IBIM_F_NZ * c + p], t[IBIM_F_W * c + p], t[IBIM_F_REG * c + p]};
}

// 1. g = E^H e and the evaluator's derivatives w.r.t. y*_j, n_j, w_j, k.
__global__ void __launch_bounds__(kT) k_adj_eval(const double* __restrict__ tube, int count, ibim_params prm,
                                                 int mode, const double2* __restrict__ phi,
                                                 const double* __restrict__ xyz, const double2* __restrict__ e,
                                                 int nrecv, double2* __restrict__ g, double* __restrict__ adj) {
    const int j = blockIdx.x * kT + threadIdx.x;
    if (j >= count) return;
    const size_t c = (size_t)count;
    const Pt b = load_pt(tube, c, j);
    const double2 ph = phi[j];
    const double k = prm.k;
    double gr = 0, gi = 0, ap[3] = {0, 0, 0}, an[3] = {0, 0, 0}, aw = 0, ak = 0;
    for (int q = 0; q < nrecv; ++q) {
        const double r[3] = {xyz[3 * q] - b.sx, xyz[3 * q + 1] - b.sy, xyz[3 * q + 2] - b.sz};
        const double rho = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        if (!(rho > prm.tau)) continue;
        const double2 eq = e[q];
        const double br = 2.0 * (eq.x * ph.x + eq.y * ph.y), bi = 2.0 * (eq.x * ph.y - eq.y * ph.x);
        const CD cd = cder(k, rho);
        double Er, Ei;
        if (mode == IBIM_EVAL_DOUBLE) {
            const double sd = r[0] * b.nx + r[1] * b.ny + r[2] * b.nz;
            Er = b.w * cd.cr * sd;
            Ei = b.w * cd.ci * sd;
            const double nn[3] = {b.nx, b.ny, b.nz};
            for (int a = 0; a < 3; ++a) {
                ap[a] += re_mul(br, bi, -b.w * (cd.pr * sd * r[a] / rho + cd.cr * nn[a]),
                                -b.w * (cd.pi * sd * r[a] / rho + cd.ci * nn[a]));
                an[a] += re_mul(br, bi, b.w * cd.cr * r[a], b.w * cd.ci * r[a]);
            }
            aw += re_mul(br, bi, cd.cr * sd, cd.ci * sd);
            ak += re_mul(br, bi, b.w * cd.kr * sd, b.w * cd.ki * sd);
        } else {
            double sv, cv;
            sincos(k * rho, &sv, &cv);
            const double pr = cv / (4.0 * kPi * rho), pim = sv / (4.0 * kPi * rho);
            Er = -b.w * pr;
            Ei = -b.w * pim;
            for (int a = 0; a < 3; ++a) ap[a] += re_mul(br, bi, b.w * cd.cr * r[a], b.w * cd.ci * r[a]);
            aw += re_mul(br, bi, -pr, -pim);
            ak += re_mul(br, bi, b.w * sv / (4.0 * kPi), -b.w * cv / (4.0 * kPi));
        }
        gr += Er * eq.x + Ei * eq.y;
        gi += Er * eq.y - Ei * eq.x;
    }
    g[j] = make_double2(gr, gi);
    adj[A_X * c + j] = ap[0];
    adj[A_Y * c + j] = ap[1];
    adj[A_Z * c + j] = ap[2];
    adj[A_NX * c + j] = an[0];
    adj[A_NY * c + j] = an[1];
    adj[A_NZ * c + j] = an[2];
    adj[A_W * c + j] = aw;
    adj[A_REG * c + j] = 0.0;
    adj[A_K * c + j] = ak;
}

// 3. Operator derivatives. alpha_ij = -2 conj(lambda_i) phi_j. One thread per
// output point (row pass: point = i; column pass: point = j), the other index
// streamed through LDS in chunks.
template <int Rows>
__global__ void __launch_bounds__(kT) k_adj_operator(const double* __restrict__ tube, int count, ibim_params prm,
                                                     const double2* __restrict__ phi,
                                                     const double2* __restrict__ lam, double* __restrict__ adj) {
    __shared__ Pt sp[kStageA];
    __shared__ double2 sv[kStageA];
    const int me = blockIdx.x * kT + threadIdx.x;
    const size_t c = (size_t)count;
    const bool active = me < count;
    const Pt P = active ? load_pt(tube, c, me) : Pt{};
    // the row point pairs with lambda, the column point with phi
    const double2 mine = active ? (Rows ? lam[me] : phi[me]) : make_double2(0.0, 0.0);
    const double k = prm.k, sig = prm.interior ? -1.0 : 1.0;
    const bool dir = prm.dirichlet != 0;
    double ap[3] = {0, 0, 0}, an[3] = {0, 0, 0}, aw = 0, areg = 0, ak = 0;
    for (int base = 0; base < count; base += kStageA) {
        const int ns = min(kStageA, count - base);
        __syncthreads();
        for (int t = threadIdx.x; t < ns; t += kT) {
            sp[t] = load_pt(tube, c, base + t);
            sv[t] = Rows ? phi[base + t] : lam[base + t];
        }
        __syncthreads();
        if (!active) continue;
        for (int s = 0; s < ns; ++s) {
            const int other = base + s;
            const Pt& O = sp[s];
            const Pt& A = Rows ? P : O;  // row point (x*_i)
            const Pt& B = Rows ? O : P;  // column point (y*_j)
            const double2 lm = Rows ? mine : sv[s];
            const double2 ph = Rows ? sv[s] : mine;
            const double ar = -2.0 * (lm.x * ph.x + lm.y * ph.y), ai = -2.0 * (lm.x * ph.y - lm.y * ph.x);
            const double r[3] = {A.sx - B.sx, A.sy - B.sy, A.sz - B.sz};
            const double rho = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            if (rho > prm.tau) {
                const double q[3] = {dir ? B.nx : A.nx, dir ? B.ny : A.ny, dir ? B.nz : A.nz};
                const double sd = r[0] * q[0] + r[1] * q[1] + r[2] * q[2];
                const CD cd = cder(k, rho);
                // dA/dx_i = sig w_j (c' sd r / rho + c q); dA/dy_j = -dA/dx_i
                const double ws = sig * B.w * (Rows ? 1.0 : -1.0);
                for (int a = 0; a < 3; ++a)
                    ap[a] += re_mul(ar, ai, ws * (cd.pr * sd * r[a] / rho + cd.cr * q[a]),
                                    ws * (cd.pi * sd * r[a] / rho + cd.ci * q[a]));
                if (Rows != dir)  // dA/dq = sig w_j c r, q the row normal (Neumann) or column normal (Dirichlet)
                    for (int a = 0; a < 3; ++a)
                        an[a] += re_mul(ar, ai, sig * B.w * cd.cr * r[a], sig * B.w * cd.ci * r[a]);
                if (Rows) {
                    ak += re_mul(ar, ai, sig * B.w * cd.kr * sd, sig * B.w * cd.ki * sd);
                } else {
                    aw += re_mul(ar, ai, sig * cd.cr * sd, sig * cd.ci * sd);
                }
            } else {
                // near field: A_ij = w_j reg_q, real
                const double regq = dir ? B.reg : A.reg;
                if (!Rows) aw += ar * regq;
                if (Rows != dir) areg += ar * B.w;
            }
            (void)other;
        }
    }
    if (!active) return;
    adj[A_X * c + me] += ap[0];
    adj[A_Y * c + me] += ap[1];
    adj[A_Z * c + me] += ap[2];
    adj[A_NX * c + me] += an[0];
    adj[A_NY * c + me] += an[1];
    adj[A_NZ * c + me] += an[2];
    adj[A_W * c + me] += aw;
    adj[A_REG * c + me] += areg;
    adj[A_K * c + me] += ak;
}

// 4. 2 Re[conj(lambda_j) db_j].
__global__ void __launch_bounds__(kT) k_adj_rhs(const double* __restrict__ tube, int count, double k, int kind,
                                                double x0, double y0, double z0, double amp, double dx, double dy,
                                                double dz, const double2* __restrict__ lam, double* __restrict__ adj) {
    const int j = blockIdx.x * kT + threadIdx.x;
    if (j >= count) return;
    const size_t c = (size_t)count;
    const double p[3] = {tube[IBIM_F_SX * c + j], tube[IBIM_F_SY * c + j], tube[IBIM_F_SZ * c + j]};
    const double2 l = lam[j];
    double gr[3], gi[3], kr, ki;
    if (kind == IBIM_RHS_POINT_SOURCE) {
        const double r[3] = {p[0] - x0, p[1] - y0, p[2] - z0};
        const double R = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        const CD cd = cder(k, R);
        for (int a = 0; a < 3; ++a) {
            gr[a] = cd.cr * r[a];
            gi[a] = cd.ci * r[a];
        }
        double sv, cv;
        sincos(k * R, &sv, &cv);
        kr = -sv / (4.0 * kPi);
        ki = cv / (4.0 * kPi);
    } else {  // -amp e^{ik dir.x}
        const double d[3] = {dx, dy, dz};
        const double dp = p[0] * dx + p[1] * dy + p[2] * dz;
        double sv, cv;
        sincos(k * dp, &sv, &cv);
        for (int a = 0; a < 3; ++a) {
            gr[a] = amp * k * d[a] * sv;
            gi[a] = -amp * k * d[a] * cv;
        }
        kr = amp * dp * sv;
        ki = -amp * dp * cv;
    }
    adj[A_X * c + j] += 2.0 * (l.x * gr[0] + l.y * gi[0]);
    adj[A_Y * c + j] += 2.0 * (l.x * gr[1] + l.y * gi[1]);
    adj[A_Z * c + j] += 2.0 * (l.x * gr[2] + l.y * gi[2]);
    adj[A_K * c + j] += 2.0 * (l.x * kr + l.y * ki);
}

// Trilinear sample with its gradient in x (corner weights as interpolate3D.c:49-52).
struct Tri {
    int i, j, k;
    double lx, ly, lz, h;
    __device__ double weight(int a, int b, int cc) const {  // a: x (j), b: y (i), cc: z (k)
        return (a ? lx : h - lx) * (b ? ly : h - ly) * (cc ? lz : h - lz) / (h * h * h);
    }
    __device__ void dweight(int a, int b, int cc, double* g) const {
        const double wx = a ? lx : h - lx, wy = b ? ly : h - ly, wz = cc ? lz : h - lz, h3 = h * h * h;
        g[0] = (a ? 1.0 : -1.0) * wy * wz / h3;
        g[1] = wx * (b ? 1.0 : -1.0) * wz / h3;
        g[2] = wx * wy * (cc ? 1.0 : -1.0) / h3;
    }
};

__device__ __forceinline__ Tri make_tri(const double* x, double lo, double h, double nudge) {
    Tri t;
    t.j = (int)((x[0] + nudge - lo) / h);
    t.i = (int)((x[1] + nudge - lo) / h);
    t.k = (int)((x[2] + nudge - lo) / h);
    t.lx = x[0] - lo - (double)t.j * h;
    t.ly = x[1] - lo - (double)t.i * h;
    t.lz = x[2] - lo - (double)t.k * h;
    t.h = h;
    return t;
}

// 5. Tube data -> grid fields (level d, gradient, H, K), scattered with atomics.
__global__ void __launch_bounds__(kT) k_tube_backprop(const double* __restrict__ u, const double* __restrict__ gx,
                                                      const double* __restrict__ gy, const double* __restrict__ gz,
                                                      const double* __restrict__ mc, const double* __restrict__ gc,
                                                      int n, double lo, double h, ibim_params prm,
                                                      const int* __restrict__ node, int count,
                                                      double* __restrict__ adj, double* __restrict__ gadj) {
    const int p = blockIdx.x * kT + threadIdx.x;
    if (p >= count) return;
    const int S = n + 1;
    const size_t nodes = (size_t)S * S * S, c = (size_t)count;
    double* dbar = gadj;
    double* gbar[3] = {gadj + nodes, gadj + 2 * nodes, gadj + 3 * nodes};
    double* hbar = gadj + 4 * nodes;
    double* kbar = gadj + 5 * nodes;
    const size_t id = (size_t)node[p];
    const int k = (int)(id % S);
    const size_t t = id / S;
    const int j = (int)(t % S), i = (int)(t / S);
    const double z[3] = {lo + h * (double)j, lo + h * (double)i, lo + h * (double)k};
    const double d = u[id];
    const double G[3] = {gx[id], gy[id], gz[id]};
    const double x[3] = {z[0] - d * G[0], z[1] - d * G[1], z[2] - d * G[2]};
    double xb[3] = {adj[A_X * c + p], adj[A_Y * c + p], adj[A_Z * c + p]};
    const double nb[3] = {adj[A_NX * c + p], adj[A_NY * c + p], adj[A_NZ * c + p]};
    const double wb = adj[A_W * c + p], rb = adj[A_REG * c + p];
    const double* grids[3] = {gx, gy, gz};

    // regularisation value from trilinear H*, K* (interpolate3D.c:569-605, with its nudge)
    if (rb != 0.0) {
        const Tri T = make_tri(x, lo, h, 1e-19);
        double Hs = 0, Ks = 0, dH[3] = {0, 0, 0}, dK[3] = {0, 0, 0};
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int cc = 0; cc < 2; ++cc) {
                    const size_t q = node_at(S, T.i + b, T.j + a, T.k + cc);
                    const double w = T.weight(a, b, cc);
                    double dw[3];
                    T.dweight(a, b, cc, dw);
                    Hs += mc[q] * w;
                    Ks += gc[q] * w;
                    for (int e = 0; e < 3; ++e) {
                        dH[e] += mc[q] * dw[e];
                        dK[e] += gc[q] * dw[e];
                    }
                }
        const double tau = prm.tau, sg = prm.interior ? 1.0 : -1.0;
        const double dregH = sg * (1.0 / (4. * kPi * tau) - 10. * tau * (12. * Hs * Hs - Ks) / (512. * kPi) -
                                   50. * tau * Ks / (1536. * kPi) + 2. * prm.k * prm.k * tau / (48. * kPi));
        const double dregK = sg * (10. * tau * Hs / (512. * kPi) - 50. * tau * Hs / (1536. * kPi));
        const double dregk = sg * (4. * prm.k * tau * Hs / (48. * kPi));
        adj[A_K * c + p] += rb * dregk;
        const double Hb = rb * dregH, Kb = rb * dregK;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int cc = 0; cc < 2; ++cc) {
                    const size_t q = node_at(S, T.i + b, T.j + a, T.k + cc);
                    const double w = T.weight(a, b, cc);
                    atomicAdd(&hbar[q], Hb * w);
                    atomicAdd(&kbar[q], Kb * w);
                }
        for (int e = 0; e < 3; ++e) xb[e] += Hb * dH[e] + Kb * dK[e];
    }

    // normal: n = N(N(v)), v the trilinear gradient at x* (interpolate3D.c:41-101 and
    // the caller's second normalisation)
    if (nb[0] != 0.0 || nb[1] != 0.0 || nb[2] != 0.0) {
        const Tri T = make_tri(x, lo, h, 0.0);
        double v[3] = {0, 0, 0}, dv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int cc = 0; cc < 2; ++cc) {
                    const size_t q = node_at(S, T.i + b, T.j + a, T.k + cc);
                    const double w = T.weight(a, b, cc);
                    double dw[3];
                    T.dweight(a, b, cc, dw);
                    for (int m = 0; m < 3; ++m) {
                        v[m] += grids[m][q] * w;
                        for (int e = 0; e < 3; ++e) dv[m][e] += grids[m][q] * dw[e];
                    }
                }
        const double nv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        const double m1[3] = {v[0] / nv, v[1] / nv, v[2] / nv};
        const double nm = sqrt(m1[0] * m1[0] + m1[1] * m1[1] + m1[2] * m1[2]);
        const double nf[3] = {m1[0] / nm, m1[1] / nm, m1[2] / nm};
        const double nbd = nb[0] * nf[0] + nb[1] * nf[1] + nb[2] * nf[2];
        double mb[3], vb[3];
        for (int m = 0; m < 3; ++m) mb[m] = (nb[m] - nbd * nf[m]) / nm;
        const double mbd = mb[0] * m1[0] + mb[1] * m1[1] + mb[2] * m1[2];
        for (int m = 0; m < 3; ++m) vb[m] = (mb[m] - mbd * m1[m]) / nv;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int cc = 0; cc < 2; ++cc) {
                    const size_t q = node_at(S, T.i + b, T.j + a, T.k + cc);
                    const double w = T.weight(a, b, cc);
                    for (int m = 0; m < 3; ++m) atomicAdd(&gbar[m][q], vb[m] * w);
                }
        for (int e = 0; e < 3; ++e) xb[e] += vb[0] * dv[0][e] + vb[1] * dv[1][e] + vb[2] * dv[2][e];
    }

    // weight w = h^3 delta(d) J, J = 1 + 2 d H + d^2 K at the node (large3Dexample2.c:4377-4447)
    double db = 0.0;
    if (wb != 0.0 && fabs(d) <= prm.eps) {
        const double eps = prm.eps, two = prm.one_sided ? 2.0 : 1.0;
        const double del = two * (1.0 / (2.0 * eps)) * (1.0 + cos(kPi * d / eps));
        const double ddel = two * (-kPi / (2.0 * eps * eps)) * sin(kPi * d / eps);
        const double Hn = mc[id], Kn = gc[id], Jn = 1 + 2. * d * Hn + d * d * Kn;
        const double h3 = h * h * h;
        db += wb * h3 * (ddel * Jn + del * (2. * Hn + 2. * d * Kn));
        atomicAdd(&hbar[id], wb * h3 * del * 2. * d);
        atomicAdd(&kbar[id], wb * h3 * del * d * d);
    }
    // projection x* = z - d grad d
    db += -(xb[0] * G[0] + xb[1] * G[1] + xb[2] * G[2]);
    for (int m = 0; m < 3; ++m) atomicAdd(&gbar[m][id], -d * xb[m]);
    atomicAdd(&dbar[id], db);
}

// Dual-number accessors: a line seeds slot m - c + 3, a plane slot (p+2)*5 + q+2.
struct DLine {
    const double* u;
    size_t base, stride;
    int c;
    __device__ Dual<7> operator()(int m) const { return Dual<7>::seed(u[base + (size_t)m * stride], m - c + 3); }
};

struct DPlane {
    const double* u;
    ptrdiff_t o, sa, sb;
    __device__ Dual<25> operator()(int p, int q) const {
        return Dual<25>::seed(u[o + (ptrdiff_t)p * sa + (ptrdiff_t)q * sb], (p + 2) * 5 + (q + 2));
    }
};

// 6. Grid fields -> level-set values.
__global__ void __launch_bounds__(128) k_grid_backprop(const double* __restrict__ u, int n, double h, int order,
                                                       double* __restrict__ gadj) {
    const int S = n + 1;
    const size_t nodes = (size_t)S * S * S;
    const size_t id = (size_t)blockIdx.x * 128 + threadIdx.x;
    if (id >= nodes) return;
    double* dbar = gadj;
    const double gb[3] = {gadj[nodes + id], gadj[2 * nodes + id], gadj[3 * nodes + id]};
    const double Hb = gadj[4 * nodes + id], Kb = gadj[5 * nodes + id];
    if (gb[0] == 0.0 && gb[1] == 0.0 && gb[2] == 0.0 && Hb == 0.0 && Kb == 0.0) return;
    const int k = (int)(id % S);
    const size_t t = id / S;
    const int j = (int)(t % S), i = (int)(t / S);
    const size_t bases[3] = {node_at(S, i, 0, k), node_at(S, 0, j, k), node_at(S, i, j, 0)};
    const size_t strides[3] = {(size_t)S, (size_t)S * S, 1};
    const int cs[3] = {j, i, k};
    // gradient: Godunov(WENO) per axis (distance_application3D.c:84-112)
    for (int ax = 0; ax < 3; ++ax) {
        if (gb[ax] == 0.0) continue;
        const int half = order == 5 ? 3 : 2;
        Dual<7> f[7];
        for (int o = -half; o <= half; ++o)
            f[o + half] = Dual<7>::seed(u[bases[ax] + (size_t)mirror(cs[ax] + o, n) * strides[ax]], o + half);
        Dual<7> dp, dm;
        weno_1d(f, h, order, &dp, &dm);
        const Dual<7> g = godunov_pick(dp, dm, u[id]);
        for (int o = -half; o <= half; ++o)
            if (g.d[o + half] != 0.0)
                atomicAdd(&dbar[bases[ax] + (size_t)mirror(cs[ax] + o, n) * strides[ax]], gb[ax] * g.d[o + half]);
    }
    if (Hb == 0.0 && Kb == 0.0) return;
    // curvatures: H = -(xx+yy+zz)/2, K = xx yy + yy zz + xx zz - (xy^2 + yz^2 + xz^2)
    Dual<7> dd[3];
    for (int ax = 0; ax < 3; ++ax)
        dd[ax] = d2_line(DLine{u, bases[ax], strides[ax], cs[ax]}, cs[ax], n, h, ax == 0 ? 0 : 1);
    const ptrdiff_t o = (ptrdiff_t)id;
    const Dual<25> xy = dmixed(DPlane{u, o, (ptrdiff_t)S * S, (ptrdiff_t)S}, i, j, n, h);
    const Dual<25> xz = dmixed(DPlane{u, o, (ptrdiff_t)S, 1}, j, k, n, h);
    const Dual<25> yz = dmixed(DPlane{u, o, (ptrdiff_t)S * S, 1}, i, k, n, h);
    const double xx = dd[0].v, yy = dd[1].v, zz = dd[2].v;
    const double sb[3] = {-0.5 * Hb + Kb * (yy + zz), -0.5 * Hb + Kb * (xx + zz), -0.5 * Hb + Kb * (xx + yy)};
    for (int ax = 0; ax < 3; ++ax)
        for (int sl = 0; sl < 7; ++sl)
            if (dd[ax].d[sl] != 0.0)
                atomicAdd(&dbar[bases[ax] + (size_t)(cs[ax] + sl - 3) * strides[ax]], sb[ax] * dd[ax].d[sl]);
    const Dual<25>* mx[3] = {&xy, &xz, &yz};
    const double mb[3] = {-2.0 * Kb * xy.v, -2.0 * Kb * xz.v, -2.0 * Kb * yz.v};
    const ptrdiff_t sa[3] = {(ptrdiff_t)S * S, (ptrdiff_t)S, (ptrdiff_t)S * S};
    const ptrdiff_t sbb[3] = {(ptrdiff_t)S, 1, 1};
    for (int pr = 0; pr < 3; ++pr) {
        if (mb[pr] == 0.0) continue;
        for (int p = -2; p <= 2; ++p)
            for (int q = -2; q <= 2; ++q) {
                const double dv = mx[pr]->d[(p + 2) * 5 + (q + 2)];
                if (dv != 0.0) atomicAdd(&dbar[o + p * sa[pr] + q * sbb[pr]], mb[pr] * dv);
            }
    }
}

}  // namespace

int run_gradient(Context* ctx, int col, int mode, const ibim_solve_opts* opts, double* J, double* dlevel,
                 double* dk) {
    const int count = ctx->count, nrecv = ctx->nrecv, S = ctx->n + 1;
    const size_t c = (size_t)count, nodes = (size_t)S * S * S;
    const size_t adj_doubles = (size_t)A_NF * c + 4 * (size_t)nrecv + 4 * c;
    if (ctx->adj.ensure(sizeof(double) * adj_doubles) || ctx->gadj.ensure(sizeof(double) * 6 * nodes))
        return IBIM_ERR_NOMEM;
    double* adj = ctx->adj.as<double>();
    double2* u = (double2*)(adj + (size_t)A_NF * c);
    double2* e = u + nrecv;
    double2* g = e + nrecv;
    double2* lam = g + count;
    const double* xyz = ctx->recv.as<double>();
    const double2* tgt = (const double2*)(xyz + 3 * (size_t)nrecv);
    const double2* phi = ctx->density.as<double2>() + c * col;

    // objective and residual e = u - t
    int rc = launch_eval(ctx, mode, col, nrecv, xyz, (double*)u);
    if (rc) return rc;
    std::vector<double2> hu(nrecv), ht(nrecv);
    IBIM_HIP(hipMemcpyAsync(hu.data(), u, sizeof(double2) * nrecv, hipMemcpyDeviceToHost, ctx->stream));
    IBIM_HIP(hipMemcpyAsync(ht.data(), tgt, sizeof(double2) * nrecv, hipMemcpyDeviceToHost, ctx->stream));
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    double Jv = 0.0;
    for (int q = 0; q < nrecv; ++q) {
        hu[q].x -= ht[q].x;
        hu[q].y -= ht[q].y;
        Jv += hu[q].x * hu[q].x + hu[q].y * hu[q].y;
    }
    if (J) *J = Jv;
    IBIM_HIP(hipMemcpyAsync(e, hu.data(), sizeof(double2) * nrecv, hipMemcpyHostToDevice, ctx->stream));
    const unsigned nb = (unsigned)((count + kT - 1) / kT);
    TimedScope ts(ctx, "adjoint");
    hipLaunchKernelGGL(k_adj_eval, dim3(nb), dim3(kT), 0, ctx->stream, ctx->tube.as<double>(), count, ctx->prm, mode,
                       phi, xyz, e, nrecv, g, adj);
    IBIM_HIP(hipGetLastError());

    // adjoint solve A^H lambda = E^H e
    const int op = fp64_operator(ctx);
    ibim_solve_info info{};
    // IBIM_SOLVE_BLOCK_GMRES: enlarged GMRES on the 16-column A^H passes where
    // they exist (pair storage or dense, traditional tube), else the reference's BiCG
    const bool enlarged = opts->method == IBIM_SOLVE_BLOCK_GMRES &&
                          (ctx->kind == IBIM_TUBE_TRADITIONAL || ctx->mat_precision == IBIM_C128);
    rc = enlarged ? enlarged_solve_op(ctx, op, true, g, lam, opts->maxit, opts->tol, &info)
                  : bicg_solve_op(ctx, op, g, lam, opts->maxit, opts->tol, opts->chunk, &info, nullptr, true, false);
    if (rc) return rc;
    // an unconverged lambda gives a wrong gradient: fail rather than return it
    // (status 4, a zero E^H e, has the exact zero lambda)
    if (info.status != 0 && info.status != 4) {
        set_error("ibim_gradient: the adjoint solve A^H lambda = E^H e did not converge (status %d after %d "
                  "iterations, residual %.3e, tol %.3e)",
                  info.status, info.iterations, info.rel_residual, opts->tol);
        return IBIM_ERR_NOCONV;
    }

    hipLaunchKernelGGL(k_adj_operator<1>, dim3(nb), dim3(kT), 0, ctx->stream, ctx->tube.as<double>(), count,
                       ctx->prm, phi, lam, adj);
    hipLaunchKernelGGL(k_adj_operator<0>, dim3(nb), dim3(kT), 0, ctx->stream, ctx->tube.as<double>(), count,
                       ctx->prm, phi, lam, adj);
    const Context::RhsDesc rd = (size_t)col < ctx->rhs_desc.size() ? ctx->rhs_desc[col] : Context::RhsDesc{};
    if (rd.kind == IBIM_RHS_POINT_SOURCE || rd.kind == IBIM_RHS_PLANE_DIRICHLET)
        hipLaunchKernelGGL(k_adj_rhs, dim3(nb), dim3(kT), 0, ctx->stream, ctx->tube.as<double>(), count, ctx->prm.k,
                           rd.kind, rd.x0[0], rd.x0[1], rd.x0[2], rd.amp, rd.dir[0], rd.dir[1], rd.dir[2], lam, adj);
    else if (rd.kind == IBIM_RHS_PLANE_NEUMANN) {
        set_error("ibim_gradient: Neumann plane-wave data is not differentiated");
        return IBIM_ERR_ARG;
    }
    IBIM_HIP(hipGetLastError());

    IBIM_HIP(hipMemsetAsync(ctx->gadj.ptr, 0, sizeof(double) * 6 * nodes, ctx->stream));
    hipLaunchKernelGGL(k_tube_backprop, dim3(nb), dim3(kT), 0, ctx->stream, ctx->level.as<double>(),
                       ctx->gx.as<double>(), ctx->gy.as<double>(), ctx->gz.as<double>(), ctx->mc.as<double>(),
                       ctx->gc.as<double>(), ctx->n, ctx->lo, ctx->h, ctx->prm, ctx->node.as<int>(), count, adj,
                       ctx->gadj.as<double>());
    hipLaunchKernelGGL(k_grid_backprop, dim3((unsigned)((nodes + 127) / 128)), dim3(128), 0, ctx->stream,
                       ctx->level.as<double>(), ctx->n, ctx->h, ctx->weno_order, ctx->gadj.as<double>());
    IBIM_HIP(hipGetLastError());

    std::vector<double> kparts(c);
    IBIM_HIP(hipMemcpyAsync(kparts.data(), adj + (size_t)A_K * c, sizeof(double) * c, hipMemcpyDeviceToHost,
                            ctx->stream));
    if (dlevel)
        IBIM_HIP(hipMemcpyAsync(dlevel, ctx->gadj.ptr, sizeof(double) * nodes, hipMemcpyDeviceToHost, ctx->stream));
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    double ks = 0.0;
    for (size_t p = 0; p < c; ++p) ks += kparts[p];
    if (dk) *dk = ks;
    return IBIM_OK;
}

}  // namespace ibim
