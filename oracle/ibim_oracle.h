/*
 * ibim_oracle.h -- CPU FP64 restatement of the reference's implicit boundary
 * integral (IBIM) hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library checks the results of the HIP
 * library: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product path never does.
 *
 * Layout conventions (shared with the HIP library):
 *   grid    (n+1)^3 nodes over [lo, lo + n h]^3, node (i, j, k) = (y, x, z)
 *           stored flat at ((i*S)+j)*S+k with S = n+1 (k fastest), matching the
 *           reference's double*** u[i=y][j=x][k=z] (large3Dexample2.c:413-421).
 *   tube    IBO_NF fields of length count, field-major: out[f*count + p]; points
 *           in raster (i, j, k) order like the reference's tube loops.
 *   matrix  row-major split planes are[i*n + j], aim[i*n + j].
 */
#ifndef IBIM_ORACLE_H
#define IBIM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Tube fields (field-major). */
enum {
    IBO_ZX = 0, IBO_ZY, IBO_ZZ,   /* grid node z                          */
    IBO_SX, IBO_SY, IBO_SZ,       /* closest point z* = z - d grad d      */
    IBO_DIST,                     /* d(z)                                 */
    IBO_NX, IBO_NY, IBO_NZ,       /* unit normal at z* (inward, grad d)   */
    IBO_DELTA,                    /* delta_eps(d) or combo weight W       */
    IBO_JAC,                      /* J = 1 + 2 d H + d^2 K at the node    */
    IBO_W,                        /* quadrature weight h^3 delta J (or R) */
    IBO_REG,                      /* near-field regularisation value      */
    IBO_HM, IBO_GK,               /* mean / Gaussian curvature at z*      */
    IBO_NF
};

typedef struct {
    double k;          /* wavenumber (WAVE)                                   */
    double eps;        /* tube half width epsilon                            */
    double tau;        /* near-field radius tau                              */
    double delta_in;   /* inner cut-off of the combo tube (delta0)           */
    double eta;        /* combined-field coupling (DEFAULT_ETA)              */
    double sine_root;  /* inner root of the Sine weight (0.1 in the reference)*/
    int interior;      /* INTERIOR_MODE                                      */
    int dirichlet;     /* DIRICHLET                                          */
    int one_sided;     /* TRAD_ONESIDED / COMBO_ONESIDED                     */
    int combined;      /* exterior Dirichlet: add -i eta S (combined field)  */
} ibo_params;

/* L2: WENO and Godunov gradient, limited second derivatives, curvatures. */
void ibo_weno(const double *f, double dx, int order, double *dplus, double *dminus);
void ibo_weno_axis(const double *u, int n, double h, int i, int j, int k, int axis, int order, double *dplus,
                   double *dminus);
void ibo_gradient(const double *u, int n, double h, int order, double *gx, double *gy, double *gz);
double ibo_d2(const double *u, int n, double h, int i, int j, int k, int axis);
double ibo_dmixed(const double *u, int n, double h, int i, int j, int k, int pair);
void ibo_curvatures(const double *u, int n, double h, double *mc, double *gc);

/* Trilinear sampling at a point. */
int ibo_interp_grad(int n, double h, double lo, const double *gx, const double *gy, const double *gz, const double *p,
                    double *out);
void ibo_interp_curv(int n, double h, double lo, const double *mc, const double *gc, const double *p, double *out);
double ibo_interp_dist(const double *u, int n, double h, double lo, const double *p);

/* L1 primitives. */
double ibo_delta_cos(double eps, double x);
double ibo_sine_weight(double t, double tau, double root);
double ibo_poly_weight(double t, double tau);

/* Tubes. kind 0 = traditional (|d| <= eps), 1 = combo (delta_in <= |d| <= eps). */
int ibo_tube_count(const double *u, int n, const ibo_params *prm, int kind);
int ibo_tube_build(const double *u, const double *gx, const double *gy, const double *gz, const double *mc,
                   const double *gc, int n, double h, double lo, const ibo_params *prm, int kind, int count,
                   int *node, double *out);

/* Kernel rows [row0, row0 + nrows) of the count x count matrix. */
void ibo_kernel_trad_rows(int count, const double *tube, const ibo_params *prm, int row0, int nrows, double *are,
                          double *aim);
int ibo_kernel_combo_rows(int count, const double *tube, const ibo_params *prm, int row0, int nrows, double *are,
                          double *aim);

/* L4: complex BiCG of the reference (maxit iterations, stop on tol/NaN/stagnation).
 * hist (maxit entries, may be NULL) receives the relative residual per iteration.
 * Returns the iteration count. */
int ibo_bicg(int n, const double *are, const double *aim, const double *bre, const double *bim, double *xre,
             double *xim, int maxit, double tol, double *hist);
/* The same iteration with OpenMP matrix sweeps (row-blocked A p, column-order
 * A^H p~, as in the reference's parallel solver). */
int ibo_bicg_omp(int n, const double *are, const double *aim, const double *bre, const double *bim, double *xre,
                 double *xim, int maxit, double tol, double *hist);

/* One reference-style sweep over a block of nrows x ncols rows of A (split
 * planes): v = A p over the rows, w += A^H q restricted to the block (the loop
 * nests of large3Dexample2.c:3262-3309, OpenMP rows / columns). For timing. */
void ibo_sweep_rows(int nrows, int ncols, const double *are, const double *aim, const double *pr, const double *pi,
                    const double *qr, const double *qi, double *vr, double *vi, double *wr, double *wi);

/* Right-hand sides at the tube's closest points. */
void ibo_rhs_point_source(int count, const double *tube, double k, const double *x0, double *bre, double *bim);
void ibo_rhs_plane_neumann(int count, const double *tube, const double *interp_n, double k, double amp,
                           const double *dir, double *bre, double *bim);
void ibo_rhs_plane_dirichlet(int count, const double *tube, double k, double amp, const double *dir, double *bre,
                             double *bim);

/* Field evaluation at receivers (mode 0 single layer, 1 double layer, 2 combo,
 * 3 double layer - i eta single layer: the combined-field potential). */
void ibo_eval(int count, const double *tube, const ibo_params *prm, const double *phre, const double *phim, int nrecv,
              const double *xyz, int mode, double *ure, double *uim);

#ifdef __cplusplus
}
#endif
#endif
