/*
 * ibim.h -- C ABI of libibim.so, the MI355X (gfx950) implementation of the
 * implicit boundary integral (IBIM) hot path of moorhsum/Helmholtz-3D-torrone.
 *
 * Two layers:
 *
 *  1. Drop-in entry points with the reference's names and argument lists
 *     (HelmholtzKernel.h:20-39, BICGSTAB.h:15-17, weno.h:26-31,
 *     distance_application3D.h:9-12). They take the reference's host data
 *     structures (double*** grids, double** row-pointer matrices), run the work
 *     on the GPU and return the same outputs. libibim.so exports each of them
 *     twice: with C linkage (declared here, for C / ctypes / FFI callers) and with
 *     the C++ linkage the reference's own build produces (its .c files are
 *     compiled as C++, Makefile:1,10-11), so a program built from the
 *     reference's headers links against libibim.so unchanged. Like the
 *     reference, these read the caller-defined globals N, SIZE, H, HCUBED, WAVE
 *     (HelmholtzKernel.h:16-18); libibim.so carries weak definitions that a
 *     caller's own definitions override.
 *
 *  2. A device-resident handle API (ibim_*) for callers that keep the problem
 *     in HBM: level set -> geometry -> tube -> operator -> right-hand sides ->
 *     solve -> field evaluation, with every array resident on the GPU between
 *     calls. The reference has no such API (its driver, large3Dexample2.c:97-1215,
 *     calls file-static functions); it is what the drop-in entries are built on.
 *
 * All functions return 0 (IBIM_OK) on success and a negative IBIM_ERR_* code
 * otherwise -- where the reference prints and calls exit(0) -- with a message
 * available from ibim_last_error(). Calls on one handle are not thread-safe;
 * each handle owns one HIP stream on the device current at ibim_create().
 */
#ifndef IBIM_H
#define IBIM_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- errors --- */
#define IBIM_OK 0
#define IBIM_ERR_ARG (-1)        /* bad argument or call order                     */
#define IBIM_ERR_HIP (-2)        /* HIP runtime error                              */
#define IBIM_ERR_SIZE (-3)       /* size != tube count (HelmholtzKernel.c:113-117) */
#define IBIM_ERR_GEOMETRY (-4)   /* closest point outside the grid / NaN
                                    (interpolate3D.c:31-36, 81-85)                 */
#define IBIM_ERR_NOMEM (-5)      /* device allocation failed                       */
#define IBIM_ERR_NODEVICE (-6)   /* no HIP device                                  */
#define IBIM_ERR_NOCONV (-7)     /* a solve inside the call (ibim_gradient's       */
                                 /* adjoint solve) did not converge                */

/* ------------------------------------------------------------ data model --- */
/* Tube fields, field-major (field f of point p at fields[f * count + p]). */
enum ibim_field {
    IBIM_F_ZX = 0, IBIM_F_ZY, IBIM_F_ZZ, /* grid node z                        */
    IBIM_F_SX, IBIM_F_SY, IBIM_F_SZ,     /* closest point z* = z - d grad d    */
    IBIM_F_DIST,                         /* d(z)                               */
    IBIM_F_NX, IBIM_F_NY, IBIM_F_NZ,     /* unit normal at z* (grad d, inward) */
    IBIM_F_DELTA,                        /* delta_eps(d), or the combo weight W */
    IBIM_F_JAC,                          /* J = 1 + 2 d H + d^2 K at z         */
    IBIM_F_W,                            /* quadrature weight h^3 delta J, or R */
    IBIM_F_REG,                          /* near-field value                   */
    IBIM_F_HM, IBIM_F_GK,                /* mean / Gaussian curvature at z*    */
    IBIM_NF
};

/* Run-time replacement of the reference's compile-time switches (PDE.h:6-30,
 * functions.h:20-30) plus the tube / regularisation sizes it passes around. */
typedef struct ibim_params {
    double k;          /* wavenumber (WAVE)                                   */
    double eps;        /* tube half width epsilon (FACTOR * H)                */
    double tau;        /* near-field radius (TAU_FACTOR * H)                  */
    double delta_in;   /* inner cut-off of the combo tube (delta0)            */
    double eta;        /* combined-field coupling (DEFAULT_ETA)               */
    double sine_root;  /* inner root of the Sine weight; 0.1 = the reference  */
    int interior;      /* INTERIOR_MODE                                       */
    int dirichlet;     /* DIRICHLET: double-layer normal at the source point  */
    int one_sided;     /* TRAD_ONESIDED / COMBO_ONESIDED                      */
    int combined;      /* traditional tube, exterior Dirichlet only: the      */
                       /* combined-field potential u = (D - i eta S) phi,     */
                       /* A = 1/2 I + K - i eta S (Brakhage-Werner), with the */
                       /* coupling eta above; 0 = the reference's double      */
                       /* layer 1/2 I + K                                     */
} ibim_params;

#define IBIM_TUBE_TRADITIONAL 0 /* |d| <= eps, cosine delta (large3Dexample2.c:4340-4452) */
#define IBIM_TUBE_COMBO 1       /* delta_in <= |d| <= eps, Sine weight (HelmholtzKernel.c:223-400) */

#define IBIM_C128 0 /* complex FP64 storage, 16 B per entry */
#define IBIM_C64 1  /* complex FP32 storage,  8 B per entry */
/* Traditional tube only: the radial factor sigma (c1 + i c2)(rho_ij) of each
 * unordered pair i < j stored once (complex FP64 / FP32), the entries
 * A_ij = F_ij w_j r_ij.n_q rebuilt from the points in every sweep, near-field
 * pairs (rho <= tau) in per-point neighbour lists: half the bytes of
 * IBIM_C128 / IBIM_C64 per sweep. The batched solvers and ibim_apply_batch
 * apply it 16 right-hand sides at a time, A P and A^H Q in one pass. */
#define IBIM_C128_PAIR 2
#define IBIM_C64_PAIR 3
/* Complex64 pairs for the BiCG iterations with a complex128 copy of the same
 * factors beside them: the FP64 operator of IBIM_SOLVE_RELIABLE's residual
 * replacements, IBIM_SOLVE_REFINED's residuals and ibim_gradient's adjoint
 * solve (1.5x the HBM of IBIM_C128_PAIR). */
#define IBIM_MIXED_PAIR 4

#define IBIM_RHS_POINT_SOURCE 0    /* Phi(x*, x0) = e^{ik|x*-x0|} / (4 pi |x*-x0|)        */
#define IBIM_RHS_PLANE_NEUMANN 1   /* amp d/dn e^{ik dir.x} at x* (cal_partialu*)        */
#define IBIM_RHS_PLANE_DIRICHLET 2 /* -amp e^{ik dir.x*} (sound-soft data)               */

#define IBIM_EVAL_SINGLE 0 /* sum_j -Phi(x,y*_j) phi_j w_j   (NeumannTest, large3Dexample2.c:1810-1824) */
#define IBIM_EVAL_DOUBLE 1 /* sum_j w_j grad_x Phi(x,y*_j).n_j phi_j (the Dirichlet kernel's layer)        */
#define IBIM_EVAL_COMBO 2  /* Combo_NeumannTest sample kernel (large3Dexample2.c:1980-2058)                  */
#define IBIM_EVAL_COMBINED 3 /* sum_j w_j [grad_x Phi(x,y*_j).n_j - i eta Phi(x,y*_j)] phi_j: the layer of the
                                combined-field Dirichlet kernel (ibim_params.combined)                          */

#define IBIM_SOLVE_BICG 0    /* the reference's complex BiCG on the stored operator      */
#define IBIM_SOLVE_REFINED 1 /* FP64 iterative refinement around a complex-FP32 BiCG      */
#define IBIM_SOLVE_RELIABLE 2 /* complex-FP32 BiCG with FP64 residual replacement: every */
                              /* inner_tol drop of the residual (default 1e-2) and at    */
                              /* convergence r = b - A64 x; no restarts                  */
#define IBIM_SOLVE_GMRES 3    /* FP64 iterative refinement around GMRES on the stored    */
                              /* operator (complex64 storage: the mixed-precision path;  */
                              /* complex128: FP64 GMRES; nothing stored: GMRES on the    */
                              /* mixed matrix-free operator); one product A p per        */
                              /* iteration, a restart every `chunk` products (0: 256;    */
                              /* 16-column batches, batch = 0 with several right-hand    */
                              /* sides: 128). No counterpart in the reference (BiCG only) */
#define IBIM_SOLVE_BLOCK_GMRES 4 /* the same refinement around block GMRES: one Krylov     */
                                 /* space per batch of 16 right-hand sides (batch = 0),   */
                                 /* a restart every `chunk` passes (0: 64); one column at */
                                 /* a time: enlarged GMRES (the residual split over 16    */
                                 /* pieces of the surface, one block space). No           */
                                 /* counterpart in the reference                          */

typedef struct ibim_solve_opts {
    int method;        /* IBIM_SOLVE_*                                        */
    int maxit;         /* BiCG iteration cap (1000 / 2000 in the reference)   */
    double tol;        /* relative residual target (DEFAULTACC = 1e-7)        */
    double inner_tol;  /* refinement: first inner FP32 BiCG relative tolerance; */
                       /* later ones use max(inner_tol, 0.3 tol / residual)      */
    int max_outer;     /* refinement: outer FP64 steps                        */
    int chunk;         /* iterations launched between host checks (0 = auto)  */
    int batch;         /* several right-hand sides: 0 = 16 per pass over the   */
                       /* operator (BiCG, refinement), 1 = one at a time       */
} ibim_solve_opts;

typedef struct ibim_solve_info {
    int iterations;      /* BiCG iterations, summed over refinement steps and,
                          * with several right-hand sides, over the columns */
    int outer;           /* refinement steps                                 */
    double rel_residual; /* last ||r|| / ||b|| (recursive residual for BiCG, */
                         /* true FP64 residual for refinement)               */
    int status;          /* 0 converged, 1 iteration cap, 2 NaN, 3 stagnation, 4 zero RHS */
} ibim_solve_info;

/* ------------------------------------------------------ handle lifecycle --- */
typedef struct ibim_ctx ibim_ctx;

int ibim_device_count(void);
const char *ibim_last_error(void);
int ibim_create(ibim_ctx **out);
void ibim_destroy(ibim_ctx *ctx);
int ibim_sync(ibim_ctx *ctx);

/* ------------------------------------------------------------- geometry --- */
/* Level set on the (n+1)^3 grid over [lo, lo + n h]^3, layout [i=y][j=x][k=z]
 * (large3Dexample2.c:405-424). _host copies from host memory, _device from a
 * device pointer on the handle's device. */
int ibim_set_level_host(ibim_ctx *ctx, int n, double lo, double h, const double *level);
int ibim_set_level_device(ibim_ctx *ctx, int n, double lo, double h, const double *level_dev);
/* Signed-distance level set (positive inside) of an analytic surface, built on
 * the GPU straight into the handle's grid. The reference reads a precomputed
 * distance file instead (large3Dexample2.c:405-424); its formulas assume
 * |grad d| = 1, which these are. */
#define IBIM_SHAPE_SPHERE 0    /* radius a                                            */
#define IBIM_SHAPE_TORUS 1     /* major radius a, minor radius b, axis z              */
#define IBIM_SHAPE_ELLIPSOID 2 /* semi-axes a, b, c                                    */
#define IBIM_SHAPE_STAR 3      /* r(theta, phi) = a + b sin(m theta) cos(n phi)  (variant 0)
                                  r(theta, phi) = a + b cos(m theta) sin(n phi)  (variant 1) */
typedef struct ibim_shape {
    int kind, variant, m, n;
    double center[3];
    double a, b, c;
} ibim_shape;
int ibim_set_level_shape(ibim_ctx *ctx, int n, double lo, double h, const ibim_shape *shape);
/* Replaces the grid's level set, which need not be a distance function
 * (|grad| != 1), by the signed distance to its own zero level set (the zero
 * set of its Catmull-Rom tricubic interpolant; closest points by projection
 * and Chopp's iteration), keeping its sign; nodes farther than band grid
 * spacings get +-band h. The IBIM formulas assume a distance function
 * (x* = x - d grad d): run this before ibim_geometry on any other level set,
 * with band above the tube half-width plus the 3-node WENO stencil. No
 * counterpart in the reference, which reads a precomputed distance file
 * (large3Dexample2.c:405-424) and leaves its Godunov Hamiltonian
 * (finite_difference3D.c:5-27) unused. */
int ibim_redistance(ibim_ctx *ctx, double band);
/* WENO-3/5 Godunov gradient and curvatures on every node
 * (ASSIGN_GRADIENT distance_application3D.c:56-141; BothCurvs :35-54). */
int ibim_geometry(ibim_ctx *ctx, int weno_order);
/* Host copy of the level set. */
int ibim_get_level(ibim_ctx *ctx, double *level);
/* Host copies of the grid fields (any pointer may be NULL). */
int ibim_get_grid(ibim_ctx *ctx, double *gx, double *gy, double *gz, double *mc, double *gc);
/* Supply gradient and curvatures computed elsewhere (host memory). */
int ibim_set_grid_fields_host(ibim_ctx *ctx, const double *gx, const double *gy, const double *gz, const double *mc,
                              const double *gc);

/* ----------------------------------------------------------------- tube --- */
/* Extract the tube (raster order), project, weight and regularise on the GPU. */
int ibim_tube(ibim_ctx *ctx, const ibim_params *prm, int kind, int *count);
int ibim_get_tube(ibim_ctx *ctx, int *node, double *fields);
/* Device pointer to the field-major tube fields (IBIM_NF x count doubles). */
const double *ibim_tube_device(ibim_ctx *ctx);

/* ------------------------------------------------------------- operator --- */
/* Dense count x count operator in HBM (traditional kernel for
 * IBIM_TUBE_TRADITIONAL, combined-field kernel for IBIM_TUBE_COMBO), or the
 * pair storage of the traditional kernel (IBIM_C128_PAIR / IBIM_C64_PAIR). */
int ibim_assemble(ibim_ctx *ctx, int precision);
/* Host copy of rows [row0, row0 + nrows) as split planes (row-major, count wide);
 * for pair storage the entries as a sweep rebuilds them. */
int ibim_get_rows(ibim_ctx *ctx, int row0, int nrows, double *re, double *im);
/* Upload a dense operator given as split planes (row-major, count x count). */
int ibim_set_matrix_host(ibim_ctx *ctx, int count, const double *re, const double *im, int precision);
/* Near-field entry count of the last assembly (the combo kernel's regcount). */
int ibim_regcount(ibim_ctx *ctx);

/* One application y = A p, z = A^H q (host vectors of length count, split
 * re / im) of the stored operator (op IBIM_OP_STORED) or of the entries
 * recomputed on the fly in FP64 (IBIM_OP_FREE64) or with FP32 sine and cosine
 * of an FP64 phase (IBIM_OP_FREE32, traditional kernel only; the combined-field
 * kernel falls back to FP64). On the traditional tube both recompute one
 * radial factor per unordered pair, IBIM_OP_FREE64 with the operations of the
 * complex128 pair storage (the same bits). With zre and zim both NULL only
 * y = A p is formed (pair storage: a sweep without the A^H products, y the
 * same bits). */
#define IBIM_OP_STORED 0
#define IBIM_OP_FREE64 1
#define IBIM_OP_FREE32 2
int ibim_apply(ibim_ctx *ctx, int op, const double *pre, const double *pim, const double *qre, const double *qim,
               double *yre, double *yim, double *zre, double *zim);
/* Y = A P (adjoint = 0) or Y = A^H P (adjoint = 1) for nvec <= 16 vectors at
 * once (host arrays, column-major, count per column, split re / im), on the
 * FP64 matrix cores: the product the batched BiCG applies per pass. Matrix-free
 * ops need the traditional tube; IBIM_OP_FREE64 computes each factor where
 * complex128 pair storage would read it (the same bits as that storage) after
 * building a point table and near lists on first use, kept until the tube
 * changes. No counterpart in the reference (it solves one right-hand side at a
 * time). */
int ibim_apply_batch(ibim_ctx *ctx, int op, int adjoint, int nvec, const double *pre, const double *pim, double *yre,
                     double *yim);
/* ibim_apply on device vectors (count complex FP64 values each, interleaved
 * re / im, on the handle's device) that allocates nothing: its partial planes
 * and flags live in a 64 MiB device array per GPU that the library holds until
 * the process ends, which bounds count at about 26,000 points (IBIM_ERR_SIZE
 * beyond). The matrix-free ops need their tables built beforehand (one
 * ibim_apply, or a pair assembly). Returns after the product is complete. All
 * calls on one GPU share the array: calls that overlap on a GPU race. No
 * counterpart in the reference; for callers that cannot allocate device memory
 * (HIP graph capture, tight serving loops). */
int ibim_apply_noalloc(ibim_ctx *ctx, int op, const void *p_dev, const void *q_dev, void *y_dev, void *z_dev);

/* Progress of long solves: called from ibim_solve / ibim_gradient at host
 * checks, at most every 5 s, with the iteration and the largest running
 * residual. No counterpart in the reference (it prints each iteration). */
typedef void (*ibim_progress_fn)(const char *what, int iteration, double residual, void *user);
int ibim_set_progress(ibim_ctx *ctx, ibim_progress_fn fn, void *user);

/* ------------------------------------------------- right-hand sides/solve --- */
int ibim_set_nrhs(ibim_ctx *ctx, int nrhs);
int ibim_rhs(ibim_ctx *ctx, int kind, const double *x0, double amp, const double *dir, int col);
int ibim_set_rhs_host(ibim_ctx *ctx, int col, const double *bre, const double *bim);
int ibim_get_rhs(ibim_ctx *ctx, int col, double *bre, double *bim);
int ibim_solve(ibim_ctx *ctx, const ibim_solve_opts *opts, ibim_solve_info *info, double *hist);
int ibim_get_density(ibim_ctx *ctx, int col, double *re, double *im);
/* Every right-hand side of nctx problems solved together, 16 columns per
 * batch, by opts->method IBIM_SOLVE_BICG or IBIM_SOLVE_RELIABLE: each BiCG
 * iteration applies the operators of all running batches in ONE persistent
 * launch whose workgroups take 64-row blocks from a queue (largest problems
 * first), then runs each batch's near field and vector updates. The problems
 * need pair storage of their tubes (IBIM_C128_PAIR, IBIM_C64_PAIR or
 * IBIM_MIXED_PAIR; the reliable solver: complex64 pairs), one device, the same
 * kernel (Dirichlet, Neumann or combined) and ibim_set_fp32_products setting.
 * Each problem's densities and infos[i] are what its own batched ibim_solve
 * with the same opts gives (host checks every opts->chunk iterations; 0: 2
 * for the reliable solver, and for BiCG 2 when the largest problem has more
 * than 20,000 points, else 16). The work runs on ctxs[0]'s stream; no other
 * thread may use the handles meanwhile. With ibim_set_graphs on ctxs[0] the
 * iterations replay captured HIP graphs (the same bits). No counterpart in the reference (one
 * surface and one right-hand side per run). */
int ibim_solve_group(ibim_ctx *const *ctxs, int nctx, const ibim_solve_opts *opts, ibim_solve_info *infos);
int ibim_set_density_host(ibim_ctx *ctx, int col, const double *re, const double *im);

/* ------------------------------------------------------------ evaluation --- */
/* u at nrecv receivers xyz (host, 3 per receiver) from density column col. */
int ibim_evaluate(ibim_ctx *ctx, int mode, int col, int nrecv, const double *xyz, double *ure, double *uim);

/* -------------------------------------------------------------- adjoint --- */
/* Misfit objective J = sum_r |u(x_r) - t_r|^2 of the field of density column col
 * (mode IBIM_EVAL_SINGLE or IBIM_EVAL_DOUBLE) at nrecv receivers with targets t,
 * and its gradient with respect to every level-set grid value and the
 * wavenumber k. The adjoint solves A^H lambda = E^H (u - t) with the same BiCG
 * (opts->method IBIM_SOLVE_BLOCK_GMRES: enlarged GMRES on the 16-column A^H
 * passes, traditional tube or dense complex128), differentiates the operator, evaluator and right-hand side entries exactly on
 * the fly (no stored derivative matrices), then back-propagates through the
 * quadrature weights, projection, trilinear sampling, WENO / Godunov gradient
 * and the limited curvature stencils (branch choices held fixed). Requires a
 * traditional tube, a solved column, and a right-hand side made by ibim_rhs
 * (point source or Dirichlet plane wave) or set from the host. dlevel: host
 * array of (n+1)^3 doubles (may be NULL). */
int ibim_set_receivers(ibim_ctx *ctx, int nrecv, const double *xyz, const double *tre, const double *tim);
int ibim_gradient(ibim_ctx *ctx, int col, int mode, const ibim_solve_opts *opts, double *J, double *dlevel,
                  double *dk);

/* ------------------------------------------------------------- profiling --- */
/* When on, kernels are bracketed by hipEvents on the handle's stream. */
int ibim_set_profiling(ibim_ctx *ctx, int on);
/* When on (default off), BiCG solves, single-vector and batched, capture one
 * chunk of iterations (ibim_solve_opts.chunk: the products, inner products and
 * vector updates) as a HIP graph once per solve and replay it for every full
 * chunk (batched: from the second chunk on); the results are the same bits as
 * the launches one by one. Ignored while profiling is on. No counterpart in
 * the reference (launch overhead). */
int ibim_set_graphs(ibim_ctx *ctx, int on);
/* When on (default off), the 16-column passes over complex64 pair storage
 * (IBIM_C64_PAIR, IBIM_MIXED_PAIR: the mixed solvers' inner operator, and
 * ibim_apply_batch on it) multiply on the FP32 matrix cores: entries formed in
 * FP32 from the stored factors (the separation of the points in FP64), FP32
 * copies of the vectors, products summed in FP32 over 64 columns and in FP64
 * beyond. The operator then depends on the vector at the FP32 rounding level,
 * which the mixed solvers' FP64 residuals correct. No counterpart in the
 * reference. */
int ibim_set_fp32_products(ibim_ctx *ctx, int on);
/* The FP64 matrix-free operator (a traditional tube with no operator stored:
 * OP_FREE64, BiCG without an assembly, the mixed solvers' FP64 residuals)
 * keeps the complex128 pairs of its first row blocks, up to max_bytes of HBM,
 * and computes the others' factors in every sweep, as the reference keeps the
 * operator rows that fit in its cache and recomputes the rest (FASTSIZEBOUND,
 * large3Dexample2.c:27,3905-3926). The cached pairs are the bits the sweeps
 * compute, so results do not change. Built on first use after each tube; 0
 * (the default) frees it. */
int ibim_set_free_cache(ibim_ctx *ctx, double max_bytes);
/* Total milliseconds and launch count of a kernel class ("matvec", "assemble", ...). */
int ibim_kernel_time(ibim_ctx *ctx, const char *name, double *ms_total, long *launches);
int ibim_reset_timers(ibim_ctx *ctx);

/* ------------------------------------------------- reference file formats --- */
/* Host-only readers / writers of the reference driver's text files, so its
 * inputs and outputs interoperate with this library:
 * distance<N>.txt (large3Dexample2.c:405-424), Rel_bN<N>k<k>.txt and
 * Rel_bimgN<N>k<k>.txt (print_f :1235-1335, read at :3108-3127),
 * Rel_DensityN<N>k<k>.txt and Rel_DensityimgN<N>k<k>.txt (:3446-3474, :858-866). */
int ibim_read_distance_file(const char *path, int n, double *level);
int ibim_write_distance_file(const char *path, int n, const double *level);
int ibim_read_rhs_text(const char *path_re, const char *path_im, int count, double *bre, double *bim);
int ibim_write_rhs_text(const char *path_re, const char *path_im, int count, const double *bre, const double *bim);
int ibim_read_density_text(const char *path_re, const char *path_im, int count, double *re, double *im);
int ibim_write_density_text(const char *path_re, const char *path_im, int count, const double *re, const double *im);

/* ------------------------------------------------ reference drop-in layer --- */
/* weno.h:26-31 */
void WENO_DX(double ***uvalue, const int i, const int j, const int k, const int N, const double H, double *ux_p,
             double *ux_n, const int order);
void WENO_DY(double ***uvalue, const int i, const int j, const int k, const int N, const double H, double *ux_p,
             double *ux_n, const int order);
void WENO_DZ(double ***uvalue, const int i, const int j, const int k, const int N, const double H, double *ux_p,
             double *ux_n, const int order);
void weno(double *dfp, double *dfn, double fi[], const double dx, const int order);

/* distance_application3D.h:9-12 (ASSIGN_GRADIENT runs on the GPU) */
double GaussianCurv(double ***dist, int i, int j, int k, int N, double H);
double MeanCurv(double ***dist, int i, int j, int k, int N, double H);
int BothCurvs(double ***dist, int i, int j, int k, int N, double H, double *meancurv, double *gausscurv);
int ASSIGN_GRADIENT(int N, double H, double ***dist, double ***graddx, double ***graddy, double ***graddz,
                    int weno_mode);

/* BICGSTAB.h:15-17 (complex BiCG on the GPU; chunk is accepted and unused) */
int CompBiCG(int size, double *density, double *densityimg, double **Kernel, double **Kernelimg, double *b,
             double *bimg);
int omp_CompBiCG(int size, double *density, double *densityimg, double **Kernel, double **Kernelimg, double *b,
                 double *bimg, int chunk);

/* HelmholtzKernel.h:20-39 (tube build and N^2 fill on the GPU) */
int new_HelmholtzKernel(int size, double delta, double epsilon, double *zxlist, double *zylist, double *zzlist,
                        double *zxstarlist, double *zystarlist, double *zzstarlist, double *dist, double *gradx,
                        double *grady, double *gradz, double *curv, double *Jacobian, double *PolyW, double *SineW,
                        double *PolyResult, double *SineResult, double **PolyKernelH, double **PolyKernelHimg,
                        double **SineKernelH, double **SineKernelHimg);
int new_HelmholtzKernelCombo(int size, double delta, double epsilon, double ***level, double ***leveldx,
                             double ***leveldy, double ***leveldz, double ***levelmcurv, double ***levelgcurv,
                             double *zxlist, double *zylist, double *zzlist, double *zxstarlist, double *zystarlist,
                             double *zzstarlist, double *dist, double *gradx, double *grady, double *gradz,
                             double *Jacobian, double *regfactor, double *regfactorimg, double *PolyW, double *SineW,
                             double *PolyResult, double *SineResult, double **PolyKernelH, double **PolyKernelHimg,
                             double **SineKernelH, double **SineKernelHimg, double wavenum, double eta);
int DoubleLayerKernel(int size, double epsilon, double ***level, double ***leveldx, double ***leveldy,
                      double ***leveldz, double ***levelmcurv, double ***levelgcurv, double *zxlist, double *zylist,
                      double *zzlist, double *zxstarlist, double *zystarlist, double *zzstarlist, double *dist,
                      double *gradx, double *grady, double *gradz, double *delta, double *Jacobian, double *regfactor,
                      double *regfactorimg, double *result, double **Kernel, double **Kernelimg, double wavenum);

#ifdef __cplusplus
}
#endif
#endif /* IBIM_H */
