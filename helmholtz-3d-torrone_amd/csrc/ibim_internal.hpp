// ibim_internal.hpp -- device-side state of one IBIM problem and the launch
// helpers shared by the translation units of libibim.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <deque>
#include <string>
#include <vector>

#include "../../include/ibim.h"

namespace ibim {

void set_error(const char* fmt, ...);

#define IBIM_HIP(call)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            ::ibim::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                              __LINE__);                                                        \
            return IBIM_ERR_HIP;                                                                \
        }                                                                                       \
    } while (0)

// A device buffer that only grows.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t want);
    void release();
    template <class T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

// Per-iteration scalars of the device-resident BiCG (one record per iteration).
struct BicgScalars {
    double rho_re, rho_im;    // <r, r~> after the iteration
    double alpha_re, alpha_im;
    double beta_re, beta_im;
    double err;               // ||r|| / ||b||
    int stop;                 // 1 once a stop rule fired at this iteration
    int pad;
};

// Timing record of one kernel class (hipEvent pairs on the context stream).
struct KernelTimer {
    std::string name;
    std::vector<hipEvent_t> begin, end;
    double ms_total = 0.0;
    long launches = 0;
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;

    // grid
    int n = 0;
    double lo = -2.0, h = 0.0;
    DevBuf level, gx, gy, gz, mc, gc;
    bool have_geometry = false;
    int weno_order = 0;  // order of the WENO gradient on the grid, 0 when set from the host

    // tube
    ibim_params prm{};
    int kind = 0;
    int count = 0;
    // count and the tube fields describe the current operator (false after
    // ibim_set_matrix_host or a drop-in solve on a host matrix, which give an
    // operator without points)
    bool have_tube = false;
    DevBuf node, tube, blockcnt, err_flag;

    // operator
    int mat_precision = -1;  // IBIM_C128 / IBIM_C64 (element type), -1 when not assembled
    bool mat_pair = false;   // pair storage (ibim_pair.hip) instead of dense rows
    bool mat_pair64 = false; // complex64 pairs with a complex128 copy in mat64 (IBIM_MIXED_PAIR)
    size_t ld = 0;
    DevBuf mat, mat64;
    int regcount = 0;
    // pair storage: tile table (row block, tile), first tile of each row block,
    // point table, near-neighbour lists (offsets, indices, build cursors)
    int pair_ntiles = 0;
    DevBuf pair_tiles, pair_blk, pair_pts, near_off, near_idx, near_cur;
    // the five tables above match the current tube and parameters (built by a
    // pair assembly, or alone for the matrix-free pair pass)
    bool pair_aux = false;
    // tuned workgroup shape (waves) of the 16-column pair passes per factor
    // source (complex64, complex128, complex128 copy, computed FP64, computed
    // mixed) and product kind, 0 until timed; reset with every assembly
    int batch_waves[15] = {};
    // untuned 16-column passes take 64 rows per workgroup without timing: set
    // around passes a solve applies once or twice (refinement residuals), where
    // timing the shapes would cost more than any shape could save
    bool batch_untimed = false;

    // the FP64 matrix-free operator's cache (ibim_set_free_cache): complex128
    // pairs of the first free_cached_blocks row blocks (free_cached_tiles
    // tiles), at most free_cache_max bytes; valid for the current tube
    size_t free_cache_max = 0;
    DevBuf free_cache;
    int free_cached_blocks = 0, free_cached_tiles = 0;
    bool free_cache_valid = false;

    // 16-column passes over complex64 pairs multiply on the FP32 matrix cores
    // (ibim_set_fp32_products)
    bool fp32_products = false;

    // right-hand sides and densities: nrhs columns of length count (complex FP64)
    int nrhs = 0;
    DevBuf rhs, density;

    // how each right-hand-side column was made (the adjoint needs db / dgeometry)
    struct RhsDesc {
        int kind = -1;  // IBIM_RHS_*, or -1 when set from the host
        double x0[3] = {0, 0, 0}, amp = 1.0, dir[3] = {0, 0, 0};
    };
    std::vector<RhsDesc> rhs_desc;

    // receivers and targets of the misfit objective, adjoint workspace
    int nrecv = 0;
    DevBuf recv, adj, gadj;

    // solver workspace
    DevBuf work, partial, scal, done;
    DevBuf xfer;    // receivers and other small transfers
    DevBuf refine;  // refinement vectors
    DevBuf batch;   // batched (16 right-hand sides) BiCG workspace
    DevBuf group;   // ibim_solve_group with this context first: every member's workspace
    int last_iters = 0;

    // progress reports from long solves (at host checks, at most every few seconds)
    ibim_progress_fn progress = nullptr;
    void* progress_user = nullptr;
    double progress_last = 0.0;

    // single-vector BiCG: replay one captured chunk of iterations as a HIP graph
    // (ibim_set_graphs; not while profiling)
    bool use_graphs = false;

    // profiling
    bool profile = false;
    std::deque<KernelTimer> timers;  // stable addresses for open TimedScopes

    // scratch host copies
    std::vector<double> host_tmp;

    // pinned host staging of operator transfers: two blocks in flight, each
    // with the event of the copy that last used it
    void* stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    size_t stage_bytes = 0;
};

// Host rows of the operator, real and imaginary parts apart: row pointers
// (the reference's double**) or contiguous rows at a stride.
struct HostRows {
    double* const* re_rows = nullptr;
    double* const* im_rows = nullptr;
    double* re = nullptr;
    double* im = nullptr;
    size_t stride = 0;
    double* re_row(int r) const { return re_rows ? re_rows[r] : re + stride * (size_t)r; }
    double* im_row(int r) const { return im_rows ? im_rows[r] : im + stride * (size_t)r; }
};

// Timed launch on ctx->stream: records hipEvents around the launch when
// profiling is on, so the per-kernel average comes from the stream the kernel
// runs on.
KernelTimer* timer_for(Context* ctx, const char* name);
void timer_begin(Context* ctx, KernelTimer* t);
void timer_end(Context* ctx, KernelTimer* t);

struct TimedScope {
    Context* ctx;
    KernelTimer* t;
    TimedScope(Context* c, const char* name) : ctx(c), t(c->profile ? timer_for(c, name) : nullptr) {
        if (t) timer_begin(ctx, t);
    }
    ~TimedScope() {
        if (t) timer_end(ctx, t);
    }
};

// Kernel launchers (defined in the .hip translation units).
// Calls the progress callback if one is set and enough time has passed.
void report_progress(Context* ctx, const char* what, int iteration, double residual);

int launch_sdf(Context* ctx, const ibim_shape* shape);
// Level set -> signed distance to its zero set within band h (ibim_sdf.hip).
int launch_redistance(Context* ctx, double band);
int launch_gradient(Context* ctx, int order);
int launch_curvatures(Context* ctx);
int launch_tube(Context* ctx, int* count_out);
int launch_assemble(Context* ctx, int precision);
// Pair storage of the traditional operator (precision IBIM_C128 / IBIM_C64 as the element type).
// with64: keep a complex128 copy of the factors (the FP64 operator of the mixed solvers).
int launch_assemble_pair(Context* ctx, int precision, bool with64);
// fp64_copy: sweep the complex128 copy (OP_PAIR64) instead of the primary factors.
// with_z = false: the v planes (A p) alone.
int sweep_pair_launch(Context* ctx, bool fp64_copy, const void* p, const void* q, void* ypart, void* zpart,
                      const int* done, bool with_z = true);
// Dense rows of the pair operator into out_dev (nrows x count complex FP64).
int pair_rows(Context* ctx, int row0, int nrows, double2* out_dev);
// Y = A P and / or Z = A^H Q for 16 RHS-interleaved vectors over the pair
// storage (the complex128 copy when fp64_copy); a null Y or Z skips that
// product; kernels return at once while *stop is set.
int pair_batch_apply(Context* ctx, bool fp64_copy, const double2* P, const double2* Q, double2* Y, double2* Z,
                     const int* stop);
// The near-field part of pair_batch_apply alone (added to Y and Z after the far
// field; both products when Y and Z are given).
int pair_batch_near(Context* ctx, const double2* P, const double2* Q, double2* Y, double2* Z, const int* stop);

// One member of a grouped 16-column pass: one pair-stored operator and one
// batch of its right-hand sides (device-side table entry).
struct GroupMember {
    const void* F;       // pair storage factors (complex64 or complex128, the same for every member)
    const int* blk;      // first tile of each row block
    const void* pts;     // point table
    const double2* P;    // 16-column interleaved vectors in
    const double2* Q;
    double2* Y;          // A P and A^H Q out (far field)
    double2* Z;
    const int* stop;     // the batch's all-done flag: its items are skipped once set
    ibim_params prm;
    int count;           // points of the operator
    int item0;           // first work item (64-row block) of the member
};
// Factor sources a group pass can read: complex64 pairs with FP64 or FP32
// products, complex128 pairs.
enum GroupSource { GROUP_C64 = 0, GROUP_C64_FP32 = 1, GROUP_C128 = 2 };
// Checks that the contexts can share one pass (pair storage, one device, one
// pair kernel, one factor source; mixed: complex64 pairs) and returns the source.
int pair_group_source(Context* const* ctxs, int n, bool mixed, int* source);
// The far field of every member in one persistent launch on lead's stream:
// workgroups take 64-row work items from the queue counter (reset first) in
// item order; dev_members is a device copy of the table. with_z false: Y = A P
// alone (the members' Q and Z unused).
int pair_group_pass(Context* lead, int source, const GroupMember* dev_members, int nmembers, int nitems,
                    int* queue, bool with_z = true);
// Points every context's work at the first one's stream for a grouped solve
// (their own streams drained first), and back when it goes out of scope.
struct SharedStream {
    std::vector<Context*> ctxs;
    std::vector<hipStream_t> own;
    int share(Context* const* c, int n) {
        for (int i = 0; i < n; ++i) IBIM_HIP(hipStreamSynchronize(c[i]->stream));
        for (int i = 0; i < n; ++i) {
            ctxs.push_back(c[i]);
            own.push_back(c[i]->stream);
            c[i]->stream = c[0]->stream;
        }
        return IBIM_OK;
    }
    ~SharedStream() {
        if (!ctxs.empty()) (void)hipStreamSynchronize(ctxs[0]->stream);
        for (size_t c = 0; c < ctxs.size(); ++c) ctxs[c]->stream = own[c];
    }
};
// Grouped FP64 refinement around GMRES (ibim_solve_group with IBIM_SOLVE_GMRES,
// ibim_gmres.hip).
int run_group_gmres(Context* const* ctxs, int n, const ibim_solve_opts* opts, ibim_solve_info* infos);
// Batched BiCG (IBIM_SOLVE_BICG) or reliable BiCG (IBIM_SOLVE_RELIABLE) of the
// right-hand sides of several contexts together (ibim_batch.hip).
int run_group(Context* const* ctxs, int n, int method, const ibim_solve_opts* opts, ibim_solve_info* infos);

// The same products of the matrix-free operators (traditional tube): each
// factor computed where the complex128 pair storage would read it -- FP64
// (that storage's bits) or, with fast, the mixed-precision factors. Builds the
// point table and near lists on first use (host synchronisation: not inside a
// stream capture).
int pair_free_batch_apply(Context* ctx, bool fast, const double2* P, const double2* Q, double2* Y, double2* Z,
                          const int* stop);
// The single-vector sweep of the same operator into the pair planes (one
// factor evaluation per pair; the complex128 pair sweep's bits), or with fast
// the mixed-precision matrix-free operator: FP32 sine and cosine of an FP64
// phase in the factors, FP64 everywhere else. with_z false: A p alone.
int sweep_pair_free_launch(Context* ctx, bool fast, const void* p, const void* q, void* ypart, void* zpart,
                           const int* done, bool with_z = true);
// Point table and near lists of the matrix-free pair operator (built once per
// tube; host synchronisation, so before any stream capture).
int prepare_pair_free(Context* ctx);

// Operator rows between the host and HBM (ibim_hostio.hip). upload_rows fills
// ctx->mat (count rows at ctx->ld, precision IBIM_C128 or IBIM_C64, allocated
// by the caller); download_rows reads rows [row0, row0 + nrows) of the stored
// operator, dense or pair. Both return when the host side is complete.
int upload_rows(Context* ctx, const HostRows& src, int precision);
int download_rows(Context* ctx, int row0, int nrows, const HostRows& dst);
void release_staging(Context* ctx);
int launch_rhs(Context* ctx, int kind, const double* x0, double amp, const double* dir, int col);
int launch_eval(Context* ctx, int mode, int col, int nrecv, const double* xyz_dev, double* out_dev);
int run_gradient(Context* ctx, int col, int mode, const ibim_solve_opts* opts, double* J, double* dlevel, double* dk);
int run_bicg(Context* ctx, const ibim_solve_opts* opts, ibim_solve_info* info, double* hist);
int run_refined(Context* ctx, const ibim_solve_opts* opts, ibim_solve_info* info, double* hist);
int run_reliable(Context* ctx, const ibim_solve_opts* opts, ibim_solve_info* info, double* hist);
// FP64 refinement around GMRES on the stored operator, column by column (ibim_gmres.hip).
int run_gmres(Context* ctx, const ibim_solve_opts* opts, ibim_solve_info* info, double* hist);
int run_bicg_batched(Context* ctx, const ibim_solve_opts* opts, ibim_solve_info* info, double* hist);
int run_refined_batched(Context* ctx, const ibim_solve_opts* opts, ibim_solve_info* info);
int run_reliable_batched(Context* ctx, const ibim_solve_opts* opts, ibim_solve_info* info, double* hist);
// Y = op(A) P (adjoint: A^H) for 16 right-hand-side-interleaved vectors of ld rows.
int batch_apply(Context* ctx, int op, bool adjoint, const double2* P, double2* Y, const void* state);
int batch_apply_host(Context* ctx, int op, bool adjoint, int nvec, const double* pre, const double* pim, double* yre,
                     double* yim);
int matvec_dense(Context* ctx, const void* p_dev, const void* q_dev, void* y_dev, void* z_dev);

// Operator used by a sweep: the stored matrix, or entries recomputed on the fly.
// OP_PAIR64: the complex128 copy kept beside complex64 pair storage.
enum OpKind { OP_DENSE = 0, OP_FREE64 = 1, OP_FREE32 = 2, OP_PAIR64 = 3 };
// The FP64 operator the mixed solvers check against: the stored complex128
// operator, the complex128 copy of mixed pair storage, else recomputed entries.
inline int fp64_operator(const Context* ctx) {
    if (ctx->mat_pair64) return OP_PAIR64;
    return ctx->mat_precision == IBIM_C128 ? OP_DENSE : OP_FREE64;
}
// The matrix-free operators of a traditional tube sweep as pair storage with
// computed factors (sweep_pair_free_launch): the pair plane layout.
inline bool free_pair_op(const Context* ctx, int op) {
    return (op == OP_FREE64 || op == OP_FREE32) && ctx->kind == IBIM_TUBE_TRADITIONAL;
}
// Sweeps of op write the pair storage's partial-plane layout.
inline bool pair_layout_op(const Context* ctx, int op) {
    return (op == OP_DENSE && ctx->mat_pair) || op == OP_PAIR64 || free_pair_op(ctx, op);
}
// Tile sizes shared by every sweep (the partial planes have these shapes).
constexpr int kSweepCols = 1024;
constexpr int kSweepRows = 512;
int sweep_free_launch(Context* ctx, int op, const void* p, const void* q, void* ypart, void* zpart, const int* done);
int matvec_op(Context* ctx, int op, const void* p_dev, const void* q_dev, void* y_dev, void* z_dev);
// y = A p alone (pair storage: a sweep without the A^H products; other
// operators: the full sweep, A^H p discarded).
int matvec_op_y(Context* ctx, int op, const void* p_dev, void* y_dev);
// The same product with its workspace (partial planes, block partials, a flag)
// in caller-provided device memory of matvec_scratch_bytes(ctx, op) bytes:
// allocates nothing.
size_t matvec_scratch_bytes(const Context* ctx, int op);
int matvec_op_scratch(Context* ctx, int op, const void* p_dev, const void* q_dev, void* y_dev, void* z_dev,
                      void* scratch);
// Device address of the no-allocation entry point's scratch array on `device`
// (resolved once per device; ibim_noalloc.hip).
int noalloc_scratch(int device, void** ptr);
// ref_guards = false for correction / adjoint solves (only exact zeros are zero data).
int bicg_solve_op(Context* ctx, int op, const void* b_dev, void* x_dev, int maxit, double tol, int chunk,
                  ibim_solve_info* info, double* hist, bool adjoint = false, bool ref_guards = true);
// The same system, op(A) x = b (op(A) = A^H when adjoint), by enlarged GMRES on
// the 16-column passes of operator `op` (ibim_gmres.hip): b split over 16
// pieces of the surface, one block Krylov space, restarted with the true
// residual until it is below tol ||b||.
int enlarged_solve_op(Context* ctx, int op, bool adjoint, const void* b_dev, void* x_dev, int maxit, double tol,
                      ibim_solve_info* info);

}  // namespace ibim

// Handle type of the public API.
struct ibim_ctx : ibim::Context {};
