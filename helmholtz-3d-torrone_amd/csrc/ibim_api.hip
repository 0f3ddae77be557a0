// ibim_api.hip -- the device-resident handle API of libibim.so (include/ibim.h).
#include <chrono>
#include <cstdarg>
#include <cstring>
#include <new>
#include <vector>

#include "ibim_internal.hpp"

namespace ibim {

namespace {
thread_local char g_err[512] = "";
}

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int DevBuf::ensure(size_t want) {
    if (want <= bytes && ptr) return 0;
    release();
    if (want == 0) want = 16;
    hipError_t e = hipMalloc(&ptr, want);
    if (e != hipSuccess) {
        ptr = nullptr;
        bytes = 0;
        set_error("hipMalloc(%zu bytes) failed: %s", want, hipGetErrorString(e));
        return 1;
    }
    bytes = want;
    return 0;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

KernelTimer* timer_for(Context* ctx, const char* name) {
    for (auto& t : ctx->timers)
        if (t.name == name) return &t;
    ctx->timers.push_back(KernelTimer{});
    ctx->timers.back().name = name;
    return &ctx->timers.back();
}

void report_progress(Context* ctx, const char* what, int iteration, double residual) {
    if (!ctx->progress) return;
    const double now =
        std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    if (ctx->progress_last != 0.0 && now - ctx->progress_last < 5.0) return;
    ctx->progress_last = now;
    ctx->progress(what, iteration, residual, ctx->progress_user);
}

void timer_begin(Context* ctx, KernelTimer* t) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, ctx->stream);
    t->begin.push_back(e);
}

void timer_end(Context* ctx, KernelTimer* t) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, ctx->stream);
    t->end.push_back(e);
}

// Fold recorded event pairs into the totals (waits for the stream).
static void collect_timers(Context* ctx) {
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& t : ctx->timers) {
        size_t n = t.begin.size() < t.end.size() ? t.begin.size() : t.end.size();
        for (size_t i = 0; i < n; ++i) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, t.begin[i], t.end[i]) == hipSuccess) t.ms_total += ms;
            (void)hipEventDestroy(t.begin[i]);
            (void)hipEventDestroy(t.end[i]);
        }
        t.launches += (long)n;
        t.begin.clear();
        t.end.clear();
    }
}

}  // namespace ibim

using namespace ibim;

namespace {

int check_ctx(ibim_ctx* ctx) {
    if (!ctx) {
        set_error("null handle");
        return IBIM_ERR_ARG;
    }
    return (int)hipSetDevice(ctx->device) == (int)hipSuccess ? IBIM_OK : IBIM_ERR_HIP;
}

#define CHECK(ctx)                  \
    do {                            \
        int rc_ = check_ctx(ctx);   \
        if (rc_) return rc_;        \
    } while (0)

}  // namespace

extern "C" {

int ibim_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char* ibim_last_error(void) { return g_err; }

int ibim_create(ibim_ctx** out) {
    if (!out) return IBIM_ERR_ARG;
    *out = nullptr;
    if (ibim_device_count() <= 0) {
        set_error("no HIP device");
        return IBIM_ERR_NODEVICE;
    }
    ibim_ctx* ctx = new (std::nothrow) ibim_ctx();
    if (!ctx) return IBIM_ERR_NOMEM;
    IBIM_HIP(hipGetDevice(&ctx->device));
    // resolve the no-allocation entry point's scratch array on this device now
    // (a failure resurfaces, with its message, from ibim_apply_noalloc)
    void* scratch_ptr = nullptr;
    (void)noalloc_scratch(ctx->device, &scratch_ptr);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreate failed");
        delete ctx;
        return IBIM_ERR_HIP;
    }
    *out = ctx;
    return IBIM_OK;
}

void ibim_destroy(ibim_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    collect_timers(ctx);
    for (DevBuf* b : {&ctx->level, &ctx->gx, &ctx->gy, &ctx->gz, &ctx->mc, &ctx->gc, &ctx->node, &ctx->tube,
                      &ctx->blockcnt, &ctx->err_flag, &ctx->mat, &ctx->mat64, &ctx->rhs, &ctx->density, &ctx->work, &ctx->partial,
                      &ctx->scal, &ctx->done, &ctx->xfer, &ctx->refine, &ctx->recv, &ctx->adj, &ctx->gadj, &ctx->batch, &ctx->group,
                      &ctx->pair_tiles, &ctx->pair_blk, &ctx->pair_pts, &ctx->near_off, &ctx->near_idx,
                      &ctx->near_cur, &ctx->free_cache})
        b->release();
    release_staging(ctx);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int ibim_sync(ibim_ctx* ctx) {
    CHECK(ctx);
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    return IBIM_OK;
}

static int set_grid(ibim_ctx* ctx, int n, double lo, double h) {
    if (n < 4 || h <= 0.0) {
        set_error("grid needs n >= 4 and h > 0");
        return IBIM_ERR_ARG;
    }
    const size_t nodes = (size_t)(n + 1) * (n + 1) * (n + 1);
    for (DevBuf* b : {&ctx->level, &ctx->gx, &ctx->gy, &ctx->gz, &ctx->mc, &ctx->gc})
        if (b->ensure(sizeof(double) * nodes)) return IBIM_ERR_NOMEM;
    ctx->n = n;
    ctx->lo = lo;
    ctx->h = h;
    ctx->have_geometry = false;
    ctx->weno_order = 0;
    ctx->count = 0;
    ctx->have_tube = false;
    ctx->mat_precision = -1;
    ctx->mat_pair = ctx->mat_pair64 = ctx->pair_aux = ctx->free_cache_valid = false;
    return IBIM_OK;
}

int ibim_set_level_host(ibim_ctx* ctx, int n, double lo, double h, const double* level) {
    CHECK(ctx);
    int rc = set_grid(ctx, n, lo, h);
    if (rc) return rc;
    const size_t nodes = (size_t)(n + 1) * (n + 1) * (n + 1);
    IBIM_HIP(hipMemcpyAsync(ctx->level.ptr, level, sizeof(double) * nodes, hipMemcpyHostToDevice, ctx->stream));
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    return IBIM_OK;
}

int ibim_set_level_device(ibim_ctx* ctx, int n, double lo, double h, const double* level_dev) {
    CHECK(ctx);
    int rc = set_grid(ctx, n, lo, h);
    if (rc) return rc;
    const size_t nodes = (size_t)(n + 1) * (n + 1) * (n + 1);
    IBIM_HIP(hipMemcpyAsync(ctx->level.ptr, level_dev, sizeof(double) * nodes, hipMemcpyDeviceToDevice, ctx->stream));
    return IBIM_OK;
}

int ibim_set_level_shape(ibim_ctx* ctx, int n, double lo, double h, const ibim_shape* shape) {
    CHECK(ctx);
    if (!shape || shape->kind < IBIM_SHAPE_SPHERE || shape->kind > IBIM_SHAPE_STAR) return IBIM_ERR_ARG;
    int rc = set_grid(ctx, n, lo, h);
    if (rc) return rc;
    return launch_sdf(ctx, shape);
}

int ibim_redistance(ibim_ctx* ctx, double band) {
    CHECK(ctx);
    if (!ctx->n || !(band >= 1.0 && band <= 1e6)) {
        set_error("ibim_redistance: needs a level set and a band of at least one grid spacing");
        return IBIM_ERR_ARG;
    }
    return launch_redistance(ctx, band);
}

int ibim_geometry(ibim_ctx* ctx, int weno_order) {
    CHECK(ctx);
    if (!ctx->n) {
        set_error("ibim_geometry: no level set");
        return IBIM_ERR_ARG;
    }
    if (weno_order != 3 && weno_order != 5) {
        set_error("weno_order must be 3 or 5");
        return IBIM_ERR_ARG;
    }
    int rc = launch_gradient(ctx, weno_order);
    if (!rc) rc = launch_curvatures(ctx);
    if (!rc) ctx->have_geometry = true;
    ctx->weno_order = rc ? 0 : weno_order;
    return rc;
}

int ibim_get_level(ibim_ctx* ctx, double* level) {
    CHECK(ctx);
    if (!ctx->n || !level) return IBIM_ERR_ARG;
    const size_t bytes = sizeof(double) * (size_t)(ctx->n + 1) * (ctx->n + 1) * (ctx->n + 1);
    IBIM_HIP(hipMemcpyAsync(level, ctx->level.ptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    return IBIM_OK;
}

int ibim_get_grid(ibim_ctx* ctx, double* gx, double* gy, double* gz, double* mc, double* gc) {
    CHECK(ctx);
    const size_t bytes = sizeof(double) * (size_t)(ctx->n + 1) * (ctx->n + 1) * (ctx->n + 1);
    double* dst[5] = {gx, gy, gz, mc, gc};
    DevBuf* src[5] = {&ctx->gx, &ctx->gy, &ctx->gz, &ctx->mc, &ctx->gc};
    for (int i = 0; i < 5; ++i)
        if (dst[i]) IBIM_HIP(hipMemcpyAsync(dst[i], src[i]->ptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    return IBIM_OK;
}

int ibim_set_grid_fields_host(ibim_ctx* ctx, const double* gx, const double* gy, const double* gz, const double* mc,
                              const double* gc) {
    CHECK(ctx);
    if (!ctx->n) return IBIM_ERR_ARG;
    const size_t bytes = sizeof(double) * (size_t)(ctx->n + 1) * (ctx->n + 1) * (ctx->n + 1);
    const double* srcs[5] = {gx, gy, gz, mc, gc};
    DevBuf* dst[5] = {&ctx->gx, &ctx->gy, &ctx->gz, &ctx->mc, &ctx->gc};
    for (int i = 0; i < 5; ++i) {
        if (!srcs[i]) {
            set_error("all five grid fields are required");
            return IBIM_ERR_ARG;
        }
        IBIM_HIP(hipMemcpyAsync(dst[i]->ptr, srcs[i], bytes, hipMemcpyHostToDevice, ctx->stream));
    }
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    ctx->have_geometry = true;
    ctx->weno_order = 0;
    return IBIM_OK;
}

int ibim_tube(ibim_ctx* ctx, const ibim_params* prm, int kind, int* count) {
    CHECK(ctx);
    if (!prm || !count || (kind != IBIM_TUBE_TRADITIONAL && kind != IBIM_TUBE_COMBO)) return IBIM_ERR_ARG;
    if (!ctx->have_geometry) {
        set_error("ibim_tube: run ibim_geometry first");
        return IBIM_ERR_ARG;
    }
    if (prm->combined && (kind != IBIM_TUBE_TRADITIONAL || !prm->dirichlet || prm->interior)) {
        set_error("ibim_tube: the combined-field operator is the exterior Dirichlet problem on the traditional tube");
        return IBIM_ERR_ARG;
    }
    ctx->prm = *prm;
    ctx->kind = kind;
    ctx->mat_precision = -1;
    ctx->mat_pair = ctx->mat_pair64 = ctx->pair_aux = ctx->free_cache_valid = false;
    ctx->count = 0;  // a failed build leaves no stale tube behind
    ctx->have_tube = false;
    int c = 0;
    int rc = launch_tube(ctx, &c);
    if (rc) return rc;
    ctx->count = c;
    ctx->have_tube = true;
    ctx->ld = ((size_t)c + 63) / 64 * 64;
    *count = c;
    return ibim_set_nrhs(ctx, ctx->nrhs > 0 ? ctx->nrhs : 1);
}

int ibim_get_tube(ibim_ctx* ctx, int* node, double* fields) {
    CHECK(ctx);
    if (!ctx->count) return IBIM_ERR_ARG;
    if (node)
        IBIM_HIP(hipMemcpyAsync(node, ctx->node.ptr, sizeof(int) * ctx->count, hipMemcpyDeviceToHost, ctx->stream));
    if (fields)
        IBIM_HIP(hipMemcpyAsync(fields, ctx->tube.ptr, sizeof(double) * IBIM_NF * ctx->count, hipMemcpyDeviceToHost,
                                ctx->stream));
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    return IBIM_OK;
}

const double* ibim_tube_device(ibim_ctx* ctx) { return ctx ? ctx->tube.as<double>() : nullptr; }

int ibim_assemble(ibim_ctx* ctx, int precision) {
    CHECK(ctx);
    if (!ctx->count || precision < IBIM_C128 || precision > IBIM_MIXED_PAIR) {
        set_error("ibim_assemble: needs a tube and a precision");
        return IBIM_ERR_ARG;
    }
    if (precision == IBIM_C128_PAIR) return launch_assemble_pair(ctx, IBIM_C128, false);
    if (precision == IBIM_C64_PAIR) return launch_assemble_pair(ctx, IBIM_C64, false);
    if (precision == IBIM_MIXED_PAIR) return launch_assemble_pair(ctx, IBIM_C64, true);
    return launch_assemble(ctx, precision);
}

int ibim_regcount(ibim_ctx* ctx) { return ctx ? ctx->regcount : 0; }

int ibim_get_rows(ibim_ctx* ctx, int row0, int nrows, double* re, double* im) {
    CHECK(ctx);
    const int n = ctx->count;
    if (ctx->mat_precision < 0 || row0 < 0 || nrows < 0 || row0 + nrows > n || (nrows && (!re || !im)))
        return IBIM_ERR_ARG;
    HostRows dst;
    dst.re = re;
    dst.im = im;
    dst.stride = (size_t)n;
    return download_rows(ctx, row0, nrows, dst);
}

int ibim_set_matrix_host(ibim_ctx* ctx, int count, const double* re, const double* im, int precision) {
    CHECK(ctx);
    if (count <= 0 || !re || !im || (precision != IBIM_C128 && precision != IBIM_C64)) return IBIM_ERR_ARG;
    // the operator no longer belongs to a tube: matrix-free operators, the
    // FP64 operators of the mixed solvers and the evaluation need one
    ctx->have_tube = false;
    ctx->count = count;
    ctx->ld = ((size_t)count + 63) / 64 * 64;
    const size_t esz = precision == IBIM_C128 ? 16 : 8;
    if (ctx->mat.ensure(esz * ctx->ld * (size_t)count)) return IBIM_ERR_NOMEM;
    HostRows src;
    src.re = const_cast<double*>(re);
    src.im = const_cast<double*>(im);
    src.stride = (size_t)count;
    ctx->mat_precision = -1;  // not a valid operator until the upload completes
    if (int rc = upload_rows(ctx, src, precision)) return rc;
    ctx->mat_precision = precision;
    ctx->mat_pair = ctx->mat_pair64 = ctx->pair_aux = ctx->free_cache_valid = false;
    return ibim_set_nrhs(ctx, ctx->nrhs > 0 ? ctx->nrhs : 1);
}

int ibim_apply(ibim_ctx* ctx, int op, const double* pre, const double* pim, const double* qre, const double* qim,
               double* yre, double* yim, double* zre, double* zim) {
    CHECK(ctx);
    const int n = ctx->count;
    if (!n || op < IBIM_OP_STORED || op > IBIM_OP_FREE32 || (op == IBIM_OP_STORED && ctx->mat_precision < 0))
        return IBIM_ERR_ARG;
    if (op != IBIM_OP_STORED && !ctx->have_tube) {
        set_error("ibim_apply: matrix-free operators need a tube (the operator was set from the host)");
        return IBIM_ERR_ARG;
    }
    if (ctx->xfer.ensure(sizeof(double2) * 4 * (size_t)n)) return IBIM_ERR_NOMEM;
    double2* d = ctx->xfer.as<double2>();
    std::vector<double> t(4 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        t[2 * i] = pre ? pre[i] : 0.0;
        t[2 * i + 1] = pim ? pim[i] : 0.0;
        t[2 * (size_t)n + 2 * i] = qre ? qre[i] : 0.0;
        t[2 * (size_t)n + 2 * i + 1] = qim ? qim[i] : 0.0;
    }
    IBIM_HIP(hipMemcpyAsync(d, t.data(), sizeof(double) * 4 * n, hipMemcpyHostToDevice, ctx->stream));
    const int kinds[3] = {OP_DENSE, OP_FREE64, OP_FREE32};
    // no A^H outputs: the product A p alone (pair storage sweeps skip A^H q)
    const bool y_only = !zre && !zim;
    int rc = y_only ? matvec_op_y(ctx, kinds[op], d, d + 2 * (size_t)n)
                    : matvec_op(ctx, kinds[op], d, d + n, d + 2 * (size_t)n, d + 3 * (size_t)n);
    if (rc) return rc;
    IBIM_HIP(hipMemcpyAsync(t.data(), d + 2 * (size_t)n, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, ctx->stream));
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; ++i) {
        if (yre) yre[i] = t[2 * i];
        if (yim) yim[i] = t[2 * i + 1];
        if (zre) zre[i] = t[2 * (size_t)n + 2 * i];
        if (zim) zim[i] = t[2 * (size_t)n + 2 * i + 1];
    }
    return IBIM_OK;
}

int ibim_apply_batch(ibim_ctx* ctx, int op, int adjoint, int nvec, const double* pre, const double* pim, double* yre,
                     double* yim) {
    CHECK(ctx);
    if (!ctx->count || nvec < 1 || nvec > 16 || !pre || !yre || !yim || op < IBIM_OP_STORED || op > IBIM_OP_FREE32)
        return IBIM_ERR_ARG;
    if (op == IBIM_OP_STORED && ctx->mat_precision < 0) {
        set_error("ibim_apply_batch: no stored operator (assemble one, or use a matrix-free op)");
        return IBIM_ERR_ARG;
    }
    if (op != IBIM_OP_STORED && (ctx->kind != IBIM_TUBE_TRADITIONAL || !ctx->have_tube)) {
        set_error("ibim_apply_batch: matrix-free products need the traditional tube");
        return IBIM_ERR_ARG;
    }
    const int kinds[3] = {OP_DENSE, OP_FREE64, OP_FREE32};
    return batch_apply_host(ctx, kinds[op], adjoint != 0, nvec, pre, pim, yre, yim);
}

int ibim_set_nrhs(ibim_ctx* ctx, int nrhs) {
    CHECK(ctx);
    if (nrhs <= 0 || !ctx->count) return IBIM_ERR_ARG;
    const size_t bytes = sizeof(double2) * (size_t)ctx->count * nrhs;
    if (ctx->rhs.ensure(bytes) || ctx->density.ensure(bytes)) return IBIM_ERR_NOMEM;
    IBIM_HIP(hipMemsetAsync(ctx->rhs.ptr, 0, bytes, ctx->stream));
    IBIM_HIP(hipMemsetAsync(ctx->density.ptr, 0, bytes, ctx->stream));
    ctx->nrhs = nrhs;
    return IBIM_OK;
}

int ibim_rhs(ibim_ctx* ctx, int kind, const double* x0, double amp, const double* dir, int col) {
    CHECK(ctx);
    if (!ctx->count || col < 0 || col >= ctx->nrhs || kind < 0 || kind > 2) return IBIM_ERR_ARG;
    if (!ctx->have_tube) {
        set_error("ibim_rhs: no tube for the current operator (set with ibim_set_matrix_host)");
        return IBIM_ERR_ARG;
    }
    ctx->rhs_desc.resize(ctx->nrhs);
    auto& d = ctx->rhs_desc[col];
    d.kind = kind;
    d.amp = amp;
    for (int i = 0; i < 3; ++i) {
        d.x0[i] = x0 ? x0[i] : 0.0;
        d.dir[i] = dir ? dir[i] : 0.0;
    }
    return launch_rhs(ctx, kind, x0, amp, dir, col);
}

static int copy_col(ibim_ctx* ctx, DevBuf& buf, int col, double* re, double* im, bool to_device) {
    const int n = ctx->count;
    std::vector<double> tmp(2 * (size_t)n);
    double2* d = buf.as<double2>() + (size_t)col * n;
    if (to_device) {
        for (int i = 0; i < n; ++i) {
            tmp[2 * i] = re ? re[i] : 0.0;
            tmp[2 * i + 1] = im ? im[i] : 0.0;
        }
        IBIM_HIP(hipMemcpyAsync(d, tmp.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice, ctx->stream));
        IBIM_HIP(hipStreamSynchronize(ctx->stream));
    } else {
        IBIM_HIP(hipMemcpyAsync(tmp.data(), d, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, ctx->stream));
        IBIM_HIP(hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < n; ++i) {
            if (re) re[i] = tmp[2 * i];
            if (im) im[i] = tmp[2 * i + 1];
        }
    }
    return IBIM_OK;
}

int ibim_set_rhs_host(ibim_ctx* ctx, int col, const double* bre, const double* bim) {
    CHECK(ctx);
    if (col < 0 || col >= ctx->nrhs) return IBIM_ERR_ARG;
    ctx->rhs_desc.resize(ctx->nrhs);
    ctx->rhs_desc[col].kind = -1;
    return copy_col(ctx, ctx->rhs, col, (double*)bre, (double*)bim, true);
}

int ibim_get_rhs(ibim_ctx* ctx, int col, double* bre, double* bim) {
    CHECK(ctx);
    if (col < 0 || col >= ctx->nrhs) return IBIM_ERR_ARG;
    return copy_col(ctx, ctx->rhs, col, bre, bim, false);
}

int ibim_get_density(ibim_ctx* ctx, int col, double* re, double* im) {
    CHECK(ctx);
    if (col < 0 || col >= ctx->nrhs) return IBIM_ERR_ARG;
    return copy_col(ctx, ctx->density, col, re, im, false);
}

int ibim_set_density_host(ibim_ctx* ctx, int col, const double* re, const double* im) {
    CHECK(ctx);
    if (col < 0 || col >= ctx->nrhs) return IBIM_ERR_ARG;
    return copy_col(ctx, ctx->density, col, (double*)re, (double*)im, true);
}

int ibim_solve(ibim_ctx* ctx, const ibim_solve_opts* opts, ibim_solve_info* info, double* hist) {
    CHECK(ctx);
    if (!opts || !info || opts->maxit <= 0 || !ctx->count) return IBIM_ERR_ARG;
    memset(info, 0, sizeof *info);
    // without a tube only the stored operator exists: BiCG on it (the mixed
    // solvers recompute their FP32 or FP64 operator from the points)
    if (!ctx->have_tube && (ctx->mat_precision < 0 || opts->method != IBIM_SOLVE_BICG)) {
        set_error("ibim_solve: the operator was set from the host; without a tube only BiCG on it is possible");
        return IBIM_ERR_ARG;
    }
    const bool batched = ctx->nrhs > 1 && opts->batch == 0;
    if (opts->method == IBIM_SOLVE_BICG)
        return batched ? run_bicg_batched(ctx, opts, info, hist) : run_bicg(ctx, opts, info, hist);
    if (opts->method == IBIM_SOLVE_REFINED)
        return batched && (ctx->kind == IBIM_TUBE_TRADITIONAL || ctx->mat_precision == IBIM_C64)
                   ? run_refined_batched(ctx, opts, info)
                   : run_refined(ctx, opts, info, hist);
    if (opts->method == IBIM_SOLVE_RELIABLE)
        return batched ? run_reliable_batched(ctx, opts, info, hist) : run_reliable(ctx, opts, info, hist);
    if (opts->method == IBIM_SOLVE_GMRES || opts->method == IBIM_SOLVE_BLOCK_GMRES)
        return run_gmres(ctx, opts, info, hist);
    return IBIM_ERR_ARG;
}

int ibim_solve_group(ibim_ctx* const* ctxs, int nctx, const ibim_solve_opts* opts, ibim_solve_info* infos) {
    if (!ctxs || nctx <= 0 || !opts || !infos || opts->maxit <= 0) return IBIM_ERR_ARG;
    for (int c = 0; c < nctx; ++c) {
        if (!ctxs[c]) {
            set_error("ibim_solve_group: null handle %d", c);
            return IBIM_ERR_ARG;
        }
        memset(&infos[c], 0, sizeof infos[c]);
    }
    CHECK(ctxs[0]);
    std::vector<Context*> v(ctxs, ctxs + nctx);
    return run_group(v.data(), nctx, opts->method, opts, infos);
}

int ibim_evaluate(ibim_ctx* ctx, int mode, int col, int nrecv, const double* xyz, double* ure, double* uim) {
    CHECK(ctx);
    if (!ctx->count || nrecv <= 0 || !xyz || col < 0 || col >= ctx->nrhs || mode < 0 || mode > IBIM_EVAL_COMBINED)
        return IBIM_ERR_ARG;
    if (!ctx->have_tube) {
        set_error("ibim_evaluate: no tube for the current operator (set with ibim_set_matrix_host)");
        return IBIM_ERR_ARG;
    }
    if (ctx->xfer.ensure(sizeof(double) * 5 * (size_t)nrecv)) return IBIM_ERR_NOMEM;
    double* dxyz = ctx->xfer.as<double>();
    double* dout = dxyz + 3 * (size_t)nrecv;
    IBIM_HIP(hipMemcpyAsync(dxyz, xyz, sizeof(double) * 3 * nrecv, hipMemcpyHostToDevice, ctx->stream));
    int rc = launch_eval(ctx, mode, col, nrecv, dxyz, dout);
    if (rc) return rc;
    std::vector<double> tmp(2 * (size_t)nrecv);
    IBIM_HIP(hipMemcpyAsync(tmp.data(), dout, sizeof(double) * 2 * nrecv, hipMemcpyDeviceToHost, ctx->stream));
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < nrecv; ++i) {
        if (ure) ure[i] = tmp[2 * i];
        if (uim) uim[i] = tmp[2 * i + 1];
    }
    return IBIM_OK;
}

int ibim_set_receivers(ibim_ctx* ctx, int nrecv, const double* xyz, const double* tre, const double* tim) {
    CHECK(ctx);
    if (nrecv <= 0 || !xyz) return IBIM_ERR_ARG;
    if (ctx->recv.ensure(sizeof(double) * 5 * (size_t)nrecv)) return IBIM_ERR_NOMEM;
    std::vector<double> t(5 * (size_t)nrecv);
    for (int i = 0; i < 3 * nrecv; ++i) t[i] = xyz[i];
    for (int i = 0; i < nrecv; ++i) {
        t[3 * (size_t)nrecv + 2 * i] = tre ? tre[i] : 0.0;
        t[3 * (size_t)nrecv + 2 * i + 1] = tim ? tim[i] : 0.0;
    }
    IBIM_HIP(hipMemcpyAsync(ctx->recv.ptr, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice, ctx->stream));
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    ctx->nrecv = nrecv;
    return IBIM_OK;
}

int ibim_gradient(ibim_ctx* ctx, int col, int mode, const ibim_solve_opts* opts, double* J, double* dlevel,
                  double* dk) {
    CHECK(ctx);
    if (!ctx->count || col < 0 || col >= ctx->nrhs || !ctx->nrecv || !opts || (mode != IBIM_EVAL_SINGLE &&
        mode != IBIM_EVAL_DOUBLE) || ctx->kind != IBIM_TUBE_TRADITIONAL || !ctx->weno_order || !ctx->have_tube ||
        ctx->prm.combined) {
        set_error("ibim_gradient: needs a solved traditional problem (not the combined-field operator) on WENO "
                  "geometry, receivers and a single/double-layer mode");
        return IBIM_ERR_ARG;
    }
    return run_gradient(ctx, col, mode, opts, J, dlevel, dk);
}

int ibim_set_progress(ibim_ctx* ctx, ibim_progress_fn fn, void* user) {
    CHECK(ctx);
    ctx->progress = fn;
    ctx->progress_user = user;
    ctx->progress_last = 0.0;
    return IBIM_OK;
}

int ibim_set_profiling(ibim_ctx* ctx, int on) {
    CHECK(ctx);
    ctx->profile = on != 0;
    return IBIM_OK;
}

int ibim_set_free_cache(ibim_ctx* ctx, double max_bytes) {
    CHECK(ctx);
    if (!(max_bytes >= 0.0) || max_bytes > 1e15) return IBIM_ERR_ARG;
    IBIM_HIP(hipStreamSynchronize(ctx->stream));  // no pass may still read the old cache
    ctx->free_cache_max = (size_t)max_bytes;
    ctx->free_cache_valid = false;
    if (ctx->free_cache_max == 0) ctx->free_cache.release();
    return IBIM_OK;
}

int ibim_set_fp32_products(ibim_ctx* ctx, int on) {
    CHECK(ctx);
    ctx->fp32_products = on != 0;
    return IBIM_OK;
}

int ibim_set_graphs(ibim_ctx* ctx, int on) {
    CHECK(ctx);
    ctx->use_graphs = on != 0;
    return IBIM_OK;
}

int ibim_kernel_time(ibim_ctx* ctx, const char* name, double* ms_total, long* launches) {
    CHECK(ctx);
    collect_timers(ctx);
    for (auto& t : ctx->timers)
        if (t.name == name) {
            if (ms_total) *ms_total = t.ms_total;
            if (launches) *launches = t.launches;
            return IBIM_OK;
        }
    if (ms_total) *ms_total = 0.0;
    if (launches) *launches = 0;
    return IBIM_OK;
}

int ibim_reset_timers(ibim_ctx* ctx) {
    CHECK(ctx);
    collect_timers(ctx);
    ctx->timers.clear();
    return IBIM_OK;
}

}  // extern "C"
