// ibim_hostio.hip -- the operator between host row arrays and HBM.
//
// The reference's kernel builders and solvers hand the operator over as
// separate real and imaginary rows of doubles (double** Kernel / Kernelimg,
// HelmholtzKernel.h:20-39, BICGSTAB.h:15-17); in HBM it is interleaved complex
// (double2 / float2 rows at the leading dimension ld) or pair storage whose
// rows are rebuilt on the device. Both directions stream through two pinned
// staging blocks: while one block crosses PCIe, host threads (de)interleave the
// other, so a transfer runs near the link's rate instead of one core's copy
// rate, and row-pointer (double**) arrays are read or written in place.
#include <algorithm>
#include <thread>
#include <vector>

#include "ibim_hoststage.hpp"
#include "ibim_internal.hpp"

namespace ibim {

namespace {

using namespace hoststage;

constexpr size_t kStageBytes = size_t(32) << 20;  // per staging block

int ensure_staging(Context* ctx, size_t row_bytes) {
    const size_t want = std::max(kStageBytes, row_bytes);
    if (ctx->stage_bytes >= want) return IBIM_OK;
    release_staging(ctx);
    for (int s = 0; s < 2; ++s) {
        IBIM_HIP(hipHostMalloc(&ctx->stage[s], want, hipHostMallocDefault));
        IBIM_HIP(hipEventCreateWithFlags(&ctx->stage_ev[s], hipEventDisableTiming));
    }
    ctx->stage_bytes = want;
    return IBIM_OK;
}

}  // namespace

void release_staging(Context* ctx) {
    for (int s = 0; s < 2; ++s) {
        if (ctx->stage_ev[s]) (void)hipEventDestroy(ctx->stage_ev[s]);
        if (ctx->stage[s]) (void)hipHostFree(ctx->stage[s]);
        ctx->stage[s] = nullptr;
        ctx->stage_ev[s] = nullptr;
    }
    ctx->stage_bytes = 0;
}

int upload_rows(Context* ctx, const HostRows& src, int precision) {
    const int count = ctx->count;
    const size_t ld = ctx->ld;
    const bool c128 = precision == IBIM_C128;
    const size_t row_bytes = (c128 ? sizeof(double2) : sizeof(float2)) * ld;
    if (int rc = ensure_staging(ctx, row_bytes)) return rc;
    const int per = (int)(ctx->stage_bytes / row_bytes);
    char* dst = ctx->mat.as<char>();
    for (int b = 0, r0 = 0; r0 < count; ++b, r0 += per) {
        const int s = b & 1, nr = std::min(per, count - r0);
        if (b >= 2) IBIM_HIP(hipEventSynchronize(ctx->stage_ev[s]));  // the copy that last read this block is done
        char* stage = static_cast<char*>(ctx->stage[s]);
        for_rows(nr, row_bytes * nr, [&](int r) {
            void* d = stage + row_bytes * r;
            if (c128)
                interleave_row((double*)d, src.re_row(r0 + r), src.im_row(r0 + r), count, ld);
            else
                interleave_row((float*)d, src.re_row(r0 + r), src.im_row(r0 + r), count, ld);
        });
        IBIM_HIP(hipMemcpyAsync(dst + row_bytes * r0, stage, row_bytes * nr, hipMemcpyHostToDevice, ctx->stream));
        IBIM_HIP(hipEventRecord(ctx->stage_ev[s], ctx->stream));
    }
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    return IBIM_OK;
}

int download_rows(Context* ctx, int row0, int nrows, const HostRows& dst) {
    const int n = ctx->count;
    const bool pair = ctx->mat_pair;
    // pair storage: rows rebuilt as complex FP64 at stride n into two halves of
    // xfer; dense: complex rows of the stored precision at stride ld
    const bool c128 = pair || ctx->mat_precision == IBIM_C128;
    const size_t row_bytes = (c128 ? sizeof(double2) : sizeof(float2)) * (pair ? (size_t)n : ctx->ld);
    if (int rc = ensure_staging(ctx, row_bytes)) return rc;
    const int per = (int)(ctx->stage_bytes / row_bytes);
    if (pair && ctx->xfer.ensure(2 * row_bytes * (size_t)per)) return IBIM_ERR_NOMEM;
    const int nb = (nrows + per - 1) / per;
    auto issue = [&](int b) -> int {
        const int r0 = b * per, nr = std::min(per, nrows - r0), s = b & 1;
        const char* src;
        if (pair) {
            double2* half = reinterpret_cast<double2*>(ctx->xfer.as<char>() + row_bytes * (size_t)per * s);
            if (int rc = pair_rows(ctx, row0 + r0, nr, half)) return rc;
            src = reinterpret_cast<const char*>(half);
        } else {
            src = ctx->mat.as<char>() + row_bytes * (size_t)(row0 + r0);
        }
        IBIM_HIP(hipMemcpyAsync(ctx->stage[s], src, row_bytes * nr, hipMemcpyDeviceToHost, ctx->stream));
        IBIM_HIP(hipEventRecord(ctx->stage_ev[s], ctx->stream));
        return IBIM_OK;
    };
    if (nb > 0)
        if (int rc = issue(0)) return rc;
    for (int b = 0; b < nb; ++b) {
        // the next block's copy goes into the other stage, whose rows were split last step
        if (b + 1 < nb)
            if (int rc = issue(b + 1)) return rc;
        const int r0 = b * per, nr = std::min(per, nrows - r0), s = b & 1;
        IBIM_HIP(hipEventSynchronize(ctx->stage_ev[s]));
        const char* stage = static_cast<const char*>(ctx->stage[s]);
        for_rows(nr, row_bytes * nr, [&](int r) {
            const void* p = stage + row_bytes * r;
            if (c128)
                split_row((const double*)p, dst.re_row(r0 + r), dst.im_row(r0 + r), n);
            else
                split_row((const float*)p, dst.re_row(r0 + r), dst.im_row(r0 + r), n);
        });
    }
    return IBIM_OK;
}

}  // namespace ibim
