// ibim_noalloc.hip -- ibim_apply_noalloc: ibim_apply's product y = A p,
// z = A^H q on device vectors, with its workspace in one file-scope device
// array instead of device memory the library allocates. For callers that
// cannot allocate: code under HIP graph capture, tight serving loops, hosts
// that forbid hipMalloc.
//
// The array (kScratchBytes per GPU, held until the process ends) is shared by
// every call on a GPU and keeps the last call's contents; the product writes
// every plane element it reads and clears its flag with a kernel. Calls that
// overlap on one GPU (two handles, two host threads) race on it.
#include <mutex>

#include "ibim_internal.hpp"

namespace ibim {

namespace {

// 64 MiB: the partial planes of a product over up to ~26,000 tube points
// (pair layout: 2 (ceil(N / 1024) + ceil(N / 512) + 1) planes of N complex
// FP64 values), plus the plane-summing kernel's block partials and flag.
constexpr size_t kScratchBytes = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

}  // namespace

// The device address of `scratch` on `device`, which must be current: each
// GPU holds its own copy at its own address, resolved once and kept.
int noalloc_scratch(int device, void** ptr) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("device %d out of range for the no-allocation scratch table", device);
        return IBIM_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            set_error("hipGetSymbolAddress(scratch) on device %d failed: %s", device, hipGetErrorString(e));
            return IBIM_ERR_HIP;
        }
        g_ptr[device] = p;
    }
    *ptr = g_ptr[device];
    return IBIM_OK;
}

}  // namespace ibim

using namespace ibim;

extern "C" int ibim_apply_noalloc(ibim_ctx* ctx, int op, const void* p_dev, const void* q_dev, void* y_dev,
                                  void* z_dev) {
    if (!ctx) {
        set_error("null handle");
        return IBIM_ERR_ARG;
    }
    if (hipSetDevice(ctx->device) != hipSuccess) return IBIM_ERR_HIP;
    const int n = ctx->count;
    if (!n || !p_dev || !q_dev || !y_dev || !z_dev || op < IBIM_OP_STORED || op > IBIM_OP_FREE32 ||
        (op == IBIM_OP_STORED && ctx->mat_precision < 0)) {
        set_error("ibim_apply_noalloc: needs an operator, an op and four device vectors");
        return IBIM_ERR_ARG;
    }
    const int kinds[3] = {OP_DENSE, OP_FREE64, OP_FREE32};
    const int kop = kinds[op];
    if (op != IBIM_OP_STORED) {
        if (!ctx->have_tube) {
            set_error("ibim_apply_noalloc: matrix-free operators need a tube");
            return IBIM_ERR_ARG;
        }
        // the traditional tube's point table and near lists are allocated on
        // first use: that must have happened already
        if (free_pair_op(ctx, kop) && !ctx->pair_aux) {
            set_error("ibim_apply_noalloc: the matrix-free tables are not built (one ibim_apply or ibim_assemble of "
                      "pair storage builds them)");
            return IBIM_ERR_ARG;
        }
        // and so is the FP64 operator's cache (ibim_set_free_cache)
        if (kop == OP_FREE64 && free_pair_op(ctx, kop) && ctx->free_cache_max > 0 && !ctx->free_cache_valid) {
            set_error("ibim_apply_noalloc: the FP64 matrix-free cache is not built (one ibim_apply with OP_FREE64 "
                      "builds it)");
            return IBIM_ERR_ARG;
        }
    }
    const size_t need = matvec_scratch_bytes(ctx, kop);
    if (need > kScratchBytes) {
        set_error("ibim_apply_noalloc: %d points need %zu bytes of scratch, above the %zu-byte limit", n, need,
                  kScratchBytes);
        return IBIM_ERR_SIZE;
    }
    void* ptr = nullptr;
    if (int rc = noalloc_scratch(ctx->device, &ptr)) return rc;
    // no timing events either: the product runs unprofiled
    const bool prof = ctx->profile;
    ctx->profile = false;
    const int rc = matvec_op_scratch(ctx, kop, p_dev, q_dev, y_dev, z_dev, ptr);
    ctx->profile = prof;
    if (rc) return rc;
    IBIM_HIP(hipStreamSynchronize(ctx->stream));
    return IBIM_OK;
}
