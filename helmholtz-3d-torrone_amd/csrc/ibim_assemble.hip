// ibim_assemble.hip -- dense operator assembly in HBM.
//
// The operator is stored row-major as interleaved complex values (double2 for
// IBIM_C128, float2 for IBIM_C64) with a leading dimension ld = count rounded up
// to 64 entries, so every row starts on a 1 KiB (c128) boundary; padding columns
// hold zeros. One thread owns one column of a kAssembleCols-wide strip and walks
// kAssembleRows rows: the column point's data stays in registers, the row
// point's data is a wave-uniform (scalar) load, and each row's stores are one
// contiguous 1 KiB (c128) burst per wavefront.
#include "ibim_entries.hpp"
#include "ibim_internal.hpp"

namespace ibim {

namespace {

constexpr int kAssembleCols = 256;
constexpr int kAssembleRows = 8;

template <class Store>
__device__ __forceinline__ void store_entry(Store* dst, double re, double im);

template <>
__device__ __forceinline__ void store_entry<double2>(double2* dst, double re, double im) {
    *dst = make_double2(re, im);
}

template <>
__device__ __forceinline__ void store_entry<float2>(float2* dst, double re, double im) {
    *dst = make_float2((float)re, (float)im);
}

template <class Store, int Kind>
__global__ void __launch_bounds__(kAssembleCols) k_assemble(const double* __restrict__ tube, int count, size_t ld,
                                                            ibim_params prm, Store* __restrict__ mat,
                                                            int* __restrict__ regs) {
    const int j = blockIdx.x * kAssembleCols + threadIdx.x;
    const int i0 = blockIdx.y * kAssembleRows;
    const size_t c = (size_t)count;
    PointData b{};
    if (j < count) b = load_point(tube, c, j);
    int near = 0;
    for (int r = 0; r < kAssembleRows; ++r) {
        const int i = i0 + r;
        if (i >= count) break;
        double re = 0.0, im = 0.0;
        if (j < count) {
            PointData a = load_point(tube, c, i);
            if (Kind == IBIM_TUBE_TRADITIONAL)
                trad_entry(a, b, i == j, prm.k, prm.tau, prm.dirichlet, prm.interior, prm.combined, prm.eta, &re,
                           &im);
            else
                near += combo_entry(a, b, i == j, prm.k, prm.tau, prm.eta, &re, &im);
        }
        if ((size_t)j < ld) store_entry<Store>(&mat[(size_t)i * ld + j], re, im);
    }
    if (Kind != IBIM_TUBE_TRADITIONAL) {
        for (int off = 32; off > 0; off >>= 1) near += __shfl_xor(near, off);
        if ((threadIdx.x & 63) == 0 && near) atomicAdd(regs, near);
    }
}

template <class Store>
int assemble_as(Context* ctx) {
    const int count = ctx->count;
    dim3 grid((unsigned)((ctx->ld + kAssembleCols - 1) / kAssembleCols),
              (unsigned)((count + kAssembleRows - 1) / kAssembleRows));
    Store* mat = ctx->mat.as<Store>();
    int* regs = ctx->err_flag.as<int>();
    IBIM_HIP(hipMemsetAsync(regs, 0, sizeof(int), ctx->stream));
    {
        TimedScope ts(ctx, "assemble");
        if (ctx->kind == IBIM_TUBE_TRADITIONAL)
            hipLaunchKernelGGL((k_assemble<Store, IBIM_TUBE_TRADITIONAL>), grid, dim3(kAssembleCols), 0, ctx->stream,
                               ctx->tube.as<double>(), count, ctx->ld, ctx->prm, mat, regs);
        else
            hipLaunchKernelGGL((k_assemble<Store, IBIM_TUBE_COMBO>), grid, dim3(kAssembleCols), 0, ctx->stream,
                               ctx->tube.as<double>(), count, ctx->ld, ctx->prm, mat, regs);
    }
    IBIM_HIP(hipGetLastError());
    if (ctx->kind != IBIM_TUBE_TRADITIONAL) {
        IBIM_HIP(hipMemcpyAsync(&ctx->regcount, regs, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        IBIM_HIP(hipStreamSynchronize(ctx->stream));
    } else {
        ctx->regcount = 0;
    }
    return IBIM_OK;
}

}  // namespace

int launch_assemble(Context* ctx, int precision) {
    const size_t esz = precision == IBIM_C128 ? sizeof(double2) : sizeof(float2);
    ctx->ld = ((size_t)ctx->count + 63) / 64 * 64;
    if (ctx->mat.ensure(esz * ctx->ld * (size_t)ctx->count)) return IBIM_ERR_NOMEM;
    if (ctx->err_flag.ensure(sizeof(int))) return IBIM_ERR_NOMEM;
    ctx->mat_precision = precision;
    ctx->mat_pair = ctx->mat_pair64 = false;
    return precision == IBIM_C128 ? assemble_as<double2>(ctx) : assemble_as<float2>(ctx);
}

}  // namespace ibim
