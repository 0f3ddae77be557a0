// ibim_wave.hpp -- wavefront reductions shared by the sweep kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace ibim {

// Sum 8 per-lane values over the 64-lane wavefront, transposed: three halving
// exchanges (xor 32, 16, 8) leave each lane with one value index, then a
// three-step butterfly (xor 4, 2, 1) completes it. Lane l ends with the full
// sum of value ((l >> 5) & 1) * 4 + ((l >> 4) & 1) * 2 + ((l >> 3) & 1):
// 10 shuffles instead of 48 for eight independent butterflies.
__device__ __forceinline__ double wave_sum8_transposed(const double (&v)[8], int lane) {
    const bool h5 = lane & 32, h4 = lane & 16, h3 = lane & 8;
    double w[4], u[2];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double send = h5 ? v[i] : v[i + 4];
        const double keep = h5 ? v[i + 4] : v[i];
        w[i] = keep + __shfl_xor(send, 32);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const double send = h4 ? w[i] : w[i + 2];
        const double keep = h4 ? w[i + 2] : w[i];
        u[i] = keep + __shfl_xor(send, 16);
    }
    double z = (h3 ? u[1] : u[0]) + __shfl_xor(h3 ? u[0] : u[1], 8);
    z += __shfl_xor(z, 4);
    z += __shfl_xor(z, 2);
    z += __shfl_xor(z, 1);
    return z;
}

// The same transposed sum on gfx950's lane-exchange hardware instead of the LDS
// crossbar (ds_bpermute). xor 32 and xor 16: v_permlane32_swap /
// v_permlane16_swap exchange half-waves / odd and even 16-lane rows of two
// registers, so one swap of (v[i], v[i + 4]) followed by an add leaves every
// lane with (own + partner) of the value its half keeps -- no selects. xor 8:
// DPP row_ror:8. The last three steps (DPP row_shl:4, quad_perm xor 2, xor 1)
// are only exact in lanes with (lane & 7) == 0, the lanes that hold a
// result; their additions are grouped as in wave_sum8_transposed, so those
// lanes get the same bits.
template <int Ctrl>
__device__ __forceinline__ double dpp_f64(double v) {
    // bound_ctrl: lanes without a source read 0, so no old value is needed
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, Ctrl, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), Ctrl, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double swap32_sum(double x, double y) {
    const long long xb = __double_as_longlong(x), yb = __double_as_longlong(y);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)xb, (unsigned)yb, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(xb >> 32), (unsigned)(yb >> 32), false, false);
    return __longlong_as_double(((long long)hi[0] << 32) | lo[0]) +
           __longlong_as_double(((long long)hi[1] << 32) | lo[1]);
}

__device__ __forceinline__ double swap16_sum(double x, double y) {
    const long long xb = __double_as_longlong(x), yb = __double_as_longlong(y);
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)xb, (unsigned)yb, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(xb >> 32), (unsigned)(yb >> 32), false, false);
    return __longlong_as_double(((long long)hi[0] << 32) | lo[0]) +
           __longlong_as_double(((long long)hi[1] << 32) | lo[1]);
}

// Result valid in lanes with (lane & 7) == 0 only.
__device__ __forceinline__ double wave_sum8_transposed_x(const double (&v)[8], int lane) {
    double w[4], u[2];
#pragma unroll
    for (int i = 0; i < 4; ++i) w[i] = swap32_sum(v[i], v[i + 4]);
#pragma unroll
    for (int i = 0; i < 2; ++i) u[i] = swap16_sum(w[i], w[i + 2]);
    const bool h3 = lane & 8;
    double z = (h3 ? u[1] : u[0]) + dpp_f64<0x128>(h3 ? u[0] : u[1]);  // row_ror:8 = xor 8
    z += dpp_f64<0x104>(z);                                            // row_shl:4: lane l <- l + 4
    z += dpp_f64<0x4E>(z);                                             // quad_perm [2,3,0,1]
    z += dpp_f64<0xB1>(z);                                             // quad_perm [1,0,3,2]
    return z;
}

// Four per-lane values summed as wave_sum8_transposed_x sums values 0, 1, 4, 5
// (the same pairings: xor 32, 16, 8, 4, 2, 1, so the same bits), for sweeps
// without their second pair of sums. Lane l ends with value
// ((l >> 5) & 1) * 2 + ((l >> 4) & 1); valid in lanes with (lane & 15) == 0.
__device__ __forceinline__ double wave_sum4_transposed_x(const double (&v)[4]) {
    const double w0 = swap32_sum(v[0], v[2]), w1 = swap32_sum(v[1], v[3]);
    double z = swap16_sum(w0, w1);
    z += dpp_f64<0x128>(z);  // row_ror:8 = xor 8
    z += dpp_f64<0x104>(z);  // row_shl:4: lane l <- l + 4
    z += dpp_f64<0x4E>(z);   // quad_perm [2,3,0,1]
    z += dpp_f64<0xB1>(z);   // quad_perm [1,0,3,2]
    return z;
}

// Streaming (non-temporal) loads of a stored operator: read once per sweep, it
// should not displace the vectors and point data from the caches (the pair
// sweeps ran 10 % faster with them on 64,676 points).
template <class MT>
__device__ __forceinline__ MT load_stream(const MT* p);

template <>
__device__ __forceinline__ float2 load_stream<float2>(const float2* p) {
    const long long v = __builtin_nontemporal_load((const long long*)p);
    float2 r;
    __builtin_memcpy(&r, &v, sizeof r);
    return r;
}

template <>
__device__ __forceinline__ double2 load_stream<double2>(const double2* p) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2 v = __builtin_nontemporal_load((const d2*)p);
    return make_double2(v.x, v.y);
}

// The value index lane `lane` holds after wave_sum8_transposed(_x).
__device__ __forceinline__ int wave_sum8_index(int lane) {
    return ((lane >> 5) & 1) * 4 + ((lane >> 4) & 1) * 2 + ((lane >> 3) & 1);
}

}  // namespace ibim
