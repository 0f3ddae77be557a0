// ibim_mfma.hpp -- complex products on the FP64 matrix cores (16 right-hand
// sides per pass), shared by the batched passes over dense and pair storage.
#pragma once

#include <hip/hip_runtime.h>

namespace ibim {

constexpr int kBatchRHS = 16;  // right-hand sides per batched pass: the MFMA N dimension

typedef double d4 __attribute__((ext_vector_type(4)));

// D = A B + C for a 16x4 A, 4x16 B, 16x16 D. Lane l holds A[l % 16][l / 16],
// B[l / 16][l % 16], and D[4 v + l / 16][l % 16] in component v.
__device__ __forceinline__ d4 mfma(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Complex accumulate with three real products (Gauss): for a = ar + i ai and
// b = br + i bi, k1 += br (ar + ai), k2 += ar (bi - br), k3 += ai (br + bi);
// then re = k1 - k3, im = k1 + k2. Three independent accumulators also keep
// consecutive matrix-core instructions free of dependences.
struct Acc3 {
    d4 k1 = {0, 0, 0, 0}, k2 = {0, 0, 0, 0}, k3 = {0, 0, 0, 0};
    __device__ __forceinline__ void add(double ar, double ai, double br, double bi) {
        k1 = mfma(ar + ai, br, k1);
        k2 = mfma(ar, bi - br, k2);
        k3 = mfma(ai, br + bi, k3);
    }
    __device__ __forceinline__ double re(int v) const { return k1[v] - k3[v]; }
    __device__ __forceinline__ double im(int v) const { return k1[v] + k2[v]; }
};

}  // namespace ibim
