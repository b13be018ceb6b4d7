// Power-of-two scaling and int8 slicing shared by the integer-MFMA kernels.
//
// A value a in a row (or column) whose largest magnitude is below 2^e is
// written as a = 2^e sum_i t_i 2^-(6 + 7 i), i < kSlices, t_i integer in
// [-64, 64]: the first slice takes 6 bits of a 2^-e, every later one the
// rounded remainder (every step exact in FP64), so the representation error
// is below 2^(e - 6 - 7 kSlices).
#pragma once

#include <hip/hip_runtime.h>

#include "mgw_kernels.h"

namespace mgw {

// scale 2^-e with max|v| < 2^e (1 for an all-zero vector)
__device__ inline double slice_scale(double mx) {
    if (!(mx > 0.0)) return 1.0;
    int e;
    frexp(mx, &e);                  // mx = f 2^e, f in [0.5, 1)
    return ldexp(1.0, -e);
}

// r in (-1, 1) -> kSlices slices
__device__ inline void slice_value(double r, int8_t* t) {
    double x = r * 64.0;
    double q = rint(x);
    t[0] = (int8_t)q;
#pragma unroll
    for (int i = 1; i < kSlices; ++i) {
        x = (x - q) * 128.0;
        q = rint(x);
        t[i] = (int8_t)q;
    }
}

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ inline i32x16 mfma_i8(u32x4 a, u32x4 b, i32x16 c) {
    return __builtin_amdgcn_mfma_i32_32x32x32_i8(__builtin_bit_cast(i32x4, a),
                                                  __builtin_bit_cast(i32x4, b), c, 0, 0, 0);
}
// result d[r] of the 32x32 MFMA forms is D[row32(lane, r)][lane % 32]
__device__ inline int row32(int lane, int r) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// sum over groups g < kSlices of acc[g] 2^-(12 + 7 g) (Horner in 2^-7)
template <class Acc>
__device__ inline double combine_groups(const Acc* acc, int r) {
    double z = (double)acc[kSlices - 1][r];
#pragma unroll
    for (int g = kSlices - 2; g >= 0; --g) z = z * 0x1p-7 + (double)acc[g][r];
    return z * 0x1p-12;
}

}  // namespace mgw
