// Power-of-two scaling and int8 digit splitting shared by the integer-MFMA
// kernels.
//
// A value a of a row (or column) scaled by 2^-e is written with S signed
// digits of B bits; two schemes:
//
//   B = 7: a = 2^e sum_i t_i 2^-(6 + 7 i), t_i integer in [-64, 64], with
//          2^e above the row's largest magnitude: the first digit takes 6
//          bits of a 2^-e, every later one the rounded remainder (each step
//          exact in FP64); representation error below 2^(e - 6 - 7 S).
//   B = 8: a = 2^e sum_i d_i 2^-(7 + 8 i), d_i integer in [-128, 127]: the
//          balanced base-256 digits of the integer A = rint(a 2^(8 S - 1 - e)),
//          cut from the bottom (each digit the low byte read as signed, i.e.
//          ((A + 128) mod 256) - 128, then the exact shift (A - d) / 256).
//          2^e >= 1.015625 x the largest magnitude keeps |A| below
//          2^(8 S - 1) x 127/128, so the top digit lies in [-127, 127];
//          representation error below 2^(e - 8 S).
//
// A product of two such values is sum over digit pairs (i, j) of
// t_i u_j 2^-(2 (B - 1) + B (i + j)); pairs with equal i + j share one int32
// accumulator (a "group") and only the groups i + j < S are formed, which
// truncates the product at the order of the representation error.
#pragma once

#include <hip/hip_runtime.h>

#include "mgw_kernels.h"

namespace mgw {

// scale 2^-e of a row whose largest magnitude is mx (1 for an all-zero row)
template <int B>
__device__ inline double digit_scale(double mx) {
    if (!(mx > 0.0)) return 1.0;
    int e;
    frexp(B == 8 ? mx * 1.015625 : mx, &e);   // x = f 2^e, f in [0.5, 1)
    return ldexp(1.0, -e);
}

// r = a 2^-e (|r| < 1, and <= 1/1.015625 for B = 8) -> S digits, most significant first
template <int S, int B>
__device__ inline void digit_split(double r, int8_t* t) {
    static_assert(B == 7 || B == 8, "digit width");
    if constexpr (B == 7) {
        double x = r * 64.0;
        double q = rint(x);
        t[0] = (int8_t)q;
#pragma unroll
        for (int i = 1; i < S; ++i) {
            x = (x - q) * 128.0;
            q = rint(x);
            t[i] = (int8_t)q;
        }
    } else {
        static_assert(8 * S - 1 <= 62, "digit count");
        // exact: |r 2^(8S-1)| < 2^(8S-1) <= 2^53 needs S <= 6; rint rounds the
        // bits beyond the last digit
        long long a = (long long)rint(r * (double)(1ll << (8 * S - 1)));
#pragma unroll
        for (int i = S - 1; i > 0; --i) {
            const int d = (int)((a + 128) & 0xff) - 128;
            t[i] = (int8_t)d;
            a = (a - d) >> 8;                 // exact: a - d is a multiple of 256
        }
        t[0] = (int8_t)a;
    }
}

// weight of group 0 (2^-(2 (B - 1))) and the ratio between groups (2^-B)
template <int B>
__device__ constexpr double group0_weight() { return B == 8 ? 0x1p-14 : 0x1p-12; }
template <int B>
__device__ constexpr double group_step() { return B == 8 ? 0x1p-8 : 0x1p-7; }

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ inline i32x16 mfma_i8(u32x4 a, u32x4 b, i32x16 c) {
    return __builtin_amdgcn_mfma_i32_32x32x32_i8(__builtin_bit_cast(i32x4, a),
                                                  __builtin_bit_cast(i32x4, b), c, 0, 0, 0);
}
// result d[r] of the 32x32 MFMA forms is D[row32(lane, r)][lane % 32]
__device__ inline int row32(int lane, int r) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

}  // namespace mgw
