// Mixed-precision contraction: 16-bit hi/lo operand splits on MFMA.
//
// q_c = -|| M S x_c ||^2 with M S and X each written as hi + lo in a 16-bit
// format; Z = Mh Xh + Mh Xl + Ml Xh (the lo*lo term, ~u^2 relative, is
// dropped) runs on v_mfma_f32_16x16x32_{bf16,f16} with FP32 accumulation; the
// row scale is undone and the squares are summed in FP64 in the epilogue.
//
// Operand fragments (wave64, 16 x 16 x 32): lane l holds 8 consecutive k of
// row (A) / column (B) l % 16, k-group l / 16.  A and B use the same k
// assignment, so the product does not depend on how the hardware orders k
// inside the instruction; results: D[4 * (l / 16) + r][l % 16].
#include "mgw_kernels.h"

#include <hip/hip_runtime.h>

namespace mgw {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

__device__ inline uint16_t to_bf16_bits(float f) {
    uint32_t b = __float_as_uint(f);
    b += 0x7FFFu + ((b >> 16) & 1u);          // round to nearest even
    return (uint16_t)(b >> 16);
}
__device__ inline float from_bf16_bits(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline uint16_t to_f16_bits(float f) {
    const _Float16 h = (_Float16)f;
    return __builtin_bit_cast(uint16_t, h);
}
__device__ inline float from_f16_bits(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }

template <Split16 F>
__device__ inline void split_value(double x, uint16_t& hi, uint16_t& lo) {
    if constexpr (F == Split16::BF16) {
        hi = to_bf16_bits((float)x);
        lo = to_bf16_bits((float)(x - (double)from_bf16_bits(hi)));
    } else {
        hi = to_f16_bits((float)x);
        lo = to_f16_bits((float)(x - (double)from_f16_bits(hi)));
    }
}

// power of two s with max * s <= 2^14 (FP16 headroom); 1 for an all-zero vector
__device__ inline double scale_for_max(double mx) {
    if (!(mx > 0.0)) return 1.0;
    int e;
    frexp(mx, &e);                   // mx = f 2^e, f in [0.5, 1)
    return ldexp(1.0, 14 - e);
}

__device__ inline double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// One wave per column c of X.
template <Split16 F>
__global__ __launch_bounds__(256) void k_split_x(const double* __restrict__ src,
                                                 const int64_t* __restrict__ pkrow,
                                                 const int* __restrict__ colidx,
                                                 uint16_t* __restrict__ xhi,
                                                 uint16_t* __restrict__ xlo,
                                                 double* __restrict__ xscale, int n, int np,
                                                 int mpad) {
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (c >= mpad) return;
    const int64_t base = pkrow[c];
    double scale = 1.0;
    if constexpr (F == Split16::FP16) {
        double mx = 0.0;
        for (int k = lane; k < n; k += 64) mx = fmax(mx, fabs(src[base + colidx[k]]));
        scale = scale_for_max(wave_max(mx));
    }
    uint16_t* oh = xhi + (size_t)c * np;
    uint16_t* ol = xlo + (size_t)c * np;
    for (int k = lane; k < np; k += 64) {
        const double x = k < n ? src[base + colidx[k]] * scale : 0.0;
        uint16_t h, l;
        split_value<F>(x, h, l);
        oh[k] = h;
        ol[k] = l;
    }
    if (lane == 0) xscale[c] = scale;
}

// One wave per row r of M_w S_w (lower triangle, zeros above).
template <Split16 F>
__global__ __launch_bounds__(256) void k_split_m(const double* __restrict__ M,
                                                 const double* __restrict__ S,
                                                 uint16_t* __restrict__ mhi,
                                                 uint16_t* __restrict__ mlo,
                                                 double* __restrict__ mscale, int np) {
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int w = blockIdx.y;
    if (r >= np) return;
    const double* row = M + (size_t)w * np * np + (size_t)r * np;
    const double* Sw = S + (size_t)w * np;
    double scale = 1.0;
    if constexpr (F == Split16::FP16) {
        double mx = 0.0;
        for (int k = lane; k <= r; k += 64) mx = fmax(mx, fabs(row[k] * Sw[k]));
        scale = scale_for_max(wave_max(mx));
    }
    uint16_t* oh = mhi + (size_t)w * np * np + (size_t)r * np;
    uint16_t* ol = mlo + (size_t)w * np * np + (size_t)r * np;
    for (int k = lane; k < np; k += 64) {
        const double x = k <= r ? (row[k] * Sw[k]) * scale : 0.0;
        uint16_t h, l;
        split_value<F>(x, h, l);
        oh[k] = h;
        ol[k] = l;
    }
    if (lane == 0) mscale[(size_t)w * np + r] = scale;
}

template <Split16 F>
__device__ inline f4 mfma16(uint4 a, uint4 b, f4 c) {
    if constexpr (F == Split16::BF16) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                       __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
    } else {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a),
                                                      __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
    }
}

// Tile: rows [r0, r0+128) x columns [c0, c0+128); 4 waves in 2 x 2 of 64 x 64.
// Operand fragments are loaded straight from HBM/L2 into registers (16 B per lane).
template <Split16 F>
__global__ __launch_bounds__(256) void k_contract_split(
    const uint16_t* __restrict__ mhi, const uint16_t* __restrict__ mlo,
    const double* __restrict__ mscale, const uint16_t* __restrict__ xhi,
    const uint16_t* __restrict__ xlo, double* __restrict__ qpart, int np, int mpad) {
    __shared__ double red[2][kCT];
    const int w = blockIdx.z, rt = blockIdx.y, r0 = rt * kCT, c0 = blockIdx.x * kCT;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int li = lane & 15, lk = lane >> 4;
    const size_t mo = (size_t)w * np * np;

    const uint16_t* arow_h[4];
    const uint16_t* arow_l[4];
    const uint16_t* bcol_h[4];
    const uint16_t* bcol_l[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const size_t ro = mo + (size_t)(r0 + wm + q * 16 + li) * np + 8 * lk;
        const size_t co = (size_t)(c0 + wn + q * 16 + li) * np + 8 * lk;
        arow_h[q] = mhi + ro;
        arow_l[q] = mlo + ro;
        bcol_h[q] = xhi + co;
        bcol_l[q] = xlo + co;
    }
    f4 acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = f4{0.f, 0.f, 0.f, 0.f};

    const int kend = min(np, r0 + kCT);
    for (int k0 = 0; k0 < kend; k0 += 32) {
        uint4 ah[4], al[4], bh[4], bl[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            ah[q] = *reinterpret_cast<const uint4*>(arow_h[q] + k0);
            al[q] = *reinterpret_cast<const uint4*>(arow_l[q] + k0);
            bh[q] = *reinterpret_cast<const uint4*>(bcol_h[q] + k0);
            bl[q] = *reinterpret_cast<const uint4*>(bcol_l[q] + k0);
        }
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                acc[x][y] = mfma16<F>(ah[x], bh[y], acc[x][y]);
                acc[x][y] = mfma16<F>(ah[x], bl[y], acc[x][y]);
                acc[x][y] = mfma16<F>(al[x], bh[y], acc[x][y]);
            }
    }

    double inv_rs[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            inv_rs[x][r] = 1.0 / mscale[(size_t)w * np + r0 + wm + x * 16 + 4 * lk + r];
#pragma unroll
    for (int y = 0; y < 4; ++y) {
        double s = 0.0;
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double z = (double)acc[x][y][r] * inv_rs[x][r];
                s += z * z;
            }
        s += __shfl_xor(s, 16);
        s += __shfl_xor(s, 32);
        if (lk == 0) red[wave >> 1][wn + y * 16 + li] = s;
    }
    __syncthreads();
    if (t < kCT) qpart[((size_t)w * (np / kCT) + rt) * mpad + c0 + t] = red[0][t] + red[1][t];
}

// Diagnostic: one 16x16x32 MFMA per wave; a/b: [t][lane][8] 16-bit patterns.
template <Split16 F>
__global__ __launch_bounds__(64) void k_probe16(const uint4* __restrict__ a,
                                                const uint4* __restrict__ b,
                                                float* __restrict__ d) {
    const int t = blockIdx.x, lane = threadIdx.x;
    f4 acc = f4{0.f, 0.f, 0.f, 0.f};
    acc = mfma16<F>(a[t * 64 + lane], b[t * 64 + lane], acc);
    for (int r = 0; r < 4; ++r) d[((size_t)t * 64 + lane) * 4 + r] = acc[r];
}

__global__ void k_reduce_q_scaled(const double* __restrict__ qpart,
                                  const double* __restrict__ xscale, double* __restrict__ q,
                                  int nrt, int m, int mpad, int f0) {
    const int w = blockIdx.y;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < m; c += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int rt = 0; rt < nrt; ++rt) s += qpart[((size_t)w * nrt + rt) * mpad + c];
        const double xs = xscale[c];
        q[(size_t)(f0 + w) * m + c] = -s / (xs * xs);
    }
}

}  // namespace

hipError_t launch_split_x(Split16 fmt, const double* src, const int64_t* pkrow, const int* colidx,
                          uint16_t* xhi, uint16_t* xlo, double* xscale, int n, int np, int mpad,
                          hipStream_t s) {
    const dim3 grid((mpad + 3) / 4);
    if (fmt == Split16::BF16)
        hipLaunchKernelGGL(k_split_x<Split16::BF16>, grid, dim3(256), 0, s, src, pkrow, colidx,
                           xhi, xlo, xscale, n, np, mpad);
    else
        hipLaunchKernelGGL(k_split_x<Split16::FP16>, grid, dim3(256), 0, s, src, pkrow, colidx,
                           xhi, xlo, xscale, n, np, mpad);
    return hipGetLastError();
}

hipError_t launch_split_m(Split16 fmt, const double* M, const double* S, uint16_t* mhi,
                          uint16_t* mlo, double* mscale, int np, int nb, hipStream_t s) {
    const dim3 grid((np + 3) / 4, nb);
    if (fmt == Split16::BF16)
        hipLaunchKernelGGL(k_split_m<Split16::BF16>, grid, dim3(256), 0, s, M, S, mhi, mlo, mscale,
                           np);
    else
        hipLaunchKernelGGL(k_split_m<Split16::FP16>, grid, dim3(256), 0, s, M, S, mhi, mlo, mscale,
                           np);
    return hipGetLastError();
}

hipError_t launch_contract_split(Split16 fmt, const uint16_t* mhi, const uint16_t* mlo,
                                 const double* mscale, const uint16_t* xhi, const uint16_t* xlo,
                                 double* qpart, int np, int mpad, int nb, hipStream_t s) {
    const dim3 grid(mpad / kCT, np / kCT, nb);
    if (fmt == Split16::BF16)
        hipLaunchKernelGGL(k_contract_split<Split16::BF16>, grid, dim3(256), 0, s, mhi, mlo, mscale,
                           xhi, xlo, qpart, np, mpad);
    else
        hipLaunchKernelGGL(k_contract_split<Split16::FP16>, grid, dim3(256), 0, s, mhi, mlo, mscale,
                           xhi, xlo, qpart, np, mpad);
    return hipGetLastError();
}

hipError_t launch_probe16(Split16 fmt, const uint16_t* a, const uint16_t* b, float* d,
                         int ntests, hipStream_t s) {
    const uint4* a4 = reinterpret_cast<const uint4*>(a);
    const uint4* b4 = reinterpret_cast<const uint4*>(b);
    if (fmt == Split16::BF16)
        hipLaunchKernelGGL(k_probe16<Split16::BF16>, dim3(ntests), dim3(64), 0, s, a4, b4, d);
    else
        hipLaunchKernelGGL(k_probe16<Split16::FP16>, dim3(ntests), dim3(64), 0, s, a4, b4, d);
    return hipGetLastError();
}

hipError_t launch_reduce_q_scaled(const double* qpart, const double* xscale, double* q, int np,
                                  int m, int mpad, int f0, int nb, hipStream_t s) {
    const int gx = (int)((m + 255) / 256 < 4096 ? (m + 255) / 256 : 4096);
    hipLaunchKernelGGL(k_reduce_q_scaled, dim3(gx > 0 ? gx : 1, nb), dim3(256), 0, s, qpart, xscale,
                       q, np / kCT, m, mpad, f0);
    return hipGetLastError();
}

}  // namespace mgw
