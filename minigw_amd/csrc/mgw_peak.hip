// Peak calibration kernels: MFMA issue rate on register-resident operands and
// HBM streaming bandwidth.  Used to state measured roofline peaks beside the
// vendor figures.
#include "mgw_kernels.h"

#include <hip/hip_runtime.h>

namespace mgw {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));

// kind 0: v_mfma_f64_16x16x4 (2048 flop; k_mfma_peak_f64), 1: f16 16x16x32 (16384),
// 2: bf16 16x16x32 (16384), 3: f16 32x32x16 (32768, 4 accumulators), 4: bf16 32x32x16,
// 5: i8 32x32x32 (65536 op, 4 accumulators), 6: i8 16x16x64 (32768 op, 8 accumulators)
template <int KIND>
__global__ __launch_bounds__(256) void k_mfma_peak(int iters, float* sink) {
    const int lane = threadIdx.x & 63;
    if constexpr (KIND == 5) {
        typedef int i16v __attribute__((ext_vector_type(16)));
        typedef int i4v __attribute__((ext_vector_type(4)));
        i16v acc[4];
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 16; ++j) acc[i][j] = 0;
        const i4v av = {0x01010101 + lane, 0x01010101, 0x01010101, 0x01010101};
        const i4v bv = {0x01010101, 0x01010101 + lane, 0x01010101, 0x01010101};
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int i = 0; i < 4; ++i) acc[i] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, bv, acc[i], 0, 0, 0);
        int s = 0;
        for (int i = 0; i < 4; ++i) s += acc[i][0] + acc[i][15];
        if (s == 12345) sink[threadIdx.x] = (float)s;
    } else if constexpr (KIND == 6) {
        typedef int i4v __attribute__((ext_vector_type(4)));
        i4v acc[8];
        for (int i = 0; i < 8; ++i) acc[i] = i4v{0, 0, 0, 0};
        const i4v av = {0x01010101 + lane, 0x01010101, 0x01010101, 0x01010101};
        const i4v bv = {0x01010101, 0x01010101 + lane, 0x01010101, 0x01010101};
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int i = 0; i < 8; ++i) acc[i] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, bv, acc[i], 0, 0, 0);
        int s = 0;
        for (int i = 0; i < 8; ++i) s += acc[i][0] + acc[i][3];
        if (s == 12345) sink[threadIdx.x] = (float)s;
    } else if constexpr (KIND >= 3) {
        typedef float f16v __attribute__((ext_vector_type(16)));
        f16v acc[4];
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 16; ++j) acc[i][j] = 0.f;
        uint4 av = make_uint4(0x3c003c00u + lane, 0x3c003c00u, 0x3c003c00u, 0x3c003c00u);
        uint4 bv = make_uint4(0x3c003c00u, 0x3c003c00u + lane, 0x3c003c00u, 0x3c003c00u);
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if constexpr (KIND == 3)
                    acc[i] = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(h8, av),
                                                                    __builtin_bit_cast(h8, bv), acc[i], 0, 0, 0);
                else
                    acc[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(b8, av),
                                                                     __builtin_bit_cast(b8, bv), acc[i], 0, 0, 0);
            }
        float s = 0;
        for (int i = 0; i < 4; ++i) s += acc[i][0] + acc[i][15];
        if (s == 12345.0f) sink[threadIdx.x] = s;
    } else {
        f4 acc[8];
        for (int i = 0; i < 8; ++i) acc[i] = f4{0, 0, 0, 0};
        uint4 av = make_uint4(0x3c003c00u + lane, 0x3c003c00u, 0x3c003c00u, 0x3c003c00u);
        uint4 bv = make_uint4(0x3c003c00u, 0x3c003c00u + lane, 0x3c003c00u, 0x3c003c00u);
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                if constexpr (KIND == 1)
                    acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8, av),
                                                                    __builtin_bit_cast(h8, bv), acc[i], 0, 0, 0);
                else
                    acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8, av),
                                                                     __builtin_bit_cast(b8, bv), acc[i], 0, 0, 0);
            }
        float s = 0;
        for (int i = 0; i < 8; ++i) s += acc[i][0] + acc[i][3];
        if (s == 12345.0f) sink[threadIdx.x] = s;
    }
}

// Hash of (block, lane, index) to a pseudo-random 32-bit word.
__device__ inline unsigned mix32(unsigned x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// The same MFMAs on random operands: eight (a, b) fragment pairs per lane,
// cycled so that consecutive MFMAs see different data, as a contraction
// does.  MFMA throughput on gfx950 depends on the operand data (power): the
// constant-operand rates above are an upper bound that real slices (uniform
// in [-64, 64]) or FP64 values do not reach.
// kind 7: i8 32x32x32 with bytes uniform in [-64, 64] (kind 8, FP64 with
// values in [-1, 1), is k_mfma_peak_f64<true>).
template <int KIND>
__global__ __launch_bounds__(256) void k_mfma_peak_random(int iters, float* sink) {
    const int lane = threadIdx.x & 63;
    const unsigned seed = (blockIdx.x * 256u + threadIdx.x) * 2654435761u;
    if constexpr (KIND == 7) {
        typedef int i16v __attribute__((ext_vector_type(16)));
        typedef int i4v __attribute__((ext_vector_type(4)));
        i4v av[8], bv[8];
        for (int r = 0; r < 8; ++r)
            for (int q = 0; q < 4; ++q) {
                unsigned wa = 0, wb = 0;
                for (int by = 0; by < 4; ++by) {
                    const unsigned ha = mix32(seed + 97u * (r * 16 + q * 4 + by)) % 129u;
                    const unsigned hb = mix32(~seed + 89u * (r * 16 + q * 4 + by)) % 129u;
                    wa |= ((unsigned)(int8_t)((int)ha - 64) & 0xffu) << (8 * by);
                    wb |= ((unsigned)(int8_t)((int)hb - 64) & 0xffu) << (8 * by);
                }
                av[r][q] = (int)wa;
                bv[r][q] = (int)wb;
            }
        i16v acc[4];
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 16; ++j) acc[i][j] = 0;
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int r = 0; r < 8; ++r)
                acc[r & 3] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av[r], bv[r], acc[r & 3], 0, 0, 0);
        int s = 0;
        for (int i = 0; i < 4; ++i) s += acc[i][0] + acc[i][15];
        if (s == 12345 + lane) sink[threadIdx.x] = (float)s;
    }
}

// FP64 MFMA as a GEMM tile issues it: a 4 x 4 outer product of distinct
// operand registers per step (16 accumulators), two waves per SIMD; random
// operands in [-1, 1) when RANDOM.  The same MFMA issued on one operand pair
// reaches only ~60 % of this rate.
template <bool RANDOM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_mfma_peak_f64(
    int iters, float* sink) {
    const int lane = threadIdx.x & 63;
    const unsigned seed = (blockIdx.x * 256u + threadIdx.x) * 2654435761u;
    double a[4], b[4];
    for (int i = 0; i < 4; ++i) {
        a[i] = RANDOM ? (double)mix32(seed + 131u * i) * 0x1p-31 - 1.0 : 1.0 + (lane + 7 * i) * 1e-3;
        b[i] = RANDOM ? (double)mix32(~seed + 137u * i) * 0x1p-31 - 1.0 : 1.0 - (lane + 5 * i) * 1e-3;
    }
    d4 acc[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) acc[i][j] = d4{0, 0, 0, 0};
    for (int it = 0; it < iters; ++it)
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y)
                acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[y], acc[x][y], 0, 0, 0);
    double s = 0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) s += acc[i][j][0] + acc[i][j][3];
    if (s == 12345.0 + lane) sink[threadIdx.x] = (float)s;
}

__global__ void k_stream_copy(const double4* __restrict__ src, double4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace

hipError_t launch_mfma_peak(int kind, int blocks, int iters, float* sink, hipStream_t s) {
    if (kind == 0) hipLaunchKernelGGL(k_mfma_peak_f64<false>, dim3(blocks), dim3(256), 0, s, iters, sink);
    else if (kind == 1) hipLaunchKernelGGL(k_mfma_peak<1>, dim3(blocks), dim3(256), 0, s, iters, sink);
    else if (kind == 2) hipLaunchKernelGGL(k_mfma_peak<2>, dim3(blocks), dim3(256), 0, s, iters, sink);
    else if (kind == 3) hipLaunchKernelGGL(k_mfma_peak<3>, dim3(blocks), dim3(256), 0, s, iters, sink);
    else if (kind == 4) hipLaunchKernelGGL(k_mfma_peak<4>, dim3(blocks), dim3(256), 0, s, iters, sink);
    else if (kind == 5) hipLaunchKernelGGL(k_mfma_peak<5>, dim3(blocks), dim3(256), 0, s, iters, sink);
    else if (kind == 6) hipLaunchKernelGGL(k_mfma_peak<6>, dim3(blocks), dim3(256), 0, s, iters, sink);
    else if (kind == 7)
        hipLaunchKernelGGL(k_mfma_peak_random<7>, dim3(blocks), dim3(256), 0, s, iters, sink);
    else hipLaunchKernelGGL(k_mfma_peak_f64<true>, dim3(blocks), dim3(256), 0, s, iters, sink);
    return hipGetLastError();
}

hipError_t launch_stream_copy(const double* src, double* dst, size_t count, hipStream_t s) {
    const size_t n4 = count / 4;
    hipLaunchKernelGGL(k_stream_copy, dim3(4096), dim3(256), 0, s,
                       reinterpret_cast<const double4*>(src), reinterpret_cast<double4*>(dst), n4);
    return hipGetLastError();
}

}  // namespace mgw
