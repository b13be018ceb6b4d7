// Mixed-precision contraction: 16-bit hi/lo operand splits on MFMA.
//
// q_c = -|| M S x_c ||^2 with M S and X each written as hi + lo in a 16-bit
// format; Z = Mh Xh + Mh Xl + Ml Xh (the lo*lo term, ~u^2 relative, is
// dropped) runs on v_mfma_f32_32x32x16_{bf16,f16} with FP32 accumulation; the
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
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
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

// Plane layout: blocks of 256 rows x 32 k (8192 halves, row-major inside),
// blocks ordered by row tile then k block, so one K step of a 256-row tile is
// one contiguous 16 KB block per plane.
__device__ inline size_t tiled_index(int row, int k, int np) {
    return ((size_t)(row >> 8) * (np >> 5) + (k >> 5)) * 8192 + (size_t)(row & 255) * 32 + (k & 31);
}

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
    for (int k = lane; k < np; k += 64) {
        const double x = k < n ? src[base + colidx[k]] * scale : 0.0;
        uint16_t h, l;
        split_value<F>(x, h, l);
        const size_t o = tiled_index(c, k, np);
        xhi[o] = h;
        xlo[o] = l;
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
    uint16_t* oh = mhi + (size_t)w * np * np;
    uint16_t* ol = mlo + (size_t)w * np * np;
    for (int k = lane; k < np; k += 64) {
        const double x = k <= r ? (row[k] * Sw[k]) * scale : 0.0;
        uint16_t h, l;
        split_value<F>(x, h, l);
        const size_t o = tiled_index(r, k, np);
        oh[o] = h;
        ol[o] = l;
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

typedef float f16acc __attribute__((ext_vector_type(16)));

// v_mfma_f32_32x32x16_{bf16,f16}: a / b as for the 16x16 form but with 32 rows
// (columns) per operand (lane % 32) and k-group lane / 32; the result d[r] is
// D[row32(lane, r)][lane % 32].  On gfx950 this form issues at twice the rate
// of the 16x16x32 one (measured by mgw_hip_measure_peaks).
template <Split16 F>
__device__ inline f16acc mfma32(uint4 a, uint4 b, f16acc c) {
    if constexpr (F == Split16::BF16) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                       __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
    } else {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a),
                                                      __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
    }
}
__device__ inline int row32(int lane, int r) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// Workgroup tile: rows [r0, r0+256) x columns [c0, c0+256), 512 threads = 8
// waves in 4 (rows) x 2 (columns), each wave 64 x 128 (2 x 4 blocks of the
// 32x32x16 MFMA).
// Per K step of 32 the four operand planes (M hi/lo rows, X hi/lo columns,
// 16 B per thread-chunk, 8 chunks per thread) are loaded into registers one
// step ahead and stored to LDS (row stride 40 halves: 80 B keeps the 16 B
// fragment reads of a 16-lane group on distinct banks).  M is lower
// triangular, so the K loop stops at r0 + 256.  Tiles are visited large-K
// first (bottom row panels), so the short tiles fill the tail of the grid.
constexpr int kSBM = 256, kSBN = 256, kSBK = 32, kSLd = kSBK + 8;

// Workgroup -> tile map.  Workgroups are dispatched round-robin over the 8
// XCDs (workgroup L runs on XCD L % 8), each XCD with its own L2 and 32 CUs.
// The tiles are cut into units of uf frequencies x uc column tiles of one row
// panel (uf * uc <= 32).  The workgroups of a unit land on one XCD and run
// side by side, so every K step brings uf M slabs and uc X slabs into that L2
// (12 for 4 x 8) instead of the ~33 of a row-major order.  Units run row panel
// by row panel, large K first; the grid is padded to whole rounds of 8 units
// and the tiles beyond the edges exit at once.
struct SplitTile {
    int rt, ct, w;
    bool valid;
};
__device__ inline SplitTile split_tile(int L, int nrow, int ncol, int nb, int uf, int uc) {
    const int us = uf * uc;
    const int xcd = L & 7, s = L >> 3;
    const int u = (s / us) * 8 + xcd, j = s % us;
    const int ncg = (ncol + uc - 1) / uc, nfg = (nb + uf - 1) / uf;
    const int cg = u % ncg, fg = (u / ncg) % nfg, rr = u / (ncg * nfg);
    SplitTile tl;
    tl.rt = nrow - 1 - rr;
    tl.w = fg * uf + j / uc;
    tl.ct = cg * uc + j % uc;
    tl.valid = rr < nrow && tl.w < nb && tl.ct < ncol;
    return tl;
}

template <Split16 F>
__global__ __launch_bounds__(512) void k_contract_split(
    const uint16_t* __restrict__ mhi, const uint16_t* __restrict__ mlo,
    const double* __restrict__ mscale, const uint16_t* __restrict__ xhi,
    const uint16_t* __restrict__ xlo, double* __restrict__ qpart, int np, int mpad, int nb,
    int uf, int uc) {
    __shared__ __attribute__((aligned(16))) uint16_t lds[4][kSBM][kSLd];
    const SplitTile tile = split_tile(blockIdx.x, np / kSBM, mpad / kSBN, nb, uf, uc);
    if (!tile.valid) return;                      // uniform over the workgroup
    const int w = tile.w, rt = tile.rt;
    const int r0 = rt * kSBM, c0 = tile.ct * kSBN;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 128;

    // loader: chunk u of thread t -> plane u / 2, row (t + 512 (u % 2)) / 4, k (idx % 4) * 8
    const uint16_t* src[8];
    int lrow[8], lkq[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int plane = u >> 1, idx = t + 512 * (u & 1);
        lrow[u] = idx >> 2;
        lkq[u] = (idx & 3) * 8;
        const uint16_t* base = plane == 0 ? mhi : plane == 1 ? mlo : plane == 2 ? xhi : xlo;
        const size_t off = plane < 2 ? (size_t)w * np * np + tiled_index(r0 + lrow[u], lkq[u], np)
                                     : tiled_index(c0 + lrow[u], lkq[u], np);
        src[u] = base + off;
    }
    u32x4 pre[8];      // a native vector type: an array of HIP's uint4 struct is kept in scratch
#pragma unroll
    for (int u = 0; u < 8; ++u) pre[u] = *reinterpret_cast<const u32x4*>(src[u]);

    f16acc acc[2][4];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[x][y][r] = 0.f;

    const int l32 = lane & 31, kg = lane >> 5;
    const int kend = min(np, r0 + kSBM);
    for (int k0 = 0; k0 < kend; k0 += kSBK) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 8; ++u)
            *reinterpret_cast<u32x4*>(&lds[u >> 1][lrow[u]][lkq[u]]) = pre[u];
        __syncthreads();
        {
            // Unconditional prefetch (the last step re-reads its own stage,
            // which stays in bounds since kend >= kSBK).
            const int kn = min(k0 + kSBK, kend - kSBK);
#pragma unroll
            for (int u = 0; u < 8; ++u)
                pre[u] = *reinterpret_cast<const u32x4*>(src[u] + (size_t)(kn >> 5) * 8192);
        }
#pragma unroll
        for (int kc = 0; kc < kSBK; kc += 16) {
            uint4 ah[2], al[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                ah[x] = *reinterpret_cast<const uint4*>(&lds[0][wm + x * 32 + l32][kc + 8 * kg]);
                al[x] = *reinterpret_cast<const uint4*>(&lds[1][wm + x * 32 + l32][kc + 8 * kg]);
            }
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                const uint4 bh =
                    *reinterpret_cast<const uint4*>(&lds[2][wn + y * 32 + l32][kc + 8 * kg]);
                const uint4 bl =
                    *reinterpret_cast<const uint4*>(&lds[3][wn + y * 32 + l32][kc + 8 * kg]);
#pragma unroll
                for (int x = 0; x < 2; ++x) acc[x][y] = mfma32<F>(ah[x], bh, acc[x][y]);
#pragma unroll
                for (int x = 0; x < 2; ++x) acc[x][y] = mfma32<F>(ah[x], bl, acc[x][y]);
#pragma unroll
                for (int x = 0; x < 2; ++x) acc[x][y] = mfma32<F>(al[x], bh, acc[x][y]);
            }
        }
    }

    // epilogue: undo the row scale, sum squares per column over the tile's rows
    __syncthreads();
    double* red = reinterpret_cast<double*>(&lds[0][0][0]);      // [4][256] doubles
    const double* ms = mscale + (size_t)w * np + r0 + wm;
    double colsum[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const double inv = 1.0 / ms[x * 32 + row32(lane, r)];
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                const double z = (double)acc[x][y][r] * inv;
                colsum[y] += z * z;
            }
        }
#pragma unroll
    for (int y = 0; y < 4; ++y) {
        const double sum = colsum[y] + __shfl_xor(colsum[y], 32);
        if (kg == 0) red[(wave >> 1) * kSBN + wn + y * 32 + l32] = sum;
    }
    __syncthreads();
    if (t < kSBN) {
        const double v = red[t] + red[kSBN + t] + red[2 * kSBN + t] + red[3 * kSBN + t];
        qpart[((size_t)w * (np / kSBM) + rt) * mpad + c0 + t] = v;
    }
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

template <Split16 F>
__global__ __launch_bounds__(64) void k_probe32(const uint4* __restrict__ a,
                                                const uint4* __restrict__ b,
                                                float* __restrict__ d) {
    const int t = blockIdx.x, lane = threadIdx.x;
    f16acc acc;
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    acc = mfma32<F>(a[t * 64 + lane], b[t * 64 + lane], acc);
    for (int r = 0; r < 16; ++r) d[((size_t)t * 64 + lane) * 16 + r] = acc[r];
}

__global__ void k_reduce_q_scaled(const double* __restrict__ qpart,
                                  const double* __restrict__ xscale, const int* __restrict__ colmap,
                                  double* __restrict__ q, int nrt, int m, int mpad,
                                  const int* __restrict__ f0p) {
    const int w = blockIdx.y, f0 = *f0p;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < m; c += gridDim.x * blockDim.x) {
        const int cu = colmap ? colmap[c] : c;
        double s = 0.0;
        for (int rt = 0; rt < nrt; ++rt) s += qpart[((size_t)w * nrt + rt) * mpad + cu];
        const double xs = xscale[cu];
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
                                 double* qpart, int np, int mpad, int nb, int uf, hipStream_t s) {
    if (np % kSBM != 0 || mpad % kSBN != 0 || nb <= 0) return hipErrorInvalidValue;
    if (uf != 1 && uf != 2 && uf != 4 && uf != 8 && uf != 16 && uf != 32) return hipErrorInvalidValue;
    const int ncol = mpad / kSBN;
    int uc = 32 / uf;
    while (uf > 1 && uf / 2 >= nb) uf >>= 1;          // no wider than the batch
    while (uc > 1 && uc / 2 >= ncol) uc >>= 1;        // nor than the columns
    const long units = (long)(np / kSBM) * ((nb + uf - 1) / uf) * ((ncol + uc - 1) / uc);
    const long wgs = (units + 7) / 8 * 8 * uf * uc;
    if (wgs > 0x7fffffffL) return hipErrorInvalidValue;
    const dim3 grid((unsigned)wgs);
    if (fmt == Split16::BF16)
        hipLaunchKernelGGL(k_contract_split<Split16::BF16>, grid, dim3(512), 0, s, mhi, mlo, mscale,
                           xhi, xlo, qpart, np, mpad, nb, uf, uc);
    else
        hipLaunchKernelGGL(k_contract_split<Split16::FP16>, grid, dim3(512), 0, s, mhi, mlo, mscale,
                           xhi, xlo, qpart, np, mpad, nb, uf, uc);
    return hipGetLastError();
}

int contract_split_row_tiles(int np) { return np / kSBM; }

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

hipError_t launch_probe32(Split16 fmt, const uint16_t* a, const uint16_t* b, float* d,
                         int ntests, hipStream_t s) {
    const uint4* a4 = reinterpret_cast<const uint4*>(a);
    const uint4* b4 = reinterpret_cast<const uint4*>(b);
    if (fmt == Split16::BF16)
        hipLaunchKernelGGL(k_probe32<Split16::BF16>, dim3(ntests), dim3(64), 0, s, a4, b4, d);
    else
        hipLaunchKernelGGL(k_probe32<Split16::FP16>, dim3(ntests), dim3(64), 0, s, a4, b4, d);
    return hipGetLastError();
}

hipError_t launch_reduce_q_scaled(const double* qpart, const double* xscale, const int* colmap,
                                  double* q, int nrt, int m, int mpad, const int* f0p, int nb,
                                  hipStream_t s) {
    const int gx = (int)((m + 255) / 256 < 4096 ? (m + 255) / 256 : 4096);
    hipLaunchKernelGGL(k_reduce_q_scaled, dim3(gx > 0 ? gx : 1, nb), dim3(256), 0, s, qpart, xscale,
                       colmap, q, nrt, m, mpad, f0p);
    return hipGetLastError();
}

}  // namespace mgw
