// Mixed-precision contraction on int8 MFMA with exact integer accumulation.
//
// q_c = -|| M x_c ||^2 with M = L^-1 S (lower triangular) and x_c an ERI
// column.  Each row of M and each column of X is written as a power of two
// times S signed digits of B bits (mgw_slices.h: five 7-bit digits, or four
// balanced 8-bit ones).  z = M x is then the sum over digit pairs (i, j),
// i + j <= S - 1 (0-based), of 2^-(2 (B - 1) + B (i + j)) (T_i U_j): the
// products T_i U_j run on v_mfma_i32_32x32x32_i8 and are exact in int32,
// pairs of equal i + j share one accumulator, and the FP64 epilogue combines
// the S groups, undoes the row scale and sums the squares of each column.
// Unlike the 16-bit hi/lo split (22 bits per element, FP32 accumulation) the
// only error is the truncation of the digits: ~2^-35 of a row's largest
// element for 5 x 7 bits (15 products), ~2^-32 for 4 x 8 bits (10 products).
//
// Bounds: a group holds at most S pairs and K <= i8_max_k rows, so a group
// sum stays below 5 * 2^12 * 2^16 (7-bit) or 4 * 2^14 * 2^15 (8-bit) <= 2^31.
#include "mgw_kernels.h"
#include "mgw_slices.h"

#include <hip/hip_runtime.h>

namespace mgw {
namespace {


constexpr int kIBM = 128, kIBN = 128, kIBK = 32;   // workgroup tile and K step

// LDS byte offset of the 16 B half `kh` (0 or 16) of row `row` in a 128 x 32
// slice block: rows of 32 B, halves swapped in rows 8-15 of every 16, so that
// the 16 lanes served together by a 16 B access (8 rows x 2 halves when a
// block is stored, 16 rows x 1 half when a fragment is read) hit 16 distinct
// groups of 4 of the 64 banks (measured: SQ_LDS_BANK_CONFLICT).
__device__ inline int lds_off(int row, int kh) { return row * 32 + (kh ^ (((row >> 3) & 1) << 4)); }

// Plane layout: blocks of 128 rows x 32 k (4096 bytes, row-major inside),
// ordered by row block then k block; one K step of a 128-row tile is one
// contiguous 4 KB block per slice.
__device__ inline size_t slice_index(int row, int k, int np) {
    return ((size_t)(row >> 7) * (np >> 5) + (k >> 5)) * 4096 + (size_t)(row & 127) * 32 + (k & 31);
}


__device__ inline double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}


// One wave per column c of X: gather, scale, split into digits.
template <int kS, int B>
__global__ __launch_bounds__(256) void k_slice_x(const double* __restrict__ src,
                                                 const int64_t* __restrict__ pkrow,
                                                 const int* __restrict__ colidx,
                                                 int8_t* __restrict__ planes,
                                                 double* __restrict__ xscale, int n, int np,
                                                 int mpad) {
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (c >= mpad) return;
    const int64_t base = pkrow[c];
    double mx = 0.0;
    for (int k = lane; k < n; k += 64) mx = fmax(mx, fabs(src[base + colidx[k]]));
    const double scale = digit_scale<B>(wave_max(mx));
    const size_t plane = (size_t)mpad * np;
    for (int k = lane; k < np; k += 64) {
        const double x = k < n ? src[base + colidx[k]] * scale : 0.0;
        int8_t t[kS];
        digit_split<kS, B>(x, t);
        const size_t o = slice_index(c, k, np);
#pragma unroll
        for (int i = 0; i < kS; ++i) planes[i * plane + o] = t[i];
    }
    if (lane == 0) xscale[c] = scale;
}

// One wave per row r of M_w S_w (lower triangle, zeros above it up to the end of
// the row's 128-row tile; the planes beyond, which the contraction never reads,
// are left as they are).
template <int kS, int B>
__global__ __launch_bounds__(256) void k_slice_m(const double* __restrict__ M,
                                                 const double* __restrict__ S,
                                                 int8_t* __restrict__ planes,
                                                 double* __restrict__ mscale, int np) {
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int w = blockIdx.y;
    if (r >= np) return;
    const double* row = M + (size_t)w * np * np + (size_t)r * np;
    const double* Sw = S + (size_t)w * np;
    double mx = 0.0;
    for (int k = lane; k <= r; k += 64) mx = fmax(mx, fabs(row[k] * Sw[k]));
    const double scale = digit_scale<B>(wave_max(mx));
    const size_t plane = (size_t)np * np;
    int8_t* out = planes + (size_t)w * kS * plane;
    // the contraction reads a row tile's K range up to its own diagonal tile
    const int kend = min(np, (r / kIBM + 1) * kIBM);
    for (int k = lane; k < kend; k += 64) {
        const double x = k <= r ? (row[k] * Sw[k]) * scale : 0.0;
        int8_t t[kS];
        digit_split<kS, B>(x, t);
        const size_t o = slice_index(r, k, np);
#pragma unroll
        for (int i = 0; i < kS; ++i) out[i * plane + o] = t[i];
    }
    if (lane == 0) mscale[(size_t)w * np + r] = scale;
}


// Workgroup -> tile map: units of uf frequencies x uc column tiles of one row
// panel land on one XCD (workgroup L runs on XCD L % 8) and share its L2;
// row panels run large K first; tiles beyond the edges exit at once.
struct I8Tile {
    int rt, ct, w;
    bool valid;
};
__device__ inline I8Tile i8_tile(int L, int nrow, int ncol, int nb, int uf, int uc) {
    const int us = uf * uc;
    const int xcd = L & 7, s = L >> 3;
    const int u = (s / us) * 8 + xcd, j = s % us;
    const int ncg = (ncol + uc - 1) / uc, nfg = (nb + uf - 1) / uf;
    const int cg = u % ncg, fg = (u / ncg) % nfg, rr = u / (ncg * nfg);
    I8Tile tl;
    tl.rt = nrow - 1 - rr;
    tl.w = fg * uf + j / uc;
    tl.ct = cg * uc + j % uc;
    tl.valid = rr < nrow && tl.w < nb && tl.ct < ncol;
    return tl;
}

// 128 x 128 tile, 512 threads = 8 waves in 4 (rows) x 2 (columns), each wave
// 32 rows x 64 columns (two 32 x 32 blocks) with kS group accumulators per
// block.  Per K step of 32 the 2 x kS digit blocks (8 kS KB) are loaded into
// registers two steps ahead (kS x 16 B per thread) and stored to the LDS
// buffer the next step reads while the current one is computed from the
// other (two buffers, one barrier per step; rows swizzled by lds_off).  M is
// lower triangular, so the K loop stops at r0 + 128.
template <int kS, int B>
__global__ __launch_bounds__(512) void k_contract_i8(
    const int8_t* __restrict__ mplanes, const double* __restrict__ mscale,
    const int8_t* __restrict__ xplanes, double* __restrict__ qpart, int np, int mpad, int nb,
    int uf, int uc) {
    __shared__ __attribute__((aligned(16))) int8_t lds[2][2][kS][kIBM * kIBK];
    const I8Tile tile = i8_tile(blockIdx.x, np / kIBM, mpad / kIBN, nb, uf, uc);
    if (!tile.valid) return;                      // uniform over the workgroup
    const int w = tile.w, rt = tile.rt;
    const int r0 = rt * kIBM, c0 = tile.ct * kIBN;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 64;

    // loader: chunk u of thread t -> global chunk g = t + 512 u of 512 kS:
    // operand g / (256 kS) (0 = M, 1 = X), digit (g % (256 kS)) / 256,
    // row (g % 256) / 2, k half (g % 2) * 16
    constexpr int kOpChunks = 256 * kS;
    const size_t mplane = (size_t)np * np, xplane = (size_t)mpad * np;
    const int8_t* src[kS];
    int lop[kS], lsl[kS], lrow[kS], lkh[kS], lofs[kS];
#pragma unroll
    for (int u = 0; u < kS; ++u) {
        const int g = t + 512 * u;
        lop[u] = g / kOpChunks;
        lsl[u] = (g % kOpChunks) / 256;
        lrow[u] = (g % 256) >> 1;
        lkh[u] = (g & 1) * 16;
        lofs[u] = lds_off(lrow[u], lkh[u]);
        src[u] = lop[u] == 0
                     ? mplanes + ((size_t)w * kS + lsl[u]) * mplane + slice_index(r0 + lrow[u], lkh[u], np)
                     : xplanes + (size_t)lsl[u] * xplane + slice_index(c0 + lrow[u], lkh[u], np);
    }
    const int kend = min(np, r0 + kIBM);            // multiple of 32, >= 128
    const int nsteps = kend / kIBK;
    u32x4 pre[kS];
#pragma unroll
    for (int u = 0; u < kS; ++u) pre[u] = *reinterpret_cast<const u32x4*>(src[u]);
#pragma unroll
    for (int u = 0; u < kS; ++u)
        *reinterpret_cast<u32x4*>(&lds[0][lop[u]][lsl[u]][lofs[u]]) = pre[u];
#pragma unroll
    for (int u = 0; u < kS; ++u) pre[u] = *reinterpret_cast<const u32x4*>(src[u] + 4096);
    __syncthreads();

    i32x16 acc[kS][2];
#pragma unroll
    for (int g = 0; g < kS; ++g)
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[g][x][r] = 0;

    const int l32 = lane & 31, kg = (lane >> 5) * 16;
    const int oa = lds_off(wr + l32, kg), ob0 = lds_off(wc + l32, kg),
              ob1 = lds_off(wc + 32 + l32, kg);
    for (int st = 0; st < nsteps; ++st) {
        const int cur = st & 1;
        u32x4 a[kS];
#pragma unroll
        for (int i = 0; i < kS; ++i)
            a[i] = *reinterpret_cast<const u32x4*>(&lds[cur][0][i][oa]);
#pragma unroll
        for (int j = 0; j < kS; ++j) {
            const u32x4 b0 = *reinterpret_cast<const u32x4*>(&lds[cur][1][j][ob0]);
            const u32x4 b1 = *reinterpret_cast<const u32x4*>(&lds[cur][1][j][ob1]);
#pragma unroll
            for (int i = 0; i + j < kS; ++i) {
                acc[i + j][0] = mfma_i8(a[i], b0, acc[i + j][0]);
                acc[i + j][1] = mfma_i8(a[i], b1, acc[i + j][1]);
            }
        }
        // stage st + 1 into the other buffer (its readers finished at the
        // previous barrier), then fetch stage st + 2 (the last steps re-read
        // the final stage, in bounds)
#pragma unroll
        for (int u = 0; u < kS; ++u)
            *reinterpret_cast<u32x4*>(&lds[cur ^ 1][lop[u]][lsl[u]][lofs[u]]) = pre[u];
        const int kn = min(st + 2, nsteps - 1);
#pragma unroll
        for (int u = 0; u < kS; ++u)
            pre[u] = *reinterpret_cast<const u32x4*>(src[u] + (size_t)kn * 4096);
        __syncthreads();
    }

    // epilogue: combine the groups (Horner in 2^-B), undo the row scale, sum
    // the squares of each column over the tile's rows
    __syncthreads();
    double* red = reinterpret_cast<double*>(&lds[0][0][0][0]);      // [4][128] doubles
    const double* ms = mscale + (size_t)w * np + r0 + wr;
    double colsum[2] = {0.0, 0.0};
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const double inv = 1.0 / ms[row32(lane, r)];
#pragma unroll
        for (int x = 0; x < 2; ++x) {
            double z = (double)acc[kS - 1][x][r];
#pragma unroll
            for (int g = kS - 2; g >= 0; --g) z = z * group_step<B>() + (double)acc[g][x][r];
            z *= group0_weight<B>() * inv;
            colsum[x] += z * z;
        }
    }
#pragma unroll
    for (int x = 0; x < 2; ++x) {
        const double sum = colsum[x] + __shfl_xor(colsum[x], 32);
        if (lane < 32) red[(wave >> 1) * kIBN + wc + x * 32 + l32] = sum;
    }
    __syncthreads();
    if (t < kIBN) {
        const double v = red[t] + red[kIBN + t] + red[2 * kIBN + t] + red[3 * kIBN + t];
        qpart[((size_t)w * (np / kIBM) + rt) * mpad + c0 + t] = v;
    }
}

// Diagnostic: one v_mfma_i32_32x32x32_i8 per wave; a/b: [t][lane][16] bytes.
__global__ __launch_bounds__(64) void k_probe_i8(const u32x4* __restrict__ a,
                                                 const u32x4* __restrict__ b,
                                                 int* __restrict__ d) {
    const int t = blockIdx.x, lane = threadIdx.x;
    i32x16 acc;
    for (int r = 0; r < 16; ++r) acc[r] = 0;
    acc = mfma_i8(a[t * 64 + lane], b[t * 64 + lane], acc);
    for (int r = 0; r < 16; ++r) d[((size_t)t * 64 + lane) * 16 + r] = acc[r];
}

}  // namespace

hipError_t launch_slice_x(const double* src, const int64_t* pkrow, const int* colidx,
                          int8_t* planes, double* xscale, int n, int np, int mpad, int scheme,
                          hipStream_t s) {
    if (np % kIBM != 0 || mpad % kIBN != 0) return hipErrorInvalidValue;
    auto kern = scheme == kDigits4x8 ? k_slice_x<4, 8> : k_slice_x<5, 7>;
    hipLaunchKernelGGL(kern, dim3((mpad + 3) / 4), dim3(256), 0, s, src, pkrow, colidx, planes,
                       xscale, n, np, mpad);
    return hipGetLastError();
}

hipError_t launch_slice_m(const double* M, const double* S, int8_t* planes, double* mscale, int np,
                          int nb, int scheme, hipStream_t s) {
    if (np % kIBM != 0) return hipErrorInvalidValue;
    auto kern = scheme == kDigits4x8 ? k_slice_m<4, 8> : k_slice_m<5, 7>;
    hipLaunchKernelGGL(kern, dim3((np + 3) / 4, nb), dim3(256), 0, s, M, S, planes, mscale, np);
    return hipGetLastError();
}

hipError_t launch_contract_i8(const int8_t* mplanes, const double* mscale, const int8_t* xplanes,
                              double* qpart, int np, int mpad, int nb, int uf, int scheme,
                              hipStream_t s) {
    if (np % kIBM != 0 || mpad % kIBN != 0 || nb <= 0 || np > i8_max_k(scheme))
        return hipErrorInvalidValue;
    if (uf != 1 && uf != 2 && uf != 4 && uf != 8 && uf != 16 && uf != 32) return hipErrorInvalidValue;
    const int ncol = mpad / kIBN;
    int uc = 32 / uf;
    while (uf > 1 && uf / 2 >= nb) uf >>= 1;          // no wider than the batch
    while (uc > 1 && uc / 2 >= ncol) uc >>= 1;        // nor than the columns
    const long units = (long)(np / kIBM) * ((nb + uf - 1) / uf) * ((ncol + uc - 1) / uc);
    const long wgs = (units + 7) / 8 * 8 * uf * uc;
    if (wgs > 0x7fffffffL) return hipErrorInvalidValue;
    auto kern = scheme == kDigits4x8 ? k_contract_i8<4, 8> : k_contract_i8<5, 7>;
    hipLaunchKernelGGL(kern, dim3((unsigned)wgs), dim3(512), 0, s, mplanes, mscale, xplanes, qpart,
                       np, mpad, nb, uf, uc);
    return hipGetLastError();
}

int contract_i8_row_tiles(int np) { return np / kIBM; }

hipError_t launch_probe_i8(const int8_t* a, const int8_t* b, int* d, int ntests, hipStream_t s) {
    hipLaunchKernelGGL(k_probe_i8, dim3(ntests), dim3(64), 0, s, reinterpret_cast<const u32x4*>(a),
                       reinterpret_cast<const u32x4*>(b), d);
    return hipGetLastError();
}

}  // namespace mgw
