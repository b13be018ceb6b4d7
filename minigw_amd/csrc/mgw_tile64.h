// FP64 64 x 64 MFMA tile core shared by the factorization kernels
// (mgw_kernels.hip, mgw_ozfact.hip): v_mfma_f64_16x16x4_f64 blocks, K slabs
// staged through LDS, 256 threads = 4 waves in 2 x 2 of 32 x 32.
//
// MFMA operand/result lane layout (wave64, checked on the device):
//   a: A[i][k], i = lane % 16, k = lane / 16      (16 x 4 block)
//   b: B[k][j], j = lane % 16, k = lane / 16      (4 x 16 block)
//   d[r]: D[i][j], i = lane / 16 + 4 * r, j = lane % 16
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace mgw {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ inline d4 mfma_f64(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int kBK = 16;          // K slab staged through LDS per step
constexpr int kLdsPad = 4;       // doubles of padding per LDS row

// ------------------------------------------------------- 64 x 64 MFMA tile core
// LDS slabs: As[k][m], Bs[k][n] for k < kBK.  256 threads = 4 waves in 2 x 2,
// each wave a 32 x 32 sub-tile (2 x 2 MFMA blocks).
struct Tile64 {
    d4 acc[2][2];
    __device__ void zero() {
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    }
    __device__ void step(const double (*As)[64 + kLdsPad], const double (*Bs)[64 + kLdsPad]) {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
        const int li = lane & 15, lk = lane >> 4;
#pragma unroll
        for (int kk = 0; kk < kBK; kk += 4) {
            const double a0 = As[kk + lk][wm + li], a1 = As[kk + lk][wm + 16 + li];
            const double b0 = Bs[kk + lk][wn + li], b1 = Bs[kk + lk][wn + 16 + li];
            acc[0][0] = mfma_f64(a0, b0, acc[0][0]);
            acc[0][1] = mfma_f64(a0, b1, acc[0][1]);
            acc[1][0] = mfma_f64(a1, b0, acc[1][0]);
            acc[1][1] = mfma_f64(a1, b1, acc[1][1]);
        }
    }
    // visit(row, col, value) over this thread's 16 results
    template <class F>
    __device__ void for_each(F f) const {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
        for (int bi = 0; bi < 2; ++bi)
            for (int bj = 0; bj < 2; ++bj)
                for (int r = 0; r < 4; ++r)
                    f(wm + bi * 16 + (lane >> 4) + 4 * r, wn + bj * 16 + (lane & 15), acc[bi][bj][r]);
    }
};

// As[k][m] = src[(row0 + m) * ld + k0 + k], m < 64, k < kBK (row-major operand, K along a row)
__device__ inline void load_rows_kmajor(double (*As)[64 + kLdsPad], const double* __restrict__ src,
                                        size_t ld, int row0, int k0) {
    const int t = threadIdx.x, m = t >> 2, kq = (t & 3) * 4;
    const double* p = src + (size_t)(row0 + m) * ld + k0 + kq;
#pragma unroll
    for (int u = 0; u < 4; ++u) As[kq + u][m] = p[u];
}

// Bs[k][n] = src[(k0 + k) * ld + col0 + n], k < kBK, n < 64 (row-major operand, K down the rows)
__device__ inline void load_kn(double (*Bs)[64 + kLdsPad], const double* __restrict__ src,
                               size_t ld, int k0, int col0) {
    const int t = threadIdx.x, k = t >> 4, nq = (t & 15) * 4;
    const double* p = src + (size_t)(k0 + k) * ld + col0 + nq;
#pragma unroll
    for (int u = 0; u < 4; ++u) Bs[k][nq + u] = p[u];
}

// acc += A[0:64, kbeg:kend] * op(B) with A row-major (lda), op(B) = B^T for a
// row-major 64-row B (NT) or B[kbeg:kend, 0:64] row-major (NN).  The next K
// slab is fetched into registers while the current one is multiplied.
template <bool NN>
__device__ inline void gemm64(const double* __restrict__ a, size_t lda,
                              const double* __restrict__ b, size_t ldb, int kbeg, int kend,
                              Tile64& tile, double (*As)[64 + kLdsPad],
                              double (*Bs)[64 + kLdsPad]) {
    const int t = threadIdx.x;
    const int am = t >> 2, akq = (t & 3) * 4;
    const int bk = t >> 4, bnq = (t & 15) * 4;
    double2 ra[2], rb[2];
    auto fetch = [&](int k0) {
        const double2* pa = reinterpret_cast<const double2*>(a + (size_t)am * lda + k0 + akq);
        ra[0] = pa[0];
        ra[1] = pa[1];
        const double2* pb =
            NN ? reinterpret_cast<const double2*>(b + (size_t)(k0 + bk) * ldb + bnq)
               : reinterpret_cast<const double2*>(b + (size_t)am * ldb + k0 + akq);
        rb[0] = pb[0];
        rb[1] = pb[1];
    };
    if (kbeg >= kend) return;
    fetch(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += kBK) {
        __syncthreads();
        As[akq + 0][am] = ra[0].x;
        As[akq + 1][am] = ra[0].y;
        As[akq + 2][am] = ra[1].x;
        As[akq + 3][am] = ra[1].y;
        if (NN) {
            Bs[bk][bnq + 0] = rb[0].x;
            Bs[bk][bnq + 1] = rb[0].y;
            Bs[bk][bnq + 2] = rb[1].x;
            Bs[bk][bnq + 3] = rb[1].y;
        } else {
            Bs[akq + 0][am] = rb[0].x;
            Bs[akq + 1][am] = rb[0].y;
            Bs[akq + 2][am] = rb[1].x;
            Bs[akq + 3][am] = rb[1].y;
        }
        __syncthreads();
        if (k0 + kBK < kend) fetch(k0 + kBK);
        tile.step(As, Bs);
    }
    __syncthreads();
}

// gemm64 over exactly one 64-wide K block (a and b already offset to it): the
// whole of both operands is fetched into registers at once, so the block
// waits for memory once instead of once per K slab, then staged through LDS
// slab by slab.  For the single-block products of the panel kernels, whose
// time is the latency of these loads.
template <bool NN>
__device__ inline void gemm64_k64(const double* __restrict__ a, size_t lda,
                                  const double* __restrict__ b, size_t ldb, Tile64& tile,
                                  double (*As)[64 + kLdsPad], double (*Bs)[64 + kLdsPad]) {
    constexpr int kSlabs = 64 / kBK;
    const int t = threadIdx.x;
    const int am = t >> 2, akq = (t & 3) * 4;
    const int bk = t >> 4, bnq = (t & 15) * 4;
    double2 ra[kSlabs][2], rb[kSlabs][2];
#pragma unroll
    for (int s = 0; s < kSlabs; ++s) {
        const int k0 = s * kBK;
        const double2* pa = reinterpret_cast<const double2*>(a + (size_t)am * lda + k0 + akq);
        ra[s][0] = pa[0];
        ra[s][1] = pa[1];
        const double2* pb =
            NN ? reinterpret_cast<const double2*>(b + (size_t)(k0 + bk) * ldb + bnq)
               : reinterpret_cast<const double2*>(b + (size_t)am * ldb + k0 + akq);
        rb[s][0] = pb[0];
        rb[s][1] = pb[1];
    }
#pragma unroll
    for (int s = 0; s < kSlabs; ++s) {
        __syncthreads();
        As[akq + 0][am] = ra[s][0].x;
        As[akq + 1][am] = ra[s][0].y;
        As[akq + 2][am] = ra[s][1].x;
        As[akq + 3][am] = ra[s][1].y;
        if (NN) {
            Bs[bk][bnq + 0] = rb[s][0].x;
            Bs[bk][bnq + 1] = rb[s][0].y;
            Bs[bk][bnq + 2] = rb[s][1].x;
            Bs[bk][bnq + 3] = rb[s][1].y;
        } else {
            Bs[akq + 0][am] = rb[s][0].x;
            Bs[akq + 1][am] = rb[s][0].y;
            Bs[akq + 2][am] = rb[s][1].x;
            Bs[akq + 3][am] = rb[s][1].y;
        }
        __syncthreads();
        tile.step(As, Bs);
    }
    __syncthreads();
}

}  // namespace
}  // namespace mgw
