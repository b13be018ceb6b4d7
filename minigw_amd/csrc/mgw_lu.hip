// Pivoted-LU screening for a symmetrised dielectric matrix that is not
// positive definite (an ERI whose V_ph is not positive semidefinite: noisy or
// truncated integrals).  The Cholesky path stops at a non-positive pivot;
// this path proceeds as the reference's LU with partial pivoting does
// (zgetrf + zgetrs against the identity, src/linalg/linalg_blas_lapack.cpp:106-140,
// src/workspace/device_screening_workspace.cu:319-335), in FP64:
//
//   E = I + S V S = P L U                  (blocked right-looking, 64-wide panels)
//   W = E^-1 = U^-1 L^-1 P^T               (block forward / backward sweeps on P^T I)
//   q_c = -(S x_c)^T W (S x_c)             (dense FP64 MFMA contraction, dot epilogue)
//
// One matrix (frequency) per blockIdx.y; the frequencies of a batch whose
// Cholesky broke down run here after the batch.
#include "mgw_kernels.h"
#include "mgw_tile64.h"

#include <hip/hip_runtime.h>

namespace mgw {
namespace {

// Panel k of the LU: rows >= k0 = 64 k, columns [k0, k0 + 64), one workgroup
// per matrix.  Column by column: the pivot is the first row of largest
// magnitude at or below the diagonal (LAPACK idamax), its panel row is swapped
// with the diagonal row (the rest of the two rows is swapped by k_lu_swap),
// the column below is divided by the pivot and the panel's trailing columns
// take the rank-1 update.  piv[c] = the swapped row; info = first zero pivot
// (1-based).  Then the inverses of the unit lower L_kk and the upper U_kk
// (64 x 64, by substitution) go to dl / du.
__global__ __launch_bounds__(256) void k_lu_panel(double* __restrict__ A, int* __restrict__ piv,
                                                  int* __restrict__ info, double* __restrict__ dl,
                                                  double* __restrict__ du, int np, int k) {
    __shared__ double rowc[64];
    __shared__ double red_v[256];
    __shared__ int red_i[256];
    __shared__ double blk[64][65];
    const int w = blockIdx.y, t = threadIdx.x, nblk = np / kNB;
    const int k0 = k * kNB;
    double* Aw = A + (size_t)w * np * np;
    int* pw = piv + (size_t)w * np;
    for (int j = 0; j < 64; ++j) {
        const int c = k0 + j;
        double best = -1.0;
        int bi = np;
        for (int r = c + t; r < np; r += 256) {
            const double v = fabs(Aw[(size_t)r * np + c]);
            if (v > best) {           // rows visited in increasing order: the first maximum
                best = v;
                bi = r;
            }
        }
        red_v[t] = best;
        red_i[t] = bi;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (t < s) {
                const double v2 = red_v[t + s];
                const int i2 = red_i[t + s];
                if (v2 > red_v[t] || (v2 == red_v[t] && i2 < red_i[t])) {
                    red_v[t] = v2;
                    red_i[t] = i2;
                }
            }
            __syncthreads();
        }
        const int p = red_i[0];
        if (t == 0) pw[c] = p;
        if (t < 64 && p != c) {
            double* rc = Aw + (size_t)c * np + k0 + t;
            double* rp = Aw + (size_t)p * np + k0 + t;
            const double a = *rc, b = *rp;
            *rc = b;
            *rp = a;
        }
        __syncthreads();
        if (t < 64) rowc[t] = Aw[(size_t)c * np + k0 + t];
        __syncthreads();
        const double pv = rowc[j];
        if (pv == 0.0) {
            if (t == 0 && info[w] == 0) info[w] = c + 1;
        } else {
            const double inv = 1.0 / pv;
            for (int r = c + 1 + t; r < np; r += 256) {
                double* row = Aw + (size_t)r * np + k0;
                const double l = row[j] * inv;
                row[j] = l;
                for (int jj = j + 1; jj < 64; ++jj) row[jj] = fma(-l, rowc[jj], row[jj]);
            }
        }
        __syncthreads();
    }
    // inverses of the diagonal block's unit lower and upper factors
    for (int e = t; e < 4096; e += 256) blk[e >> 6][e & 63] = Aw[(size_t)(k0 + (e >> 6)) * np + k0 + (e & 63)];
    __syncthreads();
    double* DL = dl + ((size_t)w * nblk + k) * 4096;
    double* DU = du + ((size_t)w * nblk + k) * 4096;
    if (t < 64) {                       // column t of L_kk^-1: L x = e_t
        double x[64];
        for (int r = 0; r < 64; ++r) {
            double v = r == t ? 1.0 : 0.0;
            for (int s = 0; s < r; ++s) v = fma(-blk[r][s], x[s], v);
            x[r] = v;
        }
        for (int r = 0; r < 64; ++r) DL[r * 64 + t] = x[r];
    } else if (t < 128) {               // column c of U_kk^-1: U x = e_c
        const int cc = t - 64;
        double x[64];
        for (int r = 63; r >= 0; --r) {
            double v = r == cc ? 1.0 : 0.0;
            for (int s = r + 1; s < 64; ++s) v = fma(-blk[r][s], x[s], v);
            x[r] = blk[r][r] != 0.0 ? v / blk[r][r] : 0.0;
        }
        for (int r = 0; r < 64; ++r) DU[r * 64 + cc] = x[r];
    }
}

// The panel's 64 row interchanges on every column outside it.
__global__ void k_lu_swap(double* __restrict__ A, const int* __restrict__ piv, int np, int k) {
    const int w = blockIdx.y, k0 = k * kNB;
    const int col0 = blockIdx.x * blockDim.x + threadIdx.x;
    const int col = col0 < k0 ? col0 : col0 + kNB;       // skip the panel's columns
    if (col >= np) return;
    double* Aw = A + (size_t)w * np * np;
    const int* pw = piv + (size_t)w * np;
    for (int j = 0; j < 64; ++j) {
        const int c = k0 + j, p = pw[c];
        if (p != c) {
            const double a = Aw[(size_t)c * np + col];
            Aw[(size_t)c * np + col] = Aw[(size_t)p * np + col];
            Aw[(size_t)p * np + col] = a;
        }
    }
}

// Block (k, j), j = k + 1 + blockIdx.x, of U: U_kj = L_kk^-1 A_kj.
__global__ __launch_bounds__(256) void k_lu_u12(double* __restrict__ A, const double* __restrict__ dl,
                                                int np, int k) {
    __shared__ double As[kBK][64 + kLdsPad];
    __shared__ double Bs[kBK][64 + kLdsPad];
    const int w = blockIdx.y, nblk = np / kNB, j = k + 1 + blockIdx.x;
    double* Aw = A + (size_t)w * np * np;
    double* blkp = Aw + (size_t)k * kNB * np + (size_t)j * kNB;
    Tile64 tile;
    tile.zero();
    gemm64<true>(dl + ((size_t)w * nblk + k) * 4096, 64, blkp, np, 0, kNB, tile, As, Bs);
    tile.for_each([&](int r, int c, double v) { blkp[(size_t)r * np + c] = v; });
}

// Trailing update A_ij -= L_ik U_kj for i, j > k.
__global__ __launch_bounds__(256) void k_lu_update(double* __restrict__ A, int np, int k) {
    __shared__ double As[kBK][64 + kLdsPad];
    __shared__ double Bs[kBK][64 + kLdsPad];
    const int w = blockIdx.y, t = np / kNB - k - 1;
    const int i = k + 1 + blockIdx.x / t, j = k + 1 + blockIdx.x % t;
    double* Aw = A + (size_t)w * np * np;
    Tile64 tile;
    tile.zero();
    gemm64<true>(Aw + (size_t)i * kNB * np + (size_t)k * kNB, np,
                 Aw + (size_t)k * kNB * np + (size_t)j * kNB, np, 0, kNB, tile, As, Bs);
    double* out = Aw + (size_t)i * kNB * np + (size_t)j * kNB;
    tile.for_each([&](int r, int c, double v) { out[(size_t)r * np + c] -= v; });
}

// Z = P^T I: row r is the unit row perm[r].
__global__ void k_perm_identity(double* __restrict__ Z, const int* __restrict__ perm, int np) {
    const int w = blockIdx.y;
    const size_t total = (size_t)np * np;
    double* Zw = Z + (size_t)w * total;
    const int* pw = perm + (size_t)w * np;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(e / np), c = (int)(e % np);
        Zw[e] = c == pw[r] ? 1.0 : 0.0;
    }
}

// One block row i of the sweeps on Z (column tile blockIdx.x), in place:
// LOWER: Z_i = L_ii^-1 (Z_i - sum_{j < i} L_ij Z_j)   (unit L below the diagonal of A)
// upper: Z_i = U_ii^-1 (Z_i - sum_{j > i} U_ij Z_j)   (U on and above it)
template <bool LOWER>
__global__ __launch_bounds__(256) void k_lu_sweep(const double* __restrict__ A,
                                                  const double* __restrict__ D,
                                                  double* __restrict__ Z, int np, int i) {
    __shared__ double As[kBK][64 + kLdsPad];
    __shared__ double Bs[kBK][64 + kLdsPad];
    __shared__ __attribute__((aligned(16))) double y[64][66];
    const int w = blockIdx.y, nblk = np / kNB, ct = blockIdx.x;
    const double* Aw = A + (size_t)w * np * np;
    double* Zw = Z + (size_t)w * np * np;
    double* Zi = Zw + (size_t)i * kNB * np + (size_t)ct * kNB;
    Tile64 tile;
    tile.zero();
    const int kbeg = LOWER ? 0 : (i + 1) * kNB, kend = LOWER ? i * kNB : np;
    gemm64<true>(Aw + (size_t)i * kNB * np, np, Zw + (size_t)ct * kNB, np, kbeg, kend, tile, As, Bs);
    tile.for_each([&](int r, int c, double v) { y[r][c] = Zi[(size_t)r * np + c] - v; });
    __syncthreads();
    tile.zero();
    gemm64<true>(D + ((size_t)w * nblk + i) * 4096, 64, &y[0][0], 66, 0, kNB, tile, As, Bs);
    tile.for_each([&](int r, int c, double v) { Zi[(size_t)r * np + c] = v; });
}

// W_c = -S W S for the boundary's download.
__global__ void k_wc_from_dense(const double* __restrict__ W, const double* __restrict__ S,
                                double* __restrict__ wc, int n, int np) {
    const size_t total = (size_t)n * n;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(e / n), c = (int)(e % n);
        wc[e] = -(S[r] * W[(size_t)r * np + c]) * S[c];
    }
}

// q partials of a dense W: per 128 x 128 tile of W (S x) (rows r0.., columns
// c0..), the column sums of (S x)_r (W S x)_r over the tile's rows ->
// qpart[w][row tile][column]; 4 waves in 2 x 2, each 64 x 64.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_contract_dense(
    const double* __restrict__ W, const double* __restrict__ S, const double* __restrict__ xp,
    double* __restrict__ qpart, int np, int mpad) {
    __shared__ double As[kBK][kCT + kLdsPad];
    __shared__ double Bs[kBK][kCT + kLdsPad];
    __shared__ double red[2][kCT];
    const int w = blockIdx.z, rt = blockIdx.y, r0 = rt * kCT, c0 = blockIdx.x * kCT;
    const double* Ww = W + (size_t)w * np * np;
    const double* Sw = S + (size_t)w * np;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int li = lane & 15, lk = lane >> 4;
    const int lrow = t >> 1, lkq = (t & 1) * 8;
    const double* arow = Ww + (size_t)(r0 + lrow) * np;
    const double* bsrc = xp + (size_t)(c0 + lrow) * np;
    d4 acc[4][4];
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < np; k0 += kBK) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int kk = k0 + lkq + u;
            As[lkq + u][lrow] = arow[kk];
            Bs[lkq + u][lrow] = bsrc[kk] * Sw[kk];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kBK; kk += 4) {
            double a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a[q] = As[kk + lk][wm + q * 16 + li];
                b[q] = Bs[kk + lk][wn + q * 16 + li];
            }
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) acc[x][y] = mfma_f64(a[x], b[y], acc[x][y]);
        }
    }
    __syncthreads();
    // d[r] of block (x, y): row r0 + wm + 16 x + lk + 4 r, column c0 + wn + 16 y + li
    for (int y = 0; y < 4; ++y) {
        const int col = c0 + wn + y * 16 + li;
        double s = 0.0;
        for (int x = 0; x < 4; ++x)
            for (int r = 0; r < 4; ++r) {
                const int row = r0 + wm + x * 16 + lk + 4 * r;
                s += acc[x][y][r] * (xp[(size_t)col * np + row] * Sw[row]);
            }
        s += __shfl_xor(s, 16);
        s += __shfl_xor(s, 32);
        if (lk == 0) red[wave >> 1][wn + y * 16 + li] = s;
    }
    __syncthreads();
    if (t < kCT) qpart[((size_t)w * (np / kCT) + rt) * mpad + c0 + t] = red[0][t] + red[1][t];
}

// all of A_w = I + S V S (V as given, not assumed symmetric: the LU takes
// the general matrix, like the reference's eps = I - diag(Pi0) V), one matrix
// per blockIdx.y
__global__ void k_dielectric_full(const double* __restrict__ v, const double* __restrict__ S,
                                  double* __restrict__ A, int np) {
    const int w = blockIdx.y;
    const double* Sw = S + (size_t)w * np;
    double* Aw = A + (size_t)w * np * np;
    const size_t total = (size_t)np * np;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(e / np), c = (int)(e % np);
        Aw[e] = (r == c ? 1.0 : 0.0) + (Sw[r] * v[e]) * Sw[c];
    }
}

unsigned grid_of(size_t work) {
    const size_t g = (work + 255) / 256;
    return (unsigned)(g < 4096 ? (g > 0 ? g : 1) : 4096);
}

}  // namespace

hipError_t launch_dielectric_full(const double* v, const double* S, double* A, int np, int nb,
                                  hipStream_t s) {
    hipLaunchKernelGGL(k_dielectric_full, dim3(grid_of((size_t)np * np), nb), dim3(256), 0, s, v, S,
                       A, np);
    return hipGetLastError();
}

hipError_t launch_lu_factor(double* A, int* piv, int* info, double* dl, double* du, int np, int nb,
                            hipStream_t s) {
    if (np % kNB != 0) return hipErrorInvalidValue;
    const int nblk = np / kNB;
    for (int k = 0; k < nblk; ++k) {
        hipLaunchKernelGGL(k_lu_panel, dim3(1, nb), dim3(256), 0, s, A, piv, info, dl, du, np, k);
        hipLaunchKernelGGL(k_lu_swap, dim3((np - kNB + 255) / 256, nb), dim3(256), 0, s, A, piv, np,
                           k);
        const int t = nblk - k - 1;
        if (t > 0) {
            hipLaunchKernelGGL(k_lu_u12, dim3(t, nb), dim3(256), 0, s, A, dl, np, k);
            hipLaunchKernelGGL(k_lu_update, dim3(t * t, nb), dim3(256), 0, s, A, np, k);
        }
    }
    return hipGetLastError();
}

hipError_t launch_lu_inverse(const double* A, const double* dl, const double* du, const int* perm,
                             double* Z, int np, int nb, hipStream_t s) {
    if (np % kNB != 0) return hipErrorInvalidValue;
    const int nblk = np / kNB;
    hipLaunchKernelGGL(k_perm_identity, dim3(grid_of((size_t)np * np), nb), dim3(256), 0, s, Z, perm,
                       np);
    for (int i = 0; i < nblk; ++i)
        hipLaunchKernelGGL(k_lu_sweep<true>, dim3(nblk, nb), dim3(256), 0, s, A, dl, Z, np, i);
    for (int i = nblk - 1; i >= 0; --i)
        hipLaunchKernelGGL(k_lu_sweep<false>, dim3(nblk, nb), dim3(256), 0, s, A, du, Z, np, i);
    return hipGetLastError();
}

hipError_t launch_contract_dense(const double* W, const double* S, const double* xp, double* qpart,
                                 int np, int mpad, int nb, hipStream_t s) {
    hipLaunchKernelGGL(k_contract_dense, dim3(mpad / kCT, np / kCT, nb), dim3(256), 0, s, W, S, xp,
                       qpart, np, mpad);
    return hipGetLastError();
}

hipError_t launch_wc_from_dense(const double* W, const double* S, double* wc, int n, int np,
                                hipStream_t s) {
    hipLaunchKernelGGL(k_wc_from_dense, dim3(grid_of((size_t)n * n)), dim3(256), 0, s, W, S, wc, n,
                       np);
    return hipGetLastError();
}

}  // namespace mgw
