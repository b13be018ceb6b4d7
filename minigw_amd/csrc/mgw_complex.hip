// Complex screening for eta != 0 (FP64, vector ALU).
//
// Pi0 = 2 [1/(i w - D + i eta) - 1/(i w + D - i eta)] (src/gw/sigma.cpp:41-54)
// has a negative real part, so S = sqrt(-Pi0) (principal branch) exists and
// E = S^-1 eps S = I + S V S is complex symmetric (not Hermitian).  With the
// factorization E = L L^T (transpose, no conjugate), W_c = eps^-1 diag(Pi0)
// = -(L^-1 S)^T (L^-1 S), so q_c = x_c^T W_c x_c = -sum_r z_rc^2 with
// Z = L^-1 S X: the reference's q (no conjugate,
// src/workspace/device_screening_workspace.cu:245-273) for real x_c.
//
// E is factored by a left-looking blocked LL^T (64 x 64 blocks; the diagonal
// block in LDS, its inverse kept), and Z by blocked forward substitution on
// the columns (Z_i = L_ii^-1 (S X_i - sum_k<i L_ik Z_k)); every block step is
// one complex GEMM-tile kernel.  Without pivoting: E is a small perturbation of
// the eta = 0 matrix, which is symmetric positive definite with eigenvalues >= 1;
// a zero or non-finite pivot is reported through info.
#include "mgw_kernels.h"

#include <hip/hip_runtime.h>

namespace mgw {
namespace {

struct __attribute__((aligned(16))) zd {
    double re, im;
};
__device__ inline zd zmul(zd a, zd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline zd zsub(zd a, zd b) { return {a.re - b.re, a.im - b.im}; }
// Smith's division
__device__ inline zd zdiv(zd a, zd b) {
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re, d = b.re + b.im * r;
        return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    }
    const double r = b.re / b.im, d = b.re * r + b.im;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
// principal square root
__device__ inline zd zsqrt(zd z) {
    if (z.re == 0.0 && z.im == 0.0) return {0.0, 0.0};
    const double r = hypot(z.re, z.im);
    if (z.re >= 0.0) {
        const double t = sqrt(0.5 * (r + z.re));
        return {t, z.im / (2.0 * t)};
    }
    const double t = sqrt(0.5 * (r - z.re));
    return {fabs(z.im) / (2.0 * t), copysign(t, z.im)};
}
__device__ inline zd pi0_of(double w, double d, double eta) {
    const zd one{1.0, 0.0};
    const zd a = zdiv(one, zd{-d, w + eta});       // 1 / (i w - D + i eta)
    const zd b = zdiv(one, zd{d, w - eta});        // 1 / (i w + D - i eta)
    return {2.0 * (a.re - b.re), 2.0 * (a.im - b.im)};
}

// S[w][r] = sqrt(-Pi0(omega[*f0p + w], delta[r], eta)), 0 in the padding.
__global__ void k_scale_grid_c(const double* __restrict__ delta, const double* __restrict__ omega,
                               double eta, zd* __restrict__ S, int n, int np,
                               const int* __restrict__ f0p) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x, w = blockIdx.y;
    if (r >= np) return;
    zd s{0.0, 0.0};
    if (r < n) {
        const zd p = pi0_of(omega[*f0p + w], delta[r], eta);
        s = zsqrt(zd{-p.re, -p.im});
    }
    S[(size_t)w * np + r] = s;
}

__global__ void k_scale_pi0_c(const double* __restrict__ pre, const double* __restrict__ pim,
                              zd* __restrict__ S, int n, int np) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= np) return;
    S[r] = r < n ? zsqrt(zd{-pre[r], -(pim ? pim[r] : 0.0)}) : zd{0.0, 0.0};
}

// Lower triangle of E_w = I + S_w V S_w.
__global__ void k_dielectric_c(const double* __restrict__ V, const zd* __restrict__ S,
                               zd* __restrict__ E, int np) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y, w = blockIdx.z;
    if (j > i || j >= np) return;
    const zd si = S[(size_t)w * np + i], sj = S[(size_t)w * np + j];
    zd e = zmul(zmul(si, zd{V[(size_t)i * np + j], 0.0}), sj);
    if (i == j) e.re += 1.0;
    E[((size_t)w * np + i) * np + j] = e;
}

// One 64 x 64 tile of C = beta C + alpha sum_kk A(i, kk) B(kk, j) per
// workgroup (256 threads, 4 x 4 outputs each, K chunks of 16 through LDS).
// A(i, kk) = TA ? A[kk lda + i] : A[i lda + kk]; B(kk, j) = TB ? B[j ldb + kk]
// : B[kk ldb + j].  Tile (bx, by) of batch member z starts at row 64 by,
// column 64 bx of its C; A, B, C advance by their strides per z.  C may
// alias the A or B tile it reads (all reads finish before the writes).
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_cgemm(const zd* __restrict__ A, int lda, long sa,
                                               const zd* __restrict__ B, int ldb, long sb, zd* C,
                                               int ldc, long sc, int K, double alpha, int beta) {
    __shared__ zd As[16][64];
    __shared__ zd Bs[16][64];
    const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
    const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
    A += blockIdx.z * sa;
    B += blockIdx.z * sb;
    C += blockIdx.z * sc;
    zd acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = zd{0.0, 0.0};
    for (int k0 = 0; k0 < K; k0 += 16) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = t + 256 * u;           // 1024 elements per operand chunk
            const int kk = e >> 6, x = e & 63;   // (kk, row / column)
            As[kk][x] = TA ? A[(size_t)(k0 + kk) * lda + i0 + x] : A[(size_t)(i0 + x) * lda + k0 + kk];
            Bs[kk][x] = TB ? B[(size_t)(j0 + x) * ldb + k0 + kk] : B[(size_t)(k0 + kk) * ldb + j0 + x];
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < 16; ++kk) {
            zd a[4], b[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = As[kk][ty * 4 + r];
#pragma unroll
            for (int c = 0; c < 4; ++c) b[c] = Bs[kk][tx * 4 + c];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    acc[r][c].re = fma(a[r].re, b[c].re, fma(-a[r].im, b[c].im, acc[r][c].re));
                    acc[r][c].im = fma(a[r].re, b[c].im, fma(a[r].im, b[c].re, acc[r][c].im));
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            zd* o = C + (size_t)(i0 + ty * 4 + r) * ldc + j0 + tx * 4 + c;
            zd v{alpha * acc[r][c].re, alpha * acc[r][c].im};
            if (beta) v = zd{o->re + v.re, o->im + v.im};
            *o = v;
        }
}

// Diagonal block k of E_w: L_kk L_kk^T in place (lower) and L_kk^-1 into
// dinv[w][k]; info[w] = 1-based row of the first zero or non-finite pivot.
__global__ __launch_bounds__(256) void k_potf2_c(zd* __restrict__ E, zd* __restrict__ dinv,
                                                 int* __restrict__ info, int np, int k) {
    __shared__ zd a[64][65];
    const int w = blockIdx.x, t = threadIdx.x;
    const int nblk = np / 64;
    zd* blk = E + ((size_t)w * np + (size_t)k * 64) * np + (size_t)k * 64;
    for (int e = t; e < 64 * 64; e += 256) {
        const int r = e >> 6, c = e & 63;
        a[r][c] = c <= r ? blk[(size_t)r * np + c] : zd{0.0, 0.0};
    }
    __syncthreads();
    for (int j = 0; j < 64; ++j) {
        if (t == 0) {
            zd l = zsqrt(a[j][j]);
            if (!(l.re != 0.0 || l.im != 0.0) || !isfinite(l.re) || !isfinite(l.im)) {
                if (info[w] == 0) info[w] = k * 64 + j + 1;
                l = zd{1.0, 0.0};
            }
            a[j][j] = l;
        }
        __syncthreads();
        const zd l = a[j][j];
        for (int i = j + 1 + t; i < 64; i += 256) a[i][j] = zdiv(a[i][j], l);
        __syncthreads();
        for (int e = t; e < 64 * 64; e += 256) {
            const int i = e >> 6, c = e & 63;
            if (i > j && c > j && c <= i) a[i][c] = zsub(a[i][c], zmul(a[i][j], a[c][j]));
        }
        __syncthreads();
    }
    for (int e = t; e < 64 * 64; e += 256) {
        const int r = e >> 6, c = e & 63;
        if (c <= r) blk[(size_t)r * np + c] = a[r][c];
    }
    // inverse of the lower factor, one column per thread: L x = e_c
    zd* inv = dinv + ((size_t)w * nblk + k) * 64 * 64;
    if (t < 64) {
        const int c = t;
        zd x[64];
        for (int r = 0; r < 64; ++r) {
            if (r < c) {
                x[r] = zd{0.0, 0.0};
                continue;
            }
            zd s{r == c ? 1.0 : 0.0, 0.0};
            for (int q = c; q < r; ++q) s = zsub(s, zmul(a[r][q], x[q]));
            x[r] = zdiv(s, a[r][r]);
        }
        for (int r = 0; r < 64; ++r) inv[(size_t)r * 64 + c] = x[r];
    }
}

// Y_w[r][c] = S_w[r] X[r][c] with X packed per column (xp[c][r], real).
__global__ void k_make_y(const zd* __restrict__ S, const double* __restrict__ xp,
                         zd* __restrict__ Y, int np, int mpad) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y, w = blockIdx.z;
    if (c >= mpad) return;
    const zd s = S[(size_t)w * np + r];
    const double x = xp[(size_t)c * np + r];
    Y[((size_t)w * np + r) * mpad + c] = zd{s.re * x, s.im * x};
}

// q[f0 + w][c] = -sum_r Z_w[r][colmap[c]]^2 (complex square, no conjugate).
__global__ void k_sumsq_c(const zd* __restrict__ Z, const int* __restrict__ colmap,
                          double* __restrict__ qre, double* __restrict__ qim, int np, int m,
                          int mpad, const int* __restrict__ f0p) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x, w = blockIdx.y;
    if (c >= m) return;
    const int cu = colmap ? colmap[c] : c;
    const zd* z = Z + (size_t)w * np * mpad + cu;
    double sr = 0.0, si = 0.0;
    for (int r = 0; r < np; ++r) {
        const zd v = z[(size_t)r * mpad];
        sr += v.re * v.re - v.im * v.im;
        si += 2.0 * v.re * v.im;
    }
    const size_t o = (size_t)(*f0p + w) * m + c;
    qre[o] = -sr;
    qim[o] = -si;
}

// W = -Z^T Z for the first n rows / columns (download of W_c).
__global__ void k_copy_c(const zd* __restrict__ W, double* __restrict__ re,
                         double* __restrict__ im, int n, int ld) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= n) return;
    const zd v = W[(size_t)i * ld + j];
    re[(size_t)i * n + j] = v.re;
    im[(size_t)i * n + j] = v.im;
}

// Y = diag(S) (np x np) for the explicit W_c.
__global__ void k_diag_c(const zd* __restrict__ S, zd* __restrict__ Y, int np) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= np) return;
    Y[(size_t)i * np + j] = i == j ? S[i] : zd{0.0, 0.0};
}

template <bool TA, bool TB>
hipError_t cgemm(const zd* A, int lda, long sa, const zd* B, int ldb, long sb, zd* C, int ldc,
                 long sc, int rows, int cols, int K, double alpha, int beta, int nb, hipStream_t s) {
    if (rows % 64 || cols % 64 || K % 16 || rows <= 0 || cols <= 0 || K <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_cgemm<TA, TB>), dim3(cols / 64, rows / 64, nb), dim3(256), 0, s, A, lda, sa,
                       B, ldb, sb, C, ldc, sc, K, alpha, beta);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_scale_grid_c(const double* delta, const double* omega, double eta, double* S,
                               int n, int np, const int* f0p, int nb, hipStream_t s) {
    hipLaunchKernelGGL(k_scale_grid_c, dim3((np + 255) / 256, nb), dim3(256), 0, s, delta, omega,
                       eta, reinterpret_cast<zd*>(S), n, np, f0p);
    return hipGetLastError();
}

hipError_t launch_scale_pi0_c(const double* pre, const double* pim, double* S, int n, int np,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_scale_pi0_c, dim3((np + 255) / 256), dim3(256), 0, s, pre, pim,
                       reinterpret_cast<zd*>(S), n, np);
    return hipGetLastError();
}

hipError_t launch_dielectric_c(const double* V, const double* S, double* E, int np, int nb,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_dielectric_c, dim3((np + 255) / 256, np, nb), dim3(256), 0, s, V,
                       reinterpret_cast<const zd*>(S), reinterpret_cast<zd*>(E), np);
    return hipGetLastError();
}

hipError_t launch_potrf_c(double* E_, double* dinv_, int* info, int np, int nb, hipStream_t s) {
    if (np % 64 != 0) return hipErrorInvalidValue;
    zd* E = reinterpret_cast<zd*>(E_);
    zd* dinv = reinterpret_cast<zd*>(dinv_);
    const int nblk = np / 64;
    const long se = (long)np * np, sd = (long)nblk * 64 * 64;
    for (int k = 0; k < nblk; ++k) {
        zd* col = E + (size_t)k * 64 * np + (size_t)k * 64;     // block (k, k)
        if (k > 0) {
            // E[k:, k] -= L[k:, :k] L[k, :k]^T
            hipError_t e = cgemm<false, true>(E + (size_t)k * 64 * np, np, se, E + (size_t)k * 64 * np,
                                              np, se, col, np, se, np - k * 64, 64, k * 64, -1.0, 1,
                                              nb, s);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(k_potf2_c, dim3(nb), dim3(256), 0, s, E, dinv, info, np, k);
        if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;
        if (k + 1 < nblk) {
            // L[k+1:, k] = E[k+1:, k] L_kk^-T
            zd* below = col + (size_t)64 * np;
            hipError_t e = cgemm<false, true>(below, np, se, dinv + (size_t)k * 64 * 64, 64, sd, below,
                                              np, se, np - (k + 1) * 64, 64, 64, 1.0, 0, nb, s);
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

hipError_t launch_make_y(const double* S, const double* xp, double* Y, int np, int mpad, int nb,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_make_y, dim3((mpad + 255) / 256, np, nb), dim3(256), 0, s,
                       reinterpret_cast<const zd*>(S), xp, reinterpret_cast<zd*>(Y), np, mpad);
    return hipGetLastError();
}

hipError_t launch_fsub_c(const double* L_, const double* dinv_, double* Y_, int np, int mpad, int nb,
                         hipStream_t s) {
    if (np % 64 != 0 || mpad % 64 != 0) return hipErrorInvalidValue;
    const zd* L = reinterpret_cast<const zd*>(L_);
    const zd* dinv = reinterpret_cast<const zd*>(dinv_);
    zd* Y = reinterpret_cast<zd*>(Y_);
    const int nblk = np / 64;
    const long sl = (long)np * np, sd = (long)nblk * 64 * 64, sy = (long)np * mpad;
    for (int i = 0; i < nblk; ++i) {
        zd* yi = Y + (size_t)i * 64 * mpad;
        if (i > 0) {
            // Y_i -= L[i, :i] Z[:i]
            hipError_t e = cgemm<false, false>(L + (size_t)i * 64 * np, np, sl, Y, mpad, sy, yi, mpad,
                                               sy, 64, mpad, i * 64, -1.0, 1, nb, s);
            if (e != hipSuccess) return e;
        }
        // Z_i = L_ii^-1 Y_i
        hipError_t e = cgemm<false, false>(dinv + (size_t)i * 64 * 64, 64, sd, yi, mpad, sy, yi, mpad,
                                           sy, 64, mpad, 64, 1.0, 0, nb, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_sumsq_c(const double* Z, const int* colmap, double* qre, double* qim, int np,
                          int m, int mpad, const int* f0p, int nb, hipStream_t s) {
    hipLaunchKernelGGL(k_sumsq_c, dim3((m + 255) / 256, nb), dim3(256), 0, s,
                       reinterpret_cast<const zd*>(Z), colmap, qre, qim, np, m, mpad, f0p);
    return hipGetLastError();
}

hipError_t launch_wc_explicit_c(const double* L, const double* dinv, const double* S, double* Y,
                                double* W, double* re, double* im, int n, int np, hipStream_t s) {
    hipLaunchKernelGGL(k_diag_c, dim3((np + 255) / 256, np), dim3(256), 0, s,
                       reinterpret_cast<const zd*>(S), reinterpret_cast<zd*>(Y), np);
    if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;
    if (hipError_t e = launch_fsub_c(L, dinv, Y, np, np, 1, s); e != hipSuccess) return e;
    // W = -Z^T Z
    hipError_t e = cgemm<true, false>(reinterpret_cast<const zd*>(Y), np, 0,
                                      reinterpret_cast<const zd*>(Y), np, 0,
                                      reinterpret_cast<zd*>(W), np, 0, np, np, np, -1.0, 0, 1, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_copy_c, dim3((n + 255) / 256, n), dim3(256), 0, s,
                       reinterpret_cast<const zd*>(W), re, im, n, np);
    return hipGetLastError();
}

}  // namespace mgw
