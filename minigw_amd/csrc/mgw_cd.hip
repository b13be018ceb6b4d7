// Cholesky-decomposed ERI form of the screening path (FP64 throughout).
//
// The ERI as a matrix over unique orbital pairs u = (p <= q),
//   M[u][v] = (p_u q_u | p_v q_v),
// is symmetric positive semidefinite (a Coulomb Gram matrix).  A pivoted
// Cholesky decomposition M ~ L L^T, stopped once the largest residual
// diagonal is <= tol * max_u M[u][u], gives every pair an r-vector l_u
// (r = the numerical rank; every residual entry is bounded by that diagonal).
// With B = [l_ia] over the particle-hole pairs (n x r), V = B B^T and
// x_pk = B l_pk, and the Woodbury identity moves the screening into the
// r-dimensional pair space:
//   E = I + S V S = I + G G^T,  G = S B,  C = I + G^T G = I + B^T S^2 B  (r x r)
//   q_pk = -(S x_pk)^T E^-1 (S x_pk) = -l_pk^T (I - C^-1) l_pk
//        = || L_C^-1 l_pk ||^2 - || l_pk ||^2,   C = L_C L_C^T.
// So each frequency factors an r x r matrix instead of an n x n one, and the
// contraction runs over r instead of n.  The same auxiliary-basis form is the
// one the reference's PySCF companion evaluates (pyscf_prep/gw_ac.py:169-219:
// Pi_PQ in the density-fitting basis, (1 - Pi)^-1 - 1 contracted with the pair
// vectors); here the pair basis comes from the ERI itself.
//
// Decomposition on the device: L is column-major (column j of all pairs
// contiguous, Npad = N rounded up to 64), one launch per pivot.  Each launch
// finds its pivot as the argmax of the previous launch's per-block maxima of
// the residual diagonal, so no host round trip is needed between pivots.
#include "mgw_kernels.h"

#include <hip/hip_runtime.h>

namespace mgw {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ inline d4 mfma_f64(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int kBK = 16;          // K slab staged through LDS per step
constexpr int kLdsPad = 4;       // doubles of padding per LDS row
constexpr int kCdU = 16;         // pairs per decomposition workgroup
constexpr int kCdP = 256 / kCdU; // partitions of the j sum per workgroup

// (v, i) <- the larger of (v, i) and (v2, i2); ties go to the smaller index,
// so every lane and every workgroup picks the same pivot.
__device__ inline void argmax_merge(double& v, int& i, double v2, int i2) {
    if (v2 > v || (v2 == v && i2 < i)) {
        v = v2;
        i = i2;
    }
}

__device__ inline void wave_argmax(double& v, int& i) {
    for (int o = 32; o > 0; o >>= 1) {
        const double v2 = __shfl_xor(v, o);
        const int i2 = __shfl_xor(i, o);
        argmax_merge(v, i, v2, i2);
    }
}

// D[u] = M[u][u]; per-block maxima into bm[1] (the "previous step" of step 0).
// One wave per block of kCdU pairs (lanes >= kCdU idle).
__global__ __launch_bounds__(64) void k_cd_init(const double* __restrict__ eri,
                                                const int* __restrict__ pq, int nmo, int N,
                                                double* __restrict__ D, double2* __restrict__ bm,
                                                int* __restrict__ rank) {
    const int u = blockIdx.x * kCdU + threadIdx.x;
    const size_t nmo2 = (size_t)nmo * nmo;
    double d = -1.0;
    if (threadIdx.x < kCdU && u < N) {
        d = eri[(size_t)pq[u] * nmo2 + pq[u]];
        D[u] = d;
    }
    int i = u;
    wave_argmax(d, i);
    if (threadIdx.x == 0) bm[gridDim.x + blockIdx.x] = make_double2(d, (double)i);
    if (blockIdx.x == 0 && threadIdx.x == 0) *rank = -1;
}

// Pivot k.  Every workgroup reduces the previous step's block maxima to the
// pivot (v, P); if v <= tol * (first pivot) the decomposition has converged
// and rank = k.  Otherwise column k:
//   L[k][u] = (M[P][u] - sum_{j<k} L[j][u] L[j][P]) / sqrt(v),
//   D[u] -= L[k][u]^2  (D[P] = 0),
// with the j sum split kCdP ways (kCdU consecutive pairs per partition, one
// 128-byte segment per column), and the block's new maximum into bm[k & 1].
__global__ __launch_bounds__(256) void k_cd_step(const double* __restrict__ eri,
                                                 const int* __restrict__ pq, int nmo, int N,
                                                 int Npad, double* __restrict__ Lc,
                                                 double* __restrict__ D, double2* __restrict__ bm,
                                                 int* __restrict__ piv, double* __restrict__ pval,
                                                 int* __restrict__ rank, int k, double tol) {
    __shared__ double part[kCdP][kCdU];
    __shared__ double rv[4];
    __shared__ int ri[4];
    if (*rank >= 0) return;                       // converged at an earlier pivot
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, nblk = gridDim.x;
    const double2* prev = bm + (size_t)((k + 1) & 1) * nblk;
    double v = -1.0;
    int vi = 0x7fffffff;
    for (int e = t; e < nblk; e += 256) {
        const double2 x = prev[e];
        argmax_merge(v, vi, x.x, (int)x.y);
    }
    wave_argmax(v, vi);
    if (lane == 0) {
        rv[wave] = v;
        ri[wave] = vi;
    }
    __syncthreads();
    v = rv[0];
    vi = ri[0];
    for (int q = 1; q < 4; ++q) argmax_merge(v, vi, rv[q], ri[q]);
    const double thr = tol * (k == 0 ? v : pval[0]);
    if (!(v > thr) || vi < 0 || vi >= N) {        // converged (or nothing to decompose)
        if (blockIdx.x == 0 && t == 0) *rank = k;
        return;
    }

    const int ul = t % kCdU, jp = t / kCdU;
    const double* lu = Lc + blockIdx.x * kCdU + ul;
    const double* lp = Lc + vi;
    double s = 0.0;
    for (int j = jp; j < k; j += kCdP) s = fma(lu[(size_t)j * Npad], lp[(size_t)j * Npad], s);
    part[jp][ul] = s;
    __syncthreads();
    if (wave != 0) return;
    const int u = blockIdx.x * kCdU + lane;       // lanes >= kCdU carry no pair
    const bool mine = lane < kCdU;
    s = 0.0;
    if (mine) {
#pragma unroll
        for (int q = 0; q < kCdP; ++q) s += part[q][lane];
    }
    double l = 0.0, d = -1.0;
    if (mine && u < N) {
        const double x = eri[(size_t)pq[vi] * ((size_t)nmo * nmo) + pq[u]];
        l = (x - s) / sqrt(v);
        d = (u == vi) ? 0.0 : D[u] - l * l;
        D[u] = d;
    }
    if (mine) Lc[(size_t)k * Npad + u] = l;
    int di = u;
    wave_argmax(d, di);
    if (lane == 0) bm[(size_t)(k & 1) * nblk + blockIdx.x] = make_double2(d, (double)di);
    if (blockIdx.x == 0 && lane == 0) {
        piv[k] = vi;
        pval[k] = v;
    }
}

// out[0] = min_u D[u], out[1] = number of non-finite D[u] (one workgroup).
__global__ __launch_bounds__(256) void k_cd_residual(const double* __restrict__ D, int N,
                                                     double* __restrict__ out) {
    __shared__ double smin[256];
    __shared__ int sbad[256];
    double m = 0.0;
    int bad = 0;
    for (int u = threadIdx.x; u < N; u += 256) {
        const double d = D[u];
        if (!isfinite(d)) ++bad;
        else m = fmin(m, d);
    }
    smin[threadIdx.x] = m;
    sbad[threadIdx.x] = bad;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + o]);
            sbad[threadIdx.x] += sbad[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = smin[0];
        out[1] = (double)sbad[0];
    }
}

// Bt[j][ph] = L[j][phu[ph]] for j < r, ph < n; 0 elsewhere (rpad x np).
__global__ void k_cd_gather_b(const double* __restrict__ Lc, int Npad, const int* __restrict__ phu,
                              int r, int n, double* __restrict__ Bt, int rpad, int np) {
    const size_t total = (size_t)rpad * np;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)(e / np), ph = (int)(e % np);
        Bt[e] = (j < r && ph < n) ? Lc[(size_t)j * Npad + phu[ph]] : 0.0;
    }
}

// Lx[c][j] = L[j][colu[c]] for j < r, c < mu (0 elsewhere; mpad x rpad) and
// nrm[c] = sum_j Lx[c][j]^2; one wave per column.
__global__ __launch_bounds__(256) void k_cd_gather_x(const double* __restrict__ Lc, int Npad,
                                                     const int* __restrict__ colu, int r, int mu,
                                                     double* __restrict__ Lx, double* __restrict__ nrm,
                                                     int rpad, int mpad) {
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (c >= mpad) return;
    const bool live = c < mu;
    const int cu = live ? colu[c] : 0;
    double s = 0.0;
    for (int j = lane; j < rpad; j += 64) {
        const double x = (live && j < r) ? Lc[(size_t)j * Npad + cu] : 0.0;
        Lx[(size_t)c * rpad + j] = x;
        s = fma(x, x, s);
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) nrm[c] = s;
}

// C_w[i][j] = delta_ij + sum_k Bt[i][k] S_w[k]^2 Bt[j][k] on the 128 x 128
// tiles on and below the diagonal (blockIdx.x = lower-triangular tile index,
// blockIdx.y = w).  256 threads = 4 waves of 64 x 64 (4 x 4 MFMA blocks of
// v_mfma_f64_16x16x4f64); K slabs of 16 staged through LDS, the next slab
// fetched into registers while the current one is multiplied.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_aux_dielectric(const double* __restrict__ Bt,
                                                        const double* __restrict__ S,
                                                        double* __restrict__ C, int rpad, int np) {
    __shared__ double As[kBK][kCT + kLdsPad];
    __shared__ double Bs[kBK][kCT + kLdsPad];
    const int w = blockIdx.y;
    int lin = blockIdx.x, ti = 0;
    while (lin > ti) {
        lin -= ti + 1;
        ++ti;
    }
    const int tj = lin;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int li = lane & 15, lk = lane >> 4;
    const int lrow = t >> 1, lkq = (t & 1) * 8;
    const double* arow = Bt + (size_t)(ti * kCT + lrow) * np + lkq;
    const double* brow = Bt + (size_t)(tj * kCT + lrow) * np + lkq;
    const double* sw = S + (size_t)w * np + lkq;
    double ra[8], rb[8];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 8; u += 2) {
            const double2 a = *reinterpret_cast<const double2*>(arow + k0 + u);
            const double2 s = *reinterpret_cast<const double2*>(sw + k0 + u);
            const double2 b = *reinterpret_cast<const double2*>(brow + k0 + u);
            ra[u] = a.x * (s.x * s.x);
            ra[u + 1] = a.y * (s.y * s.y);
            rb[u] = b.x;
            rb[u + 1] = b.y;
        }
    };
    d4 acc[4][4];
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
    fetch(0);
    for (int k0 = 0; k0 < np; k0 += kBK) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            As[lkq + u][lrow] = ra[u];
            Bs[lkq + u][lrow] = rb[u];
        }
        __syncthreads();
        if (k0 + kBK < np) fetch(k0 + kBK);
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
    double* Cw = C + (size_t)w * rpad * rpad;
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = ti * kCT + wm + x * 16 + lk + 4 * r;
                const int col = tj * kCT + wn + y * 16 + li;
                Cw[(size_t)row * rpad + col] = (row == col ? 1.0 : 0.0) + acc[x][y][r];
            }
}

// q[*f0p + w][c] = sum_rt qpart[w][rt][colmap[c]] - nrm[colmap[c]].
__global__ void k_reduce_q_cd(const double* __restrict__ qpart, const double* __restrict__ nrm,
                              const int* __restrict__ colmap, double* __restrict__ q, int nrt,
                              int m, int mpad, const int* __restrict__ f0p) {
    const int w = blockIdx.y, f0 = *f0p;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < m; c += gridDim.x * blockDim.x) {
        const int cu = colmap[c];
        double s = 0.0;
        for (int rt = 0; rt < nrt; ++rt) s += qpart[((size_t)w * nrt + rt) * mpad + cu];
        q[(size_t)(f0 + w) * m + c] = s - nrm[cu];
    }
}

__device__ inline void atomic_max_nonneg(unsigned long long* p, double v) {
    atomicMax(p, (unsigned long long)__double_as_longlong(v));   // order-preserving for v >= 0
}

// Largest asymmetries the decomposed form relies on: out[0] = max |M[u][v] -
// M[v][u]| over unique pairs, out[1] = max |(pk|ia) - (kp|ia)| over the
// particle-hole columns, out[2] = max |M[u][v]| (as double bit patterns;
// out zeroed by the caller).
__global__ __launch_bounds__(256) void k_cd_asymmetry(const double* __restrict__ eri,
                                                      const int* __restrict__ pq, int nmo, int nocc,
                                                      int N, unsigned long long* __restrict__ out) {
    const size_t nmo2 = (size_t)nmo * nmo, total = (size_t)N * N;
    const int nvir = nmo - nocc;
    const size_t total2 = nmo2 * (size_t)nocc * nvir;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const int u = (int)(e / N), v = (int)(e % N);
        if (v < u) continue;
        const double x = eri[(size_t)pq[u] * nmo2 + pq[v]];
        a2 = fmax(a2, fabs(x));
        if (v > u) a0 = fmax(a0, fabs(x - eri[(size_t)pq[v] * nmo2 + pq[u]]));
    }
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total2; e += stride) {
        const int a = (int)(e % nvir);
        const int i = (int)((e / nvir) % nocc);
        const size_t pk = e / ((size_t)nocc * nvir);
        const int p = (int)(pk / nmo), k = (int)(pk % nmo);
        if (p >= k) continue;
        const size_t col = (size_t)i * nmo + nocc + a;
        a1 = fmax(a1, fabs(eri[((size_t)p * nmo + k) * nmo2 + col] -
                           eri[((size_t)k * nmo + p) * nmo2 + col]));
    }
    for (int o = 32; o > 0; o >>= 1) {
        a0 = fmax(a0, __shfl_xor(a0, o));
        a1 = fmax(a1, __shfl_xor(a1, o));
        a2 = fmax(a2, __shfl_xor(a2, o));
    }
    if ((threadIdx.x & 63) == 0) {
        atomic_max_nonneg(out + 0, a0);
        atomic_max_nonneg(out + 1, a1);
        atomic_max_nonneg(out + 2, a2);
    }
}

}  // namespace

int cd_blocks(int N) { return round_up(N, kCdU) / kCdU; }
int cd_npad(int N) { return cd_blocks(N) * kCdU; }

hipError_t launch_cd_init(const double* eri, const int* pq, int nmo, int N, double* D, double2* bm,
                          int* rank, hipStream_t s) {
    if (N <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_cd_init, dim3(cd_blocks(N)), dim3(64), 0, s, eri, pq, nmo, N, D, bm,
                       rank);
    return hipGetLastError();
}

hipError_t launch_cd_steps(const double* eri, const int* pq, int nmo, int N, double* Lc, double* D,
                           double2* bm, int* piv, double* pval, int* rank, int k0, int k1,
                           double tol, hipStream_t s) {
    if (N <= 0 || k0 < 0 || k1 > N) return hipErrorInvalidValue;
    const int nblk = cd_blocks(N), Npad = nblk * kCdU;
    for (int k = k0; k < k1; ++k)
        hipLaunchKernelGGL(k_cd_step, dim3(nblk), dim3(256), 0, s, eri, pq, nmo, N, Npad, Lc, D, bm,
                           piv, pval, rank, k, tol);
    return hipGetLastError();
}

hipError_t launch_cd_asymmetry(const double* eri, const int* pq, int nmo, int nocc, int N,
                               unsigned long long* out, hipStream_t s) {
    const size_t total = std::max((size_t)N * N, (size_t)nmo * nmo * nocc * (nmo - nocc));
    const unsigned g = (unsigned)std::min<size_t>((total + 255) / 256, 16384);
    hipLaunchKernelGGL(k_cd_asymmetry, dim3(g > 0 ? g : 1), dim3(256), 0, s, eri, pq, nmo, nocc, N,
                       out);
    return hipGetLastError();
}

hipError_t launch_cd_residual(const double* D, int N, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_cd_residual, dim3(1), dim3(256), 0, s, D, N, out);
    return hipGetLastError();
}

hipError_t launch_cd_gather(const double* Lc, int N, const int* phu, const int* colu, int r, int n,
                            int mu, double* Bt, double* Lx, double* nrm, int rpad, int np, int mpad,
                            hipStream_t s) {
    if (rpad % kCT != 0 || np % kBK != 0 || r > rpad) return hipErrorInvalidValue;
    const int Npad = cd_blocks(N) * kCdU;
    const size_t total = (size_t)rpad * np;
    const unsigned g = (unsigned)std::min<size_t>((total + 255) / 256, 65535 * 8);
    hipLaunchKernelGGL(k_cd_gather_b, dim3(g), dim3(256), 0, s, Lc, Npad, phu, r, n, Bt, rpad, np);
    hipLaunchKernelGGL(k_cd_gather_x, dim3((mpad + 3) / 4), dim3(256), 0, s, Lc, Npad, colu, r, mu,
                       Lx, nrm, rpad, mpad);
    return hipGetLastError();
}

hipError_t launch_aux_dielectric(const double* Bt, const double* S, double* C, int rpad, int np,
                                 int nb, hipStream_t s) {
    if (rpad % kCT != 0 || np % kBK != 0 || nb <= 0) return hipErrorInvalidValue;
    const int nt = rpad / kCT;
    hipLaunchKernelGGL(k_aux_dielectric, dim3(nt * (nt + 1) / 2, nb), dim3(256), 0, s, Bt, S, C,
                       rpad, np);
    return hipGetLastError();
}

hipError_t launch_reduce_q_cd(const double* qpart, const double* nrm, const int* colmap, double* q,
                              int nrt, int m, int mpad, const int* f0p, int nb, hipStream_t s) {
    const unsigned g = (unsigned)std::min(4096, (m + 255) / 256);
    hipLaunchKernelGGL(k_reduce_q_cd, dim3(g > 0 ? g : 1, nb), dim3(256), 0, s, qpart, nrm, colmap,
                       q, nrt, m, mpad, f0p);
    return hipGetLastError();
}

}  // namespace mgw
