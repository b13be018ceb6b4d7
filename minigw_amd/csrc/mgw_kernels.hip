// CDNA4 (gfx950) kernels of the screening + Sigma_c path, FP64 form.
//
// Method (per imaginary frequency w', docs/formulation.md, reference
// src/gw/sigma_c_driver.cpp:119-169 and src/workspace/screening_workspace.cpp:21-102):
//   eps = I - diag(Pi0) V,  W_c = eps^-1 diag(Pi0),  q_pk = x_pk^T W_c x_pk.
// At eta = 0, Pi0 = -S^2 with S = sqrt(4D/(w^2+D^2)) > 0, and eps is similar to
// the symmetrised dielectric matrix  E = S^-1 eps S = I + S V S  (symmetric
// positive definite, V being a Coulomb Gram matrix).  Hence
//   W_c = eps^-1 diag(Pi0) = -S E^-1 S = -(L^-1 S)^T (L^-1 S),  E = L L^T,
//   q_pk = -|| L^-1 S x_pk ||^2.
// The kernels factor E by blocked Cholesky (potrf), invert the factor (trtri),
// and contract M = L^-1 S against X streamed from the resident ERI, with the
// sum of squares taken in the GEMM epilogue (Y = W_c X is never written).
//
// MFMA: v_mfma_f64_16x16x4_f64.  Operand/result lane layout (wave64):
//   a: A[i][k], i = lane % 16, k = lane / 16      (16 x 4 block)
//   b: B[k][j], j = lane % 16, k = lane / 16      (4 x 16 block)
//   d[r]: D[i][j], i = lane / 16 + 4 * r, j = lane % 16   (checked on the device)
#include "mgw_kernels.h"
#include "mgw_tile64.h"

#include <hip/hip_runtime.h>

namespace mgw {
namespace {

// ----------------------------------------------------------------- simple maps
__global__ void k_build_v(const double* __restrict__ eri, const int64_t* __restrict__ rowbase,
                          const int* __restrict__ colidx, double* __restrict__ v, int n, int np) {
    const size_t total = (size_t)np * np;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(e / np), c = (int)(e % np);
        v[e] = (r < n && c < n) ? eri[rowbase[r] + colidx[c]] : 0.0;
    }
}

__global__ void k_pad_square(const double* __restrict__ src, double* __restrict__ dst, int n,
                             int np) {
    const size_t total = (size_t)np * np;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(e / np), c = (int)(e % np);
        dst[e] = (r < n && c < n) ? src[(size_t)r * n + c] : 0.0;
    }
}

__global__ void k_scale_grid(const double* __restrict__ delta, const double* __restrict__ omega,
                             double* __restrict__ S, int n, int np, const int* __restrict__ f0p,
                             int nb) {
    const int f0 = *f0p;
    const int total = np * nb;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        const int w = e / np, r = e % np;
        double s = 0.0;
        if (r < n) {
            const double d = delta[r], om = omega[f0 + w];
            s = sqrt(4.0 * d / (om * om + d * d));
        }
        S[e] = s;
    }
}

__global__ void k_scale_pi0(const double* __restrict__ pi0, double* __restrict__ S, int n, int np) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < np; r += gridDim.x * blockDim.x)
        S[r] = r < n ? sqrt(-pi0[r]) : 0.0;
}

// lower triangle (incl. diagonal) of A_w = I + S V S; one matrix per blockIdx.y
__global__ void k_dielectric(const double* __restrict__ v, const double* __restrict__ S,
                             double* __restrict__ A, int np) {
    const int w = blockIdx.y;
    const double* Sw = S + (size_t)w * np;
    double* Aw = A + (size_t)w * np * np;
    const size_t total = (size_t)np * np;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(e / np), c = (int)(e % np);
        if (c > r) continue;
        Aw[e] = (r == c ? 1.0 : 0.0) + (Sw[r] * v[e]) * Sw[c];
    }
}

// ------------------------------------------------------------------ potrf pieces
// Diagonal block k of every matrix: Cholesky and the inverse of its factor,
// held in registers: thread t owns row i = t / 4, columns c0 = 16 (t % 4) ..
// c0 + 15 of the working block T (lower part of A_kk) and of B (starts as I,
// ends as L^-1 up to row scaling).  Step j: the owners publish column j of T
// (zero for rows i <= j; the pivot p_j = T[j][j] goes to its own slot) and row
// j of B to LDS (double-buffered, one barrier per step); every row then
// applies the rank-1 update with the unscaled column
//   T[i][l] -= (T[i][j] / p_j) T[l][j],  B[i][l] -= (T[i][j] / p_j) B[j][l]
// to all its 32 registers as plain FMAs.  The published zeros make it exact
// where it must not act: rows i <= j have multiplier 0, columns l <= j of T
// take 0 (so finished columns and the pivot column keep their values), and
// B[j][l] is exactly 0 for l > j (B stays lower triangular).  The upper part
// of T (l > i) collects values that nothing reads: rows publish 0 once the
// step passes them and the final stores keep only l <= i.  Steps are
// unrolled by 16 so the published register is fixed at compile time.
// Finally L[i][c] = T[i][c] / sqrt(p_c) (c <= i) and L^-1[i][c] = B[i][c] /
// sqrt(p_i).
// pre != 0: the block first takes the product of block column k - 1 that the
// calling sweep left out, A_kk -= L_k,k-1 L_k,k-1^T (FP64 MFMA tile, staged
// through LDS).
__global__ __launch_bounds__(256) void k_potf2(double* __restrict__ A, double* __restrict__ dinv,
                                               int* __restrict__ info, int np, int k, int pre) {
    __shared__ double col[2][64], row[2][64];
    __shared__ double pivs[64];                  // p_j, then 1 / sqrt(p_j)
    __shared__ double As[kBK][64 + kLdsPad];
    __shared__ double Bs[kBK][64 + kLdsPad];
    __shared__ double upd[64][65];
    const int w = blockIdx.x, t = threadIdx.x;
    const int i = t >> 2, q = t & 3, c0 = q * 16;
    const int nblk = np / kNB;
    double* Aw = A + (size_t)w * np * np;
    const size_t off = (size_t)k * kNB * np + (size_t)k * kNB;
    if (pre) {
        Tile64 tile;
        tile.zero();
        const double* Lk = Aw + (size_t)k * kNB * np;
        gemm64_k64<false>(Lk + (size_t)(k - 1) * kNB, np, Lk + (size_t)(k - 1) * kNB, np, tile, As,
                          Bs);
        tile.for_each([&](int r, int c, double v) { upd[r][c] = v; });
        __syncthreads();
    }
    double T[16], B[16];
    {
        // the whole 64 x 64 block is in bounds: load it unconditionally, mask after
        const double2* src = reinterpret_cast<const double2*>(Aw + off + (size_t)i * np + c0);
        double v[16];
#pragma unroll
        for (int c = 0; c < 16; c += 2) {
            const double2 x = src[c / 2];
            v[c] = x.x;
            v[c + 1] = x.y;
        }
        if (pre) {
#pragma unroll
            for (int c = 0; c < 16; ++c) v[c] -= upd[i][c0 + c];
        }
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            T[c] = c0 + c <= i ? v[c] : 0.0;
            B[c] = c0 + c == i ? 1.0 : 0.0;
        }
    }
    int bad = 0;
    for (int jb = 0; jb < 4; ++jb) {
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) {
            const int j = 16 * jb + jj, buf = jj & 1;
            if (q == jb) {
                col[buf][i] = i > j ? T[jj] : 0.0;
                if (i == j) pivs[j] = T[jj];
            }
            if (i == j) {
#pragma unroll
                for (int c = 0; c < 16; c += 2)
                    *reinterpret_cast<double2*>(&row[buf][c0 + c]) = make_double2(B[c], B[c + 1]);
            }
            __syncthreads();
            const double piv = pivs[j];
            if (!(piv > 0.0) && !bad) bad = j + 1;
            const double ip = 1.0 / (piv > 0.0 ? piv : 1.0);
            const double a = col[buf][i] * ip;
#pragma unroll
            for (int c = 0; c < 16; c += 2) {
                const double2 x = *reinterpret_cast<const double2*>(&col[buf][c0 + c]);
                const double2 y = *reinterpret_cast<const double2*>(&row[buf][c0 + c]);
                T[c] = fma(-a, x.x, T[c]);
                T[c + 1] = fma(-a, x.y, T[c + 1]);
                B[c] = fma(-a, y.x, B[c]);
                B[c + 1] = fma(-a, y.y, B[c + 1]);
            }
        }
    }
    __syncthreads();                             // every wave has read pivs[63]
    if (t < 64) {
        const double p = pivs[t];
        pivs[t] = 1.0 / sqrt(p > 0.0 ? p : 1.0);
    }
    __syncthreads();
    if (bad && t == 0 && info[w] == 0) info[w] = k * kNB + bad;
    double* D = dinv + ((size_t)w * nblk + k) * 64 * 64;
    double* dst = Aw + off + (size_t)i * np + c0;
    const double ri = pivs[i];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const int l = c0 + c;
        dst[c] = l <= i ? T[c] * pivs[l] : 0.0;        // zeros above the diagonal
        D[i * 64 + l] = l <= i ? B[c] * ri : 0.0;
    }
}

// Panel: A_ik <- A_ik * Dinv_k^T for i > k.
__global__ __launch_bounds__(256) void k_trsm_panel(double* __restrict__ A,
                                                    const double* __restrict__ dinv, int np,
                                                    int k) {
    __shared__ double As[kBK][64 + kLdsPad];
    __shared__ double Bs[kBK][64 + kLdsPad];
    const int w = blockIdx.y, i = k + 1 + blockIdx.x;
    const int nblk = np / kNB;
    double* Aw = A + (size_t)w * np * np;
    const double* D = dinv + ((size_t)w * nblk + k) * 64 * 64;
    const double* Aik = Aw + (size_t)i * kNB * np + (size_t)k * kNB;
    Tile64 tile;
    tile.zero();
    gemm64_k64<false>(Aik, np, D, 64, tile, As, Bs);
    double* out = Aw + (size_t)i * kNB * np + (size_t)k * kNB;
    tile.for_each([&](int r, int c, double v) { out[(size_t)r * np + c] = v; });
}

// Left-looking update of block column k: A_ik -= sum_{j<k} L_ij L_kj^T, i >= k
// (one 64 x 64 tile per workgroup, K = 64 k).  Reads only finished columns of
// L, so each trailing element is updated once per block column instead of
// being re-read and re-written after every panel.
__global__ __launch_bounds__(256) void k_potrf_update(double* __restrict__ A, int np, int k) {
    __shared__ double As[kBK][64 + kLdsPad];
    __shared__ double Bs[kBK][64 + kLdsPad];
    const int w = blockIdx.y, i = k + blockIdx.x;
    double* Aw = A + (size_t)w * np * np;
    const double* Li = Aw + (size_t)i * kNB * np;
    const double* Lk = Aw + (size_t)k * kNB * np;
    Tile64 tile;
    tile.zero();
    gemm64<false>(Li, np, Lk, np, 0, k * kNB, tile, As, Bs);
    double* out = Aw + (size_t)i * kNB * np + (size_t)k * kNB;
    tile.for_each([&](int r, int c, double v) { out[(size_t)r * np + c] -= v; });
}

// ------------------------------------------------------------------ trtri pieces
// T_i = sum_{k=j+1}^{i} M_ik L_kj, i > j (M = already inverted columns > j).
__global__ __launch_bounds__(256) void k_trtri_t(const double* __restrict__ A,
                                                 double* __restrict__ tbuf, int np, int j) {
    __shared__ double As[kBK][64 + kLdsPad];
    __shared__ double Bs[kBK][64 + kLdsPad];
    const int w = blockIdx.y, i = j + 1 + blockIdx.x;
    const double* Aw = A + (size_t)w * np * np;
    const int kbeg = (j + 1) * kNB, kend = (i + 1) * kNB;
    const double* Arow = Aw + (size_t)i * kNB * np;     // rows of block i
    const double* Bcol = Aw + (size_t)j * kNB;          // column block j
    Tile64 tile;
    tile.zero();
    gemm64<true>(Arow, np, Bcol, np, kbeg, kend, tile, As, Bs);
    double* out = tbuf + (size_t)w * np * kNB + (size_t)i * kNB * kNB;
    tile.for_each([&](int r, int c, double v) { out[r * kNB + c] = v; });
}

// M_jj = Dinv_j ; M_ij = -T_i Dinv_j for i > j.
__global__ __launch_bounds__(256) void k_trtri_m(double* __restrict__ A,
                                                 const double* __restrict__ dinv,
                                                 const double* __restrict__ tbuf, int np, int j) {
    __shared__ double As[kBK][64 + kLdsPad];
    __shared__ double Bs[kBK][64 + kLdsPad];
    const int w = blockIdx.y, nblk = np / kNB;
    double* Aw = A + (size_t)w * np * np;
    const double* D = dinv + ((size_t)w * nblk + j) * 64 * 64;
    if (blockIdx.x == 0) {
        double* out = Aw + (size_t)j * kNB * np + (size_t)j * kNB;
        for (int e = threadIdx.x; e < 64 * 64; e += 256) out[(size_t)(e >> 6) * np + (e & 63)] = D[e];
        return;
    }
    const int i = j + blockIdx.x;
    const double* T = tbuf + (size_t)w * np * kNB + (size_t)i * kNB * kNB;
    Tile64 tile;
    tile.zero();
    for (int k0 = 0; k0 < kNB; k0 += kBK) {
        load_rows_kmajor(As, T, kNB, 0, k0);
        load_kn(Bs, D, 64, k0, 0);
        __syncthreads();
        tile.step(As, Bs);
        __syncthreads();
    }
    double* out = Aw + (size_t)i * kNB * np + (size_t)j * kNB;
    tile.for_each([&](int r, int c, double v) { out[(size_t)r * np + c] = -v; });
}

// --------------------------------------------------------- fused contraction FP64
// One workgroup: rows [r0, r0+128) of Z = M S X and columns [c0, c0+128);
// K runs over [0, r0+128) (M lower triangular).  Epilogue: column sums of Z^2
// over the tile's rows -> qpart[w][row tile][column].
// 4 waves in 2 x 2, each a 64 x 64 sub-tile (4 x 4 MFMA blocks).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_contract_f64(
    const double* __restrict__ M, const double* __restrict__ S, const double* __restrict__ xp,
    double* __restrict__ qpart, int np, int mpad) {
    __shared__ double As[kBK][kCT + kLdsPad];
    __shared__ double Bs[kBK][kCT + kLdsPad];
    __shared__ double red[2][kCT];
    const int w = blockIdx.z;
    // large-K row tiles first; the column tiles of one row panel run together
    const int ncol = gridDim.x, nrow = gridDim.y;
    int lin = blockIdx.x + ncol * blockIdx.y;
    const int rt = nrow - 1 - lin / ncol, r0 = rt * kCT, c0 = (lin % ncol) * kCT;
    const double* Mw = M + (size_t)w * np * np;
    const double* Sw = S ? S + (size_t)w * np : nullptr;     // NULL: no column scaling
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int li = lane & 15, lk = lane >> 4;

    // loader coordinates: 128 rows/cols x 16 k, 8 consecutive k per thread;
    // the next slab is fetched into registers while the current one is used
    const int lrow = t >> 1, lkq = (t & 1) * 8;
    const double* arow = Mw + (size_t)(r0 + lrow) * np;
    const double* bsrc = xp + (size_t)(c0 + lrow) * np + lkq;
    double ra[8], rb[8];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int kk = k0 + lkq + u;
            ra[u] = (kk <= r0 + lrow) ? (S ? arow[kk] * Sw[kk] : arow[kk]) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; u += 2) {
            const double2 v = *reinterpret_cast<const double2*>(bsrc + k0 + u);
            rb[u] = v.x;
            rb[u + 1] = v.y;
        }
    };

    d4 acc[4][4];
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

    const int kend = min(np, r0 + kCT);
    fetch(0);
    for (int k0 = 0; k0 < kend; k0 += kBK) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            As[lkq + u][lrow] = ra[u];
            Bs[lkq + u][lrow] = rb[u];
        }
        __syncthreads();
        if (k0 + kBK < kend) fetch(k0 + kBK);
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

    // epilogue: per-column sum of squares over this wave's 64 rows
    for (int y = 0; y < 4; ++y) {
        double s = 0.0;
        for (int x = 0; x < 4; ++x)
            for (int r = 0; r < 4; ++r) s += acc[x][y][r] * acc[x][y][r];
        s += __shfl_xor(s, 16);
        s += __shfl_xor(s, 32);
        if (lk == 0) red[wave >> 1][wn + y * 16 + li] = s;
    }
    __syncthreads();
    if (t < kCT) {
        const double v = red[0][t] + red[1][t];
        qpart[((size_t)w * (np / kCT) + rt) * mpad + c0 + t] = v;
    }
}

// xp[c][k] = src[pkrow[c] + colidx[k]] for k < n, 0 for k >= n (mpad x np).
__global__ void k_pack_x(const double* __restrict__ src, const int64_t* __restrict__ pkrow,
                         const int* __restrict__ colidx, double* __restrict__ xp, int n, int np,
                         int mpad) {
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (c >= mpad) return;
    const int64_t base = pkrow[c];
    double* o = xp + (size_t)c * np;
    for (int k = lane; k < np; k += 64) o[k] = k < n ? src[base + colidx[k]] : 0.0;
}

// pkrow[c] = offset of row (p, k0 + c) of the held ERI (rows of rs values) for
// c < width, 0 in the padding
__global__ void k_pk_rows(int64_t* __restrict__ pkrow, int p, int k0, int width, int nmo, int mpad,
                          int64_t rs) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= mpad) return;
    pkrow[c] = c < width ? ((int64_t)p * nmo + k0 + c) * rs : 0;
}

__global__ void k_reduce_q(const double* __restrict__ qpart, const int* __restrict__ colmap,
                           double* __restrict__ q, int nrt, int m, int mpad,
                           const int* __restrict__ f0p) {
    const int w = blockIdx.y, f0 = *f0p;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < m; c += gridDim.x * blockDim.x) {
        const int cu = colmap ? colmap[c] : c;
        double s = 0.0;
        for (int rt = 0; rt < nrt; ++rt) s += qpart[((size_t)w * nrt + rt) * mpad + cu];
        q[(size_t)(f0 + w) * m + c] = -s;
    }
}

// flag = 1 if (pk|ia) != (kp|ia) for any p, k and occupied-virtual (i, a).
// (pk|ia) at eri[(p nmo + k) rs + i cs + co + a]
__global__ void k_check_pk_symmetry(const double* __restrict__ eri, int nmo, int nocc, int64_t rs,
                                    int cs, int co, int* __restrict__ flag) {
    const int nvir = nmo - nocc;
    const size_t total = (size_t)nmo * nmo * nocc * nvir;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int a = (int)(e % nvir);
        const int i = (int)((e / nvir) % nocc);
        const size_t pk = e / ((size_t)nocc * nvir);
        const int p = (int)(pk / nmo), k = (int)(pk % nmo);
        if (p >= k) continue;
        const size_t col = (size_t)i * cs + co + a;
        if (eri[((size_t)p * nmo + k) * rs + col] != eri[((size_t)k * nmo + p) * rs + col]) *flag = 1;
    }
}

// One slab of ERI rows p in [p0, p0 + rows) (rows (p, k) of nmo^2 values, C
// order) -> the held parts: eph[(p nmo + k) n + i nvir + a] = (pk|ia) and
// sxd[p nocc + i] = (pi|pi), i < nocc <= a.
__global__ void k_gather_ph(const double* __restrict__ slab, int p0, int rows, int nmo, int nocc,
                            double* __restrict__ eph, double* __restrict__ sxd) {
    const int nvir = nmo - nocc, n = nocc * nvir;
    const size_t n2 = (size_t)nmo * nmo;
    const size_t total = (size_t)rows * nmo * n;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int ia = (int)(e % n);
        const size_t pk = e / n;                       // slab row (p - p0) nmo + k
        const int i = ia / nvir, a = ia % nvir;
        eph[(size_t)p0 * nmo * n + e] = slab[pk * n2 + (size_t)i * nmo + nocc + a];
    }
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < rows * nocc) {
        const int pr = t / nocc, i = t % nocc, p = p0 + pr;
        sxd[(size_t)p * nocc + i] = slab[((size_t)pr * nmo + i) * n2 + (size_t)p * nmo + i];
    }
}

// W_c = -(M S)^T (M S), plain FMA (diagnostic / host-panel path, small n).
__global__ void k_wc_explicit(const double* __restrict__ M, const double* __restrict__ S,
                              double* __restrict__ wc, int n, int np) {
    const size_t total = (size_t)n * n;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(e / n), c = (int)(e % n);
        const int t0 = r > c ? r : c;
        double s = 0.0;
        for (int t = t0; t < n; ++t) s += M[(size_t)t * np + r] * M[(size_t)t * np + c];
        wc[e] = -(S[r] * s) * S[c];
    }
}

// ------------------------------------------------------------- Sigma_c, Sigma_x
// Complex helpers with the textbook formulas of the host complex type and
// Smith's division, evaluated without FMA contraction.
struct cplx {
    double re, im;
};
__device__ inline cplx cmul(cplx a, cplx b) {
#pragma clang fp contract(off)
    return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ inline cplx cdiv(cplx a, cplx b) {
#pragma clang fp contract(off)
    // Smith: |c| < |d| branch as libgcc's __divdc3 / CPython's complex division
    if (fabs(b.re) < fabs(b.im)) {
        const double ratio = b.re / b.im;
        const double denom = (b.re * ratio) + b.im;
        return cplx{((a.re * ratio) + a.im) / denom, ((a.im * ratio) - a.re) / denom};
    }
    const double ratio = b.im / b.re;
    const double denom = (b.im * ratio) + b.re;
    return cplx{((a.im * ratio) + a.re) / denom, (a.im - (a.re * ratio)) / denom};
}
__device__ inline double cabs_(cplx a) { return hypot(a.re, a.im); }

// g[(n * nfreq + f) * nmo + k] = d / (d^2 + w_f^2), d = (fermi - e_k) + i w_n: the
// Green's-function factor of the quadrature, which does not depend on the state.
// Every operation rounds as the host's d / (d * d - (i w_f) * (i w_f)) does
// (sigma_c_driver.cpp:154-160): Re(d^2) + w_f^2 cancels almost exactly at
// w_f ~ w_n >> |e_k - fermi|, and an FMA there (one rounding instead of two)
// moved Sigma_c by 7e-10 of its largest value on a molecule with a 1e-4 Ha
// HOMO-LUMO gap.
__global__ void k_sigma_g(const double* __restrict__ mo, const double* __restrict__ omega,
                          double fermi, int nmo, int nfreq, double2* __restrict__ g) {
#pragma clang fp contract(off)
    const size_t total = (size_t)nfreq * nfreq * nmo;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(e % nmo);
        const size_t nf = e / nmo;
        const int f = (int)(nf % nfreq), n = (int)(nf / nfreq);
        const double wn = omega[n], wp = omega[f];
        const cplx d{fermi - mo[k], wn};
        const cplx dd = cmul(d, d);
        const cplx den{dd.re + wp * wp, dd.im};
        const cplx gg = cdiv(d, den);
        g[e] = make_double2(gg.re, gg.im);
    }
}

// Sigma_c(s, i w_n) = -(1/pi) sum_f sum_k (g(n, f, k) q[f][s][k]) w_f for complex q
// (eta != 0), f outer and k inner for every (s, n) as in the reference's loop;
// 16 n x 8 states per workgroup, g and q staged through LDS in chunks of 64 k.
constexpr int kSgN = 16, kSgS = 8, kSgK = 64;
__global__ __launch_bounds__(128) void k_sigma_c(const double* __restrict__ q,
                                                 const double* __restrict__ qim,
                                                 const double2* __restrict__ g,
                                                 const double* __restrict__ weight, int nmo,
                                                 int nstates, int nfreq, double* __restrict__ sre,
                                                 double* __restrict__ sim) {
#pragma clang fp contract(off)       // the host's complex products round each term
    __shared__ double gre[kSgN][kSgK + 1], gim[kSgN][kSgK + 1];
    __shared__ double qs[kSgS][kSgK + 1], qis[kSgS][kSgK + 1];
    const int t = threadIdx.x, tn = t & (kSgN - 1), ts = t / kSgN;
    const int n0 = blockIdx.x * kSgN, s0 = blockIdx.y * kSgS;
    const size_t m = (size_t)nstates * nmo;
    double are = 0.0, aim = 0.0;
    for (int f = 0; f < nfreq; ++f) {
        const double wf = weight[f];
        for (int k0 = 0; k0 < nmo; k0 += kSgK) {
            const int kc = min(kSgK, nmo - k0);
            __syncthreads();
            for (int e = t; e < kSgN * kSgK; e += 128) {
                const int r = e / kSgK, kk = e % kSgK;
                double2 v = make_double2(0.0, 0.0);
                if (kk < kc && n0 + r < nfreq) v = g[((size_t)(n0 + r) * nfreq + f) * nmo + k0 + kk];
                gre[r][kk] = v.x;
                gim[r][kk] = v.y;
            }
            for (int e = t; e < kSgS * kSgK; e += 128) {
                const int r = e / kSgK, kk = e % kSgK;
                double v = 0.0, vi = 0.0;
                if (kk < kc && s0 + r < nstates) {
                    const size_t o = (size_t)f * m + (size_t)(s0 + r) * nmo + k0 + kk;
                    v = q[o];
                    if (qim) vi = qim[o];
                }
                qs[r][kk] = v;
                qis[r][kk] = vi;
            }
            __syncthreads();
            if (qim) {      // complex q (eta != 0): (g q) w
                for (int kk = 0; kk < kc; ++kk) {
                    const double gr = gre[tn][kk], gi = gim[tn][kk], qk = qs[ts][kk], qi = qis[ts][kk];
                    are -= (gr * qk - gi * qi) * wf;
                    aim -= (gr * qi + gi * qk) * wf;
                }
            } else {
                for (int kk = 0; kk < kc; ++kk) {
                    const double gr = gre[tn][kk], gi = gim[tn][kk], qk = qs[ts][kk];
                    are -= (gr * qk) * wf;
                    aim -= (gi * qk) * wf;
                }
            }
        }
    }
    const int n = n0 + tn, s = s0 + ts;
    if (n >= nfreq || s >= nstates) return;
    const double pi = 3.141592653589793238462643383279502884;
    sre[(size_t)s * nfreq + n] = are / pi;
    sim[(size_t)s * nfreq + n] = aim / pi;
}

// Sigma_c for real q (eta = 0) on FP64 MFMA: the sum over (f, k) of
// w_f q[f][s][k] g(n, f, k) is a product of the states x (f, k) matrix
// w_f q with the (f, k) x (Re g, Im g) matrix.  One workgroup: 64 states x
// 32 frequencies n (columns c < 32: Re g(n0 + c), c >= 32: Im g(n0 + c - 32))
// over the kSgF frequencies f of split blockIdx.y, K staged through LDS in
// slabs of kBK values of k (zero beyond nmo); the splits' partial sums go to
// part[split][s][2 n + {0, 1}] and k_sigma_reduce adds them in split order.
constexpr int kSgF = 8;
__global__ __launch_bounds__(256) void k_sigma_mm(const double* __restrict__ q,
                                                  const double2* __restrict__ g,
                                                  const double* __restrict__ weight, int nmo,
                                                  int nstates, int nfreq, double* __restrict__ part) {
    __shared__ double As[kBK][64 + kLdsPad];
    __shared__ double Bs[kBK][64 + kLdsPad];
    const int t = threadIdx.x;
    const int tiles_s = (nstates + 63) / 64;
    const int s0 = (blockIdx.x % tiles_s) * 64, n0 = (blockIdx.x / tiles_s) * 32;
    const int f0 = blockIdx.y * kSgF, f1 = min(nfreq, f0 + kSgF);
    const size_t m = (size_t)nstates * nmo;
    const int am = t >> 2, akq = (t & 3) * 4;          // A: state s0 + am, k slab offsets akq..+3
    const int bj = t >> 3, bkq = (t & 7) * 2;          // B: frequency n0 + bj, k slab offsets bkq..+1
    Tile64 tile;
    tile.zero();
    for (int f = f0; f < f1; ++f) {
        const double wf = weight[f];
        for (int k0 = 0; k0 < nmo; k0 += kBK) {
            __syncthreads();                           // the previous slab's reads are done
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int k = k0 + akq + u;
                As[akq + u][am] = s0 + am < nstates && k < nmo
                                      ? wf * q[(size_t)f * m + (size_t)(s0 + am) * nmo + k]
                                      : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int k = k0 + bkq + u;
                double2 v = make_double2(0.0, 0.0);
                if (n0 + bj < nfreq && k < nmo) v = g[((size_t)(n0 + bj) * nfreq + f) * nmo + k];
                Bs[bkq + u][bj] = v.x;
                Bs[bkq + u][32 + bj] = v.y;
            }
            __syncthreads();
            tile.step(As, Bs);
        }
    }
    double* out = part + (size_t)blockIdx.y * nstates * 2 * nfreq;
    tile.for_each([&](int r, int c, double v) {
        const int s = s0 + r, n = n0 + (c & 31);
        if (s < nstates && n < nfreq) out[(size_t)s * 2 * nfreq + 2 * n + (c >> 5)] = v;
    });
}

__global__ void k_sigma_reduce(const double* __restrict__ part, int nsplit, int nstates, int nfreq,
                               double* __restrict__ sre, double* __restrict__ sim) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nstates * nfreq) return;
    const int s = e / nfreq, n = e % nfreq;
    double re = 0.0, im = 0.0;
    for (int y = 0; y < nsplit; ++y) {
        const double* p = part + ((size_t)y * nstates + s) * 2 * nfreq + 2 * n;
        re += p[0];
        im += p[1];
    }
    const double pi = 3.141592653589793238462643383279502884;
    sre[e] = -re / pi;
    sim[e] = -im / pi;
}

// sx[p] = -sum_{i < nocc} (pi|pi), (pi|pi) at src[p sp + i si]
__global__ void k_sigma_x(const double* __restrict__ src, double* __restrict__ sx, int nmo,
                          int nocc, int64_t sp, int64_t si) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nmo) return;
    double v = 0.0;
    for (int i = 0; i < nocc; ++i) v -= src[(size_t)p * sp + (size_t)i * si];
    sx[p] = v;
}

// Thiele continued-fraction Pade (src/gw/pade.cpp:10-70): coefficients of
// one state's Sigma_c row from min(npar, nfreq) sampled points (0, s, 2s, ...,
// nfreq-1), reciprocal differences with the 1e-20 guard, in the table g (as
// many entries as points).  Returns the number of points.
__device__ int pade_fit(const double* __restrict__ row_re, const double* __restrict__ row_im,
                        const double* __restrict__ omega, int nfreq, int npar, cplx* xs,
                        cplx* coef, cplx* g) {
#pragma clang fp contract(off)
    const int npa = npar < nfreq ? npar : nfreq;
    const int step = npa > 1 ? nfreq / (npa - 1) : 0;
    for (int i = 0; i < npa; ++i) {
        int idx;
        if (npa == 1) idx = 0;
        else if (step == 0) idx = i;
        else idx = (i < npa - 1) ? i * step : nfreq - 1;
        xs[i] = cplx{0.0, omega[idx]};
        g[i] = cplx{row_re[idx], row_im[idx]};
    }
    coef[0] = g[0];
    for (int ip = 1; ip < npa; ++ip) {
        const cplx piv = g[ip - 1];
        for (int id = ip; id < npa; ++id) {
            const cplx num{piv.re - g[id].re, piv.im - g[id].im};
            const cplx dx{xs[id].re - xs[ip - 1].re, xs[id].im - xs[ip - 1].im};
            const cplx den = cmul(dx, g[id]);
            g[id] = (cabs_(den) < 1e-20) ? cdiv(num, cplx{1e-20, 0.0}) : cdiv(num, den);
        }
        coef[ip] = g[ip];
    }
    return npa;
}

// Backward evaluation of the continued fraction at a real z
// (src/gw/pade.cpp:72-101), 1e-20 guards.
__device__ cplx pade_eval(const cplx* coef, const cplx* xs, int npa, double z) {
#pragma clang fp contract(off)
    if (npa == 1) return coef[0];
    cplx val{1.0, 0.0};
    for (int k = npa - 2; k >= 0; --k) {
        const cplx num = cmul(coef[k + 1], cplx{z - xs[k].re, 0.0 - xs[k].im});
        if (cabs_(val) < 1e-20) val = cplx{1e-20, 0.0};
        const cplx qv = cdiv(num, val);
        val = cplx{1.0 + qv.re, 0.0 + qv.im};
    }
    return (cabs_(val) < 1e-20) ? cdiv(coef[0], cplx{1e-20, 0.0}) : cdiv(coef[0], val);
}

// Pade + QP fixed point (src/gw/qp_solver.cpp:10-47), one thread per state.
// Up to kMaxPade points the tables live in the thread's private arrays; with
// more (kGlobal) in `work`, 3 min(npar, nfreq) complex entries per state.
template <bool kGlobal>
__global__ void k_pade_qp(const double* __restrict__ sre, const double* __restrict__ sim,
                          const double* __restrict__ omega, const double* __restrict__ mo,
                          const double* __restrict__ sx, const double* __restrict__ vxc,
                          const int* __restrict__ states, double fermi, int nstates, int nfreq,
                          int npar, double* __restrict__ qp, int* __restrict__ iters,
                          double2* __restrict__ work) {
#pragma clang fp contract(off)
    const int si = blockIdx.x * blockDim.x + threadIdx.x;
    if (si >= nstates) return;
    const int state = states[si];
    constexpr int kLocal = kGlobal ? 1 : kMaxPade;
    cplx xs_l[kLocal], coef_l[kLocal], g_l[kLocal];
    const int nt = npar < nfreq ? npar : nfreq;
    cplx* xs = kGlobal ? reinterpret_cast<cplx*>(work) + (size_t)si * 3 * nt : xs_l;
    cplx* coef = kGlobal ? xs + nt : coef_l;
    cplx* g = kGlobal ? coef + nt : g_l;
    const int npa = pade_fit(sre + (size_t)si * nfreq, sim + (size_t)si * nfreq, omega, nfreq, npar,
                             xs, coef, g);
    const double e0 = mo[state];
    double cur = e0;
    int used = 200;
    for (int it = 0; it < 200; ++it) {
        const cplx res = pade_eval(coef, xs, npa, cur - fermi);
        const double nq = e0 + res.re + sx[state] - vxc[state];
        if (fabs(nq - cur) < 1e-6) {
            cur = nq;
            used = it + 1;
            break;
        }
        cur = nq;
    }
    qp[state] = cur;
    iters[state] = used;
}

// Real-axis continuation of one state's Sigma_c (the values of E_c.out,
// src/io.cpp:94-108): point i at e_i = x_start + i * interval - fermi.
__global__ void k_pade_real_axis(const double* __restrict__ sre, const double* __restrict__ sim,
                                 const double* __restrict__ omega, int nfreq, int npar,
                                 double fermi, double x_start, double interval, int npoints,
                                 double* __restrict__ out_re, double* __restrict__ out_im) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoints) return;
    cplx xs[kMaxPade], coef[kMaxPade], g[kMaxPade];
    const int npa = pade_fit(sre, sim, omega, nfreq, npar, xs, coef, g);
    const double e = x_start + static_cast<double>(i) * interval - fermi;
    const cplx v = pade_eval(coef, xs, npa, e);
    out_re[i] = v.re;
    out_im[i] = v.im;
}

// The same continuation with more than kMaxPade points: one thread fits the
// fraction into `work` (xs, coef, g: 3 min(npar, nfreq) entries) ...
__global__ void k_pade_fit_one(const double* __restrict__ sre, const double* __restrict__ sim,
                               const double* __restrict__ omega, int nfreq, int npar,
                               double2* __restrict__ work) {
    const int nt = npar < nfreq ? npar : nfreq;
    cplx* xs = reinterpret_cast<cplx*>(work);
    if (blockIdx.x == 0 && threadIdx.x == 0) pade_fit(sre, sim, omega, nfreq, npar, xs, xs + nt, xs + 2 * nt);
}

// ... and every point evaluates it from there.
__global__ void k_pade_eval_axis(const double2* __restrict__ work, int nfreq, int npar, double fermi,
                                 double x_start, double interval, int npoints,
                                 double* __restrict__ out_re, double* __restrict__ out_im) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoints) return;
    const int nt = npar < nfreq ? npar : nfreq;
    const cplx* xs = reinterpret_cast<const cplx*>(work);
    const double e = x_start + static_cast<double>(i) * interval - fermi;
    const cplx v = pade_eval(xs + nt, xs, nt, e);
    out_re[i] = v.re;
    out_im[i] = v.im;
}

// ERI from its factor: eri[P][R] = sum_K B[P][K] B[R][K] for P, R < N = nmo^2,
// B padded to Np x Kp (rows and K zero filled, Np % 64 == 0, Kp % 16 == 0).
// Only tiles with I <= J are computed; each writes its block and the mirror
// block, so eri[P][R] == eri[R][P] bit for bit.
__global__ __launch_bounds__(256) void k_eri_from_factor(const double* __restrict__ b, int N,
                                                         int Kp, double* __restrict__ eri) {
    __shared__ double As[kBK][64 + kLdsPad];
    __shared__ double Bs[kBK][64 + kLdsPad];
    int I, J;
    {
        // upper-triangular tile index t -> (I, J), I <= J, row-major over I
        const int nt = (N + 63) / 64;
        int t = blockIdx.x, i = 0;
        while (t >= nt - i) {
            t -= nt - i;
            ++i;
        }
        I = i;
        J = i + t;
    }
    Tile64 tile;
    tile.zero();
    gemm64<false>(b + (size_t)I * 64 * Kp, Kp, b + (size_t)J * 64 * Kp, Kp, 0, Kp, tile, As, Bs);
    tile.for_each([&](int r, int c, double v) {
        const int P = I * 64 + r, R = J * 64 + c;
        if (P < N && R < N) {
            eri[(size_t)P * N + R] = v;
            eri[(size_t)R * N + P] = v;
        }
    });
}

// Rows P in [P0, P0 + rows) of the same product, every column R < N:
// out[(P - P0) N + R].  Tile (I, J) takes rows P0 + 64 I.. against rows 64 J..
// of B with the K order of k_eri_from_factor; since each term B[P][K] B[R][K]
// is the same number either way round, every element equals that kernel's
// bit for bit.  B needs 64 readable rows past P0 + rows (zero padding).
__global__ __launch_bounds__(256) void k_eri_rows_from_factor(const double* __restrict__ b, int N,
                                                              int Kp, int P0, int rows,
                                                              double* __restrict__ out) {
    __shared__ double As[kBK][64 + kLdsPad];
    __shared__ double Bs[kBK][64 + kLdsPad];
    const int I = blockIdx.y, J = blockIdx.x;
    Tile64 tile;
    tile.zero();
    gemm64<false>(b + (size_t)(P0 + I * 64) * Kp, Kp, b + (size_t)J * 64 * Kp, Kp, 0, Kp, tile, As,
                  Bs);
    tile.for_each([&](int r, int c, double v) {
        const int P = I * 64 + r, R = J * 64 + c;
        if (P < rows && R < N) out[(size_t)P * N + R] = v;
    });
}

// Diagnostic: C = A B^T for 64 x 64 row-major A, B through the Tile64 path.
__global__ __launch_bounds__(256) void k_tile_selftest(const double* __restrict__ a,
                                                       const double* __restrict__ b,
                                                       double* __restrict__ c) {
    __shared__ double As[kBK][64 + kLdsPad];
    __shared__ double Bs[kBK][64 + kLdsPad];
    Tile64 tile;
    tile.zero();
    for (int k0 = 0; k0 < 64; k0 += kBK) {
        load_rows_kmajor(As, a, 64, 0, k0);
        load_rows_kmajor(Bs, b, 64, 0, k0);
        __syncthreads();
        tile.step(As, Bs);
        __syncthreads();
    }
    tile.for_each([&](int r, int col, double v) { c[r * 64 + col] = v; });
}

// Diagnostic: one v_mfma_f64_16x16x4 per wave with per-lane inputs a[t][lane],
// b[t][lane]; raw results d[t][lane][r] (lane layout check).
__global__ __launch_bounds__(64) void k_probe_f64(const double* __restrict__ a,
                                                  const double* __restrict__ b,
                                                  double* __restrict__ d) {
    const int t = blockIdx.x, lane = threadIdx.x;
    d4 acc = d4{0.0, 0.0, 0.0, 0.0};
    acc = mfma_f64(a[t * 64 + lane], b[t * 64 + lane], acc);
    for (int r = 0; r < 4; ++r) d[((size_t)t * 64 + lane) * 4 + r] = acc[r];
}

__global__ void k_set_int(int* __restrict__ p, int v) { *p = v; }

inline int grid_for(size_t total, int block = 256, int cap = 65535 * 8) {
    size_t g = (total + block - 1) / block;
    if (g < 1) g = 1;
    if (g > (size_t)cap) g = cap;
    return (int)g;
}

}  // namespace

hipError_t launch_probe_f64(const double* a, const double* b, double* d, int ntests,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_probe_f64, dim3(ntests), dim3(64), 0, s, a, b, d);
    return hipGetLastError();
}

hipError_t launch_eri_from_factor(const double* b, int N, int Kp, double* eri, hipStream_t s) {
    const int nt = (N + 63) / 64;
    hipLaunchKernelGGL(k_eri_from_factor, dim3(nt * (nt + 1) / 2), dim3(256), 0, s, b, N, Kp, eri);
    return hipGetLastError();
}

hipError_t launch_eri_rows_from_factor(const double* b, int N, int Kp, int P0, int rows,
                                      double* out, hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_eri_rows_from_factor, dim3((N + 63) / 64, (rows + 63) / 64), dim3(256), 0,
                       s, b, N, Kp, P0, rows, out);
    return hipGetLastError();
}

hipError_t launch_gather_ph(const double* slab, int p0, int rows, int nmo, int nocc, double* eph,
                            double* sxd, hipStream_t s) {
    const size_t total = (size_t)rows * nmo * nocc * (nmo - nocc);
    const int g = std::max(grid_for(total), (rows * nocc + 255) / 256);
    hipLaunchKernelGGL(k_gather_ph, dim3(g), dim3(256), 0, s, slab, p0, rows, nmo, nocc, eph, sxd);
    return hipGetLastError();
}

hipError_t launch_set_int(int* p, int v, hipStream_t s) {
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, p, v);
    return hipGetLastError();
}

hipError_t launch_tile_selftest(const double* a, const double* b, double* c, hipStream_t s) {
    hipLaunchKernelGGL(k_tile_selftest, dim3(1), dim3(256), 0, s, a, b, c);
    return hipGetLastError();
}

hipError_t launch_build_v(const double* eri, const int64_t* rowbase, const int* colidx, double* v,
                          int n, int np, hipStream_t s) {
    hipLaunchKernelGGL(k_build_v, dim3(grid_for((size_t)np * np)), dim3(256), 0, s, eri, rowbase,
                       colidx, v, n, np);
    return hipGetLastError();
}

hipError_t launch_pad_square(const double* src, double* dst, int n, int np, hipStream_t s) {
    hipLaunchKernelGGL(k_pad_square, dim3(grid_for((size_t)np * np)), dim3(256), 0, s, src, dst,
                       n, np);
    return hipGetLastError();
}

hipError_t launch_scale_from_grid(const double* delta, const double* omega, double* S, int n,
                                  int np, const int* f0p, int nb, hipStream_t s) {
    hipLaunchKernelGGL(k_scale_grid, dim3(grid_for((size_t)np * nb)), dim3(256), 0, s, delta,
                       omega, S, n, np, f0p, nb);
    return hipGetLastError();
}

hipError_t launch_scale_from_pi0(const double* pi0, double* S, int n, int np, hipStream_t s) {
    hipLaunchKernelGGL(k_scale_pi0, dim3(grid_for((size_t)np)), dim3(256), 0, s, pi0, S, n, np);
    return hipGetLastError();
}

hipError_t launch_dielectric(const double* v, const double* S, double* A, int np, int nb,
                             hipStream_t s) {
    const int gx = grid_for((size_t)np * np, 256, 2048);
    hipLaunchKernelGGL(k_dielectric, dim3(gx, nb), dim3(256), 0, s, v, S, A, np);
    return hipGetLastError();
}

hipError_t launch_potrf(double* A, double* dinv, int* info, int np, int nb, hipStream_t s) {
    const int nblk = np / kNB;
    for (int k = 0; k < nblk; ++k) {
        if (k > 0)
            hipLaunchKernelGGL(k_potrf_update, dim3(nblk - k, nb), dim3(256), 0, s, A, np, k);
        hipLaunchKernelGGL(k_potf2, dim3(nb), dim3(256), 0, s, A, dinv, info, np, k, 0);
        const int t = nblk - k - 1;
        if (t > 0) hipLaunchKernelGGL(k_trsm_panel, dim3(t, nb), dim3(256), 0, s, A, dinv, np, k);
    }
    return hipGetLastError();
}

hipError_t launch_potf2_block(double* A, double* dinv, int* info, int np, int k, int nb,
                              hipStream_t s, int pre) {
    hipLaunchKernelGGL(k_potf2, dim3(nb), dim3(256), 0, s, A, dinv, info, np, k, pre);
    return hipGetLastError();
}

hipError_t launch_trsm_block(double* A, const double* dinv, int np, int k, int nb, hipStream_t s) {
    const int t = np / kNB - k - 1;
    if (t > 0) hipLaunchKernelGGL(k_trsm_panel, dim3(t, nb), dim3(256), 0, s, A, dinv, np, k);
    return hipGetLastError();
}

hipError_t launch_trtri_block(double* A, const double* dinv, const double* tbuf, int np, int j,
                              int nb, hipStream_t s) {
    const int t = np / kNB - 1 - j;
    hipLaunchKernelGGL(k_trtri_m, dim3(t + 1, nb), dim3(256), 0, s, A, dinv, tbuf, np, j);
    return hipGetLastError();
}

hipError_t launch_trtri(double* A, const double* dinv, double* tbuf, int np, int nb,
                        hipStream_t s) {
    const int nblk = np / kNB;
    for (int j = nblk - 1; j >= 0; --j) {
        const int t = nblk - 1 - j;
        if (t > 0) hipLaunchKernelGGL(k_trtri_t, dim3(t, nb), dim3(256), 0, s, A, tbuf, np, j);
        hipLaunchKernelGGL(k_trtri_m, dim3(t + 1, nb), dim3(256), 0, s, A, dinv, tbuf, np, j);
    }
    return hipGetLastError();
}

hipError_t launch_pack_x(const double* src, const int64_t* pkrow, const int* colidx, double* xp,
                         int n, int np, int mpad, hipStream_t s) {
    hipLaunchKernelGGL(k_pack_x, dim3((mpad + 3) / 4), dim3(256), 0, s, src, pkrow, colidx, xp, n,
                       np, mpad);
    return hipGetLastError();
}

hipError_t launch_contract_f64(const double* M, const double* S, const double* xp, double* qpart,
                               int np, int mpad, int nb, hipStream_t s) {
    hipLaunchKernelGGL(k_contract_f64, dim3(mpad / kCT, np / kCT, nb), dim3(256), 0, s, M, S, xp,
                       qpart, np, mpad);
    return hipGetLastError();
}

hipError_t launch_pk_rows(int64_t* pkrow, int p, int k0, int width, int nmo, int mpad, int64_t rs,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_pk_rows, dim3((mpad + 255) / 256), dim3(256), 0, s, pkrow, p, k0, width,
                       nmo, mpad, rs);
    return hipGetLastError();
}

hipError_t launch_reduce_q(const double* qpart, const int* colmap, double* q, int nrt, int m,
                           int mpad, const int* f0p, int nb, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_q, dim3(grid_for((size_t)m, 256, 4096), nb), dim3(256), 0, s,
                       qpart, colmap, q, nrt, m, mpad, f0p);
    return hipGetLastError();
}

hipError_t launch_check_pk_symmetry(const double* eri, int nmo, int nocc, int64_t rs, int cs,
                                    int co, int* flag, hipStream_t s) {
    const size_t total = (size_t)nmo * nmo * nocc * (nmo - nocc);
    hipLaunchKernelGGL(k_check_pk_symmetry, dim3(grid_for(total, 256, 65536)), dim3(256), 0, s, eri,
                       nmo, nocc, rs, cs, co, flag);
    return hipGetLastError();
}

hipError_t launch_wc_explicit(const double* M, const double* S, double* wc, int n, int np,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_wc_explicit, dim3(grid_for((size_t)n * n)), dim3(256), 0, s, M, S, wc, n,
                       np);
    return hipGetLastError();
}

size_t sigma_workspace_bytes(int nmo, int nstates, int nfreq) {
    const size_t nsplit = (size_t)(nfreq + kSgF - 1) / kSgF;
    return (size_t)nfreq * nfreq * nmo * sizeof(double2) + nsplit * nstates * 2 * nfreq * sizeof(double);
}

hipError_t launch_sigma_c(const double* q, const double* q_im, const double* mo_energy,
                          const double* omega, const double* weight, double fermi, int nmo,
                          int nstates, int nfreq, double2* gbuf, double* sig_re, double* sig_im,
                          hipStream_t s) {
    if (nmo <= 0 || nstates <= 0 || nfreq <= 0) return hipErrorInvalidValue;
    const size_t total = (size_t)nfreq * nfreq * nmo;
    hipLaunchKernelGGL(k_sigma_g, dim3(grid_for(total)), dim3(256), 0, s, mo_energy, omega, fermi,
                       nmo, nfreq, gbuf);
    if (q_im) {      // complex q (eta != 0)
        hipLaunchKernelGGL(k_sigma_c, dim3((nfreq + kSgN - 1) / kSgN, (nstates + kSgS - 1) / kSgS),
                           dim3(128), 0, s, q, q_im, gbuf, weight, nmo, nstates, nfreq, sig_re,
                           sig_im);
        return hipGetLastError();
    }
    // partial sums after the quadrature factors in the same workspace
    double* part = reinterpret_cast<double*>(gbuf + total);
    const int nsplit = (nfreq + kSgF - 1) / kSgF;
    const int tiles = ((nstates + 63) / 64) * ((nfreq + 31) / 32);
    hipLaunchKernelGGL(k_sigma_mm, dim3(tiles, nsplit), dim3(256), 0, s, q, gbuf, weight, nmo,
                       nstates, nfreq, part);
    hipLaunchKernelGGL(k_sigma_reduce, dim3((nstates * nfreq + 255) / 256), dim3(256), 0, s, part,
                       nsplit, nstates, nfreq, sig_re, sig_im);
    return hipGetLastError();
}

hipError_t launch_sigma_x(const double* src, double* sx, int nmo, int nocc, int64_t sp, int64_t si,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_sigma_x, dim3((nmo + 63) / 64), dim3(64), 0, s, src, sx, nmo, nocc, sp, si);
    return hipGetLastError();
}

size_t pade_work_bytes(int nfreq, int npar, int nstates) {
    const int nt = npar < nfreq ? npar : nfreq;
    return nt <= kMaxPade ? 0 : (size_t)nstates * 3 * nt * sizeof(double2);
}

hipError_t launch_pade_real_axis(const double* sig_re, const double* sig_im, const double* omega,
                                 int nfreq, int npar, double fermi, double x_start, double interval,
                                 int npoints, double* out_re, double* out_im, double* work,
                                 hipStream_t s) {
    if (pade_work_bytes(nfreq, npar, 1) == 0) {
        hipLaunchKernelGGL(k_pade_real_axis, dim3((npoints + 127) / 128), dim3(128), 0, s, sig_re,
                           sig_im, omega, nfreq, npar, fermi, x_start, interval, npoints, out_re,
                           out_im);
        return hipGetLastError();
    }
    if (!work) return hipErrorInvalidValue;
    double2* w = reinterpret_cast<double2*>(work);
    hipLaunchKernelGGL(k_pade_fit_one, dim3(1), dim3(64), 0, s, sig_re, sig_im, omega, nfreq, npar, w);
    hipLaunchKernelGGL(k_pade_eval_axis, dim3((npoints + 127) / 128), dim3(128), 0, s,
                       static_cast<const double2*>(w), nfreq, npar, fermi, x_start, interval, npoints,
                       out_re, out_im);
    return hipGetLastError();
}

hipError_t launch_pade_qp(const double* sig_re, const double* sig_im, const double* omega,
                          const double* mo_energy, const double* sx, const double* vxc_diag,
                          const int* states, double fermi, int nstates, int nfreq, int npar,
                          double* qp, int* iters, double* work, hipStream_t s) {
    const bool global = pade_work_bytes(nfreq, npar, nstates) > 0;
    if (global && !work) return hipErrorInvalidValue;
    hipLaunchKernelGGL(global ? k_pade_qp<true> : k_pade_qp<false>, dim3((nstates + 63) / 64),
                       dim3(64), 0, s, sig_re, sig_im, omega, mo_energy, sx, vxc_diag, states, fermi,
                       nstates, nfreq, npar, qp, iters, reinterpret_cast<double2*>(work));
    return hipGetLastError();
}

}  // namespace mgw
