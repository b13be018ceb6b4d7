// Device kernels of the screening + Sigma_c path and their launchers.
// All launchers enqueue on the given stream and return hipGetLastError().
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace mgw {

constexpr int kNB = 64;    // factorization block (potrf / trtri tiles)
constexpr int kCT = 128;   // FP64 contraction tile: rows of M and columns of X per workgroup
constexpr int kPad = 256;  // np and the contraction column count are padded to this
constexpr int kMaxPade = 64;
// Digit schemes of the integer-MFMA kernels (mgw_slices.h): five signed 7-bit
// digits per operand (15 digit products, ~2^-35 of a row's largest element)
// or four balanced 8-bit digits (10 products, ~2^-32).
enum I8Digits : int { kDigits5x7 = 0, kDigits4x8 = 1 };
inline int i8_digits(int scheme) { return scheme == kDigits4x8 ? 4 : 5; }
// largest np of the integer contraction: keeps every int32 group sum exact
// (5 x 2^12 x K < 2^31 for 7-bit digits, 4 x 2^14 x K < 2^31 for 8-bit ones)
inline int i8_max_k(int scheme) { return scheme == kDigits4x8 ? 32512 : 65536; }

inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

// Diagnostic: C = A B^T (64 x 64 row-major) through the factorization's MFMA tile.
hipError_t launch_tile_selftest(const double* a, const double* b, double* c, hipStream_t s);

// Diagnostics: raw MFMA results for per-lane inputs (lane layout checks).
hipError_t launch_probe_f64(const double* a, const double* b, double* d, int ntests,
                           hipStream_t s);

// eri[P][R] = sum_K B[P][K] B[R][K] (P, R < N) from B padded to ((N+63)/64*64) x Kp,
// Kp % 16 == 0, extra rows and columns zero; exactly symmetric in (P, R).
hipError_t launch_eri_from_factor(const double* b, int N, int Kp, double* eri, hipStream_t s);
// Rows P0 .. P0 + rows - 1 of the same product, all N columns: out[(P - P0) N + R],
// bit for bit the values launch_eri_from_factor writes; B needs 64 zero rows
// past P0 + rows.
hipError_t launch_eri_rows_from_factor(const double* b, int N, int Kp, int P0, int rows,
                                      double* out, hipStream_t s);
// A slab of ERI rows p in [p0, p0 + rows) (nmo^3 values per p, C order) into
// the held ph block eph[(p nmo + k) n + i nvir + a] = (pk|ia) (n = nocc nvir)
// and the exchange diagonal sxd[p nocc + i] = (pi|pi).
hipError_t launch_gather_ph(const double* slab, int p0, int rows, int nmo, int nocc, double* eph,
                            double* sxd, hipStream_t s);

// V[r][c] = eri[rowbase[r] + colidx[c]] for r, c < n; 0 in the padding (np x np).
hipError_t launch_build_v(const double* eri, const int64_t* rowbase, const int* colidx,
                          double* v, int n, int np, hipStream_t s);

// Pad a host-provided n x n row-major matrix into np x np (zeros outside).
hipError_t launch_pad_square(const double* src, double* dst, int n, int np, hipStream_t s);

// S[w][r] = sqrt(4 D_r / (w_f^2 + D_r^2)), f = *f0p + w, for r < n; 0 for r >= n.
// The frequency offset is read from device memory so that a captured batch
// graph serves every batch.
hipError_t launch_scale_from_grid(const double* delta, const double* omega, double* S, int n,
                                  int np, const int* f0p, int nb, hipStream_t s);

// *p = v (one thread); sets the device-side batch offset outside a graph.
hipError_t launch_set_int(int* p, int v, hipStream_t s);

// S[r] = sqrt(-pi0[r]) for r < n; 0 in the padding.
hipError_t launch_scale_from_pi0(const double* pi0, double* S, int n, int np, hipStream_t s);

// Symmetrised dielectric, lower triangle: A_w = I + diag(S_w) V diag(S_w).
hipError_t launch_dielectric(const double* v, const double* S, double* A, int np, int nb,
                             hipStream_t s);

// In-place blocked Cholesky A_w = L L^T (lower) of nb matrices np x np, plus the
// inverses of its 64 x 64 diagonal blocks in dinv (nb x (np/64) x 64 x 64).
// info[w] = 1-based global row of the first non-positive pivot, 0 if none.
hipError_t launch_potrf(double* A, double* dinv, int* info, int np, int nb, hipStream_t s);

// Single steps of the blocked factorization (block column k / j): the
// diagonal block's factor and inverse (with pre != 0 after the update
// A_kk -= L_k,k-1 L_k,k-1^T), the panel solve below it, and the inverse
// factor's column j from T (tbuf) -- shared by the FP64 and the int8-update
// forms.
hipError_t launch_potf2_block(double* A, double* dinv, int* info, int np, int k, int nb,
                              hipStream_t s, int pre = 0);
hipError_t launch_trsm_block(double* A, const double* dinv, int np, int k, int nb, hipStream_t s);
hipError_t launch_trtri_block(double* A, const double* dinv, const double* tbuf, int np, int j,
                              int nb, hipStream_t s);

// Pivoted-LU screening (mgw_lu.hip) for a symmetrised dielectric matrix that
// is not positive definite: E = I + S V S in full, E = P L U in place
// (piv[w][c] = row swapped with row c, info[w] = first zero pivot, dl / du =
// inverses of the diagonal blocks' unit lower / upper factors, nb x (np/64) x
// 64 x 64 each), W = E^-1 into Z from perm (perm[w][r]: row r of P^T I is the
// unit row perm[r]), and the q partials of a dense W (as launch_contract_f64's,
// (S x)^T W (S x) summed per 128-row tile) and W_c = -S W S.
hipError_t launch_dielectric_full(const double* v, const double* S, double* A, int np, int nb,
                                  hipStream_t s);
hipError_t launch_lu_factor(double* A, int* piv, int* info, double* dl, double* du, int np, int nb,
                            hipStream_t s);
hipError_t launch_lu_inverse(const double* A, const double* dl, const double* du, const int* perm,
                             double* Z, int np, int nb, hipStream_t s);
hipError_t launch_contract_dense(const double* W, const double* S, const double* xp, double* qpart,
                                 int np, int mpad, int nb, hipStream_t s);
hipError_t launch_wc_from_dense(const double* W, const double* S, double* wc, int n, int np,
                                hipStream_t s);

// Factorization of the mixed forms (mgw_ozfact.hip): the same blocked
// Cholesky and inverse sweep with the block products on int8 MFMA.  RP/RS and
// CP/CS hold the row- and column-scaled slices of every 64 x 64 block
// (nb x fact_slice_bytes(np, scheme) bytes each: planes first, then
// nb x (np/64)^2 x 64 scale exponents (int32) at the byte offset
// nb x (np/64)^2 x i8_digits(scheme) x 4096).
// Both sweeps update two block columns per launch; launch_trtri_i8's tbuf
// holds two slots of nb x np x 64 doubles.
size_t fact_slice_bytes(int np, int scheme);
hipError_t launch_potrf_i8(double* A, double* dinv, int* info, int8_t* RP, double* RS, int8_t* CP,
                           double* CS, int np, int nb, int scheme, hipStream_t s);
hipError_t launch_trtri_i8(double* A, const double* dinv, double* tbuf, int8_t* RP, double* RS,
                           const int8_t* CP, const double* CS, int np, int nb, int scheme,
                           hipStream_t s);

// In-place inverse of the lower factor: A_w <- L_w^{-1} (lower, zeros above the
// diagonal inside diagonal blocks).  tbuf: nb x np x 64 scratch.
hipError_t launch_trtri(double* A, const double* dinv, double* tbuf, int np, int nb,
                        hipStream_t s);

// xp[c][k] = src[pkrow[c] + colidx[k]] for k < n, 0 for n <= k < np (mpad x np):
// the P panels packed once per molecule, K-contiguous per column.
hipError_t launch_pack_x(const double* src, const int64_t* pkrow, const int* colidx, double* xp,
                         int n, int np, int mpad, hipStream_t s);

// qpart[w][rt][c] = sum_{r in row tile rt} (sum_k M_w[r][k] S_w[k] xp[c][k])^2,
// M_w lower triangular, 128-row tiles (np, mpad multiples of 128); S == NULL
// means S_w = 1.
hipError_t launch_contract_f64(const double* M, const double* S, const double* xp, double* qpart,
                               int np, int mpad, int nb, hipStream_t s);

// nrt = row tiles of the contraction that wrote qpart (np / 128 for FP64,
// contract_split_row_tiles(np) for the split form).
// pkrow[c] = (p nmo + k0 + c) rs for c < width, 0 for width <= c < mpad (rs:
// values per (p, k) row of the held ERI)
hipError_t launch_pk_rows(int64_t* pkrow, int p, int k0, int width, int nmo, int mpad, int64_t rs,
                          hipStream_t s);
hipError_t launch_reduce_q(const double* qpart, const int* colmap, double* q, int nrt, int m,
                           int mpad, const int* f0p, int nb, hipStream_t s);

// flag = 1 if (pk|ia) != (kp|ia) for some p, k, i < nocc <= a (flag must start
// at 0), (pk|ia) at eri[(p nmo + k) rs + i cs + co + a - nocc] (a - nocc: the
// virtual index).
hipError_t launch_check_pk_symmetry(const double* eri, int nmo, int nocc, int64_t rs, int cs,
                                    int co, int* flag, hipStream_t s);

// Mixed precision: 16-bit hi/lo splits of the contraction operands.
//   BF16: x = hi + lo with hi = bf16(x), lo = bf16(x - hi); no scaling.
//   FP16: the same with fp16 after a power-of-two scaling per row of M S and
//         per column of X (largest entry near 2^14), undone in the epilogue.
enum class Split16 : int { BF16 = 0, FP16 = 1 };

// X[k][c] = src[pkrow[c] + colidx[k]] for k < n (0 for k >= n) -> planes
// xhi/xlo (mpad x np halves, blocks of 256 columns x 32 k, each block
// row-major, blocks ordered column tile then k); xscale[c] = applied scale.
hipError_t launch_split_x(Split16 fmt, const double* src, const int64_t* pkrow, const int* colidx,
                          uint16_t* xhi, uint16_t* xlo, double* xscale, int n, int np, int mpad,
                          hipStream_t s);

// (M_w S_w)[r][k] for k <= r (0 above) -> planes [w] (np x np halves in the
// same 256 x 32 block layout); mscale[w][r].
hipError_t launch_split_m(Split16 fmt, const double* M, const double* S, uint16_t* mhi,
                          uint16_t* mlo, double* mscale, int np, int nb, hipStream_t s);

// Same output as launch_contract_f64 but with 256-row tiles (scaled by
// xscale[c]^2 when FP16); needs np and mpad multiples of 256:
// products hi*hi + hi*lo + lo*hi on 16-bit MFMA (v_mfma_f32_32x32x16) with
// FP32 accumulation, row scale undone and squares summed in FP64.  uf (1, 2,
// 4, ..., 32) is the number of frequencies in a unit of 32 tiles that share
// one XCD's L2.
hipError_t launch_contract_split(Split16 fmt, const uint16_t* mhi, const uint16_t* mlo,
                                 const double* mscale, const uint16_t* xhi, const uint16_t* xlo,
                                 double* qpart, int np, int mpad, int nb, int uf, hipStream_t s);

hipError_t launch_probe16(Split16 fmt, const uint16_t* a, const uint16_t* b, float* d,
                         int ntests, hipStream_t s);
hipError_t launch_probe32(Split16 fmt, const uint16_t* a, const uint16_t* b, float* d,
                         int ntests, hipStream_t s);

// q[(*f0p + w) * m + c] = -(sum_rt qpart[w][rt][cu]) / xscale[cu]^2, cu = colmap[c].
hipError_t launch_reduce_q_scaled(const double* qpart, const double* xscale, const int* colmap,
                                  double* q, int nrt, int m, int mpad, const int* f0p, int nb,
                                  hipStream_t s);
int contract_split_row_tiles(int np);

// W_c = -(M S)^T (M S) for one matrix (explicit, n x n row-major output).
hipError_t launch_wc_explicit(const double* M, const double* S, double* wc, int n, int np,
                              hipStream_t s);

// sigma[s][n] = (1/pi) sum_f sum_k -(g0(k, n, f) q[f][s*nmo + k]) w_f; gbuf:
// sigma_workspace_bytes() of scratch: the state-independent g0 factors
// (nfreq^2 * nmo complex) and, for real q, the partial sums of the product
// on FP64 MFMA.
size_t sigma_workspace_bytes(int nmo, int nstates, int nfreq);
hipError_t launch_sigma_c(const double* q, const double* q_im, const double* mo_energy,
                          const double* omega, const double* weight, double fermi, int nmo,
                          int nstates, int nfreq, double2* gbuf, double* sig_re, double* sig_im,
                          hipStream_t s);

// sx[p] = -sum_{i < nocc} (pi|pi), with (pi|pi) at src[p sp + i si]
hipError_t launch_sigma_x(const double* src, double* sx, int nmo, int nocc, int64_t sp, int64_t si,
                          hipStream_t s);

// Pade fit + QP fixed point per state (FP64, no FMA contraction).  With more
// than kMaxPade points (min(npar, nfreq)) the fraction's tables go to `work`,
// pade_work_bytes(nfreq, npar, nstates) bytes (1 state for the real axis);
// otherwise work may be NULL.
size_t pade_work_bytes(int nfreq, int npar, int nstates);
// Thiele Pade of one state's row evaluated at x_start + i * interval - fermi.
hipError_t launch_pade_real_axis(const double* sig_re, const double* sig_im, const double* omega,
                                 int nfreq, int npar, double fermi, double x_start, double interval,
                                 int npoints, double* out_re, double* out_im, double* work,
                                 hipStream_t s);
hipError_t launch_pade_qp(const double* sig_re, const double* sig_im, const double* omega,
                          const double* mo_energy, const double* sx, const double* vxc_diag,
                          const int* states, double fermi, int nstates, int nfreq, int npar,
                          double* qp, int* iters, double* work, hipStream_t s);

// Peak calibration: `blocks` x 256 threads issuing 8 x iters independent MFMAs
// (kind 0 f64 16x16x4, 1 f16 16x16x32, 2 bf16 16x16x32) on registers; and a
// streaming copy of `count` doubles (count % 4 == 0).
hipError_t launch_mfma_peak(int kind, int blocks, int iters, float* sink, hipStream_t s);
hipError_t launch_stream_copy(const double* src, double* dst, size_t count, hipStream_t s);

// Integer-MFMA form of the contraction (i8_digits(scheme) int8 digits per
// operand, exact int32 products, FP64 epilogue).  X columns (as
// launch_split_x) and the rows of M_w S_w are scaled by powers of two
// (xscale, mscale) and split into one plane per digit of 128 x 32 byte
// blocks; launch_contract_i8 writes the same qpart as the other contractions
// with np / 128 row tiles (reduce with launch_reduce_q_scaled).  np and mpad
// multiples of 128, np <= i8_max_k(scheme).
hipError_t launch_slice_x(const double* src, const int64_t* pkrow, const int* colidx,
                          int8_t* planes, double* xscale, int n, int np, int mpad, int scheme,
                          hipStream_t s);
hipError_t launch_slice_m(const double* M, const double* S, int8_t* planes, double* mscale, int np,
                          int nb, int scheme, hipStream_t s);
hipError_t launch_contract_i8(const int8_t* mplanes, const double* mscale, const int8_t* xplanes,
                              double* qpart, int np, int mpad, int nb, int uf, int scheme,
                              hipStream_t s);
int contract_i8_row_tiles(int np);
hipError_t launch_probe_i8(const int8_t* a, const int8_t* b, int* d, int ntests, hipStream_t s);

// Cholesky-decomposed ERI form (mgw_cd.hip).  Pairs u < N are unique orbital
// pairs, pq[u] = p * nmo + q with p <= q.  L is column-major: column j of
// all pairs at Lc + j * Npad, Npad = cd_npad(N).  bm: 2 * cd_blocks(N)
// double2 per-block maxima.  launch_cd_steps runs pivots [k0, k1) and stops
// (rank = the pivot count) once the largest residual diagonal is <= tol times
// the first pivot; rank stays -1 while it has not converged.
int cd_blocks(int N);
int cd_npad(int N);
hipError_t launch_cd_init(const double* eri, const int* pq, int nmo, int N, double* D, double2* bm,
                          int* rank, hipStream_t s);
hipError_t launch_cd_steps(const double* eri, const int* pq, int nmo, int N, double* Lc, double* D,
                           double2* bm, int* piv, double* pval, int* rank, int k0, int k1,
                           double tol, hipStream_t s);
// out (3 x uint64, zeroed by the caller) = bit patterns of max |M[u][v] -
// M[v][u]|, max |(pk|ia) - (kp|ia)| and max |M[u][v]|.
hipError_t launch_cd_asymmetry(const double* eri, const int* pq, int nmo, int nocc, int N,
                               unsigned long long* out, hipStream_t s);
// out[0] = min residual diagonal, out[1] = count of non-finite entries.
hipError_t launch_cd_residual(const double* D, int N, double* out, hipStream_t s);
// Bt (rpad x np): row j = L column j over the ph pairs phu[ph]; Lx (mpad x
// rpad): row c = L over colu[c] for c < mu, and nrm[c] = ||Lx[c]||^2.
hipError_t launch_cd_gather(const double* Lc, int N, const int* phu, const int* colu, int r, int n,
                            int mu, double* Bt, double* Lx, double* nrm, int rpad, int np, int mpad,
                            hipStream_t s);
// Lower 128-tiles of C_w = I + Bt diag(S_w^2) Bt^T (rpad x rpad, row-major, per w).
hipError_t launch_aux_dielectric(const double* Bt, const double* S, double* C, int rpad, int np,
                                 int nb, hipStream_t s);
// q[*f0p + w][c] = sum_rt qpart[w][rt][colmap[c]] - nrm[colmap[c]].
hipError_t launch_reduce_q_cd(const double* qpart, const double* nrm, const int* colmap, double* q,
                              int nrt, int m, int mpad, const int* f0p, int nb, hipStream_t s);

// Complex path for eta != 0 (mgw_complex.hip).  Complex arrays are
// interleaved (re, im) doubles; matrices row-major np x np, panels np x mpad.
// S_w = sqrt(-Pi0(omega[*f0p + w], delta, eta)), 0 in the padding.
hipError_t launch_scale_grid_c(const double* delta, const double* omega, double eta, double* S,
                               int n, int np, const int* f0p, int nb, hipStream_t s);
// S = sqrt(-pi0) for a host-given complex Pi0 (pim may be NULL).
hipError_t launch_scale_pi0_c(const double* pre, const double* pim, double* S, int n, int np,
                              hipStream_t s);
// Lower triangle of E_w = I + S_w V S_w.
hipError_t launch_dielectric_c(const double* V, const double* S, double* E, int np, int nb,
                               hipStream_t s);
// In place E_w = L_w L_w^T (no conjugate, lower); dinv: inverses of the 64 x 64
// diagonal blocks (nb x np/64 x 64 x 64); info[w] = first bad pivot (1-based).
hipError_t launch_potrf_c(double* E, double* dinv, int* info, int np, int nb, hipStream_t s);
// Y_w = diag(S_w) X with X packed per column (xp[c * np + r], real).
hipError_t launch_make_y(const double* S, const double* xp, double* Y, int np, int mpad, int nb,
                         hipStream_t s);
// In place Y_w <- L_w^-1 Y_w (blocked forward substitution).
hipError_t launch_fsub_c(const double* L, const double* dinv, double* Y, int np, int mpad, int nb,
                         hipStream_t s);
// q[*f0p + w][c] = -sum_r Z_w[r][colmap[c]]^2 into (qre, qim).
hipError_t launch_sumsq_c(const double* Z, const int* colmap, double* qre, double* qim, int np,
                          int m, int mpad, const int* f0p, int nb, hipStream_t s);
// W_c = -(L^-1 S)^T (L^-1 S), first n rows / columns into re / im (n x n);
// Y and W: np x np complex scratch.
hipError_t launch_wc_explicit_c(const double* L, const double* dinv, const double* S, double* Y,
                                double* W, double* re, double* im, int n, int np, hipStream_t s);

}  // namespace mgw
