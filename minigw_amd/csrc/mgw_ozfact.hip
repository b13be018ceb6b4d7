// Factorization GEMMs of the mixed forms on int8 MFMA.
//
// The blocked Cholesky's trailing updates (A_ik -= sum_{j<k} L_ij L_kj^T) and
// the inverse factor's sweeps (T_i = sum_{k=j+1}^{i} M_ik L_kj) are products
// of 64 x 64 blocks.  Each finished block is cut once into S int8 digits
// (mgw_slices.h: five 7-bit digits or four balanced 8-bit ones) with a
// power-of-two scale per row (per column for the L_kj operand of the sweeps,
// stored transposed); a product over one 64-wide K block runs as the digit
// products with i + j < S on v_mfma_i32_32x32x32_i8 (exact in int32), and the
// K blocks, whose scales differ, are combined in FP64.  The diagonal blocks
// (factor and inverse in LDS), the panel solves and M_ij = -T_i L_jj^-1 stay
// FP64.  The truncation error per block product is ~2^-35 (5 x 7 bits) or
// ~2^-32 (4 x 8 bits) of the largest element of each row segment; the
// contraction that follows carries the same order.
//
// The sweeps take two block columns per launch; the one block product of the
// second column that the launch cannot include (kb = k for column k + 1 of
// the Cholesky, kb = j for column j - 1 of the inverse) runs in FP64 inside
// the next column's panel kernel, which also cuts its own output blocks into
// slices: per block column one panel launch instead of a product, a solve and
// a slicing launch.
#include "mgw_kernels.h"
#include "mgw_slices.h"
#include "mgw_tile64.h"

#include <hip/hip_runtime.h>

namespace mgw {
namespace {

// digit planes of one 64 x 64 block
template <int kS>
constexpr size_t blk_bytes() { return (size_t)kS * 4096; }

__device__ inline size_t blk_index(int w, int nblk, int bi, int bj) {
    return ((size_t)w * nblk + bi) * nblk + bj;
}

// Slices of the 64 x 64 block held in LDS (blk) at block index bix: row-scaled
// into RP/RS (when RP != NULL) and, transposed, column-scaled into CP/CS
// (when CP != NULL).  RS / CS hold, as int32, the exponents e of the inverse
// scales 2^e (the factor that restores a row from its digits).  256 threads;
// blk is complete (barrier before the call).
template <int kS, int B>
__device__ void slice_block(const double (*blk)[66], double* rsc, double* csc, size_t bix,
                            int8_t* __restrict__ RP, double* __restrict__ RS,
                            int8_t* __restrict__ CP, double* __restrict__ CS) {
    const int t = threadIdx.x;
    {
        // row (column) maxima: four consecutive lanes per row (column), 16 entries each
        const int line = t >> 2, part = (t & 3) * 16;
        double mr = 0.0, mc = 0.0;
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            mr = fmax(mr, fabs(blk[line][part + c]));
            if (CP) mc = fmax(mc, fabs(blk[part + c][line]));
        }
#pragma unroll
        for (int d = 1; d < 4; d <<= 1) {
            mr = fmax(mr, __shfl_xor(mr, d));
            if (CP) mc = fmax(mc, __shfl_xor(mc, d));
        }
        if ((t & 3) == 0) {
            rsc[line] = digit_scale<B>(mr);
            if (CP) csc[line] = digit_scale<B>(mc);
        }
    }
    __syncthreads();
    if (RP) {
        int8_t* out = RP + bix * blk_bytes<kS>();
        for (int e = t; e < 4096; e += 256) {
            const int r = e >> 6, c = e & 63;
            int8_t sv[kS];
            digit_split<kS, B>(blk[r][c] * rsc[r], sv);
#pragma unroll
            for (int i = 0; i < kS; ++i) out[(size_t)i * 4096 + e] = sv[i];
        }
        if (t < 64) reinterpret_cast<int*>(RS)[bix * 64 + t] = ilogb(1.0 / rsc[t]);   // e of 2^e
    }
    if (CP) {
        int8_t* out = CP + bix * blk_bytes<kS>();
        for (int e = t; e < 4096; e += 256) {
            const int c = e >> 6, r = e & 63;       // transposed: row c of the output
            int8_t sv[kS];
            digit_split<kS, B>(blk[r][c] * csc[c], sv);
#pragma unroll
            for (int i = 0; i < kS; ++i) out[(size_t)i * 4096 + e] = sv[i];
        }
        if (t < 64) reinterpret_cast<int*>(CS)[bix * 64 + t] = ilogb(1.0 / csc[t]);
    }
}

// Panel of Cholesky block column k, block i = k + 1 + blockIdx.x of A_w:
// with pre != 0 first A_ik -= L_i,k-1 L_k,k-1^T (FP64, the product the
// two-column sweep left out), then L_ik = A_ik Dinv_k^T written back and cut
// into row slices (RP/RS) and transposed column slices (CP/CS).  Only block
// column k is written; column k - 1 is final.
template <int kS, int B>
__global__ __launch_bounds__(256) void k_trsm_slice(double* __restrict__ A,
                                                    const double* __restrict__ dinv, int np, int k,
                                                    int pre, int8_t* __restrict__ RP,
                                                    double* __restrict__ RS,
                                                    int8_t* __restrict__ CP,
                                                    double* __restrict__ CS) {
    __shared__ double As[kBK][64 + kLdsPad];
    __shared__ double Bs[kBK][64 + kLdsPad];
    __shared__ __attribute__((aligned(16))) double blk[64][66];
    __shared__ double rsc[64], csc[64];
    const int w = blockIdx.y, nblk = np / kNB, i = k + 1 + blockIdx.x;
    double* Aw = A + (size_t)w * np * np;
    double* Aik = Aw + (size_t)i * kNB * np + (size_t)k * kNB;
    Tile64 tile;
    tile.zero();
    if (pre)
        gemm64_k64<false>(Aw + (size_t)i * kNB * np + (size_t)(k - 1) * kNB, np,
                          Aw + (size_t)k * kNB * np + (size_t)(k - 1) * kNB, np, tile, As, Bs);
    tile.for_each([&](int r, int c, double v) { blk[r][c] = Aik[(size_t)r * np + c] - v; });
    __syncthreads();
    tile.zero();
    gemm64_k64<false>(&blk[0][0], 66, dinv + ((size_t)w * nblk + k) * 4096, 64, tile, As, Bs);
    tile.for_each([&](int r, int c, double v) {
        Aik[(size_t)r * np + c] = v;
        blk[r][c] = v;
    });
    __syncthreads();
    slice_block<kS, B>(blk, rsc, csc, blk_index(w, nblk, i, k), RP, RS, CP, CS);
}

// Column j of the inverse factor M = L^-1, block i = j + blockIdx.x of A_w:
// M_jj = Dinv_j, M_ij = -(T_i + [pre] M_i,j+1 L_j+1,j) Dinv_j for i > j,
// written back and, with slice != 0, cut into row slices (RP/RS).  T_i is
// block i of tbuf (this column's slot); the FP64 term is the product the
// two-column sweep left out, its L_j+1,j read from the slot's block j (T is
// only defined below the diagonal), where the previous column's launch saved
// it (save != 0: block (j, j-1) of L into block j-1 of slot `next`) before
// this column's blocks overwrote it.
template <int kS, int B>
__global__ __launch_bounds__(256) void k_trtri_m_slice(double* __restrict__ A,
                                                       const double* __restrict__ dinv,
                                                       const double* __restrict__ tbuf,
                                                       double* __restrict__ next, int np, int j,
                                                       int pre, int save, int slice,
                                                       int8_t* __restrict__ RP,
                                                       double* __restrict__ RS) {
    __shared__ double As[kBK][64 + kLdsPad];
    __shared__ double Bs[kBK][64 + kLdsPad];
    __shared__ __attribute__((aligned(16))) double blk[64][66];
    __shared__ double rsc[64];
    const int w = blockIdx.y, nblk = np / kNB, i = j + blockIdx.x, t = threadIdx.x;
    double* Aw = A + (size_t)w * np * np;
    const double* D = dinv + ((size_t)w * nblk + j) * 4096;
    double* Aij = Aw + (size_t)i * kNB * np + (size_t)j * kNB;
    const double* Tw = tbuf + (size_t)w * np * kNB;
    if (blockIdx.x == 0) {
        if (save) {                 // L_j,j-1 for the next column's launch
            const double* src = Aw + (size_t)j * kNB * np + (size_t)(j - 1) * kNB;
            double* dst = next + (size_t)w * np * kNB + (size_t)(j - 1) * 4096;
            for (int e = t; e < 4096; e += 256) dst[e] = src[(size_t)(e >> 6) * np + (e & 63)];
        }
        for (int e = t; e < 4096; e += 256) {
            const double v = D[e];
            Aij[(size_t)(e >> 6) * np + (e & 63)] = v;
            blk[e >> 6][e & 63] = v;
        }
    } else {
        const double* T = Tw + (size_t)i * 4096;
        Tile64 tile;
        tile.zero();
        if (pre)      // M_i,j+1 (row block i, column block j+1) times the saved L_j+1,j
            gemm64_k64<true>(Aw + (size_t)i * kNB * np + (size_t)(j + 1) * kNB, np,
                             Tw + (size_t)j * 4096, 64, tile, As, Bs);
        tile.for_each([&](int r, int c, double v) { blk[r][c] = T[r * 64 + c] + v; });
        __syncthreads();
        tile.zero();
        gemm64_k64<true>(&blk[0][0], 66, D, 64, tile, As, Bs);
        tile.for_each([&](int r, int c, double v) {
            Aij[(size_t)r * np + c] = -v;
            blk[r][c] = -v;
        });
    }
    __syncthreads();
    if (slice)
        slice_block<kS, B>(blk, rsc, nullptr, blk_index(w, nblk, i, j), RP, RS, nullptr, nullptr);
}

// LDS byte offset of 16 B chunk `ch` (0..3) of row `row` in a 64 x 64 byte
// slice block: chunks rotated by row / 4, so the 16 lanes of a 16 B access
// (4 rows x 4 chunks stored, 16 rows x 1 chunk read) hit distinct bank groups.
__device__ inline int blk_off(int row, int ch) { return row * 64 + ((ch ^ ((row >> 2) & 3)) << 4); }

// Block products on int8 MFMA for row block i = rowbeg + blockIdx.x and NC
// output column blocks c_l (l < NC), 64 x 64 each:
// MODE 0 (Cholesky update): c_l = kj + l,  A_w(i, c_l) -= sum_{kb in [kb0, kb1)} L_i,kb L_c_l,kb^T
//   with A = RP(i, kb), B = RP(c_l, kb); blocks above the diagonal (i < c_l) are not written.
// MODE 1 (inverse sweep): c_l = kj - l,  T_l(i) = (or +=, accumulate) sum_{kb in [kb0, min(kb1, i + 1))}
//   M_i,kb L_kb,c_l with A = RP(i, kb) (M rows), B = CP(kb, c_l) (L columns); blocks with
//   i <= c_l are not written; T_l is the 64 x 64 row-major block i of tbuf slot c_l % 2.
// Two output columns share every A operand block, halving the left-looking
// sweeps' operand traffic.  256 * NC threads: waves 4l..4l+3 compute column
// block l as 2 x 2 waves of 32 x 32.  Per K block the (1 + NC) x kS digit
// blocks and the operands' scale exponents move from global memory straight
// into LDS (global_load_lds, no register staging) into one of two stage
// buffers while the other feeds 2 k chunks x kS (kS + 1) / 2 digit products
// per wave; then
// the int32 groups are combined and scaled into FP64 accumulators with the
// exponents read from LDS.  The two stages are separate LDS objects and the K
// loop is unrolled by two, so every LDS read names a buffer that no load in
// flight writes and needs no wait on those loads.  Two waves per SIMD
// (<= 256 registers per wave).
template <int NC, int kS>
struct I8Stage {
    int8_t s[1 + NC][kS][4096];   // digit blocks: operand, digit, swizzled 64 x 64 bytes
    int sc[1 + NC][64];           // scale exponents per row of each operand
};

template <int MODE, int NC, int kS, int B>
__global__ __launch_bounds__(256 * NC) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_gemm_i8blk(
    const int8_t* __restrict__ RP, const double* __restrict__ RS, const int8_t* __restrict__ CP,
    const double* __restrict__ CS, double* __restrict__ Amat, double* __restrict__ tbuf, int np,
    int kj, int rowbeg, int kb0, int kb1, int accumulate) {
    constexpr size_t kBlkBytes = blk_bytes<kS>();
    // exponent of the combined groups' weight: 2^-(12 + 7 x 4), 2^-(14 + 8 x 3)
    constexpr int kExpOffset = B == 7 ? 40 : 38;
    static_assert((B == 7 && kS == 5) || (B == 8 && kS == 4), "digit scheme");
    __shared__ __attribute__((aligned(16))) I8Stage<NC, kS> st0;   // 2 x 40.5 / 60.75 KB (5 digits)
    __shared__ __attribute__((aligned(16))) I8Stage<NC, kS> st1;
    const int w = blockIdx.y, t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int nblk = np / 64;
    const int i = rowbeg + blockIdx.x;
    const int kend = MODE == 0 ? kb1 : min(kb1, i + 1);
    const int cb = wave >> 2;
    const int col = MODE == 0 ? kj + cb : kj - cb;
    const int wr = ((wave >> 1) & 1) * 32, wc = (wave & 1) * 32;
    const int8_t* Bbase = MODE == 0 ? RP : CP;
    const int* Bsc = reinterpret_cast<const int*>(MODE == 0 ? RS : CS);
    auto a_blk = [&](int kb) { return blk_index(w, nblk, i, kb); };
    auto b_blk = [&](int c, int kb) {
        return MODE == 0 ? blk_index(w, nblk, c, kb) : blk_index(w, nblk, kb, c);
    };
    auto op_col = [&](int op) { return MODE == 0 ? kj + op - 1 : kj - (op - 1); };

    // loads of K block kb into stage S (two output columns): the (1 + NC) kS slice blocks in 1 KB
    // pieces (16 rows), piece p = wave + 4 NC v; lane l of a piece fills LDS
    // bytes 16 l.. of it, i.e. row l / 4, stored chunk l % 4, so it reads the
    // logical chunk (l % 4) ^ ((row / 4) % 4) (the swizzle of blk_off).  The
    // last wave's lanes < 16 (1 + NC) load the scale exponents (16 B each).
    constexpr int kPieces = (1 + NC) * kS * 4;
    constexpr int kWaves = 4 * NC;
    auto issue = [&](int kb, I8Stage<NC, kS>& S) {
        if constexpr (NC == 1) {
            // one output column: almost always a single K block (the remainder
            // of a two-column sweep), where plain loads and LDS stores beat the
            // direct loads; thread t moves chunk t % 4 of row t / 4 of every
            // slice block, threads < 32 the scale exponents
            const int lrow = t >> 2, lch = t & 3;
            u32x4 v[2 * kS], e;
#pragma unroll
            for (int bk = 0; bk < 2 * kS; ++bk) {
                const int op = bk / kS, sl = bk % kS;
                const int8_t* base =
                    op == 0 ? RP + a_blk(kb) * kBlkBytes : Bbase + b_blk(op_col(op), kb) * kBlkBytes;
                v[bk] = *reinterpret_cast<const u32x4*>(base + (size_t)sl * 4096 + lrow * 64 + lch * 16);
            }
            if (t < 32) {
                const int op = t >> 4;
                const int* src = (op == 0 ? reinterpret_cast<const int*>(RS) + a_blk(kb) * 64
                                          : Bsc + b_blk(op_col(op), kb) * 64) +
                                 (t & 15) * 4;
                e = *reinterpret_cast<const u32x4*>(src);
            }
#pragma unroll
            for (int bk = 0; bk < 2 * kS; ++bk)
                *reinterpret_cast<u32x4*>(&S.s[bk / kS][bk % kS][blk_off(lrow, lch)]) = v[bk];
            if (t < 32) *reinterpret_cast<u32x4*>(&S.sc[t >> 4][(t & 15) * 4]) = e;
            return;
        }
#pragma unroll
        for (int v = 0; v < (kPieces + kWaves - 1) / kWaves; ++v) {
            const int p = wave + kWaves * v;
            if (p < kPieces) {
                const int bk = p >> 2, rg = p & 3;
                const int op = bk / kS, sl = bk % kS;
                const int row = rg * 16 + (lane >> 2);
                const int ch = (lane & 3) ^ ((row >> 2) & 3);
                const int8_t* src =
                    (op == 0 ? RP + a_blk(kb) * kBlkBytes : Bbase + b_blk(op_col(op), kb) * kBlkBytes) +
                    (size_t)sl * 4096 + row * 64 + ch * 16;
                __builtin_amdgcn_global_load_lds(
                    src, (__attribute__((address_space(3))) void*)&S.s[op][sl][rg * 1024], 16, 0, 0);
            }
        }
        if (wave == kWaves - 1 && lane < 16 * (1 + NC)) {
            const int op = lane >> 4;
            const int* src = (op == 0 ? reinterpret_cast<const int*>(RS) + a_blk(kb) * 64
                                      : Bsc + b_blk(op_col(op), kb) * 64) +
                             (lane & 15) * 4;
            __builtin_amdgcn_global_load_lds(
                src, (__attribute__((address_space(3))) void*)&S.sc[0][0], 16, 0, 0);
        }
    };
    // this wave's loads have landed in LDS (the barrier after it: everyone's)
    auto loads_done = [] { __builtin_amdgcn_s_waitcnt(0x0F70); };   // vmcnt(0)

    double acc64[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) acc64[r] = 0.0;
    const int l32 = lane & 31, kgl = lane >> 5;
    // the K block's slice products, then its scale exponents (row
    // row32(lane, r) for r = 4q..4q+3 is 4 consecutive rows from wr + 8q + 4 kgl)
    i32x16 acc[kS];
    int ea[16];
    auto products = [&](const I8Stage<NC, kS>& S) {
#pragma unroll
        for (int g = 0; g < kS; ++g)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[g][r] = 0;
#pragma unroll
        for (int kc = 0; kc < 2; ++kc) {
            const int ch = 2 * kc + kgl;
            const int oa = blk_off(wr + l32, ch), ob = blk_off(wc + l32, ch);
            u32x4 a[kS];
#pragma unroll
            for (int s = 0; s < kS; ++s) a[s] = *reinterpret_cast<const u32x4*>(&S.s[0][s][oa]);
#pragma unroll
            for (int j = 0; j < kS; ++j) {
                const u32x4 b = *reinterpret_cast<const u32x4*>(&S.s[1 + cb][j][ob]);
#pragma unroll
                for (int s = 0; s + j < kS; ++s) acc[s + j] = mfma_i8(a[s], b, acc[s + j]);
            }
        }
        const int eb = S.sc[1 + cb][wc + l32] - kExpOffset;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const u32x4 v = *reinterpret_cast<const u32x4*>(&S.sc[0][wr + 8 * q + 4 * kgl]);
#pragma unroll
            for (int j = 0; j < 4; ++j) ea[4 * q + j] = (int)v[j] + eb;
        }
    };
    // combine the groups with the row / column scales, exactly for K = 64:
    // 5 x 7 bits: sum_g G_g 2^-(12 + 7g) = (G0 2^28 + G1 2^21 + G2 2^14 + G3 2^7 + G4) 2^-40;
    //   |G_g| <= 5 * 2^18, so hi = G0 * 128 + G1 and lo = G3 * 128 + G4 fit
    //   in int32 and z = hi 2^21 + G2 2^14 + lo < 2^47 is exact in FP64;
    // 4 x 8 bits: sum_g G_g 2^-(14 + 8g) = (G0 2^24 + G1 2^16 + G2 2^8 + G3) 2^-38;
    //   |G_g| <= 4 * 2^20, so hi = G0 * 256 + G1 fits in int32 and
    //   z = hi 2^16 + G2 2^8 + G3 < 2^47 is exact in FP64
    auto combine = [&] {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            double z;
            if constexpr (B == 7) {
                const int hi = acc[0][r] * 128 + acc[1][r];
                const int lo = acc[3][r] * 128 + acc[4][r];
                z = fma((double)hi, 0x1p21, fma((double)acc[2][r], 0x1p14, (double)lo));
            } else {
                const int hi = acc[0][r] * 256 + acc[1][r];
                z = fma((double)hi, 0x1p16, fma((double)acc[2][r], 0x1p8, (double)acc[3][r]));
            }
            acc64[r] += ldexp(z, ea[r]);
        }
    };
    // per K block: products from one stage, wait for the other stage's loads,
    // barrier (the other stage is complete, this one free), start the loads
    // of the block after next into this stage, then combine while they fly
    if constexpr (NC == 1) {
        // one stage (41.5 KB, two workgroups per CU): load, barrier, products
        for (int kb = kb0; kb < kend; ++kb) {
            if (kb > kb0) __syncthreads();      // the previous block's reads are done
            issue(kb, st0);
            __syncthreads();
            products(st0);
            combine();
        }
    } else if (kb0 < kend) {
        issue(kb0, st0);
        loads_done();
        __syncthreads();
        if (kb0 + 1 < kend) issue(kb0 + 1, st1);
        for (int kb = kb0; kb < kend; kb += 2) {
            products(st0);
            loads_done();
            __syncthreads();
            if (kb + 2 < kend) issue(kb + 2, st0);
            __builtin_amdgcn_sched_barrier(0);
            combine();
            if (kb + 1 >= kend) break;
            products(st1);
            loads_done();
            __syncthreads();
            if (kb + 3 < kend) issue(kb + 3, st1);
            __builtin_amdgcn_sched_barrier(0);
            combine();
        }
    }
    if (MODE == 0 ? i < col : i <= col) return;      // above the diagonal: not an output block
    if (MODE == 0) {
        double* out = Amat + (size_t)w * np * np + (size_t)i * 64 * np + (size_t)col * 64;
#pragma unroll
        for (int r = 0; r < 16; ++r) out[(size_t)(wr + row32(lane, r)) * np + wc + l32] -= acc64[r];
    } else {
        double* out = tbuf + (size_t)(col & 1) * gridDim.y * np * 64 + (size_t)w * np * 64 +
                      (size_t)i * 64 * 64;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            double& o = out[(wr + row32(lane, r)) * 64 + wc + l32];
            o = accumulate ? o + acc64[r] : acc64[r];
        }
    }
}

template <int MODE, int NC, int kS, int B>
void gemm_i8blk(const int8_t* RP, const double* RS, const int8_t* CP, const double* CS, double* A,
                double* tbuf, int np, int kj, int rowbeg, int rows, int kb0, int kb1, int acc, int nb,
                hipStream_t s) {
    if (rows <= 0) return;
    hipLaunchKernelGGL((k_gemm_i8blk<MODE, NC, kS, B>), dim3(rows, nb), dim3(256 * NC), 0, s, RP, RS,
                       CP, CS, A, tbuf, np, kj, rowbeg, kb0, kb1, acc);
}

// Left-looking blocked Cholesky, two block columns per sweep: the update of
// columns k and k+1 over the finished columns < k in one launch (they share
// the L_i,kb operands), then column k factored and its panel solved and
// sliced, then column k+1 factored and solved with its remaining product
// (kb = k) taken first in FP64.
template <int kS, int B>
hipError_t potrf_i8(double* A, double* dinv, int* info, int8_t* RP, double* RS, int8_t* CP,
                    double* CS, int np, int nb, hipStream_t s) {
    const int nblk = np / 64;
    auto finish = [&](int k, int pre) -> hipError_t {
        if (hipError_t e = launch_potf2_block(A, dinv, info, np, k, nb, s, pre); e != hipSuccess)
            return e;
        const int t = nblk - k - 1;
        if (t > 0)
            hipLaunchKernelGGL((k_trsm_slice<kS, B>), dim3(t, nb), dim3(256), 0, s, A, dinv, np, k,
                               pre, RP, RS, CP, CS);
        return hipGetLastError();
    };
    for (int k = 0; k < nblk; k += 2) {
        const bool two = k + 1 < nblk;
        if (k > 0) {
            if (two)
                gemm_i8blk<0, 2, kS, B>(RP, RS, CP, CS, A, nullptr, np, k, k, nblk - k, 0, k, 0, nb, s);
            else
                gemm_i8blk<0, 1, kS, B>(RP, RS, CP, CS, A, nullptr, np, k, k, nblk - k, 0, k, 0, nb, s);
        }
        if (hipError_t e = finish(k, 0); e != hipSuccess) return e;
        if (two)
            if (hipError_t e = finish(k + 1, 1); e != hipSuccess) return e;
    }
    return hipGetLastError();
}

// Inverse sweep, two block columns per step: T for columns j and j-1 over the
// inverted columns > j in one launch (shared M_i,kb operands), column j
// inverted and sliced (saving L_j,j-1), then column j-1 inverted with its
// remaining product (kb = j) added first in FP64.  tbuf holds two slots of
// nb x np x 64 (column parity).
template <int kS, int B>
hipError_t trtri_i8(double* A, const double* dinv, double* tbuf, int8_t* RP, double* RS,
                    const int8_t* CP, const double* CS, int np, int nb, hipStream_t s) {
    const int nblk = np / 64;
    const size_t slot = (size_t)nb * np * 64;
    auto column = [&](int j, int pre, int save) -> hipError_t {
        double* tj = tbuf + (size_t)(j & 1) * slot;
        double* tn = tbuf + (size_t)((j + 1) & 1) * slot;     // slot of column j - 1
        // column 0's blocks are never an operand of a later product: no slices
        hipLaunchKernelGGL((k_trtri_m_slice<kS, B>), dim3(nblk - j, nb), dim3(256), 0, s, A, dinv, tj,
                           tn, np, j, pre, save, j > 0 ? 1 : 0, RP, RS);
        return hipGetLastError();
    };
    for (int j = nblk - 1; j >= 0; j -= 2) {
        const bool two = j >= 1;
        if (two)
            gemm_i8blk<1, 2, kS, B>(RP, RS, CP, CS, A, tbuf, np, j, j, nblk - j, j + 1, nblk, 0, nb, s);
        else
            gemm_i8blk<1, 1, kS, B>(RP, RS, CP, CS, A, tbuf, np, j, j + 1, nblk - 1 - j, j + 1, nblk, 0,
                                    nb, s);
        if (hipError_t e = column(j, 0, two ? 1 : 0); e != hipSuccess) return e;
        if (two)
            if (hipError_t e = column(j - 1, 1, 0); e != hipSuccess) return e;
    }
    return hipGetLastError();
}

}  // namespace

size_t fact_slice_bytes(int np, int scheme) {
    const size_t nblk = (size_t)(np / 64);
    return nblk * nblk * ((size_t)i8_digits(scheme) * 4096 + 64 * 8);
}

hipError_t launch_potrf_i8(double* A, double* dinv, int* info, int8_t* RP, double* RS, int8_t* CP,
                           double* CS, int np, int nb, int scheme, hipStream_t s) {
    if (np % 64 != 0) return hipErrorInvalidValue;
    return scheme == kDigits4x8 ? potrf_i8<4, 8>(A, dinv, info, RP, RS, CP, CS, np, nb, s)
                                : potrf_i8<5, 7>(A, dinv, info, RP, RS, CP, CS, np, nb, s);
}

hipError_t launch_trtri_i8(double* A, const double* dinv, double* tbuf, int8_t* RP, double* RS,
                           const int8_t* CP, const double* CS, int np, int nb, int scheme,
                           hipStream_t s) {
    if (np % 64 != 0) return hipErrorInvalidValue;
    return scheme == kDigits4x8 ? trtri_i8<4, 8>(A, dinv, tbuf, RP, RS, CP, CS, np, nb, s)
                                : trtri_i8<5, 7>(A, dinv, tbuf, RP, RS, CP, CS, np, nb, s);
}

}  // namespace mgw
