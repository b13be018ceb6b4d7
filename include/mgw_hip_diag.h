/*
 * mgw_hip_diag.h -- diagnostics of the MI355X screening backend.
 *
 * None of these has a counterpart in the reference and none is on the
 * G0W0 path: they check MFMA lane layouts and the factorization's tile on a
 * new device and measure the device's peak rates, for tests and the
 * benchmark.  They live in the same library (libmgw_hip.so) and use the
 * contexts of mgw_hip.h.
 */
#ifndef MGW_HIP_DIAG_H
#define MGW_HIP_DIAG_H

#include "mgw_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Diagnostic with no counterpart in the reference: C = A B^T for 64 x 64
 * row-major A, B through the factorization kernels' FP64 MFMA tile (checks
 * the operand / result lane layout on a new device). */
int mgw_hip_selftest_tile(mgw_hip_ctx* ctx, const double* a, const double* b, double* c);

/* Diagnostic with no counterpart in the reference: ntests single-MFMA runs
 * with per-lane inputs, raw per-lane results (lane layout check).
 * kind 0: v_mfma_f64_16x16x4, a/b double[ntests][64], d double[ntests][64][4];
 * kind 1/2: v_mfma_f32_16x16x32 bf16/f16, a/b uint16[ntests][64][8],
 * d float[ntests][64][4]; kind 3/4: v_mfma_f32_32x32x16 bf16/f16, same
 * inputs, d float[ntests][64][16]; kind 5: v_mfma_i32_32x32x32_i8,
 * a/b int8[ntests][64][16], d int32[ntests][64][16]. */
int mgw_hip_probe_mfma(mgw_hip_ctx* ctx, int32_t kind, int32_t ntests, const void* a,
                       const void* b, void* d);

/* Diagnostic with no counterpart in the reference: measured peaks of this
 * device -- out[0] FP64 MFMA, out[1] FP16 MFMA, out[2] BF16 MFMA (flop/s,
 * 16x16 tiles, register-resident operands, 32 waves per CU), out[3] HBM copy
 * bandwidth (bytes read + written per second, 2 GiB buffers), out[4] FP16 and
 * out[5] BF16 MFMA with 32x32x16 tiles, out[6] int8 MFMA 32x32x32 and out[7]
 * int8 16x16x64 (op/s), all with constant operands; out[8] int8 32x32x32 and
 * out[9] FP64 16x16x4 with random operands (bytes in [-64, 64], doubles in
 * [-1, 1)): MFMA throughput depends on the operand data, and these are the
 * rates contractions on real data can approach.  out holds 10 entries. */
int mgw_hip_measure_peaks(mgw_hip_ctx* ctx, double* out);

#ifdef __cplusplus
}
#endif

#endif /* MGW_HIP_DIAG_H */
