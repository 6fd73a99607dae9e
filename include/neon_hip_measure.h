/* Measurement probes over libneon_hip.so contexts (libneon_hip_measure.so).  None of these has
 * a counterpart in the reference: they time the device kernels for tuning and for the roofline
 * figures, and calibrate the HBM traffic counters.  The product library exports none of them. */
#ifndef NEON_HIP_MEASURE_H
#define NEON_HIP_MEASURE_H

#include "neon_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Extra flags of neon_mesh_spmv_bench beside NEON_SPMV_NONTEMPORAL / SLICED / INDEX32 */
#define NEON_MEASURE_DOT 4096   /* the solvers' form, with block partials of x . y */
#define NEON_MEASURE_DIRTY 8192 /* each product follows a write of 4 vectors of n (timed apart, subtracted) */

/* Average milliseconds of one SpMV y = K x (x = 1) of the assembled tangent in the given
 * precision: the sliced copy with NEON_SPMV_SLICED, else the CSR-order values with `lanes`
 * (16, 32) lanes per block row; `blocks` workgroups (0: by size); reps timed launches after one
 * untimed. */
int neon_mesh_spmv_bench(neon_context* ctx, int32_t precision, int32_t lanes, int32_t flags, int32_t blocks,
                         int32_t reps, double* avg_ms);
/* One product y = K x of the assembled tangent in the given shape (flags and blocks as above;
 * NEON_MEASURE_DOT selects the kernel instance that also writes x . y partials); x and y are
 * host arrays of 3N, the FP32 path rounds x to FP32 and widens y. */
int neon_mesh_spmv_apply(neon_context* ctx, int32_t precision, int32_t lanes, int32_t flags, int32_t blocks,
                         const double* x, double* y);
/* Average milliseconds of one assembly (tangent + f_int, all colours) of the held state with
 * `blocks_per_cu` one-wavefront workgroups per compute unit and colour. */
int neon_mesh_assembly_bench(neon_context* ctx, int32_t precision, int32_t blocks_per_cu, int32_t reps,
                             double* avg_ms);
/* Average milliseconds of a streaming read and of a copy of `bytes` bytes. */
int neon_stream_probe(neon_context* ctx, int64_t bytes, int32_t reps, double* read_ms, double* copy_ms);

#ifdef __cplusplus
}
#endif
#endif
