/*
 * mole_gpu_debug.h -- diagnostic entry points of libmole_amd.so.
 *
 * These have no counterpart in the reference.  They exist so that tests can
 * check, element by element, that the device primitives the walker kernels
 * use (Philox draws, the u64 -> f64 maps, exp, sqrt, division) round exactly
 * like the CPU restatement they are compared with.
 */
#ifndef MOLE_GPU_DEBUG_H
#define MOLE_GPU_DEBUG_H

#include "mole_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* draws first .. first+count-1 of the u64 stream of (walker, epoch) */
mole_status mole_debug_philox(uint64_t seed, uint32_t sys, uint32_t walker, uint32_t epoch,
                              uint64_t first, size_t count, uint64_t *out);

/* Copies n_doubles doubles between two device buffers `repeats` times with a
 * plain grid-stride kernel (copy_probe_kernel) and reports the average
 * device time per copy: the known-size reference for HBM byte counters. */
mole_status mole_debug_copy(size_t n_doubles, int32_t repeats, double *ms_per_copy);

#define MOLE_DEBUG_EXP 0     /* out = exp(a)            */
#define MOLE_DEBUG_SQRT 1     /* out = sqrt(a)           */
#define MOLE_DEBUG_DIV 2      /* out = a / b             */
#define MOLE_DEBUG_STANDARD 3 /* out = Standard(bits(a)) */
#define MOLE_DEBUG_RANGE 4    /* out = Range(-b, b)(bits(a)) */
#define MOLE_DEBUG_NORM3 5    /* out = |(a, b, a*b)|      */
#define MOLE_DEBUG_LOG 6      /* out = ln a               */
#define MOLE_DEBUG_SINPI 7    /* out = sin(pi a), 0 <= a < 2 */
#define MOLE_DEBUG_COSPI 8    /* out = cos(pi a), 0 <= a < 2 */
#define MOLE_DEBUG_BOXMULLER0 9  /* out = first normal of Box-Muller(bits(a), bits(b)) */
#define MOLE_DEBUG_BOXMULLER1 10 /* out = second normal */
#define MOLE_DEBUG_BF16 11    /* out = bf16 round-to-nearest-even of (float)a */
mole_status mole_debug_math(int32_t op, size_t n, const double *a, const double *b, double *out);

/* D = A B for 16x16 row-major FP32 matrices rounded to bf16, through one
 * v_mfma_f32_16x16x16_bf16 with the lane placement the walker MLP kernels
 * use: pins that placement against a host product. */
mole_status mole_debug_mfma_bf16(const float *A, const float *B, float *D);

/* D (16x16) = A (16x32) B (32x16) through one v_mfma_f32_16x16x32_bf16 with
 * the walker MLP's placement: lane l holds A[l%16][8(l/16)+j] and
 * B[8(l/16)+j][l%16], j = 0..7. */
mole_status mole_debug_mfma_bf16_k32(const float *A, const float *B, float *D);

/* The learned proposal's network inputs (mlp_features) of n (configuration
 * [N][3], electron) samples as 16 floats each (the 13 features, then zeros):
 * pins the neighbour search and its tie order against the CPU restatement. */
mole_status mole_debug_mlp_features(const double *cfg, uint32_t N, const double *nuc_xyz, uint32_t n_nuc,
                                    const int32_t *electron, size_t n, float *out);

#ifdef __cplusplus
}
#endif
#endif
