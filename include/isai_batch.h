/*
 * isai_batch.h -- the batching form: many independent, uneven-sized systems, each
 * with its own number of right-hand sides, set up and solved together.
 *
 * No reference counterpart (the reference solves one matrix, main.cu:449-935).
 *
 * The systems are passed as ONE block-diagonal BSR matrix (global row pointer and
 * global block column indices) with the block-row offsets of the systems
 * (sys_ptr[nsys+1]).  Setup runs the ILU(0), the pair lists and the ISAI build of
 * every system in single launches over the whole batch.
 *
 * Multi-vectors: the right-hand side j of system s is the segment of length
 * 4*nb_s starting at isai_batch_vector_offset(s, j); segments are padded to a
 * multiple of 1024 doubles (the padding must be zero and stays zero).  The total
 * length is isai_batch_vector_length().
 *
 * Solve: right-preconditioned BiCGSTAB with one set of Krylov scalars per
 * (system, right-hand side) pair, each pair stopping on its own; FP64 Krylov,
 * ISAI factors in the chosen storage precision.  Two schedules compute the same
 * iterates bit for bit:
 *   - per-step launches: the steps of four iterations of the whole batch are
 *     captured once as a hipGraph and replayed, the host reading one word (pairs
 *     still running) per replay;
 *   - one persistent kernel: a workgroup per CU runs the whole solve (restarts
 *     included), the steps separated by grid barriers whose waits are bounded
 *     (isai_batch_set_spin_limit; an abandoned wait returns ISAI_ERR_TIMEOUT);
 *     the launch is captured once as a hipGraph and replayed per solve.
 * The multi-RHS BSR product loads every block once for all right-hand sides of
 * its system; its launch shape is chosen at setup by timing the candidates on the
 * batch itself (isai_batch_tuning).
 */
#ifndef ISAI_BATCH_H
#define ISAI_BATCH_H

#include "isai_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct isai_batch *isai_batch_t;

int isai_batch_setup(int nsys, const int *sys_ptr, const int *nrhs, int nnzb, const int *d_row,
                     const int *d_col, const double *d_val, int precision, isai_batch_t *batch,
                     isai_stream_t stream);
int isai_batch_vector_length(isai_batch_t batch, size_t *length);
int isai_batch_vector_offset(isai_batch_t batch, int sys, int rhs, size_t *offset);
/* Number of (system, right-hand side) pairs = sum of nrhs. */
int isai_batch_pairs(isai_batch_t batch, int *npairs);
/* Chosen SpMM variant (index into the candidates) and its time per product in us. */
int isai_batch_tuning(isai_batch_t batch, int *variant, double *us, int *ncandidates);
/* The autotuner's table for one operand (0 = A, 1 = the ISAI factors): us[c] = time
 * per product of candidate c = mapping (c / 6: 0 row-pair lanes, 1 column-split
 * lanes, 2 MFMA, 3 whole-row: row-pair lanes for FP64 operands, MFMA lanes for
 * FP32 / BF16 factors; each equals mapping 0 or 2 bit for bit),
 * workgroup size {64, 128, 256}[(c / 2) % 3], XCD remap c % 2; up to capacity
 * entries, *count = number of candidates (24). */
int isai_batch_tuning_table(isai_batch_t batch, int operand, double *us, int capacity, int *count);
/* Overrides the autotuner's choice for one operand with candidate c (numbering as
 * above); a captured iteration graph is dropped and recaptured on the next solve. */
int isai_batch_select(isai_batch_t batch, int operand, int candidate);
/* Setup device time: ILU(0) and ISAI build (both over the whole batch). */
int isai_batch_setup_times(isai_batch_t batch, double *ms_ilu, double *ms_isai);

/* Y = A X for every pair (block-diagonal product, multi-RHS). */
int isai_batch_spmm(isai_batch_t batch, const double *d_X, double *d_Y, isai_stream_t stream);

/* Solve A_s x_{s,j} = b_{s,j} for every pair; d_X holds initial guesses.
 * iters[npairs] and relres[npairs] (host arrays, may be null) receive each pair's
 * iteration count and true FP64 relative residual; *max_iters the largest count.
 * use_graph: 0 per-step launches, 1 per-step launches replayed as a captured
 * hipGraph, 2 the persistent kernel replayed as a captured hipGraph, 3 the
 * persistent kernel launched directly.  Modes 1 and 2 need a created stream. */
int isai_batch_solve(isai_batch_t batch, double tol, int maxit, const double *d_B, double *d_X, int use_graph,
                     int *iters, double *relres, int *max_iters, isai_stream_t stream);
/* Bound on one grid-barrier wait of the persistent kernel, microseconds (default
 * 5 s); a wait beyond it makes the solve return ISAI_ERR_TIMEOUT. */
int isai_batch_set_spin_limit(isai_batch_t batch, long long us);

int isai_batch_destroy(isai_batch_t batch);

#ifdef __cplusplus
}
#endif
#endif
