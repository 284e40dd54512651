/*
 * isai_adjoint.h -- the differentiation form: adjoints of x = A^{-1} b.
 *
 * No reference counterpart (the reference computes no gradients; its forward flow
 * is main.cu:449-935).  For a loss L(x) with g = dL/dx:
 *   dL/db = lambda,  where A^T lambda = g            (adjoint solve)
 *   dL/dA = -lambda x^T restricted to the BSR pattern (pattern-restricted outer product)
 *
 * The adjoint handle reuses the forward factors instead of refactoring A^T:
 *  - packed ILU(0) of A^T: from A ~ L U (scalar unit-lower L, upper U, D = diag(U)),
 *    A^T ~ (U^T D^{-1}) (D L^T), a unit-lower times an upper factor in the same
 *    packed layout, computed block by block from the forward factor; its sync-free
 *    triangular solves are the transposed SpTRSV;
 *  - ISAI: the forward preconditioner is M = M_U M_L, so the adjoint one is
 *    M^T = M_L^T M_U^T, applied as the transposed ISAI factors (M_U^T first).
 */
#ifndef ISAI_ADJOINT_H
#define ISAI_ADJOINT_H

#include "isai_solve.h"
#include "isai_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* BSR of A^T on the device: block (J, I) = (block (I, J) of A)^T, columns ascending
 * per block row.  d_trow[mb+1], d_tcol[nnzb], d_tval[nnzb*bs*bs] are caller
 * buffers; bs = 1..8.  Synchronizes (the pattern is transposed on the host). */
int isai_bsr_transpose(int mb, int nnzb, int bs, const int *d_row, const int *d_col, const double *d_val,
                       int *d_trow, int *d_tcol, double *d_tval, isai_stream_t stream);

/* Handle of the adjoint system A^T lambda = g built from the forward handle h
 * (same preconditioner kind and storage precision) without refactoring: it owns
 * A^T, the derived packed ILU(0) factor of A^T and the transposed ISAI factors.
 * Use it with isai_solve / isai_precond_apply / isai_set_preconditioner /
 * isai_handle_factors; release it with isai_destroy.  The forward handle's matrix
 * values are read once, here. */
int isai_adjoint_handle(isai_handle_t h, isai_handle_t *adjoint, isai_stream_t stream);

/* Pattern-restricted outer product: d_g[k] = alpha * lambda_I x_J^T for every
 * stored block k = (I, J) (alpha = -1 gives dL/dA).  bs = 4.  Asynchronous. */
int isai_bsr_outer_grad(int mb, int nnzb, int bs, const int *d_row, const int *d_col, const double *d_lambda,
                        const double *d_x, double alpha, double *d_g, isai_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif
