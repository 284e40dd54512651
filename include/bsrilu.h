/*
 * bsrilu.h -- block ILU(0) in place, level analysis.
 *
 * Replaces bsrilu02_inplace (reference bsrilu.h:20-66), which wraps
 * cusparseDbsrilu02 (DIRECTION_ROW, POLICY_NO_LEVEL) and creates / destroys its
 * handle and buffer on every call.  The packed result is the same layout:
 * strictly-lower blocks L_ik, diagonal blocks L_ii\U_ii (unit-lower \ upper),
 * strictly-upper blocks U_ij, with (L U)_IJ = A_IJ on every stored block.
 *
 * Zero pivots: the reference prints a warning and continues (bsrilu.h:46-60).
 * Here a structural zero (missing diagonal block) stops in the analysis and a
 * numerical zero (U(i,i) has an exact zero on its diagonal) completes the
 * factorization; both return ISAI_ERR_ZERO_PIVOT with the first block row in
 * *zero_pivot (-1 when none).
 */
#ifndef ISAI_BSRILU_H
#define ISAI_BSRILU_H

#include "isai_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct isai_ilu_plan *isai_ilu_plan_t;

/* One-shot form with the reference's arguments (device pointers; synchronizes). */
int isai_bsrilu0_inplace(int nb, int nnzb, int bs, const int *d_row, const int *d_col, double *d_val,
                         int *zero_pivot, isai_stream_t stream);

/* Split form: the analysis (pattern only: levels, dispatch order, diagonal
 * positions) is reused by every factorization of values with that pattern. */
int isai_bsrilu0_analysis(int nb, int nnzb, int bs, const int *d_row, const int *d_col,
                          isai_ilu_plan_t *plan, int *structural_zero, isai_stream_t stream);
int isai_bsrilu0_factor(isai_ilu_plan_t plan, double *d_val, int *zero_pivot, isai_stream_t stream);
int isai_bsrilu0_plan_info(isai_ilu_plan_t plan, int *nlevels, int *max_row_blocks);
int isai_bsrilu0_destroy(isai_ilu_plan_t plan);

/* Level sets (host arrays): level(i) = 1 + max level(k) over stored blocks (i,k)
 * with k < i (lower != 0) or k > i (upper, counted from the last row), 0 when
 * there is none; *nlevels = max level + 1.  cuSPARSE computes these internally
 * (no reference code); the closed forms are i+j for a 2-D 5-point grid, i+j+k for
 * a 3-D 7-point grid, i+2j (2-D 9-point) and i+2j+4k (3-D 27-point). */
int isai_bsr_levels(int nb, const int *Arow, const int *Acol, int lower, int *level, int *nlevels);

#ifdef __cplusplus
}
#endif
#endif
