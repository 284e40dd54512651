/*
 * isai_bsr.h -- ISAI (incomplete sparse approximate inverse) of the packed
 * ILU(0) factors, built column by column as in the reference.
 *
 * ISAI(L) column j solves L(S_j,S_j) m_j = e_j forward over S_j (the rows of
 * column j of L's pattern); ISAI(U) solves U(S_j,S_j) m_j = e_j backward.
 * Factor values are stored in BCSC order (Mbrow = column pointer, Mcol = rows),
 * row-major blocks, exactly as the reference's Mval.
 */
#ifndef ISAI_ISAI_BSR_H
#define ISAI_ISAI_BSR_H

#include "isai_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces build_bcsc_pattern_from_bsr_scalarILU (reference isai_bsr.h:246-284):
 * lower != 0 keeps blocks with bc <= br (else bc >= br), forces the diagonal and
 * stores the pattern column-wise with ascending, de-duplicated rows.
 * Host arrays: Mbrow[nb+1], Mcol[count]. */
int isai_bcsc_pattern_count(int nb, const int *Arow, const int *Acol, int lower, int *nnz);
int isai_bcsc_pattern(int nb, const int *Arow, const int *Acol, int lower, int *Mbrow, int *Mcol);

#ifdef __cplusplus
}
#endif
#endif
