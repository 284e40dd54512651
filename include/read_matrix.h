/*
 * read_matrix.h -- MatrixMarket input and the host residual check of the
 * reference driver.
 */
#ifndef ISAI_READ_MATRIX_H
#define ISAI_READ_MATRIX_H

#include "isai_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces read_matrix_market_to_csr (reference read_matrix.h:12-99): banner must
 * contain "MatrixMarket", '%' comment lines skipped, "m n nnz" size line, entries
 * 1-based -> 0-based, sorted by (row, col) with duplicates kept (stable, so equal
 * coordinates keep file order), the symmetric / pattern qualifiers ignored as in
 * the reference.  The three arrays are malloc'ed; free them with isai_host_free.
 * Returns ISAI_ERR_IO on open / parse failures (the reference prints and returns
 * false). */
int isai_read_matrix_market(const char *path, int *nrows, int *ncols, int *nnz, int **csrRowPtr,
                            int **csrColInd, double **csrVal);

/* Replaces compute_rel_errors_csr (reference residual_check.h:16-71) with host x:
 * rel_error_x = ||x - x_true|| / ||x_true|| (skipped when x_true is null: the
 * output is then 0) and rel_residual = ||b - A x|| / ||b||, both FP64 over a CSR A. */
int isai_rel_errors_csr(int n, const int *csrRowPtr, const int *csrColInd, const double *csrVal,
                        const double *b, const double *x_true, const double *x, double *rel_error_x,
                        double *rel_residual);

#ifdef __cplusplus
}
#endif
#endif
