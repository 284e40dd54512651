/*
 * bcsc2bcsr.h -- triangular BCSR patterns of A, the BCSC -> BCSR position map
 * and the value permutation that turns column-built ISAI factors into BCSR.
 */
#ifndef ISAI_BCSC2BCSR_H
#define ISAI_BCSC2BCSR_H

#include "isai_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces build_triangular_bcsr_pattern_from_A (reference bcsc2bcsr.h:11-30):
 * the lower (c <= r) or upper (c >= r) blocks of A, row-wise, diagonal NOT forced.
 * Host arrays; Trow[nb+1], Tcol[count]. */
int isai_tri_bcsr_pattern_count(int nb, const int *Arow, const int *Acol, int lower, int *nnz);
int isai_tri_bcsr_pattern(int nb, const int *Arow, const int *Acol, int lower,
                          int *Trow, int *Tcol);

/* Replaces build_map_bcsc2bcsr_host (reference bcsc2bcsr.h:32-60): map[k] is the
 * BCSR position of BCSC entry k.  A BCSC entry missing from the BCSR pattern
 * returns ISAI_ERR_PATTERN (the reference prints "Pattern mismatch" and exits). */
int isai_map_bcsc2bcsr(int nb, const int *bcscColPtr, const int *bcscRowInd,
                       const int *bcsrRowPtr, const int *bcsrColInd, int *map);

/* Replaces the kernel reorder_bcsc_val_to_bcsr_bs5 (reference bcsc2bcsr.h:62-80,
 * launched at main.cu:782-788): bcsr_val[map[k]] = bcsc_val[k], block-wise.
 * Device pointers; asynchronous on stream. */
int isai_reorder_bcsc_to_bcsr(int nnzb, int bs, const int *d_map, const double *d_bcsc_val,
                              double *d_bcsr_val, isai_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif
