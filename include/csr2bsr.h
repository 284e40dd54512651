/*
 * csr2bsr.h -- CSR -> BSR conversion (host).
 *
 * Replaces csr_to_bsr5 (reference csr2bsr.h:14-76), which takes std::vector
 * out-parameters and a compile-time BS (config.h:4) and exits when n % BS != 0
 * (csr2bsr.h:25-29).  Here bs is a runtime argument, the caller allocates the
 * outputs after a count pass, and n % bs != 0 returns ISAI_ERR_INVALID_ARG.
 *
 * Semantics kept bit-exact: nb = n / bs; block columns ascending within a block
 * row; an explicit zero entry still creates its block; duplicate entries are
 * summed in CSR order, the first entry of a new block being assigned into the
 * zero-filled block (csr2bsr.h:53-61); blocks row-major (csr2bsr.h:57,60,71-72).
 */
#ifndef ISAI_CSR2BSR_H
#define ISAI_CSR2BSR_H

#include "isai_types.h"

#ifdef __cplusplus
extern "C" {
#endif

int isai_csr_to_bsr_count(int n, int bs, const int *csrRowPtr, const int *csrColInd,
                          int *nb, int *nnzb);
/* bsrRowPtr[nb+1], bsrColInd[nnzb], bsrVal[nnzb*bs*bs], sized by the count pass. */
int isai_csr_to_bsr_fill(int n, int bs, const int *csrRowPtr, const int *csrColInd,
                         const double *csrVal, int *bsrRowPtr, int *bsrColInd, double *bsrVal);

#ifdef __cplusplus
}
#endif
#endif
