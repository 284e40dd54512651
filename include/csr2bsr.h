/*
 * csr2bsr.h -- CSR -> BSR conversion (host and device forms).
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

/* Device form of the same conversion, bit for bit (same block order, same
 * first-entry-assigned / later-entries-added rule, same summation order).  One
 * workgroup per block row sorts its block columns in LDS.
 * Count: d_bsrRowPtr[nb+1] receives the block row pointer, *nnzb the block count
 * (synchronizes).  Fill: d_bsrRowPtr from the count pass (a row pointer that does
 * not match is reported, nothing is written for such rows); d_bsrColInd[nnzb],
 * d_bsrVal[nnzb*bs*bs]; synchronizes.  A block row may hold at most 16384 CSR entries
 * (ISAI_ERR_UNSUPPORTED otherwise; the host form has no limit); bs = 1..8. */
int isai_csr_to_bsr_device_count(int n, int bs, const int *d_csrRowPtr, const int *d_csrColInd,
                                 int *d_bsrRowPtr, int *nnzb, isai_stream_t stream);
int isai_csr_to_bsr_device_fill(int n, int bs, const int *d_csrRowPtr, const int *d_csrColInd,
                                const double *d_csrVal, const int *d_bsrRowPtr, int *d_bsrColInd,
                                double *d_bsrVal, isai_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif
