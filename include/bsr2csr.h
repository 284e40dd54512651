/*
 * bsr2csr.h -- BSR -> CSR expansion on the device.
 *
 * Replaces bsr2csr_cusparse (reference bsr2csr.h:19-137), which wraps
 * cusparseDbsr2csr with host-pointer in/out and malloc'ed host results.  Same
 * output: every block entry is kept (nnz = nnzb*bs*bs, bsr2csr.h:43); scalar row
 * r = br*bs + lr lists its blocks in block-column order and lc = 0..bs-1 within
 * each block (CUSPARSE_DIRECTION_ROW, bsr2csr.h:68,87-95).
 * Device pointers: d_rp[mb*bs+1], d_ci/d_v[nnzb*bs*bs]; asynchronous on stream.
 */
#ifndef ISAI_BSR2CSR_H
#define ISAI_BSR2CSR_H

#include "isai_types.h"

#ifdef __cplusplus
extern "C" {
#endif

int isai_bsr_to_csr(int mb, int nnzb, int bs, const int *d_bsrRowPtr, const int *d_bsrColInd,
                    const double *d_bsrVal, int *d_csrRowPtr, int *d_csrColInd, double *d_csrVal,
                    isai_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif
