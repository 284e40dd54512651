/*
 * bsr_tensor_SpMV.h -- BSR SpMV y = A x, the ISAI apply x = M_U (M_L b), and the
 * reduced-precision factor forms.
 *
 * Replaces the reference's SpMV kernels, which the caller launches directly:
 *   bcsr_spmv                 bcsc2bcsrAndSpMV.h:48-73 (the ISAI apply of main.cu:820-848)
 *   bsr_spmv_kernel(_new)     bsr_tensor_SpMV.h:21-115
 *   bsr_spmv_bdim5_kernel(*)  bsr_tensor_SpMV.h:140-400
 *   bsr_SpMV_tensorcore       bsr_tensor_SpMV.h:402-479 (self-contained host benchmark)
 * Device pointers, asynchronous on stream unless stated.  bs = 4 runs the
 * coalesced block-row streaming kernels; other bs in 1..8 run a generic kernel.
 */
#ifndef ISAI_BSR_TENSOR_SPMV_H
#define ISAI_BSR_TENSOR_SPMV_H

#include "isai_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* y[mb*bs] = A x with FP64 values, FP64 accumulation. */
int isai_bsr_spmv(int mb, int nnzb, int bs, const int *d_rowptr, const int *d_colind,
                  const double *d_val, const double *d_x, double *d_y, isai_stream_t stream);

/* Same product with values stored in prec (FP32: float[nnzb*bs*bs], BF16:
 * uint16_t bfloat16 bits); x and y stay FP64; FP32 accumulation for FP32/BF16. */
int isai_bsr_spmv_prec(int mb, int nnzb, int bs, isai_precision_t prec, const int *d_rowptr,
                       const int *d_colind, const void *d_val, const double *d_x, double *d_y,
                       isai_stream_t stream);

/* Round FP64 block values to prec (round-to-nearest-even; BF16 via FP32). */
int isai_convert_values(size_t count, isai_precision_t prec, const double *d_src, void *d_dst,
                        isai_stream_t stream);

/* ISAI apply as in main.cu:817-849: d_y = M_L d_b, then d_x = M_U d_y.  Both
 * factors in BCSR, values in prec; d_y is caller-provided scratch (mb*bs). */
int isai_apply(int mb, int bs, isai_precision_t prec,
               const int *d_Lrow, const int *d_Lcol, const void *d_Lval, int nnzbL,
               const int *d_Urow, const int *d_Ucol, const void *d_Uval, int nnzbU,
               const double *d_b, double *d_y, double *d_x, isai_stream_t stream);

/* The FP32 product of the mixed-precision inner solve (bs = 4, no reference
 * counterpart): d_y = A d_x with d_val the FP32 block values (isai_convert_values),
 * FP32 x and y and FP32 accumulation, plus fused FP64 dot products of the result:
 * over the rows of workgroup g, d_part[g] = sum y.a and, with ndot = 2,
 * d_part[G + g] = sum y.y.  *groups receives G = min(ceil(mb / 32), 2048); d_part
 * holds 2 G doubles. */
int isai_bsr_spmv_f32_dot(int mb, int nnzb, const int *d_rowptr, const int *d_colind, const float *d_val,
                          const float *d_x, float *d_y, const float *d_a, int ndot, double *d_part, int *groups,
                          isai_stream_t stream);

/* Host-pointer benchmark mirroring bsr_SpMV_tensorcore: copies A and B in, runs
 * warmup + iters launches timed with hipEvents, copies C out (C = res_Y as in the
 * reference's two outputs) and returns the average milliseconds per launch. */
int isai_bsr_spmv_bench(const int *bsrRowPtr, const int *bsrColIdx, const double *bsrVal,
                        int mb, int nnzb, int bdim, const double *Val_B, double *Val_C,
                        int warmup, int iters, double *avg_ms);

#ifdef __cplusplus
}
#endif
#endif
