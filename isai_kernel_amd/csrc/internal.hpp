// internal.hpp -- entry points shared between translation units of the library
// (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include "cape_sptrsv.h"
#include "isai_solve.h"
#include "isai_types.h"

namespace isai {

// BSR SpMV launch; when skip is non-null the kernel returns at once if *skip != 0
// (a converged solve keeps replaying its iteration without doing work).
int spmv_launch(int mb, int nnzb, int bs, isai_precision_t prec, const int *rp, const int *ci,
                const void *val, const double *x, double *y, hipStream_t st, const int *skip);

// bs = 4 product with FP32/BF16 values; x and y each either FP64 or FP32 vectors
// (x is rounded to FP32 on load either way, so an FP32 x gives the same result as
// the FP64 vector it was rounded from).
// aligned: every row starts on a 128-byte line of values (the handle's factors,
// pad_convert_lowp), which picks the streamed shape for FP32.
int spmv_lowp_launch(int mb, long long nnzb, isai_precision_t prec, const int *rp, const int *ci, const void *val,
                     const void *x, bool x_f32, void *y, bool y_f32, hipStream_t st, const int *skip,
                     bool aligned = false);

// Line-aligned reduced-precision copy of a bs = 4 factor (FP32 or BF16 values):
// prp = padded row pointer (each row's block count rounded up to the blocks per
// 128-byte line, 2 FP32 / 4 BF16), pci / pval of prp[nb] blocks; the padding
// blocks are zeros in the row's last column.  Asynchronous.
int pad_convert_lowp(int nb, const int *rp, const int *ci, const double *val, const int *prp, int *pci, void *pval,
                     int prec, hipStream_t st);

// bs = 4 FP64 product y = A x with fused dot products of y: part[wg] = sum y.a over
// the rows of workgroup wg, and with ndot = 2 also part[G + wg] = sum y.y, where
// G = spmv_dot_groups(mb) partials per dot; nnzb picks the row batch width.
int spmv_dot_groups(int mb);
// x_f32: x holds FP32 values (widened exactly on load), e.g. the output of a
// reduced-precision ISAI apply.
int spmv_dot_launch(int mb, const int *rp, const int *ci, const double *val, const void *x, bool x_f32, double *y,
                    const double *a, int ndot, double *part, hipStream_t st, const int *skip, long long nnzb);

// Q_j = PairPtr[Mbrow[j+1]] - PairPtr[Mbrow[j]] for every column into d_Q, and the
// largest Q_j into the int at (char *)d_acc + pair_acc_max_q_offset(); d_acc is
// pair_acc_bytes() bytes of device scratch the caller zeroes beforehand.
// Asynchronous.
size_t pair_acc_bytes();
size_t pair_acc_max_q_offset();
// Device-to-device copy of `count` doubles on stream `st` (streaming copy kernel).
int copy_f64_d2d(double *dst, const double *src, size_t count, hipStream_t st);
// isai_pairs_count without the read-back of the total (left at d_PairPtr[Mnnz]).
int pairs_count_async(int nb, int Mnnz, const int *d_Arow, const int *d_Acol, const int *d_Mbrow,
                      const int *d_Mcol, int upper, int *d_PairPtr, hipStream_t st);
int pair_column_q_launch(int nb, const int *d_Mbrow, const int *d_PairPtr, int *d_Q, void *d_acc, hipStream_t st);

// One ISAI factor for the column-wise build.  order (optional) lists the columns
// in the order the workgroups take them; the values do not depend on it.
struct FactorArgs {
    const int *Mbrow, *Mcol, *PairPtr, *PairPLocal, *PairSrc;
    const double *Dinv;
    double *Mval;
    const int *order;
};

// ISAI(L) and/or ISAI(U) in one launch; Nmax = LDS Y-cache blocks per column
// (columns longer than the cache read Y back from Mval).
int build_dispatch(int nb, int bs, const double *Aval, const FactorArgs *L, const FactorArgs *U, int Nmax,
                   hipStream_t st);

// Sync-free L then U solve of a plan, with the same skip convention.
int sptrsv_solve_internal(isai_sptrsv_plan_t plan, const double *d_LU, const double *d_b, double *d_x,
                          hipStream_t st, const int *skip);

struct HandleFactors {
    int nb, nnzb, nnzbL, nnzbU;
    const int *Lrow, *Lcol, *Urow, *Ucol;
    const void *Lval, *Uval;  // in the handle's current precision
    int precision;
};
// Factors of a setup handle (isai_handle_t) in its current storage precision; for
// FP32 / BF16 these are the line-aligned copies with their own padded patterns
// (nnzbL / nnzbU count the padding blocks, which are zeros).
int handle_factors_internal(isai_handle_t h, HandleFactors *out);

// Transposes (adjoint.hip).  transpose_pattern_host: pattern of A^T and perm[d] =
// block of A stored at position d of A^T.  transpose_blocks: dst[d] = src[perm[d]]^T.
// transpose_packed_lu: packed ILU(0) factor of A^T from the factor of A (bs = 4);
// trow_of[d] = block row of position d in A^T, diag[i] = position of A's block (i, i).
int transpose_pattern_host(int nb, const int *rp, const int *ci, int *trp, int *tci, int *perm);
int transpose_blocks(long long nnzb, int bs, const int *d_perm, const double *src, double *dst, hipStream_t st);
int transpose_packed_lu(int nnzb, const int *d_trow_of, const int *d_tcol, const int *d_perm, const int *d_diag,
                        const double *d_lu, double *d_lut, hipStream_t st);

// ---- mixed-precision inner solve (mixed.hip)
// Workgroups of the FP32 A product (grid-stride over 32-row chunks): its dot
// partials, mixed_dot_groups(nb) per dot, are summed by every workgroup of the
// consuming vector kernel.
constexpr int kMixedGroups = 2048;
struct InnerScalars {
    double rho_cur, rho_prev, alpha, omega, tol2, tol_in, rnorm2, pad;
    int done, iters;  // done: 0 running, 1 converged, 3 breakdown
    int pad2[2];
};
struct InnerVecs {  // FP32 vectors of n entries, and the FP64 partial sums
    float *r, *rh, *p, *v, *s, *t, *ph, *sh, *x, *tmp;
    double *part;   // 2 * mixed_dot_groups(nb)
    double *rpart;  // 2 * 512
};
struct InnerPrecond {  // ISAI factors in BCSR with FP32 / BF16 values
    isai_precision_t prec;
    int nnzbL, nnzbU;
    const int *Lrow, *Lcol, *Urow, *Ucol;
    const void *Lval, *Uval;
};
int mixed_dot_groups(int nb);
size_t inner_part_doubles(int nb);
// Value layouts of the FP32 A product: the caller's rows (read through L2), every
// row starting on a 128-byte line, or row pairs starting on a line (both padded
// with zero blocks and streamed with non-temporal loads).
enum { kA32Unaligned = 0, kA32RowsAligned = 1, kA32PairsAligned = 2 };
struct A32Operand {  // A with FP32 values in one of those layouts
    long long nnzb;  // blocks stored, padding included
    const int *rp, *ci;
    const float *val;
    int layout;
};
// y = A x with FP32 values, FP32 x and y; part[wg] = sum y.a (and part[G + wg] =
// sum y.y for ndot = 2) over the rows of workgroup wg, in FP64; ndot = 0: no sums
// (a and part unused).  G = mixed_dot_groups(nb) for every layout.
int mixed_spmv_f32_dot(int nb, long long nnzb, const int *rp, const int *ci, const float *val, const float *x,
                       float *y, const float *a, int ndot, double *part, hipStream_t st, const int *skip, int layout);
struct InnerLog {  // outcome of one inner solve, recorded on the device
    int done, iters;
    double tol_in;
};
// r = rhs (rounded to FP32), rhat = r, x = 0; tolerance relative to ||r||:
// min(0.5, max(floor_tol, 0.5 sqrt(*tol2) / |*gj|)) with *gj the outer residual
// estimate and *tol2 = tol^2 ||b||^2; starts done when *outer_done != 0.
int inner_init(size_t n, const double *rhs, InnerVecs *w, InnerScalars *IS, const int *outer_done,
               const double *gj, const double *tol2, double floor_tol, hipStream_t st);
int inner_iteration(size_t n, int nb, const A32Operand &A, const InnerPrecond *M, InnerVecs *w, InnerScalars *IS,
                    bool first, hipStream_t st);
// z = x in FP64; *log = the inner solve's done flag, iterations and tolerance.
int inner_finish(size_t n, const InnerVecs *w, double *z, const InnerScalars *IS, InnerLog *log, hipStream_t st);

}  // namespace isai
