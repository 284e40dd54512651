/*
 * oracle.h -- FP64 CPU oracle for the block-sparse ILU(0)/ISAI preconditioning path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is linked or loaded
 * only by the test suite, by the package's smoke check and by the CPU-baseline
 * leg of the benchmark.  The product (libisai_amd.so) never calls it and has no
 * CPU fallback to it.
 *
 * Every function restates an algorithm of the reference (gmfff/isai_kernel) in
 * plain C; the file:line it follows is cited at each declaration.  Pinning (how
 * each function is checked against something other than itself) is listed in the
 * header comment of oracle.c.
 *
 * Conventions (as in the reference): 0-based int32 indices, row-major bs x bs
 * blocks, BSR = (rowptr[nb+1], colind[nnzb], val[nnzb*bs*bs]); BCSC ISAI storage
 * = (Mbrow = column pointer, Mcol = ascending row indices, row-major blocks).
 * Return codes: 0 = ok, negative = error (see ORA_ERR_*).
 */
#ifndef ISAI_ORACLE_H
#define ISAI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORA_OK 0
#define ORA_ERR_INVALID (-1)
#define ORA_ERR_PATTERN (-4)
#define ORA_ERR_ZERO_PIVOT (-5)
#define ORA_ERR_NOT_CONVERGED (-6)
#define ORA_ERR_CAPACITY (-8)

/* csr2bsr.h:14-76 (csr_to_bsr5).  Count pass: nb = n/bs, nnzb = distinct
 * (block row, block col) pairs.  Error if n % bs != 0 (reference exits). */
int ora_csr2bsr_count(int n, int bs, const int *rp, const int *ci, int *nb, int *nnzb);
/* Fill pass: columns ascending per block row, duplicates summed in CSR order,
 * the first entry of a new block assigned (csr2bsr.h:53-61). */
int ora_csr2bsr_fill(int n, int bs, const int *rp, const int *ci, const double *v,
                     int *brp, int *bci, double *bv);

/* bsr2csr.h:19-137 (cusparse bsr2csr, DIRECTION_ROW): every block entry kept,
 * nnz = nnzb*bs*bs (bsr2csr.h:43); scalar row r = br*bs+lr lists its blocks in
 * block-column order, lc = 0..bs-1 inside each block. */
int ora_bsr2csr(int nb, int bs, const int *brp, const int *bci, const double *bv,
                int *rp, int *ci, double *v);

/* isai_bsr.h:246-284 (build_bcsc_pattern_from_bsr_scalarILU).  lower != 0 keeps
 * bc <= br, else bc >= br; diagonal forced; stored column-wise, rows ascending.
 * Mcol must hold cap entries; *nnz receives the count (also when cap is short). */
int ora_bcsc_pattern(int nb, const int *Arow, const int *Acol, int lower,
                     int *Mbrow, int *Mcol, int cap, int *nnz);

/* bcsc2bcsr.h:11-30 (build_triangular_bcsr_pattern_from_A): row-wise triangle of
 * A, diagonal NOT forced. */
int ora_tri_bcsr_pattern(int nb, const int *Arow, const int *Acol, int lower,
                         int *Trow, int *Tcol, int cap, int *nnz);

/* bcsc2bcsr.h:32-60 (build_map_bcsc2bcsr_host): map[k] = BCSR position of BCSC
 * entry k; ORA_ERR_PATTERN on a missing entry ("Pattern mismatch"). */
int ora_map_bcsc2bcsr(int nb, const int *colptr, const int *rowind,
                      const int *rowptr, const int *colind, int *map);

/* bcsc2bcsr.h:62-80 (reorder_bcsc_val_to_bcsr_bs5): dst[map[k]] = src[k]. */
void ora_reorder_bcsc_to_bcsr(int nnzb, int bs, const int *map, const double *src, double *dst);

/* isai_bsr.h:1378-1424 (count_pairs_lower_kernel_v2) + main.cu:684 (exclusive
 * scan): PairPtr[Mnnz+1].  isai_bsr.h:1427-1476 (fill_pairs_lower_kernel_v2):
 * PairPLocal / PairSrc.  For ISAI(L) entry (j, idx), t = S_j[idx]: pairs are the
 * (p_local, srcA) with A having block (t, S_j[p_local]) at position srcA and
 * p_local < idx, in ascending p_local order. */
int ora_pairs_lower(int nb, const int *Arow, const int *Acol, const int *Mbrow, const int *Mcol,
                    int *PairPtr, int *PairPLocal, int *PairSrc, int cap, int *total);
/* Upper counterpart used by this build's ISAI(U) kernel: the products the
 * reference forms by binary-search lookup (isai_bsr.h:196-226) restricted to
 * blocks that exist, p_local > idx, in DESCENDING p_local order (the reference's
 * loop order p = N-1 .. idx+1). */
int ora_pairs_upper(int nb, const int *Arow, const int *Acol, const int *Mbrow, const int *Mcol,
                    int *PairPtr, int *PairPLocal, int *PairSrc, int cap, int *total);

/* analyze_pair.h:6-163 (analyze_pair_lengths) and :250-301 (analyze_column_q_bins,
 * numBins = 10).  Per column j: N_j = Mbrow[j+1]-Mbrow[j], Q_j = sum of the pair
 * counts qlen = PairPtr[g+1]-PairPtr[g] over its entries g.  Percentiles follow
 * the reference: sort ascending, take v[(size_t)((size-1) * p)].  Averages: total
 * N / nb, total Q / nb, total Q / Mnnz (0 when empty).  qlen histogram bins: 0, 1,
 * 2-3, 4-7, 8-15, 16-31, 32-63, 64-127, 128-255, >= 256.  Column bin of j:
 * j * 10 / nb (int64), capped at 9.  ints[]: nb, Mnnz, maxN, maxQ, maxQlen,
 * p50/p90/p95/p99 of N, Q and qlen (12 values), bin_cols[10], bin_maxQ[10]
 * (39 ints).  i64[]: totalQ, qlen_hist[10], bin_sumQ[10] (21).  dbl[]: avgN,
 * avgQ, avgQlen (3).  Per-column arrays colN/colQ/colMaxQlen may be NULL. */
int ora_pair_stats(int nb, const int *Mbrow, const int *PairPtr, int *ints, long long *i64, double *dbl,
                   int *colN, int *colQ, int *colMaxQlen);

/* isai_bsr.h:60-88 (invert5_dense): Gauss-Jordan without pivoting, fails when
 * |pivot| < eps.  Returns 1 on success, 0 on failure. */
int ora_invert_block(int bs, const double *A, double *Ainv, double eps);

/* isai_bsr.h:91-125 / 126-158 (build_DinvL/U_from_BSR5_kernel_fix): inverse of
 * the unit-lower (lower != 0) or upper part of the packed diagonal block, with
 * the +1e-12 diagonal retry of isai_bsr.h:114-120.  A missing diagonal block is
 * read as zero (bsr_lookup5_full, isai_bsr.h:33-57). */
void ora_build_dinv(int nb, int bs, const int *Arow, const int *Acol, const double *Aval,
                    int lower, double *Dinv);

/* isai_bsr.h:288-399 (isai_lower_bsr5_warpcol_with_pairs_kernel_cachedY):
 * column-wise ISAI of the packed unit-lower factor, forward over S_j. */
void ora_isai_lower(int nb, int bs, const double *Aval, const int *Mbrow, const int *Mcol,
                    const int *PairPtr, const int *PairPLocal, const int *PairSrc,
                    const double *DinvL, double *Mval);
/* isai_bsr.h:161-242 (isai_upper_bsr5_warpcol_with_Dinv_kernel): column-wise
 * ISAI of the upper factor, backward over S_j, block lookup per product. */
void ora_isai_upper(int nb, int bs, const int *Arow, const int *Acol, const double *Aval,
                    const int *Mbrow, const int *Mcol, const double *DinvU, double *Mval);

/* bsrilu.h:20-66 (cusparseDbsrilu02, DIRECTION_ROW, no pivoting): block ILU(0)
 * in place, packed as cuSPARSE stores it: strictly-lower blocks L_ik, diagonal
 * blocks L_ii\U_ii (unit-lower \ upper), strictly-upper blocks U_ij.
 * *zero_pivot = first block row with a missing diagonal block (structural) or a
 * zero diagonal entry of U (numerical), else -1. */
int ora_bsrilu0(int nb, int bs, const int *Arow, const int *Acol, double *Aval, int *zero_pivot);

/* Level sets of the block triangular solve / ILU(0) dependency graph:
 * level(i) = 1 + max level(k) over stored blocks (i,k), k < i (lower) or k > i
 * (upper, levels counted from the last row); 0 when there is none.
 * Returns the number of levels. */
int ora_levels(int nb, const int *Arow, const int *Acol, int lower, int *level);

/* bcsc2bcsrAndSpMV.h:48-73 / 119-137 (bcsr_spmv, cpu_bcsr_spmv): y = M x. */
void ora_bsr_spmv(int nb, int bs, const int *rp, const int *ci, const double *val,
                  const double *x, double *y);
/* FP32 / BF16 stored values, FP32 accumulation: the reduced-precision ISAI apply.
 * Values are rounded from FP64 with round-to-nearest-even before the product;
 * products accumulate in float in block-column order, x rounded to float. */
void ora_bsr_spmv_f32acc(int nb, int bs, const int *rp, const int *ci, const double *val,
                         int bf16_values, const double *x, double *y);
/* Round-to-nearest-even helpers used by the above (exposed for the tests). */
float ora_round_f32(double v);
float ora_round_bf16(double v);

/* Sequential block triangular solves with the packed ILU(0) factor (the result
 * the sync-free solve of cape_sptrsv.h:8-103 computes, in BSR form):
 * L y = b (unit diagonal), then U x = y. */
void ora_bsr_sptrsv_lu(int nb, int bs, const int *rp, const int *ci, const double *LU,
                       const double *b, double *x);
/* Scalar CSR form with cape semantics (cape_sptrsv.h:8-53 unit L, ascending
 * columns; 55-103 U divided by the diagonal entry, descending columns). */
int ora_csr_sptrsv_lu(int m, const int *rp, const int *ci, const double *v,
                      const double *b, double *x);

/* residual_check.h:36-70 (compute_rel_errors_csr) on BSR storage: returns
 * ||b - A x||_2 / ||b||_2. */
double ora_rel_residual_bsr(int nb, int bs, const int *rp, const int *ci, const double *val,
                            const double *b, const double *x);

/* Right-preconditioned BiCGSTAB (van der Vorst 1992; Saad, Iterative Methods,
 * 2nd ed., Alg. 7.7 with M^{-1} applied on the right).  No reference counterpart.
 * precond: 0 = none, 1 = ISAI apply z = M_U (M_L v) with BCSR factors,
 * 2 = exact ILU(0) solves with the packed factor LU.  Iterates until
 * ||r||/||b|| <= tol (recursive residual) or maxit.  Returns 0 or
 * ORA_ERR_NOT_CONVERGED; *iters and *relres receive the count and the TRUE
 * residual ||b - A x|| / ||b||. */
typedef struct {
    int precond;
    int nb, bs;
    const int *Lrp, *Lci; const double *Lval;   /* ISAI M_L in BCSR */
    const int *Urp, *Uci; const double *Uval;   /* ISAI M_U in BCSR */
    const int *LUrp, *LUci; const double *LUval; /* packed ILU(0) */
    int isai_bf16; int isai_f32;                /* reduced-precision apply */
} ora_precond_t;

int ora_bicgstab(int nb, int bs, const int *rp, const int *ci, const double *val,
                 const ora_precond_t *P, const double *b, double *x,
                 double tol, int maxit, int *iters, double *relres);
/* Flexible GMRES(m) with modified Gram-Schmidt (Saad Alg. 9.6). */
int ora_fgmres(int nb, int bs, const int *rp, const int *ci, const double *val,
               const ora_precond_t *P, const double *b, double *x, int m,
               double tol, int maxit, int *iters, double *relres);

/* Mixed-precision form (no reference counterpart): FP64 FGMRES(m) (modified
 * Gram-Schmidt, Saad Alg. 9.6) whose flexible preconditioner is an inner
 * right-preconditioned BiCGSTAB in FP32 storage:
 *   - A's values rounded to FP32 (RNE), products accumulated in FP32
 *     (ora_bsr_spmv_f32acc);
 *   - the ISAI apply of P (give P reduced-precision flags for FP32 / BF16 factors);
 *   - every inner vector rounded to FP32 after each update; the scalars alpha,
 *     beta, omega rounded to FP32 where they scale a vector;
 *   - dot products of the FP32 vectors accumulated in FP64.
 * Inner solve j starts from x = 0 on rhs = v_j and stops when ||r_k|| <= tol_in
 * ||r_0|| (tested from the second iteration on), rho = 0, or inner_maxit
 * iterations, with tol_in = min(0.5, max(inner_tol, 0.5 tol ||b|| / |g_j|)) and
 * |g_j| the outer residual estimate before step j.  Outer steps stop when |g_{j+1}|
 * <= tol ||b||; a cycle ends with the FP64 true residual.  *iters = inner
 * iterations (bounded by maxit), *outer = outer steps, *relres = the true FP64
 * relative residual. */
int ora_fgmres_mixed(int nb, int bs, const int *rp, const int *ci, const double *val,
                     const ora_precond_t *P, const double *b, double *x, int m, double tol,
                     int maxit, double inner_tol, int inner_maxit, int *iters, int *outer,
                     double *relres);

/* The same outer loop whose flexible preconditioner is one inner GMRES cycle of at
 * most inner_maxit steps in FP32 storage (right-preconditioned by P): q_0 =
 * fp32(v_j) with beta = 1, zh = fp32(P q_i), w = fp32(A zh) with A rounded to FP32
 * and FP32 accumulation, one classical Gram-Schmidt pass (dots in
 * FP64, w stored in FP32), FP64 Hessenberg and Givens rotations, stop when |g_{i+1}|
 * <= tol_in (same tol_in rule), z_j = P fp32(Q y).  *iters = inner Arnoldi steps. */
int ora_fgmres_mixed_gmres(int nb, int bs, const int *rp, const int *ci, const double *val,
                           const ora_precond_t *P, const double *b, double *x, int m, double tol,
                           int maxit, double inner_tol, int inner_maxit, int *iters, int *outer,
                           double *relres);

/* Apply the preconditioner once: z = P v. */
void ora_precond_apply(const ora_precond_t *P, const double *v, double *z);

/* ---- differentiation form (no reference counterpart: the reference has no
 * gradients; these restate the adjoint of x = A^{-1} b, L = L(x)). */

/* BSR of A^T: block (J, I) of the result = (block (I, J) of A)^T; columns
 * ascending per block row.  Sizes: trp[nb+1], tci[nnzb], tval[nnzb*bs*bs]. */
void ora_bsr_transpose(int nb, int bs, const int *rp, const int *ci, const double *val,
                       int *trp, int *tci, double *tval);
/* Packed ILU(0) factor of A^T derived from the packed factor of A (bs = 4 layout
 * of ora_bsrilu0, any bs): with A ~ L U (scalar unit-lower L, upper U, diagonal
 * D = diag(U)), A^T ~ L' U' with L' = U^T D^{-1} (unit lower) and U' = D L^T.
 * Output on the transposed pattern (trp, tci from ora_bsr_transpose). */
int ora_transpose_packed_lu(int nb, int bs, const int *rp, const int *ci, const double *lu,
                            const int *trp, const int *tci, double *lut);
/* Pattern-restricted outer product: g[k] = alpha * lam_I x_J^T for block k = (I, J). */
void ora_bsr_outer_grad(int nb, int bs, const int *rp, const int *ci, const double *lam,
                        const double *x, double alpha, double *g);

/* Number of OpenMP threads the oracle runs with (1 when built without OpenMP). */
int ora_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
