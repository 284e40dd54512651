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

/* Replaces the kernels build_DinvL_from_BSR5_kernel_fix / build_DinvU_...
 * (reference isai_bsr.h:91-158, launched at main.cu:659-663 and 154-159):
 * d_Dinv[k] = inverse of the unit-lower (lower != 0) or upper part of the packed
 * diagonal block of row k, Gauss-Jordan without pivoting with the reference's
 * 1e-14 pivot guard and +1e-12 diagonal retry.  Device pointers, bs in 1..8. */
int isai_build_dinv(int nb, int bs, const int *d_Arow, const int *d_Acol, const double *d_Aval,
                    int lower, double *d_Dinv, isai_stream_t stream);

/* Replaces count_pairs_lower_kernel_v2 + thrust::exclusive_scan (reference
 * isai_bsr.h:1378-1424, main.cu:665-690): d_PairPtr[Mnnz+1] becomes the
 * exclusive prefix of the per-entry pair counts and *total_pairs = d_PairPtr[Mnnz]
 * (this call synchronizes the stream to return it, as main.cu:690 does).
 * upper != 0 builds the ISAI(U) lists (p_local > idx), else ISAI(L) (p_local < idx). */
int isai_pairs_count(int nb, int Mnnz, const int *d_Arow, const int *d_Acol, const int *d_Mbrow,
                     const int *d_Mcol, int upper, int *d_PairPtr, int *total_pairs,
                     isai_stream_t stream);
/* No counterpart in the reference: isai_pairs_count for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results; the scan's scratch lives in a fixed per-device array
 * of the library (about 1 MiB, held until the process ends) instead of being
 * allocated, so Mnnz + 1 is limited to 2^28 (ISAI_ERR_CAPACITY beyond).  Calls that
 * overlap on one GPU share that array and race; serialize them. */
int isai_pairs_count_noalloc(int nb, int Mnnz, const int *d_Arow, const int *d_Acol,
                             const int *d_Mbrow, const int *d_Mcol, int upper, int *d_PairPtr,
                             int *total_pairs, isai_stream_t stream);
/* Replaces fill_pairs_lower_kernel_v2 (reference isai_bsr.h:1427-1476): PairPLocal
 * (index within S_j) and PairSrc (BSR position of A(t, S_j[p])).  Lower lists are
 * ascending in p_local, upper lists descending (the reference's ISAI(U) loop order,
 * isai_bsr.h:196). */
int isai_pairs_fill(int nb, const int *d_Arow, const int *d_Acol, const int *d_Mbrow,
                    const int *d_Mcol, int upper, const int *d_PairPtr, int *d_PairPLocal,
                    int *d_PairSrc, isai_stream_t stream);

/* Pair-list statistics of one ISAI factor: the figures that analyze_pair_lengths
 * and analyze_column_q_bins print (reference analyze_pair.h:6-163, 250-301),
 * computed on the device and returned in a struct.  Per column j: N_j =
 * Mbrow[j+1] - Mbrow[j], Q_j = the pairs of its entries (PairPtr[Mbrow[j+1]] -
 * PairPtr[Mbrow[j]]), qlen = the pairs of one entry.  Percentiles are the
 * reference's: sort ascending and take v[(size_t)((size - 1) * p)].  qlen
 * histogram bins: 0, 1, 2-3, 4-7, 8-15, 16-31, 32-63, 64-127, 128-255, >= 256.
 * Column bin of j: j * 10 / nb.  The ISAI build uses these statistics to size
 * its LDS Y cache and to order its columns (isai_handle_build_plan). */
typedef struct {
    int nb, mnnz;
    long long total_pairs;
    double n_avg, q_avg, qlen_avg;
    int n_max, n_p50, n_p90, n_p95, n_p99;
    int q_max, q_p50, q_p90, q_p95, q_p99;
    int qlen_max, qlen_p50, qlen_p90, qlen_p95, qlen_p99;
    int bin_cols[10];
    int bin_max_q[10];
    long long qlen_hist[10];
    long long bin_sum_q[10];
} isai_pair_stats_t;

/* Device pointers d_Mbrow[nb+1], d_PairPtr[Mnnz+1] (from isai_pairs_count);
 * synchronizes the stream and allocates temporaries. */
int isai_pair_stats(int nb, const int *d_Mbrow, const int *d_PairPtr, isai_pair_stats_t *stats,
                    isai_stream_t stream);
/* Per-column figures on the device (the input of analyze_column_q_distribution /
 * analyze_last_columns, analyze_pair.h:165-248, 303-327): d_N[nb], d_Q[nb] and
 * d_maxqlen[nb] (the largest qlen of the column's entries); asynchronous. */
int isai_pair_column_stats(int nb, const int *d_Mbrow, const int *d_PairPtr, int *d_N, int *d_Q,
                           int *d_maxqlen, isai_stream_t stream);

/* Replaces the ISAI(L) kernels (reference isai_bsr.h:288-399 cachedY, 869-974
 * tensor cachedY, 1189-1353 globalY, cacheY_isai.h:159-341 hybrid; launched at
 * main.cu:51-152): column-wise forward substitution over S_j with the pair lists
 * and DinvL, d_Mval in BCSC order.  Nmax = longest column (main.cu:699-703) sizes
 * the LDS Y cache; beyond the cache budget Y is read back from d_Mval. */
int isai_build_lower(int nb, int bs, const double *d_Aval, const int *d_Mbrow, const int *d_Mcol,
                     const int *d_PairPtr, const int *d_PairPLocal, const int *d_PairSrc,
                     const double *d_DinvL, double *d_Mval, int Nmax, isai_stream_t stream);
/* ISAI(U) with precomputed upper pair lists (same recurrence as
 * isai_upper_bsr5_warpcol_with_Dinv_kernel, reference isai_bsr.h:161-242). */
int isai_build_upper(int nb, int bs, const double *d_Aval, const int *d_Mbrow, const int *d_Mcol,
                     const int *d_PairPtr, const int *d_PairPLocal, const int *d_PairSrc,
                     const double *d_DinvU, double *d_Mval, int Nmax, isai_stream_t stream);
/* Drop-in form of the reference's ISAI(U) kernel arguments (isai_bsr.h:161-170 plus
 * Mnnz and Nmax): builds the upper pair lists internally (library-owned scratch,
 * freed before returning; synchronizes the stream). */
int isai_build_upper_lookup(int nb, int bs, const int *d_Arow, const int *d_Acol,
                            const double *d_Aval, const int *d_Mbrow, const int *d_Mcol, int Mnnz,
                            const double *d_DinvU, double *d_Mval, int Nmax, isai_stream_t stream);
/* Both factors in ONE launch: every column's local triangular system of ISAI(L)
 * and ISAI(U) is solved by the same batched kernel. */
int isai_build_lu(int nb, int bs, const double *d_Aval,
                  const int *d_MbrowL, const int *d_McolL, const int *d_PairPtrL,
                  const int *d_PairPLocalL, const int *d_PairSrcL, const double *d_DinvL,
                  double *d_MvalL,
                  const int *d_MbrowU, const int *d_McolU, const int *d_PairPtrU,
                  const int *d_PairPLocalU, const int *d_PairSrcU, const double *d_DinvU,
                  double *d_MvalU, int Nmax, isai_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif
