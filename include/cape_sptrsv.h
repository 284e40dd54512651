/*
 * cape_sptrsv.h -- sync-free triangular solves with the packed ILU(0) factor:
 * L y = b (unit lower), then U x = y.
 *
 * Replaces spTrSolveL / spTrSolveU / csr_cape_SpTRSV (reference
 * cape_sptrsv.h:8-179).  Semantics kept: L's diagonal is taken as 1 and never
 * read (cape_sptrsv.h:36-47); U is divided by its diagonal (cape_sptrsv.h:95-96).
 * The reference's host wrapper copies b into d_x and then zeroes d_x
 * (cape_sptrsv.h:131,155), so it solves with a zero right-hand side, and never
 * copies the result back despite its comment (cape_sptrsv.h:111); here the
 * intended behaviour is implemented: x = (LU)^{-1} b written back into b.
 *
 * No host synchronisation and no global barrier inside a solve; waits are
 * bounded and report ISAI_ERR_TIMEOUT instead of hanging.
 */
#ifndef ISAI_CAPE_SPTRSV_H
#define ISAI_CAPE_SPTRSV_H

#include "isai_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Host-pointer form with the reference's signature (csr_cape_SpTRSV,
 * cape_sptrsv.h:105-112): CSR of the packed factor (e.g. from isai_bsr_to_csr),
 * b overwritten with the solution.  Synchronous. */
int isai_csr_sptrsv_lu(int m, const int *csrRowPtr, const int *csrColIdx, const double *csrVal,
                       int nnz, double *b);
/* Device form of the same scalar CSR solve; d_work holds
 * isai_csr_sptrsv_work_size(m) bytes; chunk <= 0 picks the default.  Synchronizes
 * to report a timed-out wait. */
int isai_csr_sptrsv_work_size(int m, size_t *bytes);
int isai_csr_sptrsv_lu_device(int m, int nnz, const int *d_rowptr, const int *d_colind,
                              const double *d_val, const double *d_b, double *d_x, void *d_work,
                              int chunk, int spin_limit, isai_stream_t stream);

/* BSR-native form (bs = 4), two schedules:
 *   chunk schedule: one wave per chunk of consecutive block rows whose boundaries
 *     follow the dependency chains; dependencies outside the chunk are fetched
 *     ahead of the serial chain.  A block row may depend on at most 64 blocks
 *     outside its chunk (ISAI_ERR_UNSUPPORTED otherwise; the CSR form has no such
 *     limit);
 *   row schedule: one quad of lanes per block row, 64 rows per workgroup; no
 *     limit on the row length.  In unit order (the default) the rows are grouped
 *     into units of 64 so that most dependencies fall inside a unit: those travel
 *     through LDS, and one wave per workgroup imports the others from the solution
 *     vector.  In level order every dependency is polled from the solution vector.
 *     Level-synchronous units: units of 512 rows (8x8x8 blocks of a 3-D 7-point
 *     grid), each row sorted by its level inside its unit, and the unit's waves
 *     step through those levels together with one barrier per level; rows with
 *     at most 3 dependencies.
 * chunk = 1..128: chunk schedule with at most that many rows per chunk; chunk = -1:
 * row schedule in unit order; -2: in level order; -3: unit order without the
 * import wave (what -1 falls back to when a unit imports more than 512 rows);
 * -4: level-synchronous units (falls back to -1 when a row has more than 3
 * dependencies, a unit imports more than 512 rows or holds levels too wide for it);
 * chunk = 0: when the levels of the triangles average at least 16 block rows, -4 if
 * its units' local levels average at least 16 rows too (3-D stencils), else -1;
 * otherwise the chunk schedule with 128.  The analysis builds the tables from the
 * pattern and pre-scales the off-diagonal blocks by the inverted diagonal parts of
 * the packed factor; isai_bsr_sptrsv_update_values redoes the scaling for new
 * values on the same pattern. */
typedef struct isai_sptrsv_plan *isai_sptrsv_plan_t;
int isai_bsr_sptrsv_analysis(int nb, int nnzb, int bs, const int *d_row, const int *d_col,
                             const double *d_LU, int chunk, isai_sptrsv_plan_t *plan,
                             isai_stream_t stream);
int isai_bsr_sptrsv_update_values(isai_sptrsv_plan_t plan, const double *d_LU, isai_stream_t stream);
/* d_x = U^{-1} L^{-1} d_b (d_x may alias d_b).  Asynchronous. */
int isai_bsr_sptrsv_solve(isai_sptrsv_plan_t plan, const double *d_LU, const double *d_b, double *d_x,
                          isai_stream_t stream);
/* Bound on the polls of one dependency before the solve gives up on it. */
int isai_bsr_sptrsv_set_spin_limit(isai_sptrsv_plan_t plan, int limit);
/* Synchronizes the stream; ISAI_ERR_TIMEOUT if a wait of the previous solves
 * ran out (the error word is then cleared). */
int isai_sptrsv_status(isai_sptrsv_plan_t plan, isai_stream_t stream);
/* The schedule the plan runs, in the chunk argument's terms: -4 level-synchronous
 * units, -1 unit order with the import wave, -3 unit order through the row kernel,
 * -2 level order, or the chunk schedule's rows per chunk (> 0). */
int isai_bsr_sptrsv_schedule(isai_sptrsv_plan_t plan, int *schedule);
int isai_bsr_sptrsv_destroy(isai_sptrsv_plan_t plan);

#ifdef __cplusplus
}
#endif
#endif
