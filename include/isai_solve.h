/*
 * isai_solve.h -- the preconditioned solve that the reference's driver stops short
 * of: main.cu:449-866 factors, builds ISAI(L/U) and applies x = M_U (M_L b) once;
 * it has no Krylov method (no BiCGSTAB / GMRES anywhere in the reference).
 *
 * A handle owns everything the setup produces from a BSR matrix A on the device:
 * the packed ILU(0) factor (A is not modified), the ISAI factors in BCSR (FP64 and
 * the selected storage precision), the sync-free triangular-solve plan, and the
 * Krylov work vectors.  The Krylov methods run in FP64 (matrix products with A,
 * residuals, dot products, orthogonalisation); only the preconditioner values may
 * be stored in FP32 / BF16 (applied with FP32 accumulation).
 *
 *   BiCGSTAB: van der Vorst (1992), right preconditioning (Saad, Iterative
 *             Methods for Sparse Linear Systems, 2nd ed., Alg. 7.7).
 *   FGMRES(m): flexible GMRES (Saad Alg. 9.6) with classical Gram-Schmidt and a
 *              second pass when the first one cancelled (DGKS criterion).
 * Scalars (alpha, beta, omega, Hessenberg entries, Givens rotations) live on the
 * device; the host reads a convergence word once per few iterations.
 */
#ifndef ISAI_SOLVE_H
#define ISAI_SOLVE_H

#include "isai_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    ISAI_PRECOND_NONE = 0,
    ISAI_PRECOND_ISAI = 1,        /* z = M_U (M_L v) */
    ISAI_PRECOND_ILU_SPTRSV = 2   /* z = U^{-1} L^{-1} v, sync-free solves */
} isai_precond_kind_t;

/* ISAI_SOLVER_FGMRES_MIXED: the mixed-precision form.  An FP64 FGMRES(restart)
 * outer loop (residual r = b - A x, Arnoldi products with A, Gram-Schmidt and
 * Givens rotations in FP64) whose flexible preconditioner is an inner BiCGSTAB in
 * FP32 storage: A rounded to FP32 (FP32 accumulation), the ISAI factors in the
 * handle's storage precision (FP32 when that is FP64), FP32 vectors, dot products
 * accumulated in FP64.  Each inner solve is asked for the reduction still missing
 * (see isai_set_inner_solve); *iters counts inner iterations and maxit bounds them.
 * Needs the ISAI preconditioner.
 * ISAI_SOLVER_FGMRES_MIXED_GMRES: the same FP64 FGMRES(restart) outer loop whose
 * flexible preconditioner is one GMRES cycle of at most inner_maxit (<= 64) steps
 * in FP32 storage: FP32 basis vectors, A rounded to FP32, the ISAI factors as
 * above, one classical Gram-Schmidt pass with dots and updates accumulated in
 * FP64, FP64 Hessenberg and
 * Givens rotations; the cycle stops early on the same inner tolerance rule.  It
 * needs no host read per inner step; *iters counts inner Arnoldi steps. */
typedef enum {
    ISAI_SOLVER_BICGSTAB = 0,
    ISAI_SOLVER_FGMRES = 1,
    ISAI_SOLVER_FGMRES_MIXED = 2,
    ISAI_SOLVER_FGMRES_MIXED_GMRES = 3
} isai_solver_kind_t;

typedef struct {
    int precond;      /* isai_precond_kind_t */
    int precision;    /* isai_precision_t of the ISAI factor values */
    int sptrsv_chunk; /* schedule of the sync-free solve, as isai_bsr_sptrsv_analysis's chunk (0 = automatic) */
    int reserved[5];
} isai_setup_opts_t;

typedef struct isai_handle *isai_handle_t;

/* BiCGSTAB solves of this handle replay their iterations from a captured hipGraph
 * (three iterations per replay, captured on the solve's stream on first use and
 * again when the solution vector, the preconditioner or its precision changes)
 * instead of launching them; the iterates are the same bit for bit.  The solve's
 * stream must be a created (non-null) stream; with the null stream the iterations
 * are launched as before.  No reference counterpart (the reference has no Krylov
 * method). */
int isai_set_solve_graph(isai_handle_t handle, int use_graph);

/* Runs ILU(0) on a copy of A, the ISAI patterns / pair lists / Dinv / batched
 * ISAI build / BCSR permutation, the precision conversion and (for
 * ISAI_PRECOND_ILU_SPTRSV) the triangular-solve analysis.  Synchronous.
 * A's device arrays must stay valid for the life of the handle. */
int isai_setup(int nb, int nnzb, int bs, const int *d_row, const int *d_col, const double *d_val,
               const isai_setup_opts_t *opts, isai_handle_t *handle, isai_stream_t stream);
/* Device time of the setup phases (hipEvents): ILU(0) factorization, ISAI
 * construction (pairs + Dinv + build + permutation + conversion), total. */
int isai_setup_times(isai_handle_t handle, double *ms_ilu, double *ms_isai, double *ms_total);
/* Switch the preconditioner (kind, storage precision) without refactoring. */
int isai_set_preconditioner(isai_handle_t handle, int precond, int precision, isai_stream_t stream);
/* Sizes of the ISAI factors (block counts). */
int isai_handle_info(isai_handle_t handle, int *nnzbL, int *nnzbU, int *nlevels);
/* Device pointers of the factors (BCSR of M_L / M_U, values in FP64, and the
 * packed ILU(0) factor with A's pattern); for tests and custom pipelines. */
int isai_handle_factors(isai_handle_t handle, const int **Lrow, const int **Lcol, const double **Lval,
                        const int **Urow, const int **Ucol, const double **Uval, const double **LU);

/* d_z = M d_v with the handle's current preconditioner.  Asynchronous; with the
 * ILU(0) preconditioner a triangular-solve wait that ran out is reported by
 * isai_handle_status (and by the next isai_solve). */
int isai_precond_apply(isai_handle_t handle, const double *d_v, double *d_z, isai_stream_t stream);

/* Synchronizes the stream; ISAI_ERR_TIMEOUT if a sync-free triangular solve of the
 * handle ran out of its wait budget since the last check (the flag is then cleared). */
int isai_handle_status(isai_handle_t handle, isai_stream_t stream);

/* Bound on the polls of one dependency of the handle's sync-free triangular solves
 * (ILU(0) preconditioner); see isai_bsr_sptrsv_set_spin_limit. */
int isai_handle_set_spin_limit(isai_handle_t handle, int limit);

/* The ISAI build plan isai_setup derived from the pair statistics (the choice the
 * reference makes by hand from analyze_pair.h's output and the hybrid cache
 * switch of cacheY_isai.h:192): *ncache = Y blocks per column kept in LDS (columns
 * longer than that read Y back from global memory; 0 = every column), and
 * *ordered_L / *ordered_U = 1 when that factor's columns were scheduled by
 * descending pair count Q_j because the Q distribution is skewed. */
int isai_handle_build_plan(isai_handle_t handle, int *ncache, int *ordered_L, int *ordered_U);

/* Solve A x = b.  d_x holds the initial guess on entry.  Stops when the
 * recursive residual estimate reaches tol * ||b|| (then confirmed with the true
 * FP64 residual, iterating on if it is not met) or after maxit iterations.
 * restart = FGMRES cycle length (ignored by BiCGSTAB).  *iters = iterations
 * (matrix products with A), *relres = ||b - A x|| / ||b|| computed in FP64.
 * Returns ISAI_ERR_NOT_CONVERGED when maxit is reached, ISAI_ERR_TIMEOUT when a
 * sync-free triangular solve of the ILU(0) preconditioner ran out of its wait budget. */
int isai_solve(isai_handle_t handle, int solver, int restart, double tol, int maxit,
               const double *d_b, double *d_x, int *iters, double *relres, isai_stream_t stream);

/* Inner solve of the mixed-precision solvers: an inner solve stops when its residual
 * has fallen by max(inner_tol, 0.5 tol ||b|| / (current outer residual)) (at most
 * 0.5) or after inner_maxit iterations (GMRES cycle: steps, at most 64).  Defaults
 * 3e-5 and 40. */
int isai_set_inner_solve(isai_handle_t handle, double inner_tol, int inner_maxit);
/* Outer FGMRES steps and inner iterations of the last mixed-precision solve (0 and
 * the iteration count for the other solvers). */
int isai_last_solve_stats(isai_handle_t handle, int *outer_steps, int *inner_iters);

/* The FP32 product with A that the handle's mixed-precision solves run (bs = 4; no
 * reference counterpart): d_y = A d_x with FP32 x and y, FP32 accumulation, and the
 * fused FP64 dots of isai_bsr_spmv_f32_dot (*groups = G, d_part holds 2 G doubles;
 * ndot 0 = no dots, d_a and d_part unused).  It multiplies with the handle's FP32
 * copy of A (built on first use), whose rows are laid out on 128-byte lines and
 * streamed with non-temporal loads: one row per lane group, every row padded with
 * zero blocks to an even count, or row pairs stored from a line with the pair's
 * second row padded (*rows_per_group = 1 or 2, may be NULL).  y equals
 * isai_bsr_spmv_f32_dot's on the unpadded FP32 values; with row pairs a
 * workgroup's partials cover other rows, so the dots differ in rounding. */
int isai_handle_spmv_f32_dot(isai_handle_t handle, const float *d_x, float *d_y, const float *d_a, int ndot,
                             double *d_part, int *groups, int *rows_per_group, isai_stream_t stream);

/* Inspection of the last FGMRES cycle (no reference counterpart; for tests of the
 * Arnoldi relation A Z_k = V_{k+1} Hbar_k): *m = the cycle length of the last
 * FGMRES solve, *steps = the Arnoldi steps k of its last cycle, *d_V = the m + 1
 * basis vectors and *d_Z = the m preconditioned vectors (device, n doubles each,
 * consecutive), Hbar = the (m + 1) x m Hessenberg matrix before the Givens
 * rotations, row-major (capacity >= (m + 1) m doubles; may be null).  For the
 * mixed-precision solver Z holds the inner solutions.  Synchronizes the device. */
int isai_fgmres_arnoldi(isai_handle_t handle, int *m, int *steps, const double **d_V, const double **d_Z,
                        double *Hbar, int capacity);

int isai_destroy(isai_handle_t handle);

#ifdef __cplusplus
}
#endif
#endif
