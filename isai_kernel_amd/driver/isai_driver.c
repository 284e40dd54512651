/*
 * isai_driver.c -- C host driver over libisai_amd.so, following the reference
 * driver's flow (main.cu:449-935): read a MatrixMarket matrix (or generate a
 * synthetic stencil), CSR -> BSR, ISAI(L/U) BCSC patterns and BCSR maps, BSR ILU(0)
 * in place, DinvL + pair lists, ISAI(L), DinvU + ISAI(U), BCSC -> BCSR values,
 * one application x = M_U (M_L b) with b = A * x_true (x_true = 0.7, main.cu:795),
 * the host residual check -- and then the preconditioned Krylov solve the
 * reference does not have.
 *
 * usage: isai_driver matrix.mtx [options]
 *        isai_driver --problem P [--L L] [options]
 * options: --bs B (mtx input, default 4)  --solver bicgstab|fgmres
 *          --precision fp64|fp32|bf16  --tol T  --maxit N  --iters N (timing loops)
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bcsc2bcsr.h"
#include "bsr2csr.h"
#include "bsr_tensor_SpMV.h"
#include "bsrilu.h"
#include "csr2bsr.h"
#include "isai_bsr.h"
#include "isai_gen.h"
#include "isai_runtime.h"
#include "isai_solve.h"
#include "read_matrix.h"

#define CHECK(call)                                                                       \
    do {                                                                                  \
        int _s = (call);                                                                  \
        if (_s != ISAI_OK) {                                                              \
            fprintf(stderr, "%s failed: %s: %s (%s:%d)\n", #call, isai_status_string(_s), \
                    isai_last_error(), __FILE__, __LINE__);                               \
            exit(1);                                                                      \
        }                                                                                 \
    } while (0)

static void *dmalloc(size_t bytes) {
    void *p = NULL;
    CHECK(isai_device_malloc(&p, bytes));
    return p;
}

static void *dupload(const void *h, size_t bytes) {
    void *d = dmalloc(bytes);
    CHECK(isai_memcpy_htod(d, h, bytes));
    return d;
}

static void *hmalloc(size_t bytes) {
    void *p = malloc(bytes ? bytes : 1);
    if (!p) {
        fprintf(stderr, "out of host memory\n");
        exit(1);
    }
    return p;
}

/* Average device milliseconds of `iters` calls of fn between two events. */
typedef void (*work_fn)(void *);
static double time_ms(work_fn fn, void *ctx, int iters) {
    void *e0, *e1;
    float ms = 0.f;
    CHECK(isai_event_create(&e0));
    CHECK(isai_event_create(&e1));
    CHECK(isai_device_synchronize());
    CHECK(isai_event_record(e0, NULL));
    for (int i = 0; i < iters; ++i) fn(ctx);
    CHECK(isai_event_record(e1, NULL));
    CHECK(isai_event_synchronize(e1));
    CHECK(isai_event_elapsed_ms(&ms, e0, e1));
    isai_event_destroy(e0);
    isai_event_destroy(e1);
    return ms / iters;
}

struct lower_ctx {
    int nb, bs, Nmax;
    const double *LU;
    const int *Mb, *Mc, *pp, *pl, *ps;
    const double *Dinv;
    double *Mval;
};
static void run_lower(void *c) {
    struct lower_ctx *k = (struct lower_ctx *)c;
    CHECK(isai_build_lower(k->nb, k->bs, k->LU, k->Mb, k->Mc, k->pp, k->pl, k->ps, k->Dinv, k->Mval, k->Nmax, NULL));
}
static void run_upper(void *c) {
    struct lower_ctx *k = (struct lower_ctx *)c;
    CHECK(isai_build_upper(k->nb, k->bs, k->LU, k->Mb, k->Mc, k->pp, k->pl, k->ps, k->Dinv, k->Mval, k->Nmax, NULL));
}
struct apply_ctx {
    int nb, bs, nnzbL, nnzbU;
    const int *Lr, *Lc, *Ur, *Uc;
    const double *Lv, *Uv, *b;
    double *y, *x;
};
static void run_apply(void *c) {
    struct apply_ctx *k = (struct apply_ctx *)c;
    CHECK(isai_apply(k->nb, k->bs, ISAI_PREC_FP64, k->Lr, k->Lc, k->Lv, k->nnzbL, k->Ur, k->Uc, k->Uv, k->nnzbU,
                     k->b, k->y, k->x, NULL));
}

int main(int argc, char **argv) {
    const char *mtx = NULL;
    int problem = -1, L = 0, bs = 4, iters = 10, maxit = 1000, solver = ISAI_SOLVER_BICGSTAB, precision = 0;
    double tol = 1e-10;
    for (int a = 1; a < argc; ++a) {
        if (!strcmp(argv[a], "--problem") && a + 1 < argc) problem = atoi(argv[++a]);
        else if (!strcmp(argv[a], "--L") && a + 1 < argc) L = atoi(argv[++a]);
        else if (!strcmp(argv[a], "--bs") && a + 1 < argc) bs = atoi(argv[++a]);
        else if (!strcmp(argv[a], "--iters") && a + 1 < argc) iters = atoi(argv[++a]);
        else if (!strcmp(argv[a], "--maxit") && a + 1 < argc) maxit = atoi(argv[++a]);
        else if (!strcmp(argv[a], "--tol") && a + 1 < argc) tol = atof(argv[++a]);
        else if (!strcmp(argv[a], "--solver") && a + 1 < argc) solver = !strcmp(argv[++a], "fgmres");
        else if (!strcmp(argv[a], "--precision") && a + 1 < argc) {
            ++a;
            precision = !strcmp(argv[a], "fp32") ? ISAI_PREC_FP32 : !strcmp(argv[a], "bf16") ? ISAI_PREC_BF16 : 0;
        } else if (argv[a][0] != '-') mtx = argv[a];
        else {
            fprintf(stderr, "unknown option %s\n", argv[a]);
            return 2;
        }
    }
    if (!mtx && problem < 0) {
        fprintf(stderr, "Usage: %s matrix.mtx | --problem P [--L L] [options]\n", argv[0]);
        return 0;
    }

    /* 1-2. matrix -> BSR (host) */
    int nb = 0, nnzb = 0, n = 0;
    int *brp = NULL, *bci = NULL;
    double *bv = NULL;
    int *crp = NULL, *cci = NULL;
    double *cv = NULL;
    int csr_nnz = 0;
    if (mtx) {
        int nr, nc;
        CHECK(isai_read_matrix_market(mtx, &nr, &nc, &csr_nnz, &crp, &cci, &cv));
        if (nr != nc) {
            fprintf(stderr, "Matrix is not square, nrows=%d, ncols=%d\n", nr, nc);
            return 0;
        }
        printf("Read matrix %s of size %d x %d, nnz = %d\n", mtx, nr, nc, csr_nnz);
        n = nr;
        CHECK(isai_csr_to_bsr_count(n, bs, crp, cci, &nb, &nnzb));
        brp = hmalloc(sizeof(int) * (nb + 1));
        bci = hmalloc(sizeof(int) * nnzb);
        bv = hmalloc(sizeof(double) * (size_t)nnzb * bs * bs);
        CHECK(isai_csr_to_bsr_fill(n, bs, crp, cci, cv, brp, bci, bv));
    } else {
        bs = 4;
        CHECK(isai_gen_problem_count(problem, L, &nb, &nnzb));
        brp = hmalloc(sizeof(int) * (nb + 1));
        bci = hmalloc(sizeof(int) * nnzb);
        bv = hmalloc(sizeof(double) * (size_t)nnzb * 16);
        CHECK(isai_gen_problem_fill(problem, L, 0, brp, bci, bv));
        n = nb * bs;
        printf("Generated problem %d (L=%d) of size %d x %d\n", problem, L, n, n);
    }
    const size_t b2 = (size_t)bs * bs;
    printf("Converted to BSR with block size %d, nb = %d, nnzb(blocks) = %d\n", bs, nb, nnzb);

    /* 3-4. ISAI patterns (BCSC), triangular BCSR patterns, maps */
    int nnzbL, nnzbU, tL, tU;
    CHECK(isai_bcsc_pattern_count(nb, brp, bci, 1, &nnzbL));
    CHECK(isai_bcsc_pattern_count(nb, brp, bci, 0, &nnzbU));
    int *MbL = hmalloc(sizeof(int) * (nb + 1)), *McL = hmalloc(sizeof(int) * nnzbL);
    int *MbU = hmalloc(sizeof(int) * (nb + 1)), *McU = hmalloc(sizeof(int) * nnzbU);
    CHECK(isai_bcsc_pattern(nb, brp, bci, 1, MbL, McL));
    CHECK(isai_bcsc_pattern(nb, brp, bci, 0, MbU, McU));
    printf("ISAI(L) pattern: nnzb(blocks) = %d\nISAI(U) pattern: nnzb(blocks) = %d\n", nnzbL, nnzbU);
    CHECK(isai_tri_bcsr_pattern_count(nb, brp, bci, 1, &tL));
    CHECK(isai_tri_bcsr_pattern_count(nb, brp, bci, 0, &tU));
    if (tL != nnzbL || tU != nnzbU) {
        fprintf(stderr, "pattern mismatch: BCSR nnzb L/U = %d/%d, BCSC nnzb L/U = %d/%d\n", tL, tU, nnzbL, nnzbU);
        return 1;
    }
    int *TrL = hmalloc(sizeof(int) * (nb + 1)), *TcL = hmalloc(sizeof(int) * tL);
    int *TrU = hmalloc(sizeof(int) * (nb + 1)), *TcU = hmalloc(sizeof(int) * tU);
    CHECK(isai_tri_bcsr_pattern(nb, brp, bci, 1, TrL, TcL));
    CHECK(isai_tri_bcsr_pattern(nb, brp, bci, 0, TrU, TcU));
    int *mapL = hmalloc(sizeof(int) * nnzbL), *mapU = hmalloc(sizeof(int) * nnzbU);
    CHECK(isai_map_bcsc2bcsr(nb, MbL, McL, TrL, TcL, mapL));
    CHECK(isai_map_bcsc2bcsr(nb, MbU, McU, TrU, TcU, mapU));
    int Nmax = 0;
    for (int j = 0; j < nb; ++j) {
        if (MbL[j + 1] - MbL[j] > Nmax) Nmax = MbL[j + 1] - MbL[j];
        if (MbU[j + 1] - MbU[j] > Nmax) Nmax = MbU[j + 1] - MbU[j];
    }

    /* 5. A to the device; the original values are kept for the residual */
    int *d_Ar = dupload(brp, sizeof(int) * (nb + 1));
    int *d_Ac = dupload(bci, sizeof(int) * nnzb);
    double *d_A = dupload(bv, sizeof(double) * b2 * nnzb);
    double *d_LU = dupload(bv, sizeof(double) * b2 * nnzb);

    /* 6. BSR ILU(0) in place */
    int zp = -1;
    int st = isai_bsrilu0_inplace(nb, nnzb, bs, d_Ar, d_Ac, d_LU, &zp, NULL);
    if (st == ISAI_ERR_ZERO_PIVOT) fprintf(stderr, "[bsrilu0] WARNING zero pivot at block row %d\n", zp);
    else CHECK(st);

    /* 7-8. DinvL, pair lists, ISAI(L); DinvU, ISAI(U) */
    int *d_MbL = dupload(MbL, sizeof(int) * (nb + 1)), *d_McL = dupload(McL, sizeof(int) * nnzbL);
    int *d_MbU = dupload(MbU, sizeof(int) * (nb + 1)), *d_McU = dupload(McU, sizeof(int) * nnzbU);
    double *d_DinvL = dmalloc(sizeof(double) * b2 * nb), *d_DinvU = dmalloc(sizeof(double) * b2 * nb);
    double *d_MvL = dmalloc(sizeof(double) * b2 * nnzbL), *d_MvU = dmalloc(sizeof(double) * b2 * nnzbU);
    CHECK(isai_build_dinv(nb, bs, d_Ar, d_Ac, d_LU, 1, d_DinvL, NULL));
    CHECK(isai_build_dinv(nb, bs, d_Ar, d_Ac, d_LU, 0, d_DinvU, NULL));
    int *d_ppL = dmalloc(sizeof(int) * (nnzbL + 1)), *d_ppU = dmalloc(sizeof(int) * (nnzbU + 1));
    int totL = 0, totU = 0;
    CHECK(isai_pairs_count(nb, nnzbL, d_Ar, d_Ac, d_MbL, d_McL, 0, d_ppL, &totL, NULL));
    CHECK(isai_pairs_count(nb, nnzbU, d_Ar, d_Ac, d_MbU, d_McU, 1, d_ppU, &totU, NULL));
    int *d_plL = dmalloc(sizeof(int) * (totL + 1)), *d_psL = dmalloc(sizeof(int) * (totL + 1));
    int *d_plU = dmalloc(sizeof(int) * (totU + 1)), *d_psU = dmalloc(sizeof(int) * (totU + 1));
    CHECK(isai_pairs_fill(nb, d_Ar, d_Ac, d_MbL, d_McL, 0, d_ppL, d_plL, d_psL, NULL));
    CHECK(isai_pairs_fill(nb, d_Ar, d_Ac, d_MbU, d_McU, 1, d_ppU, d_plU, d_psU, NULL));
    printf("Nmax = %d\nMnnz: %d total pairs L: %d, U: %d\n", Nmax, nnzbL, totL, totU);
    struct lower_ctx lc = {nb, bs, Nmax, d_LU, d_MbL, d_McL, d_ppL, d_plL, d_psL, d_DinvL, d_MvL};
    struct lower_ctx uc = {nb, bs, Nmax, d_LU, d_MbU, d_McU, d_ppU, d_plU, d_psU, d_DinvU, d_MvU};
    double t_low = time_ms(run_lower, &lc, iters);
    double t_up = time_ms(run_upper, &uc, iters);
    printf("ISAI(L) kernel avg time: %g ms\nISAI(U) kernel avg time: %g ms\n", t_low, t_up);

    /* 12. BCSC values -> BCSR values */
    int *d_TrL = dupload(TrL, sizeof(int) * (nb + 1)), *d_TcL = dupload(TcL, sizeof(int) * nnzbL);
    int *d_TrU = dupload(TrU, sizeof(int) * (nb + 1)), *d_TcU = dupload(TcU, sizeof(int) * nnzbU);
    int *d_mapL = dupload(mapL, sizeof(int) * nnzbL), *d_mapU = dupload(mapU, sizeof(int) * nnzbU);
    double *d_BvL = dmalloc(sizeof(double) * b2 * nnzbL), *d_BvU = dmalloc(sizeof(double) * b2 * nnzbU);
    CHECK(isai_reorder_bcsc_to_bcsr(nnzbL, bs, d_mapL, d_MvL, d_BvL, NULL));
    CHECK(isai_reorder_bcsc_to_bcsr(nnzbU, bs, d_mapU, d_MvU, d_BvU, NULL));

    /* 13. x = M_U (M_L b) with b = A * 0.7 */
    double *xt = hmalloc(sizeof(double) * n), *hb = hmalloc(sizeof(double) * n), *hx = hmalloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) xt[i] = 0.7;
    double *d_xt = dupload(xt, sizeof(double) * n);
    double *d_b = dmalloc(sizeof(double) * n), *d_y = dmalloc(sizeof(double) * n), *d_x = dmalloc(sizeof(double) * n);
    CHECK(isai_bsr_spmv(nb, nnzb, bs, d_Ar, d_Ac, d_A, d_xt, d_b, NULL));
    CHECK(isai_memcpy_dtoh(hb, d_b, sizeof(double) * n));
    if (bs == 4) {
        struct apply_ctx ac = {nb, bs, nnzbL, nnzbU, d_TrL, d_TcL, d_TrU, d_TcU, d_BvL, d_BvU, d_b, d_y, d_x};
        double t_apply = time_ms(run_apply, &ac, iters);
        printf("ISAI apply (M_L then M_U) time: %g ms\n", t_apply);
    } else {
        CHECK(isai_bsr_spmv(nb, nnzbL, bs, d_TrL, d_TcL, d_BvL, d_b, d_y, NULL));
        CHECK(isai_bsr_spmv(nb, nnzbU, bs, d_TrU, d_TcU, d_BvU, d_y, d_x, NULL));
    }
    CHECK(isai_memcpy_dtoh(hx, d_x, sizeof(double) * n));

    /* 14. host residual check (CSR of A) */
    if (!crp) {
        csr_nnz = nnzb * (int)b2;
        crp = hmalloc(sizeof(int) * (n + 1));
        cci = hmalloc(sizeof(int) * csr_nnz);
        cv = hmalloc(sizeof(double) * csr_nnz);
        int *d_crp = dmalloc(sizeof(int) * (n + 1)), *d_cci = dmalloc(sizeof(int) * csr_nnz);
        double *d_cv = dmalloc(sizeof(double) * csr_nnz);
        CHECK(isai_bsr_to_csr(nb, nnzb, bs, d_Ar, d_Ac, d_A, d_crp, d_cci, d_cv, NULL));
        CHECK(isai_memcpy_dtoh(crp, d_crp, sizeof(int) * (n + 1)));
        CHECK(isai_memcpy_dtoh(cci, d_cci, sizeof(int) * csr_nnz));
        CHECK(isai_memcpy_dtoh(cv, d_cv, sizeof(double) * csr_nnz));
        isai_device_free(d_crp);
        isai_device_free(d_cci);
        isai_device_free(d_cv);
    }
    double ex, rr;
    CHECK(isai_rel_errors_csr(n, crp, cci, cv, hb, xt, hx, &ex, &rr));
    printf("[M_U M_L apply] rel_error_x = %g\n[M_U M_L apply] rel_residual = %g\n", ex, rr);

    /* beyond the reference: the preconditioned solve A x = b */
    if (bs == 4) {
        isai_setup_opts_t opts;
        memset(&opts, 0, sizeof(opts));
        opts.precond = ISAI_PRECOND_ISAI;
        opts.precision = precision;
        isai_handle_t h;
        CHECK(isai_setup(nb, nnzb, bs, d_Ar, d_Ac, d_A, &opts, &h, NULL));
        double ms_ilu, ms_isai, ms_tot;
        CHECK(isai_setup_times(h, &ms_ilu, &ms_isai, &ms_tot));
        CHECK(isai_fill_f64_async(d_x, 0.0, (size_t)n, NULL));
        int it = 0;
        double relres = 0;
        void *e0, *e1;
        float ms = 0.f;
        CHECK(isai_event_create(&e0));
        CHECK(isai_event_create(&e1));
        CHECK(isai_event_record(e0, NULL));
        st = isai_solve(h, solver, 30, tol, maxit, d_b, d_x, &it, &relres, NULL);
        CHECK(isai_event_record(e1, NULL));
        CHECK(isai_event_synchronize(e1));
        CHECK(isai_event_elapsed_ms(&ms, e0, e1));
        if (st != ISAI_OK) fprintf(stderr, "solve: %s: %s\n", isai_status_string(st), isai_last_error());
        CHECK(isai_memcpy_dtoh(hx, d_x, sizeof(double) * n));
        CHECK(isai_rel_errors_csr(n, crp, cci, cv, hb, xt, hx, &ex, &rr));
        printf("[%s, ISAI %s] iterations = %d, solve time = %g ms (setup: ILU(0) %g ms, ISAI %g ms)\n",
               solver ? "FGMRES(30)" : "BiCGSTAB", precision == 1 ? "fp32" : precision == 2 ? "bf16" : "fp64", it,
               ms, ms_ilu, ms_isai);
        printf("[solve] rel_error_x = %g\n[solve] rel_residual = %g\n", ex, rr);
        isai_destroy(h);
        isai_event_destroy(e0);
        isai_event_destroy(e1);
    }
    /* (device buffers are released with the process, as the reference's
     * cudaDeviceReset at main.cu:933 does) */
    free(brp); free(bci); free(bv); free(MbL); free(McL); free(MbU); free(McU);
    free(TrL); free(TcL); free(TrU); free(TcU); free(mapL); free(mapU); free(xt); free(hb); free(hx);
    isai_host_free(crp); isai_host_free(cci); isai_host_free(cv);
    return 0;
}
