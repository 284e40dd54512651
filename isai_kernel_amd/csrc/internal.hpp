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

// Sync-free L then U solve of a plan, with the same skip convention.
int sptrsv_solve_internal(isai_sptrsv_plan_t plan, const double *d_LU, const double *d_b, double *d_x,
                          hipStream_t st, const int *skip);

struct HandleFactors {
    int nb, nnzb, nnzbL, nnzbU;
    const int *Lrow, *Lcol, *Urow, *Ucol;
    const void *Lval, *Uval;  // in the handle's current precision
    int precision;
};
// Factors of a setup handle (isai_handle_t) in its current storage precision.
int handle_factors_internal(isai_handle_t h, HandleFactors *out);

}  // namespace isai
