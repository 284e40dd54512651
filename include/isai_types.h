/*
 * isai_types.h -- status codes, precision enum and stream handle shared by every
 * entry point of libisai_amd.so.
 *
 * The reference mixes three error styles (exit() in isai_bsr.h:7-15,
 * bsrilu.h:9-16, csr2bsr.h:28, bcsc2bcsr.h:55; return -1 in
 * csr_cusp_sptrsv.h:17-37; print-and-continue in bsrilu.h:46-60).  Every entry
 * point here returns an int status instead and never exits; isai_last_error()
 * gives the message of the last failure on the calling thread.
 *
 * Conventions of the reference kept unchanged: 0-based int32 indices, row-major
 * bs x bs blocks, BCSC ISAI storage as (Mbrow = column pointer, Mcol = row
 * index), packed ILU(0) in A's BSR arrays in place, caller-owned device buffers.
 */
#ifndef ISAI_TYPES_H
#define ISAI_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    ISAI_OK = 0,
    ISAI_ERR_INVALID_ARG = -1,
    ISAI_ERR_HIP = -2,
    ISAI_ERR_ALLOC = -3,
    ISAI_ERR_PATTERN = -4,       /* bcsc2bcsr.h:53-56 "Pattern mismatch" */
    ISAI_ERR_ZERO_PIVOT = -5,    /* bsrilu.h:46-60 structural / numerical zero */
    ISAI_ERR_NOT_CONVERGED = -6,
    ISAI_ERR_TIMEOUT = -7,       /* a bounded spin of a sync-free kernel ran out */
    ISAI_ERR_CAPACITY = -8,      /* a fixed-size buffer would overflow */
    ISAI_ERR_UNSUPPORTED = -9,   /* e.g. a block size without a kernel */
    ISAI_ERR_IO = -10            /* read_matrix.h:19-40 open / parse failures */
} isai_status_t;

/* Storage precision of ISAI factor values (the reference is FP64 only under
 * USE_DOUBLE, common.h:18-22).  Vectors stay FP64 in every form. */
typedef enum {
    ISAI_PREC_FP64 = 0,
    ISAI_PREC_FP32 = 1,
    ISAI_PREC_BF16 = 2
} isai_precision_t;

/* A hipStream_t (NULL = the default stream), declared without HIP headers so
 * that plain C callers can use the library. */
typedef struct ihipStream_t *isai_stream_t;

/* Message of the last failing call on this thread ("" when none). */
const char *isai_last_error(void);
/* Name of a status code. */
const char *isai_status_string(int status);

#ifdef __cplusplus
}
#endif
#endif
