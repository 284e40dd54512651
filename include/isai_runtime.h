/*
 * isai_runtime.h -- device memory, stream and event plumbing for C and ctypes
 * callers (the reference does this inline with cudaMalloc/cudaMemcpy/cudaEvent in
 * main.cu:465-474,556-652,822-841).  No reference counterpart as an API.
 */
#ifndef ISAI_RUNTIME_H
#define ISAI_RUNTIME_H

#include "isai_types.h"

#ifdef __cplusplus
extern "C" {
#endif

int isai_device_count(int *count);
/* Name and compute-unit count of the current device (buf gets at most len bytes). */
int isai_device_info(char *name, int len, int *cu_count, int *xcc_count);
int isai_set_device(int device);

int isai_device_malloc(void **ptr, size_t bytes);
int isai_device_free(void *ptr);
int isai_host_free(void *ptr); /* frees host arrays the library malloc'ed */

int isai_memcpy_htod(void *dst, const void *src, size_t bytes);
int isai_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int isai_memcpy_dtod_async(void *dst, const void *src, size_t bytes, isai_stream_t stream);
/* Fills with a kernel (not hipMemset), so it is safe inside a captured graph. */
int isai_fill_u32_async(void *dst, uint32_t value, size_t count, isai_stream_t stream);
int isai_fill_f64_async(double *dst, double value, size_t count, isai_stream_t stream);
/* dst[0..count) = src[0..count) with a plain streaming kernel (count even). */
int isai_copy_f64_async(double *dst, const double *src, size_t count, isai_stream_t stream);

int isai_device_synchronize(void);
int isai_stream_create(isai_stream_t *stream);
int isai_stream_destroy(isai_stream_t stream);
int isai_stream_synchronize(isai_stream_t stream);

int isai_event_create(void **event);
int isai_event_destroy(void *event);
int isai_event_record(void *event, isai_stream_t stream);
int isai_event_synchronize(void *event);
int isai_event_elapsed_ms(float *ms, void *start, void *stop);

#ifdef __cplusplus
}
#endif
#endif
