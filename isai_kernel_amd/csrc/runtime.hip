// runtime.hip -- error state, device memory / stream / event plumbing, fill kernels.
#include <cstdlib>
#include <cstring>

#include "common.hpp"
#include "internal.hpp"
#include "isai_runtime.h"

namespace isai {

static thread_local char g_last_error[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

void clear_error() { g_last_error[0] = '\0'; }

__global__ void fill_u32_kernel(uint32_t *dst, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = v;
}

__global__ void fill_f64_kernel(double *dst, double v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = v;
}

// Plain streaming copy, 16 B per lane: the reference point for byte counters
// and the empirical bandwidth ceiling.
__global__ void copy_f64_kernel(v2d *__restrict__ dst, const v2d *__restrict__ src, size_t n2) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n2; i += stride) dst[i] = src[i];
}

// Device-to-device copy of doubles with the streaming copy kernel (the runtime's
// blit took 286 us for the 785 MB ILU factor); odd counts or pointers not
// 16-byte aligned go through hipMemcpyAsync.
int copy_f64_d2d(double *dst, const double *src, size_t count, hipStream_t st) {
    if (!count) return ISAI_OK;
    if (count % 2 || ((uintptr_t)dst | (uintptr_t)src) % 16) {
        ISAI_HIP(hipMemcpyAsync(dst, src, sizeof(double) * count, hipMemcpyDeviceToDevice, st));
        return ISAI_OK;
    }
    const size_t n2 = count / 2;
    size_t g = (n2 + 255) / 256;
    if (g > 16384) g = 16384;
    copy_f64_kernel<<<(unsigned)g, 256, 0, st>>>((v2d *)dst, (const v2d *)src, n2);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

static unsigned fill_grid(size_t n) {
    size_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    return (unsigned)(g ? g : 1);
}

}  // namespace isai

using namespace isai;

extern "C" {

const char *isai_last_error(void) { return g_last_error; }

const char *isai_status_string(int s) {
    switch (s) {
        case ISAI_OK: return "ISAI_OK";
        case ISAI_ERR_INVALID_ARG: return "ISAI_ERR_INVALID_ARG";
        case ISAI_ERR_HIP: return "ISAI_ERR_HIP";
        case ISAI_ERR_ALLOC: return "ISAI_ERR_ALLOC";
        case ISAI_ERR_PATTERN: return "ISAI_ERR_PATTERN";
        case ISAI_ERR_ZERO_PIVOT: return "ISAI_ERR_ZERO_PIVOT";
        case ISAI_ERR_NOT_CONVERGED: return "ISAI_ERR_NOT_CONVERGED";
        case ISAI_ERR_TIMEOUT: return "ISAI_ERR_TIMEOUT";
        case ISAI_ERR_CAPACITY: return "ISAI_ERR_CAPACITY";
        case ISAI_ERR_UNSUPPORTED: return "ISAI_ERR_UNSUPPORTED";
        case ISAI_ERR_IO: return "ISAI_ERR_IO";
        default: return "ISAI_ERR_UNKNOWN";
    }
}

int isai_device_count(int *count) {
    ISAI_REQUIRE(count, "isai_device_count: null output");
    ISAI_HIP(hipGetDeviceCount(count));
    return ISAI_OK;
}

int isai_device_info(char *name, int len, int *cu_count, int *xcc_count) {
    int dev = 0;
    ISAI_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    ISAI_HIP(hipGetDeviceProperties(&p, dev));
    if (name && len > 0) {
        snprintf(name, (size_t)len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (xcc_count) *xcc_count = 8;
    return ISAI_OK;
}

int isai_set_device(int device) {
    ISAI_HIP(hipSetDevice(device));
    return ISAI_OK;
}

int isai_device_malloc(void **ptr, size_t bytes) {
    ISAI_REQUIRE(ptr, "isai_device_malloc: null output");
    *ptr = nullptr;
    if (bytes == 0) bytes = 1;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        return ISAI_ERR_ALLOC;
    }
    return ISAI_OK;
}

int isai_device_free(void *ptr) {
    if (ptr) ISAI_HIP(hipFree(ptr));
    return ISAI_OK;
}

int isai_host_free(void *ptr) {
    free(ptr);
    return ISAI_OK;
}

int isai_memcpy_htod(void *dst, const void *src, size_t bytes) {
    if (bytes) ISAI_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return ISAI_OK;
}

int isai_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    if (bytes) ISAI_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return ISAI_OK;
}

int isai_memcpy_dtod_async(void *dst, const void *src, size_t bytes, isai_stream_t stream) {
    if (bytes) ISAI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return ISAI_OK;
}

int isai_fill_u32_async(void *dst, uint32_t value, size_t count, isai_stream_t stream) {
    if (!count) return ISAI_OK;
    ISAI_REQUIRE(dst, "isai_fill_u32_async: null destination");
    fill_u32_kernel<<<fill_grid(count), 256, 0, (hipStream_t)stream>>>((uint32_t *)dst, value, count);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

int isai_fill_f64_async(double *dst, double value, size_t count, isai_stream_t stream) {
    if (!count) return ISAI_OK;
    ISAI_REQUIRE(dst, "isai_fill_f64_async: null destination");
    fill_f64_kernel<<<fill_grid(count), 256, 0, (hipStream_t)stream>>>(dst, value, count);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

int isai_copy_f64_async(double *dst, const double *src, size_t count, isai_stream_t stream) {
    if (!count) return ISAI_OK;
    ISAI_REQUIRE(dst && src && count % 2 == 0, "isai_copy_f64_async: null pointer or odd count");
    const size_t n2 = count / 2;
    size_t g = (n2 + 255) / 256;
    if (g > 16384) g = 16384;
    copy_f64_kernel<<<(unsigned)g, 256, 0, (hipStream_t)stream>>>((v2d *)dst, (const v2d *)src, n2);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

int isai_device_synchronize(void) {
    ISAI_HIP(hipDeviceSynchronize());
    return ISAI_OK;
}

int isai_stream_create(isai_stream_t *stream) {
    ISAI_REQUIRE(stream, "isai_stream_create: null output");
    hipStream_t s;
    ISAI_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (isai_stream_t)s;
    return ISAI_OK;
}

int isai_stream_destroy(isai_stream_t stream) {
    if (stream) ISAI_HIP(hipStreamDestroy((hipStream_t)stream));
    return ISAI_OK;
}

int isai_stream_synchronize(isai_stream_t stream) {
    ISAI_HIP(hipStreamSynchronize((hipStream_t)stream));
    return ISAI_OK;
}

int isai_event_create(void **event) {
    ISAI_REQUIRE(event, "isai_event_create: null output");
    hipEvent_t e;
    ISAI_HIP(hipEventCreate(&e));
    *event = (void *)e;
    return ISAI_OK;
}

int isai_event_destroy(void *event) {
    if (event) ISAI_HIP(hipEventDestroy((hipEvent_t)event));
    return ISAI_OK;
}

int isai_event_record(void *event, isai_stream_t stream) {
    ISAI_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return ISAI_OK;
}

int isai_event_synchronize(void *event) {
    ISAI_HIP(hipEventSynchronize((hipEvent_t)event));
    return ISAI_OK;
}

int isai_event_elapsed_ms(float *ms, void *start, void *stop) {
    ISAI_REQUIRE(ms, "isai_event_elapsed_ms: null output");
    ISAI_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return ISAI_OK;
}

}  // extern "C"
