// scan.hip -- exclusive prefix sum of int32 counts (replaces thrust::exclusive_scan,
// main.cu:684, and cub::DeviceScan::ExclusiveSum, bcsc2bcsrAndSpMV.h:101-103).
//
// Three passes over tiles of 2048 elements: tile sums, a recursive scan of the
// tile sums, then an in-tile scan plus the tile offset.  Integer addition is
// exact, so the result is bit-identical to any sequential prefix sum.
#include "common.hpp"
#include "scan.hpp"

namespace isai {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ long long block_exclusive_scan(long long v, long long *total) {
    __shared__ long long warp_sums[kScanThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    long long incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        long long o = __shfl_up(incl, off);
        if (lane >= off) incl += o;
    }
    if (lane == 63) warp_sums[wid] = incl;
    __syncthreads();
    long long base = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
        if (w < wid) base += warp_sums[w];
        all += warp_sums[w];
    }
    __syncthreads();
    *total = all;
    return base + incl - v;
}

__global__ __launch_bounds__(kScanThreads) void tile_sum_kernel(const int *__restrict__ in, long long n,
                                                                long long *__restrict__ sums) {
    const long long base = (long long)blockIdx.x * kScanTile;
    long long s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const long long i = base + (long long)threadIdx.x * kScanItems + k;
        if (i < n) s += in[i];
    }
    long long total;
    block_exclusive_scan(s, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanThreads) void tile_scan_i64_kernel(long long *__restrict__ data, long long n,
                                                                     long long *__restrict__ sums) {
    const long long base = (long long)blockIdx.x * kScanTile;
    long long v[kScanItems];
    long long s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const long long i = base + (long long)threadIdx.x * kScanItems + k;
        v[k] = i < n ? data[i] : 0;
        s += v[k];
    }
    long long total;
    long long off = block_exclusive_scan(s, &total);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const long long i = base + (long long)threadIdx.x * kScanItems + k;
        if (i < n) data[i] = off;
        off += v[k];
    }
    if (threadIdx.x == 0 && sums) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanThreads) void add_tile_offsets_i64_kernel(long long *__restrict__ data, long long n,
                                                                            const long long *__restrict__ offs) {
    const long long base = (long long)blockIdx.x * kScanTile;
    const long long o = offs[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const long long i = base + (long long)threadIdx.x * kScanItems + k;
        if (i < n) data[i] += o;
    }
}

__global__ __launch_bounds__(kScanThreads) void final_scan_kernel(const int *__restrict__ in, int *__restrict__ out,
                                                                  long long n, const long long *__restrict__ offs,
                                                                  int *__restrict__ overflow) {
    const long long base = (long long)blockIdx.x * kScanTile;
    int v[kScanItems];
    long long s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const long long i = base + (long long)threadIdx.x * kScanItems + k;
        v[k] = i < n ? in[i] : 0;
        s += v[k];
    }
    long long total;
    long long off = block_exclusive_scan(s, &total) + offs[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const long long i = base + (long long)threadIdx.x * kScanItems + k;
        if (i < n) {
            if (off > 0x7fffffffLL) *overflow = 1;
            out[i] = (int)off;
        }
        off += v[k];
    }
}

// Exclusive scan of a device array of tile sums, in place (int64), recursively.
static int scan_i64_inplace(long long *d, long long n, hipStream_t st) {
    if (n <= 1) {
        if (n == 1) ISAI_HIP(hipMemsetAsync(d, 0, sizeof(long long), st));
        return ISAI_OK;
    }
    const long long tiles = (n + kScanTile - 1) / kScanTile;
    long long *sums = nullptr;
    if (tiles > 1) ISAI_HIP(hipMallocAsync((void **)&sums, sizeof(long long) * tiles, st));
    tile_scan_i64_kernel<<<(unsigned)tiles, kScanThreads, 0, st>>>(d, n, sums);
    ISAI_LAUNCH_CHECK();
    if (tiles > 1) {
        int s = scan_i64_inplace(sums, tiles, st);
        if (s != ISAI_OK) return s;
        add_tile_offsets_i64_kernel<<<(unsigned)tiles, kScanThreads, 0, st>>>(d, n, sums);
        ISAI_LAUNCH_CHECK();
        ISAI_HIP(hipFreeAsync(sums, st));
    }
    return ISAI_OK;
}

// ------------------------------------------------ scan with caller-owned scratch
// The same three passes with every tile-sum level in a workspace the caller owns
// (scan_ws_longs(n) int64 words) and the overflow flag at *overflow: no allocation,
// no memset (a fill kernel clears the flag and a one-tile level).
__global__ void fill_i32_kernel(int *p, long long n, int v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void zero_i64_kernel(long long *p) { *p = 0; }

long long scan_ws_longs(long long n) {
    long long total = 0;
    for (long long t = (n + kScanTile - 1) / kScanTile;; t = (t + kScanTile - 1) / kScanTile) {
        total += t;
        if (t <= 1) break;
    }
    return total;
}

static int scan_i64_inplace_ws(long long *d, long long n, long long *ws, hipStream_t st) {
    if (n <= 1) {
        if (n == 1) {
            zero_i64_kernel<<<1, 1, 0, st>>>(d);
            ISAI_LAUNCH_CHECK();
        }
        return ISAI_OK;
    }
    const long long tiles = (n + kScanTile - 1) / kScanTile;
    long long *sums = tiles > 1 ? ws : nullptr;
    tile_scan_i64_kernel<<<(unsigned)tiles, kScanThreads, 0, st>>>(d, n, sums);
    ISAI_LAUNCH_CHECK();
    if (tiles > 1) {
        int s = scan_i64_inplace_ws(sums, tiles, ws + tiles, st);
        if (s != ISAI_OK) return s;
        add_tile_offsets_i64_kernel<<<(unsigned)tiles, kScanThreads, 0, st>>>(d, n, sums);
        ISAI_LAUNCH_CHECK();
    }
    return ISAI_OK;
}

int exclusive_scan_i32_ws(const int *d_in, int *d_out, long long n, long long *ws, int *overflow, hipStream_t st) {
    if (n <= 0) return ISAI_OK;
    const long long tiles = (n + kScanTile - 1) / kScanTile;
    fill_i32_kernel<<<1, 64, 0, st>>>(overflow, 1, 0);
    ISAI_LAUNCH_CHECK();
    tile_sum_kernel<<<(unsigned)tiles, kScanThreads, 0, st>>>(d_in, n, ws);
    ISAI_LAUNCH_CHECK();
    int s = scan_i64_inplace_ws(ws, tiles, ws + tiles, st);
    if (s != ISAI_OK) return s;
    final_scan_kernel<<<(unsigned)tiles, kScanThreads, 0, st>>>(d_in, d_out, n, ws, overflow);
    ISAI_LAUNCH_CHECK();
    int h_overflow = 0;
    ISAI_HIP(hipMemcpyAsync(&h_overflow, overflow, sizeof(int), hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipStreamSynchronize(st));
    if (h_overflow) {
        set_error("exclusive scan: prefix exceeds int32");
        return ISAI_ERR_CAPACITY;
    }
    return ISAI_OK;
}

int exclusive_scan_i32(const int *d_in, int *d_out, long long n, hipStream_t st) {
    if (n <= 0) return ISAI_OK;
    const long long tiles = (n + kScanTile - 1) / kScanTile;
    long long *sums = nullptr;
    int *overflow = nullptr;
    ISAI_HIP(hipMallocAsync((void **)&sums, sizeof(long long) * tiles, st));
    ISAI_HIP(hipMallocAsync((void **)&overflow, sizeof(int), st));
    ISAI_HIP(hipMemsetAsync(overflow, 0, sizeof(int), st));
    tile_sum_kernel<<<(unsigned)tiles, kScanThreads, 0, st>>>(d_in, n, sums);
    ISAI_LAUNCH_CHECK();
    int s = scan_i64_inplace(sums, tiles, st);
    if (s != ISAI_OK) return s;
    final_scan_kernel<<<(unsigned)tiles, kScanThreads, 0, st>>>(d_in, d_out, n, sums, overflow);
    ISAI_LAUNCH_CHECK();
    int h_overflow = 0;
    ISAI_HIP(hipMemcpyAsync(&h_overflow, overflow, sizeof(int), hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipStreamSynchronize(st));
    ISAI_HIP(hipFreeAsync(sums, st));
    ISAI_HIP(hipFreeAsync(overflow, st));
    if (h_overflow) {
        set_error("exclusive scan: prefix exceeds int32");
        return ISAI_ERR_CAPACITY;
    }
    return ISAI_OK;
}

}  // namespace isai
