// scan.hpp -- device exclusive prefix sum used by the pair-list and level passes.
#pragma once

#include <hip/hip_runtime.h>

namespace isai {

// d_out[i] = sum(d_in[0..i)) for i < n (in place allowed).  Synchronizes the
// stream to report an int32 overflow of the prefix as ISAI_ERR_CAPACITY.
int exclusive_scan_i32(const int *d_in, int *d_out, long long n, hipStream_t st);

// The same scan with caller-owned scratch: ws holds scan_ws_longs(n) int64 words
// (every tile-sum level), *overflow one int; nothing is allocated.  Synchronizes
// the stream like exclusive_scan_i32.
long long scan_ws_longs(long long n);
int exclusive_scan_i32_ws(const int *d_in, int *d_out, long long n, long long *ws, int *overflow, hipStream_t st);
// d[0..n) = v
__global__ void fill_i32_kernel(int *p, long long n, int v);

}  // namespace isai
