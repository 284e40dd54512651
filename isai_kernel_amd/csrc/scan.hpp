// scan.hpp -- device exclusive prefix sum used by the pair-list and level passes.
#pragma once

#include <hip/hip_runtime.h>

namespace isai {

// d_out[i] = sum(d_in[0..i)) for i < n (in place allowed).  Synchronizes the
// stream to report an int32 overflow of the prefix as ISAI_ERR_CAPACITY.
int exclusive_scan_i32(const int *d_in, int *d_out, long long n, hipStream_t st);

}  // namespace isai
