// common.hpp -- shared host/device helpers of libisai_amd.so (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "isai_types.h"

namespace isai {

// Clang vector types (the HIP double2/float4 structs are not accepted by the
// non-temporal load builtins).
typedef double v2d __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

// Records the message of the last failure on this thread (isai_last_error()).
void set_error(const char *fmt, ...);
void clear_error();

inline int hip_fail(hipError_t e, const char *what, const char *file, int line) {
    set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    return ISAI_ERR_HIP;
}

#define ISAI_HIP(call)                                                     \
    do {                                                                   \
        hipError_t _e = (call);                                            \
        if (_e != hipSuccess) return ::isai::hip_fail(_e, #call, __FILE__, __LINE__); \
    } while (0)

#define ISAI_LAUNCH_CHECK()                                                \
    do {                                                                   \
        hipError_t _e = hipGetLastError();                                 \
        if (_e != hipSuccess) return ::isai::hip_fail(_e, "kernel launch", __FILE__, __LINE__); \
    } while (0)

#define ISAI_REQUIRE(cond, ...)                                            \
    do {                                                                   \
        if (!(cond)) {                                                     \
            ::isai::set_error(__VA_ARGS__);                                \
            return ISAI_ERR_INVALID_ARG;                                   \
        }                                                                  \
    } while (0)

// MI355X has 8 XCDs; hardware dispatches consecutive workgroups round-robin over
// them.  Remapping gives each XCD one contiguous range of the logical index
// space, so neighbouring block rows (and the x entries they share) meet in the
// same L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
    constexpr int kXcd = 8;
    const int q = nblocks / kXcd, r = nblocks % kXcd;
    const int xcd = bid % kXcd, local = bid / kXcd;
    return xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

// BF16 halves of a 32-bit word as floats.
__device__ __forceinline__ float bf16_lo(unsigned int w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(unsigned int w) { return __uint_as_float(w & 0xffff0000u); }

// Lane K of each quad to all four lanes of the quad, and the swap of the two lanes
// of each pair (DPP quad_perm [1,0,3,2]).
template <int K>
__device__ __forceinline__ float quad_bcast_f(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), K | (K << 2) | (K << 4) | (K << 6), 0xF, 0xF, false));
}
__device__ __forceinline__ float pair_swap_f(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}

// Block rows per 8-lane group of the reduced-precision products for rows of at most
// two 128-byte lines of values (several rows keep enough loads in flight).
constexpr int kLowpRowsF32 = 3, kLowpRowsBF16 = 4;

inline unsigned grid_for(long long work, int per_block) {
    return (unsigned)((work + per_block - 1) / per_block);
}

}  // namespace isai
