// spmv.hip -- BSR SpMV, ISAI apply, reduced-precision factor storage, BCSC->BCSR
// value permutation and BSR->CSR expansion for gfx950.
//
// bs = 4 layout: one 4x4 FP64 block = 128 B, row-major.  The kernels stream the
// value array with 16-byte lane loads so that the lanes serving one block row
// read whole 128-byte lines:
//   FP64  8 lanes per block row, lane s holds row s/2, columns 2*(s%2)..+1
//   FP32  8 lanes per block row = 2 blocks per line, lane s holds row s%4 of one
//   BF16  8 lanes per block row = 4 blocks per line, lane s holds rows 2(s%2), +1
// Values are read with non-temporal loads (streamed once per product); the x
// gathers use the normal cached path, since each x block is reused by every
// block row that references it.
#include <hip/hip_bf16.h>

#include "bsr2csr.h"
#include "bcsc2bcsr.h"
#include "bsr_tensor_SpMV.h"
#include "common.hpp"
#include "internal.hpp"
#include "isai_runtime.h"

namespace isai {

constexpr int kSpmvBlock = 256;
constexpr int kSpmvF64Block = 128;  // plain FP64 product: 16 block rows per workgroup

// ---------------------------------------------------------------- FP64, bs = 4
// One block row for the 8 lanes s = 0..7 of a lane group (lane s holds row s/2 of
// every block, columns 2(s%2)..+1).  The row is taken KB blocks at a time: lane s
// loads the column index of block t+s (and t+8+s for KB = 16) -- one coalesced load
// per group instead of one per lane and block -- the indices reach the other lanes
// by shuffle, and every value line and x gather of the batch is issued before the
// first product, so that KB lines per lane group are in flight.  All eight lanes of
// a group share the row and take the same path, so the shuffles only read active
// lanes.  Products accumulate block by block into two partial sums per lane, the
// same order for every KB.
// x2: the FP64 vector as pairs; with XF32 it holds FP32 values (pairs of floats),
// each widened exactly to FP64 on load.
template <int KB, bool XF32 = false>
__device__ __forceinline__ double bsr4_row_f64(int rs, int len, int s, const int *__restrict__ ci,
                                               const v2d *__restrict__ val2, const void *__restrict__ x2v) {
    static_assert(KB == 4 || KB == 8 || KB == 16, "KB");
    const v2d *x2 = (const v2d *)x2v;
    const float2 *xf2 = (const float2 *)x2v;
    const int h = s & 1;
    double acc0 = 0.0, acc1 = 0.0;
    for (int t = 0; t < len; t += KB) {
        const int clo = s < KB && t + s < len ? ci[rs + t + s] : 0;
        const int chi = KB > 8 && t + 8 + s < len ? ci[rs + t + 8 + s] : 0;
        v2d a[KB], xv[KB];
        float2 xf[KB];
#pragma unroll
        for (int u = 0; u < KB; ++u)
            if (t + u < len) a[u] = __builtin_nontemporal_load(&val2[(size_t)(rs + t + u) * 8 + s]);
        // every gather of the batch is issued before any of them is used
#pragma unroll
        for (int u = 0; u < KB; ++u) {
            const int c = u < 8 ? __shfl(clo, u, 8) : __shfl(chi, u - 8, 8);
            if (t + u < len) {
                if (XF32) xf[u] = xf2[(size_t)c * 2 + h];
                else xv[u] = x2[(size_t)c * 2 + h];
            }
        }
#pragma unroll
        for (int u = 0; u < KB; ++u)
            if (t + u < len) {
                if (XF32) xv[u] = v2d{(double)xf[u].x, (double)xf[u].y};
                acc0 = fma(a[u].x, xv[u].x, acc0);
                acc1 = fma(a[u].y, xv[u].y, acc1);
            }
    }
    double sum = acc0 + acc1;
    sum += __shfl_xor(sum, 1);
    return sum;
}

// Batch width from the average blocks per row: a batch wider than the rows only
// costs registers (measured on MI355X: 4 for 4-block rows, 8 for 7, 16 for 14 and 27).
inline int f64_batch(long long nnzb, int mb) {
    const long long avg = mb > 0 ? (nnzb + mb - 1) / mb : 0;
    return avg <= 4 ? 4 : avg <= 8 ? 8 : 16;
}

// Workgroups take consecutive block rows in launch order (no XCD remap): streaming
// neighbouring rows from all XCDs at once measured 10-15 % faster than giving each
// XCD a contiguous eighth of the matrix.
template <int KB>
__global__ __launch_bounds__(kSpmvF64Block) void bsr4_spmv_f64_kernel(
    int nb, const int *__restrict__ rp, const int *__restrict__ ci,
    const v2d *__restrict__ val2, const v2d *__restrict__ x2, double *__restrict__ y,
    const int *__restrict__ skip) {
    if (skip && *skip) return;
    const int br = blockIdx.x * (kSpmvF64Block / 8) + (threadIdx.x >> 3);
    const int s = threadIdx.x & 7;
    const int rs = br < nb ? rp[br] : 0;
    const int len = br < nb ? rp[br + 1] - rs : 0;
    const double sum = bsr4_row_f64<KB>(rs, len, s, ci, val2, x2);
    if (br < nb && (s & 1) == 0) y[(size_t)br * 4 + (s >> 1)] = sum;
}

// FP64 product with fused dot products of the result (the Krylov loop's v = A p^
// followed by rhat.v, and t = A s^ followed by t.s and t.t): every workgroup writes
// its partial sums to part[wg] (and part[G + wg]), in a fixed order, so the final
// sum over the G = gridDim.x partials is deterministic.  32 block rows per workgroup
// keep the partial count at nb / 32.
template <int NDOT, int KB, bool XF32>
__global__ __launch_bounds__(kSpmvBlock) void bsr4_spmv_f64_dot_kernel(
    int nb, const int *__restrict__ rp, const int *__restrict__ ci, const v2d *__restrict__ val2,
    const void *__restrict__ x2, double *__restrict__ y, const double *__restrict__ a, double *__restrict__ part,
    const int *__restrict__ skip) {
    if (skip && *skip) return;
    const int wg = blockIdx.x;
    const int s = threadIdx.x & 7;
    const int br = wg * (kSpmvBlock / 8) + (threadIdx.x >> 3);
    const int rs = br < nb ? rp[br] : 0;
    const int len = br < nb ? rp[br + 1] - rs : 0;
    const double sum = bsr4_row_f64<KB, XF32>(rs, len, s, ci, val2, x2);
    double d0 = 0.0, d1 = 0.0;
    if (br < nb && (s & 1) == 0) {
        const size_t row = (size_t)br * 4 + (s >> 1);
        y[row] = sum;
        d0 = sum * a[row];
        if (NDOT == 2) d1 = sum * sum;
    }
    __shared__ double red[2][kSpmvBlock / 64];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        d0 += __shfl_xor(d0, off);
        if (NDOT == 2) d1 += __shfl_xor(d1, off);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = d0;
        red[1][threadIdx.x >> 6] = d1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t0 = 0.0, t1 = 0.0;
#pragma unroll
        for (int w = 0; w < kSpmvBlock / 64; ++w) {
            t0 += red[0][w];
            t1 += red[1][w];
        }
        part[wg] = t0;
        if (NDOT == 2) part[gridDim.x + wg] = t1;
    }
}

int spmv_dot_groups(int mb) { return (int)grid_for(mb, kSpmvBlock / 8); }

template <int NDOT, bool XF32>
static void dot_launch(unsigned g, int kb, int mb, const int *rp, const int *ci, const v2d *val, const void *x,
                       double *y, const double *a, double *part, hipStream_t st, const int *skip) {
    if (kb == 4)
        bsr4_spmv_f64_dot_kernel<NDOT, 4, XF32><<<g, kSpmvBlock, 0, st>>>(mb, rp, ci, val, x, y, a, part, skip);
    else if (kb == 8)
        bsr4_spmv_f64_dot_kernel<NDOT, 8, XF32><<<g, kSpmvBlock, 0, st>>>(mb, rp, ci, val, x, y, a, part, skip);
    else
        bsr4_spmv_f64_dot_kernel<NDOT, 16, XF32><<<g, kSpmvBlock, 0, st>>>(mb, rp, ci, val, x, y, a, part, skip);
}

int spmv_dot_launch(int mb, const int *rp, const int *ci, const double *val, const void *x, bool x_f32, double *y,
                    const double *a, int ndot, double *part, hipStream_t st, const int *skip, long long nnzb) {
    const unsigned g = (unsigned)spmv_dot_groups(mb);
    const int kb = f64_batch(nnzb, mb);
    const v2d *v = (const v2d *)val;
    if (ndot == 1 && x_f32) dot_launch<1, true>(g, kb, mb, rp, ci, v, x, y, a, part, st, skip);
    else if (ndot == 1) dot_launch<1, false>(g, kb, mb, rp, ci, v, x, y, a, part, st, skip);
    else if (x_f32) dot_launch<2, true>(g, kb, mb, rp, ci, v, x, y, a, part, st, skip);
    else dot_launch<2, false>(g, kb, mb, rp, ci, v, x, y, a, part, st, skip);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

// ------------------------------------------------- FP32 / BF16 values, bs = 4
// Eight lanes per block row, each lane 16 B of the row's value array, so one load
// instruction of the eight lanes covers one whole 128-byte line: BPB = 2 (FP32)
// or 4 (BF16) consecutive blocks of the row.  Lane s holds block k + s / LPB and,
// within it, element run (s % LPB) * VPL .. + VPL - 1 (row-major).  Products
// accumulate in FP32; lanes that share a row are summed with cross-lane adds at
// the end (x is FP64 in memory and rounded to FP32 on load).

// x of one block column rounded to FP32, from an FP64 vector (two 16-B loads) or
// from an FP32 vector (one 16-B load); the two give the same floats.
template <bool XF32>
__device__ __forceinline__ void load_x4(const void *__restrict__ x, int c, float f[4]) {
    if (XF32) {
        const v4f v = ((const v4f *)x)[c];
        f[0] = v.x, f[1] = v.y, f[2] = v.z, f[3] = v.w;
    } else {
        const v2d *x2 = (const v2d *)x;
        const v2d x01 = x2[(size_t)c * 2 + 0];
        const v2d x23 = x2[(size_t)c * 2 + 1];
        f[0] = (float)x01.x, f[1] = (float)x01.y, f[2] = (float)x23.x, f[3] = (float)x23.y;
    }
}

// The same four floats for the LPB lanes that share a block (lanes of one quad, or
// one pair of it): each lane loads only its own part of the block column and the
// lanes exchange the rest with DPP moves, instead of every lane loading all of it.
template <int LPB, bool XF32>
__device__ __forceinline__ void load_x4_split(const void *__restrict__ x, int c, int sub, float f[4]) {
    if (LPB == 4) {  // lane sub loads component sub
        const float v = XF32 ? ((const float *)x)[(size_t)c * 4 + sub] : (float)((const double *)x)[(size_t)c * 4 + sub];
        f[0] = quad_bcast_f<0>(v), f[1] = quad_bcast_f<1>(v), f[2] = quad_bcast_f<2>(v), f[3] = quad_bcast_f<3>(v);
    } else {  // LPB == 2: lane sub loads components 2 sub, 2 sub + 1
        float a, b;
        if (XF32) {
            const float2 t = ((const float2 *)x)[(size_t)c * 2 + sub];
            a = t.x, b = t.y;
        } else {
            const v2d t = ((const v2d *)x)[(size_t)c * 2 + sub];
            a = (float)t.x, b = (float)t.y;
        }
        const float pa = pair_swap_f(a), pb = pair_swap_f(b);
        f[0] = sub ? pa : a, f[1] = sub ? pb : b, f[2] = sub ? a : pa, f[3] = sub ? b : pb;
    }
}

template <bool YF32>
__device__ __forceinline__ void store_y(void *__restrict__ y, size_t i, float v) {
    if (YF32)
        ((float *)y)[i] = v;
    else
        ((double *)y)[i] = (double)v;
}

// Each 8-lane group serves kLowpRows block rows (32 consecutive rows of the
// workgroup per step, so x/y accesses stay coalesced) and issues the loads of all
// of them before accumulating: a reduced-precision row carries too few bytes to keep
// enough loads in flight on its own.  Per row the lanes accumulate in the same
// order as with one row per group, so results do not depend on kLowpRows.
// Measured on the 96^3 7-pt ISAI factors (4 blocks per row), M_L and M_U together:
// BF16 60.5 us with 4 rows (63.5 with 2, 66.9 with 8); FP32 on line-aligned rows
// 95.5 us with 3 rows and streamed loads (105.3 with 2 rows through L2, 117.3 with
// 1), on unaligned rows 110.5 us with 2 rows through L2 (126.9 streamed, 121.1 with
// 1 row).  On the 64^3 27-pt factors (14 blocks per row) one
// row already keeps enough loads in flight and several rows per group are slower, so
// the launcher uses several rows only for rows of at most two lines of values
// (kLowpRowsBF16 / kLowpRowsF32).
template <bool BF16, int kLowpRows, bool XF32, bool YF32, bool NT>
__global__ __launch_bounds__(kSpmvBlock) void bsr4_spmv_lowp_kernel(
    int nb, const int *__restrict__ rp, const int *__restrict__ ci, const v4u *__restrict__ val16,
    const void *__restrict__ x, void *__restrict__ y, const int *__restrict__ skip) {
    if (skip && *skip) return;
    constexpr int BPB = BF16 ? 4 : 2;  // blocks per 128-byte line
    constexpr int LPB = 8 / BPB;       // lanes per block
    constexpr int G = kSpmvBlock / 8;  // lane groups per workgroup
    // BF16 workgroups are spread over the XCDs in contiguous runs of rows (58.9 us
    // against 60.3 us in launch order); FP32 runs in launch order (108.9 us against
    // 116.5 us remapped), as the FP64 product does.
    const int wg = BF16 ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int row0 = wg * G * kLowpRows + (threadIdx.x >> 3);
    const int s = threadIdx.x & 7;
    const int b = s / LPB, sub = s % LPB;
    int rs[kLowpRows], re[kLowpRows], lmax = 0;
#pragma unroll
    for (int r = 0; r < kLowpRows; ++r) {
        const int br = row0 + r * G;
        rs[r] = br < nb ? rp[br] : 0;
        re[r] = br < nb ? rp[br + 1] : 0;
        lmax = max(lmax, re[r] - rs[r]);
    }
    float acc0[kLowpRows], acc1[kLowpRows];
#pragma unroll
    for (int r = 0; r < kLowpRows; ++r) acc0[r] = acc1[r] = 0.0f;
    for (int t = 0; t < lmax; t += BPB) {
        v4u w[kLowpRows];
        float f[kLowpRows][4];
        bool live[kLowpRows];
        int c[kLowpRows];
        // With several rows per group the loads are unconditional (a finished row
        // reads block 0 and discards it): loads under a per-row branch made the
        // compiler wait for each row's column index and x in turn, one row after
        // another.  With one row there is nothing to overlap, and the branch saves
        // the loads past the row's end.  The lanes of one block share `live`, so
        // the DPP exchange of x stays inside active lanes.
#pragma unroll
        for (int r = 0; r < kLowpRows; ++r) {
            const int kb = rs[r] + t + b;
            live[r] = kb < re[r];
        }
        if (kLowpRows > 1 || live[0]) {
#pragma unroll
            for (int r = 0; r < kLowpRows; ++r) {
                const int kc = live[r] ? rs[r] + t + b : 0;
                c[r] = ci[kc];
                // 16 B per lane.  NT (streamed): rows start on a line (BF16 rows
                // fill whole lines; the handle's factors are line-aligned by
                // pad_convert_lowp).  Otherwise a 64-B FP32 block may share its
                // line with a neighbour another wave reads, and the line stays in L2.
                w[r] = NT ? __builtin_nontemporal_load(&val16[(size_t)kc * LPB + sub]) : val16[(size_t)kc * LPB + sub];
            }
#pragma unroll
            for (int r = 0; r < kLowpRows; ++r) load_x4_split<LPB, XF32>(x, c[r], sub, f[r]);
        }
#pragma unroll
        for (int r = 0; r < kLowpRows; ++r) {
            if (!live[r]) continue;
            if (BF16) {  // rows 2*sub, 2*sub+1
                acc0[r] = fmaf(bf16_lo(w[r].x), f[r][0], acc0[r]);
                acc0[r] = fmaf(bf16_hi(w[r].x), f[r][1], acc0[r]);
                acc0[r] = fmaf(bf16_lo(w[r].y), f[r][2], acc0[r]);
                acc0[r] = fmaf(bf16_hi(w[r].y), f[r][3], acc0[r]);
                acc1[r] = fmaf(bf16_lo(w[r].z), f[r][0], acc1[r]);
                acc1[r] = fmaf(bf16_hi(w[r].z), f[r][1], acc1[r]);
                acc1[r] = fmaf(bf16_lo(w[r].w), f[r][2], acc1[r]);
                acc1[r] = fmaf(bf16_hi(w[r].w), f[r][3], acc1[r]);
            } else {  // row sub
                acc0[r] = fmaf(__uint_as_float(w[r].x), f[r][0], acc0[r]);
                acc0[r] = fmaf(__uint_as_float(w[r].y), f[r][1], acc0[r]);
                acc0[r] = fmaf(__uint_as_float(w[r].z), f[r][2], acc0[r]);
                acc0[r] = fmaf(__uint_as_float(w[r].w), f[r][3], acc0[r]);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < kLowpRows; ++r) {
        // sum the lanes holding the same row(s): same sub, different b
#pragma unroll
        for (int off = LPB; off < 8; off <<= 1) {
            acc0[r] += __shfl_xor(acc0[r], off);
            if (BF16) acc1[r] += __shfl_xor(acc1[r], off);
        }
        const int br = row0 + r * G;
        if (br < nb && b == 0) {
            if (BF16) {
                store_y<YF32>(y, (size_t)br * 4 + 2 * sub, acc0[r]);
                store_y<YF32>(y, (size_t)br * 4 + 2 * sub + 1, acc1[r]);
            } else {
                store_y<YF32>(y, (size_t)br * 4 + sub, acc0[r]);
            }
        }
    }
}

// BF16 rows longer than two lines: one block row per lane group, NL lines of the
// row per step (KB = 4 NL blocks), with the column indices loaded once per group
// and shuffled as in the FP64 product.  Lane s accumulates blocks b, b + 4, b + 8,
// ... of the row in order, as with one row per group, so results are identical.
// Measured on the 64^3 27-pt ISAI factor (14 blocks per row): 33.3 us with NL = 4
// against 43.4 us with one line per step; on 4-block rows it is slower than the
// several-rows form above, which the launcher keeps for them.
template <int NL, bool XF32, bool YF32>
__global__ __launch_bounds__(kSpmvF64Block) void bsr4_spmv_bf16_batch_kernel(
    int nb, const int *__restrict__ rp, const int *__restrict__ ci, const v4u *__restrict__ val16,
    const void *__restrict__ x, void *__restrict__ y, const int *__restrict__ skip) {
    if (skip && *skip) return;
    constexpr int KB = NL * 4;
    static_assert(KB <= 16, "KB");
    const int br = blockIdx.x * (kSpmvF64Block / 8) + (threadIdx.x >> 3);
    const int s = threadIdx.x & 7;
    const int b = s >> 1, sub = s & 1;
    const int rs = br < nb ? rp[br] : 0;
    const int len = br < nb ? rp[br + 1] - rs : 0;
    float acc0 = 0.0f, acc1 = 0.0f;
    for (int t = 0; t < len; t += KB) {
        const int clo = s < KB && t + s < len ? ci[rs + t + s] : 0;
        const int chi = KB > 8 && t + 8 + s < len ? ci[rs + t + 8 + s] : 0;
        v4u w[NL];
        float f[NL][4];
#pragma unroll
        for (int l = 0; l < NL; ++l) {
            const int kk = t + l * 4 + b;
            if (kk < len) w[l] = __builtin_nontemporal_load(&val16[(size_t)(rs + kk) * 2 + sub]);
        }
#pragma unroll
        for (int l = 0; l < NL; ++l) {
            const int u = l * 4 + b;
            const int clo_u = __shfl(clo, u & 7, 8);
            const int chi_u = KB > 8 ? __shfl(chi, u & 7, 8) : 0;
            const int c = u < 8 ? clo_u : chi_u;
            if (t + u < len) load_x4<XF32>(x, c, f[l]);
        }
#pragma unroll
        for (int l = 0; l < NL; ++l) {
            if (t + l * 4 + b >= len) continue;
            acc0 = fmaf(bf16_lo(w[l].x), f[l][0], acc0);
            acc0 = fmaf(bf16_hi(w[l].x), f[l][1], acc0);
            acc0 = fmaf(bf16_lo(w[l].y), f[l][2], acc0);
            acc0 = fmaf(bf16_hi(w[l].y), f[l][3], acc0);
            acc1 = fmaf(bf16_lo(w[l].z), f[l][0], acc1);
            acc1 = fmaf(bf16_hi(w[l].z), f[l][1], acc1);
            acc1 = fmaf(bf16_lo(w[l].w), f[l][2], acc1);
            acc1 = fmaf(bf16_hi(w[l].w), f[l][3], acc1);
        }
    }
#pragma unroll
    for (int off = 2; off < 8; off <<= 1) {
        acc0 += __shfl_xor(acc0, off);
        acc1 += __shfl_xor(acc1, off);
    }
    if (br < nb && b == 0) {
        store_y<YF32>(y, (size_t)br * 4 + 2 * sub, acc0);
        store_y<YF32>(y, (size_t)br * 4 + 2 * sub + 1, acc1);
    }
}

// ---------------------------------------------------- generic bs (1..8), FP64
__global__ void bsr_spmv_generic_kernel(int nb, int bs, const int *__restrict__ rp,
                                        const int *__restrict__ ci, const double *__restrict__ val,
                                        const double *__restrict__ x, double *__restrict__ y,
                                        const int *__restrict__ skip) {
    if (skip && *skip) return;
    const long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= (long long)nb * bs) return;
    const int br = (int)(row / bs), r = (int)(row % bs), bs2 = bs * bs;
    double sum = 0.0;
    for (int k = rp[br]; k < rp[br + 1]; ++k) {
        const double *blk = val + (size_t)k * bs2 + r * bs;
        const double *xb = x + (size_t)ci[k] * bs;
        for (int c = 0; c < bs; ++c) sum = fma(blk[c], xb[c], sum);
    }
    y[row] = sum;
}

// ------------------------------------------------ value conversion (RNE)
__device__ __forceinline__ unsigned short f32_to_bf16_rne(float f) {
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (unsigned short)((u | 0x00400000u) >> 16);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__global__ void convert_f32_kernel(size_t n, const double *__restrict__ src, float *__restrict__ dst) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = (float)src[i];
}

__global__ void convert_bf16_kernel(size_t n, const double *__restrict__ src, unsigned short *__restrict__ dst) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = f32_to_bf16_rne((float)src[i]);
}

// ------------------------------------------------ line-aligned reduced-precision copy
// Every row of the copy starts on a 128-byte line: its block count is padded up to
// a multiple of the blocks per line (align = 2 FP32, 4 BF16) with blocks of zeros
// whose column repeats the row's last column (any padded row pointer works: the
// FP32 copy of A pads row pairs instead).  A zero block adds exact zeros, so a
// product with the copy equals the product with the unpadded factor; without the
// padding a row starting mid-line shares that line with its neighbour and the line
// is fetched twice.  One 16-lane group per row, lane e converting element e of each
// block (64-B / 32-B coalesced stores).
template <bool BF16>
__global__ void pad_convert_kernel(int nb, const int *__restrict__ rp, const int *__restrict__ ci,
                                   const double *__restrict__ val, const int *__restrict__ prp, int *__restrict__ pci,
                                   void *__restrict__ pval) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int row = (int)(gid >> 4), e = (int)(gid & 15);
    if (row >= nb) return;
    const int s = rp[row], len = rp[row + 1] - s, ps = prp[row], plen = prp[row + 1] - ps;
    for (int j = 0; j < plen; ++j) {
        const bool real = j < len;
        // padding repeats the row's last column (the row index for an empty row)
        if (e == 0) pci[ps + j] = real ? ci[s + j] : len > 0 ? ci[s + len - 1] : row;
        const double v = real ? val[(size_t)(s + j) * 16 + e] : 0.0;
        if (BF16) ((unsigned short *)pval)[(size_t)(ps + j) * 16 + e] = f32_to_bf16_rne((float)v);
        else ((float *)pval)[(size_t)(ps + j) * 16 + e] = (float)v;
    }
}

int pad_convert_lowp(int nb, const int *rp, const int *ci, const double *val, const int *prp, int *pci, void *pval,
                     int prec, hipStream_t st) {
    ISAI_REQUIRE(prec == ISAI_PREC_FP32 || prec == ISAI_PREC_BF16, "pad_convert_lowp: precision %d", prec);
    if (nb == 0) return ISAI_OK;
    const unsigned g = grid_for((long long)nb * 16, 256);
    if (prec == ISAI_PREC_BF16) pad_convert_kernel<true><<<g, 256, 0, st>>>(nb, rp, ci, val, prp, pci, pval);
    else pad_convert_kernel<false><<<g, 256, 0, st>>>(nb, rp, ci, val, prp, pci, pval);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

// ------------------------------------------------ BCSC -> BCSR value permutation
// One thread per (entry, element): reads are fully coalesced, each block's
// writes form one contiguous bs*bs run.
__global__ void reorder_kernel(long long total, int bs2, const int *__restrict__ map,
                               const double *__restrict__ src, double *__restrict__ dst) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const long long k = i / bs2;
    const int e = (int)(i - k * bs2);
    dst[(size_t)map[k] * bs2 + e] = src[i];
}

// ------------------------------------------------ BSR -> CSR (all block entries)
// Scalar row r = br*bs + lr holds its blocks in block-column order and lc =
// 0..bs-1 inside each block; nnz = nnzb*bs*bs.  The CSR entries of block row br
// form one contiguous range [brp[br]*bs2, brp[br+1]*bs2), so one wave per block row
// writes it in order, one entry per lane: the stores are coalesced (one thread per
// scalar row, each writing its own run, took 2.4 ms on the 96^3 7-pt matrix).
constexpr int kB2cBlock = 256;
__global__ __launch_bounds__(kB2cBlock) void bsr2csr_kernel(int nb, int bs, const int *__restrict__ brp,
                                                            const int *__restrict__ bci, const double *__restrict__ bv,
                                                            int *__restrict__ rp, int *__restrict__ ci,
                                                            double *__restrict__ v) {
    const int br = blockIdx.x * (kB2cBlock / 64) + (int)(threadIdx.x >> 6);
    if (br >= nb) return;
    const int lane = threadIdx.x & 63, bs2 = bs * bs;
    const int b0 = brp[br], len = brp[br + 1] - b0;
    const int rowlen = len * bs;  // entries per scalar row of this block row
    const size_t out0 = (size_t)b0 * bs2;
    if (lane < bs) rp[(size_t)br * bs + lane + 1] = (int)(out0 + (size_t)(lane + 1) * rowlen);
    if (br == 0 && lane == 0) rp[0] = 0;
    for (int o = lane; o < len * bs2; o += 64) {
        const int lr = o / rowlen, rem = o - lr * rowlen, q = rem / bs, lc = rem - q * bs;
        ci[out0 + o] = bci[b0 + q] * bs + lc;
        v[out0 + o] = bv[(size_t)(b0 + q) * bs2 + lr * bs + lc];
    }
}

template <bool XF32, bool YF32>
static void lowp_launch(int mb, long long nnzb, bool bf16, bool aligned, const int *rp, const int *ci,
                        const v4u *v16, const void *x, void *y, hipStream_t st, const int *skip) {
    const int bpb = bf16 ? 4 : 2;  // blocks per 128-byte line
    const bool short_rows = nnzb <= (long long)2 * bpb * mb;
    const int rows = short_rows ? (bf16 ? kLowpRowsBF16 : (aligned ? kLowpRowsF32 : 2)) : 1;
    const unsigned g = grid_for(mb, rows * kSpmvBlock / 8);
    if (bf16 && short_rows)
        bsr4_spmv_lowp_kernel<true, kLowpRowsBF16, XF32, YF32, true>
            <<<g, kSpmvBlock, 0, st>>>(mb, rp, ci, v16, x, y, skip);
    else if (bf16)
        bsr4_spmv_bf16_batch_kernel<4, XF32, YF32><<<grid_for(mb, kSpmvF64Block / 8), kSpmvF64Block, 0, st>>>(
            mb, rp, ci, v16, x, y, skip);
    else if (short_rows && aligned)
        bsr4_spmv_lowp_kernel<false, kLowpRowsF32, XF32, YF32, true>
            <<<g, kSpmvBlock, 0, st>>>(mb, rp, ci, v16, x, y, skip);
    else if (short_rows)
        bsr4_spmv_lowp_kernel<false, 2, XF32, YF32, false><<<g, kSpmvBlock, 0, st>>>(mb, rp, ci, v16, x, y, skip);
    else
        bsr4_spmv_lowp_kernel<false, 1, XF32, YF32, false><<<g, kSpmvBlock, 0, st>>>(mb, rp, ci, v16, x, y, skip);
}

int spmv_lowp_launch(int mb, long long nnzb, isai_precision_t prec, const int *rp, const int *ci, const void *val,
                     const void *x, bool x_f32, void *y, bool y_f32, hipStream_t st, const int *skip,
                     bool aligned) {
    ISAI_REQUIRE(prec == ISAI_PREC_FP32 || prec == ISAI_PREC_BF16, "bsr_spmv: reduced precision expected (got %d)",
                 (int)prec);
    if (mb == 0) return ISAI_OK;
    const bool bf16 = prec == ISAI_PREC_BF16;
    const v4u *v16 = (const v4u *)val;
    if (x_f32 && y_f32)
        lowp_launch<true, true>(mb, nnzb, bf16, aligned, rp, ci, v16, x, y, st, skip);
    else if (x_f32)
        lowp_launch<true, false>(mb, nnzb, bf16, aligned, rp, ci, v16, x, y, st, skip);
    else if (y_f32)
        lowp_launch<false, true>(mb, nnzb, bf16, aligned, rp, ci, v16, x, y, st, skip);
    else
        lowp_launch<false, false>(mb, nnzb, bf16, aligned, rp, ci, v16, x, y, st, skip);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

int spmv_launch(int mb, int nnzb, int bs, isai_precision_t prec, const int *rp, const int *ci,
                const void *val, const double *x, double *y, hipStream_t st, const int *skip) {
    ISAI_REQUIRE(mb >= 0 && nnzb >= 0, "bsr_spmv: negative size (mb=%d nnzb=%d)", mb, nnzb);
    ISAI_REQUIRE(bs >= 1 && bs <= 8, "bsr_spmv: bs=%d outside 1..8", bs);
    if (mb == 0) return ISAI_OK;
    ISAI_REQUIRE(rp && ci && val && x && y, "bsr_spmv: null device pointer");
    if (bs == 4) {
        if (prec == ISAI_PREC_FP64) {
            const unsigned g = grid_for(mb, kSpmvF64Block / 8);
            const v2d *v2 = (const v2d *)val, *x2 = (const v2d *)x;
            const int kb = f64_batch(nnzb, mb);
            if (kb == 4)
                bsr4_spmv_f64_kernel<4><<<g, kSpmvF64Block, 0, st>>>(mb, rp, ci, v2, x2, y, skip);
            else if (kb == 8)
                bsr4_spmv_f64_kernel<8><<<g, kSpmvF64Block, 0, st>>>(mb, rp, ci, v2, x2, y, skip);
            else
                bsr4_spmv_f64_kernel<16><<<g, kSpmvF64Block, 0, st>>>(mb, rp, ci, v2, x2, y, skip);
        } else if (prec == ISAI_PREC_FP32 || prec == ISAI_PREC_BF16) {
            return spmv_lowp_launch(mb, nnzb, prec, rp, ci, val, x, false, y, false, st, skip);
        } else {
            set_error("bsr_spmv: unknown precision %d", (int)prec);
            return ISAI_ERR_INVALID_ARG;
        }
    } else {
        if (prec != ISAI_PREC_FP64) {
            set_error("bsr_spmv: reduced precision needs bs = 4 (got %d)", bs);
            return ISAI_ERR_UNSUPPORTED;
        }
        const unsigned g = grid_for((long long)mb * bs, 256);
        bsr_spmv_generic_kernel<<<g, 256, 0, st>>>(mb, bs, rp, ci, (const double *)val, x, y, skip);
    }
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

}  // namespace isai

using namespace isai;

extern "C" {

int isai_bsr_spmv(int mb, int nnzb, int bs, const int *d_rowptr, const int *d_colind,
                  const double *d_val, const double *d_x, double *d_y, isai_stream_t stream) {
    return spmv_launch(mb, nnzb, bs, ISAI_PREC_FP64, d_rowptr, d_colind, d_val, d_x, d_y,
                       (hipStream_t)stream, nullptr);
}

int isai_bsr_spmv_prec(int mb, int nnzb, int bs, isai_precision_t prec, const int *d_rowptr,
                       const int *d_colind, const void *d_val, const double *d_x, double *d_y,
                       isai_stream_t stream) {
    return spmv_launch(mb, nnzb, bs, prec, d_rowptr, d_colind, d_val, d_x, d_y, (hipStream_t)stream, nullptr);
}

int isai_convert_values(size_t count, isai_precision_t prec, const double *d_src, void *d_dst,
                        isai_stream_t stream) {
    if (count == 0) return ISAI_OK;
    ISAI_REQUIRE(d_src && d_dst, "isai_convert_values: null pointer");
    size_t g = (count + 255) / 256;
    if (g > 65536) g = 65536;
    hipStream_t st = (hipStream_t)stream;
    if (prec == ISAI_PREC_FP32) {
        convert_f32_kernel<<<(unsigned)g, 256, 0, st>>>(count, d_src, (float *)d_dst);
    } else if (prec == ISAI_PREC_BF16) {
        convert_bf16_kernel<<<(unsigned)g, 256, 0, st>>>(count, d_src, (unsigned short *)d_dst);
    } else if (prec == ISAI_PREC_FP64) {
        ISAI_HIP(hipMemcpyAsync(d_dst, d_src, count * sizeof(double), hipMemcpyDeviceToDevice, st));
        return ISAI_OK;
    } else {
        set_error("isai_convert_values: unknown precision %d", (int)prec);
        return ISAI_ERR_INVALID_ARG;
    }
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

int isai_apply(int mb, int bs, isai_precision_t prec, const int *d_Lrow, const int *d_Lcol,
               const void *d_Lval, int nnzbL, const int *d_Urow, const int *d_Ucol,
               const void *d_Uval, int nnzbU, const double *d_b, double *d_y, double *d_x,
               isai_stream_t stream) {
    hipStream_t st = (hipStream_t)stream;
    int s = spmv_launch(mb, nnzbL, bs, prec, d_Lrow, d_Lcol, d_Lval, d_b, d_y, st, nullptr);
    if (s != ISAI_OK) return s;
    return spmv_launch(mb, nnzbU, bs, prec, d_Urow, d_Ucol, d_Uval, d_y, d_x, st, nullptr);
}

int isai_reorder_bcsc_to_bcsr(int nnzb, int bs, const int *d_map, const double *d_src,
                              double *d_dst, isai_stream_t stream) {
    ISAI_REQUIRE(nnzb >= 0 && bs >= 1 && bs <= 8, "isai_reorder_bcsc_to_bcsr: bad size");
    if (nnzb == 0) return ISAI_OK;
    ISAI_REQUIRE(d_map && d_src && d_dst, "isai_reorder_bcsc_to_bcsr: null pointer");
    const long long total = (long long)nnzb * bs * bs;
    reorder_kernel<<<grid_for(total, 256), 256, 0, (hipStream_t)stream>>>(total, bs * bs, d_map, d_src, d_dst);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

int isai_bsr_to_csr(int mb, int nnzb, int bs, const int *d_brp, const int *d_bci, const double *d_bv,
                    int *d_rp, int *d_ci, double *d_v, isai_stream_t stream) {
    ISAI_REQUIRE(mb >= 0 && nnzb >= 0 && bs >= 1 && bs <= 8, "isai_bsr_to_csr: bad size");
    ISAI_REQUIRE((long long)nnzb * bs * bs < (1LL << 31), "isai_bsr_to_csr: nnz overflows int32");
    if (mb == 0) return ISAI_OK;
    bsr2csr_kernel<<<grid_for(mb, kB2cBlock / 64), kB2cBlock, 0, (hipStream_t)stream>>>(mb, bs, d_brp, d_bci, d_bv,
                                                                                        d_rp, d_ci, d_v);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

int isai_bsr_spmv_bench(const int *bsrRowPtr, const int *bsrColIdx, const double *bsrVal, int mb,
                        int nnzb, int bdim, const double *Val_B, double *Val_C, int warmup,
                        int iters, double *avg_ms) {
    ISAI_REQUIRE(bsrRowPtr && bsrColIdx && bsrVal && Val_B && Val_C && avg_ms,
                 "isai_bsr_spmv_bench: null pointer");
    ISAI_REQUIRE(mb > 0 && nnzb >= 0 && bdim >= 1 && bdim <= 8 && iters > 0 && warmup >= 0,
                 "isai_bsr_spmv_bench: bad size");
    const size_t bs2 = (size_t)bdim * bdim;
    int *d_rp = nullptr, *d_ci = nullptr;
    double *d_v = nullptr, *d_x = nullptr, *d_y = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int status = ISAI_OK;
    float ms = 0.0f;
#define BENCH_TRY(call)                                                                \
    do {                                                                               \
        hipError_t _e = (call);                                                        \
        if (_e != hipSuccess) { status = hip_fail(_e, #call, __FILE__, __LINE__); goto done; } \
    } while (0)
    BENCH_TRY(hipMalloc(&d_rp, sizeof(int) * (mb + 1)));
    BENCH_TRY(hipMalloc(&d_ci, sizeof(int) * (nnzb ? nnzb : 1)));
    BENCH_TRY(hipMalloc(&d_v, sizeof(double) * (nnzb ? nnzb : 1) * bs2));
    BENCH_TRY(hipMalloc(&d_x, sizeof(double) * mb * bdim));
    BENCH_TRY(hipMalloc(&d_y, sizeof(double) * mb * bdim));
    BENCH_TRY(hipMemcpy(d_rp, bsrRowPtr, sizeof(int) * (mb + 1), hipMemcpyHostToDevice));
    BENCH_TRY(hipMemcpy(d_ci, bsrColIdx, sizeof(int) * nnzb, hipMemcpyHostToDevice));
    BENCH_TRY(hipMemcpy(d_v, bsrVal, sizeof(double) * nnzb * bs2, hipMemcpyHostToDevice));
    BENCH_TRY(hipMemcpy(d_x, Val_B, sizeof(double) * mb * bdim, hipMemcpyHostToDevice));
    BENCH_TRY(hipEventCreate(&e0));
    BENCH_TRY(hipEventCreate(&e1));
    for (int i = 0; i < warmup && status == ISAI_OK; ++i)
        status = spmv_launch(mb, nnzb, bdim, ISAI_PREC_FP64, d_rp, d_ci, d_v, d_x, d_y, nullptr, nullptr);
    if (status != ISAI_OK) goto done;
    BENCH_TRY(hipDeviceSynchronize());
    BENCH_TRY(hipEventRecord(e0, nullptr));
    for (int i = 0; i < iters && status == ISAI_OK; ++i)
        status = spmv_launch(mb, nnzb, bdim, ISAI_PREC_FP64, d_rp, d_ci, d_v, d_x, d_y, nullptr, nullptr);
    if (status != ISAI_OK) goto done;
    BENCH_TRY(hipEventRecord(e1, nullptr));
    BENCH_TRY(hipEventSynchronize(e1));
    BENCH_TRY(hipEventElapsedTime(&ms, e0, e1));
    *avg_ms = (double)ms / iters;
    BENCH_TRY(hipMemcpy(Val_C, d_y, sizeof(double) * mb * bdim, hipMemcpyDeviceToHost));
#undef BENCH_TRY
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(d_rp);
    (void)hipFree(d_ci);
    (void)hipFree(d_v);
    (void)hipFree(d_x);
    (void)hipFree(d_y);
    return status;
}

}  // extern "C"
