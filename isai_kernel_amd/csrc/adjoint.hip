// adjoint.hip -- kernels of the differentiation form: BSR transposes (matrix and
// ISAI factors), the packed ILU(0) factor of A^T derived from the factor of A,
// and the pattern-restricted outer-product gradient.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "internal.hpp"
#include "isai_adjoint.h"

namespace isai {

// Counting sort of the blocks by column: trp/tci is the pattern of A^T and
// perm[d] the block of A that lands at position d (rows ascending, so the
// columns of every row of A^T come out ascending).
int transpose_pattern_host(int nb, const int *rp, const int *ci, int *trp, int *tci, int *perm) {
    std::vector<int> next(nb + 1, 0);
    for (int i = 0; i < nb; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            if (ci[k] < 0 || ci[k] >= nb) {
                set_error("transpose: block (%d,%d) has a column outside [0,%d)", i, ci[k], nb);
                return ISAI_ERR_PATTERN;
            }
            next[ci[k] + 1]++;
        }
    trp[0] = 0;
    for (int j = 0; j < nb; ++j) trp[j + 1] = trp[j] + next[j + 1];
    for (int j = 0; j < nb; ++j) next[j] = trp[j];
    for (int i = 0; i < nb; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const int d = next[ci[k]]++;
            tci[d] = i;
            perm[d] = k;
        }
    return ISAI_OK;
}

// dst block d = (src block perm[d])^T; one thread per destination entry, so the
// writes are contiguous.
__global__ void transpose_blocks_kernel(long long total, int bs, const int *__restrict__ perm,
                                        const double *__restrict__ src, double *__restrict__ dst) {
    const int b2 = bs * bs;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long d = e / b2;
        const int w = (int)(e % b2), r = w / bs, c = w % bs;
        dst[e] = src[(size_t)perm[d] * b2 + c * bs + r];
    }
}

int transpose_blocks(long long nnzb, int bs, const int *d_perm, const double *src, double *dst, hipStream_t st) {
    const long long total = nnzb * bs * bs;
    if (total == 0) return ISAI_OK;
    const unsigned grid = (unsigned)std::min<long long>((total + 255) / 256, 1 << 20);
    transpose_blocks_kernel<<<grid, 256, 0, st>>>(total, bs, d_perm, src, dst);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

// Packed ILU(0) factor of A^T from the packed factor P of A (bs = 4), one 16-lane
// group per destination block:
//   I > J: L'_IJ = U_JI^T D_J^{-1}      -> (r,c) = P_JI[c][r] / P_JJ[c][c]
//   I < J: U'_IJ = D_I L_JI^T           -> (r,c) = P_II[r][r] * P_JI[c][r]
//   I = J: strict lower P[c][r] / P[c][c], diagonal P[r][r], strict upper P[r][r] * P[c][r]
__global__ void transpose_packed_lu_kernel(int nnzb, const int *__restrict__ trow_of, const int *__restrict__ tcol,
                                           const int *__restrict__ perm, const int *__restrict__ diag,
                                           const double *__restrict__ lu, double *__restrict__ lut) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)nnzb * 16) return;
    const int d = (int)(e >> 4), w = (int)(e & 15), r = w >> 2, c = w & 3;
    const int I = trow_of[d], J = tcol[d];
    const double *P = lu + (size_t)perm[d] * 16;  // block (J, I) of the factor of A
    double v;
    if (I > J) {
        v = P[c * 4 + r] / lu[(size_t)diag[J] * 16 + c * 5];
    } else if (I < J) {
        v = lu[(size_t)diag[I] * 16 + r * 5] * P[c * 4 + r];
    } else {
        v = r > c ? P[c * 4 + r] / P[c * 5] : (r == c ? P[r * 5] : P[r * 5] * P[c * 4 + r]);
    }
    lut[e] = v;
}

int transpose_packed_lu(int nnzb, const int *d_trow_of, const int *d_tcol, const int *d_perm, const int *d_diag,
                        const double *d_lu, double *d_lut, hipStream_t st) {
    const long long total = (long long)nnzb * 16;
    transpose_packed_lu_kernel<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(nnzb, d_trow_of, d_tcol, d_perm,
                                                                               d_diag, d_lu, d_lut);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

// g_k = alpha * lam_I x_J^T: eight lanes per block row, lane l writes row l/2,
// columns 2(l%2)..+1 of every block of the row (a full 128-byte line per block).
constexpr int kGradBlock = 256;
constexpr int kGradBatch = 8;

__global__ __launch_bounds__(kGradBlock) void outer_grad_kernel(int nb, const int *__restrict__ rp,
                                                                const int *__restrict__ ci,
                                                                const double *__restrict__ lam,
                                                                const double *__restrict__ x, double alpha,
                                                                v2d *__restrict__ g) {
    const int wg = xcd_remap(blockIdx.x, gridDim.x);
    const int I = wg * (kGradBlock / 8) + (threadIdx.x >> 3);
    if (I >= nb) return;
    const int l = threadIdx.x & 7, r = l >> 1, h = l & 1;
    const double a = alpha * lam[(size_t)I * 4 + r];
    const v2d *x2 = (const v2d *)x;
    const int ke = rp[I + 1];
    // eight blocks per pass: their column indices, then their x, all in flight
    // together (one block at a time the gathers waited on each index in turn);
    // positions past the row's end re-read its last block and store nothing
    for (int k0 = rp[I]; k0 < ke; k0 += kGradBatch) {
        int c[kGradBatch];
        v2d xv[kGradBatch];
#pragma unroll
        for (int j = 0; j < kGradBatch; ++j) c[j] = ci[min(k0 + j, ke - 1)];
#pragma unroll
        for (int j = 0; j < kGradBatch; ++j) xv[j] = x2[(size_t)c[j] * 2 + h];
#pragma unroll
        for (int j = 0; j < kGradBatch; ++j)
            if (k0 + j < ke) {
                v2d o;
                o.x = a * xv[j].x;
                o.y = a * xv[j].y;
                __builtin_nontemporal_store(o, &g[(size_t)(k0 + j) * 8 + l]);
            }
    }
}

}  // namespace isai

using namespace isai;

extern "C" {

int isai_bsr_transpose(int mb, int nnzb, int bs, const int *d_row, const int *d_col, const double *d_val,
                       int *d_trow, int *d_tcol, double *d_tval, isai_stream_t stream) {
    ISAI_REQUIRE(d_row && d_col && d_val && d_trow && d_tcol && d_tval, "isai_bsr_transpose: null argument");
    ISAI_REQUIRE(mb > 0 && nnzb >= 0 && bs >= 1 && bs <= 8, "isai_bsr_transpose: bad size mb=%d nnzb=%d bs=%d", mb,
                 nnzb, bs);
    hipStream_t st = (hipStream_t)stream;
    std::vector<int> rp(mb + 1), ci(nnzb), trp(mb + 1), tci(nnzb), perm(nnzb);
    ISAI_HIP(hipMemcpyAsync(rp.data(), d_row, sizeof(int) * (mb + 1), hipMemcpyDeviceToHost, st));
    if (nnzb) ISAI_HIP(hipMemcpyAsync(ci.data(), d_col, sizeof(int) * nnzb, hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipStreamSynchronize(st));
    ISAI_REQUIRE(rp[0] == 0 && rp[mb] == nnzb, "isai_bsr_transpose: row pointer spans [%d,%d), expected [0,%d)",
                 rp[0], rp[mb], nnzb);
    int s = transpose_pattern_host(mb, rp.data(), ci.data(), trp.data(), tci.data(), perm.data());
    if (s != ISAI_OK) return s;
    int *d_perm = nullptr;
    if (hipMalloc(&d_perm, sizeof(int) * std::max(nnzb, 1)) != hipSuccess) {
        set_error("isai_bsr_transpose: allocation failed");
        return ISAI_ERR_ALLOC;
    }
    hipError_t e = hipMemcpyAsync(d_trow, trp.data(), sizeof(int) * (mb + 1), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nnzb) e = hipMemcpyAsync(d_tcol, tci.data(), sizeof(int) * nnzb, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nnzb) e = hipMemcpyAsync(d_perm, perm.data(), sizeof(int) * nnzb, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) s = transpose_blocks(nnzb, bs, d_perm, d_val, d_tval, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d_perm);
    if (e != hipSuccess) return hip_fail(e, "isai_bsr_transpose", __FILE__, __LINE__);
    return s;
}

int isai_bsr_outer_grad(int mb, int nnzb, int bs, const int *d_row, const int *d_col, const double *d_lambda,
                        const double *d_x, double alpha, double *d_g, isai_stream_t stream) {
    ISAI_REQUIRE(d_row && d_col && d_lambda && d_x && d_g, "isai_bsr_outer_grad: null argument");
    ISAI_REQUIRE(mb > 0 && nnzb >= 0, "isai_bsr_outer_grad: bad size mb=%d nnzb=%d", mb, nnzb);
    if (bs != 4) {
        set_error("isai_bsr_outer_grad: bs = 4 only (got %d)", bs);
        return ISAI_ERR_UNSUPPORTED;
    }
    const unsigned grid = grid_for(mb, kGradBlock / 8);
    outer_grad_kernel<<<grid, kGradBlock, 0, (hipStream_t)stream>>>(mb, d_row, d_col, d_lambda, d_x, alpha,
                                                                    (v2d *)d_g);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

}  // extern "C"
