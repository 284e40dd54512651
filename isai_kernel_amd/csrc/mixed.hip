// mixed.hip -- the inner solve of the mixed-precision form on gfx950: BiCGSTAB
// with FP32 storage (A rounded to FP32, the handle's FP32 / BF16 ISAI factors,
// FP32 vectors) and FP64 dot products, used as the flexible preconditioner of an
// FP64 FGMRES outer loop.  No reference counterpart: the reference applies its
// ISAI factors once (main.cu:817-849) and has no Krylov method.
//
// Scalars without single-block kernels: the A products run on a fixed grid of
// kMixedGroups workgroups (grid-stride over 32-row chunks), so each product leaves
// kMixedGroups partial sums per dot, and the consumer kernel of a scalar (alpha in
// the s update, omega in the x/r update, rho and beta in the p update) has every
// workgroup sum those partials itself, in the same fixed order, before it streams
// its vectors.  Every workgroup therefore computes bit-identical scalars; the
// first workgroup also records them for the later kernels.  A scalar record field
// is only written by a kernel that does not read it.
#include "bsr_tensor_SpMV.h"
#include "common.hpp"
#include "internal.hpp"

namespace isai {

constexpr int kMixedBlock = 256;       // A products: 32 block rows per chunk
constexpr int kMixedVecThreads = 1024; // vector kernels
constexpr int kMixedVecBlocks = 512;

// -------------------------------------------------- FP32 A product, bs = 4
// Eight lanes per lane group; lane s loads 16 B = row s%4 of block 2j + s/4 of the
// batch, so one load instruction of the eight lanes reads two consecutive blocks
// (one 128-byte line when the pair is line-aligned).  The lanes of one block (a
// quad) each gather one FP32 component of the block's x column and exchange the
// rest with DPP moves.  The batch's column indices are loaded once per lane group
// and handed out by shuffle.  All value loads and x gathers of a batch are issued
// before the first product and are unconditional (a lane past the range's end
// reads block 0 and discards it): loads under a per-block branch were waited for
// one at a time.  The grid strides over chunks of 32 R block rows.  FP32
// accumulation: lane (b, sub) sums the blocks of a row at even (b = 0) or odd
// (b = 1) positions of its range, in order, and the two lanes are added at the end.
//
// R = rows per lane group, and the value layout:
//  * R = 1, NT = false: any row layout (the caller's arrays); values through L2.
//  * R = 1, NT = true: every row starts on a 128-byte line (padded with zero
//    blocks to an even count); values streamed with non-temporal loads.
//  * R = 2, NT = true: a lane group takes a row pair stored contiguously, the pair
//    starting on a line and its second row padded so the pair fills whole lines.
//    A block counts to the first row before position lenA of the pair's range and
//    to the second after it; the second row's blocks land on the lanes of their
//    parity relative to the pair, so each lane still sums every other block of the
//    row in order and the row's result equals the R = 1 result.
// A line that two lane groups share is read twice when streamed non-temporally
// (+30 % reads on 7-block rows), and padding every 7-block row to 8 blocks adds
// 13.5 % of the blocks, which row pairs avoid (0.3 %).  Measured (FP32 values, two
// dots): 96^3 7-pt 86.0 us (R = 1 through L2), 83.1 us (R = 1 aligned), 78.5 us
// (R = 2); 27-pt 64^3 83.4, 76.9 (3.5 % padding) and 80.2 us.  Pipelining the
// next chunk's row pointers or whole first batch changed nothing (85.7-86.4 us),
// and staging a chunk's contiguous blocks through LDS was slower (129 us).
template <int NDOT, int KB, int R, bool NT>
__global__ __launch_bounds__(kMixedBlock) void bsr4_spmv_f32_dot_kernel(
    int nb, const int *__restrict__ rp, const int *__restrict__ ci, const v4f *__restrict__ val4,
    const float *__restrict__ x, float *__restrict__ y, const float *__restrict__ a, double *__restrict__ part,
    const int *__restrict__ skip) {
    static_assert(KB == 4 || KB == 8 || KB == 16 || KB == 32, "KB");
    static_assert(R == 1 || R == 2, "R");
    if (skip && *skip) return;
    constexpr int kJ = KB / 2, kC = KB < 8 ? 1 : KB / 8, kRows = R * (kMixedBlock / 8);
    const int s = threadIdx.x & 7;
    const int b = s >> 2, sub = s & 3;
    double d0 = 0.0, d1 = 0.0;
    for (int chunk = blockIdx.x; chunk * kRows < nb; chunk += gridDim.x) {
        const int r0 = chunk * kRows + R * (threadIdx.x >> 3);  // the group's first row
        const int rs = r0 < nb ? rp[r0] : 0;
        const int mid = r0 < nb ? rp[r0 + 1] : 0;
        const int re = R == 2 && r0 + 1 < nb ? rp[r0 + 2] : mid;
        const int len = re - rs, lenA = mid - rs;
        float accA = 0.0f, accB = 0.0f;
        for (int t = 0; t < len; t += KB) {
            // lane s holds the column of blocks t + 8q + s
            int cv[kC];
#pragma unroll
            for (int q = 0; q < kC; ++q) cv[q] = s < KB && t + 8 * q + s < len ? ci[rs + t + 8 * q + s] : 0;
            v4f w[kJ];
            float xv[kJ];
#pragma unroll
            for (int j = 0; j < kJ; ++j) {
                const int k = t + 2 * j + b;
                const size_t e = (size_t)(k < len ? rs + k : 0) * 4 + sub;
                w[j] = NT ? __builtin_nontemporal_load(&val4[e]) : val4[e];
            }
#pragma unroll
            for (int j = 0; j < kJ; ++j) {
                const int c = __shfl(cv[j >> 2], (2 * j + b) & 7, 8);  // block 2j + b
                xv[j] = x[(size_t)c * 4 + sub];
            }
#pragma unroll
            for (int j = 0; j < kJ; ++j) {
                const int v = __float_as_int(xv[j]);
                const float f0 = __int_as_float(__builtin_amdgcn_mov_dpp(v, 0x00, 0xF, 0xF, false));
                const float f1 = __int_as_float(__builtin_amdgcn_mov_dpp(v, 0x55, 0xF, 0xF, false));
                const float f2 = __int_as_float(__builtin_amdgcn_mov_dpp(v, 0xAA, 0xF, 0xF, false));
                const float f3 = __int_as_float(__builtin_amdgcn_mov_dpp(v, 0xFF, 0xF, 0xF, false));
                const int k = t + 2 * j + b;
                if (k >= len) continue;
                const bool first = R == 1 || k < lenA;  // a select, not a branch: both rows share the lanes
                float acc = first ? accA : accB;
                acc = fmaf(w[j].x, f0, acc);
                acc = fmaf(w[j].y, f1, acc);
                acc = fmaf(w[j].z, f2, acc);
                acc = fmaf(w[j].w, f3, acc);
                accA = first ? acc : accA;
                if (R == 2) accB = first ? accB : acc;
            }
        }
        const float tA = accA + __shfl_xor(accA, 4);
        const float tB = R == 2 ? accB + __shfl_xor(accB, 4) : 0.0f;
        // R = 1: lanes b = 0 write the row; R = 2: lanes b write row r0 + b
        const int row = r0 + (R == 2 ? b : 0);
        if (row < nb && (R == 2 || b == 0)) {
            const float res = R == 2 && b ? tB : tA;
            const size_t i = (size_t)row * 4 + sub;
            y[i] = res;
            if (NDOT >= 1) d0 += (double)res * (double)a[i];
            if (NDOT == 2) d1 += (double)res * (double)res;
        }
    }
    if (NDOT == 0) return;
    __shared__ double red[2][kMixedBlock / 64];
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
        for (int w = 0; w < kMixedBlock / 64; ++w) {
            t0 += red[0][w];
            t1 += red[1][w];
        }
        part[blockIdx.x] = t0;
        if (NDOT == 2) part[gridDim.x + blockIdx.x] = t1;
    }
}

int mixed_dot_groups(int nb) {
    const int chunks = (int)grid_for(nb, kMixedBlock / 8);
    return chunks < kMixedGroups ? (chunks > 0 ? chunks : 1) : kMixedGroups;
}

namespace {
template <int ND, int R, bool NT>
void launch_f32_dot(int kb, unsigned g, int nb, const int *rp, const int *ci, const v4f *v, const float *x, float *y,
                    const float *a, double *part, hipStream_t st, const int *skip) {
    if (kb == 4) bsr4_spmv_f32_dot_kernel<ND, 4, R, NT><<<g, kMixedBlock, 0, st>>>(nb, rp, ci, v, x, y, a, part, skip);
    else if (kb == 8) bsr4_spmv_f32_dot_kernel<ND, 8, R, NT><<<g, kMixedBlock, 0, st>>>(nb, rp, ci, v, x, y, a, part, skip);
    else if (kb == 16) bsr4_spmv_f32_dot_kernel<ND, 16, R, NT><<<g, kMixedBlock, 0, st>>>(nb, rp, ci, v, x, y, a, part, skip);
    else bsr4_spmv_f32_dot_kernel<ND, 32, R, NT><<<g, kMixedBlock, 0, st>>>(nb, rp, ci, v, x, y, a, part, skip);
}
template <int ND>
void launch_f32_dot_layout(int layout, int kb, unsigned g, int nb, const int *rp, const int *ci, const v4f *v,
                           const float *x, float *y, const float *a, double *part, hipStream_t st, const int *skip) {
    if (layout == kA32RowsAligned) launch_f32_dot<ND, 1, true>(kb, g, nb, rp, ci, v, x, y, a, part, st, skip);
    else if (layout == kA32PairsAligned) launch_f32_dot<ND, 2, true>(kb, g, nb, rp, ci, v, x, y, a, part, st, skip);
    else launch_f32_dot<ND, 1, false>(kb, g, nb, rp, ci, v, x, y, a, part, st, skip);
}
}  // namespace

int mixed_spmv_f32_dot(int nb, long long nnzb, const int *rp, const int *ci, const float *val, const float *x,
                       float *y, const float *a, int ndot, double *part, hipStream_t st, const int *skip, int layout) {
    if (nb <= 0) return ISAI_OK;
    // Every workgroup writes its partials, also one left without a chunk (row pairs
    // halve the chunk count): G = mixed_dot_groups(nb) for every layout.
    const unsigned g = (unsigned)mixed_dot_groups(nb);
    const long long avg = (nnzb + nb - 1) / nb;
    const int kb = avg <= 4 ? 4 : avg <= 8 ? 8 : avg <= 16 ? 16 : 32;
    const v4f *v = (const v4f *)val;
    if (ndot == 0) launch_f32_dot_layout<0>(layout, kb, g, nb, rp, ci, v, x, y, a, part, st, skip);
    else if (ndot == 1) launch_f32_dot_layout<1>(layout, kb, g, nb, rp, ci, v, x, y, a, part, st, skip);
    else launch_f32_dot_layout<2>(layout, kb, g, nb, rp, ci, v, x, y, a, part, st, skip);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

// ---------------------------------------------------- inner BiCGSTAB kernels
template <int T>
__device__ __forceinline__ double mixed_block_sum(double v) {
    __shared__ double ws[T / 64];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < T / 64; ++k) t += ws[k];
    return t;
}

// Sum of part[0..G) in a fixed order (thread-strided, then the block tree): every
// workgroup that calls it gets the same bits.
__device__ __forceinline__ double partial_sum(const double *__restrict__ part, int G) {
    double v = 0.0;
    for (int k = threadIdx.x; k < G; k += kMixedVecThreads) v += part[k];
    return mixed_block_sum<kMixedVecThreads>(v);
}

// r = rhs (FP64 input rounded to FP32), rhat = r, x = 0 (p and v are first read
// after the first iteration wrote them); rpart[0] = ||r||^2, rpart[1] = rhat.r
// (= ||r||^2).  The tolerance asks for the reduction still missing in the outer
// loop, tol_in = min(0.5, max(floor, 0.5 tol ||b|| / |g_j|)), from the outer
// residual estimate |g_j| and tol2 = tol^2 ||b||^2 on the device.  The solve starts
// done when the outer loop is done.
__global__ __launch_bounds__(kMixedVecThreads) void inner_init_kernel(
    size_t n, const double *__restrict__ rhs, float *__restrict__ r, float *__restrict__ rh, float *__restrict__ x,
    double *__restrict__ rpart, InnerScalars *IS, const int *__restrict__ outer_done,
    const double *__restrict__ gj, const double *__restrict__ tol2, double floor_tol) {
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * kMixedVecThreads + threadIdx.x; i < n;
         i += (size_t)kMixedVecBlocks * kMixedVecThreads) {
        const float ri = (float)rhs[i];
        r[i] = ri;
        rh[i] = ri;
        x[i] = 0.0f;
        acc += (double)ri * (double)ri;
    }
    acc = mixed_block_sum<kMixedVecThreads>(acc);
    if (threadIdx.x == 0) {
        rpart[blockIdx.x] = acc;
        rpart[kMixedVecBlocks + blockIdx.x] = acc;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const double g = fabs(*gj);
        double t = g > 0.0 ? 0.5 * sqrt(*tol2) / g : 0.5;
        t = fmin(0.5, fmax(floor_tol, t));
        IS->done = (outer_done && *outer_done) ? 1 : 0;
        IS->iters = 0;
        IS->tol_in = t;
    }
}

// The vector kernels below stream FP32 vectors 16 B per lane (n is a multiple of
// four) and issue the loads of their first element before the scalar prologue
// (the fixed-order sums of the previous kernel's partials), so that the memory
// latency overlaps it.
constexpr size_t kVecStride = (size_t)kMixedVecBlocks * kMixedVecThreads;

// Start of an iteration: rho = rhat.r and ||r||^2 from the previous kernel's
// partials; convergence test ||r|| <= tol_in ||r_0||; then p = r (first
// iteration) or p = r + beta (p - omega v), beta = (rho / rho_prev)(alpha / omega).
// `first` comes from the host (it knows which iteration it launches): a flag in the
// record would be written by this kernel's first workgroup while others read it.
__global__ __launch_bounds__(kMixedVecThreads) void inner_p_kernel(size_t n, const float *__restrict__ r,
                                                                   float *__restrict__ p, const float *__restrict__ v,
                                                                   const double *__restrict__ rpart, InnerScalars *IS,
                                                                   int first) {
    if (IS->done) return;
    const v4f *r4 = (const v4f *)r, *v4 = (const v4f *)v;
    v4f *p4 = (v4f *)p;
    const size_t n4 = n / 4;
    size_t i = (size_t)blockIdx.x * kMixedVecThreads + threadIdx.x;
    v4f ra = {}, pa = {}, va = {};
    if (i < n4) {
        ra = r4[i];
        if (!first) pa = p4[i], va = v4[i];
    }
    const double r2 = partial_sum(rpart, kMixedVecBlocks);
    const double rho = partial_sum(rpart + kMixedVecBlocks, kMixedVecBlocks);
    const double tol2 = first ? IS->tol_in * IS->tol_in * r2 : IS->tol2;
    int stop = 0;
    if (!first && r2 <= tol2) stop = 1;
    else if (rho == 0.0) stop = 3;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (first) IS->tol2 = tol2;
        IS->rnorm2 = r2;
        if (stop) IS->done = stop;
        else {
            IS->rho_cur = rho;
            IS->iters++;  // only this thread reads or writes the count
        }
    }
    if (stop) return;
    const float beta = first ? 0.0f : (float)((rho / IS->rho_prev) * (IS->alpha / IS->omega));
    const float omega = first ? 0.0f : (float)IS->omega;
    for (; i < n4; i += kVecStride) {
        if (first) p4[i] = ra;
        else p4[i] = ra + beta * (pa - omega * va);
        const size_t nx = i + kVecStride;
        if (nx < n4) {
            ra = r4[nx];
            if (!first) pa = p4[nx], va = v4[nx];
        }
    }
}

// alpha = rho / rhat.v (partials of the v = A p^ product); s = r - alpha v.
__global__ __launch_bounds__(kMixedVecThreads) void inner_s_kernel(size_t n, const float *__restrict__ r,
                                                                   const float *__restrict__ v, float *__restrict__ s,
                                                                   const double *__restrict__ part, int G,
                                                                   InnerScalars *IS) {
    if (IS->done) return;
    const v4f *r4 = (const v4f *)r, *v4 = (const v4f *)v;
    v4f *s4 = (v4f *)s;
    const size_t n4 = n / 4;
    size_t i = (size_t)blockIdx.x * kMixedVecThreads + threadIdx.x;
    v4f ra = {}, va = {};
    if (i < n4) ra = r4[i], va = v4[i];
    const double d = partial_sum(part, G);
    const double rho = IS->rho_cur;
    if (d == 0.0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) IS->done = 3;
        return;
    }
    const double alpha = rho / d;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        IS->alpha = alpha;
        IS->rho_prev = rho;
    }
    const float af = (float)alpha;
    for (; i < n4; i += kVecStride) {
        s4[i] = ra - af * va;
        const size_t nx = i + kVecStride;
        if (nx < n4) ra = r4[nx], va = v4[nx];
    }
}

// omega = t.s / t.t (partials of the t = A s^ product); x += alpha p^ + omega s^;
// r = s - omega t; rpart[0] = ||r||^2, rpart[1] = rhat.r.  One element per lane
// step: the 16-byte form (six vectors in flight, eight FP64 accumulations per
// step) measured 29.0 us against 24.2 us at 96^3.
__global__ __launch_bounds__(kMixedVecThreads) void inner_xr_kernel(
    size_t n, float *__restrict__ x, float *__restrict__ r, const float *__restrict__ ph, const float *__restrict__ sh,
    const float *__restrict__ s, const float *__restrict__ t, const float *__restrict__ rh,
    const double *__restrict__ part, int G, double *__restrict__ rpart, InnerScalars *IS) {
    if (IS->done) return;
    const double ts = partial_sum(part, G);
    const double tt = partial_sum(part + G, G);
    if (tt == 0.0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) IS->done = 3;
        return;
    }
    const double omega = ts / tt;
    if (blockIdx.x == 0 && threadIdx.x == 0) IS->omega = omega;
    const float af = (float)IS->alpha, of = (float)omega;
    double acc = 0.0, accr = 0.0;
    for (size_t i = (size_t)blockIdx.x * kMixedVecThreads + threadIdx.x; i < n; i += kVecStride) {
        x[i] = x[i] + (af * ph[i] + of * sh[i]);
        const float ri = s[i] - of * t[i];
        r[i] = ri;
        acc += (double)ri * (double)ri;
        accr += (double)rh[i] * (double)ri;
    }
    acc = mixed_block_sum<kMixedVecThreads>(acc);
    accr = mixed_block_sum<kMixedVecThreads>(accr);
    if (threadIdx.x == 0) {
        rpart[blockIdx.x] = acc;
        rpart[kMixedVecBlocks + blockIdx.x] = accr;
    }
}

// z = x widened to FP64 (the flexible preconditioner's output); the first thread
// records the inner solve's outcome in *log for the host's next read.
__global__ void inner_finish_kernel(size_t n, const float *__restrict__ x, double *__restrict__ z,
                                    const InnerScalars *IS, InnerLog *log) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        log->done = IS->done;
        log->iters = IS->iters;
        log->tol_in = IS->tol_in;
    }
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        z[i] = (double)x[i];
}

int inner_init(size_t n, const double *rhs, InnerVecs *w, InnerScalars *IS, const int *outer_done,
               const double *gj, const double *tol2, double floor_tol, hipStream_t st) {
    inner_init_kernel<<<kMixedVecBlocks, kMixedVecThreads, 0, st>>>(n, rhs, w->r, w->rh, w->x, w->rpart, IS,
                                                                    outer_done, gj, tol2, floor_tol);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

// One iteration: p update, p^ = M p, v = A p^, s update, s^ = M s, t = A s^, x/r update.
int inner_iteration(size_t n, int nb, const A32Operand &A, const InnerPrecond *M, InnerVecs *w, InnerScalars *IS,
                    bool first, hipStream_t st) {
    const int *skip = &IS->done;
    const int G = mixed_dot_groups(nb);
    inner_p_kernel<<<kMixedVecBlocks, kMixedVecThreads, 0, st>>>(n, w->r, w->p, w->v, w->rpart, IS, first ? 1 : 0);
    ISAI_LAUNCH_CHECK();
    int s = spmv_lowp_launch(nb, M->nnzbL, M->prec, M->Lrow, M->Lcol, M->Lval, w->p, true, w->tmp, true, st, skip,
                             true);
    if (s == ISAI_OK)
        s = spmv_lowp_launch(nb, M->nnzbU, M->prec, M->Urow, M->Ucol, M->Uval, w->tmp, true, w->ph, true, st, skip,
                             true);
    if (s == ISAI_OK) s = mixed_spmv_f32_dot(nb, A.nnzb, A.rp, A.ci, A.val, w->ph, w->v, w->rh, 1, w->part, st, skip, A.layout);
    if (s != ISAI_OK) return s;
    inner_s_kernel<<<kMixedVecBlocks, kMixedVecThreads, 0, st>>>(n, w->r, w->v, w->s, w->part, G, IS);
    ISAI_LAUNCH_CHECK();
    s = spmv_lowp_launch(nb, M->nnzbL, M->prec, M->Lrow, M->Lcol, M->Lval, w->s, true, w->tmp, true, st, skip,
                         true);
    if (s == ISAI_OK)
        s = spmv_lowp_launch(nb, M->nnzbU, M->prec, M->Urow, M->Ucol, M->Uval, w->tmp, true, w->sh, true, st, skip,
                             true);
    if (s == ISAI_OK) s = mixed_spmv_f32_dot(nb, A.nnzb, A.rp, A.ci, A.val, w->sh, w->t, w->s, 2, w->part, st, skip, A.layout);
    if (s != ISAI_OK) return s;
    inner_xr_kernel<<<kMixedVecBlocks, kMixedVecThreads, 0, st>>>(n, w->x, w->r, w->ph, w->sh, w->s, w->t, w->rh,
                                                                  w->part, G, w->rpart, IS);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

int inner_finish(size_t n, const InnerVecs *w, double *z, const InnerScalars *IS, InnerLog *log, hipStream_t st) {
    size_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    inner_finish_kernel<<<(unsigned)g, 256, 0, st>>>(n, w->x, z, IS, log);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

size_t inner_part_doubles(int nb) { return 2 * (size_t)mixed_dot_groups(nb) + 2 * (size_t)kMixedVecBlocks; }

}  // namespace isai

extern "C" int isai_bsr_spmv_f32_dot(int mb, int nnzb, const int *d_rowptr, const int *d_colind, const float *d_val,
                                     const float *d_x, float *d_y, const float *d_a, int ndot, double *d_part,
                                     int *groups, isai_stream_t stream) {
    ISAI_REQUIRE(mb >= 0 && nnzb >= 0 && (ndot == 1 || ndot == 2), "isai_bsr_spmv_f32_dot: bad size or ndot");
    ISAI_REQUIRE(mb == 0 || (d_rowptr && d_colind && d_val && d_x && d_y && d_a && d_part),
                 "isai_bsr_spmv_f32_dot: null pointer");
    if (groups) *groups = isai::mixed_dot_groups(mb);
    return isai::mixed_spmv_f32_dot(mb, nnzb, d_rowptr, d_colind, d_val, d_x, d_y, d_a, ndot, d_part,
                                    (hipStream_t)stream, nullptr, isai::kA32Unaligned);
}
