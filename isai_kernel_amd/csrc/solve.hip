// solve.hip -- setup handle and the FP64 Krylov methods (BiCGSTAB, FGMRES(m))
// around the ILU(0)/ISAI preconditioner, for gfx950.
//
// The reference's driver ends with one application x = M_U (M_L b) and a host
// residual print (main.cu:817-866); the Krylov loop here is new.  All scalar
// recurrences stay on the device: every reduction is a fixed-shape two-stage
// sum (deterministic), and the small scalar updates run in the final stage of the
// reduction that produces them.  A device word `done` makes every later kernel of
// an iteration return at once after convergence, so the host only reads that
// word once per iteration (BiCGSTAB) or per Arnoldi step (FGMRES).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "bcsc2bcsr.h"
#include "bsr_tensor_SpMV.h"
#include "bsrilu.h"
#include "cape_sptrsv.h"
#include "common.hpp"
#include "internal.hpp"
#include "levels.hpp"
#include "isai_adjoint.h"
#include "isai_bsr.h"
#include "isai_solve.h"

namespace isai {

constexpr int kRedBlocks = 512;
// 512 x 1024 threads keep 32 waves per CU streaming the vector kernels (512 x 256
// measured 4.6 TB/s on s = r - alpha v); the multi-dot, with eight accumulators per
// thread, is faster at 256 threads per block.
constexpr int kRedThreads = 1024;
constexpr int kMdotThreads = 256;
constexpr int kMaxRestart = 64;

// Device scalar record of a solve.
struct Scalars {
    double rho, rho_old, alpha, omega, beta;
    double bnorm2, rnorm2, snorm2, tol2;
    double H[(kMaxRestart + 1) * kMaxRestart];
    double Hbar[(kMaxRestart + 1) * kMaxRestart];  // H before the Givens rotations (Arnoldi relation)
    double cs[kMaxRestart], sn[kMaxRestart], g[kMaxRestart + 1], y[kMaxRestart];
    double hcur[kMaxRestart + 1];  // Gram-Schmidt coefficients of the current pass
    // 0 running, 1 converged, 2 converged at s (BiCGSTAB), 3 breakdown, 4 iteration limit
    int done;
    int first;
    int iters, maxit;  // BiCGSTAB iteration count and limit (device side)
    int jend;          // FGMRES: Arnoldi steps of the current cycle
    int reorth;        // FGMRES: 1 when the current step needs the second Gram-Schmidt pass
    int xzero;         // mixed FGMRES: 1 when the initial guess is zero everywhere
    int pad;
    double wnorm2;     // FGMRES: ||w||^2 before the first Gram-Schmidt pass
    double partial[2][kRedBlocks];
    double mpart[(kMaxRestart + 1) * kRedBlocks];  // multi-dot partials, [vector][block]
};
constexpr int kStateInts = 8;  // done .. pad, read by the host in one copy

template <int T = kRedThreads>
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double ws[T / 64];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) ws[w] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < T / 64; ++k) t += ws[k];
    return t;
}

// N block sums with one pair of barriers: on return thread u < N holds the block
// total of v[u] (each in block_sum's order: wave butterfly, then the waves in order).
template <int N, int T>
__device__ __forceinline__ double block_sum_many(double (&v)[N]) {
    __shared__ double ws[T / 64][N];
#pragma unroll
    for (int u = 0; u < N; ++u)
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v[u] += __shfl_xor(v[u], off);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int u = 0; u < N; ++u) ws[w][u] = v[u];
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x < N)
#pragma unroll
        for (int k = 0; k < T / 64; ++k) t += ws[k][threadIdx.x];
    return t;
}

// Sum of partial[slot][0..kRedBlocks) in a fixed order (one block).
__device__ __forceinline__ double final_sum(const Scalars *S, int slot) {
    double v = 0.0;
    for (int k = threadIdx.x; k < kRedBlocks; k += kRedThreads) v += S->partial[slot][k];
    return block_sum(v);
}

// final_sum of both slots in one pass (each sum in final_sum's order).
__device__ __forceinline__ void final_sum2(const Scalars *S, double &s0, double &s1) {
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < kRedBlocks; k += kRedThreads) {
        a += S->partial[0][k];
        b += S->partial[1][k];
    }
    __shared__ double wa[kRedThreads / 64], wb[kRedThreads / 64];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        wa[threadIdx.x >> 6] = a;
        wb[threadIdx.x >> 6] = b;
    }
    __syncthreads();
    s0 = s1 = 0.0;
#pragma unroll
    for (int k = 0; k < kRedThreads / 64; ++k) {
        s0 += wa[k];
        s1 += wb[k];
    }
}

__device__ __forceinline__ bool skipped(const Scalars *S) { return S->done != 0; }

// partial[0] += a.b ; partial[1] += a.c (when c != null)
__global__ __launch_bounds__(kRedThreads) void dot_kernel(size_t n, const double *__restrict__ a,
                                                          const double *__restrict__ b, const double *__restrict__ c,
                                                          Scalars *S, int honor_done) {
    if (honor_done && skipped(S)) return;
    double s0 = 0.0, s1 = 0.0;
    for (size_t i = (size_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (size_t)kRedBlocks * kRedThreads) {
        const double av = a[i];
        s0 += av * b[i];
        if (c) s1 += av * c[i];
    }
    s0 = block_sum(s0);
    if (c) s1 = block_sum(s1);
    if (threadIdx.x == 0) {
        S->partial[0][blockIdx.x] = s0;
        S->partial[1][blockIdx.x] = c ? s1 : 0.0;
    }
}

// ------------------------------------------------------------------ BiCGSTAB
// init: r = b - Ax (r holds Ax), rhat = r, p = v = 0; partial ||b||^2, ||r||^2
__global__ __launch_bounds__(kRedThreads) void bicg_init_kernel(size_t n, const double *__restrict__ b, double *r,
                                                                double *rhat, double *p, double *v, Scalars *S) {
    double sb = 0.0, sr = 0.0;
    for (size_t i = (size_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (size_t)kRedBlocks * kRedThreads) {
        const double bi = b[i], ri = bi - r[i];
        r[i] = ri;
        rhat[i] = ri;
        p[i] = 0.0;
        v[i] = 0.0;
        sb += bi * bi;
        sr += ri * ri;
    }
    sb = block_sum(sb);
    sr = block_sum(sr);
    if (threadIdx.x == 0) {
        S->partial[0][blockIdx.x] = sb;
        S->partial[1][blockIdx.x] = sr;
    }
}

__global__ void bicg_start_kernel(Scalars *S, int maxit) {
    S->iters = 0;
    S->maxit = maxit;
}

// rhat = r, so the first rho = ||r||^2
__global__ __launch_bounds__(kRedThreads) void bicg_init_final(Scalars *S, double tol) {
    double b2, r2;
    final_sum2(S, b2, r2);
    if (threadIdx.x == 0) {
        S->bnorm2 = b2 > 0.0 ? b2 : 1.0;
        S->rnorm2 = r2;
        S->tol2 = tol * tol * S->bnorm2;
        S->rho_old = S->alpha = S->omega = 1.0;
        S->rho = r2;
        S->first = 1;
        S->done = (r2 <= S->tol2) ? 1 : (S->iters >= S->maxit ? 4 : 0);
    }
}

// Sum of G workgroup partials in a fixed order by one block of kWgSumThreads;
// eight independent accumulators per thread keep eight loads in flight.
constexpr int kWgSumThreads = 1024;

__device__ __forceinline__ double wg_partial_sum(const double *part, int G) {
    double v[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int k = threadIdx.x;
    for (; k + 7 * kWgSumThreads < G; k += 8 * kWgSumThreads) {
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] += part[k + u * kWgSumThreads];
    }
    for (; k < G; k += kWgSumThreads) v[0] += part[k];
    double t = ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
    __shared__ double ws[kWgSumThreads / 64];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) t += __shfl_xor(t, off);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = t;
    __syncthreads();
    double sum = 0.0;
#pragma unroll
    for (int w = 0; w < kWgSumThreads / 64; ++w) sum += ws[w];
    return sum;
}

// The same sum for two partial arrays in one pass (each in the order of
// wg_partial_sum, so the results are identical): 16 loads in flight per thread.
__device__ __forceinline__ void wg_partial_sum2(const double *pa, const double *pb, int G, double &sa, double &sb) {
    double v[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, w[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int k = threadIdx.x;
    for (; k + 7 * kWgSumThreads < G; k += 8 * kWgSumThreads) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            v[u] += pa[k + u * kWgSumThreads];
            w[u] += pb[k + u * kWgSumThreads];
        }
    }
    for (; k < G; k += kWgSumThreads) {
        v[0] += pa[k];
        w[0] += pb[k];
    }
    double ta = ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
    double tb = ((w[0] + w[1]) + (w[2] + w[3])) + ((w[4] + w[5]) + (w[6] + w[7]));
    __shared__ double wsa[kWgSumThreads / 64], wsb[kWgSumThreads / 64];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        ta += __shfl_xor(ta, off);
        tb += __shfl_xor(tb, off);
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        wsa[threadIdx.x >> 6] = ta;
        wsb[threadIdx.x >> 6] = tb;
    }
    __syncthreads();
    sa = sb = 0.0;
#pragma unroll
    for (int q = 0; q < kWgSumThreads / 64; ++q) {
        sa += wsa[q];
        sb += wsb[q];
    }
}

// alpha = rho / rhat.v, with rhat.v from the product's fused partials
__global__ __launch_bounds__(kWgSumThreads) void bicg_alpha_final_wg(Scalars *S, const double *part, int G) {
    if (skipped(S)) return;
    const double d = wg_partial_sum(part, G);
    if (threadIdx.x == 0) {
        S->iters++;
        if (d == 0.0) S->done = 3;
        else S->alpha = S->rho / d;
        S->first = 0;
    }
}

// omega = t.s / t.t from the product's fused partials
__global__ __launch_bounds__(kWgSumThreads) void bicg_omega_final_wg(Scalars *S, const double *part, int G) {
    if (skipped(S)) return;
    double ts, tt;
    wg_partial_sum2(part, part + G, G, ts, tt);
    if (threadIdx.x == 0) {
        if (tt == 0.0) S->done = 3;
        else S->omega = ts / tt;
    }
}

__global__ void bicg_p_kernel(size_t n, const double *__restrict__ r, double *p, const double *__restrict__ v,
                              const Scalars *S) {
    if (skipped(S)) return;
    const bool first = S->first;
    const double beta = S->beta, omega = S->omega;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = first ? r[i] : r[i] + beta * (p[i] - omega * v[i]);
}

// s = r - alpha v; partial ||s||^2
__global__ __launch_bounds__(kRedThreads) void bicg_s_kernel(size_t n, const double *__restrict__ r,
                                                             const double *__restrict__ v, double *__restrict__ s,
                                                             Scalars *S) {
    if (skipped(S)) return;
    const double alpha = S->alpha;
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (size_t)kRedBlocks * kRedThreads) {
        const double si = r[i] - alpha * v[i];
        s[i] = si;
        acc += si * si;
    }
    acc = block_sum(acc);
    if (threadIdx.x == 0) S->partial[0][blockIdx.x] = acc;
}

__global__ __launch_bounds__(kRedThreads) void bicg_snorm_final(Scalars *S) {
    if (skipped(S)) return;
    const double s2 = final_sum(S, 0);
    if (threadIdx.x == 0) {
        S->snorm2 = s2;
        if (s2 <= S->tol2) S->done = 2;
    }
}

// x += alpha phat + omega shat; r = s - omega t; partials ||r||^2 and rhat.r (the
// next iteration's rho).  When the solve converged on s (done = 2) it only adds
// alpha phat.  HatT: double, or float when the preconditioner's output is held in
// FP32 (a reduced-precision ISAI apply, whose values are FP32 numbers either way).
template <typename HatT>
__global__ __launch_bounds__(kRedThreads) void bicg_xr_kernel(size_t n, double *x, double *r,
                                                              const HatT *__restrict__ phat,
                                                              const HatT *__restrict__ shat,
                                                              const double *__restrict__ s,
                                                              const double *__restrict__ t,
                                                              const double *__restrict__ rhat, Scalars *S) {
    const int done = S->done;
    if (done == 2) {
        const double alpha = S->alpha;
        for (size_t i = (size_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (size_t)kRedBlocks * kRedThreads)
            x[i] += alpha * (double)phat[i];
        return;
    }
    if (done) return;
    const double alpha = S->alpha, omega = S->omega;
    double acc = 0.0, accr = 0.0;
    for (size_t i = (size_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (size_t)kRedBlocks * kRedThreads) {
        x[i] += alpha * (double)phat[i] + omega * (double)shat[i];
        const double ri = s[i] - omega * t[i];
        r[i] = ri;
        acc += ri * ri;
        accr += rhat[i] * ri;
    }
    acc = block_sum(acc);
    accr = block_sum(accr);
    if (threadIdx.x == 0) {
        S->partial[0][blockIdx.x] = acc;
        S->partial[1][blockIdx.x] = accr;
    }
}

// End of an iteration: convergence test, next rho and beta, iteration limit.  A
// solve that converged on s (done = 2, x already updated) becomes done = 1 here, so
// iterations launched after convergence do nothing.
__global__ __launch_bounds__(kRedThreads) void bicg_rnorm_final(Scalars *S) {
    if (S->done == 2) {
        __syncthreads();
        if (threadIdx.x == 0) S->done = 1;
        return;
    }
    if (skipped(S)) return;
    double r2, rho;
    final_sum2(S, r2, rho);
    if (threadIdx.x == 0) {
        S->rnorm2 = r2;
        if (r2 <= S->tol2) {
            S->done = 1;
        } else if (rho == 0.0) {
            S->done = 3;
        } else {
            S->beta = (rho / S->rho) * (S->alpha / S->omega);
            S->rho_old = S->rho;
            S->rho = rho;
            if (S->iters >= S->maxit) S->done = 4;
        }
    }
}

// ------------------------------------------------------------- residual
// r = b - r (r holds A x); partial ||r||^2, ||b||^2
// r = b - r (r holds A x); with xz set (x was zero everywhere, so the product that
// would have written r was skipped) r = b, the same bits b - A*0 gives
__global__ __launch_bounds__(kRedThreads) void residual_kernel(size_t n, const double *__restrict__ b, double *r,
                                                               Scalars *S, const int *xz = nullptr) {
    const bool zero = xz && *xz;
    double sr = 0.0, sb = 0.0;
    for (size_t i = (size_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (size_t)kRedBlocks * kRedThreads) {
        const double bi = b[i], ri = zero ? bi : bi - r[i];
        r[i] = ri;
        sr += ri * ri;
        sb += bi * bi;
    }
    sr = block_sum(sr);
    sb = block_sum(sb);
    if (threadIdx.x == 0) {
        S->partial[0][blockIdx.x] = sr;
        S->partial[1][blockIdx.x] = sb;
    }
}

// S->xzero = 1 when x is zero everywhere (-0 counts as zero), else 0: a one-thread
// reset, then a scan in which any wave that meets a nonzero stores 0
__global__ void xzero_reset_kernel(Scalars *S) { S->xzero = 1; }
__global__ __launch_bounds__(256) void xzero_scan_kernel(size_t n, const double *__restrict__ x, Scalars *S) {
    bool nz = false;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) nz |= x[i] != 0.0;
    if (__ballot(nz) && (threadIdx.x & 63) == 0) S->xzero = 0;
}

__global__ __launch_bounds__(kRedThreads) void residual_final(Scalars *S, double *out2) {
    double r2, b2;
    final_sum2(S, r2, b2);
    if (threadIdx.x == 0) {
        out2[0] = r2;
        out2[1] = b2;
    }
}

// ------------------------------------------------------------------- FGMRES
// cycle start: w holds A x -> w = b - Ax, partial ||w||^2, ||b||^2
__global__ __launch_bounds__(kRedThreads) void fg_beta_final(Scalars *S, double tol) {
    double w2, b2;
    final_sum2(S, w2, b2);
    if (threadIdx.x == 0) {
        S->bnorm2 = b2 > 0.0 ? b2 : 1.0;
        S->tol2 = tol * tol * S->bnorm2;
        const double beta = sqrt(w2);
        S->beta = beta;
        for (int k = 0; k <= kMaxRestart; ++k) S->g[k] = 0.0;
        S->g[0] = beta;
        S->jend = 0;
        S->done = (w2 <= S->tol2) ? 1 : 0;
    }
}

__global__ void fg_scale_kernel(size_t n, const double *__restrict__ w, double *__restrict__ v, const Scalars *S,
                                int hrow, int hcol, int use_beta) {
    if (S->done) return;
    const double d = use_beta ? S->beta : S->H[hrow * kMaxRestart + hcol];
    if (d == 0.0) return;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        v[i] = w[i] / d;
}

// Classical Gram-Schmidt with a conditional second pass (DGKS criterion, Daniel,
// Gragg, Kaufman and Stewart 1976): each pass computes h = V^T w with one
// multi-dot (up to kDotChunk vectors per block column of the grid), sums the
// partials per vector in a fixed order, and subtracts V h from w in one pass.  The
// second pass runs only when the first shrank ||w|| below ||w_before|| / sqrt(2);
// its kernels return at once otherwise.  Pass 0 also measures ||w_before||^2.
constexpr int kDotChunk = 8;

__device__ __forceinline__ bool pass_skipped(const Scalars *S, int pass) {
    return S->done != 0 || (pass == 1 && !S->reorth);
}

// T = double (FP64 FGMRES) or float (the inner GMRES cycle of the mixed form: FP32
// storage, FP64 accumulation).
// Pass 1 decides itself whether it is needed: every block sums pass 0's ||w||^2
// partials in the same order and compares with ||w_before||^2 / 2 (block (0, 0)
// records the decision for the pass's other kernels), which saves a launch.
template <class T>
__global__ __launch_bounds__(kMdotThreads) void fg_mdot_kernel(size_t n, const T *__restrict__ w,
                                                              const T *__restrict__ V, int nv, Scalars *S, int pass) {
    if (S->done) return;
    if (pass == 1) {
        double p = 0.0;
        for (int k = threadIdx.x; k < kRedBlocks; k += kMdotThreads) p += S->partial[0][k];
        const bool again = block_sum<kMdotThreads>(p) < 0.5 * S->wnorm2;
        if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) S->reorth = again ? 1 : 0;
        if (!again) return;
    }
    const int i0 = blockIdx.y * kDotChunk;
    const bool self = pass == 0 && blockIdx.y == 0;  // this block column also sums w.w
    double acc[kDotChunk], ww = 0.0;
#pragma unroll
    for (int u = 0; u < kDotChunk; ++u) acc[u] = 0.0;
    // The loads of a column are unconditional (a block column past the last vector
    // re-reads the last one and discards it): loads under a per-vector branch were
    // waited for one at a time.  Each vector's partial sum keeps its order.  FP32
    // vectors go 16 B per lane (four entries; n is a multiple of four).
    if constexpr (sizeof(T) == 4) {
        const size_t n4 = n / 4;
        const v4f *w4 = (const v4f *)w, *V4 = (const v4f *)V;
#pragma unroll 2
        for (size_t k = (size_t)blockIdx.x * kMdotThreads + threadIdx.x; k < n4;
             k += (size_t)kRedBlocks * kMdotThreads) {
            const v4f wk = w4[k];
            v4f vv[kDotChunk];
#pragma unroll
            for (int u = 0; u < kDotChunk; ++u) vv[u] = V4[(size_t)min(i0 + u, nv - 1) * n4 + k];
            if (self)
                ww += (double)wk.x * wk.x + (double)wk.y * wk.y + (double)wk.z * wk.z + (double)wk.w * wk.w;
#pragma unroll
            for (int u = 0; u < kDotChunk; ++u) {
                const double d = (double)vv[u].x * wk.x + (double)vv[u].y * wk.y + (double)vv[u].z * wk.z +
                                 (double)vv[u].w * wk.w;
                acc[u] = i0 + u < nv ? acc[u] + d : acc[u];
            }
        }
    } else {
        for (size_t k = (size_t)blockIdx.x * kMdotThreads + threadIdx.x; k < n;
             k += (size_t)kRedBlocks * kMdotThreads) {
            const double wk = w[k];
            double vv[kDotChunk];
#pragma unroll
            for (int u = 0; u < kDotChunk; ++u) vv[u] = V[(size_t)min(i0 + u, nv - 1) * n + k];
            if (self) ww += wk * wk;
#pragma unroll
            for (int u = 0; u < kDotChunk; ++u) acc[u] = i0 + u < nv ? acc[u] + vv[u] * wk : acc[u];
        }
    }
    double sums[kDotChunk + 1];
#pragma unroll
    for (int u = 0; u < kDotChunk; ++u) sums[u] = acc[u];
    sums[kDotChunk] = ww;
    const double t = block_sum_many<kDotChunk + 1, kMdotThreads>(sums);
    const int u = threadIdx.x;
    if (u < kDotChunk && i0 + u < nv) S->mpart[(size_t)(i0 + u) * kRedBlocks + blockIdx.x] = t;
    if (self && u == kDotChunk) S->partial[1][blockIdx.x] = t;
}

// block i: h_i = sum of its partials; pass 0 sets H[i][j] (and block 0 stores
// ||w_before||^2), pass 1 adds to H[i][j]
__global__ __launch_bounds__(kRedThreads) void fg_hsum_final(Scalars *S, int j, int pass) {
    if (pass_skipped(S, pass)) return;
    const int i = blockIdx.x;
    double v = 0.0;
    for (int k = threadIdx.x; k < kRedBlocks; k += kRedThreads) v += S->mpart[(size_t)i * kRedBlocks + k];
    v = block_sum(v);
    double wb = 0.0;
    if (pass == 0 && i == 0) wb = final_sum(S, 1);
    if (threadIdx.x == 0) {
        S->hcur[i] = v;
        if (pass == 0) S->H[i * kMaxRestart + j] = v;
        else S->H[i * kMaxRestart + j] += v;
        if (pass == 0 && i == 0) S->wnorm2 = wb;
    }
}

// w -= sum_i hcur[i] V_i (i ascending, FP64), stored as T; partial ||w||^2 of the
// stored values
template <class T>
__global__ __launch_bounds__(kRedThreads) void fg_axpy_multi_kernel(size_t n, T *w, const T *__restrict__ V, int nv,
                                                                    Scalars *S, int pass) {
    if (pass_skipped(S, pass)) return;
    // the coefficients in LDS; each thread issues the loads of eight basis vectors
    // before it uses them (a clamped index re-reads the last vector and is not
    // used), so a step waits once per eight vectors instead of once per vector
    __shared__ double hs[kMaxRestart + 1];
    for (int i = threadIdx.x; i < nv; i += kRedThreads) hs[i] = S->hcur[i];
    __syncthreads();
    double acc = 0.0;
    if constexpr (sizeof(T) == 4) {  // 16 B per lane
        const size_t n4 = n / 4;
        v4f *w4 = (v4f *)w;
        const v4f *V4 = (const v4f *)V;
        for (size_t k = (size_t)blockIdx.x * kRedThreads + threadIdx.x; k < n4;
             k += (size_t)kRedBlocks * kRedThreads) {
            const v4f w0 = w4[k];
            double a = w0.x, b = w0.y, c = w0.z, d = w0.w;
            for (int i0 = 0; i0 < nv; i0 += kDotChunk) {
                v4f q[kDotChunk];
#pragma unroll
                for (int u = 0; u < kDotChunk; ++u) q[u] = V4[(size_t)min(i0 + u, nv - 1) * n4 + k];
#pragma unroll
                for (int u = 0; u < kDotChunk; ++u) {
                    if (i0 + u < nv) {
                        const double hi = hs[i0 + u];
                        a -= hi * q[u].x;
                        b -= hi * q[u].y;
                        c -= hi * q[u].z;
                        d -= hi * q[u].w;
                    }
                }
            }
            const v4f r = v4f{(float)a, (float)b, (float)c, (float)d};
            w4[k] = r;
            acc += (double)r.x * r.x + (double)r.y * r.y + (double)r.z * r.z + (double)r.w * r.w;
        }
    } else {
        for (size_t k = (size_t)blockIdx.x * kRedThreads + threadIdx.x; k < n;
             k += (size_t)kRedBlocks * kRedThreads) {
            double wk = w[k];
            for (int i0 = 0; i0 < nv; i0 += kDotChunk) {
                double q[kDotChunk];
#pragma unroll
                for (int u = 0; u < kDotChunk; ++u) q[u] = V[(size_t)min(i0 + u, nv - 1) * n + k];
#pragma unroll
                for (int u = 0; u < kDotChunk; ++u)
                    if (i0 + u < nv) wk -= hs[i0 + u] * q[u];
            }
            w[k] = wk;
            acc += wk * wk;
        }
    }
    acc = block_sum(acc);
    if (threadIdx.x == 0) S->partial[0][blockIdx.x] = acc;
}

// ||w|| -> H[j+1][j]; Givens rotations on column j; residual estimate |g[j+1]|.
// The column, its copy into Hbar and the rotations' cosines and sines are staged in
// LDS by the whole block; one thread then runs the chain of rotations with the
// running entry in a register (the same operations in the same order as a direct
// walk over H, without a global-memory round trip per rotation).
__global__ __launch_bounds__(kRedThreads) void fg_givens_final(Scalars *S, int j) {
    if (skipped(S)) return;
    const double w2 = final_sum(S, 0);
    constexpr int m = kMaxRestart;
    __shared__ double col[kMaxRestart + 1], c_[kMaxRestart], s_[kMaxRestart];
    const double hn = sqrt(w2);
    for (int i = threadIdx.x; i <= j; i += blockDim.x) {
        const double v = S->H[i * m + j];
        col[i] = v;
        S->Hbar[i * m + j] = v;
        if (i < j) {
            c_[i] = S->cs[i];
            s_[i] = S->sn[i];
        }
    }
    if (threadIdx.x == 0) {
        S->H[(j + 1) * m + j] = hn;
        S->Hbar[(j + 1) * m + j] = hn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double cur = col[0];
        for (int i = 0; i < j; ++i) {
            const double c = col[i + 1];
            col[i] = c_[i] * cur + s_[i] * c;
            cur = -s_[i] * cur + c_[i] * c;
        }
        const double a = cur;
        const double den = sqrt(a * a + hn * hn);
        const double cj = den == 0.0 ? 1.0 : a / den;
        const double sj = den == 0.0 ? 0.0 : hn / den;
        S->cs[j] = cj;
        S->sn[j] = sj;
        col[j] = cj * a + sj * hn;
        const double gj = S->g[j];
        S->g[j + 1] = -sj * gj;
        S->g[j] = cj * gj;
        // keep the unrotated norm for the basis scaling
        S->rnorm2 = hn;
        S->jend = j + 1;
        if (S->g[j + 1] * S->g[j + 1] <= S->tol2 || hn == 0.0) S->done = 1;
    }
    __syncthreads();
    for (int i = threadIdx.x; i <= j; i += blockDim.x) S->H[i * m + j] = col[i];
}

template <class T>
__global__ void fg_scale_next_kernel(size_t n, const T *__restrict__ w, T *__restrict__ v, const Scalars *S) {
    if (S->done) return;
    const double d = S->rnorm2;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        v[i] = (T)((double)w[i] / d);
}

// ------------------------------------------- inner GMRES cycle (mixed form)
// Start of one cycle on A z = v: q_0 = v rounded to FP32 with beta = 1 (v is a unit
// vector of the outer basis), g = (1, 0, ...), and the cycle stops when |g_{i+1}|
// <= tol_in, tol_in = min(0.5, max(floor_tol, 0.5 sqrt(*tol2) / |*gj|)) from the
// outer residual estimate |g_j| and tol2 = tol^2 ||b||^2 (the inner BiCGSTAB's
// rule).  The cycle starts done when the outer loop is done.
__global__ void ig_start_kernel(size_t n, const double *__restrict__ v, float *__restrict__ q0, Scalars *Si,
                                const int *__restrict__ outer_done, const double *__restrict__ gj,
                                const double *__restrict__ tol2, double floor_tol) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const double g = fabs(*gj);
        double t = g > 0.0 ? 0.5 * sqrt(*tol2) / g : 0.5;
        t = fmin(0.5, fmax(floor_tol, t));
        Si->tol2 = t * t;
        Si->beta = 1.0;
        for (int k = 0; k <= kMaxRestart; ++k) Si->g[k] = 0.0;
        Si->g[0] = 1.0;
        Si->jend = 0;
        Si->reorth = 0;
        Si->done = *outer_done ? 1 : 0;
    }
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        q0[i] = (float)v[i];
}

// u = sum_{k < jend} y_k q_k (FP64 accumulation, stored in FP32); the first thread
// records the cycle's outcome (done, steps, tolerance) for the host's next read.
__global__ void ig_combine_kernel(size_t n, const float *__restrict__ Q, const Scalars *Si, float *__restrict__ u,
                                  InnerLog *log) {
    const int jend = Si->jend;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        log->done = Si->done;
        log->iters = jend;
        log->tol_in = sqrt(Si->tol2);
    }
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int k0 = 0; k0 < jend; k0 += 8) {  // eight loads in flight, then their terms in order
            float q[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) q[t] = Q[(size_t)min(k0 + t, jend - 1) * n + i];
#pragma unroll
            for (int t = 0; t < 8; ++t)
                if (k0 + t < jend) acc += Si->y[k0 + t] * (double)q[t];
        }
        u[i] = (float)acc;
    }
}

// Back substitution H(0:jend,0:jend) y = g.
__global__ void fg_solve_y_kernel(Scalars *S) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int jend = S->jend;
    constexpr int m = kMaxRestart;
    for (int i = jend - 1; i >= 0; --i) {
        double s = S->g[i];
        for (int k = i + 1; k < jend; ++k) s -= S->H[i * m + k] * S->y[k];
        S->y[i] = s / S->H[i * m + i];
    }
}

// x += sum_k y_k Z_k
__global__ void fg_update_x_kernel(size_t n, double *x, const double *__restrict__ Z, const Scalars *S) {
    const int jend = S->jend;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double acc = x[i];
        for (int k0 = 0; k0 < jend; k0 += 8) {  // eight loads in flight, then their terms in order
            double z[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) z[t] = Z[(size_t)min(k0 + t, jend - 1) * n + i];
#pragma unroll
            for (int t = 0; t < 8; ++t)
                if (k0 + t < jend) acc += S->y[k0 + t] * z[t];
        }
        x[i] = acc;
    }
}

static unsigned vec_grid(size_t n) {
    size_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    return (unsigned)(g ? g : 1);
}

}  // namespace isai

using namespace isai;

struct isai_handle {
    int nb = 0, nnzb = 0, bs = 0;
    size_t n = 0;
    const int *d_row = nullptr, *d_col = nullptr;
    const double *d_val = nullptr;
    double *d_LU = nullptr;
    int nnzbL = 0, nnzbU = 0, nlevels = 0;
    int plan_ncache = 0, plan_ordered[2] = {0, 0};  // ISAI build plan (isai_handle_build_plan)
    int *d_Lrow = nullptr, *d_Lcol = nullptr, *d_Urow = nullptr, *d_Ucol = nullptr;
    double *d_Lval = nullptr, *d_Uval = nullptr;
    // reduced-precision copies, each row padded to start on a 128-byte line
    // (pad_convert_lowp), with their own patterns for the current line alignment
    void *d_Llp = nullptr, *d_Ulp = nullptr;
    int *d_Lrow_lp = nullptr, *d_Lcol_lp = nullptr, *d_Urow_lp = nullptr, *d_Ucol_lp = nullptr;
    int nnzbL_lp = 0, nnzbU_lp = 0, lp_align = 0;
    int lp_prec = -1;
    int precond = ISAI_PRECOND_ISAI, precision = ISAI_PREC_FP64, chunk = 0;
    isai_sptrsv_plan_t sv = nullptr;
    double ms_ilu = 0, ms_isai = 0, ms_total = 0;
    double *work = nullptr;
    size_t work_vecs = 0;
    Scalars *S = nullptr;
    double *d_res2 = nullptr;  // residual output pair
    int *h_done = nullptr;      // pinned, kStateInts words (done, first, iters, maxit, jend, ...)
    double *d_wgpart = nullptr; // fused-dot partials of the A products, 2 * spmv_dot_groups(nb)
    double *h_res2 = nullptr;   // pinned
    // matrix arrays owned by the handle (adjoint handles own A^T)
    int *own_row = nullptr, *own_col = nullptr;
    double *own_val = nullptr;
    // mixed-precision form: A rounded to FP32, the inner solve's FP32 vectors and
    // FP64 partials, its scalar record, and the inner-solve parameters
    // (A32 is laid out on 128-byte lines with its own pattern: build_a32)
    float *d_A32 = nullptr;
    int *d_A32row = nullptr, *d_A32col = nullptr;
    long long nnzb32 = 0;
    int a32_layout = kA32Unaligned;
    float *inner_vecs = nullptr;
    double *inner_part = nullptr;
    InnerScalars *IS = nullptr;
    int *h_inner = nullptr;        // pinned: inner done, iters (the inner solve's own read loop)
    InnerLog *d_log = nullptr;     // outcome of each inner solve since the last host read
    InnerLog *h_log = nullptr;     // pinned copy
    // inner GMRES cycle (ISAI_SOLVER_FGMRES_MIXED_GMRES): basis, work vectors and
    // scalar record; capacity in steps
    float *ig_vecs = nullptr;
    Scalars *Sin = nullptr;
    int ig_cap = 0;
    double inner_tol = 3e-5;       // smallest relative reduction asked of one inner solve
    int inner_maxit = 40;          // inner iterations per outer step
    double inner_rate = 1.6;       // inner iterations per decade of reduction seen so far
    int inner_hist[64] = {};       // iterations the inner solve of outer step j took last time
    int outer_pred = 0;            // outer steps the last converged solve took (0 = unknown)
    int last_outer = 0, last_inner = 0;
    int last_restart = 0;          // FGMRES cycle length of the last FGMRES solve
    // BiCGSTAB iterations captured as a hipGraph (isai_set_solve_graph) and the
    // configuration they were captured with
    bool use_graph = false;
    hipGraph_t it_graph = nullptr;
    hipGraphExec_t it_exec = nullptr;
    const void *it_key[4] = {nullptr, nullptr, nullptr, nullptr};
    int it_keyi[2] = {-1, -1};
};

namespace {

int ensure_work(isai_handle *h, size_t vecs) {
    if (!h->d_wgpart && hipMalloc(&h->d_wgpart, sizeof(double) * 2 * spmv_dot_groups(h->nb)) != hipSuccess) {
        h->d_wgpart = nullptr;
        set_error("isai_solve: cannot allocate the reduction partials");
        return ISAI_ERR_ALLOC;
    }
    if (h->work_vecs >= vecs) return ISAI_OK;
    if (h->work) (void)hipFree(h->work);
    h->work = nullptr;
    h->work_vecs = 0;
    if (hipMalloc(&h->work, sizeof(double) * h->n * vecs) != hipSuccess) {
        set_error("isai_solve: cannot allocate %zu work vectors of %zu doubles", vecs, h->n);
        return ISAI_ERR_ALLOC;
    }
    h->work_vecs = vecs;
    return ISAI_OK;
}

void free_lp(isai_handle *h) {
    (void)hipFree(h->d_Llp);
    (void)hipFree(h->d_Ulp);
    (void)hipFree(h->d_Lrow_lp);
    (void)hipFree(h->d_Lcol_lp);
    (void)hipFree(h->d_Urow_lp);
    (void)hipFree(h->d_Ucol_lp);
    h->d_Llp = h->d_Ulp = nullptr;
    h->d_Lrow_lp = h->d_Lcol_lp = h->d_Urow_lp = h->d_Ucol_lp = nullptr;
    h->nnzbL_lp = h->nnzbU_lp = h->lp_align = 0;
    h->lp_prec = -1;
}

// Padded row pointer of one factor (row lengths rounded up to `align` blocks).
int padded_rows(int nb, const int *d_row, int align, std::vector<int> &prp) {
    std::vector<int> rp(nb + 1);
    ISAI_HIP(hipMemcpy(rp.data(), d_row, sizeof(int) * (nb + 1), hipMemcpyDeviceToHost));
    prp.assign(nb + 1, 0);
    long long tot = 0;
    for (int i = 0; i < nb; ++i) {
        const int len = rp[i + 1] - rp[i];
        tot += (len + align - 1) / align * align;
        if (tot > INT_MAX) {
            set_error("isai: the padded reduced-precision factor exceeds %d blocks", INT_MAX);
            return ISAI_ERR_CAPACITY;
        }
        prp[i + 1] = (int)tot;
    }
    return ISAI_OK;
}

// Reduced-precision copies of the ISAI factors in their line-aligned layout.
int convert_lp(isai_handle *h, int prec, hipStream_t st) {
    if (prec == ISAI_PREC_FP64 || h->lp_prec == prec) return ISAI_OK;
    const int align = prec == ISAI_PREC_FP32 ? 2 : 4;  // blocks per 128-byte line
    if (h->lp_align != align) {
        free_lp(h);
        std::vector<int> pl, pu;
        int s = padded_rows(h->nb, h->d_Lrow, align, pl);
        if (s == ISAI_OK) s = padded_rows(h->nb, h->d_Urow, align, pu);
        if (s != ISAI_OK) return s;
        const size_t nl = (size_t)pl[h->nb], nu = (size_t)pu[h->nb];
        if (hipMalloc(&h->d_Lrow_lp, sizeof(int) * (h->nb + 1)) != hipSuccess ||
            hipMalloc(&h->d_Urow_lp, sizeof(int) * (h->nb + 1)) != hipSuccess ||
            hipMalloc(&h->d_Lcol_lp, sizeof(int) * (nl ? nl : 1)) != hipSuccess ||
            hipMalloc(&h->d_Ucol_lp, sizeof(int) * (nu ? nu : 1)) != hipSuccess ||
            hipMalloc(&h->d_Llp, 4 * 16 * (nl ? nl : 1)) != hipSuccess ||
            hipMalloc(&h->d_Ulp, 4 * 16 * (nu ? nu : 1)) != hipSuccess) {
            (void)hipGetLastError();
            free_lp(h);
            set_error("isai: cannot allocate reduced-precision factors");
            return ISAI_ERR_ALLOC;
        }
        if (hipMemcpy(h->d_Lrow_lp, pl.data(), sizeof(int) * (h->nb + 1), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(h->d_Urow_lp, pu.data(), sizeof(int) * (h->nb + 1), hipMemcpyHostToDevice) != hipSuccess) {
            free_lp(h);
            set_error("isai: uploading the reduced-precision patterns failed");
            return ISAI_ERR_HIP;
        }
        h->nnzbL_lp = (int)nl;
        h->nnzbU_lp = (int)nu;
        h->lp_align = align;
    }
    int s = pad_convert_lowp(h->nb, h->d_Lrow, h->d_Lcol, h->d_Lval, h->d_Lrow_lp, h->d_Lcol_lp, h->d_Llp, prec, st);
    if (s == ISAI_OK)
        s = pad_convert_lowp(h->nb, h->d_Urow, h->d_Ucol, h->d_Uval, h->d_Urow_lp, h->d_Ucol_lp, h->d_Ulp, prec, st);
    if (s == ISAI_OK) h->lp_prec = prec;
    return s;
}

// True when the preconditioner's results are FP32 numbers (the reduced-precision
// ISAI forms), so that they can be held in FP32 without changing a bit.
bool apply_out_f32(const isai_handle *h) {
    return h->precond == ISAI_PRECOND_ISAI && h->precision != ISAI_PREC_FP64 && h->bs == 4;
}

// z = M v; tmp is a work vector (ISAI form).  out_f32 (only when apply_out_f32):
// z is written as FP32 values.
int precond_apply(isai_handle *h, const double *v, double *z, double *tmp, hipStream_t st, const int *skip,
                  bool out_f32) {
    switch (h->precond) {
        case ISAI_PRECOND_NONE:
            ISAI_HIP(hipMemcpyAsync(z, v, sizeof(double) * h->n, hipMemcpyDeviceToDevice, st));
            return ISAI_OK;
        case ISAI_PRECOND_ISAI: {
            const isai_precision_t p = (isai_precision_t)h->precision;
            if (p != ISAI_PREC_FP64) {
                // line-aligned reduced-precision copies; M_L v is kept in FP32, the
                // precision the M_U product reads it in: half the intermediate's
                // bytes, same result
                int s = spmv_lowp_launch(h->nb, h->nnzbL_lp, p, h->d_Lrow_lp, h->d_Lcol_lp, h->d_Llp, v, false, tmp,
                                         true, st, skip, true);
                if (s == ISAI_OK)
                    s = spmv_lowp_launch(h->nb, h->nnzbU_lp, p, h->d_Urow_lp, h->d_Ucol_lp, h->d_Ulp, tmp, true, z,
                                         out_f32, st, skip, true);
                return s;
            }
            int s = spmv_launch(h->nb, h->nnzbL, h->bs, p, h->d_Lrow, h->d_Lcol, h->d_Lval, v, tmp, st, skip);
            if (s == ISAI_OK)
                s = spmv_launch(h->nb, h->nnzbU, h->bs, p, h->d_Urow, h->d_Ucol, h->d_Uval, tmp, z, st, skip);
            return s;
        }
        case ISAI_PRECOND_ILU_SPTRSV:
            return sptrsv_solve_internal(h->sv, h->d_LU, v, z, st, skip);
        default:
            set_error("isai: unknown preconditioner %d", h->precond);
            return ISAI_ERR_INVALID_ARG;
    }
}

int true_residual(isai_handle *h, const double *b, const double *x, double *r, double *relres, hipStream_t st) {
    int s = spmv_launch(h->nb, h->nnzb, h->bs, ISAI_PREC_FP64, h->d_row, h->d_col, h->d_val, x, r, st, nullptr);
    if (s != ISAI_OK) return s;
    residual_kernel<<<kRedBlocks, kRedThreads, 0, st>>>(h->n, b, r, h->S);
    residual_final<<<1, kRedThreads, 0, st>>>(h->S, h->d_res2);
    ISAI_LAUNCH_CHECK();
    ISAI_HIP(hipMemcpyAsync(h->h_res2, h->d_res2, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipStreamSynchronize(st));
    const double b2 = h->h_res2[1] > 0.0 ? h->h_res2[1] : 1.0;
    *relres = std::sqrt(h->h_res2[0]) / std::sqrt(b2);
    return ISAI_OK;
}

// Copies done, first, iters, maxit, jend of the device record to the host.
int read_state(isai_handle *h, hipStream_t st) {
    ISAI_HIP(hipMemcpyAsync(h->h_done, &h->S->done, sizeof(int) * kStateInts, hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipStreamSynchronize(st));
    return ISAI_OK;
}

int read_done(isai_handle *h, hipStream_t st, int *done) {
    const int s = read_state(h, st);
    *done = h->h_done[0];
    return s;
}

// Iterations launched between two host reads of the device state; iterations
// launched after convergence return at once.
constexpr int kItersPerRead = 3;

#define TRY(x)                          \
    do {                                \
        int _s = (x);                   \
        if (_s != ISAI_OK) return _s;   \
    } while (0)

void drop_iter_graph(isai_handle *h) {
    if (h->it_exec) (void)hipGraphExecDestroy(h->it_exec);
    if (h->it_graph) (void)hipGraphDestroy(h->it_graph);
    h->it_exec = nullptr;
    h->it_graph = nullptr;
    h->it_keyi[0] = h->it_keyi[1] = -1;
}

int bicgstab(isai_handle *h, double tol, int maxit, const double *b, double *x, int *iters, double *relres,
             hipStream_t st) {
    TRY(ensure_work(h, 9));
    const size_t n = h->n;
    double *r = h->work, *rh = r + n, *p = rh + n, *v = p + n, *ph = v + n, *s = ph + n, *sh = s + n, *t = sh + n,
           *tmp = t + n;
    const int *skip = &h->S->done;
    const int G = spmv_dot_groups(h->nb);
    double *part = h->d_wgpart;
    // p^ and s^ held in FP32 when the ISAI apply produces FP32 numbers: the A
    // products and the x update read the same values with half the bytes
    const bool hat32 = apply_out_f32(h);
    // one iteration: kernels only (each returns at once once the device record says
    // done, so iterations past convergence or the limit change nothing)
    auto iteration = [&]() -> int {
        bicg_p_kernel<<<vec_grid(n), 256, 0, st>>>(n, r, p, v, h->S);
        TRY(precond_apply(h, p, ph, tmp, st, skip, hat32));
        TRY(spmv_dot_launch(h->nb, h->d_row, h->d_col, h->d_val, ph, hat32, v, rh, 1, part, st, skip, h->nnzb));
        bicg_alpha_final_wg<<<1, kWgSumThreads, 0, st>>>(h->S, part, G);
        bicg_s_kernel<<<kRedBlocks, kRedThreads, 0, st>>>(n, r, v, s, h->S);
        bicg_snorm_final<<<1, kRedThreads, 0, st>>>(h->S);
        TRY(precond_apply(h, s, sh, tmp, st, skip, hat32));
        TRY(spmv_dot_launch(h->nb, h->d_row, h->d_col, h->d_val, sh, hat32, t, s, 2, part, st, skip, h->nnzb));
        bicg_omega_final_wg<<<1, kWgSumThreads, 0, st>>>(h->S, part, G);
        if (hat32)
            bicg_xr_kernel<float><<<kRedBlocks, kRedThreads, 0, st>>>(n, x, r, (const float *)ph, (const float *)sh, s,
                                                                     t, rh, h->S);
        else
            bicg_xr_kernel<double><<<kRedBlocks, kRedThreads, 0, st>>>(n, x, r, ph, sh, s, t, rh, h->S);
        bicg_rnorm_final<<<1, kRedThreads, 0, st>>>(h->S);
        ISAI_LAUNCH_CHECK();
        return ISAI_OK;
    };
    // graph mode: kItersPerRead iterations captured once per configuration and
    // replayed; the device record stops at maxit, so a replay past it is a no-op
    const bool graph = h->use_graph && st != nullptr;
    if (graph) {
        const void *key[4] = {x, h->work, h->d_Llp, h->sv};
        const bool same = h->it_exec && std::equal(key, key + 4, h->it_key) && h->it_keyi[0] == h->precond &&
                          h->it_keyi[1] == h->precision;
        if (!same) {
            drop_iter_graph(h);
            ISAI_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            int cs = ISAI_OK;
            for (int k = 0; k < kItersPerRead && cs == ISAI_OK; ++k) cs = iteration();
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(st, &g);
            if (cs != ISAI_OK) {
                if (g) (void)hipGraphDestroy(g);
                return cs;
            }
            ISAI_HIP(e);
            h->it_graph = g;
            ISAI_HIP(hipGraphInstantiate(&h->it_exec, g, nullptr, nullptr, 0));
            std::copy(key, key + 4, h->it_key);
            h->it_keyi[0] = h->precond;
            h->it_keyi[1] = h->precision;
        }
    }
    *iters = 0;
    bicg_start_kernel<<<1, 1, 0, st>>>(h->S, maxit);
    for (int restart = 0; restart < 8; ++restart) {
        TRY(spmv_launch(h->nb, h->nnzb, h->bs, ISAI_PREC_FP64, h->d_row, h->d_col, h->d_val, x, r, st, nullptr));
        bicg_init_kernel<<<kRedBlocks, kRedThreads, 0, st>>>(n, b, r, rh, p, v, h->S);
        bicg_init_final<<<1, kRedThreads, 0, st>>>(h->S, tol);
        ISAI_LAUNCH_CHECK();
        TRY(read_state(h, st));
        int done = h->h_done[0];
        for (int launched = h->h_done[2]; !done && launched < maxit;) {
            if (graph) {
                ISAI_HIP(hipGraphLaunch(h->it_exec, st));
                launched += kItersPerRead;
            } else {
                for (int k = 0; k < kItersPerRead && launched < maxit; ++k, ++launched) TRY(iteration());
            }
            TRY(read_state(h, st));
            done = h->h_done[0];
        }
        *iters = h->h_done[2];
        TRY(true_residual(h, b, x, r, relres, st));
        if (*relres <= tol) return ISAI_OK;
        if (*iters >= maxit) break;
        // recursive residual drifted from the true one, or a breakdown: restart
    }
    set_error("bicgstab: not converged after %d iterations (relres %.3e)", *iters, *relres);
    return ISAI_ERR_NOT_CONVERGED;
}

int fgmres(isai_handle *h, int m, double tol, int maxit, const double *b, double *x, int *iters, double *relres,
           hipStream_t st) {
    if (m <= 0) m = 30;
    if (m > kMaxRestart) {
        set_error("fgmres: restart %d above %d", m, kMaxRestart);
        return ISAI_ERR_INVALID_ARG;
    }
    TRY(ensure_work(h, (size_t)2 * m + 3));
    h->last_restart = m;
    const size_t n = h->n;
    double *V = h->work, *Z = V + n * (m + 1), *w = Z + n * m, *tmp = w + n;
    const int *skip = &h->S->done;
    int it = 0;
    *relres = INFINITY;
    while (it < maxit) {
        TRY(spmv_launch(h->nb, h->nnzb, h->bs, ISAI_PREC_FP64, h->d_row, h->d_col, h->d_val, x, w, st, nullptr));
        residual_kernel<<<kRedBlocks, kRedThreads, 0, st>>>(n, b, w, h->S);
        fg_beta_final<<<1, kRedThreads, 0, st>>>(h->S, tol);
        fg_scale_kernel<<<vec_grid(n), 256, 0, st>>>(n, w, V, h->S, 0, 0, 1);
        ISAI_LAUNCH_CHECK();
        int done = 0;
        TRY(read_done(h, st, &done));
        if (done) break;
        const int steps = std::min(m, maxit - it);
        for (int j = 0; j < steps; ++j) {
            double *vj = V + n * j, *zj = Z + n * j;
            TRY(precond_apply(h, vj, zj, tmp, st, skip, false));
            TRY(spmv_launch(h->nb, h->nnzb, h->bs, ISAI_PREC_FP64, h->d_row, h->d_col, h->d_val, zj, w, st, skip));
            const int nv = j + 1;
            const dim3 dgrid(kRedBlocks, (nv + kDotChunk - 1) / kDotChunk);
            for (int pass = 0; pass < 2; ++pass) {
                fg_mdot_kernel<<<dgrid, kMdotThreads, 0, st>>>(n, w, V, nv, h->S, pass);
                fg_hsum_final<<<nv, kRedThreads, 0, st>>>(h->S, j, pass);
                fg_axpy_multi_kernel<<<kRedBlocks, kRedThreads, 0, st>>>(n, w, V, nv, h->S, pass);
            }
            fg_givens_final<<<1, kRedThreads, 0, st>>>(h->S, j);
            fg_scale_next_kernel<<<vec_grid(n), 256, 0, st>>>(n, w, V + n * (j + 1), h->S);
            ISAI_LAUNCH_CHECK();
            if ((j + 1) % kItersPerRead == 0 || j + 1 == steps) {
                TRY(read_done(h, st, &done));
                if (done) break;
            }
        }
        TRY(read_state(h, st));
        it += h->h_done[4];  // Arnoldi steps the device took in this cycle
        fg_solve_y_kernel<<<1, 64, 0, st>>>(h->S);
        fg_update_x_kernel<<<vec_grid(n), 256, 0, st>>>(n, x, Z, h->S);
        ISAI_LAUNCH_CHECK();
        TRY(true_residual(h, b, x, w, relres, st));
        if (*relres <= tol) break;
    }
    *iters = it;
    if (*relres <= tol) return ISAI_OK;
    set_error("fgmres: not converged after %d iterations (relres %.3e)", it, *relres);
    return ISAI_ERR_NOT_CONVERGED;
}

// ------------------------------------------------- mixed-precision form
constexpr int kLogSteps = 16;  // inner outcomes held on the device between two host reads

// The FP32 copy of A that the inner solves multiply with (converted once per
// handle).  The FP32 product streams its values with non-temporal loads, which
// read a 128-byte line twice when two lane groups share it, so the copy is laid
// out on lines: every row padded with zero blocks to an even block count when that
// adds at most 5 % of the blocks (the 27-pt stencil: 3.5 %), otherwise row pairs
// stored contiguously from a line and the pair's second row padded (7-block rows:
// 13.5 % as single rows, 0.3 % as pairs).  The products equal those with the
// unpadded values: a zero block adds exact zeros.
int build_a32(isai_handle *h, hipStream_t st) {
    const int nb = h->nb;
    std::vector<int> rp(nb + 1), prp(nb + 1, 0);
    ISAI_HIP(hipMemcpy(rp.data(), h->d_row, sizeof(int) * (nb + 1), hipMemcpyDeviceToHost));
    long long odd = 0;
    for (int i = 0; i < nb; ++i) odd += (rp[i + 1] - rp[i]) & 1;
    const int layout = odd * 20 <= h->nnzb ? kA32RowsAligned : kA32PairsAligned;
    long long tot = 0;
    for (int i = 0; i < nb && tot <= INT_MAX;) {
        const long long la = rp[i + 1] - rp[i];
        if (layout == kA32RowsAligned || i + 1 == nb) {
            tot += (la + 1) / 2 * 2;
            prp[++i] = (int)std::min(tot, (long long)INT_MAX);
        } else {
            const long long lb = rp[i + 2] - rp[i + 1];
            prp[i + 1] = (int)std::min(tot + la, (long long)INT_MAX);
            tot += (la + lb + 1) / 2 * 2;
            prp[i + 2] = (int)std::min(tot, (long long)INT_MAX);
            i += 2;
        }
    }
    if (tot > INT_MAX) {
        set_error("isai: the line-aligned FP32 copy of A exceeds %d blocks", INT_MAX);
        return ISAI_ERR_CAPACITY;
    }
    if (hipMalloc(&h->d_A32row, sizeof(int) * (nb + 1)) != hipSuccess ||
        hipMalloc(&h->d_A32col, sizeof(int) * (size_t)std::max(tot, 1LL)) != hipSuccess ||
        hipMalloc(&h->d_A32, sizeof(float) * 16 * (size_t)std::max(tot, 1LL)) != hipSuccess) {
        (void)hipFree(h->d_A32row);
        (void)hipFree(h->d_A32col);
        (void)hipFree(h->d_A32);
        h->d_A32row = h->d_A32col = nullptr;
        h->d_A32 = nullptr;
        set_error("isai_solve: cannot allocate the FP32 copy of A (%lld blocks)", tot);
        return ISAI_ERR_ALLOC;
    }
    h->nnzb32 = tot;
    h->a32_layout = layout;
    ISAI_HIP(hipMemcpyAsync(h->d_A32row, prp.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice, st));
    TRY(pad_convert_lowp(nb, h->d_row, h->d_col, h->d_val, h->d_A32row, h->d_A32col, h->d_A32, ISAI_PREC_FP32, st));
    // prp is a host temporary: wait for its upload before it goes out of scope
    ISAI_HIP(hipStreamSynchronize(st));
    return ISAI_OK;
}

A32Operand a32_operand(const isai_handle *h) {
    A32Operand A;
    A.nnzb = h->nnzb32;
    A.rp = h->d_A32row;
    A.ci = h->d_A32col;
    A.val = h->d_A32;
    A.layout = h->a32_layout;
    return A;
}

// Buffers of the inner solve: the FP32 copy of A (build_a32), ten FP32 vectors, the
// partial sums, the scalar record and the outcome log.
int ensure_inner(isai_handle *h, hipStream_t st) {
    if (h->IS) return ISAI_OK;
    const size_t n = h->n;
    if (!h->d_A32) TRY(build_a32(h, st));
    if (hipMalloc(&h->inner_vecs, sizeof(float) * 10 * n) != hipSuccess ||
        hipMalloc(&h->inner_part, sizeof(double) * inner_part_doubles(h->nb)) != hipSuccess ||
        hipMalloc(&h->IS, sizeof(InnerScalars)) != hipSuccess ||
        hipMalloc(&h->d_log, sizeof(InnerLog) * kLogSteps) != hipSuccess ||
        hipHostMalloc(&h->h_inner, sizeof(int) * 2) != hipSuccess ||
        hipHostMalloc(&h->h_log, sizeof(InnerLog) * kLogSteps) != hipSuccess) {
        (void)hipFree(h->inner_vecs);
        (void)hipFree(h->inner_part);
        (void)hipFree(h->IS);
        (void)hipFree(h->d_log);
        if (h->h_inner) (void)hipHostFree(h->h_inner);
        if (h->h_log) (void)hipHostFree(h->h_log);
        h->inner_vecs = nullptr, h->inner_part = nullptr, h->IS = nullptr;
        h->d_log = nullptr, h->h_inner = nullptr, h->h_log = nullptr;
        set_error("isai_solve: cannot allocate the mixed-precision inner solve");
        return ISAI_ERR_ALLOC;
    }
    return ISAI_OK;
}

InnerVecs inner_vecs(isai_handle *h) {
    const size_t n = h->n;
    float *b = h->inner_vecs;
    InnerVecs w;
    w.r = b, w.rh = b + n, w.p = b + 2 * n, w.v = b + 3 * n, w.s = b + 4 * n, w.t = b + 5 * n;
    w.ph = b + 6 * n, w.sh = b + 7 * n, w.x = b + 8 * n, w.tmp = b + 9 * n;
    w.part = h->inner_part;
    w.rpart = h->inner_part + 2 * (size_t)mixed_dot_groups(h->nb);
    return w;
}

// z = (inner BiCGSTAB on A z = v_j, tolerance from the outer residual estimate g_j,
// at most budget iterations); its outcome goes to log.  When the inner solve of
// the same outer step has run before, the host launches the iteration count it
// needed then and does not wait: the outer loop is flexible, so an inner solve
// that stops short only costs accuracy, and the outcome is read back later with
// the outer state (inner_settle adjusts the next prediction).  Otherwise the host
// launches an estimate from the iterations per decade seen so far and tops up one
// at a time, reading the inner state after each batch.  Iterations launched after
// convergence return at once.  A host read idles the GPU for about 30 us; a
// skipped iteration costs about 40 us of empty launches.
int inner_solve(isai_handle *h, const InnerPrecond *M, const double *v, double *z, const double *gj, int budget,
                int step, InnerLog *log, int *launched_out, hipStream_t st) {
    InnerVecs w = inner_vecs(h);
    TRY(inner_init(h->n, v, &w, h->IS, &h->S->done, gj, &h->S->tol2, h->inner_tol, st));
    const int hist = step < 64 ? h->inner_hist[step] : 0;
    int batch = hist > 0 ? hist : (int)std::ceil(h->inner_rate * -std::log10(h->inner_tol));
    batch = std::max(1, std::min(budget, batch));
    int launched = 0;
    while (true) {
        for (int k = 0; k < batch; ++k, ++launched)
            TRY(inner_iteration(h->n, h->nb, a32_operand(h), M, &w, h->IS, launched == 0, st));
        if (hist > 0) break;
        ISAI_HIP(hipMemcpyAsync(h->h_inner, &h->IS->done, sizeof(int) * 2, hipMemcpyDeviceToHost, st));
        ISAI_HIP(hipStreamSynchronize(st));
        if (h->h_inner[0] || launched >= budget) break;
        batch = 1;
    }
    TRY(inner_finish(h->n, &w, z, h->IS, log, st));
    *launched_out = launched;
    return ISAI_OK;
}

// Buffers of the inner GMRES cycle: k + 1 basis vectors and three work vectors in
// FP32, and its own scalar record.
int ensure_inner_gmres(isai_handle *h, int k) {
    if (h->Sin && h->ig_cap >= k) return ISAI_OK;
    (void)hipFree(h->ig_vecs);
    h->ig_vecs = nullptr;
    h->ig_cap = 0;
    if (!h->Sin && hipMalloc(&h->Sin, sizeof(Scalars)) != hipSuccess) {
        h->Sin = nullptr;
        set_error("isai_solve: cannot allocate the inner GMRES record");
        return ISAI_ERR_ALLOC;
    }
    if (hipMalloc(&h->ig_vecs, sizeof(float) * (size_t)(k + 4) * h->n) != hipSuccess) {
        h->ig_vecs = nullptr;
        set_error("isai_solve: cannot allocate the inner GMRES basis (%d vectors)", k + 4);
        return ISAI_ERR_ALLOC;
    }
    h->ig_cap = k;
    return ISAI_OK;
}

// z = P u with u from one inner GMRES cycle of at most k steps on A z = v, in FP32
// storage: q_i in FP32, zh = M q_i (the handle's reduced-precision ISAI factors,
// FP32 accumulation), w = A32 zh (FP32 A), one classical Gram-Schmidt pass (FP64
// dot products and updates, w stored in FP32), FP64 Hessenberg and Givens rotations
// in the cycle's own scalar record.  The cycle needs no host
// read: its kernels return at once after it converged, and its final steps (y, u =
// Q y, z = M u with z in FP64) run unless the outer loop is done.  *log receives the
// steps taken.
int inner_gmres(isai_handle *h, const InnerPrecond *M, const double *v, double *z, const double *gj, int k,
                InnerLog *log, hipStream_t st) {
    const size_t n = h->n;
    float *Q = h->ig_vecs, *w = Q + n * (size_t)(k + 1), *tmp = w + n, *zh = tmp + n;
    Scalars *Si = h->Sin;
    const int *skip = &Si->done;
    ig_start_kernel<<<vec_grid(n), 256, 0, st>>>(n, v, Q, Si, &h->S->done, gj, &h->S->tol2, h->inner_tol);
    ISAI_LAUNCH_CHECK();
    for (int i = 0; i < k; ++i) {
        const float *qi = Q + n * (size_t)i;
        TRY(spmv_lowp_launch(h->nb, M->nnzbL, M->prec, M->Lrow, M->Lcol, M->Lval, qi, true, tmp, true, st, skip,
                             true));
        TRY(spmv_lowp_launch(h->nb, M->nnzbU, M->prec, M->Urow, M->Ucol, M->Uval, tmp, true, zh, true, st, skip,
                             true));
        TRY(mixed_spmv_f32_dot(h->nb, h->nnzb32, h->d_A32row, h->d_A32col, h->d_A32, zh, w, nullptr, 0, nullptr, st,
                               skip, h->a32_layout));
        // one classical Gram-Schmidt pass: the cycle only has to reach the inner
        // tolerance, and a second pass changed neither its steps nor the outer steps
        const int nv = i + 1;
        const dim3 dgrid(kRedBlocks, (nv + kDotChunk - 1) / kDotChunk);
        fg_mdot_kernel<<<dgrid, kMdotThreads, 0, st>>>(n, (const float *)w, (const float *)Q, nv, Si, 0);
        fg_hsum_final<<<nv, kRedThreads, 0, st>>>(Si, i, 0);
        fg_axpy_multi_kernel<<<kRedBlocks, kRedThreads, 0, st>>>(n, w, (const float *)Q, nv, Si, 0);
        fg_givens_final<<<1, kRedThreads, 0, st>>>(Si, i);
        fg_scale_next_kernel<<<vec_grid(n), 256, 0, st>>>(n, (const float *)w, Q + n * (size_t)(i + 1), Si);
        ISAI_LAUNCH_CHECK();
    }
    fg_solve_y_kernel<<<1, 64, 0, st>>>(Si);
    ig_combine_kernel<<<vec_grid(n), 256, 0, st>>>(n, Q, Si, zh, log);
    ISAI_LAUNCH_CHECK();
    const int *oskip = &h->S->done;
    TRY(spmv_lowp_launch(h->nb, M->nnzbL, M->prec, M->Lrow, M->Lcol, M->Lval, zh, true, tmp, true, st, oskip, true));
    TRY(spmv_lowp_launch(h->nb, M->nnzbU, M->prec, M->Urow, M->Ucol, M->Uval, tmp, true, z, false, st, oskip, true));
    return ISAI_OK;
}

// An inner outcome read back: a solve that converged sets the next prediction for
// its outer step to its count, one that ran out of launched iterations without
// converging to one more.
void inner_settle(isai_handle *h, int step, int launched, const InnerLog &lg) {
    if (lg.done == 1 && lg.iters > 0) {
        h->inner_rate = 0.5 * h->inner_rate + 0.5 * (lg.iters / -std::log10(lg.tol_in));
        if (step < 64) h->inner_hist[step] = lg.iters;
    } else if (lg.done == 0 && step < 64 && h->inner_hist[step] > 0) {
        h->inner_hist[step] = std::min(h->inner_maxit, launched + 1);
    }
}

// FP64 FGMRES(m) whose flexible preconditioner is the inner mixed-precision solve:
// z_j = inner(v_j).  The residual r = b - A x, the Arnoldi products w = A z_j, the
// Gram-Schmidt orthogonalisation and the Givens rotations are FP64.  The inner
// tolerance for step j asks for the reduction still missing, 0.5 tol ||b|| /
// |g_j| (|g_j| = the outer residual estimate), between inner_tol and 0.5,
// computed on the device.  When the last converged solve took P outer steps, the
// host launches the first P steps without reading the state in between (steps
// launched after convergence return at once), then reads the convergence word and
// the inner outcomes once.  *iters = inner BiCGSTAB iterations, bounded by maxit.
// gmres_inner: the inner solve is one GMRES cycle of at most inner_maxit steps
// (capped at kMaxRestart) instead, launched without host reads; *iters = its steps.
int fgmres_mixed(isai_handle *h, int m, double tol, int maxit, const double *b, double *x, int *iters,
                 double *relres, hipStream_t st, bool gmres_inner) {
    if (m <= 0) m = 10;
    if (m > kMaxRestart) {
        set_error("fgmres: restart %d above %d", m, kMaxRestart);
        return ISAI_ERR_INVALID_ARG;
    }
    if (h->precond != ISAI_PRECOND_ISAI) {
        set_error("mixed-precision solve: the inner solve needs the ISAI preconditioner");
        return ISAI_ERR_UNSUPPORTED;
    }
    // the inner ISAI apply runs in the handle's storage precision (FP32 for FP64)
    const int lp = h->precision == ISAI_PREC_FP64 ? ISAI_PREC_FP32 : h->precision;
    TRY(convert_lp(h, lp, st));
    TRY(ensure_work(h, (size_t)2 * m + 3));
    TRY(ensure_inner(h, st));
    const int kcap = std::min(h->inner_maxit, kMaxRestart);
    if (gmres_inner) TRY(ensure_inner_gmres(h, kcap));
    h->last_restart = m;
    const InnerPrecond M{(isai_precision_t)lp, h->nnzbL_lp, h->nnzbU_lp, h->d_Lrow_lp, h->d_Lcol_lp, h->d_Urow_lp,
                         h->d_Ucol_lp, h->d_Llp, h->d_Ulp};
    const size_t n = h->n;
    double *V = h->work, *Z = V + n * (m + 1), *w = Z + n * m;
    const int *skip = &h->S->done;
    int inner_total = 0, steps = 0, pending = 0, launched_est = 0, real_steps = 0;
    int launched[kLogSteps] = {};
    const int pred = h->outer_pred;
    bool converged = false;
    *relres = INFINITY;
    // reads the convergence word and the outcomes of the pending inner solves
    auto read_back = [&](int *done) -> int {
        if (pending) ISAI_HIP(hipMemcpyAsync(h->h_log, h->d_log, sizeof(InnerLog) * pending, hipMemcpyDeviceToHost, st));
        TRY(read_done(h, st, done));
        for (int k = 0; k < pending; ++k) {
            inner_total += h->h_log[k].iters;
            real_steps += h->h_log[k].iters > 0;  // steps launched after convergence did nothing
            if (!gmres_inner) inner_settle(h, steps - pending + k, launched[k], h->h_log[k]);
        }
        pending = 0;
        launched_est = inner_total;
        return ISAI_OK;
    };
    bool first_cycle = true;
    while (launched_est < maxit) {
        // the first cycle's product is skipped on the device when x is zero (r = b)
        const int *xz = first_cycle ? &h->S->xzero : nullptr;
        if (first_cycle) {
            xzero_reset_kernel<<<1, 1, 0, st>>>(h->S);
            xzero_scan_kernel<<<vec_grid(n), 256, 0, st>>>(n, x, h->S);
        }
        first_cycle = false;
        TRY(spmv_launch(h->nb, h->nnzb, h->bs, ISAI_PREC_FP64, h->d_row, h->d_col, h->d_val, x, w, st, xz));
        residual_kernel<<<kRedBlocks, kRedThreads, 0, st>>>(n, b, w, h->S, xz);
        fg_beta_final<<<1, kRedThreads, 0, st>>>(h->S, tol);
        fg_scale_kernel<<<vec_grid(n), 256, 0, st>>>(n, w, V, h->S, 0, 0, 1);
        ISAI_LAUNCH_CHECK();
        int done = 0;
        if (pred <= 0 || steps >= pred) {
            TRY(read_done(h, st, &done));
            if (done) {
                converged = true;
                break;
            }
        }
        int jend = 0;
        for (int j = 0; j < m && launched_est < maxit; ++j) {
            double *vj = V + n * j, *zj = Z + n * j;
            const int step = steps++;
            // g_j is the residual estimate before step j (beta for j = 0)
            const double *gj = j == 0 ? &h->S->beta : &h->S->g[j];
            if (gmres_inner) {
                launched[pending] = std::min(kcap, maxit - launched_est);
                TRY(inner_gmres(h, &M, vj, zj, gj, launched[pending], &h->d_log[pending], st));
            } else {
                TRY(inner_solve(h, &M, vj, zj, gj, std::min(h->inner_maxit, maxit - launched_est), step,
                                &h->d_log[pending], &launched[pending], st));
            }
            launched_est += launched[pending];
            ++pending;
            TRY(spmv_launch(h->nb, h->nnzb, h->bs, ISAI_PREC_FP64, h->d_row, h->d_col, h->d_val, zj, w, st, skip));
            const int nv = j + 1;
            const dim3 dgrid(kRedBlocks, (nv + kDotChunk - 1) / kDotChunk);
            for (int pass = 0; pass < 2; ++pass) {
                fg_mdot_kernel<<<dgrid, kMdotThreads, 0, st>>>(n, w, V, nv, h->S, pass);
                fg_hsum_final<<<nv, kRedThreads, 0, st>>>(h->S, j, pass);
                fg_axpy_multi_kernel<<<kRedBlocks, kRedThreads, 0, st>>>(n, w, V, nv, h->S, pass);
            }
            fg_givens_final<<<1, kRedThreads, 0, st>>>(h->S, j);
            fg_scale_next_kernel<<<vec_grid(n), 256, 0, st>>>(n, w, V + n * (j + 1), h->S);
            ISAI_LAUNCH_CHECK();
            jend = j + 1;
            if (pred <= 0 || steps >= pred || pending == kLogSteps || j + 1 == m) {
                TRY(read_back(&done));
                if (done) break;
            }
        }
        if (pending) {
            int d = 0;
            TRY(read_back(&d));
        }
        if (jend == 0) break;
        fg_solve_y_kernel<<<1, 64, 0, st>>>(h->S);
        fg_update_x_kernel<<<vec_grid(n), 256, 0, st>>>(n, x, Z, h->S);
        ISAI_LAUNCH_CHECK();
        TRY(true_residual(h, b, x, w, relres, st));
        if (*relres <= tol) {
            converged = true;
            break;
        }
    }
    *iters = inner_total;
    h->last_outer = real_steps;
    h->last_inner = inner_total;
    if (converged && !std::isfinite(*relres)) TRY(true_residual(h, b, x, w, relres, st));
    if (*relres <= tol) {
        h->outer_pred = real_steps;
        return ISAI_OK;
    }
    h->outer_pred = 0;
    if (!std::isfinite(*relres)) TRY(true_residual(h, b, x, w, relres, st));
    if (*relres <= tol) return ISAI_OK;
    set_error("mixed-precision fgmres: not converged after %d inner iterations (relres %.3e)", inner_total, *relres);
    return ISAI_ERR_NOT_CONVERGED;
}

}  // namespace

namespace isai {

int handle_factors_internal(isai_handle_t h, HandleFactors *o) {
    ISAI_REQUIRE(h && o, "handle_factors_internal: null argument");
    o->nb = h->nb;
    o->nnzb = h->nnzb;
    o->nnzbL = h->nnzbL;
    o->nnzbU = h->nnzbU;
    o->Lrow = h->d_Lrow;
    o->Lcol = h->d_Lcol;
    o->Urow = h->d_Urow;
    o->Ucol = h->d_Ucol;
    o->precision = h->precision;
    o->Lval = h->d_Lval;
    o->Uval = h->d_Uval;
    if (h->precision != ISAI_PREC_FP64) {  // the line-aligned copies and their patterns
        o->nnzbL = h->nnzbL_lp;
        o->nnzbU = h->nnzbU_lp;
        o->Lrow = h->d_Lrow_lp;
        o->Lcol = h->d_Lcol_lp;
        o->Urow = h->d_Urow_lp;
        o->Ucol = h->d_Ucol_lp;
        o->Lval = h->d_Llp;
        o->Uval = h->d_Ulp;
    }
    return ISAI_OK;
}

}  // namespace isai

extern "C" {

int isai_destroy(isai_handle_t h) {
    if (!h) return ISAI_OK;
    drop_iter_graph(h);
    (void)hipFree(h->d_LU);
    (void)hipFree(h->d_Lrow);
    (void)hipFree(h->d_Lcol);
    (void)hipFree(h->d_Urow);
    (void)hipFree(h->d_Ucol);
    (void)hipFree(h->d_Lval);
    (void)hipFree(h->d_Uval);
    free_lp(h);
    (void)hipFree(h->work);
    (void)hipFree(h->S);
    (void)hipFree(h->d_res2);
    (void)hipFree(h->d_wgpart);
    if (h->h_done) (void)hipHostFree(h->h_done);
    if (h->h_res2) (void)hipHostFree(h->h_res2);
    if (h->sv) isai_bsr_sptrsv_destroy(h->sv);
    (void)hipFree(h->own_row);
    (void)hipFree(h->own_col);
    (void)hipFree(h->own_val);
    (void)hipFree(h->d_A32);
    (void)hipFree(h->d_A32row);
    (void)hipFree(h->d_A32col);
    (void)hipFree(h->ig_vecs);
    (void)hipFree(h->Sin);
    (void)hipFree(h->inner_vecs);
    (void)hipFree(h->inner_part);
    (void)hipFree(h->IS);
    (void)hipFree(h->d_log);
    if (h->h_inner) (void)hipHostFree(h->h_inner);
    if (h->h_log) (void)hipHostFree(h->h_log);
    delete h;
    return ISAI_OK;
}

int isai_setup(int nb, int nnzb, int bs, const int *d_row, const int *d_col, const double *d_val,
               const isai_setup_opts_t *opts, isai_handle_t *out, isai_stream_t stream) {
    ISAI_REQUIRE(out && d_row && d_col && d_val, "isai_setup: null argument");
    ISAI_REQUIRE(nb > 0 && nnzb >= nb, "isai_setup: bad size nb=%d nnzb=%d", nb, nnzb);
    if (bs != 4) {
        set_error("isai_setup: the solve path is built for bs = 4 (got %d)", bs);
        return ISAI_ERR_UNSUPPORTED;
    }
    *out = nullptr;
    hipStream_t st = (hipStream_t)stream;
    auto *h = new isai_handle;
    h->nb = nb;
    h->nnzb = nnzb;
    h->bs = bs;
    h->n = (size_t)nb * bs;
    h->d_row = d_row;
    h->d_col = d_col;
    h->d_val = d_val;
    if (opts) {
        h->precond = opts->precond;
        h->precision = opts->precision;
        h->chunk = opts->sptrsv_chunk;
    }
    int s = ISAI_OK;
    auto fail = [&](int code) {
        isai_destroy(h);
        return code;
    };
    const size_t b2 = (size_t)bs * bs;
    if (hipMalloc(&h->d_LU, sizeof(double) * b2 * nnzb) != hipSuccess || hipMalloc(&h->S, sizeof(Scalars)) != hipSuccess ||
        hipMalloc(&h->d_res2, 2 * sizeof(double)) != hipSuccess ||
        hipHostMalloc(&h->h_done, sizeof(int) * kStateInts) != hipSuccess || hipHostMalloc(&h->h_res2, 2 * sizeof(double)) != hipSuccess) {
        set_error("isai_setup: allocation failed");
        return fail(ISAI_ERR_ALLOC);
    }
    hipEvent_t e0, e1, e1b, e2;
    ISAI_HIP(hipEventCreate(&e0));
    ISAI_HIP(hipEventCreate(&e1));
    ISAI_HIP(hipEventCreate(&e1b));
    ISAI_HIP(hipEventCreate(&e2));

    // --- host pattern work (analysis): BCSC / BCSR patterns and maps
    std::vector<int> rp(nb + 1), ci(nnzb);
    ISAI_HIP(hipMemcpyAsync(rp.data(), d_row, sizeof(int) * (nb + 1), hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipMemcpyAsync(ci.data(), d_col, sizeof(int) * nnzb, hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipStreamSynchronize(st));
    struct Fac {
        std::vector<int> mb, mc, trp, tci, map;
        int nmax = 0;
    } F[2];
    for (int f = 0; f < 2; ++f) {
        const int lower = f == 0;
        int nnz = 0;
        if ((s = isai_bcsc_pattern_count(nb, rp.data(), ci.data(), lower, &nnz)) != ISAI_OK) return fail(s);
        F[f].mb.resize(nb + 1);
        F[f].mc.resize(nnz);
        if ((s = isai_bcsc_pattern(nb, rp.data(), ci.data(), lower, F[f].mb.data(), F[f].mc.data())) != ISAI_OK)
            return fail(s);
        int tn = 0;
        isai_tri_bcsr_pattern_count(nb, rp.data(), ci.data(), lower, &tn);
        if (tn != nnz) {
            set_error("isai_setup: %s pattern mismatch: BCSR nnzb = %d, BCSC nnzb = %d (missing diagonal block?)",
                      lower ? "L" : "U", tn, nnz);
            return fail(ISAI_ERR_PATTERN);
        }
        F[f].trp.resize(nb + 1);
        F[f].tci.resize(nnz);
        isai_tri_bcsr_pattern(nb, rp.data(), ci.data(), lower, F[f].trp.data(), F[f].tci.data());
        F[f].map.resize(nnz);
        if ((s = isai_map_bcsc2bcsr(nb, F[f].mb.data(), F[f].mc.data(), F[f].trp.data(), F[f].tci.data(),
                                    F[f].map.data())) != ISAI_OK)
            return fail(s);
        for (int j = 0; j < nb; ++j) F[f].nmax = std::max(F[f].nmax, F[f].mb[j + 1] - F[f].mb[j]);
    }
    h->nnzbL = (int)F[0].mc.size();
    h->nnzbU = (int)F[1].mc.size();

    // --- ISAI build plan, part 1: the LDS Y cache from the column lengths N_j of
    // both factors (the statistics analyze_pair.h prints; the per-column switch of
    // cacheY_isai.h:192).  Every column is cached when the longest fits
    // kCacheCap blocks; otherwise the cache covers 99 % of the columns and the
    // rest read Y back from global memory; otherwise no column is cached.  A
    // cache of kCacheCap blocks keeps a 64-lane workgroup at 16 KB of LDS.
    {
        constexpr int kCacheCap = 31;
        const int nmax = std::max(F[0].nmax, F[1].nmax);
        std::vector<int> cnt((size_t)nmax + 1, 0);
        for (int f = 0; f < 2; ++f)
            for (int j = 0; j < nb; ++j) cnt[F[f].mb[j + 1] - F[f].mb[j]]++;
        const long long k99 = (long long)(size_t)((double)(2LL * nb - 1) * 0.99);
        long long cum = 0;
        int p99 = nmax;
        for (int v = 0; v <= nmax; ++v) {
            cum += cnt[v];
            if (cum > k99) {
                p99 = v;
                break;
            }
        }
        h->plan_ncache = nmax <= kCacheCap ? nmax : (p99 <= kCacheCap ? p99 : 0);
    }

    // --- ILU(0) on a copy of A
    isai_ilu_plan_t ilu = nullptr;
    int zp = -1;
    if ((s = copy_f64_d2d(h->d_LU, d_val, b2 * nnzb, st)) != ISAI_OK) return fail(s);
    if ((s = isai_bsrilu0_analysis(nb, nnzb, bs, d_row, d_col, &ilu, &zp, stream)) != ISAI_OK) return fail(s);
    isai_bsrilu0_plan_info(ilu, &h->nlevels, nullptr);
    ISAI_HIP(hipEventRecord(e0, st));
    s = isai_bsrilu0_factor(ilu, h->d_LU, &zp, stream);
    ISAI_HIP(hipEventRecord(e1, st));
    isai_bsrilu0_destroy(ilu);
    if (s != ISAI_OK) return fail(s);

    // --- ISAI construction
    int *d_buf[2][5] = {};  // mb, mc, pp, pl, ps
    double *d_dinv[2] = {}, *d_mval[2] = {};
    int *d_map[2] = {}, *d_order[2] = {}, *d_colq = nullptr, *d_pairs = nullptr;
    void *d_acc = nullptr;
    int totals[2] = {0, 0};
    auto cleanup = [&]() {
        (void)hipFree(d_pairs);  // d_buf[f][3], d_buf[f][4] point into it
        for (int f = 0; f < 2; ++f) {
            for (int k = 0; k < 3; ++k) (void)hipFree(d_buf[f][k]);
            (void)hipFree(d_dinv[f]);
            (void)hipFree(d_mval[f]);
            (void)hipFree(d_map[f]);
            (void)hipFree(d_order[f]);
        }
        (void)hipFree(d_colq);
        (void)hipFree(d_acc);
    };
    for (int f = 0; f < 2; ++f) {
        const size_t nnz = F[f].mc.size();
        bool ok = hipMalloc(&d_buf[f][0], sizeof(int) * (nb + 1)) == hipSuccess &&
                  hipMalloc(&d_buf[f][1], sizeof(int) * nnz) == hipSuccess &&
                  hipMalloc(&d_buf[f][2], sizeof(int) * (nnz + 1)) == hipSuccess &&
                  hipMalloc(&d_dinv[f], sizeof(double) * b2 * nb) == hipSuccess &&
                  hipMalloc(&d_mval[f], sizeof(double) * b2 * nnz) == hipSuccess &&
                  hipMalloc(&d_map[f], sizeof(int) * nnz) == hipSuccess;
        if (!ok) {
            cleanup();
            set_error("isai_setup: allocation failed");
            return fail(ISAI_ERR_ALLOC);
        }
        ISAI_HIP(hipMemcpyAsync(d_buf[f][0], F[f].mb.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice, st));
        ISAI_HIP(hipMemcpyAsync(d_buf[f][1], F[f].mc.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, st));
        ISAI_HIP(hipMemcpyAsync(d_map[f], F[f].map.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, st));
    }
    int *factor_rows[2] = {nullptr, nullptr}, *factor_cols[2] = {nullptr, nullptr};
    double *factor_vals[2] = {nullptr, nullptr};
    for (int f = 0; f < 2; ++f) {
        const size_t nnz = F[f].mc.size();
        if (hipMalloc(&factor_rows[f], sizeof(int) * (nb + 1)) != hipSuccess ||
            hipMalloc(&factor_cols[f], sizeof(int) * nnz) != hipSuccess ||
            hipMalloc(&factor_vals[f], sizeof(double) * b2 * nnz) != hipSuccess) {
            cleanup();
            set_error("isai_setup: allocation failed");
            return fail(ISAI_ERR_ALLOC);
        }
        ISAI_HIP(hipMemcpyAsync(factor_rows[f], F[f].trp.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice, st));
        ISAI_HIP(hipMemcpyAsync(factor_cols[f], F[f].tci.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, st));
    }
    h->d_Lrow = factor_rows[0];
    h->d_Lcol = factor_cols[0];
    h->d_Lval = factor_vals[0];
    h->d_Urow = factor_rows[1];
    h->d_Ucol = factor_cols[1];
    h->d_Uval = factor_vals[1];
    // build-plan scratch: per-column pair counts and one statistics record per factor
    const size_t acc_bytes = pair_acc_bytes();
    if (hipMalloc(&d_colq, sizeof(int) * 2 * (size_t)std::max(nb, 1)) != hipSuccess ||
        hipMalloc(&d_acc, 2 * acc_bytes) != hipSuccess) {
        cleanup();
        set_error("isai_setup: allocation failed");
        return fail(ISAI_ERR_ALLOC);
    }
    ISAI_HIP(hipStreamSynchronize(st));
    ISAI_HIP(hipEventRecord(e1b, st));
    // Pair counts of both factors and the build plan's per-column statistics, then
    // one read-back of the two totals and the two largest Q_j (the reference reads
    // each total with its own synchronisation, main.cu:690).
    int qmax2[2] = {0, 0};
    ISAI_HIP(hipMemsetAsync(d_acc, 0, 2 * acc_bytes, st));
    for (int f = 0; f < 2 && s == ISAI_OK; ++f)
        s = pairs_count_async(nb, (int)F[f].mc.size(), d_row, d_col, d_buf[f][0], d_buf[f][1], f == 1, d_buf[f][2], st);
    for (int f = 0; f < 2 && s == ISAI_OK && nb > 0; ++f)
        s = pair_column_q_launch(nb, d_buf[f][0], d_buf[f][2], d_colq + (size_t)f * nb,
                                 (char *)d_acc + f * acc_bytes, st);
    if (s != ISAI_OK) { cleanup(); return fail(s); }
    for (int f = 0; f < 2; ++f) {
        ISAI_HIP(hipMemcpyAsync(&totals[f], d_buf[f][2] + F[f].mc.size(), sizeof(int), hipMemcpyDeviceToHost, st));
        ISAI_HIP(hipMemcpyAsync(&qmax2[f], (char *)d_acc + f * acc_bytes + pair_acc_max_q_offset(), sizeof(int),
                                hipMemcpyDeviceToHost, st));
    }
    ISAI_HIP(hipStreamSynchronize(st));
    // one allocation holds both factors' pair lists (PairPLocal, PairSrc)
    {
        const size_t t0 = (size_t)std::max(totals[0], 1), t1 = (size_t)std::max(totals[1], 1);
        if (hipMalloc(&d_pairs, sizeof(int) * 2 * (t0 + t1)) != hipSuccess) {
            cleanup();
            set_error("isai_setup: allocation failed");
            return fail(ISAI_ERR_ALLOC);
        }
        d_buf[0][3] = d_pairs;
        d_buf[0][4] = d_pairs + t0;
        d_buf[1][3] = d_pairs + 2 * t0;
        d_buf[1][4] = d_pairs + 2 * t0 + t1;
    }
    for (int f = 0; f < 2 && s == ISAI_OK; ++f) {
        s = isai_pairs_fill(nb, d_row, d_col, d_buf[f][0], d_buf[f][1], f == 1, d_buf[f][2], d_buf[f][3], d_buf[f][4], stream);
        if (s == ISAI_OK) s = isai_build_dinv(nb, bs, d_row, d_col, h->d_LU, f == 0, d_dinv[f], stream);
    }
    // --- ISAI build plan, part 2: a column's build time grows with its pair count
    // Q_j, and the four columns of a wave run in lockstep.  When a factor's Q
    // distribution is skewed (largest Q_j above kOrderSkew times the mean), its
    // columns are scheduled by descending Q_j (stable counting sort): the longest
    // columns start first and share waves with columns of similar length.
    if (s == ISAI_OK && nb > 0) {
        constexpr double kOrderSkew = 8.0;
        for (int f = 0; f < 2 && s == ISAI_OK; ++f) {
            const int qmax = qmax2[f];
            const double qavg = (double)totals[f] / nb;
            if (qmax <= kOrderSkew * std::max(qavg, 1.0)) continue;
            std::vector<int> q(nb), order(nb);
            std::vector<long long> start((size_t)qmax + 2, 0);
            ISAI_HIP(hipMemcpyAsync(q.data(), d_colq + (size_t)f * nb, sizeof(int) * nb, hipMemcpyDeviceToHost, st));
            ISAI_HIP(hipStreamSynchronize(st));
            for (int j = 0; j < nb; ++j) start[qmax - q[j] + 1]++;  // bucket 0 = largest Q
            for (int v = 0; v <= qmax; ++v) start[v + 1] += start[v];
            for (int j = 0; j < nb; ++j) order[start[qmax - q[j]]++] = j;
            if (hipMalloc(&d_order[f], sizeof(int) * nb) != hipSuccess) {
                cleanup();
                set_error("isai_setup: allocation failed");
                return fail(ISAI_ERR_ALLOC);
            }
            ISAI_HIP(hipMemcpyAsync(d_order[f], order.data(), sizeof(int) * nb, hipMemcpyHostToDevice, st));
            ISAI_HIP(hipStreamSynchronize(st));
            h->plan_ordered[f] = 1;
        }
    }
    if (s == ISAI_OK) {
        FactorArgs Lf{d_buf[0][0], d_buf[0][1], d_buf[0][2], d_buf[0][3], d_buf[0][4], d_dinv[0], d_mval[0], d_order[0]};
        FactorArgs Uf{d_buf[1][0], d_buf[1][1], d_buf[1][2], d_buf[1][3], d_buf[1][4], d_dinv[1], d_mval[1], d_order[1]};
        s = build_dispatch(nb, bs, h->d_LU, &Lf, &Uf, h->plan_ncache, st);
    }
    for (int f = 0; f < 2 && s == ISAI_OK; ++f)
        s = isai_reorder_bcsc_to_bcsr((int)F[f].mc.size(), bs, d_map[f], d_mval[f], factor_vals[f], stream);
    if (s == ISAI_OK && h->precision != ISAI_PREC_FP64) s = convert_lp(h, h->precision, st);
    ISAI_HIP(hipEventRecord(e2, st));
    ISAI_HIP(hipStreamSynchronize(st));
    cleanup();
    if (s != ISAI_OK) return fail(s);
    float t_ilu = 0.f, t_isai = 0.f;
    ISAI_HIP(hipEventElapsedTime(&t_ilu, e0, e1));
    ISAI_HIP(hipEventElapsedTime(&t_isai, e1b, e2));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e1b);
    (void)hipEventDestroy(e2);
    h->ms_ilu = t_ilu;
    h->ms_isai = t_isai;
    h->ms_total = t_ilu + t_isai;
    if (h->precond == ISAI_PRECOND_ILU_SPTRSV) {
        s = isai_bsr_sptrsv_analysis(nb, nnzb, bs, d_row, d_col, h->d_LU, h->chunk, &h->sv, stream);
        if (s != ISAI_OK) return fail(s);
    }
    *out = h;
    return ISAI_OK;
}

int isai_adjoint_handle(isai_handle_t h, isai_handle_t *out, isai_stream_t stream) {
    ISAI_REQUIRE(h && out, "isai_adjoint_handle: null argument");
    ISAI_REQUIRE(h->bs == 4, "isai_adjoint_handle: bs = 4 only");
    *out = nullptr;
    hipStream_t st = (hipStream_t)stream;
    const int nb = h->nb, nnzb = h->nnzb;
    auto *g = new isai_handle;
    g->nb = nb;
    g->nnzb = nnzb;
    g->bs = h->bs;
    g->n = h->n;
    g->precond = h->precond;
    g->precision = h->precision;
    g->chunk = h->chunk;
    g->nnzbL = h->nnzbU;  // M_U^T is lower triangular and is applied first
    g->nnzbU = h->nnzbL;
    std::vector<void *> tmp;
    auto fail = [&](int code) {
        for (void *p : tmp) (void)hipFree(p);
        isai_destroy(g);
        return code;
    };
    auto dmalloc = [&](void **p, size_t bytes) { return hipMalloc(p, bytes ? bytes : 1) == hipSuccess; };
    const size_t b16 = 16 * sizeof(double);
    if (!dmalloc((void **)&g->S, sizeof(Scalars)) || !dmalloc((void **)&g->d_res2, 2 * sizeof(double)) ||
        hipHostMalloc(&g->h_done, sizeof(int) * kStateInts) != hipSuccess || hipHostMalloc(&g->h_res2, 2 * sizeof(double)) != hipSuccess ||
        !dmalloc((void **)&g->own_row, sizeof(int) * (nb + 1)) || !dmalloc((void **)&g->own_col, sizeof(int) * nnzb) ||
        !dmalloc((void **)&g->own_val, b16 * nnzb) || !dmalloc((void **)&g->d_LU, b16 * nnzb)) {
        set_error("isai_adjoint_handle: allocation failed");
        return fail(ISAI_ERR_ALLOC);
    }
    g->d_row = g->own_row;
    g->d_col = g->own_col;
    g->d_val = g->own_val;
    // upload a host int array into a temporary device buffer
    auto upload = [&](const std::vector<int> &v, int **d) -> bool {
        if (!dmalloc((void **)d, sizeof(int) * v.size())) return false;
        tmp.push_back(*d);
        return v.empty() || hipMemcpyAsync(*d, v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice, st) == hipSuccess;
    };
    // --- A^T and the packed factor of A^T
    std::vector<int> rp(nb + 1), ci(nnzb), trp(nb + 1), tci(nnzb), perm(nnzb), trow_of(nnzb), diag(nb, -1);
    ISAI_HIP(hipMemcpyAsync(rp.data(), h->d_row, sizeof(int) * (nb + 1), hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipMemcpyAsync(ci.data(), h->d_col, sizeof(int) * nnzb, hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipStreamSynchronize(st));
    int s = transpose_pattern_host(nb, rp.data(), ci.data(), trp.data(), tci.data(), perm.data());
    if (s != ISAI_OK) return fail(s);
    for (int i = 0; i < nb; ++i) {
        for (int d = trp[i]; d < trp[i + 1]; ++d) trow_of[d] = i;
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] == i) diag[i] = k;
        if (diag[i] < 0) {
            set_error("isai_adjoint_handle: block row %d has no diagonal block", i);
            return fail(ISAI_ERR_PATTERN);
        }
    }
    int *d_perm = nullptr, *d_trow_of = nullptr, *d_diag = nullptr;
    if (!upload(perm, &d_perm) || !upload(trow_of, &d_trow_of) || !upload(diag, &d_diag)) {
        set_error("isai_adjoint_handle: allocation / copy failed");
        return fail(ISAI_ERR_ALLOC);
    }
    ISAI_HIP(hipMemcpyAsync(g->own_row, trp.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice, st));
    ISAI_HIP(hipMemcpyAsync(g->own_col, tci.data(), sizeof(int) * nnzb, hipMemcpyHostToDevice, st));
    if ((s = transpose_blocks(nnzb, 4, d_perm, h->d_val, g->own_val, st)) != ISAI_OK) return fail(s);
    if ((s = transpose_packed_lu(nnzb, d_trow_of, g->own_col, d_perm, d_diag, h->d_LU, g->d_LU, st)) != ISAI_OK)
        return fail(s);
    // --- transposed ISAI factors: slot 0 (applied first) = M_U^T, slot 1 = M_L^T
    struct Src {
        const int *row, *col;
        const double *val;
        int nnz;
    } src[2] = {{h->d_Urow, h->d_Ucol, h->d_Uval, h->nnzbU}, {h->d_Lrow, h->d_Lcol, h->d_Lval, h->nnzbL}};
    int **dst_row[2] = {&g->d_Lrow, &g->d_Urow}, **dst_col[2] = {&g->d_Lcol, &g->d_Ucol};
    double **dst_val[2] = {&g->d_Lval, &g->d_Uval};
    for (int f = 0; f < 2; ++f) {
        const int nz = src[f].nnz;
        std::vector<int> frp(nb + 1), fci(nz), ftrp(nb + 1), ftci(nz), fperm(nz);
        ISAI_HIP(hipMemcpyAsync(frp.data(), src[f].row, sizeof(int) * (nb + 1), hipMemcpyDeviceToHost, st));
        ISAI_HIP(hipMemcpyAsync(fci.data(), src[f].col, sizeof(int) * nz, hipMemcpyDeviceToHost, st));
        ISAI_HIP(hipStreamSynchronize(st));
        if ((s = transpose_pattern_host(nb, frp.data(), fci.data(), ftrp.data(), ftci.data(), fperm.data())) != ISAI_OK)
            return fail(s);
        int *d_fperm = nullptr;
        if (!dmalloc((void **)dst_row[f], sizeof(int) * (nb + 1)) || !dmalloc((void **)dst_col[f], sizeof(int) * nz) ||
            !dmalloc((void **)dst_val[f], b16 * nz) || !upload(fperm, &d_fperm)) {
            set_error("isai_adjoint_handle: allocation failed");
            return fail(ISAI_ERR_ALLOC);
        }
        ISAI_HIP(hipMemcpyAsync(*dst_row[f], ftrp.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice, st));
        ISAI_HIP(hipMemcpyAsync(*dst_col[f], ftci.data(), sizeof(int) * nz, hipMemcpyHostToDevice, st));
        if ((s = transpose_blocks(nz, 4, d_fperm, src[f].val, *dst_val[f], st)) != ISAI_OK) return fail(s);
    }
    if (g->precision != ISAI_PREC_FP64 && (s = convert_lp(g, g->precision, st)) != ISAI_OK) return fail(s);
    ISAI_HIP(hipStreamSynchronize(st));
    for (void *p : tmp) (void)hipFree(p);
    tmp.clear();
    int nlev = 0;
    {
        std::vector<int> lev(nb);
        if (compute_levels(nb, trp.data(), tci.data(), true, lev.data(), &nlev) == ISAI_OK) g->nlevels = nlev;
    }
    if (g->precond == ISAI_PRECOND_ILU_SPTRSV) {
        s = isai_bsr_sptrsv_analysis(nb, nnzb, 4, g->d_row, g->d_col, g->d_LU, g->chunk, &g->sv, stream);
        if (s != ISAI_OK) return fail(s);
    }
    *out = g;
    return ISAI_OK;
}

int isai_setup_times(isai_handle_t h, double *ms_ilu, double *ms_isai, double *ms_total) {
    ISAI_REQUIRE(h, "isai_setup_times: null handle");
    if (ms_ilu) *ms_ilu = h->ms_ilu;
    if (ms_isai) *ms_isai = h->ms_isai;
    if (ms_total) *ms_total = h->ms_total;
    return ISAI_OK;
}

int isai_handle_info(isai_handle_t h, int *nnzbL, int *nnzbU, int *nlevels) {
    ISAI_REQUIRE(h, "isai_handle_info: null handle");
    if (nnzbL) *nnzbL = h->nnzbL;
    if (nnzbU) *nnzbU = h->nnzbU;
    if (nlevels) *nlevels = h->nlevels;
    return ISAI_OK;
}

int isai_handle_factors(isai_handle_t h, const int **Lrow, const int **Lcol, const double **Lval, const int **Urow,
                        const int **Ucol, const double **Uval, const double **LU) {
    ISAI_REQUIRE(h, "isai_handle_factors: null handle");
    if (Lrow) *Lrow = h->d_Lrow;
    if (Lcol) *Lcol = h->d_Lcol;
    if (Lval) *Lval = h->d_Lval;
    if (Urow) *Urow = h->d_Urow;
    if (Ucol) *Ucol = h->d_Ucol;
    if (Uval) *Uval = h->d_Uval;
    if (LU) *LU = h->d_LU;
    return ISAI_OK;
}

int isai_set_solve_graph(isai_handle_t h, int use_graph) {
    ISAI_REQUIRE(h, "isai_set_solve_graph: null handle");
    h->use_graph = use_graph != 0;
    if (!h->use_graph) drop_iter_graph(h);
    return ISAI_OK;
}

int isai_set_preconditioner(isai_handle_t h, int precond, int precision, isai_stream_t stream) {
    ISAI_REQUIRE(h, "isai_set_preconditioner: null handle");
    ISAI_REQUIRE(precond >= 0 && precond <= 2 && precision >= 0 && precision <= 2,
                 "isai_set_preconditioner: bad kind %d / precision %d", precond, precision);
    hipStream_t st = (hipStream_t)stream;
    drop_iter_graph(h);  // the captured iterations name the old factors
    if (precond == ISAI_PRECOND_ILU_SPTRSV && !h->sv) {
        int s = isai_bsr_sptrsv_analysis(h->nb, h->nnzb, h->bs, h->d_row, h->d_col, h->d_LU, h->chunk, &h->sv, stream);
        if (s != ISAI_OK) return s;
    }
    if (precond == ISAI_PRECOND_ISAI && precision != ISAI_PREC_FP64) {
        int s = convert_lp(h, precision, st);
        if (s != ISAI_OK) return s;
    }
    h->precond = precond;
    h->precision = precision;
    return ISAI_OK;
}

int isai_precond_apply(isai_handle_t h, const double *d_v, double *d_z, isai_stream_t stream) {
    ISAI_REQUIRE(h && d_v && d_z, "isai_precond_apply: null argument");
    int s = ensure_work(h, 9);
    if (s != ISAI_OK) return s;
    double *tmp = h->work + 8 * h->n;
    return precond_apply(h, d_v, d_z, tmp, (hipStream_t)stream, nullptr, false);
}

int isai_solve(isai_handle_t h, int solver, int restart, double tol, int maxit, const double *d_b, double *d_x,
               int *iters, double *relres, isai_stream_t stream) {
    ISAI_REQUIRE(h && d_b && d_x && iters && relres, "isai_solve: null argument");
    ISAI_REQUIRE(tol > 0.0 && maxit >= 0, "isai_solve: bad tol / maxit");
    *iters = 0;
    *relres = INFINITY;
    hipStream_t st = (hipStream_t)stream;
    int s;
    if (solver == ISAI_SOLVER_BICGSTAB) {
        s = bicgstab(h, tol, maxit, d_b, d_x, iters, relres, st);
    } else if (solver == ISAI_SOLVER_FGMRES) {
        s = fgmres(h, restart, tol, maxit, d_b, d_x, iters, relres, st);
    } else if (solver == ISAI_SOLVER_FGMRES_MIXED || solver == ISAI_SOLVER_FGMRES_MIXED_GMRES) {
        s = fgmres_mixed(h, restart, tol, maxit, d_b, d_x, iters, relres, st, solver == ISAI_SOLVER_FGMRES_MIXED_GMRES);
    } else {
        set_error("isai_solve: unknown solver %d", solver);
        return ISAI_ERR_INVALID_ARG;
    }
    if (solver != ISAI_SOLVER_FGMRES_MIXED && solver != ISAI_SOLVER_FGMRES_MIXED_GMRES) {
        h->last_outer = 0;
        h->last_inner = *iters;
    }
    // a triangular-solve wait that gave up makes the iterates meaningless
    const int t = isai_handle_status(h, stream);
    return t != ISAI_OK ? t : s;
}

int isai_handle_spmv_f32_dot(isai_handle_t h, const float *d_x, float *d_y, const float *d_a, int ndot,
                             double *d_part, int *groups, int *rows_per_group, isai_stream_t stream) {
    ISAI_REQUIRE(h && d_x && d_y, "isai_handle_spmv_f32_dot: null argument");
    ISAI_REQUIRE(ndot >= 0 && ndot <= 2, "isai_handle_spmv_f32_dot: ndot %d", ndot);
    ISAI_REQUIRE(ndot == 0 || (d_a && d_part), "isai_handle_spmv_f32_dot: null dot operand or partials");
    hipStream_t st = (hipStream_t)stream;
    if (!h->d_A32) TRY(build_a32(h, st));
    if (groups) *groups = mixed_dot_groups(h->nb);
    if (rows_per_group) *rows_per_group = h->a32_layout == kA32PairsAligned ? 2 : 1;
    return mixed_spmv_f32_dot(h->nb, h->nnzb32, h->d_A32row, h->d_A32col, h->d_A32, d_x, d_y, d_a, ndot, d_part, st,
                              nullptr, h->a32_layout);
}

int isai_set_inner_solve(isai_handle_t h, double inner_tol, int inner_maxit) {
    ISAI_REQUIRE(h, "isai_set_inner_solve: null handle");
    ISAI_REQUIRE(inner_tol > 0.0 && inner_tol < 1.0 && inner_maxit > 0,
                 "isai_set_inner_solve: bad inner_tol %g / inner_maxit %d", inner_tol, inner_maxit);
    h->inner_tol = inner_tol;
    h->inner_maxit = inner_maxit;
    return ISAI_OK;
}

int isai_last_solve_stats(isai_handle_t h, int *outer_steps, int *inner_iters) {
    ISAI_REQUIRE(h, "isai_last_solve_stats: null handle");
    if (outer_steps) *outer_steps = h->last_outer;
    if (inner_iters) *inner_iters = h->last_inner;
    return ISAI_OK;
}

int isai_fgmres_arnoldi(isai_handle_t h, int *m, int *steps, const double **d_V, const double **d_Z, double *Hbar,
                        int capacity) {
    ISAI_REQUIRE(h && m && steps, "isai_fgmres_arnoldi: null argument");
    ISAI_REQUIRE(h->last_restart > 0 && h->work, "isai_fgmres_arnoldi: the handle has run no FGMRES solve");
    const int mm = h->last_restart;
    *m = mm;
    ISAI_HIP(hipDeviceSynchronize());
    ISAI_HIP(hipMemcpy(steps, &h->S->jend, sizeof(int), hipMemcpyDeviceToHost));
    if (d_V) *d_V = h->work;
    if (d_Z) *d_Z = h->work + h->n * (mm + 1);
    if (Hbar) {
        ISAI_REQUIRE(capacity >= (mm + 1) * mm, "isai_fgmres_arnoldi: Hbar needs (m + 1) m = %d doubles", (mm + 1) * mm);
        std::vector<double> full((size_t)(kMaxRestart + 1) * kMaxRestart);
        ISAI_HIP(hipMemcpy(full.data(), h->S->Hbar, sizeof(double) * full.size(), hipMemcpyDeviceToHost));
        for (int i = 0; i <= mm; ++i)
            for (int j = 0; j < mm; ++j) Hbar[i * mm + j] = i <= j + 1 ? full[(size_t)i * kMaxRestart + j] : 0.0;
    }
    return ISAI_OK;
}

int isai_handle_status(isai_handle_t h, isai_stream_t stream) {
    ISAI_REQUIRE(h, "isai_handle_status: null handle");
    if (h->sv) return isai_sptrsv_status(h->sv, stream);
    ISAI_HIP(hipStreamSynchronize((hipStream_t)stream));
    return ISAI_OK;
}

int isai_handle_build_plan(isai_handle_t h, int *ncache, int *ordered_L, int *ordered_U) {
    ISAI_REQUIRE(h, "isai_handle_build_plan: null handle");
    if (ncache) *ncache = h->plan_ncache;
    if (ordered_L) *ordered_L = h->plan_ordered[0];
    if (ordered_U) *ordered_U = h->plan_ordered[1];
    return ISAI_OK;
}

int isai_handle_set_spin_limit(isai_handle_t h, int limit) {
    ISAI_REQUIRE(h && limit > 0, "isai_handle_set_spin_limit: bad argument");
    ISAI_REQUIRE(h->sv, "isai_handle_set_spin_limit: the handle has no triangular-solve plan "
                        "(set the ILU(0) preconditioner first)");
    return isai_bsr_sptrsv_set_spin_limit(h->sv, limit);
}

}  // extern "C"
