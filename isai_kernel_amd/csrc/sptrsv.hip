// sptrsv.hip -- sync-free triangular solves with the packed ILU(0) factor on gfx950
// (replaces spTrSolveL / spTrSolveU / csr_cape_SpTRSV, reference cape_sptrsv.h:8-179).
//
// The reference gives every scalar row one thread that spins on a volatile bool
// flag of each dependency and relies on independent thread scheduling inside a
// warp (cape_sptrsv.h:36-51).  A wave64 has no independent thread scheduling, so
// a lane spinning on a value another lane of its own wave must produce would
// never finish.  Here:
//   * one wave owns a CHUNK of consecutive rows (ascending for L, descending for
//     U) taken from an atomic ticket, and walks them in order: dependencies inside
//     the chunk are read from LDS, never waited on;
//   * dependencies on other chunks are waited on with per-row ready flags that
//     live in the solution vector itself: it is pre-filled with a NaN sentinel
//     and every value is published with an agent-scope atomic store, so a reader
//     polling with agent-scope atomic loads sees either the sentinel or the final
//     value -- no fence, no separate flag array, no host synchronisation and no
//     grid-wide barrier;
//   * tickets are handed out in row order, so a wave only ever waits on chunks
//     already held by running waves (forward progress under any dispatch order);
//   * every wait is bounded; an exhausted wait sets an error word instead of
//     hanging, which the host reads with isai_sptrsv_status().
// The last wave of a solve resets the ticket counters, so consecutive solves
// (or graph replays) need no memset.
#include <climits>
#include <vector>

#include "cape_sptrsv.h"
#include "common.hpp"
#include "internal.hpp"
#include "isai_bsr.h"

namespace isai {

constexpr unsigned long long kSentinel = 0x7FF4DEADBEEF0001ull;  // signalling-NaN payload
constexpr int kMaxChunk = 256;
constexpr int kCsrMaxChunk = 1024;

struct Counters {
    unsigned int ticket_l, done_l, ticket_u, done_u;
    int err;
    int spin_limit;
    int pad[2];
};

__device__ __forceinline__ double poll(const double *p, int limit, bool &timed_out) {
    unsigned long long *q = (unsigned long long *)(p);
    unsigned long long v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int spins = 0;
    while (v == kSentinel) {
        if (++spins > limit) {
            timed_out = true;
            break;
        }
        __builtin_amdgcn_s_sleep(1);
        v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return __longlong_as_double((long long)v);
}

__device__ __forceinline__ void publish(double *p, double v) {
    __hip_atomic_store((unsigned long long *)p, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// -------------------------------------------------------------- BSR, bs = 4
// Lane layout per row: 4 lanes per off-diagonal block (lane r holds row r of the
// block), 16 blocks per pass; the per-row sums are reduced over the blocks with
// cross-lane adds, then multiplied by the precomputed inverse of the diagonal
// part (unit-lower for L, upper for U).
template <bool UPPER>
__global__ __launch_bounds__(64) void bsr4_sptrsv_kernel(int nchunks, const int2 *__restrict__ chunks,
                                                         const int *__restrict__ rp, const int *__restrict__ ci,
                                                         const double *__restrict__ LU,
                                                         const int *__restrict__ diagpos,
                                                         const double *__restrict__ dinv,
                                                         const double *rhs, double *out,
                                                         double *sentinel_fill, Counters *cnt,
                                                         const int *__restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double xs[kMaxChunk * 4];
    __shared__ int s_c;
    const int lane = threadIdx.x;
    const int blk = lane >> 2, r = lane & 3;
    unsigned int *ticket = UPPER ? &cnt->ticket_u : &cnt->ticket_l;
    unsigned int *done = UPPER ? &cnt->done_u : &cnt->done_l;
    const int limit = cnt->spin_limit;
    if (lane == 0) s_c = (int)atomicAdd(ticket, 1u);
    __syncthreads();
    const int2 range = chunks[s_c];  // chunks are listed in processing order
    const int lo = range.x, hi = range.y;
    bool timed_out = false;

    for (int step = 0; step < hi - lo; ++step) {
        const int i = UPPER ? hi - 1 - step : lo + step;
        const int dp = diagpos[i];
        const int pbeg = UPPER ? dp + 1 : rp[i];
        const int pend = UPPER ? rp[i + 1] : dp;
        double acc = 0.0;
        for (int base = pbeg; base < pend; base += 16) {
            const int p = base + blk;
            if (p < pend) {
                const int k = ci[p];
                double x0, x1, x2, x3;
                const bool local = UPPER ? (k < hi) : (k >= lo);
                if (local) {
                    const double *xl = xs + (k - lo) * 4;
                    x0 = xl[0]; x1 = xl[1]; x2 = xl[2]; x3 = xl[3];
                } else {
                    const double *xg = out + (size_t)k * 4;
                    x0 = poll(xg + 0, limit, timed_out);
                    x1 = poll(xg + 1, limit, timed_out);
                    x2 = poll(xg + 2, limit, timed_out);
                    x3 = poll(xg + 3, limit, timed_out);
                }
                const v2d a01 = *(const v2d *)(LU + (size_t)p * 16 + r * 4);
                const v2d a23 = *(const v2d *)(LU + (size_t)p * 16 + r * 4 + 2);
                acc = fma(a01.x, x0, acc);
                acc = fma(a01.y, x1, acc);
                acc = fma(a23.x, x2, acc);
                acc = fma(a23.y, x3, acc);
            }
        }
        acc += __shfl_xor(acc, 4);
        acc += __shfl_xor(acc, 8);
        acc += __shfl_xor(acc, 16);
        acc += __shfl_xor(acc, 32);
        const double *bi = rhs + (size_t)i * 4;
        const double z0 = bi[0] - __shfl(acc, 0);
        const double z1 = bi[1] - __shfl(acc, 1);
        const double z2 = bi[2] - __shfl(acc, 2);
        const double z3 = bi[3] - __shfl(acc, 3);
        const double *D = dinv + (size_t)i * 16 + r * 4;
        double v = D[0] * z0;
        v = fma(D[1], z1, v);
        v = fma(D[2], z2, v);
        v = fma(D[3], z3, v);
        if (lane < 4) {
            xs[(i - lo) * 4 + lane] = v;
            publish(out + (size_t)i * 4 + lane, v);
            if (sentinel_fill) sentinel_fill[(size_t)i * 4 + lane] = __longlong_as_double((long long)kSentinel);
        }
        __syncthreads();
    }
    if (timed_out) atomicOr(&cnt->err, 1);
    if (lane == 0) {
        // The last wave of this solve resets the counters for the next one.
        if (atomicAdd(done, 1u) == (unsigned)nchunks - 1) {
            atomicExch(ticket, 0u);
            atomicExch(done, 0u);
        }
    }
}

// ------------------------------------------------------------- CSR, scalar
// cape semantics: L has a unit diagonal that is skipped (cape_sptrsv.h:36-47);
// U divides by the diagonal entry (cape_sptrsv.h:95-96).  The wave's 64 lanes
// split the entries of one row; the row sum is a wave reduction.
template <bool UPPER>
__global__ __launch_bounds__(64) void csr_sptrsv_kernel(int m, int chunk, int nchunks, const int *__restrict__ rp,
                                                        const int *__restrict__ ci, const double *__restrict__ v,
                                                        const double *rhs, double *out,
                                                        double *sentinel_fill, Counters *cnt) {
    __shared__ double xs[kCsrMaxChunk];
    __shared__ int s_c;
    const int lane = threadIdx.x;
    unsigned int *ticket = UPPER ? &cnt->ticket_u : &cnt->ticket_l;
    unsigned int *done = UPPER ? &cnt->done_u : &cnt->done_l;
    const int limit = cnt->spin_limit;
    if (lane == 0) s_c = (int)atomicAdd(ticket, 1u);
    __syncthreads();
    const int t = s_c;
    const int c = UPPER ? nchunks - 1 - t : t;
    const int lo = c * chunk, hi = min(lo + chunk, m);
    bool timed_out = false;
    for (int step = 0; step < hi - lo; ++step) {
        const int i = UPPER ? hi - 1 - step : lo + step;
        double acc = 0.0, diag = 0.0;
        int has_diag = 0;
        for (int p = rp[i] + lane; p < rp[i + 1]; p += 64) {
            const int k = ci[p];
            if (k == i) {
                diag = v[p];
                has_diag = 1;
                continue;
            }
            if (UPPER ? (k < i) : (k > i)) continue;
            double xk;
            const bool local = UPPER ? (k < hi) : (k >= lo);
            if (local) xk = xs[k - lo];
            else xk = poll(out + k, limit, timed_out);
            acc = fma(v[p], xk, acc);
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
        double val = rhs[i] - acc;
        if (UPPER) {
            const unsigned long long bal = __ballot(has_diag);
            const int src = bal ? __ffsll((long long)bal) - 1 : 0;
            const double d = __shfl(diag, src);
            if (!bal) timed_out = true;  // no diagonal entry: unsolvable row
            val = val / d;
        }
        if (lane == 0) {
            xs[i - lo] = val;
            publish(out + i, val);
            if (sentinel_fill) sentinel_fill[i] = __longlong_as_double((long long)kSentinel);
        }
        __syncthreads();
    }
    if (timed_out) atomicOr(&cnt->err, 1);
    if (lane == 0) {
        if (atomicAdd(done, 1u) == (unsigned)nchunks - 1) {
            atomicExch(ticket, 0u);
            atomicExch(done, 0u);
        }
    }
}

__global__ void diag_pos_kernel(int nb, const int *__restrict__ rp, const int *__restrict__ ci, int *__restrict__ dp,
                                int *__restrict__ missing) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    int s = rp[i], e = rp[i + 1] - 1, pos = -1;
    while (s <= e) {
        const int mid = (s + e) >> 1, cc = ci[mid];
        if (cc == i) { pos = mid; break; }
        if (cc < i) s = mid + 1; else e = mid - 1;
    }
    dp[i] = pos;
    if (pos < 0) atomicMin(missing, i);
}

__global__ void fill_sentinel_kernel(double *p, size_t n, const int *__restrict__ skip) {
    if (skip && *skip) return;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        p[k] = __longlong_as_double((long long)kSentinel);
}

static unsigned fill_blocks(size_t n) {
    size_t g = (n + 255) / 256;
    return (unsigned)(g > 8192 ? 8192 : (g ? g : 1));
}

}  // namespace isai

using namespace isai;

struct isai_sptrsv_plan {
    int nb = 0, nnzb = 0, bs = 0, chunk = 0;
    int nchunksL = 0, nchunksU = 0;
    const int *d_row = nullptr, *d_col = nullptr;
    int *d_diag = nullptr;
    int2 *d_chunksL = nullptr, *d_chunksU = nullptr;
    double *d_dinvL = nullptr, *d_dinvU = nullptr, *d_y = nullptr;
    Counters *d_cnt = nullptr;
};

namespace isai {

int sptrsv_solve_internal(isai_sptrsv_plan_t P, const double *d_LU, const double *d_b, double *d_x,
                          hipStream_t st, const int *skip) {
    const size_t n = (size_t)P->nb * 4;
    fill_sentinel_kernel<<<fill_blocks(n), 256, 0, st>>>(P->d_y, n, skip);
    bsr4_sptrsv_kernel<false><<<P->nchunksL, 64, 0, st>>>(P->nchunksL, P->d_chunksL, P->d_row, P->d_col, d_LU,
                                                          P->d_diag, P->d_dinvL, d_b, P->d_y, d_x, P->d_cnt, skip);
    bsr4_sptrsv_kernel<true><<<P->nchunksU, 64, 0, st>>>(P->nchunksU, P->d_chunksU, P->d_row, P->d_col, d_LU,
                                                         P->d_diag, P->d_dinvU, P->d_y, d_x, nullptr, P->d_cnt, skip);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

}  // namespace isai

extern "C" {

int isai_bsr_sptrsv_analysis(int nb, int nnzb, int bs, const int *d_row, const int *d_col, const double *d_LU,
                             int chunk, isai_sptrsv_plan_t *plan, isai_stream_t stream) {
    ISAI_REQUIRE(plan, "isai_bsr_sptrsv_analysis: null output");
    *plan = nullptr;
    ISAI_REQUIRE(nb > 0 && nnzb >= nb && d_row && d_col && d_LU, "isai_bsr_sptrsv_analysis: bad argument");
    if (bs != 4) {
        set_error("isai_bsr_sptrsv: the BSR-native solve is built for bs = 4 (got %d); use the CSR form", bs);
        return ISAI_ERR_UNSUPPORTED;
    }
    if (chunk <= 0) chunk = kMaxChunk;
    ISAI_REQUIRE(chunk <= kMaxChunk, "isai_bsr_sptrsv_analysis: chunk %d above %d", chunk, kMaxChunk);
    hipStream_t st = (hipStream_t)stream;
    // Chunks follow the dependency chains: a new L chunk starts at row i when row i
    // has no block in column i-1 (so it does not wait for the previous chunk's last
    // row); a new U chunk starts (walking down from the last row) when row i has
    // no block in column i+1.  Chunks are capped at `chunk` rows.
    std::vector<int> rp(nb + 1), ci(nnzb);
    ISAI_HIP(hipMemcpyAsync(rp.data(), d_row, sizeof(int) * (nb + 1), hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipMemcpyAsync(ci.data(), d_col, sizeof(int) * nnzb, hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipStreamSynchronize(st));
    auto has = [&](int i, int j) {
        for (int p = rp[i]; p < rp[i + 1]; ++p)
            if (ci[p] == j) return true;
        return false;
    };
    std::vector<int2> cl, cu;
    for (int i = 0; i < nb;) {
        int j = i + 1;
        while (j < nb && j - i < chunk && has(j, j - 1)) ++j;
        cl.push_back(make_int2(i, j));
        i = j;
    }
    for (int i = nb - 1; i >= 0;) {
        int j = i - 1;  // chunk is (j, i]
        while (j >= 0 && i - j < chunk && has(j, j + 1)) --j;
        cu.push_back(make_int2(j + 1, i + 1));
        i = j;
    }
    auto *P = new isai_sptrsv_plan;
    P->nb = nb;
    P->nnzb = nnzb;
    P->bs = bs;
    P->chunk = chunk;
    P->nchunksL = (int)cl.size();
    P->nchunksU = (int)cu.size();
    P->d_row = d_row;
    P->d_col = d_col;
    bool ok = hipMalloc(&P->d_diag, sizeof(int) * nb) == hipSuccess &&
              hipMalloc(&P->d_dinvL, sizeof(double) * 16 * nb) == hipSuccess &&
              hipMalloc(&P->d_dinvU, sizeof(double) * 16 * nb) == hipSuccess &&
              hipMalloc(&P->d_y, sizeof(double) * 4 * nb) == hipSuccess &&
              hipMalloc(&P->d_chunksL, sizeof(int2) * cl.size()) == hipSuccess &&
              hipMalloc(&P->d_chunksU, sizeof(int2) * cu.size()) == hipSuccess &&
              hipMalloc(&P->d_cnt, sizeof(Counters) + sizeof(int)) == hipSuccess;
    if (!ok) {
        isai_bsr_sptrsv_destroy(P);
        set_error("isai_bsr_sptrsv_analysis: hipMalloc failed");
        return ISAI_ERR_ALLOC;
    }
    ISAI_HIP(hipMemcpy(P->d_chunksL, cl.data(), sizeof(int2) * cl.size(), hipMemcpyHostToDevice));
    ISAI_HIP(hipMemcpy(P->d_chunksU, cu.data(), sizeof(int2) * cu.size(), hipMemcpyHostToDevice));
    Counters h{};
    h.spin_limit = 1 << 20;
    int *missing = (int *)(P->d_cnt + 1);
    const int big = INT_MAX;
    ISAI_HIP(hipMemcpy(P->d_cnt, &h, sizeof(h), hipMemcpyHostToDevice));
    ISAI_HIP(hipMemcpy(missing, &big, sizeof(int), hipMemcpyHostToDevice));
    diag_pos_kernel<<<grid_for(nb, 256), 256, 0, st>>>(nb, d_row, d_col, P->d_diag, missing);
    ISAI_LAUNCH_CHECK();
    int miss = INT_MAX;
    ISAI_HIP(hipMemcpyAsync(&miss, missing, sizeof(int), hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipStreamSynchronize(st));
    if (miss != INT_MAX) {
        isai_bsr_sptrsv_destroy(P);
        set_error("isai_bsr_sptrsv_analysis: block row %d has no diagonal block", miss);
        return ISAI_ERR_ZERO_PIVOT;
    }
    *plan = P;
    return isai_bsr_sptrsv_update_values(P, d_LU, stream);
}

int isai_bsr_sptrsv_update_values(isai_sptrsv_plan_t P, const double *d_LU, isai_stream_t stream) {
    ISAI_REQUIRE(P && d_LU, "isai_bsr_sptrsv_update_values: null argument");
    int s = isai_build_dinv(P->nb, 4, P->d_row, P->d_col, d_LU, 1, P->d_dinvL, stream);
    if (s == ISAI_OK) s = isai_build_dinv(P->nb, 4, P->d_row, P->d_col, d_LU, 0, P->d_dinvU, stream);
    return s;
}

int isai_bsr_sptrsv_solve(isai_sptrsv_plan_t P, const double *d_LU, const double *d_b, double *d_x,
                          isai_stream_t stream) {
    ISAI_REQUIRE(P && d_LU && d_b && d_x, "isai_bsr_sptrsv_solve: null argument");
    return isai::sptrsv_solve_internal(P, d_LU, d_b, d_x, (hipStream_t)stream, nullptr);
}

int isai_bsr_sptrsv_set_spin_limit(isai_sptrsv_plan_t P, int limit) {
    ISAI_REQUIRE(P && limit > 0, "isai_bsr_sptrsv_set_spin_limit: bad argument");
    ISAI_HIP(hipMemcpy(&P->d_cnt->spin_limit, &limit, sizeof(int), hipMemcpyHostToDevice));
    return ISAI_OK;
}

int isai_sptrsv_status(isai_sptrsv_plan_t P, isai_stream_t stream) {
    ISAI_REQUIRE(P, "isai_sptrsv_status: null plan");
    hipStream_t st = (hipStream_t)stream;
    int err = 0;
    ISAI_HIP(hipMemcpyAsync(&err, &P->d_cnt->err, sizeof(int), hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipStreamSynchronize(st));
    if (err) {
        const int zero = 0;
        ISAI_HIP(hipMemcpy(&P->d_cnt->err, &zero, sizeof(int), hipMemcpyHostToDevice));
        set_error("sptrsv: a dependency wait exceeded its spin limit (malformed pattern or missing diagonal)");
        return ISAI_ERR_TIMEOUT;
    }
    return ISAI_OK;
}

int isai_bsr_sptrsv_destroy(isai_sptrsv_plan_t P) {
    if (!P) return ISAI_OK;
    (void)hipFree(P->d_diag);
    (void)hipFree(P->d_dinvL);
    (void)hipFree(P->d_dinvU);
    (void)hipFree(P->d_y);
    (void)hipFree(P->d_chunksL);
    (void)hipFree(P->d_chunksU);
    (void)hipFree(P->d_cnt);
    delete P;
    return ISAI_OK;
}

int isai_csr_sptrsv_lu_device(int m, int nnz, const int *d_rp, const int *d_ci, const double *d_v,
                              const double *d_b, double *d_x, void *d_work, int chunk, int spin_limit,
                              isai_stream_t stream) {
    ISAI_REQUIRE(m > 0 && nnz >= 0 && d_rp && d_ci && d_v && d_b && d_x && d_work,
                 "isai_csr_sptrsv_lu_device: bad argument");
    if (chunk <= 0) chunk = 256;
    ISAI_REQUIRE(chunk <= kCsrMaxChunk, "isai_csr_sptrsv_lu_device: chunk above %d", kCsrMaxChunk);
    hipStream_t st = (hipStream_t)stream;
    const int nchunks = (m + chunk - 1) / chunk;
    // d_work: m doubles for the intermediate vector, then a Counters record.
    double *y = (double *)d_work;
    Counters *cnt = (Counters *)((double *)d_work + ((size_t)m + 7) / 8 * 8);
    Counters h{};
    h.spin_limit = spin_limit > 0 ? spin_limit : (1 << 20);
    ISAI_HIP(hipMemcpyAsync(cnt, &h, sizeof(h), hipMemcpyHostToDevice, st));
    fill_sentinel_kernel<<<fill_blocks(m), 256, 0, st>>>(y, (size_t)m, nullptr);
    csr_sptrsv_kernel<false><<<nchunks, 64, 0, st>>>(m, chunk, nchunks, d_rp, d_ci, d_v, d_b, y, d_x, cnt);
    csr_sptrsv_kernel<true><<<nchunks, 64, 0, st>>>(m, chunk, nchunks, d_rp, d_ci, d_v, y, d_x, nullptr, cnt);
    ISAI_LAUNCH_CHECK();
    int err = 0;
    ISAI_HIP(hipMemcpyAsync(&err, &cnt->err, sizeof(int), hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipStreamSynchronize(st));
    if (err) {
        set_error("csr sptrsv: a dependency wait exceeded its spin limit (missing diagonal or malformed pattern)");
        return ISAI_ERR_TIMEOUT;
    }
    return ISAI_OK;
}

int isai_csr_sptrsv_work_size(int m, size_t *bytes) {
    ISAI_REQUIRE(m >= 0 && bytes, "isai_csr_sptrsv_work_size: bad argument");
    *bytes = sizeof(double) * (((size_t)m + 7) / 8 * 8) + sizeof(Counters);
    return ISAI_OK;
}

int isai_csr_sptrsv_lu(int m, const int *rowptr, const int *colind, const double *val, int nnz, double *b) {
    ISAI_REQUIRE(m > 0 && nnz >= 0 && rowptr && colind && val && b, "isai_csr_sptrsv_lu: bad argument");
    int *d_rp = nullptr, *d_ci = nullptr;
    double *d_v = nullptr, *d_b = nullptr, *d_x = nullptr, *d_w = nullptr;
    size_t work = 0;
    int s = isai_csr_sptrsv_work_size(m, &work);
    if (s != ISAI_OK) return s;
    if (hipMalloc(&d_rp, sizeof(int) * (m + 1)) != hipSuccess || hipMalloc(&d_ci, sizeof(int) * (nnz ? nnz : 1)) != hipSuccess ||
        hipMalloc(&d_v, sizeof(double) * (nnz ? nnz : 1)) != hipSuccess || hipMalloc(&d_b, sizeof(double) * m) != hipSuccess ||
        hipMalloc(&d_x, sizeof(double) * m) != hipSuccess ||
        hipMalloc(&d_w, work) != hipSuccess) {
        set_error("isai_csr_sptrsv_lu: hipMalloc failed");
        s = ISAI_ERR_ALLOC;
    }
    if (s == ISAI_OK) {
        hipError_t e = hipMemcpy(d_rp, rowptr, sizeof(int) * (m + 1), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_ci, colind, sizeof(int) * nnz, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_v, val, sizeof(double) * nnz, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_b, b, sizeof(double) * m, hipMemcpyHostToDevice);
        if (e != hipSuccess) s = hip_fail(e, "hipMemcpy", __FILE__, __LINE__);
    }
    if (s == ISAI_OK) s = isai_csr_sptrsv_lu_device(m, nnz, d_rp, d_ci, d_v, d_b, d_x, d_w, 0, 0, nullptr);
    if (s == ISAI_OK) {
        hipError_t e = hipMemcpy(b, d_x, sizeof(double) * m, hipMemcpyDeviceToHost);
        if (e != hipSuccess) s = hip_fail(e, "hipMemcpy", __FILE__, __LINE__);
    }
    (void)hipFree(d_rp);
    (void)hipFree(d_ci);
    (void)hipFree(d_v);
    (void)hipFree(d_b);
    (void)hipFree(d_x);
    (void)hipFree(d_w);
    return s;
}

}  // extern "C"
