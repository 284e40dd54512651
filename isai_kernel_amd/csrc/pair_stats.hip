// pair_stats.hip -- pair-list statistics of an ISAI factor on the device (the
// figures the reference prints from analyze_pair.h:6-327) and the column
// statistics the ISAI build plan uses.
//
// Per column j (one thread): N_j = Mbrow[j+1] - Mbrow[j], Q_j = PairPtr[Mbrow[j+1]] -
// PairPtr[Mbrow[j]] and the largest per-entry pair count qlen of the column.  The
// histogram, column bins, maxima and totals are aggregated per workgroup in LDS
// and added to one global record.  Percentiles are exact: counting histograms of N,
// Q and qlen (LDS-privatised for values below 1024) are walked on the host, which
// gives the same element as sorting and indexing (size - 1) * p.
#include <algorithm>
#include <cstddef>
#include <vector>

#include "common.hpp"
#include "internal.hpp"
#include "isai_bsr.h"

namespace isai {

struct StatsAcc {
    unsigned long long total_q;
    unsigned long long qlen_hist[10];
    unsigned long long bin_sum_q[10];
    int bin_cols[10];
    int bin_max_q[10];
    int max_n, max_q, max_qlen, pad;
};

__device__ __forceinline__ int qlen_bin(int q) {
    return q == 0 ? 0 : q == 1 ? 1 : q < 4 ? 2 : q < 8 ? 3 : q < 16 ? 4 : q < 32 ? 5 : q < 64 ? 6 : q < 128 ? 7
           : q < 256 ? 8 : 9;
}

constexpr int kStatsBlock = 256;

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off));
    return v;
}

// The per-column figures are accumulated per thread over a grid-stride loop,
// summed (or maxed) over the wave with shuffles, added to LDS by one lane per wave
// and to the global record by one thread per quantity and workgroup.  Same-address
// global atomics serialise, so the grid with a record is capped at kStatsGrid
// workgroups: 3,456 workgroups of 44 atomics each took 156 us on the 96^3 7-pt
// factor's 884,736 columns.
constexpr int kStatsGrid = 256;
__global__ __launch_bounds__(kStatsBlock) void pair_col_kernel(int nb, const int *__restrict__ Mbrow,
                                                               const int *__restrict__ PairPtr, int *__restrict__ N,
                                                               int *__restrict__ Q, int *__restrict__ MQ,
                                                               StatsAcc *acc) {
    __shared__ unsigned long long s_hist[10], s_bsum[10], s_tq;
    __shared__ int s_bcols[10], s_bmax[10], s_max[3];
    if (acc) {
        if (threadIdx.x < 10) {
            s_hist[threadIdx.x] = s_bsum[threadIdx.x] = 0ull;
            s_bcols[threadIdx.x] = s_bmax[threadIdx.x] = 0;
        }
        if (threadIdx.x < 3) s_max[threadIdx.x] = 0;
        if (threadIdx.x == 0) s_tq = 0ull;
        __syncthreads();
    }
    unsigned hc[10], bc[10];
    unsigned long long bs[10], tq = 0ull;
    int bm[10], m0 = 0, m1 = 0, m2 = 0;
#pragma unroll
    for (int k = 0; k < 10; ++k) hc[k] = bc[k] = 0, bs[k] = 0ull, bm[k] = 0;
    for (int j = blockIdx.x * kStatsBlock + threadIdx.x; j < nb; j += gridDim.x * kStatsBlock) {
        const int st = Mbrow[j], en = Mbrow[j + 1];
        const int n = en - st, q = PairPtr[en] - PairPtr[st];
        int mq = 0;
        for (int g = st; g < en; ++g) {
            const int ql = PairPtr[g + 1] - PairPtr[g];
            mq = max(mq, ql);
            if (acc) {
                const int bin = qlen_bin(ql);
#pragma unroll
                for (int k = 0; k < 10; ++k) hc[k] += bin == k;
            }
        }
        if (N) N[j] = n;
        if (Q) Q[j] = q;
        if (MQ) MQ[j] = mq;
        if (acc) {
            const int b = min((int)((long long)j * 10 / nb), 9);
#pragma unroll
            for (int k = 0; k < 10; ++k) {
                bc[k] += b == k;
                bs[k] += b == k ? (unsigned long long)q : 0ull;
                bm[k] = max(bm[k], b == k ? q : 0);
            }
            tq += (unsigned long long)q;
            m0 = max(m0, n), m1 = max(m1, q), m2 = max(m2, mq);
        }
    }
    if (!acc) return;
    const bool lead = (threadIdx.x & 63) == 0;
#pragma unroll
    for (int k = 0; k < 10; ++k) {
        const unsigned long long hk = wave_sum((unsigned long long)hc[k]);
        const unsigned ck = wave_sum(bc[k]);
        const unsigned long long sk = wave_sum(bs[k]);
        const int mk = wave_max(bm[k]);
        if (lead) {
            if (hk) atomicAdd(&s_hist[k], hk);
            if (ck) {
                atomicAdd(&s_bsum[k], sk);
                atomicAdd(&s_bcols[k], (int)ck);
                atomicMax(&s_bmax[k], mk);
            }
        }
    }
    tq = wave_sum(tq);
    m0 = wave_max(m0), m1 = wave_max(m1), m2 = wave_max(m2);
    if (lead) {
        atomicAdd(&s_tq, tq);
        atomicMax(&s_max[0], m0);
        atomicMax(&s_max[1], m1);
        atomicMax(&s_max[2], m2);
    }
    __syncthreads();
    if (threadIdx.x < 10) {
        const int k = threadIdx.x;
        if (s_hist[k]) atomicAdd(&acc->qlen_hist[k], s_hist[k]);
        if (s_bcols[k]) {
            atomicAdd(&acc->bin_sum_q[k], s_bsum[k]);
            atomicAdd(&acc->bin_cols[k], s_bcols[k]);
            atomicMax(&acc->bin_max_q[k], s_bmax[k]);
        }
    }
    if (threadIdx.x == 0) {
        atomicAdd(&acc->total_q, s_tq);
        atomicMax(&acc->max_n, s_max[0]);
        atomicMax(&acc->max_q, s_max[1]);
        atomicMax(&acc->max_qlen, s_max[2]);
    }
}

// cnt[v] += number of values v among vals[0..n) (or, with pp, among the
// differences pp[i+1] - pp[i]); cnt holds cap = max value + 1 entries.
constexpr int kCountLds = 1024;

__global__ __launch_bounds__(kStatsBlock) void count_values_kernel(int n, const int *__restrict__ vals,
                                                                   const int *__restrict__ pp, int *cnt, int cap) {
    __shared__ int s_cnt[kCountLds];
    for (int k = threadIdx.x; k < kCountLds; k += kStatsBlock) s_cnt[k] = 0;
    __syncthreads();
    for (int i = blockIdx.x * kStatsBlock + threadIdx.x; i < n; i += gridDim.x * kStatsBlock) {
        const int v = pp ? pp[i + 1] - pp[i] : vals[i];
        if ((unsigned)v >= (unsigned)cap) continue;  // cannot happen for values whose maximum sized cnt
        if (v < kCountLds) atomicAdd(&s_cnt[v], 1);
        else atomicAdd(&cnt[v], 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kCountLds && k < cap; k += kStatsBlock)
        if (s_cnt[k]) atomicAdd(&cnt[k], s_cnt[k]);
}

// Element (size_t)((n - 1) * p) of the sorted values, from their counts.
static int percentile(const std::vector<int> &cnt, long long n, double p) {
    if (n <= 0) return 0;
    const long long k = (long long)(size_t)((double)(n - 1) * p);
    long long cum = 0;
    for (size_t v = 0; v < cnt.size(); ++v) {
        cum += cnt[v];
        if (cum > k) return (int)v;
    }
    return (int)cnt.size() - 1;
}

size_t pair_acc_bytes() { return sizeof(StatsAcc); }
size_t pair_acc_max_q_offset() { return offsetof(StatsAcc, max_q); }

int pair_column_q_launch(int nb, const int *d_Mbrow, const int *d_PairPtr, int *d_Q, void *d_acc, hipStream_t st) {
    if (nb == 0) return ISAI_OK;
    const unsigned g = std::min<unsigned>(grid_for(nb, kStatsBlock), kStatsGrid);
    pair_col_kernel<<<g, kStatsBlock, 0, st>>>(nb, d_Mbrow, d_PairPtr, nullptr, d_Q, nullptr, (StatsAcc *)d_acc);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

}  // namespace isai

using namespace isai;

extern "C" int isai_pair_column_stats(int nb, const int *d_Mbrow, const int *d_PairPtr, int *d_N, int *d_Q,
                                      int *d_maxqlen, isai_stream_t stream) {
    ISAI_REQUIRE(nb >= 0, "isai_pair_column_stats: nb = %d", nb);
    if (nb == 0) return ISAI_OK;
    ISAI_REQUIRE(d_Mbrow && d_PairPtr, "isai_pair_column_stats: null pattern pointer");
    hipStream_t st = (hipStream_t)stream;
    pair_col_kernel<<<grid_for(nb, kStatsBlock), kStatsBlock, 0, st>>>(nb, d_Mbrow, d_PairPtr, d_N, d_Q, d_maxqlen,
                                                                      nullptr);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

extern "C" int isai_pair_stats(int nb, const int *d_Mbrow, const int *d_PairPtr, isai_pair_stats_t *out,
                               isai_stream_t stream) {
    ISAI_REQUIRE(nb >= 1, "isai_pair_stats: nb = %d (need at least one column)", nb);
    ISAI_REQUIRE(d_Mbrow && d_PairPtr && out, "isai_pair_stats: null pointer");
    hipStream_t st = (hipStream_t)stream;
    int mb_ends[2] = {0, 0};
    ISAI_HIP(hipMemcpyAsync(&mb_ends[0], d_Mbrow, sizeof(int), hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipMemcpyAsync(&mb_ends[1], d_Mbrow + nb, sizeof(int), hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipStreamSynchronize(st));
    ISAI_REQUIRE(mb_ends[0] == 0 && mb_ends[1] >= 0, "isai_pair_stats: Mbrow[0] = %d, Mbrow[nb] = %d", mb_ends[0],
                 mb_ends[1]);
    const int mnnz = mb_ends[1];

    int *d_N = nullptr, *d_Q = nullptr, *d_cnt = nullptr;
    StatsAcc *d_acc = nullptr;
    auto release = [&]() {
        (void)hipFree(d_N);
        (void)hipFree(d_Q);
        (void)hipFree(d_cnt);
        (void)hipFree(d_acc);
    };
    if (hipMalloc(&d_N, sizeof(int) * nb) != hipSuccess || hipMalloc(&d_Q, sizeof(int) * nb) != hipSuccess ||
        hipMalloc(&d_acc, sizeof(StatsAcc)) != hipSuccess) {
        release();
        set_error("isai_pair_stats: allocation failed");
        return ISAI_ERR_ALLOC;
    }
    StatsAcc acc{};
    int s = ISAI_OK;
    auto hip = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && s == ISAI_OK) s = hip_fail(e, what, __FILE__, __LINE__);
    };
    hip(hipMemsetAsync(d_acc, 0, sizeof(StatsAcc), st), "hipMemsetAsync");
    if (s == ISAI_OK) {
        pair_col_kernel<<<std::min<unsigned>(grid_for(nb, kStatsBlock), kStatsGrid), kStatsBlock, 0, st>>>(
            nb, d_Mbrow, d_PairPtr, d_N, d_Q, nullptr, d_acc);
        hip(hipGetLastError(), "pair_col_kernel");
    }
    if (s == ISAI_OK) hip(hipMemcpyAsync(&acc, d_acc, sizeof(StatsAcc), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    if (s == ISAI_OK) hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    // exact percentiles from counting histograms of N, Q and qlen
    const size_t capN = (size_t)acc.max_n + 1, capQ = (size_t)acc.max_q + 1, capL = (size_t)acc.max_qlen + 1;
    std::vector<int> cN(capN), cQ(capQ), cL(capL);
    if (s == ISAI_OK && hipMalloc(&d_cnt, sizeof(int) * (capN + capQ + capL)) != hipSuccess) {
        set_error("isai_pair_stats: allocation failed");
        s = ISAI_ERR_ALLOC;
    }
    if (s == ISAI_OK) hip(hipMemsetAsync(d_cnt, 0, sizeof(int) * (capN + capQ + capL), st), "hipMemsetAsync");
    if (s == ISAI_OK) {
        const unsigned gc = std::min<unsigned>(grid_for(nb, kStatsBlock), 1024u);
        count_values_kernel<<<gc, kStatsBlock, 0, st>>>(nb, d_N, nullptr, d_cnt, (int)capN);
        count_values_kernel<<<gc, kStatsBlock, 0, st>>>(nb, d_Q, nullptr, d_cnt + capN, (int)capQ);
        if (mnnz > 0) {
            const unsigned gl = std::min<unsigned>(grid_for(mnnz, kStatsBlock), 1024u);
            count_values_kernel<<<gl, kStatsBlock, 0, st>>>(mnnz, nullptr, d_PairPtr, d_cnt + capN + capQ,
                                                            (int)capL);
        }
        hip(hipGetLastError(), "count_values_kernel");
    }
    if (s == ISAI_OK) {
        hip(hipMemcpyAsync(cN.data(), d_cnt, sizeof(int) * capN, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
        hip(hipMemcpyAsync(cQ.data(), d_cnt + capN, sizeof(int) * capQ, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
        hip(hipMemcpyAsync(cL.data(), d_cnt + capN + capQ, sizeof(int) * capL, hipMemcpyDeviceToHost, st),
            "hipMemcpyAsync");
        hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    }
    release();
    if (s != ISAI_OK) return s;

    isai_pair_stats_t r{};
    r.nb = nb;
    r.mnnz = mnnz;
    r.total_pairs = (long long)acc.total_q;
    r.n_avg = (double)mnnz / nb;
    r.q_avg = (double)r.total_pairs / nb;
    r.qlen_avg = mnnz ? (double)r.total_pairs / mnnz : 0.0;
    r.n_max = acc.max_n;
    r.q_max = acc.max_q;
    r.qlen_max = acc.max_qlen;
    static const double P[4] = {0.50, 0.90, 0.95, 0.99};
    int *pn[4] = {&r.n_p50, &r.n_p90, &r.n_p95, &r.n_p99};
    int *pq[4] = {&r.q_p50, &r.q_p90, &r.q_p95, &r.q_p99};
    int *pl[4] = {&r.qlen_p50, &r.qlen_p90, &r.qlen_p95, &r.qlen_p99};
    for (int k = 0; k < 4; ++k) {
        *pn[k] = percentile(cN, nb, P[k]);
        *pq[k] = percentile(cQ, nb, P[k]);
        *pl[k] = percentile(cL, mnnz, P[k]);
    }
    for (int b = 0; b < 10; ++b) {
        r.qlen_hist[b] = (long long)acc.qlen_hist[b];
        r.bin_sum_q[b] = (long long)acc.bin_sum_q[b];
        r.bin_cols[b] = acc.bin_cols[b];
        r.bin_max_q[b] = acc.bin_max_q[b];
    }
    *out = r;
    return ISAI_OK;
}
