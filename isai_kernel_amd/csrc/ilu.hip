// ilu.hip -- block ILU(0) in place on gfx950 (replaces cusparseDbsrilu02 as called
// by bsrilu02_inplace, reference bsrilu.h:20-66).
//
// Result layout = cuSPARSE's DIRECTION_ROW packed factor: strictly-lower blocks
// hold L_ik, diagonal blocks hold L_ii\U_ii (unit-lower \ upper), strictly-upper
// blocks hold U_ij, with A = L U on every stored block.
//
// Schedule: the analysis (host) computes the level of every block row and the
// rows in (level, row) order.  The factor kernel gives each block row one wave64;
// a workgroup takes the next W rows of that order from an atomic ticket, so a row
// only ever waits on rows whose waves already hold a ticket (no global barrier, no
// host round trip between levels).  Finished rows are published into a buffer
// pre-filled with a NaN sentinel, with relaxed agent-scope atomic stores, and
// readers poll each value they need with relaxed agent-scope atomic loads until
// it is no longer the sentinel: every value is its own ready flag.  Agent-scope
// release/acquire fences are avoided on purpose -- on gfx950 they write back or
// invalidate a whole XCD L2 per use, which serialises one such operation per row.
//
// Per row i:
//   for each stored k < i, ascending:   L_ik = A'_ik inv(U_kk)  (right substitution)
//       for each stored j > k with (k,j) stored:  A'_ij -= L_ik U_kj
//   A'_ii = L_ii U_ii (scalar LU, no pivoting, FP64)
//   for each stored j > i:   U_ij = inv(L_ii) A'_ij            (forward substitution)
// Every scalar operation is the oracle's, in the same order, without contraction,
// so the factor is bit-identical to the sequential restatement.
#include <algorithm>
#include <climits>
#include <vector>

#include "bsrilu.h"
#include "common.hpp"
#include "internal.hpp"
#include "levels.hpp"

struct isai_ilu_plan {
    int nb = 0, nnzb = 0, bs = 0, nlevels = 0;
    const int *d_row = nullptr, *d_col = nullptr;
    int *d_order = nullptr, *d_diag = nullptr;
    int *d_upd_ptr = nullptr, *d_upd_q = nullptr, *d_upd_src = nullptr;  // update lists
    int max_row = 0;    // blocks of the longest row
    int max_stage = 0;  // largest staged set of one row (diagonal blocks read + update sources), <= kIluStage
};

namespace isai {

constexpr int kIluMaxRow = 64;       // blocks of one row staged in LDS
constexpr int kIluSpinLimit = 1 << 20;
constexpr unsigned long long kIluSentinel = 0x7FF4DEADBEEF0002ull;  // signalling-NaN payload

// Value of another row's factor: relaxed agent-scope loads until it is published.
__device__ __forceinline__ double ilu_ready(const double *p, bool &timed_out) {
    const unsigned long long *q = (const unsigned long long *)p;
    unsigned long long v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int spins = 0;
    while (v == kIluSentinel) {
        if (++spins > kIluSpinLimit) {
            timed_out = true;
            break;
        }
        __builtin_amdgcn_s_sleep(1);
        v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return __longlong_as_double((long long)v);
}

__global__ void ilu_sentinel_kernel(double *p, size_t n) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        p[k] = __longlong_as_double((long long)kIluSentinel);
}
constexpr int kIluStage = 32;     // blocks of earlier rows staged in LDS per row

template <int BS>
struct IluShape {
    static constexpr int B2 = BS * BS;
    static constexpr int G = B2 <= 16 ? 16 : (B2 <= 32 ? 32 : 64);
    static constexpr int C = 64 / G;
};

__device__ __forceinline__ int ilu_find(const int *__restrict__ rp, const int *__restrict__ ci, int br, int bc) {
    int s = rp[br], e = rp[br + 1] - 1;
    while (s <= e) {
        const int m = (s + e) >> 1, c = ci[m];
        if (c == bc) return m;
        if (c < bc) s = m + 1; else e = m - 1;
    }
    return -1;
}

// Per-wave LDS, sized by the analysis: the row being factored (the longest row of
// the matrix) and a staging area for the blocks of earlier rows it reads (the
// largest such set, at most kIluStage blocks).  A workgroup takes one ticket for
// all its waves (rows), and tickets are same-address atomics, which serialise:
// when the levels are wide (kIluWideLevel rows on average) 16 rows share a ticket
// (96^3 7-pt: 2.88 ms against 3.45 ms with 4), else 4 (2-D 256^2 9-pt, 85 rows per
// level: 5.05 ms with 4 against 5.86 ms with 16), as long as their LDS fits in 64 KB.
// Narrow level structures also keep the fixed per-wave layout (kIluMaxRow +
// kIluStage blocks): its LDS holds a CU to fewer resident rows, and with narrow
// levels more resident rows only add waiting ones (5.56 ms with the plan-sized
// layout on the 2-D case).
constexpr int kIluMaxWaves = 16, kIluWideLevel = 512;
struct IluLaunch {
    int waves, row_blocks, stage_blocks;
};
static IluLaunch ilu_launch_shape(const isai_ilu_plan *P) {
    const bool wide = (long long)P->nb >= (long long)kIluWideLevel * std::max(P->nlevels, 1);
    IluLaunch L{wide ? kIluMaxWaves : 4, wide ? P->max_row : kIluMaxRow, wide ? P->max_stage : kIluStage};
    const size_t wave = (size_t)(L.row_blocks + L.stage_blocks) * P->bs * P->bs * sizeof(double);
    while (L.waves > 1 && wave * L.waves > 65536) L.waves >>= 1;
    return L;
}

// Orders the LDS accesses of one wave (its lanes exchange data through LDS).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// One ticket per workgroup covers W consecutive positions of the level order; wave
// w factors position base + w.  Every row waits only on smaller positions, held by
// earlier tickets, so the schedule cannot deadlock.
//
// The analysis lists, for every stored lower block (i, k) at position p, the
// updates A'_ij -= L_ik U_kj it causes: upd_q = local index of (i, j) in row i,
// upd_src = position of (k, j).  Once the rows the row depends on are ready, the
// wave gathers every block it will read -- the diagonal blocks of its k's, then
// the U_kj of its updates in list order -- into LDS in one parallel pass, so the
// sequential part of the factorization runs from LDS.  Rows whose list exceeds
// the staging area read those blocks from global memory instead.
template <int BS>
__global__ __launch_bounds__(64 * kIluMaxWaves) void ilu0_kernel(
    int nb, const int *__restrict__ rp, const int *__restrict__ ci, const double *__restrict__ val,
    double *fac, const int *__restrict__ order, const int *__restrict__ diagpos, const int *__restrict__ upd_ptr,
    const int *__restrict__ upd_q, const int *__restrict__ upd_src, unsigned int *__restrict__ ticket,
    int *__restrict__ zero_pivot, int *__restrict__ err, int max_row, int max_stage) {
    using S = IluShape<BS>;
    constexpr int B2 = S::B2;
    constexpr int CB = 64 / BS;  // blocks per pass when one lane owns one row or column
    extern __shared__ double ilu_lds[];
    __shared__ int s_base;
    const int W = blockDim.x >> 6;
    if (threadIdx.x == 0) s_base = (int)atomicAdd(ticket, (unsigned)W);
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int pos = s_base + w;
    if (pos >= nb) return;
    double *rowv = ilu_lds + (size_t)w * (max_row + max_stage) * B2;
    double *stage = rowv + (size_t)max_row * B2;
    const int i = order[pos];
    const int rs = rp[i], re = rp[i + 1], len = re - rs;
    const int dp = diagpos[i];
    const int nlow = dp - rs;
    const int u0 = upd_ptr[rs], nupd = upd_ptr[dp] - u0;
    const int nstage = nlow + nupd;
    const bool staged = nstage <= max_stage;
    const int g = lane / S::G, e = lane % S::G;
    const bool ok = e < B2;
    const int r = ok ? e / BS : 0, c = ok ? e % BS : 0;

    for (int x = lane; x < len * B2; x += 64) rowv[x] = val[(size_t)rs * B2 + x];

    // Gather the published blocks this row reads (waiting on each value).
    bool timed_out = false;
    if (staged) {
        for (int x = lane; x < nstage * B2; x += 64) {
            const int slot = x / B2, el = x - slot * B2;
            const int src = slot < nlow ? diagpos[ci[rs + slot]] : upd_src[u0 + slot - nlow];
            stage[x] = ilu_ready(fac + (size_t)src * B2 + el, timed_out);
        }
    }
    wave_lds_sync();

    for (int p = rs; p < dp; ++p) {
        const int k = ci[p];
        const int lp = p - rs;
        // L_ik = A'_ik inv(U_kk) by right substitution with the upper part of the
        // packed diagonal block of row k; lane r owns row r.
        if (lane < BS) {
            double Dk[B2];
            if (staged) {
#pragma unroll
                for (int t = 0; t < B2; ++t) Dk[t] = stage[lp * B2 + t];
            } else {
#pragma unroll
                for (int t = 0; t < B2; ++t) Dk[t] = ilu_ready(fac + (size_t)diagpos[k] * B2 + t, timed_out);
            }
            double *X = rowv + lp * B2 + lane * BS;
#pragma unroll
            for (int cc = 0; cc < BS; ++cc) {
                double sacc = X[cc];
#pragma unroll
                for (int m = 0; m < cc; ++m) sacc = sacc - X[m] * Dk[m * BS + cc];
                X[cc] = sacc / Dk[cc * BS + cc];
            }
        }
        wave_lds_sync();
        // A'_ij -= L_ik U_kj for the updates of (i, k); one lane per element (r, c),
        // G lanes per block.
        const int ue = upd_ptr[p + 1];
        for (int ub = upd_ptr[p]; ub < ue; ub += S::C) {
            const int u = ub + g;
            if (u < ue && ok) {
                const double *Lr = rowv + lp * B2 + r * BS;
                double sacc = 0.0;
                if (staged) {
                    const double *Ukj = stage + (nlow + u - u0) * B2;
#pragma unroll
                    for (int m = 0; m < BS; ++m) sacc = sacc + Lr[m] * Ukj[m * BS + c];
                } else {
                    const double *Ukj = fac + (size_t)upd_src[u] * B2;
#pragma unroll
                    for (int m = 0; m < BS; ++m) sacc = sacc + Lr[m] * ilu_ready(Ukj + m * BS + c, timed_out);
                }
                const int q = upd_q[u];
                rowv[q * B2 + e] = rowv[q * B2 + e] - sacc;
            }
        }
        wave_lds_sync();
    }

    // Diagonal block: scalar LU without pivoting, packed L_ii\U_ii.
    if (lane == 0) {
        double *D = rowv + nlow * B2;
        bool zero = false;
        for (int cc = 0; cc < BS; ++cc) {
            if (D[cc * BS + cc] == 0.0) zero = true;
            for (int rr = cc + 1; rr < BS; ++rr) {
                D[rr * BS + cc] = D[rr * BS + cc] / D[cc * BS + cc];
                for (int m = cc + 1; m < BS; ++m) D[rr * BS + m] = D[rr * BS + m] - D[rr * BS + cc] * D[cc * BS + m];
            }
        }
        if (zero) atomicMin(zero_pivot, i);
    }
    wave_lds_sync();

    // U_ij = inv(L_ii) A'_ij for the stored j > i by forward substitution with the
    // unit-lower part of the diagonal block; lane (block, column c).
    {
        const double *D = rowv + nlow * B2;
        const int gb = lane / BS, col = lane % BS;
        for (int q0 = dp + 1; q0 < re; q0 += CB) {
            const int q = q0 + gb;
            if (gb < CB && q < re) {
                double *X = rowv + (q - rs) * B2 + col;
                for (int rr = 1; rr < BS; ++rr) {
                    double xv = X[rr * BS];
                    for (int m = 0; m < rr; ++m) xv = xv - D[rr * BS + m] * X[m * BS];
                    X[rr * BS] = xv;
                }
            }
        }
    }
    wave_lds_sync();
    // publish: every value is its own ready flag
    for (int x = lane; x < len * B2; x += 64)
        __hip_atomic_store((unsigned long long *)(fac + (size_t)rs * B2 + x), (unsigned long long)__double_as_longlong(rowv[x]),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (timed_out) atomicExch(err, 1);
}

template <int BS>
void launch_ilu(const isai_ilu_plan *P, const double *val, double *fac, unsigned *ticket, int *zp, int *err,
                hipStream_t st) {
    const IluLaunch L = ilu_launch_shape(P);
    const int W = L.waves;
    const size_t lds = (size_t)W * (L.row_blocks + L.stage_blocks) * BS * BS * sizeof(double);
    ilu0_kernel<BS><<<(P->nb + W - 1) / W, 64 * W, lds, st>>>(P->nb, P->d_row, P->d_col, val, fac, P->d_order,
                                                             P->d_diag, P->d_upd_ptr, P->d_upd_q, P->d_upd_src,
                                                             ticket, zp, err, L.row_blocks, L.stage_blocks);
}

}  // namespace isai

using namespace isai;

extern "C" {

int isai_bsrilu0_analysis(int nb, int nnzb, int bs, const int *d_row, const int *d_col, isai_ilu_plan_t *plan,
                          int *structural_zero, isai_stream_t stream) {
    ISAI_REQUIRE(plan && structural_zero, "isai_bsrilu0_analysis: null output");
    ISAI_REQUIRE(nb >= 0 && nnzb >= 0 && bs >= 1 && bs <= 8, "isai_bsrilu0_analysis: bad size");
    *plan = nullptr;
    *structural_zero = -1;
    hipStream_t st = (hipStream_t)stream;
    std::vector<int> rp(nb + 1), ci(nnzb), level(nb), order(nb), diag(nb);
    ISAI_HIP(hipMemcpyAsync(rp.data(), d_row, sizeof(int) * (nb + 1), hipMemcpyDeviceToHost, st));
    if (nnzb) ISAI_HIP(hipMemcpyAsync(ci.data(), d_col, sizeof(int) * nnzb, hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipStreamSynchronize(st));
    ISAI_REQUIRE(rp[0] == 0 && rp[nb] == nnzb, "isai_bsrilu0_analysis: row pointer does not span nnzb");
    int max_row = 0;
    for (int i = 0; i < nb; ++i) {
        diag[i] = -1;
        const int len = rp[i + 1] - rp[i];
        if (len > max_row) max_row = len;
        for (int p = rp[i]; p < rp[i + 1]; ++p) {
            if (p > rp[i] && ci[p] <= ci[p - 1]) {
                set_error("isai_bsrilu0_analysis: block columns of row %d are not strictly ascending", i);
                return ISAI_ERR_INVALID_ARG;
            }
            if (ci[p] == i) diag[i] = p;
        }
        if (diag[i] < 0 && *structural_zero < 0) *structural_zero = i;
    }
    if (*structural_zero >= 0) {
        set_error("bsrilu0: structural zero at block (%d,%d)", *structural_zero, *structural_zero);
        return ISAI_ERR_ZERO_PIVOT;
    }
    if (max_row > kIluMaxRow) {
        set_error("bsrilu0: a block row holds %d blocks (limit %d)", max_row, kIluMaxRow);
        return ISAI_ERR_UNSUPPORTED;
    }
    // update lists: for each stored lower block (i, k) at position p, the (i, j)
    // with j > k such that (k, j) is stored, in ascending j
    std::vector<int> upd_ptr(nnzb + 1, 0), upd_q, upd_src;
    for (int i = 0; i < nb; ++i) {
        for (int p = rp[i]; p < rp[i + 1]; ++p) {
            upd_ptr[p] = (int)upd_q.size();
            const int k = ci[p];
            if (k >= i) continue;
            int t = rp[k];
            for (int q = p + 1; q < rp[i + 1]; ++q) {
                while (t < rp[k + 1] && ci[t] < ci[q]) ++t;
                if (t < rp[k + 1] && ci[t] == ci[q]) {
                    upd_q.push_back(q - rp[i]);
                    upd_src.push_back(t);
                }
            }
        }
    }
    upd_ptr[nnzb] = (int)upd_q.size();
    // the largest staged set of a row that fits the staging cap (longer ones read global memory)
    int max_stage = 0;
    for (int i = 0; i < nb; ++i) {
        const int d = diag[i];
        const int n = (d - rp[i]) + (upd_ptr[d] - upd_ptr[rp[i]]);
        if (n <= kIluStage) max_stage = std::max(max_stage, n);
    }
    if (upd_q.size() > (size_t)INT_MAX / 2) {
        set_error("bsrilu0 analysis: %zu updates exceed the int32 index range", upd_q.size());
        return ISAI_ERR_CAPACITY;
    }
    int nl = 0;
    compute_levels(nb, rp.data(), ci.data(), true, level.data(), &nl);
    level_order(nb, level.data(), nl, order.data(), nullptr);
    auto *P = new isai_ilu_plan;
    P->nb = nb;
    P->nnzb = nnzb;
    P->bs = bs;
    P->nlevels = nl;
    P->d_row = d_row;
    P->d_col = d_col;
    P->max_row = max_row;
    P->max_stage = max_stage;
    const size_t nu = upd_q.size();
    if (hipMalloc(&P->d_order, sizeof(int) * (nb ? nb : 1)) != hipSuccess ||
        hipMalloc(&P->d_diag, sizeof(int) * (nb ? nb : 1)) != hipSuccess ||
        hipMalloc(&P->d_upd_ptr, sizeof(int) * (nnzb + 1)) != hipSuccess ||
        hipMalloc(&P->d_upd_q, sizeof(int) * (nu ? nu : 1)) != hipSuccess ||
        hipMalloc(&P->d_upd_src, sizeof(int) * (nu ? nu : 1)) != hipSuccess) {
        isai_bsrilu0_destroy(P);
        set_error("bsrilu0 analysis: hipMalloc failed");
        return ISAI_ERR_ALLOC;
    }
    // uploads; the plan is released if one fails
    auto upload = [&]() -> int {
        ISAI_HIP(hipMemcpyAsync(P->d_order, order.data(), sizeof(int) * nb, hipMemcpyHostToDevice, st));
        ISAI_HIP(hipMemcpyAsync(P->d_diag, diag.data(), sizeof(int) * nb, hipMemcpyHostToDevice, st));
        ISAI_HIP(hipMemcpyAsync(P->d_upd_ptr, upd_ptr.data(), sizeof(int) * (nnzb + 1), hipMemcpyHostToDevice, st));
        if (nu) {
            ISAI_HIP(hipMemcpyAsync(P->d_upd_q, upd_q.data(), sizeof(int) * nu, hipMemcpyHostToDevice, st));
            ISAI_HIP(hipMemcpyAsync(P->d_upd_src, upd_src.data(), sizeof(int) * nu, hipMemcpyHostToDevice, st));
        }
        ISAI_HIP(hipStreamSynchronize(st));
        return ISAI_OK;
    };
    if (const int us = upload(); us != ISAI_OK) {
        (void)hipStreamSynchronize(st);  // no copy may still read the host arrays
        isai_bsrilu0_destroy(P);
        return us;
    }
    *plan = P;
    return ISAI_OK;
}

int isai_bsrilu0_plan_info(isai_ilu_plan_t plan, int *nlevels, int *max_row_blocks) {
    ISAI_REQUIRE(plan, "isai_bsrilu0_plan_info: null plan");
    if (nlevels) *nlevels = plan->nlevels;
    if (max_row_blocks) *max_row_blocks = plan->max_row;
    return ISAI_OK;
}

int isai_bsrilu0_factor(isai_ilu_plan_t P, double *d_val, int *zero_pivot, isai_stream_t stream) {
    ISAI_REQUIRE(P && d_val && zero_pivot, "isai_bsrilu0_factor: null argument");
    *zero_pivot = -1;
    if (P->nb == 0) return ISAI_OK;
    hipStream_t st = (hipStream_t)stream;
    // scratch: the published factor (sentinel-filled), then ticket, zero-pivot word, error word
    char *scratch = nullptr;
    const size_t b2 = (size_t)P->bs * P->bs;
    const size_t nval = b2 * (size_t)P->nnzb;
    const size_t off_words = sizeof(double) * ((nval + 31) / 32 * 32);
    const size_t bytes = off_words + 4 * sizeof(int);
    ISAI_HIP(hipMallocAsync((void **)&scratch, bytes, st));
    double *fac = (double *)scratch;
    int *words = (int *)(scratch + off_words);  // [0] ticket, [1] zero pivot, [2] error
    const unsigned fill_grid = (unsigned)std::min<size_t>((nval + 255) / 256, 16384);
    ilu_sentinel_kernel<<<fill_grid ? fill_grid : 1, 256, 0, st>>>(fac, nval);
    ISAI_HIP(hipMemsetAsync(words, 0, 4 * sizeof(int), st));
    ISAI_HIP(hipMemsetD32Async((hipDeviceptr_t)&words[1], INT_MAX, 1, st));
    unsigned *ticket = (unsigned *)&words[0];
    switch (P->bs) {
        case 1: launch_ilu<1>(P, d_val, fac, ticket, &words[1], &words[2], st); break;
        case 2: launch_ilu<2>(P, d_val, fac, ticket, &words[1], &words[2], st); break;
        case 3: launch_ilu<3>(P, d_val, fac, ticket, &words[1], &words[2], st); break;
        case 4: launch_ilu<4>(P, d_val, fac, ticket, &words[1], &words[2], st); break;
        case 5: launch_ilu<5>(P, d_val, fac, ticket, &words[1], &words[2], st); break;
        case 6: launch_ilu<6>(P, d_val, fac, ticket, &words[1], &words[2], st); break;
        case 7: launch_ilu<7>(P, d_val, fac, ticket, &words[1], &words[2], st); break;
        case 8: launch_ilu<8>(P, d_val, fac, ticket, &words[1], &words[2], st); break;
    }
    ISAI_LAUNCH_CHECK();
    int h[4];
    ISAI_HIP(hipMemcpyAsync(h, words, sizeof(h), hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipStreamSynchronize(st));
    const int cs = h[2] ? ISAI_OK : copy_f64_d2d(d_val, fac, nval, st);
    ISAI_HIP(hipFreeAsync(scratch, st));
    if (cs != ISAI_OK) return cs;
    if (h[2]) {
        set_error("bsrilu0: a dependency wait exceeded its spin limit");
        return ISAI_ERR_TIMEOUT;
    }
    if (h[1] != INT_MAX) {
        *zero_pivot = h[1];
        set_error("bsrilu0: numerical zero at U(%d,%d)", h[1], h[1]);
        return ISAI_ERR_ZERO_PIVOT;
    }
    return ISAI_OK;
}

int isai_bsrilu0_destroy(isai_ilu_plan_t P) {
    if (!P) return ISAI_OK;
    (void)hipFree(P->d_order);
    (void)hipFree(P->d_diag);
    (void)hipFree(P->d_upd_ptr);
    (void)hipFree(P->d_upd_q);
    (void)hipFree(P->d_upd_src);
    delete P;
    return ISAI_OK;
}

int isai_bsrilu0_inplace(int nb, int nnzb, int bs, const int *d_row, const int *d_col, double *d_val,
                         int *zero_pivot, isai_stream_t stream) {
    ISAI_REQUIRE(zero_pivot, "isai_bsrilu0_inplace: null zero_pivot");
    isai_ilu_plan_t P = nullptr;
    int s = isai_bsrilu0_analysis(nb, nnzb, bs, d_row, d_col, &P, zero_pivot, stream);
    if (s != ISAI_OK) return s;
    s = isai_bsrilu0_factor(P, d_val, zero_pivot, stream);
    isai_bsrilu0_destroy(P);
    return s;
}

}  // extern "C"
