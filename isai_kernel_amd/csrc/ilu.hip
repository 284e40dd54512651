// ilu.hip -- block ILU(0) in place on gfx950 (replaces cusparseDbsrilu02 as called
// by bsrilu02_inplace, reference bsrilu.h:20-66).
//
// Result layout = cuSPARSE's DIRECTION_ROW packed factor: strictly-lower blocks
// hold L_ik, diagonal blocks hold L_ii\U_ii (unit-lower \ upper), strictly-upper
// blocks hold U_ij, with A = L U on every stored block.
//
// Schedule: the analysis (host) computes the level of every block row and the
// rows in (level, row) order.  The factor kernel gives each block row one wave64
// workgroup; a workgroup takes the next row of that order from an atomic ticket,
// waits on the per-row ready flags of the rows it depends on (no global barrier,
// no host round trip between levels), factors its row with the row staged in LDS,
// and publishes its flag with release semantics.  Because tickets follow level
// order, a row only ever waits on rows whose workgroups already hold a ticket.
//
// Per row i:
//   for each stored k < i, ascending:   L_ik = A'_ik inv(U_kk)  (right substitution)
//       for each stored j > k with (k,j) stored:  A'_ij -= L_ik U_kj
//   A'_ii = L_ii U_ii (scalar LU, no pivoting, FP64)
//   for each stored j > i:   U_ij = inv(L_ii) A'_ij            (forward substitution)
// Every scalar operation is the oracle's, in the same order, without contraction,
// so the factor is bit-identical to the sequential restatement.
#include <climits>
#include <vector>

#include "bsrilu.h"
#include "common.hpp"
#include "levels.hpp"

struct isai_ilu_plan {
    int nb = 0, nnzb = 0, bs = 0, nlevels = 0;
    const int *d_row = nullptr, *d_col = nullptr;
    int *d_order = nullptr, *d_diag = nullptr;
    int max_row = 0;
};

namespace isai {

constexpr int kIluMaxRow = 64;       // blocks of one row staged in LDS
constexpr int kIluSpinLimit = 1 << 20;

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

template <int BS>
__global__ __launch_bounds__(64) void ilu0_kernel(int nb, const int *__restrict__ rp, const int *__restrict__ ci,
                                                  double *__restrict__ val, const int *__restrict__ order,
                                                  const int *__restrict__ diagpos, int *__restrict__ flags,
                                                  unsigned int *__restrict__ ticket, int *__restrict__ zero_pivot,
                                                  int *__restrict__ err) {
    using S = IluShape<BS>;
    constexpr int B2 = S::B2;
    constexpr int CB = 64 / BS;  // blocks per pass when one lane owns one row or column
    __shared__ double rowv[kIluMaxRow * B2];
    __shared__ int s_row;
    const int lane = threadIdx.x;
    if (lane == 0) s_row = order[atomicAdd(ticket, 1u)];
    __syncthreads();
    const int i = s_row;
    const int rs = rp[i], re = rp[i + 1], len = re - rs;
    const int dp = diagpos[i];
    const int g = lane / S::G, e = lane % S::G;
    const bool ok = e < B2;
    const int r = ok ? e / BS : 0, c = ok ? e % BS : 0;

    for (int x = lane; x < len * B2; x += 64) rowv[x] = val[(size_t)rs * B2 + x];

    // Wait for every row this row depends on (stored k < i).
    bool timed_out = false;
    for (int p = rs + lane; p < dp; p += 64) {
        const int k = ci[p];
        int spins = 0;
        while (__hip_atomic_load(&flags[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
            if (++spins > kIluSpinLimit) { timed_out = true; break; }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    if (timed_out) atomicExch(err, 1);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    __syncthreads();

    for (int p = rs; p < dp; ++p) {
        const int k = ci[p];
        const int lp = p - rs;
        // L_ik = A'_ik inv(U_kk) by right substitution with the upper part of the
        // packed diagonal block of row k; lane r owns row r.
        if (lane < BS) {
            const double *Dk = val + (size_t)diagpos[k] * B2;
            double *X = rowv + lp * B2 + lane * BS;
#pragma unroll
            for (int cc = 0; cc < BS; ++cc) {
                double sacc = X[cc];
#pragma unroll
                for (int m = 0; m < cc; ++m) sacc = sacc - X[m] * Dk[m * BS + cc];
                X[cc] = sacc / Dk[cc * BS + cc];
            }
        }
        __syncthreads();
        // A'_ij -= L_ik U_kj for the stored j > k with (k, j) stored; one lane per
        // element (r, c), G lanes per block.
        for (int q0 = p + 1; q0 < re; q0 += S::C) {
            const int q = q0 + g;
            if (q < re && ok) {
                const int pkj = ilu_find(rp, ci, k, ci[q]);
                if (pkj >= 0) {
                    const double *Ukj = val + (size_t)pkj * B2;
                    const double *Lr = rowv + lp * B2 + r * BS;
                    double sacc = 0.0;
#pragma unroll
                    for (int m = 0; m < BS; ++m) sacc = sacc + Lr[m] * Ukj[m * BS + c];
                    rowv[(q - rs) * B2 + e] = rowv[(q - rs) * B2 + e] - sacc;
                }
            }
        }
        __syncthreads();
    }

    // Diagonal block: scalar LU without pivoting, packed L_ii\U_ii.
    if (lane == 0) {
        double *D = rowv + (dp - rs) * B2;
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
    __syncthreads();

    // U_ij = inv(L_ii) A'_ij for the stored j > i by forward substitution with the
    // unit-lower part of the diagonal block; lane (block, column c).
    {
        const double *D = rowv + (dp - rs) * B2;
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
    __syncthreads();
    for (int x = lane; x < len * B2; x += 64) val[(size_t)rs * B2 + x] = rowv[x];
    __syncthreads();
    if (lane == 0) __hip_atomic_store(&flags[i], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

template <int BS>
void launch_ilu(const isai_ilu_plan *P, double *val, int *flags, unsigned *ticket, int *zp, int *err,
                hipStream_t st) {
    ilu0_kernel<BS><<<P->nb, 64, 0, st>>>(P->nb, P->d_row, P->d_col, val, P->d_order, P->d_diag, flags, ticket,
                                          zp, err);
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
    if (hipMalloc(&P->d_order, sizeof(int) * (nb ? nb : 1)) != hipSuccess ||
        hipMalloc(&P->d_diag, sizeof(int) * (nb ? nb : 1)) != hipSuccess) {
        (void)hipFree(P->d_order);
        delete P;
        set_error("bsrilu0 analysis: hipMalloc failed");
        return ISAI_ERR_ALLOC;
    }
    ISAI_HIP(hipMemcpyAsync(P->d_order, order.data(), sizeof(int) * nb, hipMemcpyHostToDevice, st));
    ISAI_HIP(hipMemcpyAsync(P->d_diag, diag.data(), sizeof(int) * nb, hipMemcpyHostToDevice, st));
    ISAI_HIP(hipStreamSynchronize(st));
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
    // scratch: ready flags, then ticket, zero-pivot word, error word
    char *scratch = nullptr;
    const size_t off_flags = 0;
    const size_t off_words = sizeof(int) * ((P->nb + 63) / 64 * 64);
    const size_t bytes = off_words + 4 * sizeof(int);
    ISAI_HIP(hipMallocAsync((void **)&scratch, bytes, st));
    int *flags = (int *)(scratch + off_flags);
    int *words = (int *)(scratch + off_words);  // [0] ticket, [1] zero pivot, [2] error
    ISAI_HIP(hipMemsetAsync(flags, 0, bytes - off_flags, st));
    ISAI_HIP(hipMemsetD32Async((hipDeviceptr_t)&words[1], INT_MAX, 1, st));
    unsigned *ticket = (unsigned *)&words[0];
    switch (P->bs) {
        case 1: launch_ilu<1>(P, d_val, flags, ticket, &words[1], &words[2], st); break;
        case 2: launch_ilu<2>(P, d_val, flags, ticket, &words[1], &words[2], st); break;
        case 3: launch_ilu<3>(P, d_val, flags, ticket, &words[1], &words[2], st); break;
        case 4: launch_ilu<4>(P, d_val, flags, ticket, &words[1], &words[2], st); break;
        case 5: launch_ilu<5>(P, d_val, flags, ticket, &words[1], &words[2], st); break;
        case 6: launch_ilu<6>(P, d_val, flags, ticket, &words[1], &words[2], st); break;
        case 7: launch_ilu<7>(P, d_val, flags, ticket, &words[1], &words[2], st); break;
        case 8: launch_ilu<8>(P, d_val, flags, ticket, &words[1], &words[2], st); break;
    }
    ISAI_LAUNCH_CHECK();
    int h[4];
    ISAI_HIP(hipMemcpyAsync(h, words, sizeof(h), hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipStreamSynchronize(st));
    ISAI_HIP(hipFreeAsync(scratch, st));
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
