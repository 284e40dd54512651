// isai_build.hip -- ISAI of the packed ILU(0) factors on gfx950: diagonal-block
// inverses, pair lists, and the batched column-wise ISAI(L) / ISAI(U) build.
//
// Column j of ISAI(L) solves L(S_j,S_j) m_j = e_j by forward substitution over
// S_j (isai_bsr.h:288-399): for idx = 0..N-1, t = S_j[idx],
//     rhs = delta_tj I - sum_{pairs (p, src)} A[src] * Y_p ;  Y_idx = DinvL[t] * rhs.
// ISAI(U) runs the same recurrence backward (isai_bsr.h:161-242) with DinvU and
// the upper pair lists (the reference finds those blocks by a binary search per
// product, isai_bsr.h:207-209; here they are listed once, in the reference's
// descending p order).
//
// Launch shape: one wave64 workgroup handles several columns at once; each column
// owns a group of G = 16/32/64 lanes (bs^2 <= G), one lane per block element
// (i, jj).  Y blocks of the column stay in LDS for the later products (the
// reference's "cachedY" form); columns longer than the LDS budget read Y back from
// Mval (the "globalY" form, cacheY_isai.h:159-341 picks between the two the same
// way).  Every column of both factors is solved in a single launch.
#include <algorithm>
#include <mutex>

#include "common.hpp"
#include "internal.hpp"
#include "isai_bsr.h"
#include "scan.hpp"

namespace isai {

// ---------------------------------------------------------------- Dinv
// Gauss-Jordan without pivoting, |pivot| < eps fails (isai_bsr.h:60-88).
template <int BS>
__device__ bool gj_invert(double (&M)[BS * BS], double (&I)[BS * BS], double eps) {
#pragma unroll
    for (int e = 0; e < BS * BS; ++e) I[e] = (e % (BS + 1) == 0) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < BS; ++k) {
        const double piv = M[k * BS + k];
        if (fabs(piv) < eps) return false;
        const double invp = 1.0 / piv;
#pragma unroll
        for (int j = 0; j < BS; ++j) {
            M[k * BS + j] *= invp;
            I[k * BS + j] *= invp;
        }
#pragma unroll
        for (int i = 0; i < BS; ++i) {
            if (i == k) continue;
            const double f = M[i * BS + k];
#pragma unroll
            for (int j = 0; j < BS; ++j) {
                M[i * BS + j] -= f * M[k * BS + j];
                I[i * BS + j] -= f * I[k * BS + j];
            }
        }
    }
    return true;
}

__device__ __forceinline__ int find_in_row(const int *__restrict__ rp, const int *__restrict__ ci, int br, int bc) {
    int s = rp[br], e = rp[br + 1] - 1;
    while (s <= e) {
        const int m = (s + e) >> 1, c = ci[m];
        if (c == bc) return m;
        if (c < bc) s = m + 1; else e = m - 1;
    }
    return -1;
}

// One thread per block row: DinvL[k] = inverse of the unit-lower part of the
// packed diagonal block, DinvU[k] = inverse of its upper part (isai_bsr.h:91-158),
// with the +1e-12 diagonal retry of isai_bsr.h:114-120.
template <int BS>
__global__ void dinv_kernel(int nb, const int *__restrict__ rp, const int *__restrict__ ci,
                            const double *__restrict__ val, int lower, double *__restrict__ dinv) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    const int pos = find_in_row(rp, ci, k, k);
    double D[BS * BS], M[BS * BS], Inv[BS * BS];
#pragma unroll
    for (int e = 0; e < BS * BS; ++e) D[e] = pos >= 0 ? val[(size_t)pos * BS * BS + e] : 0.0;
#pragma unroll
    for (int i = 0; i < BS; ++i) {
#pragma unroll
        for (int j = 0; j < BS; ++j) {
            if (lower && j > i) D[i * BS + j] = 0.0;
            if (!lower && j < i) D[i * BS + j] = 0.0;
        }
        if (lower) D[i * BS + i] = 1.0;
    }
#pragma unroll
    for (int e = 0; e < BS * BS; ++e) M[e] = D[e];
    if (!gj_invert<BS>(M, Inv, 1e-14)) {
#pragma unroll
        for (int e = 0; e < BS * BS; ++e) M[e] = D[e] + ((e % (BS + 1) == 0) ? 1e-12 : 0.0);
        gj_invert<BS>(M, Inv, 1e-14);
    }
#pragma unroll
    for (int e = 0; e < BS * BS; ++e) dinv[(size_t)k * BS * BS + e] = Inv[e];
}

// ---------------------------------------------------------------- pairs
// One thread per column j; for each entry idx of S_j the merge of A's row t with
// S_j yields the pairs (isai_bsr.h:1378-1476).  Lower: p_local < idx, ascending.
// Upper: p_local > idx, stored in descending p_local order.
template <bool FILL>
__global__ void pairs_kernel(int nb, const int *__restrict__ Arow, const int *__restrict__ Acol,
                             const int *__restrict__ Mbrow, const int *__restrict__ Mcol, int upper,
                             int *__restrict__ counts_or_ptr, int *__restrict__ PairPLocal,
                             int *__restrict__ PairSrc) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nb) return;
    const int start = Mbrow[j], end = Mbrow[j + 1];
    for (int idx = 0; idx < end - start; ++idx) {
        const int t = Mcol[start + idx];
        int iA = Arow[t];
        const int a1 = Arow[t + 1];
        int iS = start, cnt = 0;
        int out = 0, n_here = 0;
        if (FILL) {
            out = counts_or_ptr[start + idx];
            n_here = counts_or_ptr[start + idx + 1] - out;
        }
        while (iA < a1 && iS < end) {
            const int colA = Acol[iA], colS = Mcol[iS];
            if (colA == colS) {
                const int pl = iS - start;
                if (upper ? (pl > idx) : (pl < idx)) {
                    if (FILL) {
                        const int slot = upper ? out + n_here - 1 - cnt : out + cnt;
                        PairPLocal[slot] = pl;
                        PairSrc[slot] = iA;
                    }
                    ++cnt;
                }
                ++iA;
                ++iS;
            } else if (colA < colS) {
                ++iA;
            } else {
                ++iS;
            }
        }
        if (!FILL) counts_or_ptr[start + idx] = cnt;
    }
}

// ---------------------------------------------------------------- ISAI build
template <int BS>
struct IsaiShape {
    static constexpr int B2 = BS * BS;
    static constexpr int G = B2 <= 16 ? 16 : (B2 <= 32 ? 32 : 64);  // lanes per column
    static constexpr int C = 64 / G;                                 // columns per wave
};

// blockIdx.x < nblk_lower builds ISAI(L) columns, the rest ISAI(U) columns.
template <int BS, bool GLOBAL_Y>
__global__ __launch_bounds__(64) void isai_build_kernel(int nb, const double *__restrict__ Aval,
                                                        FactorArgs Lf, FactorArgs Uf, int nblk_lower,
                                                        int ncache) {
    using S = IsaiShape<BS>;
    constexpr int B2 = S::B2;
    extern __shared__ double lds[];
    const bool upper = (int)blockIdx.x >= nblk_lower;
    const FactorArgs F = upper ? Uf : Lf;
    const int blk = upper ? (int)blockIdx.x - nblk_lower : (int)blockIdx.x;
    const int g = threadIdx.x / S::G;
    const int e = threadIdx.x % S::G;
    const bool lane_ok = e < B2;
    const int i = lane_ok ? e / BS : 0, jj = lane_ok ? e % BS : 0;
    // column of this lane group: position blk * C + g of the factor's schedule
    // (natural order, or the build plan's descending-Q order)
    const int pos = blk * S::C + g;
    const int j = (F.order && pos < nb) ? F.order[pos] : pos;
    double *rhsS = lds + g * (B2 + (GLOBAL_Y ? 0 : ncache * B2));
    double *Yc = rhsS + B2;

    int start = 0, N = 0;
    if (j < nb) {
        start = F.Mbrow[j];
        N = F.Mbrow[j + 1] - start;
    }
    int maxN = N;
#pragma unroll
    for (int off = S::G; off < 64; off <<= 1) maxN = max(maxN, __shfl_xor(maxN, off));
    // A column longer than the LDS cache (Nmax given too small) reads Y back from
    // Mval instead of overrunning the cache.
    const bool gy = GLOBAL_Y || maxN > ncache;

    for (int step = 0; step < maxN; ++step) {
        const bool act = step < N;
        const int idx = upper ? N - 1 - step : step;
        const int t = act ? F.Mcol[start + idx] : 0;
        double r = (act && lane_ok && t == j && i == jj) ? 1.0 : 0.0;
        int qb = 0, nq = 0;
        if (act) {
            qb = F.PairPtr[start + idx];
            nq = F.PairPtr[start + idx + 1] - qb;
        }
        int maxq = nq;
#pragma unroll
        for (int off = S::G; off < 64; off <<= 1) maxq = max(maxq, __shfl_xor(maxq, off));
        for (int q = 0; q < maxq; ++q) {
            if (q < nq && lane_ok) {
                const int pl = F.PairPLocal[qb + q];
                const int src = F.PairSrc[qb + q];
                const double *Arow = Aval + (size_t)src * B2 + i * BS;
                const double *Y = gy ? F.Mval + (size_t)(start + pl) * B2 + jj : Yc + pl * B2 + jj;
                // Same operations and order as the sequential restatement (no
                // contraction): s = sum_k A(i,k) Y(k,jj); r = r - s.
                double s = 0.0;
                if constexpr (BS == 4) {
                    const v2d a01 = *(const v2d *)Arow;
                    const v2d a23 = *(const v2d *)(Arow + 2);
                    s = s + a01.x * Y[0];
                    s = s + a01.y * Y[4];
                    s = s + a23.x * Y[8];
                    s = s + a23.y * Y[12];
                } else {
#pragma unroll
                    for (int k = 0; k < BS; ++k) s = s + Arow[k] * Y[k * BS];
                }
                r = r - s;
            }
        }
        if (lane_ok) rhsS[e] = r;
        __syncthreads();
        if (act && lane_ok) {
            const double *D = F.Dinv + (size_t)t * B2 + i * BS;
            double y = 0.0;
            if constexpr (BS == 4) {
                const v2d d01 = *(const v2d *)D;
                const v2d d23 = *(const v2d *)(D + 2);
                y = y + d01.x * rhsS[jj];
                y = y + d01.y * rhsS[4 + jj];
                y = y + d23.x * rhsS[8 + jj];
                y = y + d23.y * rhsS[12 + jj];
            } else {
#pragma unroll
                for (int k = 0; k < BS; ++k) y = y + D[k] * rhsS[k * BS + jj];
            }
            if (!gy) Yc[idx * B2 + e] = y;
            F.Mval[(size_t)(start + idx) * B2 + e] = y;
        }
        __syncthreads();
    }
}

constexpr int kLdsBudget = 64 * 1024;

template <int BS>
int launch_build(int nb, const double *Aval, const FactorArgs *L, const FactorArgs *U, int Nmax,
                 hipStream_t st) {
    using S = IsaiShape<BS>;
    const int nblk = (nb + S::C - 1) / S::C;
    const int nblk_lower = L ? nblk : 0;
    const int grid = (L ? nblk : 0) + (U ? nblk : 0);
    if (grid == 0) return ISAI_OK;
    const FactorArgs none{};
    const FactorArgs Lf = L ? *L : none, Uf = U ? *U : none;
    const size_t cached = (size_t)S::C * (S::B2 + (size_t)Nmax * S::B2) * sizeof(double);
    if (cached <= kLdsBudget) {
        isai_build_kernel<BS, false><<<grid, 64, cached, st>>>(nb, Aval, Lf, Uf, nblk_lower, Nmax);
    } else {
        const size_t lds = (size_t)S::C * S::B2 * sizeof(double);
        isai_build_kernel<BS, true><<<grid, 64, lds, st>>>(nb, Aval, Lf, Uf, nblk_lower, 0);
    }
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

int build_dispatch(int nb, int bs, const double *Aval, const FactorArgs *L, const FactorArgs *U, int Nmax,
                   hipStream_t st) {
    ISAI_REQUIRE(nb >= 0 && Nmax >= 0, "isai_build: bad size");
    if (nb == 0) return ISAI_OK;
    switch (bs) {
        case 1: return launch_build<1>(nb, Aval, L, U, Nmax, st);
        case 2: return launch_build<2>(nb, Aval, L, U, Nmax, st);
        case 3: return launch_build<3>(nb, Aval, L, U, Nmax, st);
        case 4: return launch_build<4>(nb, Aval, L, U, Nmax, st);
        case 5: return launch_build<5>(nb, Aval, L, U, Nmax, st);
        case 6: return launch_build<6>(nb, Aval, L, U, Nmax, st);
        case 7: return launch_build<7>(nb, Aval, L, U, Nmax, st);
        case 8: return launch_build<8>(nb, Aval, L, U, Nmax, st);
        default:
            set_error("isai_build: bs=%d outside 1..8", bs);
            return ISAI_ERR_UNSUPPORTED;
    }
}

}  // namespace isai

using namespace isai;

namespace isai {
// The count pass of isai_pairs_count without its read-back: PairPtr holds the
// exclusive scan of the per-entry pair counts, the total at PairPtr[Mnnz].
int pairs_count_async(int nb, int Mnnz, const int *d_Arow, const int *d_Acol, const int *d_Mbrow,
                      const int *d_Mcol, int upper, int *d_PairPtr, hipStream_t st) {
    ISAI_HIP(hipMemsetAsync(d_PairPtr, 0, sizeof(int) * ((size_t)Mnnz + 1), st));
    if (nb > 0) {
        pairs_kernel<false><<<grid_for(nb, 128), 128, 0, st>>>(nb, d_Arow, d_Acol, d_Mbrow, d_Mcol, upper,
                                                              d_PairPtr, nullptr, nullptr);
        ISAI_LAUNCH_CHECK();
    }
    return exclusive_scan_i32(d_PairPtr, d_PairPtr, (long long)Mnnz + 1, st);
}

// ---------------------------------------------------------------- allocation-free pair count
// isai_pairs_count_noalloc keeps the scan's tile sums (every level) and its overflow
// flag in this file-scope device array instead of allocating them.  It bounds
// Mnnz + 1 by kNoAllocMaxEntries (2^28 entries): 2^17 + 2^6 + 1 tile sums, one
// word for the flag, about 1 MiB per device, held until the process ends.  Every
// call on a device shares that device's copy, so calls that overlap on one GPU race.
constexpr long long kNoAllocMaxEntries = 1LL << 28;
constexpr long long kNoAllocWords = 131168;  // >= scan_ws_longs(2^28) + 1, a multiple of 32
static __device__ __align__(256) long long scratch[kNoAllocWords];

constexpr int kNoAllocMaxDevices = 64;
static long long *g_scratch_ptr[kNoAllocMaxDevices];
static std::mutex g_scratch_mu;

// Device address of the current device's copy of scratch, looked up once per
// device with that device current.
static int scratch_address(long long **out) {
    int dev = 0;
    ISAI_HIP(hipGetDevice(&dev));
    ISAI_REQUIRE(dev >= 0 && dev < kNoAllocMaxDevices, "isai_pairs_count_noalloc: device %d beyond %d", dev,
                 kNoAllocMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        ISAI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ISAI_REQUIRE(p != nullptr, "isai_pairs_count_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = static_cast<long long *>(p);
    }
    *out = g_scratch_ptr[dev];
    return ISAI_OK;
}
}  // namespace isai

extern "C" {

int isai_build_dinv(int nb, int bs, const int *d_Arow, const int *d_Acol, const double *d_Aval, int lower,
                    double *d_Dinv, isai_stream_t stream) {
    ISAI_REQUIRE(nb >= 0, "isai_build_dinv: nb < 0");
    if (nb == 0) return ISAI_OK;
    ISAI_REQUIRE(d_Arow && d_Acol && d_Aval && d_Dinv, "isai_build_dinv: null pointer");
    hipStream_t st = (hipStream_t)stream;
    const unsigned g = grid_for(nb, 128);
    switch (bs) {
        case 1: dinv_kernel<1><<<g, 128, 0, st>>>(nb, d_Arow, d_Acol, d_Aval, lower, d_Dinv); break;
        case 2: dinv_kernel<2><<<g, 128, 0, st>>>(nb, d_Arow, d_Acol, d_Aval, lower, d_Dinv); break;
        case 3: dinv_kernel<3><<<g, 128, 0, st>>>(nb, d_Arow, d_Acol, d_Aval, lower, d_Dinv); break;
        case 4: dinv_kernel<4><<<g, 128, 0, st>>>(nb, d_Arow, d_Acol, d_Aval, lower, d_Dinv); break;
        case 5: dinv_kernel<5><<<g, 128, 0, st>>>(nb, d_Arow, d_Acol, d_Aval, lower, d_Dinv); break;
        case 6: dinv_kernel<6><<<g, 128, 0, st>>>(nb, d_Arow, d_Acol, d_Aval, lower, d_Dinv); break;
        case 7: dinv_kernel<7><<<g, 128, 0, st>>>(nb, d_Arow, d_Acol, d_Aval, lower, d_Dinv); break;
        case 8: dinv_kernel<8><<<g, 128, 0, st>>>(nb, d_Arow, d_Acol, d_Aval, lower, d_Dinv); break;
        default:
            set_error("isai_build_dinv: bs=%d outside 1..8", bs);
            return ISAI_ERR_UNSUPPORTED;
    }
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

int isai_pairs_count(int nb, int Mnnz, const int *d_Arow, const int *d_Acol, const int *d_Mbrow,
                     const int *d_Mcol, int upper, int *d_PairPtr, int *total_pairs, isai_stream_t stream) {
    ISAI_REQUIRE(nb >= 0 && Mnnz >= 0 && total_pairs, "isai_pairs_count: bad argument");
    ISAI_REQUIRE(d_PairPtr, "isai_pairs_count: null PairPtr");
    hipStream_t st = (hipStream_t)stream;
    int s = isai::pairs_count_async(nb, Mnnz, d_Arow, d_Acol, d_Mbrow, d_Mcol, upper, d_PairPtr, st);
    if (s != ISAI_OK) return s;
    ISAI_HIP(hipMemcpyAsync(total_pairs, d_PairPtr + Mnnz, sizeof(int), hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipStreamSynchronize(st));
    return ISAI_OK;
}

int isai_pairs_count_noalloc(int nb, int Mnnz, const int *d_Arow, const int *d_Acol, const int *d_Mbrow,
                             const int *d_Mcol, int upper, int *d_PairPtr, int *total_pairs,
                             isai_stream_t stream) {
    ISAI_REQUIRE(nb >= 0 && Mnnz >= 0 && total_pairs, "isai_pairs_count_noalloc: bad argument");
    ISAI_REQUIRE(d_PairPtr, "isai_pairs_count_noalloc: null PairPtr");
    const long long n = (long long)Mnnz + 1;
    if (n > kNoAllocMaxEntries) {
        set_error("isai_pairs_count_noalloc: Mnnz + 1 = %lld exceeds the scratch limit of %lld entries", n,
                  kNoAllocMaxEntries);
        return ISAI_ERR_CAPACITY;
    }
    ISAI_REQUIRE(scan_ws_longs(n) + 1 <= kNoAllocWords, "isai_pairs_count_noalloc: scratch too small");
    long long *ws = nullptr;
    int s = scratch_address(&ws);
    if (s != ISAI_OK) return s;
    int *overflow = reinterpret_cast<int *>(ws + kNoAllocWords - 1);
    hipStream_t st = (hipStream_t)stream;
    const unsigned fill_grid = std::min(grid_for(n, 256), 1024u);
    fill_i32_kernel<<<fill_grid, 256, 0, st>>>(d_PairPtr, n, 0);
    ISAI_LAUNCH_CHECK();
    if (nb > 0) {
        pairs_kernel<false><<<grid_for(nb, 128), 128, 0, st>>>(nb, d_Arow, d_Acol, d_Mbrow, d_Mcol, upper,
                                                              d_PairPtr, nullptr, nullptr);
        ISAI_LAUNCH_CHECK();
    }
    s = exclusive_scan_i32_ws(d_PairPtr, d_PairPtr, n, ws, overflow, st);
    if (s != ISAI_OK) return s;
    ISAI_HIP(hipMemcpyAsync(total_pairs, d_PairPtr + Mnnz, sizeof(int), hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipStreamSynchronize(st));
    return ISAI_OK;
}

int isai_pairs_fill(int nb, const int *d_Arow, const int *d_Acol, const int *d_Mbrow, const int *d_Mcol,
                    int upper, const int *d_PairPtr, int *d_PairPLocal, int *d_PairSrc, isai_stream_t stream) {
    ISAI_REQUIRE(nb >= 0, "isai_pairs_fill: nb < 0");
    if (nb == 0) return ISAI_OK;
    pairs_kernel<true><<<grid_for(nb, 128), 128, 0, (hipStream_t)stream>>>(
        nb, d_Arow, d_Acol, d_Mbrow, d_Mcol, upper, const_cast<int *>(d_PairPtr), d_PairPLocal, d_PairSrc);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

int isai_build_lower(int nb, int bs, const double *d_Aval, const int *d_Mbrow, const int *d_Mcol,
                     const int *d_PairPtr, const int *d_PairPLocal, const int *d_PairSrc, const double *d_DinvL,
                     double *d_Mval, int Nmax, isai_stream_t stream) {
    FactorArgs L{d_Mbrow, d_Mcol, d_PairPtr, d_PairPLocal, d_PairSrc, d_DinvL, d_Mval};
    return build_dispatch(nb, bs, d_Aval, &L, nullptr, Nmax, (hipStream_t)stream);
}

int isai_build_upper(int nb, int bs, const double *d_Aval, const int *d_Mbrow, const int *d_Mcol,
                     const int *d_PairPtr, const int *d_PairPLocal, const int *d_PairSrc, const double *d_DinvU,
                     double *d_Mval, int Nmax, isai_stream_t stream) {
    FactorArgs U{d_Mbrow, d_Mcol, d_PairPtr, d_PairPLocal, d_PairSrc, d_DinvU, d_Mval};
    return build_dispatch(nb, bs, d_Aval, nullptr, &U, Nmax, (hipStream_t)stream);
}

int isai_build_lu(int nb, int bs, const double *d_Aval, const int *d_MbrowL, const int *d_McolL,
                  const int *d_PairPtrL, const int *d_PairPLocalL, const int *d_PairSrcL, const double *d_DinvL,
                  double *d_MvalL, const int *d_MbrowU, const int *d_McolU, const int *d_PairPtrU,
                  const int *d_PairPLocalU, const int *d_PairSrcU, const double *d_DinvU, double *d_MvalU,
                  int Nmax, isai_stream_t stream) {
    FactorArgs L{d_MbrowL, d_McolL, d_PairPtrL, d_PairPLocalL, d_PairSrcL, d_DinvL, d_MvalL};
    FactorArgs U{d_MbrowU, d_McolU, d_PairPtrU, d_PairPLocalU, d_PairSrcU, d_DinvU, d_MvalU};
    return build_dispatch(nb, bs, d_Aval, &L, &U, Nmax, (hipStream_t)stream);
}

int isai_build_upper_lookup(int nb, int bs, const int *d_Arow, const int *d_Acol, const double *d_Aval,
                            const int *d_Mbrow, const int *d_Mcol, int Mnnz, const double *d_DinvU,
                            double *d_Mval, int Nmax, isai_stream_t stream) {
    hipStream_t st = (hipStream_t)stream;
    int *pp = nullptr, *pl = nullptr, *ps = nullptr;
    int total = 0, s;
    ISAI_HIP(hipMalloc(&pp, sizeof(int) * ((size_t)Mnnz + 1)));
    s = isai_pairs_count(nb, Mnnz, d_Arow, d_Acol, d_Mbrow, d_Mcol, 1, pp, &total, stream);
    if (s == ISAI_OK) {
        hipError_t e1 = hipMalloc(&pl, sizeof(int) * (total ? total : 1));
        hipError_t e2 = hipMalloc(&ps, sizeof(int) * (total ? total : 1));
        if (e1 != hipSuccess || e2 != hipSuccess) {
            set_error("isai_build_upper_lookup: hipMalloc of %d pairs failed", total);
            s = ISAI_ERR_ALLOC;
        }
    }
    if (s == ISAI_OK) s = isai_pairs_fill(nb, d_Arow, d_Acol, d_Mbrow, d_Mcol, 1, pp, pl, ps, stream);
    if (s == ISAI_OK) s = isai_build_upper(nb, bs, d_Aval, d_Mbrow, d_Mcol, pp, pl, ps, d_DinvU, d_Mval, Nmax, stream);
    if (s == ISAI_OK) {
        hipError_t e = hipStreamSynchronize(st);
        if (e != hipSuccess) s = hip_fail(e, "hipStreamSynchronize", __FILE__, __LINE__);
    }
    (void)hipFree(pp);
    (void)hipFree(pl);
    (void)hipFree(ps);
    return s;
}

}  // extern "C"
