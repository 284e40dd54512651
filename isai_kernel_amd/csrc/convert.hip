// convert.hip -- device CSR -> BSR conversion (the device form of csr_to_bsr5,
// reference csr2bsr.h:14-76), bit-identical to the host form.
//
// One 64-lane workgroup per block row.
//   * The block columns of the row's entries are sorted and compacted: in
//     registers by a wave-level bitonic network (up to 512 entries, 8 per lane,
//     no LDS and no barrier), or for longer block rows in LDS.
//   * Each entry's value slot (block, row, column) is found in parallel, with the
//     block's first entry in CSR order (LDS atomicMin) and the number of entries
//     per slot (LDS atomicAdd).
//   * A slot with one entry takes it directly: the entry itself if it is its
//     block's first entry, else 0.0 + value -- what the host form's traversal gives
//     (assign the block's first entry, add every other into the zeroed block).
//   * Slots with several entries (duplicates) are filled by the host form's
//     traversal itself: lane lr walks scalar row lr in CSR order, applying only
//     those entries, so every value sees the same operations in the same order.
// LDS is sized on the host from the longest block row.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <vector>

#include "common.hpp"
#include "csr2bsr.h"
#include "scan.hpp"

namespace isai {

constexpr int kC2BMaxEntries = 16384;  // CSR entries of one block row
constexpr int kC2BWaves = 4;           // block rows per workgroup, one per wave

// Orders the LDS accesses of one wave (its lanes exchange data through LDS); the
// waves of a workgroup work on different block rows and never wait on each other.
__device__ __forceinline__ void c2b_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ int c2b_pow2(int m) {
    int p = 1;
    while (p < m) p <<= 1;
    return p;
}

// Ascending bitonic sort of keys[0, m2) by one wave (m2 a power of two).
__device__ void c2b_sort(int *keys, int m2) {
    const int lane = threadIdx.x & 63;
    for (int k = 2; k <= m2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = lane; t < m2; t += 64) {
                const int u = t ^ j;
                if (u > t) {
                    const bool up = (t & k) == 0;
                    const int a = keys[t], b = keys[u];
                    if ((a > b) == up) {
                        keys[t] = b;
                        keys[u] = a;
                    }
                }
            }
            c2b_wave_sync();
        }
    }
}

__device__ __forceinline__ int c2b_lower_bound(const int *a, int n, int key) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Loads and sorts the block columns of block row I; returns the number of unique
// block columns, compacted into ukeys[0, u).  keys and ukeys hold m2 ints.
__device__ int c2b_unique(int bs, const int *__restrict__ ci, int e0, int m, int m2, int *keys, int *ukeys) {
    const int lane = threadIdx.x & 63;
    for (int k = lane; k < m2; k += 64) keys[k] = k < m ? ci[e0 + k] / bs : INT_MAX;
    c2b_wave_sync();
    c2b_sort(keys, m2);
    // compaction: each lane takes a contiguous segment of the sorted keys
    const int seg = (m + 63) / 64;
    const int s0 = min(m, lane * seg), s1 = min(m, s0 + seg);
    int c = 0;
    for (int k = s0; k < s1; ++k) c += (k == 0 || keys[k] != keys[k - 1]);
    int incl = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(incl, off);
        if (lane >= off) incl += t;
    }
    int w = incl - c;
    for (int k = s0; k < s1; ++k)
        if (k == 0 || keys[k] != keys[k - 1]) ukeys[w++] = keys[k];
    const int u = __shfl(incl, 63);
    c2b_wave_sync();
    return u;
}

// Ascending bitonic sort of 64 * KPL keys held KPL per lane (element e = lane * KPL
// + i) by one wave: partners in other lanes are exchanged with shuffles, partners
// in the same lane swapped in registers.
template <int KPL>
__device__ __forceinline__ void c2b_sort_regs(int (&key)[KPL]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 2; k <= 64 * KPL; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= KPL) {
                const int lj = j / KPL;
                const bool lower = (lane & lj) == 0;
#pragma unroll
                for (int i = 0; i < KPL; ++i) {
                    const bool up = ((lane * KPL + i) & k) == 0;
                    const int other = __shfl_xor(key[i], lj);
                    key[i] = (lower == up) ? min(key[i], other) : max(key[i], other);
                }
            } else {
#pragma unroll
                for (int i = 0; i < KPL; ++i) {
                    if (i & j) continue;
                    const bool up = ((lane * KPL + i) & k) == 0;
                    const int a = key[i], b = key[i | j];
                    key[i] = up ? min(a, b) : max(a, b);
                    key[i | j] = up ? max(a, b) : min(a, b);
                }
            }
        }
    }
}

// c2b_unique for m <= 64 * KPL entries, sorted in registers; ukeys in LDS.
template <int KPL>
__device__ int c2b_unique_regs(int bs, const int *__restrict__ ci, int e0, int m, int *ukeys) {
    const int lane = threadIdx.x & 63;
    int key[KPL];
#pragma unroll
    for (int i = 0; i < KPL; ++i) {
        const int e = lane * KPL + i;
        key[i] = e < m ? ci[e0 + e] / bs : INT_MAX;
    }
    c2b_sort_regs<KPL>(key);
    const int prev_last = __shfl_up(key[KPL - 1], 1);
    bool flag[KPL];
    int c = 0;
#pragma unroll
    for (int i = 0; i < KPL; ++i) {
        const int prev = i ? key[i - 1] : prev_last;
        flag[i] = key[i] != INT_MAX && ((lane == 0 && i == 0) || key[i] != prev);
        c += flag[i];
    }
    int incl = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(incl, off);
        if (lane >= off) incl += t;
    }
    int w = incl - c;
#pragma unroll
    for (int i = 0; i < KPL; ++i)
        if (flag[i]) ukeys[w++] = key[i];
    const int u = __shfl(incl, 63);
    c2b_wave_sync();
    return u;
}

// Unique sorted block columns of block row I (m entries from e0) in ukeys; keys is
// LDS scratch for rows longer than the register networks take.
__device__ int c2b_unique_any(int bs, const int *__restrict__ ci, int e0, int m, int *keys, int *ukeys) {
    const int m2 = c2b_pow2(max(m, 1));
    if (m2 <= 64) return c2b_unique_regs<1>(bs, ci, e0, m, ukeys);
    if (m2 <= 128) return c2b_unique_regs<2>(bs, ci, e0, m, ukeys);
    if (m2 <= 256) return c2b_unique_regs<4>(bs, ci, e0, m, ukeys);
    if (m2 <= 512) return c2b_unique_regs<8>(bs, ci, e0, m, ukeys);
    return c2b_unique(bs, ci, e0, m, m2, keys, ukeys);
}

__global__ __launch_bounds__(64 * kC2BWaves) void c2b_count_kernel(int nb, int bs, const int *__restrict__ rp,
                                                                   const int *__restrict__ ci, int m2max,
                                                                   int *counts) {
    extern __shared__ int lds_all[];
    const int w = threadIdx.x >> 6;
    const int I = blockIdx.x * (int)(blockDim.x >> 6) + w;
    if (I >= nb) return;
    int *lds = lds_all + (size_t)w * 2 * m2max;
    const int e0 = rp[I * bs], m = rp[(I + 1) * bs] - e0;
    const int u = c2b_unique_any(bs, ci, e0, m, lds, lds + m2max);
    if ((threadIdx.x & 63) == 0) counts[I] = u;
}

// LDS ints of one wave of the fill kernel (its doubles follow, 8-byte aligned).
__host__ __device__ inline size_t c2b_fill_wave_ints(int m2max, int umax, int bs) {
    return 2 * (size_t)m2max + (((size_t)umax + (size_t)umax * bs * bs + 1) & ~(size_t)1);
}

__global__ __launch_bounds__(64 * kC2BWaves) void c2b_fill_kernel(int nb, int bs, const int *__restrict__ rp,
                                                                  const int *__restrict__ ci,
                                                                  const double *__restrict__ v,
                                                                  const int *__restrict__ brp, int *__restrict__ bci,
                                                                  double *__restrict__ bv, int m2max, int umax,
                                                                  int *err) {
    extern __shared__ int lds_all[];
    const int bs2 = bs * bs;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int I = blockIdx.x * (int)(blockDim.x >> 6) + wv;
    if (I >= nb) return;
    const size_t wave_bytes = sizeof(int) * c2b_fill_wave_ints(m2max, umax, bs) + sizeof(double) * (size_t)umax * bs2;
    int *lds = (int *)((char *)lds_all + wv * wave_bytes);
    int *keys = lds, *ukeys = lds + m2max, *first = ukeys + m2max;
    int *cnt = first + umax;  // umax * bs2 entries per slot
    double *vals = (double *)(lds + c2b_fill_wave_ints(m2max, umax, bs));
    __shared__ int s_dup_all[kC2BWaves];
    int &s_dup = s_dup_all[wv];
    const int e0 = rp[I * bs], m = rp[(I + 1) * bs] - e0;
    int rstart[8];  // first entry of each scalar row of the block row
#pragma unroll
    for (int r = 0; r < 8; ++r) rstart[r] = r < bs ? rp[I * bs + r] : INT_MAX;
    const int u = c2b_unique_any(bs, ci, e0, m, keys, ukeys);
    if (u != brp[I + 1] - brp[I]) {  // row pointer not from the count pass: write nothing
        if (lane == 0) atomicOr(err, 1);
        return;
    }
    const int nslot = u * bs2;
    for (int q = lane; q < u; q += 64) first[q] = INT_MAX;
    for (int t = lane; t < nslot; t += 64) {
        vals[t] = 0.0;
        cnt[t] = 0;
    }
    if (lane == 0) s_dup = 0;
    c2b_wave_sync();
    // slot of entry k: block q, scalar row lr (from the row starts), column c % bs
    auto slot_of = [&](int k, int &q) {
        const int c = ci[e0 + k];
        q = c2b_lower_bound(ukeys, u, c / bs);
        int lr = 0;
#pragma unroll
        for (int r = 1; r < 8; ++r) lr += rstart[r] <= e0 + k;
        return q * bs2 + lr * bs + c % bs;
    };
    for (int k = lane; k < m; k += 64) {
        int q;
        const int sl = slot_of(k, q);
        atomicMin(&first[q], e0 + k);
        atomicAdd(&cnt[sl], 1);
    }
    c2b_wave_sync();
    for (int k = lane; k < m; k += 64) {
        int q;
        const int sl = slot_of(k, q);
        if (cnt[sl] == 1) {
            const double x = v[e0 + k];
            vals[sl] = e0 + k == first[q] ? x : 0.0 + x;
        } else {
            s_dup = 1;
        }
    }
    c2b_wave_sync();
    if (s_dup && lane < bs) {
        const int lr = lane, row = I * bs + lr;
        for (int p = rp[row]; p < rp[row + 1]; ++p) {
            const int c = ci[p];
            const int q = c2b_lower_bound(ukeys, u, c / bs);
            const int sl = q * bs2 + lr * bs + c % bs;
            if (cnt[sl] == 1) continue;
            double &dst = vals[sl];
            if (p == first[q]) dst = v[p];
            else dst += v[p];
        }
    }
    c2b_wave_sync();
    const int base = brp[I];
    for (int q = lane; q < u; q += 64) bci[base + q] = ukeys[q];
    for (int t = lane; t < nslot; t += 64) bv[(size_t)base * bs2 + t] = vals[t];
}

// Per block row: CSR entries m = rp[(I+1) bs] - rp[I bs] and (with brp) unique
// blocks u = brp[I+1] - brp[I].  out[0] = max m, out[1] = max(-m) (positive when
// the row pointer decreases), out[2] = max u; out is zeroed by the caller.  Each
// workgroup reduces in registers and LDS and adds one atomic per quantity, with a
// capped grid (same-address global atomics serialise).
constexpr int kRowMaxBlock = 256, kRowMaxGrid = 256;
__global__ __launch_bounds__(kRowMaxBlock) void c2b_rowmax_kernel(int nb, int bs, const int *__restrict__ rp,
                                                                  const int *__restrict__ brp, int *out) {
    __shared__ int s_max[3];
    if (threadIdx.x < 3) s_max[threadIdx.x] = 0;
    __syncthreads();
    int a = 0, b = 0, c = 0;
    for (int I = blockIdx.x * kRowMaxBlock + threadIdx.x; I < nb; I += gridDim.x * kRowMaxBlock) {
        const int m = rp[(size_t)(I + 1) * bs] - rp[(size_t)I * bs];
        a = max(a, m);
        b = max(b, -m);
        if (brp) c = max(c, brp[I + 1] - brp[I]);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        a = max(a, __shfl_xor(a, off));
        b = max(b, __shfl_xor(b, off));
        c = max(c, __shfl_xor(c, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&s_max[0], a);
        atomicMax(&s_max[1], b);
        atomicMax(&s_max[2], c);
    }
    __syncthreads();
    if (threadIdx.x < 3) atomicMax(&out[threadIdx.x], s_max[threadIdx.x]);
}

// Longest block row (CSR entries; and, with d_brp, unique blocks) computed on the
// device, and the LDS sizes that follow from it.
static int c2b_prepare(int n, int bs, const int *d_rp, const int *d_brp, hipStream_t st, int *m2max, int *umax) {
    const int nb = n / bs;
    int *d_out = nullptr;
    int h[3] = {0, 0, 0};
    ISAI_HIP(hipMallocAsync((void **)&d_out, sizeof(int) * 3, st));
    ISAI_HIP(hipMemsetAsync(d_out, 0, sizeof(int) * 3, st));
    c2b_rowmax_kernel<<<std::min<unsigned>(grid_for(nb, kRowMaxBlock), kRowMaxGrid), kRowMaxBlock, 0, st>>>(
        nb, bs, d_rp, d_brp, d_out);
    ISAI_LAUNCH_CHECK();
    ISAI_HIP(hipMemcpyAsync(h, d_out, sizeof(h), hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipFreeAsync(d_out, st));
    ISAI_HIP(hipStreamSynchronize(st));
    if (h[1] > 0) {
        set_error("csr_to_bsr (device): the row pointer decreases within a block row");
        return ISAI_ERR_INVALID_ARG;
    }
    const int mmax = std::max(h[0], 1);
    if (mmax > kC2BMaxEntries) {
        set_error("csr_to_bsr (device): a block row holds %d CSR entries (limit %d); use the host form", mmax,
                  kC2BMaxEntries);
        return ISAI_ERR_UNSUPPORTED;
    }
    int p = 1;
    while (p < mmax) p <<= 1;
    *m2max = p;
    if (umax) *umax = std::max(h[2], 1);
    return ISAI_OK;
}

}  // namespace isai

using namespace isai;

extern "C" {

int isai_csr_to_bsr_device_count(int n, int bs, const int *d_rp, const int *d_ci, int *d_brp, int *nnzb,
                                 isai_stream_t stream) {
    ISAI_REQUIRE(d_rp && d_ci && d_brp && nnzb, "isai_csr_to_bsr_device_count: null argument");
    ISAI_REQUIRE(bs >= 1 && bs <= 8 && n >= 0, "isai_csr_to_bsr_device_count: bad size n=%d bs=%d", n, bs);
    ISAI_REQUIRE(n % bs == 0, "csr_to_bsr: n = %d is not a multiple of block size %d", n, bs);
    hipStream_t st = (hipStream_t)stream;
    const int nb = n / bs;
    *nnzb = 0;
    if (nb == 0) {
        const int zero = 0;
        ISAI_HIP(hipMemcpyAsync(d_brp, &zero, sizeof(int), hipMemcpyHostToDevice, st));
        ISAI_HIP(hipStreamSynchronize(st));
        return ISAI_OK;
    }
    int m2max = 1;
    int s = c2b_prepare(n, bs, d_rp, nullptr, st, &m2max, nullptr);
    if (s != ISAI_OK) return s;
    // counts in d_brp[0, nb), then an exclusive scan over nb + 1 entries
    // four block rows per workgroup while their LDS (sort keys) fits in 64 KB, else one
    const size_t wave_lds = sizeof(int) * 2 * (size_t)m2max;
    if (wave_lds > 64 * 1024) {
        set_error("csr_to_bsr (device): a block row needs %zu bytes of LDS (limit 65536); use the host form", wave_lds);
        return ISAI_ERR_UNSUPPORTED;
    }
    const int waves = wave_lds * kC2BWaves <= 64 * 1024 ? kC2BWaves : 1;
    c2b_count_kernel<<<(nb + waves - 1) / waves, 64 * waves, wave_lds * waves, st>>>(nb, bs, d_rp, d_ci, m2max,
                                                                                     d_brp);
    ISAI_LAUNCH_CHECK();
    const int zero = 0;
    ISAI_HIP(hipMemcpyAsync(d_brp + nb, &zero, sizeof(int), hipMemcpyHostToDevice, st));
    s = exclusive_scan_i32(d_brp, d_brp, (long long)nb + 1, st);
    if (s != ISAI_OK) return s;
    ISAI_HIP(hipMemcpyAsync(nnzb, d_brp + nb, sizeof(int), hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipStreamSynchronize(st));
    return ISAI_OK;
}

int isai_csr_to_bsr_device_fill(int n, int bs, const int *d_rp, const int *d_ci, const double *d_v,
                                const int *d_brp, int *d_bci, double *d_bv, isai_stream_t stream) {
    ISAI_REQUIRE(d_rp && d_ci && d_v && d_brp && d_bci && d_bv, "isai_csr_to_bsr_device_fill: null argument");
    ISAI_REQUIRE(bs >= 1 && bs <= 8 && n >= 0, "isai_csr_to_bsr_device_fill: bad size n=%d bs=%d", n, bs);
    ISAI_REQUIRE(n % bs == 0, "csr_to_bsr: n = %d is not a multiple of block size %d", n, bs);
    hipStream_t st = (hipStream_t)stream;
    const int nb = n / bs;
    if (nb == 0) return ISAI_OK;
    int m2max = 1, umax = 1;
    int s = c2b_prepare(n, bs, d_rp, d_brp, st, &m2max, &umax);
    if (s != ISAI_OK) return s;
    const size_t bs2 = (size_t)bs * bs;
    const size_t wave_lds = sizeof(int) * c2b_fill_wave_ints(m2max, umax, bs) + sizeof(double) * (size_t)umax * bs2;
    if (wave_lds > 64 * 1024) {
        set_error("csr_to_bsr (device): a block row needs %zu bytes of LDS (limit 65536); use the host form", wave_lds);
        return ISAI_ERR_UNSUPPORTED;
    }
    // four block rows per workgroup while their LDS fits in 64 KB, else one
    const int waves = wave_lds * kC2BWaves <= 64 * 1024 ? kC2BWaves : 1;
    int *d_err = nullptr;
    ISAI_HIP(hipMallocAsync((void **)&d_err, sizeof(int), st));
    ISAI_HIP(hipMemsetAsync(d_err, 0, sizeof(int), st));
    c2b_fill_kernel<<<(nb + waves - 1) / waves, 64 * waves, wave_lds * waves, st>>>(nb, bs, d_rp, d_ci, d_v, d_brp,
                                                                                    d_bci, d_bv, m2max, umax, d_err);
    ISAI_LAUNCH_CHECK();
    int err = 0;
    ISAI_HIP(hipMemcpyAsync(&err, d_err, sizeof(int), hipMemcpyDeviceToHost, st));
    ISAI_HIP(hipStreamSynchronize(st));
    ISAI_HIP(hipFreeAsync(d_err, st));
    if (err) {
        set_error("csr_to_bsr (device): the block row pointer does not match the matrix (use the count pass)");
        return ISAI_ERR_INVALID_ARG;
    }
    return ISAI_OK;
}

}  // extern "C"
