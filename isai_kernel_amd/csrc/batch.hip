// batch.hip -- the batching form: uneven block-diagonal systems with 1..16
// right-hand sides each, one batched setup, one batched BiCGSTAB whose
// iterations are captured as a hipGraph and replayed.
//
// Layout: pair p = (system s, right-hand side j) owns a contiguous vector segment
// of 4*nb_s doubles, padded to a multiple of kTile = 1024; the segments of one
// system follow each other with a fixed stride, and tile t of the multi-vector
// belongs to one pair (tile_pair[t]).  Reductions are per tile (one wave, fixed
// shape), then per pair over its contiguous tile range in order, so every result
// is deterministic and independent of how many iterations one replay holds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "internal.hpp"
#include "isai_batch.h"
#include "isai_solve.h"

namespace isai {

constexpr int kTile = 1024;
constexpr int kMaxRhs = 16;
constexpr int kItersPerReplay = 4;  // iterations in one captured graph
constexpr int kMaxRestarts = 8;
constexpr int kSpmmMaxWg = 256;  // largest SpMM workgroup the autotuner tries

// Chunked XCD remap: the hardware deals consecutive workgroups round-robin over
// the 8 XCDs; this hands each XCD runs of kXcdChunk consecutive logical workgroups
// in turn (all XCDs still stream neighbouring regions at once, and the x entries a
// run shares stay in one L2).  The tail past the last whole round keeps its order.
constexpr int kXcdChunk = 64;
__device__ __forceinline__ int xcd_chunk_remap(int bid, int nblocks) {
    constexpr int kXcd = 8;
    const int round = kXcd * kXcdChunk;
    if (bid >= nblocks / round * round) return bid;
    const int xcd = bid % kXcd, local = bid / kXcd;
    return ((local / kXcdChunk) * kXcd + xcd) * kXcdChunk + local % kXcdChunk;
}

// status: 0 running, 1 converged (recursive residual), 2 converged on s with the
// x update pending in this iteration, 3 breakdown, 4 done (true residual met),
// 5 iteration limit reached
struct PairState {
    double rho, rho_old, alpha, omega, beta, bnorm2, tol2, pad;
    int status, iters, first, pad2;
};

// nactive: pairs still running after an iteration (zeroed per iteration by the
// graph path; a running total in the persistent kernel); nrestart: running total
// of pairs sent back by a true-residual check; arrive / abort / spin_us: the
// persistent kernel's grid barrier (arrivals, abandoned wait, wait bound in us).
struct Ctl {
    int nactive, maxit, nrestart, abort;
    unsigned long long arrive;
    long long spin_us;
};

struct BatchDev {
    const int *sys_of_row;      // [nb]
    const int *sys_ptr;         // [nsys+1] block rows
    const int *sys_nrhs;        // [nsys]
    const int *sys_pair;        // [nsys] first pair of the system
    const long long *sys_vec;   // [nsys] offset of the system's first segment
    const long long *sys_step;  // [nsys] padded segment length
    const int *tile_pair;       // [ntiles]
    const int *ptile;           // [npairs+1]
    PairState *ps;              // [npairs]
    double *part0, *part1;      // [ntiles]
    Ctl *ctl;
    int ntiles, npairs;
};

__device__ __forceinline__ unsigned active_mask(const BatchDev &B, int p0, int r, int all) {
    unsigned m = 0;
    for (int j = 0; j < r; ++j)
        if (all || B.ps[p0 + j].status == 0) m |= 1u << j;
    return m;
}

// --------------------------------------------------------------- multi-RHS SpMM
// Eight lanes per block row (one 128-byte line of FP64 block values per
// instruction, as in the single-system product); every block is loaded once and
// applied to all R right-hand sides of its system.  R is a template parameter so
// the R vector loads of a block are issued back to back before the FMAs; pairs
// that have stopped are skipped by the caller only when the whole row is idle,
// and their outputs are never written.
struct RowCtx {
    int br, base, rs, re, lane;
    unsigned mask;
    long long v0, step;
};

// Block row br handled by an 8-lane group; lane = position in the group.
__device__ __forceinline__ bool row_context(const BatchDev &B, const int *rp, int nb, int br, int lane, int all,
                                            RowCtx &q, int &r) {
    q.br = br;
    if (q.br >= nb) return false;  // whole 8-lane groups leave together
    const int sys = B.sys_of_row[q.br];
    r = B.sys_nrhs[sys];
    q.mask = active_mask(B, B.sys_pair[sys], r, all);
    if (!q.mask) return false;
    q.base = B.sys_ptr[sys];
    q.v0 = B.sys_vec[sys];
    q.step = B.sys_step[sys];
    q.rs = rp[q.br];
    q.re = rp[q.br + 1];
    q.lane = lane;
    return true;
}

// Multi-vectors are FP64, or FP32 where they hold the outputs of a reduced-precision
// ISAI product (FP32 accumulations, so the FP32 copy is exact): loads widen, stores
// of FP32 results narrow without rounding.
template <class T>
__device__ __forceinline__ v2d load2(const T *p) {
    if constexpr (sizeof(T) == 8) {
        return *(const v2d *)p;
    } else {
        const float2 f = *(const float2 *)p;
        return v2d{(double)f.x, (double)f.y};
    }
}
// four consecutive entries (16-byte aligned) rounded to FP32
template <class T>
__device__ __forceinline__ v4f load4f(const T *p) {
    if constexpr (sizeof(T) == 8) {
        const v2d a = *(const v2d *)p, b = *(const v2d *)(p + 2);
        return v4f{(float)a.x, (float)a.y, (float)b.x, (float)b.y};
    } else {
        return *(const v4f *)p;
    }
}
template <class T>
__device__ __forceinline__ void store4f(T *p, v4f v) {
    if constexpr (sizeof(T) == 8) {
        *(v2d *)p = v2d{(double)v.x, (double)v.y};
        *(v2d *)(p + 2) = v2d{(double)v.z, (double)v.w};
    } else {
        *(v4f *)p = v;
    }
}

// FP64 values: lane l holds row l/2, columns 2(l%2)..+1; FMA chain per lane, then
// the two lanes of a row are added.
template <int R, class XT>
__device__ __forceinline__ void f64_row(const RowCtx &q, const int *__restrict__ ci, const v2d *__restrict__ val,
                                        const XT *__restrict__ X, double *__restrict__ Y) {
    constexpr int KU = R <= 2 ? 4 : (R <= 6 ? 2 : 1);
    const int h = q.lane & 1;
    double acc0[R], acc1[R];
#pragma unroll
    for (int j = 0; j < R; ++j) acc0[j] = acc1[j] = 0.0;
    const XT *xb = X + q.v0 + 2 * h;
#pragma unroll KU
    for (int k = q.rs; k < q.re; ++k) {
        const int c = ci[k] - q.base;
        const v2d a = __builtin_nontemporal_load(&val[(size_t)k * 8 + q.lane]);
        v2d x[R];
#pragma unroll
        for (int j = 0; j < R; ++j) x[j] = load2(xb + (size_t)c * 4 + j * q.step);
#pragma unroll
        for (int j = 0; j < R; ++j) {
            acc0[j] = fma(a.x, x[j].x, acc0[j]);
            acc1[j] = fma(a.y, x[j].y, acc1[j]);
        }
    }
    double *yr = Y + q.v0 + (size_t)(q.br - q.base) * 4 + (q.lane >> 1);
#pragma unroll
    for (int j = 0; j < R; ++j) {
        double s = acc0[j] + acc1[j];
        s += __shfl_xor(s, 1);
        if ((q.mask & (1u << j)) && h == 0) yr[j * q.step] = s;
    }
}

// FP32 / BF16 factor values (the single-system lane layouts): FP32 lane l holds row
// l%4 of block l/4 (two blocks per line); BF16 lane l holds rows 2(l%2), 2(l%2)+1 of
// block l/2 (four blocks per line).  FP32 accumulation, x rounded to FP32 on load.
__device__ __forceinline__ float b16_lo(unsigned int w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float b16_hi(unsigned int w) { return __uint_as_float(w & 0xffff0000u); }

template <bool BF16, int R, class XT, class YT>
__device__ __forceinline__ void lowp_row(const RowCtx &q, const int *__restrict__ ci, const v4u *__restrict__ val16,
                                         const XT *__restrict__ X, YT *__restrict__ Y) {
    constexpr int BPB = BF16 ? 4 : 2;
    constexpr int LPB = 8 / BPB;
    constexpr int KU = R <= 4 ? 2 : 1;
    const int b = q.lane / LPB, sub = q.lane % LPB;
    float acc0[R], acc1[R];
#pragma unroll
    for (int j = 0; j < R; ++j) acc0[j] = acc1[j] = 0.0f;
#pragma unroll KU
    for (int k = q.rs; k < q.re; k += BPB) {
        const int kb = k + b;
        if (kb < q.re) {
            const int c = ci[kb] - q.base;
            const v4u w = BF16 ? __builtin_nontemporal_load(&val16[(size_t)kb * LPB + sub])
                               : val16[(size_t)kb * LPB + sub];
            const XT *xc = X + q.v0 + (size_t)c * 4;
            v4f xf[R];
#pragma unroll
            for (int j = 0; j < R; ++j) xf[j] = load4f(xc + j * q.step);
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const float f0 = xf[j].x, f1 = xf[j].y, f2 = xf[j].z, f3 = xf[j].w;
                if (BF16) {
                    acc0[j] = fmaf(b16_lo(w.x), f0, acc0[j]);
                    acc0[j] = fmaf(b16_hi(w.x), f1, acc0[j]);
                    acc0[j] = fmaf(b16_lo(w.y), f2, acc0[j]);
                    acc0[j] = fmaf(b16_hi(w.y), f3, acc0[j]);
                    acc1[j] = fmaf(b16_lo(w.z), f0, acc1[j]);
                    acc1[j] = fmaf(b16_hi(w.z), f1, acc1[j]);
                    acc1[j] = fmaf(b16_lo(w.w), f2, acc1[j]);
                    acc1[j] = fmaf(b16_hi(w.w), f3, acc1[j]);
                } else {
                    acc0[j] = fmaf(__uint_as_float(w.x), f0, acc0[j]);
                    acc0[j] = fmaf(__uint_as_float(w.y), f1, acc0[j]);
                    acc0[j] = fmaf(__uint_as_float(w.z), f2, acc0[j]);
                    acc0[j] = fmaf(__uint_as_float(w.w), f3, acc0[j]);
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
#pragma unroll
        for (int off = LPB; off < 8; off <<= 1) {
            acc0[j] += __shfl_xor(acc0[j], off);
            if (BF16) acc1[j] += __shfl_xor(acc1[j], off);
        }
    }
    if (b != 0) return;
    YT *yr = Y + q.v0 + (size_t)(q.br - q.base) * 4;
#pragma unroll
    for (int j = 0; j < R; ++j) {
        if (!(q.mask & (1u << j))) continue;
        if (BF16) {
            yr[j * q.step + 2 * sub] = (YT)acc0[j];
            yr[j * q.step + 2 * sub + 1] = (YT)acc1[j];
        } else {
            yr[j * q.step + sub] = (YT)acc0[j];
        }
    }
}

// Column-split mapping: lane l of a block row's eight owns block column c = l % 4
// and the right-hand sides of parity l / 4.  It loads its column of every block
// (4 values) and one x value per right-hand side it handles, so the eight lanes
// never load the same x value twice; the four column lanes add their partial row
// sums at the end and lane c writes row c.  PREC: 0 FP64 values and accumulation,
// 1 FP32 values, 2 BF16 values (FP32 accumulation, x rounded to FP32 on load).
template <int PREC>
__device__ __forceinline__ auto load_val(const void *__restrict__ val, size_t idx) {
    if constexpr (PREC == 0) {
        return __builtin_nontemporal_load((const double *)val + idx);
    } else if constexpr (PREC == 1) {
        return ((const float *)val)[idx];
    } else {
        return __uint_as_float((unsigned)((const unsigned short *)val)[idx] << 16);
    }
}

template <int PREC, int R, class XT, class YT>
__device__ __forceinline__ void col_row(const RowCtx &q, const int *__restrict__ ci, const void *__restrict__ val,
                                        const XT *__restrict__ X, YT *__restrict__ Y) {
    using Acc = typename std::conditional<PREC == 0, double, float>::type;
    constexpr int RH = (R + 1) / 2;
    constexpr int KU = R <= 4 ? 2 : 1;
    const int c = q.lane & 3, par = q.lane >> 2;
    // the lane's right-hand sides are j = 2u + par; when R is odd the last one of
    // the odd lanes does not exist: it reloads j = R - 1 and is never written
    const bool last_ok = 2 * (RH - 1) + par < R;
    Acc acc[RH][4];
#pragma unroll
    for (int u = 0; u < RH; ++u)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[u][r] = Acc(0);
    const XT *xb = X + q.v0 + c;
#pragma unroll KU
    for (int k = q.rs; k < q.re; ++k) {
        const int col = ci[k] - q.base;
        Acc a[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) a[r] = load_val<PREC>(val, (size_t)k * 16 + r * 4 + c);
        Acc xv[RH];
#pragma unroll
        for (int u = 0; u < RH; ++u) {
            const int j = (u == RH - 1 && !last_ok) ? R - 1 : 2 * u + par;
            xv[u] = (Acc)xb[(size_t)col * 4 + j * q.step];
        }
#pragma unroll
        for (int u = 0; u < RH; ++u)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[u][r] = fma(a[r], xv[u], acc[u][r]);
    }
#pragma unroll
    for (int u = 0; u < RH; ++u)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            acc[u][r] += __shfl_xor(acc[u][r], 1);
            acc[u][r] += __shfl_xor(acc[u][r], 2);
        }
    YT *yr = Y + q.v0 + (size_t)(q.br - q.base) * 4 + c;
#pragma unroll
    for (int u = 0; u < RH; ++u) {
        const int j = 2 * u + par;
        const Acc out = c == 0 ? acc[u][0] : (c == 1 ? acc[u][1] : (c == 2 ? acc[u][2] : acc[u][3]));
        if (j < R && (q.mask & (1u << j))) yr[j * q.step] = (YT)out;
    }
}

// Whole-row mapping (FP64 values): the row-pair lane layout of f64_row, with the
// row's column indices and value lines loaded once, up front -- lane s loads the
// indices of blocks s and s + 8 (passed to the other lanes by shuffle) and every
// value line of the row -- so a lane group holds a whole row of up to KB blocks in
// flight instead of one block.  The right-hand sides are then taken four per pass
// with one gather per block: lane s (row s/2, columns 2h, 2h + 1 with h = s % 2)
// loads its half of block column c for right-hand side j0 + s/2, and takes the
// half of right-hand side j0 + u from lane 2u + h of its group by shuffle.  f64_row
// issues one gather per block and right-hand side, which the texture address path
// does not sustain (four of the eight lanes load the same 8 bytes); here each lane
// loads distinct bytes and the gathers per block fall four-fold.  Registers do not
// grow with R.  Per (row, right-hand side) the products accumulate block by block
// into the same two partial sums as f64_row, so the results equal the row-pair
// mapping's bit for bit.  X4 = false takes one right-hand side per pass instead
// (fewer registers).  Rows longer than KB take the right-hand sides one at a time,
// block by block, reloading the values per right-hand side.
template <int KB, class XT, bool X4 = true>
__device__ __forceinline__ void f64_rowc(const RowCtx &q, int R, const int *__restrict__ ci,
                                         const v2d *__restrict__ val, const XT *__restrict__ X,
                                         double *__restrict__ Y) {
    static_assert(KB <= 16, "two index loads per lane");
    using XP = typename std::conditional<sizeof(XT) == 8, v2d, v2f>::type;
    const int h = q.lane & 1;
    const int len = q.re - q.rs;
    // 32-bit element offsets from the uniform bases X and val (the launch checks
    // that every multi-vector offset and value index fits), so the gathers take
    // the scalar-base addressing form instead of a 64-bit address per load
    const unsigned xo = (unsigned)q.v0 + 2u * h, st = (unsigned)q.step;
    double *yr = Y + q.v0 + (size_t)(q.br - q.base) * 4 + (q.lane >> 1);
    if (len <= KB) {
        const int g0 = (int)(threadIdx.x & 63) & ~7;  // first lane of the group in the wave
        const int c0 = q.lane < len ? ci[q.rs + q.lane] - q.base : 0;
        const int c1 = (KB > 8 && q.lane + 8 < len) ? ci[q.rs + 8 + q.lane] - q.base : 0;
        v2d a[KB];
        const unsigned vo = (unsigned)q.rs * 8u + q.lane;
#pragma unroll
        for (int t = 0; t < KB; ++t)
            a[t] = t < len ? __builtin_nontemporal_load(&val[vo + 8u * t]) : v2d{0.0, 0.0};
        unsigned col[KB];
#pragma unroll
        for (int t = 0; t < KB; ++t) col[t] = xo + 4u * (unsigned)__shfl(t < 8 ? c0 : c1, g0 + (t & 7));
        if (!X4) {
            // one right-hand side per pass, every gather of the pass issued before its
            // products (fewer registers: the persistent kernel's form)
            for (int j = 0; j < R; ++j) {
                XP x[KB];
#pragma unroll
                for (int t = 0; t < KB; ++t) x[t] = t < len ? *(const XP *)&X[col[t] + (unsigned)j * st] : XP{0, 0};
                double acc0 = 0.0, acc1 = 0.0;
#pragma unroll
                for (int t = 0; t < KB; ++t) {
                    if (t < len) {
                        acc0 = fma(a[t].x, (double)x[t].x, acc0);
                        acc1 = fma(a[t].y, (double)x[t].y, acc1);
                    }
                }
                double sum = acc0 + acc1;
                sum += __shfl_xor(sum, 1);
                if ((q.mask & (1u << j)) && h == 0) yr[(size_t)j * q.step] = sum;
            }
            return;
        }
        for (int j0 = 0; j0 < R; j0 += 4) {
            const int jl = j0 + (q.lane >> 1);  // the right-hand side this lane gathers
            XP pc[KB];                          // held in storage precision until the products
#pragma unroll
            for (int t = 0; t < KB; ++t)
                pc[t] = (t < len && jl < R) ? *(const XP *)&X[col[t] + (unsigned)jl * st] : XP{0, 0};
            double acc0[4], acc1[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) acc0[u] = acc1[u] = 0.0;
#pragma unroll
            for (int t = 0; t < KB; ++t) {
                if (t < len) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int src = g0 + 2 * u + h;
                        const double vx = (double)__shfl(pc[t].x, src), vy = (double)__shfl(pc[t].y, src);
                        acc0[u] = fma(a[t].x, vx, acc0[u]);
                        acc1[u] = fma(a[t].y, vy, acc1[u]);
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                double sum = acc0[u] + acc1[u];
                sum += __shfl_xor(sum, 1);
                const int j = j0 + u;
                if (j < R && (q.mask & (1u << j)) && h == 0) yr[(size_t)j * q.step] = sum;
            }
        }
        return;
    }
    for (int j = 0; j < R; ++j) {
        double acc0 = 0.0, acc1 = 0.0;
        for (int k = q.rs; k < q.re; ++k) {
            const unsigned c = xo + 4u * (unsigned)(ci[k] - q.base);
            const v2d av = val[(unsigned)k * 8u + q.lane];
            const XP x = *(const XP *)&X[c + (unsigned)j * st];
            acc0 = fma(av.x, (double)x.x, acc0);
            acc1 = fma(av.y, (double)x.y, acc1);
        }
        double sum = acc0 + acc1;
        sum += __shfl_xor(sum, 1);
        if ((q.mask & (1u << j)) && h == 0) yr[(size_t)j * q.step] = sum;
    }
}

// MFMA mapping.  v_mfma_f64_4x4x4f64 computes four independent 4x4x4 products; its
// operands (measured on gfx950) sit at lane x + 4b + 16y with A[b][i=x][k=y],
// B[b][k=y][j=x] and C[b][i=y][j=x].  One wave takes four block rows (b), every
// step multiplies one block of each row by the 4x4 slice of X for four
// right-hand sides (one MFMA per group of four), so each lane loads one value of
// A and one x value per group: no value is loaded twice.  Rows shorter than the
// longest of the four, and right-hand sides beyond a row's count, feed zeros.
constexpr int kMfmaF64Rows = 4;    // block rows per wave, FP64 values
constexpr int kMfmaLowpRows = 16;  // block rows per wave, FP32 / BF16 values

struct MfmaRow {
    int br, R, rs, len, base;
    unsigned mask;
    long long v0, step;
};

__device__ __forceinline__ void mfma_row(const BatchDev &B, const int *rp, int nb, int br, int all, MfmaRow &m) {
    m.br = br;
    m.R = m.rs = m.len = m.base = 0;
    m.mask = 0;
    m.v0 = m.step = 0;
    if (br >= nb) return;
    const int sys = B.sys_of_row[br];
    m.R = B.sys_nrhs[sys];
    m.mask = active_mask(B, B.sys_pair[sys], m.R, all);
    if (!m.mask) return;
    m.base = B.sys_ptr[sys];
    m.v0 = B.sys_vec[sys];
    m.step = B.sys_step[sys];
    m.rs = rp[br];
    m.len = rp[br + 1] - m.rs;
}

// Wave unit u: block rows u * kMfmaF64Rows .. + 3 (lane = lane in the wave).
template <class XT>
__device__ __forceinline__ void mfma_f64_unit(int nb, const int *__restrict__ rp, const int *__restrict__ ci,
                                              const double *__restrict__ val, const BatchDev &B,
                                              const XT *__restrict__ X, double *__restrict__ Y, int u, int lane,
                                              int all) {
    const int x = lane & 3, b = (lane >> 2) & 3, y = lane >> 4;
    MfmaRow m;
    mfma_row(B, rp, nb, u * kMfmaF64Rows + b, all, m);
    // wave-uniform bounds: the longest row and the most right-hand-side groups
    int L = m.len, G = m.len ? (m.R + 3) >> 2 : 0;
    L = max(L, __shfl_xor(L, 4));
    L = max(L, __shfl_xor(L, 8));
    G = max(G, __shfl_xor(G, 4));
    G = max(G, __shfl_xor(G, 8));
    if (L == 0) return;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const XT *xb = X + m.v0 + y;
    // operands of step t (one block of each of the four rows); zero when past the row
    auto load_step = [&](int t, double &a, double (&bv)[4]) {
        const bool live = t < m.len;
        int c = 0;
        a = 0.0;
        if (live) {
            const int k = m.rs + t;
            a = __builtin_nontemporal_load(&val[(size_t)k * 16 + x * 4 + y]);
            c = ci[k] - m.base;
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int n = 4 * g + x;
            bv[g] = (g < G && live && n < m.R) ? (double)xb[(size_t)c * 4 + n * m.step] : 0.0;
        }
    };
    int t = 0;
    for (; t + 2 <= L; t += 2) {  // two steps' loads in flight before their MFMAs
        double a0, a1, b0[4], b1[4];
        load_step(t, a0, b0);
        load_step(t + 1, a1, b1);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            if (g < G) {
                acc[g] = __builtin_amdgcn_mfma_f64_4x4x4f64(a0, b0[g], acc[g], 0, 0, 0);
                acc[g] = __builtin_amdgcn_mfma_f64_4x4x4f64(a1, b1[g], acc[g], 0, 0, 0);
            }
        }
    }
    if (t < L) {
        double a0, b0[4];
        load_step(t, a0, b0);
#pragma unroll
        for (int g = 0; g < 4; ++g)
            if (g < G) acc[g] = __builtin_amdgcn_mfma_f64_4x4x4f64(a0, b0[g], acc[g], 0, 0, 0);
    }
    if (!m.len) return;
    double *yr = Y + m.v0 + (size_t)(m.br - m.base) * 4 + y;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const int n = 4 * g + x;
        if (g < G && n < m.R && (m.mask & (1u << n))) yr[n * m.step] = acc[g];
    }
}

template <class XT>
__global__ __launch_bounds__(kSpmmMaxWg) void bspmm_mfma_f64_kernel(int nb, const int *__restrict__ rp,
                                                                    const int *__restrict__ ci,
                                                                    const double *__restrict__ val, BatchDev B,
                                                                    const XT *__restrict__ X,
                                                                    double *__restrict__ Y, int remap, int all) {
    const int wg = remap ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    mfma_f64_unit<XT>(nb, rp, ci, val, B, X, Y, wg * (int)(blockDim.x / 64) + (int)(threadIdx.x >> 6),
                      (int)(threadIdx.x & 63), all);
}

// v_mfma_f32_4x4x1f32 (sixteen 4x4x1 products): A[b][i] at lane i + 4b, B[b][j] at
// lane j + 4b, C[b][i][j] in element i of lane j + 4b.  One wave takes sixteen
// block rows; lane (i, b) loads row i of its block (16 B FP32, 8 B BF16) and, as a
// B-operand lane (j, b), the four x values of right-hand side j of each group
// (rounded to FP32); four K = 1 MFMAs per group and block, FP32 accumulation.
// Wave unit u: block rows u * kMfmaLowpRows .. + 15.
template <bool BF16, class XT, class YT>
__device__ __forceinline__ void mfma_lowp_unit(int nb, const int *__restrict__ rp, const int *__restrict__ ci,
                                               const void *__restrict__ val, const BatchDev &B,
                                               const XT *__restrict__ X, YT *__restrict__ Y, int u, int lane,
                                               int all) {
    const int q = lane & 3, b = lane >> 2;  // q: row i for A, right-hand side j for B and C
    MfmaRow m;
    mfma_row(B, rp, nb, u * kMfmaLowpRows + b, all, m);
    int L = m.len, G = m.len ? (m.R + 3) >> 2 : 0;
#pragma unroll
    for (int off = 4; off < 64; off <<= 1) {
        L = max(L, __shfl_xor(L, off));
        G = max(G, __shfl_xor(G, off));
    }
    if (L == 0) return;
    v4f acc[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) acc[g] = v4f{0.f, 0.f, 0.f, 0.f};
    // operands of step t: row q of each block (A) and the four x values of right-hand
    // side q of each group (B), zero when past the row
    auto load_step = [&](int t, float (&a)[4], float (&xv)[4][4]) {
        const bool live = t < m.len;
        int c = 0;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) a[kk] = 0.f;
        if (live) {
            const int k = m.rs + t;
            c = ci[k] - m.base;
            if (BF16) {
                const uint2 w = ((const uint2 *)val)[(size_t)k * 4 + q];
                a[0] = __uint_as_float(w.x << 16);
                a[1] = __uint_as_float(w.x & 0xffff0000u);
                a[2] = __uint_as_float(w.y << 16);
                a[3] = __uint_as_float(w.y & 0xffff0000u);
            } else {
                const v4f w = ((const v4f *)val)[(size_t)k * 4 + q];
                a[0] = w.x;
                a[1] = w.y;
                a[2] = w.z;
                a[3] = w.w;
            }
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int n = 4 * g + q;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) xv[g][kk] = 0.f;
            if (g < G && live && n < m.R) {
                const v4f xq = load4f(X + m.v0 + n * m.step + (size_t)c * 4);
                xv[g][0] = xq.x;
                xv[g][1] = xq.y;
                xv[g][2] = xq.z;
                xv[g][3] = xq.w;
            }
        }
    };
    auto mfma_step = [&](const float (&a)[4], const float (&xv)[4][4]) {
#pragma unroll
        for (int g = 0; g < 4; ++g)
            if (g < G)
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) acc[g] = __builtin_amdgcn_mfma_f32_4x4x1f32(a[kk], xv[g][kk], acc[g], 0, 0, 0);
    };
    int t = 0;
    for (; t + 2 <= L; t += 2) {  // two steps' loads in flight before their MFMAs
        float a0[4], a1[4], x0[4][4], x1[4][4];
        load_step(t, a0, x0);
        load_step(t + 1, a1, x1);
        mfma_step(a0, x0);
        mfma_step(a1, x1);
    }
    if (t < L) {
        float a0[4], x0[4][4];
        load_step(t, a0, x0);
        mfma_step(a0, x0);
    }
    if (!m.len) return;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const int n = 4 * g + q;
        if (g < G && n < m.R && (m.mask & (1u << n))) store4f(Y + m.v0 + n * m.step + (size_t)(m.br - m.base) * 4, acc[g]);
    }
}

// Whole-row form of mfma_lowp_unit for rows of up to KB blocks (wave-wide): lane
// (i, b) loads the column indices of blocks i and i + 4 of its row (passed to the
// quad by shuffle) and row i of every block up front, then one pass per group of
// four right-hand sides issues all of the group's x loads before its MFMAs.  The
// MFMA chain of each group runs over the same steps in the same order (zero
// operands past a row's end) as mfma_lowp_unit, so the results are equal bit for
// bit; longer rows take mfma_lowp_unit.
template <bool BF16, int KB, class XT, class YT>
__device__ __forceinline__ void mfma_lowp_unit_wr(int nb, const int *__restrict__ rp, const int *__restrict__ ci,
                                                  const void *__restrict__ val, const BatchDev &B,
                                                  const XT *__restrict__ X, YT *__restrict__ Y, int u, int lane,
                                                  int all) {
    static_assert(KB <= 8, "two index loads per lane");
    const int q = lane & 3, b = lane >> 2;
    MfmaRow m;
    mfma_row(B, rp, nb, u * kMfmaLowpRows + b, all, m);
    int L = m.len, G = m.len ? (m.R + 3) >> 2 : 0;
#pragma unroll
    for (int off = 4; off < 64; off <<= 1) {
        L = max(L, __shfl_xor(L, off));
        G = max(G, __shfl_xor(G, off));
    }
    if (L == 0) return;
    if (L > KB) {
        mfma_lowp_unit<BF16, XT, YT>(nb, rp, ci, val, B, X, Y, u, lane, all);
        return;
    }
    const int c0 = q < m.len ? ci[m.rs + q] - m.base : 0;
    const int c1 = (KB > 4 && q + 4 < m.len) ? ci[m.rs + 4 + q] - m.base : 0;
    float a[KB][4];
#pragma unroll
    for (int t = 0; t < KB; ++t) {
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) a[t][kk] = 0.f;
        if (t < m.len) {
            const int k = m.rs + t;
            if (BF16) {
                const uint2 w = ((const uint2 *)val)[(size_t)k * 4 + q];
                a[t][0] = __uint_as_float(w.x << 16);
                a[t][1] = __uint_as_float(w.x & 0xffff0000u);
                a[t][2] = __uint_as_float(w.y << 16);
                a[t][3] = __uint_as_float(w.y & 0xffff0000u);
            } else {
                const v4f w = ((const v4f *)val)[(size_t)k * 4 + q];
                a[t][0] = w.x;
                a[t][1] = w.y;
                a[t][2] = w.z;
                a[t][3] = w.w;
            }
        }
    }
    int col[KB];
#pragma unroll
    for (int t = 0; t < KB; ++t) col[t] = __shfl(t < 4 ? c0 : c1, (lane & ~3) + (t & 3)) * 4;
    v4f acc[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) acc[g] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        if (g >= G) break;
        const int n = 4 * g + q;
        const XT *xn = X + m.v0 + n * m.step;
        v4f xv[KB];
#pragma unroll
        for (int t = 0; t < KB; ++t)
            xv[t] = (t < m.len && n < m.R) ? load4f(xn + col[t]) : v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t = 0; t < KB; ++t) {
            if (t < L) {
                acc[g] = __builtin_amdgcn_mfma_f32_4x4x1f32(a[t][0], xv[t].x, acc[g], 0, 0, 0);
                acc[g] = __builtin_amdgcn_mfma_f32_4x4x1f32(a[t][1], xv[t].y, acc[g], 0, 0, 0);
                acc[g] = __builtin_amdgcn_mfma_f32_4x4x1f32(a[t][2], xv[t].z, acc[g], 0, 0, 0);
                acc[g] = __builtin_amdgcn_mfma_f32_4x4x1f32(a[t][3], xv[t].w, acc[g], 0, 0, 0);
            }
        }
    }
    if (!m.len) return;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const int n = 4 * g + q;
        if (g < G && n < m.R && (m.mask & (1u << n))) store4f(Y + m.v0 + n * m.step + (size_t)(m.br - m.base) * 4, acc[g]);
    }
}

template <bool BF16, int KB, class XT, class YT>
__global__ __launch_bounds__(kSpmmMaxWg) void bspmm_mfma_lowp_wr_kernel(int nb, const int *__restrict__ rp,
                                                                        const int *__restrict__ ci,
                                                                        const void *__restrict__ val, BatchDev B,
                                                                        const XT *__restrict__ X,
                                                                        YT *__restrict__ Y, int remap, int all) {
    const int wg = remap ? xcd_chunk_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    mfma_lowp_unit_wr<BF16, KB, XT, YT>(nb, rp, ci, val, B, X, Y,
                                        wg * (int)(blockDim.x / 64) + (int)(threadIdx.x >> 6), (int)(threadIdx.x & 63),
                                        all);
}

template <bool BF16, class XT, class YT>
__global__ __launch_bounds__(kSpmmMaxWg) void bspmm_mfma_lowp_kernel(int nb, const int *__restrict__ rp,
                                                                     const int *__restrict__ ci,
                                                                     const void *__restrict__ val, BatchDev B,
                                                                     const XT *__restrict__ X,
                                                                     YT *__restrict__ Y, int remap, int all) {
    const int wg = remap ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    mfma_lowp_unit<BF16, XT, YT>(nb, rp, ci, val, B, X, Y, wg * (int)(blockDim.x / 64) + (int)(threadIdx.x >> 6),
                                 (int)(threadIdx.x & 63), all);
}

#define ISAI_RHS_CASES(CALL) \
    CALL(1) CALL(2) CALL(3) CALL(4) CALL(5) CALL(6) CALL(7) CALL(8) \
    CALL(9) CALL(10) CALL(11) CALL(12) CALL(13) CALL(14) CALL(15) CALL(16)

// MAP 0: row-pair mapping (f64_row / lowp_row); MAP 1: column-split mapping.
template <int MAP, class XT>
__device__ __forceinline__ void f64_unit(int nb, const int *__restrict__ rp, const int *__restrict__ ci,
                                         const v2d *__restrict__ val, const BatchDev &B, const XT *__restrict__ X,
                                         double *__restrict__ Y, int br, int lane, int all) {
    RowCtx q;
    int r;
    if (!row_context(B, rp, nb, br, lane, all, q, r)) return;
    switch (r) {
#define ISAI_F64_CASE(N)                              \
    case N:                                           \
        if (MAP == 0) f64_row<N, XT>(q, ci, val, X, Y); \
        else col_row<0, N>(q, ci, val, X, Y);         \
        break;
        ISAI_RHS_CASES(ISAI_F64_CASE)
#undef ISAI_F64_CASE
        default:
            break;
    }
}

template <int MAP, class XT>
__global__ __launch_bounds__(kSpmmMaxWg) void bspmm_f64_kernel(int nb, const int *__restrict__ rp,
                                                               const int *__restrict__ ci, const v2d *__restrict__ val,
                                                               BatchDev B, const XT *__restrict__ X,
                                                               double *__restrict__ Y, int remap, int all) {
    const int wg = remap ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    f64_unit<MAP, XT>(nb, rp, ci, val, B, X, Y, wg * (int)(blockDim.x / 8) + (int)(threadIdx.x >> 3),
                      (int)(threadIdx.x & 7), all);
}

// MAP 3: whole-row mapping, KB blocks up front.
template <int KB, class XT, bool X4 = true>
__device__ __forceinline__ void f64c_unit(int nb, const int *__restrict__ rp, const int *__restrict__ ci,
                                          const v2d *__restrict__ val, const BatchDev &B, const XT *__restrict__ X,
                                          double *__restrict__ Y, int br, int lane, int all) {
    RowCtx q;
    int r;
    if (!row_context(B, rp, nb, br, lane, all, q, r)) return;
    f64_rowc<KB, XT, X4>(q, r, ci, val, X, Y);
}

template <int KB, class XT>
__global__ __launch_bounds__(kSpmmMaxWg) void bspmm_f64c_kernel(int nb, const int *__restrict__ rp,
                                                                const int *__restrict__ ci,
                                                                const v2d *__restrict__ val, BatchDev B,
                                                                const XT *__restrict__ X, double *__restrict__ Y,
                                                                int remap, int all) {
    const int wg = remap ? xcd_chunk_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    f64c_unit<KB, XT>(nb, rp, ci, val, B, X, Y, wg * (int)(blockDim.x / 8) + (int)(threadIdx.x >> 3),
                         (int)(threadIdx.x & 7), all);
}

template <bool BF16, int MAP, class XT, class YT>
__device__ __forceinline__ void lowp_unit(int nb, const int *__restrict__ rp, const int *__restrict__ ci,
                                          const v4u *__restrict__ val16, const BatchDev &B, const XT *__restrict__ X,
                                          YT *__restrict__ Y, int br, int lane, int all) {
    RowCtx q;
    int r;
    if (!row_context(B, rp, nb, br, lane, all, q, r)) return;
    switch (r) {
#define ISAI_LOWP_CASE(N)                                   \
    case N:                                                 \
        if (MAP == 0) lowp_row<BF16, N>(q, ci, val16, X, Y); \
        else col_row<BF16 ? 2 : 1, N>(q, ci, val16, X, Y);  \
        break;
        ISAI_RHS_CASES(ISAI_LOWP_CASE)
#undef ISAI_LOWP_CASE
        default:
            break;
    }
}

template <bool BF16, int MAP, class XT, class YT>
__global__ __launch_bounds__(kSpmmMaxWg) void bspmm_lowp_kernel(int nb, const int *__restrict__ rp,
                                                                const int *__restrict__ ci,
                                                                const v4u *__restrict__ val16, BatchDev B,
                                                                const XT *__restrict__ X, YT *__restrict__ Y,
                                                                int remap, int all) {
    const int wg = remap ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    lowp_unit<BF16, MAP, XT, YT>(nb, rp, ci, val16, B, X, Y, wg * (int)(blockDim.x / 8) + (int)(threadIdx.x >> 3),
                                 (int)(threadIdx.x & 7), all);
}

// --------------------------------------------------------------- reductions
constexpr int kRed = 64;               // threads per tile: one wave
constexpr int kTileWg = 256;           // per-step launches: four tiles per workgroup
constexpr int kPersistThreads = 1024;  // persistent kernel: sixteen one-wave tile groups

// Sum of v over the wave that holds one tile, by a fixed butterfly (every lane gets
// the same total): a tile's partial sums need no workgroup barrier, so a workgroup's
// waves take their tiles independently.
__device__ __forceinline__ double tile_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Tile of a thread in the per-step launches.
__device__ __forceinline__ int launch_tile() { return (int)(blockIdx.x * (kTileWg / kRed) + (threadIdx.x >> 6)); }
inline unsigned tile_grid(int ntiles) { return (unsigned)((ntiles + kTileWg / kRed - 1) / (kTileWg / kRed)); }

// The per-pair finals run one wave per pair: the wave sums the pair's tile partials
// lane-strided, then by a fixed butterfly, so every lane holds the same,
// deterministic totals; lane 0 updates the pair's state.
__device__ __forceinline__ void pair_sums(const BatchDev &B, int p, int lane, double &s0, double &s1) {
    double a = 0.0, c = 0.0;
    for (int t = B.ptile[p] + lane; t < B.ptile[p + 1]; t += 64) {
        a += B.part0[t];
        c += B.part1[t];
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        a += __shfl_xor(a, off);
        c += __shfl_xor(c, off);
    }
    s0 = a;
    s1 = c;
}

// Tile t of a running pair (t may lie past the last tile: the per-step launches
// round the tile count up to whole workgroups).
__device__ __forceinline__ bool tile_live(const BatchDev &B, int t) {
    return t < B.ntiles && B.ps[B.tile_pair[t]].status == 0;
}

// part0[t] = a.b and part1[t] = a.c (c optional) over tile t, for running pairs.
__device__ __forceinline__ void tile_dot_body(int t, const double *__restrict__ a,
                                              const double *__restrict__ b, const double *__restrict__ c,
                                              const BatchDev &B) {
    const bool live = tile_live(B, t);  // the pair's finals skip it too
    const int tid = threadIdx.x & (kRed - 1);
    const size_t base = (size_t)t * kTile;
    double s0 = 0.0, s1 = 0.0;
    if (live) {
#pragma unroll
        for (int k = 0; k < kTile / kRed; ++k) {
            const size_t i = base + k * kRed + tid;
            s0 = fma(a[i], b[i], s0);
            if (c) s1 = fma(a[i], c[i], s1);
        }
    }
    s0 = tile_sum(s0);
    if (c) s1 = tile_sum(s1);
    if (live && tid == 0) {
        B.part0[t] = s0;
        B.part1[t] = s1;
    }
}

// ------------------------------------------------------------ BiCGSTAB kernels
// Right-preconditioned BiCGSTAB (Saad, Iterative Methods, Alg. 7.7 with M = M_U M_L),
// the same recurrences as the single-system solver, with one set of scalars per pair.

// r = b - r (r holds A x), rhat = r, p = v = 0 for running pairs; partials ||b||^2, ||r||^2
__device__ __forceinline__ void binit_body(int t, const double *__restrict__ b, double *__restrict__ r,
                                           double *__restrict__ rh, double *__restrict__ p, double *__restrict__ v,
                                           const BatchDev &B) {
    const bool live = tile_live(B, t);
    const int tid = threadIdx.x & (kRed - 1);
    const size_t base = (size_t)t * kTile;
    double sb = 0.0, sr = 0.0;
    if (live) {
#pragma unroll
        for (int k = 0; k < kTile / kRed; ++k) {
            const size_t i = base + k * kRed + tid;
            const double bi = b[i], ri = bi - r[i];
            r[i] = ri;
            rh[i] = ri;
            p[i] = 0.0;
            v[i] = 0.0;
            sb = fma(bi, bi, sb);
            sr = fma(ri, ri, sr);
        }
    }
    sb = tile_sum(sb);
    sr = tile_sum(sr);
    if (live && tid == 0) {
        B.part0[t] = sb;
        B.part1[t] = sr;
    }
}

__device__ __forceinline__ void binit_final_body(const BatchDev &B, int p, int lane, double tol) {
    PairState &S = B.ps[p];
    if (S.status != 0) return;
    double b2, r2;
    pair_sums(B, p, lane, b2, r2);
    if (lane != 0) return;
    S.bnorm2 = b2 > 0.0 ? b2 : 1.0;
    S.tol2 = tol * tol * S.bnorm2;
    S.rho_old = S.alpha = S.omega = 1.0;
    S.first = 1;
    if (r2 <= S.tol2) S.status = 1;
}

__device__ __forceinline__ void brho_final_body(const BatchDev &B, int p, int lane) {
    PairState &S = B.ps[p];
    if (S.status != 0) return;
    if (S.iters >= B.ctl->maxit) {
        if (lane == 0) S.status = 5;
        return;
    }
    double unused, rho;
    pair_sums(B, p, lane, unused, rho);  // part1 = rhat.r (from the r update, or ||r||^2 when rhat = r)
    if (lane != 0) return;
    S.iters++;
    S.rho = rho;
    if (rho == 0.0) S.status = 3;
    else S.beta = S.first ? 0.0 : (rho / S.rho_old) * (S.alpha / S.omega);
}

// p = r + beta (p - omega v)  (p = r on a pair's first iteration); with FP32 /
// BF16 factors also p32 = p rounded to FP32, the operand M_L reads (its products
// round x to FP32 on load, so the copy gives the same results with half the
// gathered bytes)
__device__ __forceinline__ void bp_body(int t, const double *__restrict__ r, double *__restrict__ p,
                                        const double *__restrict__ v, float *__restrict__ p32, const BatchDev &B) {
    if (!tile_live(B, t)) return;
    const PairState &S = B.ps[B.tile_pair[t]];
    const int tid = threadIdx.x & (kRed - 1);
    const size_t base = (size_t)t * kTile;
    const double beta = S.beta, omega = S.omega;
    const bool first = S.first;
    // every load of a half tile before its stores (the stores to p otherwise kept the
    // next element's loads waiting)
    constexpr int kHalf = kTile / kRed / 2;
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
        double rr[kHalf], pp[kHalf], vv[kHalf];
#pragma unroll
        for (int k = 0; k < kHalf; ++k) {
            const size_t i = base + (hh * kHalf + k) * kRed + tid;
            rr[k] = r[i];
            pp[k] = first ? 0.0 : p[i];
            vv[k] = first ? 0.0 : v[i];
        }
#pragma unroll
        for (int k = 0; k < kHalf; ++k) {
            const size_t i = base + (hh * kHalf + k) * kRed + tid;
            const double pk = first ? rr[k] : rr[k] + beta * (pp[k] - omega * vv[k]);
            p[i] = pk;
            if (p32) p32[i] = (float)pk;
        }
    }
}

__device__ __forceinline__ void balpha_final_body(const BatchDev &B, int p, int lane) {
    PairState &S = B.ps[p];
    if (S.status != 0) return;
    double d, unused;
    pair_sums(B, p, lane, d, unused);
    if (lane != 0) return;
    if (d == 0.0) S.status = 3;
    else S.alpha = S.rho / d;
    S.first = 0;
}

// s = r - alpha v (and its FP32 copy s32 with FP32 / BF16 factors, as in bp_body);
// partial ||s||^2
__device__ __forceinline__ void bs_body(int t, const double *__restrict__ r, const double *__restrict__ v,
                                        double *__restrict__ s, float *__restrict__ s32, const BatchDev &B) {
    const bool live = tile_live(B, t);
    const int tid = threadIdx.x & (kRed - 1);
    const size_t base = (size_t)t * kTile;
    double acc = 0.0;
    if (live) {
        const double alpha = B.ps[B.tile_pair[t]].alpha;
#pragma unroll
        for (int k = 0; k < kTile / kRed; ++k) {
            const size_t i = base + k * kRed + tid;
            const double si = r[i] - alpha * v[i];
            s[i] = si;
            if (s32) s32[i] = (float)si;
            acc = fma(si, si, acc);
        }
    }
    acc = tile_sum(acc);
    if (live && tid == 0) B.part0[t] = acc;
}

__device__ __forceinline__ void bsnorm_final_body(const BatchDev &B, int p, int lane) {
    PairState &S = B.ps[p];
    if (S.status != 0) return;
    double s2, unused;
    pair_sums(B, p, lane, s2, unused);
    if (lane == 0 && s2 <= S.tol2) S.status = 2;
}

// pairs converged on s: x += alpha ph
template <class PT>
__device__ __forceinline__ void bxs_body(int t, double *__restrict__ x, const PT *__restrict__ ph, const BatchDev &B) {
    if (t >= B.ntiles) return;
    const PairState &S = B.ps[B.tile_pair[t]];
    if (S.status != 2) return;
    const double alpha = S.alpha;
    const int tid = threadIdx.x & (kRed - 1);
    const size_t base = (size_t)t * kTile;
#pragma unroll
    for (int k = 0; k < kTile / kRed; ++k) {
        const size_t i = base + k * kRed + tid;
        x[i] += alpha * (double)ph[i];
    }
}

__device__ __forceinline__ void bomega_final_body(const BatchDev &B, int p, int lane) {
    PairState &S = B.ps[p];
    if (S.status != 0) return;
    double ts, tt;
    pair_sums(B, p, lane, ts, tt);  // part0 = t.s, part1 = t.t
    if (lane != 0) return;
    if (tt == 0.0) S.status = 3;
    else S.omega = ts / tt;
}

// x += alpha ph + omega sh; r = s - omega t; partials ||r||^2 and rhat.r (the next
// iteration's rho)
template <class PT>
__device__ __forceinline__ void bxr_body(int t, double *__restrict__ x, double *__restrict__ r,
                                         const double *__restrict__ rh,
                                         const PT *__restrict__ ph, const PT *__restrict__ sh,
                                         const double *__restrict__ s, const double *__restrict__ tv,
                                         const BatchDev &B) {
    const bool live = tile_live(B, t);
    const int tid = threadIdx.x & (kRed - 1);
    const size_t base = (size_t)t * kTile;
    double acc = 0.0, rho = 0.0;
    if (live) {
        const PairState &S = B.ps[B.tile_pair[t]];
        const double alpha = S.alpha, omega = S.omega;
#pragma unroll
        for (int k = 0; k < kTile / kRed; ++k) {
            const size_t i = base + k * kRed + tid;
            x[i] += alpha * (double)ph[i] + omega * (double)sh[i];
            const double ri = s[i] - omega * tv[i];
            r[i] = ri;
            acc = fma(ri, ri, acc);
            rho = fma(rh[i], ri, rho);
        }
    }
    acc = tile_sum(acc);
    rho = tile_sum(rho);
    if (live && tid == 0) {
        B.part0[t] = acc;
        B.part1[t] = rho;
    }
}

// Pairs still running after the update add one to nactive.
__device__ __forceinline__ void brnorm_final_body(const BatchDev &B, int p, int lane) {
    PairState &S = B.ps[p];
    const int st = S.status;
    if (st == 2 && lane == 0) S.status = 1;
    if (st != 0) return;
    double r2, unused;
    pair_sums(B, p, lane, r2, unused);
    if (lane != 0) return;
    S.rho_old = S.rho;
    if (r2 <= S.tol2) S.status = 1;
    else atomicAdd(&B.ctl->nactive, 1);
}

// true residual of every pair: r = b - r (r holds A x); part0 = ||r||^2, part1 = ||b||^2
__device__ __forceinline__ void bres_body(int t, const double *__restrict__ b, double *__restrict__ r,
                                          const BatchDev &B) {
    const bool live = t < B.ntiles;
    const int tid = threadIdx.x & (kRed - 1);
    const size_t base = (size_t)t * kTile;
    double sr = 0.0, sb = 0.0;
    if (live) {
#pragma unroll
        for (int k = 0; k < kTile / kRed; ++k) {
            const size_t i = base + k * kRed + tid;
            const double bi = b[i], ri = bi - r[i];
            r[i] = ri;
            sr = fma(ri, ri, sr);
            sb = fma(bi, bi, sb);
        }
    }
    sr = tile_sum(sr);
    sb = tile_sum(sb);
    if (live && tid == 0) {
        B.part0[t] = sr;
        B.part1[t] = sb;
    }
}

// relres per pair; pairs whose true residual misses tol restart while iterations
// remain (each such pair adds one to nrestart)
__device__ __forceinline__ void bres_final_body(const BatchDev &B, int p, int lane, double tol, double *relres) {
    double r2, b2;
    pair_sums(B, p, lane, r2, b2);
    if (lane != 0) return;
    const double rr = sqrt(r2) / sqrt(b2 > 0.0 ? b2 : 1.0);
    relres[p] = rr;
    PairState &S = B.ps[p];
    if (rr <= tol) {
        S.status = 4;
    } else if (S.iters < B.ctl->maxit) {
        S.status = 0;
        atomicAdd(&B.ctl->nrestart, 1);
    } else {
        S.status = 5;
    }
}

__device__ __forceinline__ void breset_body(const BatchDev &B, int p) {
    B.ps[p].status = 0;
    B.ps[p].iters = 0;
}

// ---- one launch per step (the eager and graph-replayed paths)
__global__ void zero_active_kernel(Ctl *ctl) { ctl->nactive = 0; }

__global__ __launch_bounds__(kTileWg) void tile_dot_kernel(const double *__restrict__ a, const double *__restrict__ b,
                                                        const double *__restrict__ c, BatchDev B) {
    tile_dot_body(launch_tile(), a, b, c, B);
}
__global__ __launch_bounds__(kTileWg) void binit_kernel(const double *__restrict__ b, double *r, double *rh, double *p,
                                                     double *v, BatchDev B) {
    binit_body(launch_tile(), b, r, rh, p, v, B);
}
__global__ __launch_bounds__(kTileWg) void bp_kernel(const double *__restrict__ r, double *__restrict__ p,
                                                  const double *__restrict__ v, float *p32, BatchDev B) {
    bp_body(launch_tile(), r, p, v, p32, B);
}
__global__ __launch_bounds__(kTileWg) void bs_kernel(const double *__restrict__ r, const double *__restrict__ v, double *s,
                                                  float *s32, BatchDev B) {
    bs_body(launch_tile(), r, v, s, s32, B);
}
template <class PT>
__global__ __launch_bounds__(kTileWg) void bxs_kernel(double *x, const PT *__restrict__ ph, BatchDev B) {
    bxs_body(launch_tile(), x, ph, B);
}
template <class PT>
__global__ __launch_bounds__(kTileWg) void bxr_kernel(double *x, double *r, const double *__restrict__ rh,
                                                   const PT *__restrict__ ph, const PT *__restrict__ sh,
                                                   const double *__restrict__ s, const double *__restrict__ tv,
                                                   BatchDev B) {
    bxr_body(launch_tile(), x, r, rh, ph, sh, s, tv, B);
}
__global__ __launch_bounds__(kTileWg) void bres_kernel(const double *__restrict__ b, double *r, BatchDev B) {
    bres_body(launch_tile(), b, r, B);
}
__global__ __launch_bounds__(64) void binit_final(BatchDev B, double tol) {
    binit_final_body(B, blockIdx.x, threadIdx.x, tol);
}
__global__ __launch_bounds__(64) void brho_final(BatchDev B) { brho_final_body(B, blockIdx.x, threadIdx.x); }
__global__ __launch_bounds__(64) void balpha_final(BatchDev B) { balpha_final_body(B, blockIdx.x, threadIdx.x); }
__global__ __launch_bounds__(64) void bsnorm_final(BatchDev B) { bsnorm_final_body(B, blockIdx.x, threadIdx.x); }
__global__ __launch_bounds__(64) void bomega_final(BatchDev B) { bomega_final_body(B, blockIdx.x, threadIdx.x); }
__global__ __launch_bounds__(64) void brnorm_final(BatchDev B) { brnorm_final_body(B, blockIdx.x, threadIdx.x); }
__global__ __launch_bounds__(64) void bres_final(BatchDev B, double tol, double *relres) {
    bres_final_body(B, blockIdx.x, threadIdx.x, tol, relres);
}
__global__ void breset_kernel(BatchDev B, int npairs) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < npairs) breset_body(B, p);
}

// ------------------------------------------------------- persistent kernel
// The whole batched solve in one launch: every step above becomes a phase of a
// grid-stride loop, with a grid barrier between phases.  The grid is one
// workgroup of kPersistThreads per CU, all resident (checked at launch), so the
// barrier waits only on workgroups that are running; each wait is still bounded
// (Ctl::spin_us) and an abandoned wait makes every workgroup leave, reported as
// ISAI_ERR_TIMEOUT.  Phases call the same step bodies as the per-step launches,
// so both paths compute bit-identical iterates.
//
// Grid barrier: arrivals on one counter (agent-scope release before, acquire
// after, so the phase's stores are visible to every XCD's next phase), the
// target moving up by gridDim.x per barrier.
__device__ __forceinline__ bool grid_barrier(Ctl *c, unsigned long long &target) {
    __shared__ int s_ok;
    __syncthreads();
    target += gridDim.x;
    if (threadIdx.x == 0) {
        __threadfence();
        atomicAdd(&c->arrive, 1ULL);
        int ok = 1;
        const long long limit = c->spin_us * 100;  // s_memrealtime ticks at 100 MHz
        const long long t0 = (long long)__builtin_amdgcn_s_memrealtime();
        for (;;) {
            if (__hip_atomic_load(&c->arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) break;
            if (__hip_atomic_load(&c->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                ok = 0;
                break;
            }
            if ((long long)__builtin_amdgcn_s_memrealtime() - t0 > limit) {
                __hip_atomic_store(&c->abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __threadfence();
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

struct PersistArgs {
    BatchDev B;
    int nb;
    const int *rp, *ci;
    const double *val;                 // A
    const int *Lrow, *Lcol, *Urow, *Ucol;
    const void *Lval, *Uval;           // ISAI factors in storage precision
    const double *b;                   // right-hand sides
    double *x;                         // solutions (initial guesses in)
    double *r, *rh, *p, *v, *s, *t;
    void *ph, *sh, *tmp;               // FP32 with FP32 / BF16 factors (exact copies of their results)
    size_t n;                          // multi-vector length
    double tol;
    double *relres;
};

enum Operand { kOpA = 0, kOpML = 1, kOpMU = 2 };

// Y = op X over the pairs still running (all = 1: every pair); MAP as the
// autotuner's mapping (0 row-pair, 1 column-split, 2 MFMA, 3 whole-row with KB
// blocks per row), PREC the operand's value precision.  Units: 8-lane groups (one
// block row) for MAP 0/1/3, waves of kMfmaF64Rows / kMfmaLowpRows block rows for
// MAP 2.
template <int MAP, int PREC, int KB = 8, class XT, class YT>
__device__ __forceinline__ void spmm_phase(int nb, const int *rp, const int *ci, const void *val,
                                           const BatchDev &B, const XT *X, YT *Y, int all) {
    const int lane = threadIdx.x & 63;
    if constexpr (MAP == 2) {
        const int rows = PREC == ISAI_PREC_FP64 ? kMfmaF64Rows : kMfmaLowpRows;
        const int units = (nb + rows - 1) / rows;
        const int wpg = kPersistThreads / 64;
        for (int u = blockIdx.x * wpg + (int)(threadIdx.x >> 6); u < units; u += gridDim.x * wpg) {
            if constexpr (PREC == ISAI_PREC_FP64)
                mfma_f64_unit<XT>(nb, rp, ci, (const double *)val, B, X, Y, u, lane, all);
            else
                mfma_lowp_unit<PREC == ISAI_PREC_BF16, XT, YT>(nb, rp, ci, val, B, X, Y, u, lane, all);
        }
    } else {
        const int gpg = kPersistThreads / 8;
        for (int br = blockIdx.x * gpg + (int)(threadIdx.x >> 3); br < nb; br += gridDim.x * gpg) {
            if constexpr (PREC == ISAI_PREC_FP64 && MAP == 3)
                f64c_unit<KB, XT, false>(nb, rp, ci, (const v2d *)val, B, X, Y, br, lane & 7, all);
            else if constexpr (PREC == ISAI_PREC_FP64)
                f64_unit<MAP, XT>(nb, rp, ci, (const v2d *)val, B, X, Y, br, lane & 7, all);
            else
                lowp_unit<PREC == ISAI_PREC_BF16, MAP, XT, YT>(nb, rp, ci, (const v4u *)val, B, X, Y, br, lane & 7,
                                                               all);
        }
    }
}

template <int MAPA, int MAPM, int PRECM, int KBA = 8>
__global__ __launch_bounds__(kPersistThreads) void batch_persistent_kernel(PersistArgs a) {
    const BatchDev &B = a.B;
    const int sub = threadIdx.x / kRed, lane = threadIdx.x & 63;
    const int wpg = kPersistThreads / 64, tpg = kPersistThreads / kRed;
    Ctl *ctl = B.ctl;
    unsigned long long target = 0;
    using IT = typename std::conditional<PRECM == ISAI_PREC_FP64, double, float>::type;  // ISAI outputs
    IT *const ph = (IT *)a.ph, *const sh = (IT *)a.sh, *const tmp = (IT *)a.tmp;
    // FP32 copy of p / s for M_L (second half of tmp's region), FP32 / BF16 factors only
    float *const x32 = PRECM == ISAI_PREC_FP64 ? nullptr : (float *)a.tmp + a.n;
    // tile phases: wave `sub` of the workgroup takes tiles t0 + sub (a tile's sums
    // stay inside its wave, so the waves need not step together)
    auto tiles_end = [&](int t0) { return t0 + sub < B.ntiles; };
    auto A = [&](const auto *X, double *Y, int all) {
        spmm_phase<MAPA, ISAI_PREC_FP64, KBA>(a.nb, a.rp, a.ci, a.val, B, X, Y, all);
    };
    auto ML = [&](const double *X, IT *Y) {
        // FP32 / BF16 factors: M_L reads the FP32 copy of p or s (x32, written by
        // the update that formed X), as the per-step launches do
        if constexpr (PRECM == ISAI_PREC_FP64)
            spmm_phase<MAPM, PRECM>(a.nb, a.Lrow, a.Lcol, a.Lval, B, X, Y, 0);
        else
            spmm_phase<MAPM, PRECM>(a.nb, a.Lrow, a.Lcol, a.Lval, B, (const float *)x32, Y, 0);
    };
    auto MU = [&](const IT *X, IT *Y) { spmm_phase<MAPM, PRECM>(a.nb, a.Urow, a.Ucol, a.Uval, B, X, Y, 0); };
#define ISAI_SYNC() \
    if (!grid_barrier(ctl, target)) return
#define ISAI_TILES(BODY) \
    for (int t0 = blockIdx.x * tpg; tiles_end(t0); t0 += gridDim.x * tpg) BODY
#define ISAI_PAIRS(BODY) \
    for (int p = blockIdx.x * wpg + (int)(threadIdx.x >> 6); p < B.npairs; p += gridDim.x * wpg) BODY

    ISAI_PAIRS(breset_body(B, p));
    ISAI_SYNC();
    int nrestart = 0;
    for (int round = 0; round <= kMaxRestarts; ++round) {
        // (re)start the running pairs from their current x
        A(a.x, a.r, 0);
        ISAI_SYNC();
        ISAI_TILES(binit_body(t0 + sub, a.b, a.r, a.rh, a.p, a.v, B));
        ISAI_SYNC();
        ISAI_PAIRS({
            binit_final_body(B, p, lane, a.tol);
            __threadfence_block();  // lane 0's stores before the wave reads them back
            brho_final_body(B, p, lane);
        });
        ISAI_SYNC();
        int nactive = __hip_atomic_load(&ctl->nactive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // every pair stops within maxit iterations; the bound only guards the loop
        for (int it = 0; it <= ctl->maxit; ++it) {
            ISAI_TILES(bp_body(t0 + sub, a.r, a.p, a.v, x32, B));
            ISAI_SYNC();
            ML(a.p, tmp);
            ISAI_SYNC();
            MU(tmp, ph);
            ISAI_SYNC();
            A(ph, a.v, 0);
            ISAI_SYNC();
            ISAI_TILES(tile_dot_body(t0 + sub, a.rh, a.v, nullptr, B));
            ISAI_SYNC();
            ISAI_PAIRS(balpha_final_body(B, p, lane));
            ISAI_SYNC();
            ISAI_TILES(bs_body(t0 + sub, a.r, a.v, a.s, x32, B));
            ISAI_SYNC();
            ISAI_PAIRS(bsnorm_final_body(B, p, lane));
            ISAI_SYNC();
            // x of the pairs converged on s, and M_L s of the running ones (the two
            // touch different pairs and vectors)
            ISAI_TILES(bxs_body(t0 + sub, a.x, ph, B));
            ML(a.s, tmp);
            ISAI_SYNC();
            MU(tmp, sh);
            ISAI_SYNC();
            A(sh, a.t, 0);
            ISAI_SYNC();
            ISAI_TILES(tile_dot_body(t0 + sub, a.t, a.s, a.t, B));
            ISAI_SYNC();
            ISAI_PAIRS(bomega_final_body(B, p, lane));
            ISAI_SYNC();
            ISAI_TILES(bxr_body(t0 + sub, a.x, a.r, a.rh, ph, sh, a.s, a.t, B));
            ISAI_SYNC();
            // end of this iteration and rho / beta of the next
            ISAI_PAIRS({
                brnorm_final_body(B, p, lane);
                __threadfence_block();
                brho_final_body(B, p, lane);
            });
            ISAI_SYNC();
            const int now = __hip_atomic_load(&ctl->nactive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (now == nactive) break;  // no pair ran on
            nactive = now;
        }
        // true FP64 residual of every pair; pairs that miss tol restart
        A(a.x, a.r, 1);
        ISAI_SYNC();
        ISAI_TILES(bres_body(t0 + sub, a.b, a.r, B));
        ISAI_SYNC();
        ISAI_PAIRS(bres_final_body(B, p, lane, a.tol, a.relres));
        ISAI_SYNC();
        const int now = __hip_atomic_load(&ctl->nrestart, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (now == nrestart) break;
        nrestart = now;
    }
#undef ISAI_SYNC
#undef ISAI_TILES
#undef ISAI_PAIRS
}

// Control block before a persistent launch (a kernel, so the pair can be captured).
__global__ void persist_reset_kernel(Ctl *c, int maxit) {
    c->nactive = 0;
    c->nrestart = 0;
    c->abort = 0;
    c->arrive = 0;
    c->maxit = maxit;
}

}  // namespace isai

using namespace isai;

struct isai_batch {
    int nsys = 0, nb = 0, nnzb = 0, npairs = 0, ntiles = 0, precision = 0;
    size_t len = 0;
    const int *d_row = nullptr, *d_col = nullptr;
    const double *d_val = nullptr;
    isai_handle_t pre = nullptr;
    std::vector<int> sys_ptr, nrhs, sys_pair;
    std::vector<long long> sys_vec, sys_step;
    int *d_sys_of_row = nullptr, *d_sys_ptr = nullptr, *d_sys_nrhs = nullptr, *d_sys_pair = nullptr;
    long long *d_sys_vec = nullptr, *d_sys_step = nullptr;
    int *d_tile_pair = nullptr, *d_ptile = nullptr;
    PairState *d_ps = nullptr;
    double *d_part = nullptr, *d_relres = nullptr;
    Ctl *d_ctl = nullptr, *h_ctl = nullptr;
    double *work = nullptr;  // kWork multi-vectors
    static constexpr int kWork = 9;
    // launch shapes picked by the autotuner: [0] product with A, [1] with the ISAI factors
    int wg[2] = {256, 256}, remap[2] = {1, 1}, map[2] = {2, 2};
    // whole-row mappings: blocks held per row, from the operand's longest row (A: 8 /
    // 10 / 12 / 16; the ISAI factors: 4 / 6 / 8, longer rows take the MFMA kernel)
    int kb[2] = {16, 8};
    bool idx32 = false;    // whole-row mapping: value indices and multi-vector offsets fit 32 bits
    int variant = 0, ncand = 0;
    double tuned_us = 0;
    double table[2][24] = {};  // autotuner: us per product, [operand][candidate]
    double ms_ilu = 0, ms_isai = 0;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    double *g_x = nullptr;
    // the persistent kernel's captured launch and the arguments it was captured with
    hipGraphExec_t pexec = nullptr;
    hipGraph_t pgraph = nullptr;
    const double *p_b = nullptr;
    double *p_x = nullptr;
    double p_tol = 0.0;
    int p_maxit = -1;
    int ncu = 0;                     // persistent grid: one workgroup per CU
    long long spin_us = 5000000;     // bound on one grid-barrier wait

    void drop_graphs() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (pexec) (void)hipGraphExecDestroy(pexec);
        if (pgraph) (void)hipGraphDestroy(pgraph);
        gexec = nullptr;
        graph = nullptr;
        g_x = nullptr;
        pexec = nullptr;
        pgraph = nullptr;
        p_b = nullptr;
        p_x = nullptr;
        p_maxit = -1;
    }

    BatchDev dev() const {
        BatchDev B;
        B.sys_of_row = d_sys_of_row;
        B.sys_ptr = d_sys_ptr;
        B.sys_nrhs = d_sys_nrhs;
        B.sys_pair = d_sys_pair;
        B.sys_vec = d_sys_vec;
        B.sys_step = d_sys_step;
        B.tile_pair = d_tile_pair;
        B.ptile = d_ptile;
        B.ps = d_ps;
        B.part0 = d_part;
        B.part1 = d_part + ntiles;
        B.ctl = d_ctl;
        B.ntiles = ntiles;
        B.npairs = npairs;
        return B;
    }
};

namespace {

// Y = op X over the pairs still running (all = 1: over every pair).  X and Y are
// FP64 or FP32 multi-vectors (FP32 only for the outputs of reduced-precision ISAI
// products and the products that read them back).
template <int MAP, class XT, class YT>
int launch_spmm(isai_batch *b, Operand op, int prec, const int *rp, const int *ci, const void *v, const XT *X, YT *Y,
                int all, hipStream_t st) {
    const BatchDev B = b->dev();
    const int slot = op == kOpA ? 0 : 1;
    const int threads = b->wg[slot];
    const bool mfma = MAP == 2 || (MAP == 3 && prec != ISAI_PREC_FP64);  // factor whole-row form is MFMA
    const int rows = !mfma ? threads / 8 : (threads / 64) * (prec == ISAI_PREC_FP64 ? kMfmaF64Rows : kMfmaLowpRows);
    const unsigned grid = grid_for(b->nb, rows);
    const int remap = b->remap[slot];
    if (prec == ISAI_PREC_FP64) {
        if constexpr (std::is_same<YT, double>::value) {
            if (MAP == 2) {
                bspmm_mfma_f64_kernel<XT><<<grid, threads, 0, st>>>(b->nb, rp, ci, (const double *)v, B, X, Y, remap,
                                                                    all);
            } else if (MAP == 3 && b->idx32) {
                const int kb = b->kb[slot];
                const v2d *v2 = (const v2d *)v;
                if (kb <= 8)
                    bspmm_f64c_kernel<8, XT><<<grid, threads, 0, st>>>(b->nb, rp, ci, v2, B, X, Y, remap, all);
                else if (kb <= 10)
                    bspmm_f64c_kernel<10, XT><<<grid, threads, 0, st>>>(b->nb, rp, ci, v2, B, X, Y, remap, all);
                else if (kb <= 12)
                    bspmm_f64c_kernel<12, XT><<<grid, threads, 0, st>>>(b->nb, rp, ci, v2, B, X, Y, remap, all);
                else
                    bspmm_f64c_kernel<16, XT><<<grid, threads, 0, st>>>(b->nb, rp, ci, v2, B, X, Y, remap, all);
            } else {
                // row-pair for MAP 0 and for the whole-row mappings beyond 32-bit offsets
                bspmm_f64_kernel<MAP == 1 ? 1 : 0, XT><<<grid, threads, 0, st>>>(b->nb, rp, ci, (const v2d *)v, B, X,
                                                                                Y, remap, all);
            }
        } else {
            set_error("batch spmm: an FP64 operand writes FP64 vectors");
            return ISAI_ERR_INVALID_ARG;
        }
    } else if (MAP == 2) {
        if (prec == ISAI_PREC_FP32)
            bspmm_mfma_lowp_kernel<false, XT, YT><<<grid, threads, 0, st>>>(b->nb, rp, ci, v, B, X, Y, remap, all);
        else
            bspmm_mfma_lowp_kernel<true, XT, YT><<<grid, threads, 0, st>>>(b->nb, rp, ci, v, B, X, Y, remap, all);
    } else if (MAP == 3) {
        const int kb = b->kb[1];
        const bool bf = prec == ISAI_PREC_BF16;
#define ISAI_WR_LAUNCH(KB)                                                                                         \
    if (bf)                                                                                                        \
        bspmm_mfma_lowp_wr_kernel<true, KB, XT, YT><<<grid, threads, 0, st>>>(b->nb, rp, ci, v, B, X, Y, remap, all); \
    else                                                                                                           \
        bspmm_mfma_lowp_wr_kernel<false, KB, XT, YT><<<grid, threads, 0, st>>>(b->nb, rp, ci, v, B, X, Y, remap, all);
        if (kb <= 4) {
            ISAI_WR_LAUNCH(4)
        } else if (kb <= 6) {
            ISAI_WR_LAUNCH(6)
        } else {
            ISAI_WR_LAUNCH(8)
        }
#undef ISAI_WR_LAUNCH
    } else if (prec == ISAI_PREC_FP32) {
        bspmm_lowp_kernel<false, MAP == 1 ? 1 : 0, XT, YT>
            <<<grid, threads, 0, st>>>(b->nb, rp, ci, (const v4u *)v, B, X, Y, remap, all);
    } else {
        bspmm_lowp_kernel<true, MAP == 1 ? 1 : 0, XT, YT>
            <<<grid, threads, 0, st>>>(b->nb, rp, ci, (const v4u *)v, B, X, Y, remap, all);
    }
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

template <class XT, class YT>
int spmm_typed(isai_batch *b, Operand op, const XT *X, YT *Y, int all, hipStream_t st) {
    const int *rp = b->d_row, *ci = b->d_col;
    const void *v = b->d_val;
    int prec = ISAI_PREC_FP64;
    if (op != kOpA) {
        HandleFactors f;
        const int s = handle_factors_internal(b->pre, &f);
        if (s != ISAI_OK) return s;
        rp = op == kOpML ? f.Lrow : f.Urow;
        ci = op == kOpML ? f.Lcol : f.Ucol;
        v = op == kOpML ? f.Lval : f.Uval;
        prec = f.precision;
    }
    const int map = b->map[op == kOpA ? 0 : 1];
    if (map == 0) return launch_spmm<0>(b, op, prec, rp, ci, v, X, Y, all, st);
    if (map == 1) return launch_spmm<1>(b, op, prec, rp, ci, v, X, Y, all, st);
    if (map == 3) return launch_spmm<3>(b, op, prec, rp, ci, v, X, Y, all, st);
    return launch_spmm<2>(b, op, prec, rp, ci, v, X, Y, all, st);
}

// x_f32 / y_f32: X / Y hold FP32 values.
int spmm(isai_batch *b, Operand op, const void *X, bool x_f32, void *Y, bool y_f32, int all, hipStream_t st) {
    if (x_f32 && y_f32) return spmm_typed(b, op, (const float *)X, (float *)Y, all, st);
    if (x_f32) return spmm_typed(b, op, (const float *)X, (double *)Y, all, st);
    if (y_f32) return spmm_typed(b, op, (const double *)X, (float *)Y, all, st);
    return spmm_typed(b, op, (const double *)X, (double *)Y, all, st);
}

int spmm(isai_batch *b, Operand op, const double *X, double *Y, int all, hipStream_t st) {
    return spmm(b, op, X, false, Y, false, all, st);
}

#define BTRY(x)                       \
    do {                              \
        const int _s = (x);           \
        if (_s != ISAI_OK) return _s; \
    } while (0)

// One BiCGSTAB iteration of every running pair: kernels only, so it can be captured.
// With FP32 / BF16 factors the ISAI outputs (tmp, ph, sh) are held in FP32: they are
// FP32 accumulations, so the copy is exact and the iterates are those of FP64 storage.
int batch_iteration(isai_batch *b, double *x, hipStream_t st) {
    const size_t n = b->len;
    double *w = b->work;  // r, rh, p, v, ph, s, sh, t, tmp: n doubles each
    double *r = w, *rh = w + n, *p = w + 2 * n, *v = w + 3 * n, *s = w + 5 * n, *t = w + 7 * n;
    void *ph = w + 4 * n, *sh = w + 6 * n, *tmp = w + 8 * n;
    const bool f32 = b->precision != ISAI_PREC_FP64;
    // FP32 / BF16 factors: tmp holds n floats, and the second half of its region the
    // FP32 copy of p / s that M_L reads (the same values its loads would round to)
    float *x32 = f32 ? (float *)tmp + n : nullptr;
    const BatchDev B = b->dev();
    const int np = b->npairs, T = b->ntiles;
    zero_active_kernel<<<1, 1, 0, st>>>(b->d_ctl);
    brho_final<<<np, 64, 0, st>>>(B);
    bp_kernel<<<tile_grid(T), kTileWg, 0, st>>>(r, p, v, x32, B);
    BTRY(f32 ? spmm(b, kOpML, x32, true, tmp, true, 0, st) : spmm(b, kOpML, p, false, tmp, false, 0, st));
    BTRY(spmm(b, kOpMU, tmp, f32, ph, f32, 0, st));
    BTRY(spmm(b, kOpA, ph, f32, v, false, 0, st));
    tile_dot_kernel<<<tile_grid(T), kTileWg, 0, st>>>(rh, v, nullptr, B);
    balpha_final<<<np, 64, 0, st>>>(B);
    bs_kernel<<<tile_grid(T), kTileWg, 0, st>>>(r, v, s, x32, B);
    bsnorm_final<<<np, 64, 0, st>>>(B);
    if (f32) bxs_kernel<float><<<tile_grid(T), kTileWg, 0, st>>>(x, (const float *)ph, B);
    else bxs_kernel<double><<<tile_grid(T), kTileWg, 0, st>>>(x, (const double *)ph, B);
    BTRY(f32 ? spmm(b, kOpML, x32, true, tmp, true, 0, st) : spmm(b, kOpML, s, false, tmp, false, 0, st));
    BTRY(spmm(b, kOpMU, tmp, f32, sh, f32, 0, st));
    BTRY(spmm(b, kOpA, sh, f32, t, false, 0, st));
    tile_dot_kernel<<<tile_grid(T), kTileWg, 0, st>>>(t, s, t, B);
    bomega_final<<<np, 64, 0, st>>>(B);
    if (f32)
        bxr_kernel<float><<<tile_grid(T), kTileWg, 0, st>>>(x, r, rh, (const float *)ph, (const float *)sh, s, t, B);
    else
        bxr_kernel<double><<<tile_grid(T), kTileWg, 0, st>>>(x, r, rh, (const double *)ph, (const double *)sh, s, t, B);
    brnorm_final<<<np, 64, 0, st>>>(B);
    ISAI_LAUNCH_CHECK();
    return ISAI_OK;
}

// Launch-shape autotuner: for the product with A (FP64) and with the ISAI factors
// (storage precision) time every (lane mapping, workgroup size, XCD remap)
// candidate on the batch itself, all pairs active, and keep the fastest of each.
// Candidate c: mapping c / 6 (0 row-pair lanes, 1 column-split lanes, 2 MFMA,
// 3 whole-row: row-pair lanes for A, MFMA lanes for FP32 / BF16 factors, row-pair
// for FP64 factors), workgroup kWgCand[(c / 2) % 3], remap c % 2 (3: chunked remap).
constexpr int kWgCand[3] = {64, 128, kSpmmMaxWg};
constexpr int kCandPerOp = 24;

int autotune(isai_batch *b, hipStream_t st) {
    double *X = b->work, *Y = b->work + b->len;
    hipEvent_t e0, e1;
    ISAI_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        set_error("isai_batch_setup: hipEventCreate failed");
        return ISAI_ERR_HIP;
    }
    int status = ISAI_OK;
    double best_us[2] = {1e300, 1e300};
    int best[2] = {0, 0};
    for (int slot = 0; slot < 2 && status == ISAI_OK; ++slot) {
        const Operand op = slot == 0 ? kOpA : kOpML;
        for (int c = 0; c < kCandPerOp && status == ISAI_OK; ++c) {
            b->map[slot] = c / 6;
            b->wg[slot] = kWgCand[(c / 2) % 3];
            b->remap[slot] = c % 2;
            // each product as the iteration runs it: with FP32 / BF16 factors M_L reads
            // the FP32 copy of p / s and writes FP32, and A reads the FP32 p-hat / s-hat
            const bool lp = b->precision != ISAI_PREC_FP64;
            const bool xf = lp, yf = slot == 1 && lp;
            status = spmm(b, op, X, xf, Y, yf, 1, st);  // warm-up
            if (status != ISAI_OK) break;
            if (hipEventRecord(e0, st) != hipSuccess) status = ISAI_ERR_HIP;
            for (int k = 0; k < 5 && status == ISAI_OK; ++k) status = spmm(b, op, X, xf, Y, yf, 1, st);
            float ms = 0.f;
            if (status == ISAI_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                      hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                status = ISAI_ERR_HIP;
            if (status == ISAI_ERR_HIP) set_error("isai_batch_setup: autotuner timing failed");
            const double us = ms * 1e3 / 5;
            b->table[slot][c] = us;
            if (status == ISAI_OK && us < best_us[slot]) {
                best_us[slot] = us;
                best[slot] = c;
            }
        }
        b->map[slot] = best[slot] / 6;
        b->wg[slot] = kWgCand[(best[slot] / 2) % 3];
        b->remap[slot] = best[slot] % 2;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->ncand = 2 * kCandPerOp;
    b->variant = best[0];
    b->tuned_us = best_us[0];
    return status;
}

int capture_graph(isai_batch *b, double *d_X, hipStream_t st) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->g_x = nullptr;
    ISAI_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int s = ISAI_OK;
    for (int k = 0; k < kItersPerReplay && s == ISAI_OK; ++k) s = batch_iteration(b, d_X, st);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (s != ISAI_OK) {
        if (g) (void)hipGraphDestroy(g);
        return s;
    }
    ISAI_HIP(e);
    b->graph = g;
    ISAI_HIP(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    b->g_x = d_X;
    return ISAI_OK;
}

typedef void (*PersistFn)(PersistArgs);

template <int MA, int MM, int KBA = 8>
PersistFn persist_fn_prec(int prec) {
    if (prec == ISAI_PREC_FP64) return batch_persistent_kernel<MA, MM, ISAI_PREC_FP64, KBA>;
    if (prec == ISAI_PREC_FP32) return batch_persistent_kernel<MA, MM, ISAI_PREC_FP32, KBA>;
    return batch_persistent_kernel<MA, MM, ISAI_PREC_BF16, KBA>;
}

// The persistent kernel holds the row-pair (0), MFMA (2) and whole-row (3, rows of
// up to 10 blocks) mappings for A; a column-split choice (1), or a whole-row choice
// on longer rows, runs as row-pair there, and the factors' whole-row choice as MFMA.  The whole-row products equal the row-pair
// ones bit for bit, and the whole-row A product fits the kernel's 128-register
// budget without spilling (the row-pair one spills).
PersistFn persist_fn(int map_a, int map_m, int prec, int kb_a, bool idx32) {
    // the factors' whole-row form (3) equals the MFMA one (2) bit for bit
    const bool mm = map_m == 2 || (map_m == 3 && prec != ISAI_PREC_FP64);
    if (map_a == 3 && idx32 && kb_a <= 10) {
        if (kb_a <= 8) return mm ? persist_fn_prec<3, 2, 8>(prec) : persist_fn_prec<3, 0, 8>(prec);
        return mm ? persist_fn_prec<3, 2, 10>(prec) : persist_fn_prec<3, 0, 10>(prec);
    }
    const bool ma = map_a == 2;
    if (ma && mm) return persist_fn_prec<2, 2>(prec);
    if (ma) return persist_fn_prec<2, 0>(prec);
    if (mm) return persist_fn_prec<0, 2>(prec);
    return persist_fn_prec<0, 0>(prec);
}

// Queues the whole solve as one persistent launch (after its control reset), or
// replays the captured pair when graph is set (recaptured when an argument changes).
int persistent_solve(isai_batch *b, double tol, int maxit, const double *d_B, double *d_X, bool graph,
                     hipStream_t st) {
    HandleFactors f;
    BTRY(handle_factors_internal(b->pre, &f));
    PersistArgs a;
    a.B = b->dev();
    a.nb = b->nb;
    a.rp = b->d_row;
    a.ci = b->d_col;
    a.val = b->d_val;
    a.Lrow = f.Lrow;
    a.Lcol = f.Lcol;
    a.Urow = f.Urow;
    a.Ucol = f.Ucol;
    a.Lval = f.Lval;
    a.Uval = f.Uval;
    a.b = d_B;
    a.x = d_X;
    const size_t n = b->len;
    double *w = b->work;  // r, rh, p, v, ph, s, sh, t, tmp: n doubles each
    a.r = w;
    a.rh = w + n;
    a.p = w + 2 * n;
    a.v = w + 3 * n;
    a.ph = w + 4 * n;
    a.s = w + 5 * n;
    a.sh = w + 6 * n;
    a.t = w + 7 * n;
    a.tmp = w + 8 * n;
    a.n = n;
    a.tol = tol;
    a.relres = b->d_relres;
    const PersistFn fn = persist_fn(b->map[0], b->map[1], f.precision, b->kb[0], b->idx32);
    int occ = 0;
    ISAI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *)fn, kPersistThreads, 0));
    if (occ < 1 || b->ncu < 1) {
        set_error("isai_batch_solve: the persistent kernel does not fit one workgroup per CU (occupancy %d)", occ);
        return ISAI_ERR_UNSUPPORTED;
    }
    auto launch = [&]() -> int {
        persist_reset_kernel<<<1, 1, 0, st>>>(b->d_ctl, maxit);
        ISAI_LAUNCH_CHECK();
        void *args[] = {&a};
        ISAI_HIP(hipLaunchKernel((const void *)fn, dim3(b->ncu), dim3(kPersistThreads), args, 0, st));
        return ISAI_OK;
    };
    if (!graph) return launch();
    if (!b->pexec || b->p_b != d_B || b->p_x != d_X || b->p_tol != tol || b->p_maxit != maxit) {
        if (b->pexec) (void)hipGraphExecDestroy(b->pexec);
        if (b->pgraph) (void)hipGraphDestroy(b->pgraph);
        b->pexec = nullptr;
        b->pgraph = nullptr;
        b->p_maxit = -1;
        ISAI_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int s = launch();
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (s != ISAI_OK) {
            if (g) (void)hipGraphDestroy(g);
            return s;
        }
        ISAI_HIP(e);
        b->pgraph = g;
        ISAI_HIP(hipGraphInstantiate(&b->pexec, g, nullptr, nullptr, 0));
        b->p_b = d_B;
        b->p_x = d_X;
        b->p_tol = tol;
        b->p_maxit = maxit;
    }
    ISAI_HIP(hipGraphLaunch(b->pexec, st));
    return ISAI_OK;
}

int batch_results(isai_batch *b, double tol, int maxit, const std::vector<PairState> &hs, int *iters,
                  double *relres, int *max_iters);

}  // namespace

extern "C" {

int isai_batch_destroy(isai_batch_t b) {
    if (!b) return ISAI_OK;
    b->drop_graphs();
    if (b->pre) isai_destroy(b->pre);
    (void)hipFree(b->d_sys_of_row);
    (void)hipFree(b->d_sys_ptr);
    (void)hipFree(b->d_sys_nrhs);
    (void)hipFree(b->d_sys_pair);
    (void)hipFree(b->d_sys_vec);
    (void)hipFree(b->d_sys_step);
    (void)hipFree(b->d_tile_pair);
    (void)hipFree(b->d_ptile);
    (void)hipFree(b->d_ps);
    (void)hipFree(b->d_part);
    (void)hipFree(b->d_relres);
    (void)hipFree(b->d_ctl);
    (void)hipFree(b->work);
    if (b->h_ctl) (void)hipHostFree(b->h_ctl);
    delete b;
    return ISAI_OK;
}

int isai_batch_setup(int nsys, const int *sys_ptr, const int *nrhs, int nnzb, const int *d_row, const int *d_col,
                     const double *d_val, int precision, isai_batch_t *out, isai_stream_t stream) {
    ISAI_REQUIRE(out && sys_ptr && nrhs && d_row && d_col && d_val && nsys > 0 && nnzb > 0,
                 "isai_batch_setup: bad argument");
    ISAI_REQUIRE(precision == ISAI_PREC_FP64 || precision == ISAI_PREC_FP32 || precision == ISAI_PREC_BF16,
                 "isai_batch_setup: unknown precision %d", precision);
    *out = nullptr;
    hipStream_t st = (hipStream_t)stream;
    auto *b = new isai_batch;
    auto fail = [&](int code) {
        isai_batch_destroy(b);
        return code;
    };
    b->nsys = nsys;
    b->precision = precision;
    b->sys_ptr.assign(sys_ptr, sys_ptr + nsys + 1);
    b->nrhs.assign(nrhs, nrhs + nsys);
    if (b->sys_ptr[0] != 0) {
        set_error("isai_batch_setup: sys_ptr[0] must be 0 (got %d)", b->sys_ptr[0]);
        return fail(ISAI_ERR_INVALID_ARG);
    }
    for (int s = 0; s < nsys; ++s) {
        if (b->sys_ptr[s + 1] <= b->sys_ptr[s] || b->nrhs[s] < 1 || b->nrhs[s] > kMaxRhs) {
            set_error("isai_batch_setup: system %d has %d block rows and %d right-hand sides (need >= 1 and 1..%d)",
                      s, b->sys_ptr[s + 1] - b->sys_ptr[s], b->nrhs[s], kMaxRhs);
            return fail(ISAI_ERR_INVALID_ARG);
        }
    }
    b->nb = b->sys_ptr[nsys];
    b->nnzb = nnzb;
    b->d_row = d_row;
    b->d_col = d_col;
    b->d_val = d_val;
    // every block must stay inside its system's diagonal block
    {
        std::vector<int> rp(b->nb + 1), ci(nnzb);
        if (hipMemcpy(rp.data(), d_row, sizeof(int) * (b->nb + 1), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(ci.data(), d_col, sizeof(int) * nnzb, hipMemcpyDeviceToHost) != hipSuccess) {
            set_error("isai_batch_setup: reading the block pattern failed");
            return fail(ISAI_ERR_HIP);
        }
        if (rp[0] != 0 || rp[b->nb] != nnzb) {
            set_error("isai_batch_setup: row pointer spans [%d, %d), expected [0, %d)", rp[0], rp[b->nb], nnzb);
            return fail(ISAI_ERR_INVALID_ARG);
        }
        int maxlen = 0;
        for (int i = 0; i < b->nb; ++i) maxlen = std::max(maxlen, rp[i + 1] - rp[i]);
        b->kb[0] = maxlen <= 8 ? 8 : (maxlen <= 10 ? 10 : (maxlen <= 12 ? 12 : 16));
        for (int s = 0; s < nsys; ++s)
            for (int i = b->sys_ptr[s]; i < b->sys_ptr[s + 1]; ++i)
                for (int q = rp[i]; q < rp[i + 1]; ++q)
                    if (ci[q] < b->sys_ptr[s] || ci[q] >= b->sys_ptr[s + 1]) {
                        set_error("isai_batch_setup: block (%d,%d) couples system %d to another system", i, ci[q], s);
                        return fail(ISAI_ERR_PATTERN);
                    }
    }
    // pairs, padded segments, tiles
    std::vector<int> sys_of_row(b->nb), tile_pair, ptile{0};
    long long off = 0;
    for (int s = 0; s < nsys; ++s) {
        for (int i = b->sys_ptr[s]; i < b->sys_ptr[s + 1]; ++i) sys_of_row[i] = s;
        const long long seg = 4LL * (b->sys_ptr[s + 1] - b->sys_ptr[s]);
        const long long step = (seg + kTile - 1) / kTile * kTile;
        b->sys_pair.push_back(b->npairs);
        b->sys_vec.push_back(off);
        b->sys_step.push_back(step);
        for (int j = 0; j < b->nrhs[s]; ++j) {
            for (long long t = 0; t < step / kTile; ++t) tile_pair.push_back(b->npairs);
            ptile.push_back((int)tile_pair.size());
            ++b->npairs;
            off += step;
        }
    }
    b->len = (size_t)off;
    b->ntiles = (int)tile_pair.size();
    b->idx32 = 8LL * nnzb < (1LL << 32) && off < (1LL << 32);
    auto up = [](void **d, const void *h, size_t bytes) -> bool {
        if (hipMalloc(d, bytes) != hipSuccess) return false;
        return hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    const bool ok = up((void **)&b->d_sys_of_row, sys_of_row.data(), sizeof(int) * b->nb) &&
                    up((void **)&b->d_sys_ptr, b->sys_ptr.data(), sizeof(int) * (nsys + 1)) &&
                    up((void **)&b->d_sys_nrhs, b->nrhs.data(), sizeof(int) * nsys) &&
                    up((void **)&b->d_sys_pair, b->sys_pair.data(), sizeof(int) * nsys) &&
                    up((void **)&b->d_sys_vec, b->sys_vec.data(), sizeof(long long) * nsys) &&
                    up((void **)&b->d_sys_step, b->sys_step.data(), sizeof(long long) * nsys) &&
                    up((void **)&b->d_tile_pair, tile_pair.data(), sizeof(int) * b->ntiles) &&
                    up((void **)&b->d_ptile, ptile.data(), sizeof(int) * (b->npairs + 1)) &&
                    hipMalloc(&b->d_ps, sizeof(PairState) * b->npairs) == hipSuccess &&
                    hipMalloc(&b->d_part, sizeof(double) * 2 * b->ntiles) == hipSuccess &&
                    hipMalloc(&b->d_relres, sizeof(double) * b->npairs) == hipSuccess &&
                    hipMalloc(&b->d_ctl, sizeof(Ctl)) == hipSuccess &&
                    hipHostMalloc(&b->h_ctl, sizeof(Ctl)) == hipSuccess &&
                    hipMalloc(&b->work, sizeof(double) * b->len * isai_batch::kWork) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        set_error("isai_batch_setup: allocation failed (%d pairs, %zu doubles per multi-vector)", b->npairs, b->len);
        return fail(ISAI_ERR_ALLOC);
    }
    if (hipMemsetAsync(b->work, 0, sizeof(double) * b->len * isai_batch::kWork, st) != hipSuccess ||
        hipMemsetAsync(b->d_ps, 0, sizeof(PairState) * b->npairs, st) != hipSuccess ||
        hipMemsetAsync(b->d_ctl, 0, sizeof(Ctl), st) != hipSuccess) {
        set_error("isai_batch_setup: clearing the work vectors failed");
        return fail(ISAI_ERR_HIP);
    }
    // preconditioner of the whole block-diagonal matrix: the ILU(0), the pair lists
    // and the ISAI build each run as one launch over every system
    isai_setup_opts_t opts{};
    opts.precond = ISAI_PRECOND_ISAI;
    opts.precision = precision;
    int s = isai_setup(b->nb, nnzb, 4, d_row, d_col, d_val, &opts, &b->pre, stream);
    if (s != ISAI_OK) return fail(s);
    double total = 0;
    s = isai_setup_times(b->pre, &b->ms_ilu, &b->ms_isai, &total);
    if (s != ISAI_OK) return fail(s);
    {  // the factors' longest row (their stored, possibly padded, patterns)
        HandleFactors f;
        s = handle_factors_internal(b->pre, &f);
        if (s != ISAI_OK) return fail(s);
        std::vector<int> lr(b->nb + 1), ur(b->nb + 1);
        if (hipMemcpy(lr.data(), f.Lrow, sizeof(int) * (b->nb + 1), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(ur.data(), f.Urow, sizeof(int) * (b->nb + 1), hipMemcpyDeviceToHost) != hipSuccess) {
            set_error("isai_batch_setup: reading the factor patterns failed");
            return fail(ISAI_ERR_HIP);
        }
        int maxlen = 0;
        for (int i = 0; i < b->nb; ++i) maxlen = std::max({maxlen, lr[i + 1] - lr[i], ur[i + 1] - ur[i]});
        b->kb[1] = maxlen <= 4 ? 4 : (maxlen <= 6 ? 6 : 8);
    }
    s = autotune(b, st);
    if (s != ISAI_OK) return fail(s);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipMemcpyAsync(&b->d_ctl->spin_us, &b->spin_us, sizeof(long long), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        set_error("isai_batch_setup: reading the CU count failed");
        return fail(ISAI_ERR_HIP);
    }
    *out = b;
    return ISAI_OK;
}

int isai_batch_vector_length(isai_batch_t b, size_t *length) {
    ISAI_REQUIRE(b && length, "isai_batch_vector_length: null argument");
    *length = b->len;
    return ISAI_OK;
}

int isai_batch_vector_offset(isai_batch_t b, int sys, int rhs, size_t *offset) {
    ISAI_REQUIRE(b && offset, "isai_batch_vector_offset: null argument");
    ISAI_REQUIRE(sys >= 0 && sys < b->nsys && rhs >= 0 && rhs < b->nrhs[sys],
                 "isai_batch_vector_offset: no pair (system %d, right-hand side %d)", sys, rhs);
    *offset = (size_t)(b->sys_vec[sys] + rhs * b->sys_step[sys]);
    return ISAI_OK;
}

int isai_batch_pairs(isai_batch_t b, int *npairs) {
    ISAI_REQUIRE(b && npairs, "isai_batch_pairs: null argument");
    *npairs = b->npairs;
    return ISAI_OK;
}

int isai_batch_tuning(isai_batch_t b, int *variant, double *us, int *ncand) {
    ISAI_REQUIRE(b, "isai_batch_tuning: null batch");
    if (variant) *variant = b->variant;
    if (us) *us = b->tuned_us;
    if (ncand) *ncand = b->ncand;
    return ISAI_OK;
}

int isai_batch_tuning_table(isai_batch_t b, int operand, double *us, int capacity, int *count) {
    ISAI_REQUIRE(b && (operand == 0 || operand == 1), "isai_batch_tuning_table: bad argument");
    if (count) *count = kCandPerOp;
    for (int c = 0; c < kCandPerOp && c < capacity && us; ++c) us[c] = b->table[operand][c];
    return ISAI_OK;
}

int isai_batch_select(isai_batch_t b, int operand, int candidate) {
    ISAI_REQUIRE(b && (operand == 0 || operand == 1) && candidate >= 0 && candidate < kCandPerOp,
                 "isai_batch_select: bad operand %d / candidate %d", operand, candidate);
    b->map[operand] = candidate / 6;
    b->wg[operand] = kWgCand[(candidate / 2) % 3];
    b->remap[operand] = candidate % 2;
    b->drop_graphs();
    return ISAI_OK;
}

int isai_batch_set_spin_limit(isai_batch_t b, long long us) {
    ISAI_REQUIRE(b && us > 0, "isai_batch_set_spin_limit: bad argument");
    b->spin_us = us;
    ISAI_HIP(hipMemcpy(&b->d_ctl->spin_us, &us, sizeof(long long), hipMemcpyHostToDevice));
    return ISAI_OK;
}

int isai_batch_setup_times(isai_batch_t b, double *ms_ilu, double *ms_isai) {
    ISAI_REQUIRE(b, "isai_batch_setup_times: null batch");
    if (ms_ilu) *ms_ilu = b->ms_ilu;
    if (ms_isai) *ms_isai = b->ms_isai;
    return ISAI_OK;
}

int isai_batch_spmm(isai_batch_t b, const double *d_X, double *d_Y, isai_stream_t stream) {
    ISAI_REQUIRE(b && d_X && d_Y, "isai_batch_spmm: null argument");
    return spmm(b, kOpA, d_X, d_Y, 1, (hipStream_t)stream);
}

int isai_batch_solve(isai_batch_t b, double tol, int maxit, const double *d_B, double *d_X, int use_graph, int *iters,
                     double *relres, int *max_iters, isai_stream_t stream) {
    ISAI_REQUIRE(b && d_B && d_X && tol > 0.0 && maxit >= 0, "isai_batch_solve: bad argument");
    ISAI_REQUIRE(use_graph >= 0 && use_graph <= 3, "isai_batch_solve: unknown mode %d", use_graph);
    hipStream_t st = (hipStream_t)stream;
    ISAI_REQUIRE(use_graph == 0 || use_graph == 3 || st != nullptr,
                 "isai_batch_solve: graph capture needs a created (non-null) stream");
    const size_t n = b->len;
    double *r = b->work, *rh = r + n, *p = rh + n, *v = p + n;
    const BatchDev B = b->dev();
    const int np = b->npairs, pg = (np + 255) / 256, T = b->ntiles;
    std::vector<PairState> hs(np);
    if (use_graph >= 2) {
        BTRY(persistent_solve(b, tol, maxit, d_B, d_X, use_graph == 2, st));
        Ctl hc;
        ISAI_HIP(hipMemcpyAsync(hs.data(), b->d_ps, sizeof(PairState) * np, hipMemcpyDeviceToHost, st));
        ISAI_HIP(hipMemcpyAsync(&hc, b->d_ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
        ISAI_HIP(hipStreamSynchronize(st));
        if (hc.abort) {
            set_error("isai_batch_solve: a grid-barrier wait of the persistent kernel exceeded %lld us", b->spin_us);
            return ISAI_ERR_TIMEOUT;
        }
        return batch_results(b, tol, maxit, hs, iters, relres, max_iters);
    }
    if (use_graph && (b->g_x != d_X || !b->gexec)) BTRY(capture_graph(b, d_X, st));
    *b->h_ctl = Ctl{0, maxit, 0, 0, 0ULL, b->spin_us};
    ISAI_HIP(hipMemcpyAsync(b->d_ctl, b->h_ctl, sizeof(Ctl), hipMemcpyHostToDevice, st));
    breset_kernel<<<pg, 256, 0, st>>>(B, np);
    ISAI_LAUNCH_CHECK();
    const int max_replays = maxit / kItersPerReplay + 2;
    for (int round = 0; round <= kMaxRestarts; ++round) {
        // (re)start the running pairs from their current x
        BTRY(spmm(b, kOpA, d_X, r, 0, st));
        binit_kernel<<<tile_grid(T), kTileWg, 0, st>>>(d_B, r, rh, p, v, B);
        binit_final<<<np, 64, 0, st>>>(B, tol);
        ISAI_LAUNCH_CHECK();
        for (int rep = 0; rep < max_replays; ++rep) {
            if (use_graph) {
                ISAI_HIP(hipGraphLaunch(b->gexec, st));
            } else {
                for (int k = 0; k < kItersPerReplay; ++k) BTRY(batch_iteration(b, d_X, st));
            }
            ISAI_HIP(hipMemcpyAsync(&b->h_ctl->nactive, &b->d_ctl->nactive, sizeof(int), hipMemcpyDeviceToHost, st));
            ISAI_HIP(hipStreamSynchronize(st));
            if (b->h_ctl->nactive == 0) break;
        }
        // true FP64 residual of every pair; pairs that miss tol restart
        BTRY(spmm(b, kOpA, d_X, r, 1, st));
        bres_kernel<<<tile_grid(T), kTileWg, 0, st>>>(d_B, r, B);
        bres_final<<<np, 64, 0, st>>>(B, tol, b->d_relres);
        ISAI_LAUNCH_CHECK();
        ISAI_HIP(hipMemcpyAsync(hs.data(), b->d_ps, sizeof(PairState) * np, hipMemcpyDeviceToHost, st));
        ISAI_HIP(hipStreamSynchronize(st));
        bool any = false;
        for (const PairState &q : hs) any = any || q.status == 0;
        if (!any) break;
    }
    return batch_results(b, tol, maxit, hs, iters, relres, max_iters);
}

}  // extern "C"

namespace {

// Per-pair iteration counts and true relative residuals of a finished solve.
int batch_results(isai_batch *b, double tol, int maxit, const std::vector<PairState> &hs, int *iters,
                  double *relres, int *max_iters) {
    const int np = b->npairs;
    std::vector<double> rr(np);
    ISAI_HIP(hipMemcpy(rr.data(), b->d_relres, sizeof(double) * np, hipMemcpyDeviceToHost));
    int mx = 0;
    bool all = true;
    for (int q = 0; q < np; ++q) {
        if (iters) iters[q] = hs[q].iters;
        if (relres) relres[q] = rr[q];
        mx = std::max(mx, hs[q].iters);
        all = all && rr[q] <= tol;
    }
    if (max_iters) *max_iters = mx;
    if (!all) {
        set_error("isai_batch_solve: some pairs did not reach %.1e within %d iterations", tol, maxit);
        return ISAI_ERR_NOT_CONVERGED;
    }
    return ISAI_OK;
}

}  // namespace
